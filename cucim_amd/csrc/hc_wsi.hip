// Whole-slide image reader: tiled (and stripped) TIFF / BigTIFF pyramids,
// tile decode and region assembly into host or device memory.
//
// Reference behaviour: CuImage::read_region (cpp/src/cuimage.cpp) over the
// cuslide TIFF plugin (cpp/plugins/cucim.kit.cuslide/src/cuslide/tiff/
// tiff.cpp, ifd.cpp): a region of a pyramid level is assembled from the
// tiles it overlaps, pixels outside the level are the background value (255
// for Aperio SVS / Philips TIFF, 0 otherwise), output is interleaved 8-bit
// samples (YXC).  Codecs: none, deflate (zlib), LZW, baseline JPEG (with the
// TIFF JPEGTables stream), horizontal predictor for the lossless ones.
//
// Design: the TIFF structure is parsed once on the host.  Lossless tiles are
// read and inflated on host threads.  JPEG tiles are entropy-decoded and
// inverse-transformed by the same __host__ __device__ functions either on
// host threads (device output not requested) or on the GPU: the compressed
// tiles are uploaded, one lane per tile decodes its Huffman stream into DCT
// blocks, and a second launch runs the inverse DCT, chroma upsampling and
// YCbCr -> RGB conversion straight into the device region.
#include <fcntl.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hc_common.h"

namespace hc {
namespace {

// ------------------------------------------------------------- TIFF --
struct Ifd {
  int64_t width = 0, height = 0, tile_w = 0, tile_h = 0;
  int spp = 1, bits = 8, compression = 1, photometric = 2, planar = 1, predictor = 1;
  int64_t subfile = 0;
  bool tiled = false;
  std::vector<uint64_t> offsets, counts;
  std::vector<uint8_t> jpeg_tables;
  std::string description;
  double xres = 0, yres = 0;
  int res_unit = 2;
  int64_t tiles_x() const { return (width + tile_w - 1) / tile_w; }
  int64_t tiles_y() const { return (height + tile_h - 1) / tile_h; }
};

struct Wsi {
  int fd = -1;
  bool big = false, le = true;
  std::vector<Ifd> ifds;
  std::vector<int> levels;  // IFD of each pyramid level
  std::vector<std::pair<std::string, int>> assoc;  // associated image name, IFD
  uint8_t background = 0;
  std::string vendor = "generic";
  ~Wsi() {
    if (fd >= 0) close(fd);
  }
};

bool read_at(int fd, uint64_t off, void* dst, size_t n) {
  uint8_t* p = static_cast<uint8_t*>(dst);
  while (n > 0) {
    const ssize_t r = pread(fd, p, n, (off_t)off);
    if (r <= 0) return false;
    p += r;
    n -= (size_t)r;
    off += (uint64_t)r;
  }
  return true;
}

struct Reader {
  int fd;
  bool le;
  uint64_t u(const uint8_t* p, int n) const {
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v |= (uint64_t)p[le ? i : n - 1 - i] << (8 * i);
    return v;
  }
};

int type_size(int t) {
  switch (t) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: case 13: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    default: return 0;
  }
}

// Values of one IFD entry as unsigned integers (rationals as num, den pairs).
bool entry_values(const Wsi& w, const Reader& rd, int type, uint64_t count,
                  const uint8_t* inline_val, std::vector<uint64_t>& out,
                  std::vector<uint8_t>* raw = nullptr) {
  const int ts = type_size(type);
  if (ts == 0) return true;  // unknown type: ignored
  if (count > (1ull << 32)) return false;
  const uint64_t nbytes = count * (uint64_t)ts;
  std::vector<uint8_t> buf(nbytes);
  const int inl = w.big ? 8 : 4;
  if (nbytes <= (uint64_t)inl) {
    std::memcpy(buf.data(), inline_val, nbytes);
  } else {
    const uint64_t off = rd.u(inline_val, inl);
    if (!read_at(w.fd, off, buf.data(), nbytes)) return false;
  }
  if (raw) *raw = buf;
  if (type == 5 || type == 10) {
    out.resize(2 * count);
    for (uint64_t i = 0; i < 2 * count; ++i) out[i] = rd.u(&buf[4 * i], 4);
  } else if (type == 2 || type == 7 || type == 1) {
    out.resize(count);
    for (uint64_t i = 0; i < count; ++i) out[i] = buf[i];
  } else {
    out.resize(count);
    for (uint64_t i = 0; i < count; ++i) out[i] = rd.u(&buf[i * ts], ts);
  }
  return true;
}

int parse_tiff(Wsi& w) {
  uint8_t hdr[16];
  if (!read_at(w.fd, 0, hdr, 8)) return fail(HC_ERR_VALUE, "not a TIFF file (short header)");
  if (hdr[0] == 'I' && hdr[1] == 'I') w.le = true;
  else if (hdr[0] == 'M' && hdr[1] == 'M') w.le = false;
  else return fail(HC_ERR_VALUE, "not a TIFF file (byte order mark)");
  Reader rd{w.fd, w.le};
  const uint64_t magic = rd.u(hdr + 2, 2);
  uint64_t next;
  if (magic == 42) {
    next = rd.u(hdr + 4, 4);
  } else if (magic == 43) {
    if (!read_at(w.fd, 0, hdr, 16)) return fail(HC_ERR_VALUE, "short BigTIFF header");
    if (rd.u(hdr + 4, 2) != 8) return fail(HC_ERR_VALUE, "unsupported BigTIFF offset size");
    w.big = true;
    next = rd.u(hdr + 8, 8);
  } else {
    return fail(HC_ERR_VALUE, "not a TIFF file (version)");
  }
  const int esz = w.big ? 20 : 12, csz = w.big ? 8 : 2, nsz = w.big ? 8 : 4;
  int guard = 0;
  while (next != 0) {
    if (++guard > 4096) return fail(HC_ERR_VALUE, "TIFF IFD chain does not end");
    uint8_t cb[8];
    if (!read_at(w.fd, next, cb, csz)) return fail(HC_ERR_VALUE, "truncated IFD");
    const uint64_t n = rd.u(cb, csz);
    if (n > 65535) return fail(HC_ERR_VALUE, "implausible IFD entry count");
    std::vector<uint8_t> ent(n * esz + nsz);
    if (!read_at(w.fd, next + csz, ent.data(), ent.size()))
      return fail(HC_ERR_VALUE, "truncated IFD entries");
    Ifd f;
    int64_t rows_per_strip = -1;
    for (uint64_t e = 0; e < n; ++e) {
      const uint8_t* p = &ent[e * esz];
      const int tag = (int)rd.u(p, 2), type = (int)rd.u(p + 2, 2);
      const uint64_t count = w.big ? rd.u(p + 4, 8) : rd.u(p + 4, 4);
      const uint8_t* val = p + (w.big ? 12 : 8);
      std::vector<uint64_t> v;
      std::vector<uint8_t> raw;
      const bool want_raw = tag == 270 || tag == 347;
      if (!entry_values(w, rd, type, count, val, v, want_raw ? &raw : nullptr))
        return fail(HC_ERR_VALUE, "truncated IFD value");
      if (v.empty()) continue;
      switch (tag) {
        case 254: f.subfile = (int64_t)v[0]; break;
        case 256: f.width = (int64_t)v[0]; break;
        case 257: f.height = (int64_t)v[0]; break;
        case 258: f.bits = (int)v[0]; break;
        case 259: f.compression = (int)v[0]; break;
        case 262: f.photometric = (int)v[0]; break;
        case 270: f.description.assign(raw.begin(), raw.end());
                  while (!f.description.empty() && f.description.back() == 0)
                    f.description.pop_back();
                  break;
        case 273: case 324: f.offsets = v; break;
        case 277: f.spp = (int)v[0]; break;
        case 278: rows_per_strip = (int64_t)v[0]; break;
        case 279: case 325: f.counts = v; break;
        case 282: if (v.size() >= 2 && v[1]) f.xres = (double)v[0] / (double)v[1]; break;
        case 283: if (v.size() >= 2 && v[1]) f.yres = (double)v[0] / (double)v[1]; break;
        case 284: f.planar = (int)v[0]; break;
        case 296: f.res_unit = (int)v[0]; break;
        case 317: f.predictor = (int)v[0]; break;
        case 322: f.tile_w = (int64_t)v[0]; f.tiled = true; break;
        case 323: f.tile_h = (int64_t)v[0]; break;
        case 347: f.jpeg_tables = raw; break;
        default: break;
      }
    }
    if (!f.tiled) {
      f.tile_w = f.width;
      f.tile_h = rows_per_strip > 0 ? std::min(rows_per_strip, f.height) : f.height;
    }
    if (f.width <= 0 || f.height <= 0 || f.tile_w <= 0 || f.tile_h <= 0)
      return fail(HC_ERR_VALUE, "TIFF IFD without image geometry");
    if ((int64_t)f.offsets.size() < f.tiles_x() * f.tiles_y() || f.counts.size() < f.offsets.size())
      return fail(HC_ERR_VALUE, "TIFF IFD: fewer tile offsets / byte counts than tiles");
    w.ifds.push_back(std::move(f));
    const uint8_t* np = &ent[n * esz];
    next = rd.u(np, nsz);
  }
  if (w.ifds.empty()) return fail(HC_ERR_VALUE, "TIFF file without images");
  const std::string& d = w.ifds[0].description;
  if (d.rfind("Aperio", 0) == 0) {
    w.vendor = "aperio";
    w.background = 0xFF;
  } else if (d.find("DPUfsImport") != std::string::npos ||
             d.find("<DataObject ObjectType=\"DPUfsImport\"") != std::string::npos) {
    w.vendor = "philips";
    w.background = 0xFF;
  }
  // pyramid levels (tiff.cpp:362-385, 670-700): every IFD, except that for
  // Aperio / Philips files the untiled IFDs after the first are associated
  // images (thumbnail, label, macro); sorted by width, then height, largest
  // first
  for (int i = 0; i < (int)w.ifds.size(); ++i)
    if (i == 0 || w.vendor == "generic" || w.ifds[i].tiled) w.levels.push_back(i);
  // associated images of Aperio files (tiff.cpp:670-700): an untiled IFD 1
  // of subfile type 0 is the thumbnail, subfile type 1 the label, 9 the
  // macro image.  (Philips files name theirs in the XML description, which
  // is not read here.)
  if (w.vendor == "aperio")
    for (int i = 1; i < (int)w.ifds.size(); ++i) {
      const Ifd& f = w.ifds[i];
      if (f.tiled) continue;
      const char* name = nullptr;
      if (i == 1 && f.subfile == 0) name = "thumbnail";
      else if (f.subfile == 1) name = "label";
      else if (f.subfile == 9) name = "macro";
      if (name) w.assoc.emplace_back(name, i);
    }
  std::stable_sort(w.levels.begin(), w.levels.end(), [&](int a, int b) {
    const Ifd &fa = w.ifds[a], &fb = w.ifds[b];
    return fa.width != fb.width ? fa.width > fb.width : fa.height > fb.height;
  });
  return HC_OK;
}

int check_level_codec(const Ifd& f) {
  if (f.bits != 8) return fail(HC_ERR_UNSUPPORTED, "only 8-bit samples are supported");
  if (f.planar != 1 && f.spp > 1)
    return fail(HC_ERR_UNSUPPORTED, "only interleaved (chunky) samples are supported");
  if (f.spp < 1 || f.spp > 4) return fail(HC_ERR_UNSUPPORTED, "1 to 4 samples per pixel");
  const int c = f.compression;
  if (c != 1 && c != 5 && c != 7 && c != 8 && c != 32946)
    return fail(HC_ERR_UNSUPPORTED, "compression not supported (none, LZW, JPEG, deflate)");
  if (c == 7 && f.predictor != 1) return fail(HC_ERR_UNSUPPORTED, "JPEG with a predictor");
  if (f.predictor != 1 && f.predictor != 2)
    return fail(HC_ERR_UNSUPPORTED, "predictor not supported (none, horizontal)");
  return HC_OK;
}

// Output samples per pixel of a level (JPEG YCbCr / RGB tiles give RGB).
int out_channels(const Ifd& f) {
  if (f.compression == 7) return f.spp >= 3 ? 3 : 1;
  return f.spp;
}

// ------------------------------------------------------- lossless --
// TIFF LZW (variable 9..12-bit codes, MSB first, code 256 clears, 257 ends,
// code width grows one code early).
bool lzw_decode(const uint8_t* src, size_t n, uint8_t* dst, size_t cap) {
  std::vector<uint16_t> prefix(4096);
  std::vector<uint8_t> suffix(4096), first(4096);
  std::vector<uint8_t> stack(4097);
  for (int i = 0; i < 256; ++i) {
    prefix[i] = 0xFFFF;
    suffix[i] = (uint8_t)i;
    first[i] = (uint8_t)i;
  }
  int next = 258, width = 9, prev = -1;
  uint64_t bitbuf = 0;
  int bits = 0;
  size_t pos = 0, out = 0;
  while (true) {
    while (bits < width && pos < n) {
      bitbuf = (bitbuf << 8) | src[pos++];
      bits += 8;
    }
    if (bits < width) break;
    const int code = (int)((bitbuf >> (bits - width)) & ((1u << width) - 1));
    bits -= width;
    if (code == 257) break;
    if (code == 256) {
      next = 258;
      width = 9;
      prev = -1;
      continue;
    }
    int c = code;
    int sp = 0;
    uint8_t head;
    if (code < next) {
      head = first[code];
    } else if (code == next && prev >= 0) {
      head = first[prev];
      stack[sp++] = head;  // the KwKwK case: last byte is the first of prev
      c = prev;
    } else {
      return false;
    }
    while (c >= 0 && c != 0xFFFF) {
      if (sp >= 4097) return false;
      stack[sp++] = suffix[c];
      c = c < 256 ? 0xFFFF : prefix[c];
    }
    if (out + (size_t)sp > cap) sp = (int)(cap - out);
    for (int i = sp - 1; i >= 0; --i) dst[out++] = stack[i];
    if (prev >= 0 && next < 4096) {
      prefix[next] = (uint16_t)prev;
      suffix[next] = head;
      first[next] = first[prev];
      ++next;
    }
    prev = code;
    if (next + 1 >= (1 << width) && width < 12) ++width;
    if (out >= cap) break;
  }
  return true;
}

void undo_predictor(uint8_t* buf, int64_t w, int64_t h, int spp) {
  for (int64_t y = 0; y < h; ++y) {
    uint8_t* r = buf + y * w * spp;
    for (int64_t x = 1; x < w; ++x)
      for (int c = 0; c < spp; ++c) r[x * spp + c] = (uint8_t)(r[x * spp + c] + r[(x - 1) * spp + c]);
  }
}

// ------------------------------------------------------------ JPEG --
constexpr int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Decoding tables of one Huffman table (JPEG spec F.2.2.3): maxcode /
// valptr / mincode per code length, the values, and a 9-bit lookup of the
// short codes (length << 8 | value, 0 when the code is longer).
struct HuffTab {
  int32_t maxcode[18];
  int32_t valptr[17];
  int32_t mincode[17];
  uint8_t vals[256];
  uint16_t lut[512];
};

struct JComp {
  int id, h, v, tq, td, ta;
  int bw, bh;          // blocks per row / column of the component (MCU padded)
  int64_t block0;      // index of the component's first block in the tile
};

// One JPEG tile: everything the entropy decoder and the inverse transform
// need, with table indices into the per-tile table arrays.
struct JTile {
  int64_t data_off, data_len;  // entropy-coded segment(s) in the upload
  int width, height;           // tile pixels
  int ncomp;
  JComp comp[3];
  int hmax, vmax, mcux, mcuy;
  int restart;                 // MCUs per restart interval (0: none)
  int transform;               // 1: YCbCr -> RGB, 0: components as stored
  int64_t nblocks;
  int64_t coef0;               // first DCT block in the coefficient buffer
  int tabset;                  // index of the tile's JTables
  int x0, y0;                  // tile origin in the level
  int vis_w, vis_h;            // pixels of the tile inside the level
};

// Quantization (natural order) and Huffman tables in force for a tile: a
// level's JPEGTables stream, overridden by tables inside the tile's stream.
struct JTables {
  int32_t qtab[4][64];
  HuffTab dc[2], ac[2];
};

bool build_huff(const uint8_t* counts, const uint8_t* vals, int nvals, HuffTab& t) {
  std::memset(&t, 0, sizeof(t));
  int k = 0, code = 0;
  for (int l = 1; l <= 16; ++l) {
    t.valptr[l] = k;
    t.mincode[l] = code;
    code += counts[l - 1];
    k += counts[l - 1];
    if (code > (1 << l)) return false;  // over-subscribed code lengths
    t.maxcode[l] = counts[l - 1] ? code - 1 : -1;
    code <<= 1;
  }
  t.maxcode[17] = 0x7fffffff;
  if (k > 256 || k > nvals) return false;
  std::memcpy(t.vals, vals, (size_t)k);
  // lookup of codes of up to 9 bits
  k = 0;
  code = 0;
  for (int l = 1; l <= 16; ++l) {
    for (int i = 0; i < counts[l - 1]; ++i, ++k, ++code) {
      if (l <= 9) {
        const int base = code << (9 - l);
        for (int j = 0; j < (1 << (9 - l)); ++j)
          t.lut[base + j] = (uint16_t)((l << 8) | t.vals[k]);
      }
    }
    code <<= 1;
  }
  return true;
}

// Marker-level JPEG parser.  segs() reads the segments of one stream: DQT /
// DHT into tb, the frame and scan into jt (a tables-only stream has neither);
// finish() checks the tables the scan references and derives the MCU
// geometry.  A level's JPEGTables stream is parsed once; each tile's stream
// is parsed on top of a copy of the resulting state.
struct JParser {
  JTables* tb = nullptr;
  bool have_q[4] = {false, false, false, false};
  bool have_dc[2] = {false, false}, have_ac[2] = {false, false};
  bool have_frame = false, saw_jfif = false, own_tables = false;
  int adobe = -1;

  int segs(const uint8_t* b, size_t len, bool tile, JTile& jt) {
    size_t p = 0;
    while (p + 1 < len) {
      if (b[p] != 0xFF) {
        ++p;
        continue;
      }
      const int m = b[p + 1];
      p += 2;
      if (m == 0xD8 || m == 0x01 || (m >= 0xD0 && m <= 0xD7) || m == 0xFF) continue;
      if (m == 0xD9) return HC_OK;
      if (p + 2 > len) return fail(HC_ERR_VALUE, "truncated JPEG marker");
      const size_t L = ((size_t)b[p] << 8) | b[p + 1];
      if (L < 2 || p + L > len) return fail(HC_ERR_VALUE, "truncated JPEG segment");
      const uint8_t* d = b + p + 2;
      const size_t dl = L - 2;
      if (m == 0xDB) {  // DQT
        if (tile) own_tables = true;
        size_t q = 0;
        while (q < dl) {
          const int pq = d[q] >> 4, tq = d[q] & 15;
          if (tq > 3) return fail(HC_ERR_VALUE, "JPEG quantization table id");
          if (q + 1 + (pq ? 128 : 64) > dl) return fail(HC_ERR_VALUE, "truncated JPEG DQT");
          ++q;
          for (int i = 0; i < 64; ++i) {
            int v;
            if (pq) {
              v = (d[q] << 8) | d[q + 1];
              q += 2;
            } else {
              v = d[q++];
            }
            tb->qtab[tq][kZigzag[i]] = v;
          }
          have_q[tq] = true;
        }
      } else if (m == 0xC4) {  // DHT
        if (tile) own_tables = true;
        size_t q = 0;
        while (q + 17 <= dl) {
          const int tc = d[q] >> 4, th = d[q] & 15;
          if (th > 1 || tc > 1) return fail(HC_ERR_UNSUPPORTED, "JPEG Huffman table id above 1");
          int nv = 0;
          for (int i = 0; i < 16; ++i) nv += d[q + 1 + i];
          if (q + 17 + (size_t)nv > dl) return fail(HC_ERR_VALUE, "truncated JPEG Huffman table");
          HuffTab& t = tc ? tb->ac[th] : tb->dc[th];
          if (!build_huff(d + q + 1, d + q + 17, nv, t))
            return fail(HC_ERR_VALUE, "bad JPEG Huffman table");
          (tc ? have_ac : have_dc)[th] = true;
          q += 17 + (size_t)nv;
        }
      } else if (m == 0xC0 || m == 0xC1) {  // baseline / extended sequential
        if (dl < 6) return fail(HC_ERR_VALUE, "truncated JPEG frame");
        if (d[0] != 8) return fail(HC_ERR_UNSUPPORTED, "JPEG sample precision other than 8");
        jt.height = (d[1] << 8) | d[2];
        jt.width = (d[3] << 8) | d[4];
        jt.ncomp = d[5];
        if (jt.ncomp != 1 && jt.ncomp != 3)
          return fail(HC_ERR_UNSUPPORTED, "JPEG with other than 1 or 3 components");
        if (dl < 6 + 3 * (size_t)jt.ncomp) return fail(HC_ERR_VALUE, "truncated JPEG frame");
        for (int c = 0; c < jt.ncomp; ++c) {
          jt.comp[c].id = d[6 + 3 * c];
          jt.comp[c].h = d[7 + 3 * c] >> 4;
          jt.comp[c].v = d[7 + 3 * c] & 15;
          jt.comp[c].tq = d[8 + 3 * c] & 3;
          if (jt.comp[c].h < 1 || jt.comp[c].h > 2 || jt.comp[c].v < 1 || jt.comp[c].v > 2)
            return fail(HC_ERR_UNSUPPORTED, "JPEG sampling factors other than 1 or 2");
        }
        have_frame = true;
      } else if (m >= 0xC2 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
        return fail(HC_ERR_UNSUPPORTED, "progressive, lossless or arithmetic JPEG");
      } else if (m == 0xDD) {  // DRI
        if (dl < 2) return fail(HC_ERR_VALUE, "truncated JPEG DRI");
        jt.restart = (d[0] << 8) | d[1];
      } else if (m == 0xE0 && dl >= 5 && std::memcmp(d, "JFIF", 5) == 0) {
        saw_jfif = true;
      } else if (m == 0xEE && dl >= 12 && std::memcmp(d, "Adobe", 5) == 0) {
        adobe = d[11] == 0 ? 0 : 1;
      } else if (m == 0xDA) {  // SOS
        if (!tile || !have_frame) return fail(HC_ERR_VALUE, "JPEG scan before the frame");
        const int ns = d[0];
        if (ns != jt.ncomp) return fail(HC_ERR_UNSUPPORTED, "JPEG with more than one scan");
        if (dl < 1 + 2 * (size_t)ns) return fail(HC_ERR_VALUE, "truncated JPEG scan header");
        for (int i = 0; i < ns; ++i) {
          const int cs = d[1 + 2 * i], t = d[2 + 2 * i];
          int c = 0;
          while (c < jt.ncomp && jt.comp[c].id != cs) ++c;
          if (c == jt.ncomp) return fail(HC_ERR_VALUE, "JPEG scan component not in the frame");
          jt.comp[c].td = t >> 4;
          jt.comp[c].ta = t & 15;
          if (jt.comp[c].td > 1 || jt.comp[c].ta > 1)
            return fail(HC_ERR_UNSUPPORTED, "JPEG Huffman table id above 1");
        }
        jt.data_off = (int64_t)(p + L);
        // entropy data runs to the EOI marker (restart markers inside)
        size_t e = p + L;
        while (e + 1 < len && !(b[e] == 0xFF && b[e + 1] == 0xD9)) ++e;
        if (e + 1 >= len) e = len;  // no EOI: the data runs to the end
        jt.data_len = (int64_t)(e - (p + L));
        return HC_OK;
      }
      p += L;
    }
    return tile ? fail(HC_ERR_VALUE, "JPEG tile without a scan") : HC_OK;
  }

  int finish(JTile& jt) {
    for (int c = 0; c < jt.ncomp; ++c) {
      if (!have_q[jt.comp[c].tq] || !have_dc[jt.comp[c].td] || !have_ac[jt.comp[c].ta])
        return fail(HC_ERR_VALUE, "JPEG table referenced but not defined");
    }
    // a one-component scan is non-interleaved: one block per MCU in raster
    // order whatever the frame's sampling factors (JPEG spec A.2.2)
    if (jt.ncomp == 1) jt.comp[0].h = jt.comp[0].v = 1;
    jt.hmax = jt.vmax = 1;
    for (int c = 0; c < jt.ncomp; ++c) {
      jt.hmax = std::max(jt.hmax, jt.comp[c].h);
      jt.vmax = std::max(jt.vmax, jt.comp[c].v);
    }
    jt.mcux = (jt.width + 8 * jt.hmax - 1) / (8 * jt.hmax);
    jt.mcuy = (jt.height + 8 * jt.vmax - 1) / (8 * jt.vmax);
    int64_t nb = 0;
    for (int c = 0; c < jt.ncomp; ++c) {
      JComp& k = jt.comp[c];
      k.bw = jt.mcux * k.h;
      k.bh = jt.mcuy * k.v;
      k.block0 = nb;
      nb += (int64_t)k.bw * k.bh;
    }
    jt.nblocks = nb;
    // colour space of a three-component stream as libjpeg's header defaults
    // decide: JFIF -> YCbCr; else an Adobe marker's transform (0: RGB);
    // else component ids 'R','G','B' -> RGB; else YCbCr
    if (jt.ncomp != 3) {
      jt.transform = 0;
    } else if (saw_jfif) {
      jt.transform = 1;
    } else if (adobe >= 0) {
      jt.transform = adobe;
    } else {
      const bool rgb_ids = jt.comp[0].id == 'R' && jt.comp[1].id == 'G' && jt.comp[2].id == 'B';
      jt.transform = rgb_ids ? 0 : 1;
    }
    return HC_OK;
  }
};

// Whether a tile stream carries its own DQT / DHT segments before its scan.
bool stream_has_tables(const uint8_t* b, size_t len) {
  size_t p = 2;
  while (p + 4 <= len && b[p] == 0xFF) {
    const int m = b[p + 1];
    if (m == 0xDA) return false;
    if (m == 0xDB || m == 0xC4) return true;
    if (m == 0xFF) {
      ++p;
      continue;
    }
    p += 2 + (((size_t)b[p + 2] << 8) | b[p + 3]);
  }
  return true;  // unusual layout: parse it fully
}

// Whole parse of one tile: the tables stream, then the tile's stream.
int parse_jpeg(const uint8_t* tables, size_t tlen, const uint8_t* s, size_t n, JTile& jt,
               JTables& tb) {
  std::memset(&jt, 0, sizeof(jt));
  JParser ps;
  ps.tb = &tb;
  if (tables && tlen) {
    const int rc = ps.segs(tables, tlen, false, jt);
    if (rc != HC_OK) return rc;
  }
  const int rc = ps.segs(s, n, true, jt);
  if (rc != HC_OK) return rc;
  return ps.finish(jt);
}

// Bit reader over an entropy-coded segment: 0xFF 0x00 is a data 0xFF;
// another marker ends the data (zeros are fed past it).
struct BitReader {
  const uint8_t* p;
  int64_t n, pos;
  uint32_t buf;
  int bits;
  __host__ __device__ void fill() {
    while (bits <= 24) {
      uint32_t byte = 0;
      if (pos < n) {
        byte = p[pos];
        if (byte == 0xFF) {
          const uint32_t nx = pos + 1 < n ? p[pos + 1] : 0;
          if (nx == 0x00) {
            pos += 2;
          } else {
            byte = 0;  // a marker: stop here
          }
        } else {
          ++pos;
        }
      }
      buf |= byte << (24 - bits);
      bits += 8;
    }
  }
  __host__ __device__ uint32_t peek(int k) {
    if (bits < k) fill();
    return buf >> (32 - k);
  }
  __host__ __device__ void skip(int k) {
    buf <<= k;
    bits -= k;
  }
  __host__ __device__ int get(int k) {
    if (k == 0) return 0;
    const uint32_t v = peek(k);
    skip(k);
    return (int)v;
  }
  // skip to the byte after the next RSTn marker
  __host__ __device__ void restart() {
    while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) ++pos;
    if (pos + 1 < n) pos += 2;
    buf = 0;
    bits = 0;
  }
};

__host__ __device__ inline int huff_decode(BitReader& br, const HuffTab& t) {
  const uint32_t look = br.peek(9);
  const uint16_t e = t.lut[look];
  if (e) {
    br.skip(e >> 8);
    return e & 0xFF;
  }
  int code = (int)br.peek(9);
  br.skip(9);
  int l = 9;
  while (l < 16 && code > t.maxcode[l]) {
    code = (code << 1) | br.get(1);
    ++l;
  }
  if (code > t.maxcode[l]) return 0;  // corrupt stream: treat as symbol 0
  return t.vals[t.valptr[l] + code - t.mincode[l]];
}

__host__ __device__ inline int extend(int v, int s) {
  return (s > 0 && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v;
}

// Entropy-decode one tile into quantized coefficients (natural order),
// block b of component c at coef[(block0_c + b) * 64].
__host__ __device__ void jpeg_entropy(const JTile& jt, const JTables& tb, const uint8_t* data,
                                    int16_t* coef) {
  BitReader br{data + jt.data_off, jt.data_len, 0, 0u, 0};
  int pred[3] = {0, 0, 0};
  const int64_t nmcu = (int64_t)jt.mcux * jt.mcuy;
  int64_t todo = jt.restart;
  for (int64_t m = 0; m < nmcu; ++m) {
    if (jt.restart && todo == 0) {
      br.restart();
      pred[0] = pred[1] = pred[2] = 0;
      todo = jt.restart;
    }
    const int my = (int)(m / jt.mcux), mx = (int)(m - (int64_t)my * jt.mcux);
    for (int c = 0; c < jt.ncomp; ++c) {
      const JComp& k = jt.comp[c];
      const HuffTab& dct = tb.dc[k.td];
      const HuffTab& act = tb.ac[k.ta];
      // non-interleaved single-component scans hold blocks in raster order;
      // one component with h = v = 1 is the same as interleaved order
      for (int by = 0; by < k.v; ++by)
        for (int bx = 0; bx < k.h; ++bx) {
          const int64_t b = (int64_t)(my * k.v + by) * k.bw + (mx * k.h + bx);
          int16_t* blk = coef + (k.block0 + b) * 64;
          for (int i = 0; i < 64; ++i) blk[i] = 0;
          int s = huff_decode(br, dct);
          if (s > 15) s = 0;  // not a DC category (corrupt table)
          pred[c] += extend(br.get(s), s);
          blk[0] = (int16_t)pred[c];
          for (int i = 1; i < 64;) {
            const int rs = huff_decode(br, act);
            const int r = rs >> 4, sz = rs & 15;
            if (sz == 0) {
              if (r != 15) break;
              i += 16;
              continue;
            }
            i += r;
            // libjpeg's natural-order table maps the positions past 63 of a
            // corrupt stream to 63; the loop then ends
            blk[kZigzag[i < 63 ? i : 63]] = (int16_t)extend(br.get(sz), sz);
            ++i;
          }
        }
    }
    if (jt.restart) --todo;
  }
}

// 8x8 inverse DCT of one block of quantized coefficients (natural order) to
// 8-bit samples: the integer Loeffler-Ligtenberg-Moschytz factorisation of
// IJG libjpeg's "islow" method (13-bit constants round(c * 2^13), 2 extra
// bits between the column and the row pass), the method the reference's
// TurboJPEG call selects with its default flags (cuslide/jpeg/
// libjpeg_turbo.cpp:116-121).  Dequantization is an int product, as in
// libjpeg; the final "& 1023" table lookup of libjpeg's range limiting is
// restated by idct_limit().
__host__ __device__ inline uint8_t idct_limit(int x) {
  const int m = x & 1023;
  if (m < 128) return (uint8_t)(m + 128);
  if (m < 512) return 255;
  if (m < 896) return 0;
  return (uint8_t)(m - 896);
}

__host__ __device__ void idct_islow(const int16_t* blk, const int32_t* q, uint8_t* out,
                                   int stride) {
  constexpr int kCB = 13, kP1 = 2;
  // round(c * 8192) of the rotation constants
  constexpr int32_t c0298 = 2446, c0390 = 3196, c0541 = 4433, c0765 = 6270, c0899 = 7373,
                    c1175 = 9633, c1501 = 12299, c1847 = 15137, c1961 = 16069, c2053 = 16819,
                    c2562 = 20995, c3072 = 25172;
  int32_t ws[64];
  // one 1-D pass over 8 values a[0..7] spaced by 'step'; results r[0..7]
  auto pass = [&](const int32_t* a, int32_t* r) {
    int32_t z2 = a[2], z3 = a[6];
    int32_t z1 = (z2 + z3) * c0541;
    const int32_t t2 = z1 - z3 * c1847;
    const int32_t t3 = z1 + z2 * c0765;
    const int32_t t0 = (int32_t)((uint32_t)(a[0] + a[4]) << kCB);
    const int32_t t1 = (int32_t)((uint32_t)(a[0] - a[4]) << kCB);
    const int32_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
    int32_t o0 = a[7], o1 = a[5], o2 = a[3], o3 = a[1];
    z1 = o0 + o3;
    z2 = o1 + o2;
    z3 = o0 + o2;
    int32_t z4 = o1 + o3;
    const int32_t z5 = (z3 + z4) * c1175;
    o0 *= c0298;
    o1 *= c2053;
    o2 *= c3072;
    o3 *= c1501;
    z1 *= -c0899;
    z2 *= -c2562;
    z3 = z3 * -c1961 + z5;
    z4 = z4 * -c0390 + z5;
    o0 += z1 + z3;
    o1 += z2 + z4;
    o2 += z2 + z3;
    o3 += z1 + z4;
    r[0] = e10 + o3;
    r[7] = e10 - o3;
    r[1] = e11 + o2;
    r[6] = e11 - o2;
    r[2] = e12 + o1;
    r[5] = e12 - o1;
    r[3] = e13 + o0;
    r[4] = e13 - o0;
  };
  for (int x = 0; x < 8; ++x) {  // columns: dequantize, transform, keep kP1 extra bits
    int32_t a[8], r[8];
    for (int v = 0; v < 8; ++v) a[v] = (int32_t)blk[v * 8 + x] * q[v * 8 + x];
    pass(a, r);
    constexpr int sh = kCB - kP1;
    for (int y = 0; y < 8; ++y) ws[y * 8 + x] = (r[y] + (1 << (sh - 1))) >> sh;
  }
  for (int y = 0; y < 8; ++y) {  // rows: transform, remove all scaling (incl. the 1/8)
    int32_t r[8];
    pass(ws + y * 8, r);
    constexpr int sh = kCB + kP1 + 3;
    for (int x = 0; x < 8; ++x) out[y * stride + x] = idct_limit((r[x] + (1 << (sh - 1))) >> sh);
  }
}

// One component plane sample at full resolution: libjpeg's "fancy"
// (triangle-filter) chroma upsampling for 2x1, 1x2 and 2x2 factors, which
// the reference's decoder keeps on (its TJFLAG_FASTUPSAMPLE is not set).
// plane holds the component's decoded samples, rows of 'stride' bytes; dw x
// dh are the component's own dimensions (ceil(image * factor / max factor)):
// the edge columns use the one-sided weights, and rows above the first or
// below row dh - 1 repeat the edge row, as libjpeg's context rows do.
__host__ __device__ inline int up_sample(const uint8_t* plane, int stride, int dw, int dh, int fx,
                                          int fy, int x, int y) {
  if (fx == 1 && fy == 1) return plane[(int64_t)y * stride + x];
  const int sx = x / fx, sy = y / fy;
  auto at = [&](int xx, int yy) {
    xx = xx < 0 ? 0 : (xx >= dw ? dw - 1 : xx);
    yy = yy < 0 ? 0 : (yy >= dh ? dh - 1 : yy);
    return (int)plane[(int64_t)yy * stride + xx];
  };
  if (fy == 1) {  // 2x1: 3/4 nearer + 1/4 farther sample, alternating bias
    if (x == 0) return at(0, sy);
    if (x == dw * 2 - 1) return at(dw - 1, sy);
    const int c = at(sx, sy);
    if ((x & 1) == 0) return (3 * c + at(sx - 1, sy) + 1) >> 2;
    return (3 * c + at(sx + 1, sy) + 2) >> 2;
  }
  // 2x2: vertical 3:1 column sums, then horizontal 3:1 of the sums
  const int ny = (y & 1) ? sy + 1 : sy - 1;
  auto col = [&](int xx) { return 3 * at(xx, sy) + at(xx, ny); };
  const int c = col(sx);
  if (fx == 1) return (c + ((y & 1) ? 2 : 1)) >> 2;  // 1x2 (vertical only)
  if (x == 0) return (4 * c + 8) >> 4;
  if (x == dw * 2 - 1) return (4 * c + 7) >> 4;
  if ((x & 1) == 0) return (3 * c + col(sx - 1) + 8) >> 4;
  return (3 * c + col(sx + 1) + 7) >> 4;
}

__host__ __device__ inline uint8_t clamp8(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

// YCbCr -> RGB (JFIF), 16-bit fixed point with rounding.
__host__ __device__ inline void ycc_rgb(int Y, int Cb, int Cr, uint8_t* o) {
  const int cb = Cb - 128, cr = Cr - 128;
  constexpr int kR = 91881, kGb = 22554, kGr = 46802, kB = 116130;  // 1.402, .34414, .71414, 1.772
  o[0] = clamp8(Y + ((kR * cr + 32768) >> 16));
  o[1] = clamp8(Y + ((-kGb * cb - kGr * cr + 32768) >> 16));
  o[2] = clamp8(Y + ((kB * cb + 32768) >> 16));
}

// Decoded tile -> region: planes[c] hold each component at its own
// resolution (pw = bw * 8 wide); pixel (tx, ty) of the tile goes to region
// pixel (tx + x0 - rx, ty + y0 - ry) when it lies inside the region.
// What the per-pixel colour stage needs of a decoded tile, in registers:
// each component's plane, row stride, own dimensions and upsampling factors.
struct PixCtx {
  const uint8_t* pl[3];
  int stride[3], dw[3], dh[3], fx[3], fy[3];
  int ncomp, transform;
};

__host__ __device__ inline PixCtx pix_ctx(const JTile& jt, uint8_t* const* planes) {
  PixCtx p;
  p.ncomp = jt.ncomp;
  p.transform = jt.transform;
  for (int c = 0; c < 3; ++c) {
    const int cc = c < jt.ncomp ? c : 0;
    const JComp& k = jt.comp[cc];
    p.pl[c] = planes[cc];
    p.stride[c] = k.bw * 8;
    p.fx[c] = jt.hmax / k.h;
    p.fy[c] = jt.vmax / k.v;
    p.dw[c] = (jt.width * k.h + jt.hmax - 1) / jt.hmax;
    p.dh[c] = (jt.height * k.v + jt.vmax - 1) / jt.vmax;
  }
  return p;
}

// Tile pixel (tx, ty) into dst (nch samples).
__host__ __device__ inline void jpeg_pixel(const PixCtx& p, int tx, int ty, uint8_t* dst,
                                           int nch) {
  int v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    v[c] = c < p.ncomp ? up_sample(p.pl[c], p.stride[c], p.dw[c], p.dh[c], p.fx[c], p.fy[c],
                                   tx, ty)
                       : 0;
  if (p.ncomp == 1) {
    for (int c = 0; c < nch; ++c) dst[c] = (uint8_t)v[0];
  } else if (p.transform) {
    ycc_rgb(v[0], v[1], v[2], dst);
  } else {
    dst[0] = (uint8_t)v[0];
    dst[1] = (uint8_t)v[1];
    dst[2] = (uint8_t)v[2];
  }
}

// ------------------------------------------------------- device side --
struct JDev {
  const JTile* tiles;
  const JTables* tabsets;
  const uint8_t* data;
  int16_t* coef;
  uint8_t* planes;      // decoded component planes: a tile's at coef0 * 64
  int ntiles;
  uint8_t* out;         // region, rows of rw * nch bytes
  int64_t rx, ry, rw, rh;
  int nch;
};

// Entropy decoding of one tile per work-group, from LDS.  The 64 lanes copy
// the tile's four Huffman tables into LDS and remove the byte stuffing of
// its entropy-coded segment (0xFF 0x00 -> 0xFF) while staging it in LDS
// (16 bytes per lane per round, positions by a wave prefix sum), then lane 0
// decodes: Huffman lookups and bit reads hit LDS instead of global memory.
// Coefficients were zeroed beforehand; only the non-zero ones are stored.
// Tiles with restart markers, or whose segment exceeds the LDS staging
// capacity, are decoded by lane 0 from global memory (jpeg_entropy).
constexpr int kTabBytes = 4 * (int)sizeof(HuffTab);

struct LdsBits {
  const uint8_t* s;  // destuffed bytes in LDS, 4-byte aligned, zero-padded to a word past n
  int n, pos, bits;
  uint64_t buf;      // next bits, MSB first
  // 32 more bits from two aligned LDS words (bits <= 32 on entry)
  __device__ void refill() {
    uint32_t w = 0;
    if (pos < n) {
      const uint32_t* s32 = reinterpret_cast<const uint32_t*>(s);
      const int q = pos >> 2, r = (pos & 3) * 8;
      const uint32_t lo = s32[q], hi = s32[q + 1];
      w = r ? (lo >> r) | (hi << (32 - r)) : lo;  // bytes pos..pos+3, little-endian
      w = __builtin_bswap32(w);
    }
    buf |= (uint64_t)w << (32 - bits);
    bits += 32;
    pos += 4;
  }
  __device__ uint32_t peek(int k) {
    if (bits < k) refill();
    return (uint32_t)(buf >> (64 - k));
  }
  __device__ void skip(int k) {
    buf <<= k;
    bits -= k;
  }
  __device__ int get(int k) {
    if (k == 0) return 0;
    const uint32_t v = peek(k);
    skip(k);
    return (int)v;
  }
};

__device__ inline int huff_lds(LdsBits& br, const HuffTab& t) {
  const uint32_t look = br.peek(16);
  const uint16_t e = t.lut[look >> 7];
  if (e) {
    br.skip(e >> 8);
    return e & 0xFF;
  }
  int l = 10;
  int code = (int)(look >> 6);
  while (l < 16 && code > t.maxcode[l]) {
    ++l;
    code = (int)(look >> (16 - l));
  }
  if (code > t.maxcode[l]) {
    br.skip(16);
    return 0;  // corrupt stream: symbol 0
  }
  br.skip(l);
  return t.vals[t.valptr[l] + code - t.mincode[l]];
}

// Decoder state at a codeword boundary: bit position, unit of the MCU (the
// block within it) and coefficient index (0: the DC difference is next).
struct JState {
  int64_t bit;
  int u, z;
};

// What a lane needs to decode part of one tile's destuffed segment.
struct JSeg {
  const uint8_t* s;
  int n;
  const HuffTab* tabs;  // dc0, dc1, ac0, ac1
  const uint8_t* zz;
  const int8_t *ucomp, *ubx, *uby;
  int U, mcux;
  int total;            // blocks of the tile
  int16_t* coef;
  int dtab[3], atab[3], ch[3], cv[3], cbw[3];
  int64_t cb0[3];
};

__device__ inline int16_t* jseg_block(const JSeg& sg, int g, int u) {
  const int m = g / sg.U;
  const int c = sg.ucomp[u];
  const int my = m / sg.mcux, mx = m - my * sg.mcux;
  const int64_t bidx =
      (int64_t)(my * sg.cv[c] + sg.uby[u]) * sg.cbw[c] + (mx * sg.ch[c] + sg.ubx[u]);
  return sg.coef + (sg.cb0[c] + bidx) * 64;
}

// DC coefficient of the k-th block of component c in decode order.
__device__ inline int16_t* jseg_dc(const JSeg& sg, int c, int hv, int k) {
  const int m = k / hv;
  const int r = k - m * hv;
  const int by = r / sg.ch[c], bx = r - by * sg.ch[c];
  const int my = m / sg.mcux, mx = m - my * sg.mcux;
  const int64_t bidx = (int64_t)(my * sg.cv[c] + by) * sg.cbw[c] + (mx * sg.ch[c] + bx);
  return sg.coef + (sg.cb0[c] + bidx) * 64;
}

// Decode symbols from state st until the first codeword boundary at or past
// end_bit (or until the tile's last block is complete).  g: index of the
// block in progress (decode order), nblk: blocks completed.  With WRITE the
// values are stored (DC as the difference, AC at its natural position).
template <bool WRITE>
__device__ JState jseg_run(const JSeg& sg, JState st, int64_t end_bit, int& g, int& nblk) {
  LdsBits br{sg.s, sg.n, 0, 0, 0ull};
  br.pos = (int)((st.bit >> 5) << 2);
  br.refill();
  br.refill();
  br.skip((int)(st.bit & 31));
  int u = st.u, z = st.z;
  auto tell = [&]() { return (int64_t)br.pos * 8 - br.bits; };
  int16_t* blk = nullptr;
  if (WRITE && g < sg.total) blk = jseg_block(sg, g, u);
  while (tell() < end_bit && g < sg.total) {
    const int c = sg.ucomp[u];
    if (z == 0) {
      int sdc = huff_lds(br, sg.tabs[sg.dtab[c]]);
      if (sdc > 15) sdc = 0;
      const int diff = extend(br.get(sdc), sdc);
      if (WRITE) blk[0] = (int16_t)diff;
      z = 1;
      continue;
    }
    const int rs = huff_lds(br, sg.tabs[sg.atab[c]]);
    const int r = rs >> 4, sz = rs & 15;
    bool done;
    if (sz == 0) {
      z += 16;
      done = r != 15 || z >= 64;
    } else {
      z += r;
      const int v = extend(br.get(sz), sz);
      if (WRITE) blk[sg.zz[z < 63 ? z : 63]] = (int16_t)v;
      ++z;
      done = z >= 64;
    }
    if (done) {
      z = 0;
      u = u + 1 == sg.U ? 0 : u + 1;
      ++g;
      ++nblk;
      if (WRITE && g < sg.total) blk = jseg_block(sg, g, u);
    }
  }
  return JState{tell(), u, z};
}

// Lanes of the entropy kernel's work-group: one tile's segment is split into
// kJLanes parts.
constexpr int kJWaves = 4, kJLanes = 64 * kJWaves;

__device__ inline int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive prefix sum of v over the work-group's kJLanes threads (total in
// *total); sh holds kJWaves ints.  Contains two barriers.
__device__ inline int block_excl_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int incl = wave_incl_scan(v, lane);
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kJWaves; ++k) {
    const int t = sh[k];
    off += k < w ? t : 0;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Restart intervals (DRI) decode independently: each starts byte-aligned
// after its RST marker with the DC predictors at zero.  Thread tid decodes
// intervals tid, tid + kJLanes, ... of R MCUs each; ist[k] is the destuffed
// byte position where interval k + 1 starts (nmk markers found).
__device__ void jseg_intervals(const JSeg& sg, const int* ist, int nmk, int nint, int R,
                               int tid) {
  const int nmcu = sg.total / sg.U;
  for (int i = tid; i < nint; i += kJLanes) {
    const int st = i == 0 ? 0 : (i - 1 < nmk ? ist[i - 1] : sg.n);
    LdsBits br{sg.s, sg.n, 0, 0, 0ull};
    br.pos = (st >> 2) << 2;
    br.refill();
    br.refill();
    br.skip((st & 3) * 8);
    int p0 = 0, p1 = 0, p2 = 0;
    const int m1 = min((i + 1) * R, nmcu);
    for (int m = i * R; m < m1; ++m)
      for (int u = 0; u < sg.U; ++u) {
        const int c = sg.ucomp[u];
        int16_t* blk = jseg_block(sg, m * sg.U + u, u);
        int sdc = huff_lds(br, sg.tabs[sg.dtab[c]]);
        if (sdc > 15) sdc = 0;
        const int diff = extend(br.get(sdc), sdc);
        int& pr = c == 0 ? p0 : (c == 1 ? p1 : p2);
        pr += diff;
        blk[0] = (int16_t)pr;
        const HuffTab& act = sg.tabs[sg.atab[c]];
        for (int z = 1; z < 64;) {
          const int rs = huff_lds(br, act);
          const int r = rs >> 4, sz = rs & 15;
          if (sz == 0) {
            if (r != 15) break;
            z += 16;
            continue;
          }
          z += r;
          blk[sg.zz[z < 63 ? z : 63]] = (int16_t)extend(br.get(sz), sz);
          ++z;
        }
      }
  }
}

// max_iv: entries of the restart-interval table after the staged segment in
// dynamic LDS (0: no tile of the launch has restart markers).
__global__ __launch_bounds__(kJLanes) void k_jpeg_entropy_lds(const JDev a, int cap, int max_iv) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ int scan_sh[kJWaves];
  __shared__ int64_t end_bit[kJLanes];
  __shared__ int end_u[kJLanes], end_z[kJLanes];
  const int t = blockIdx.x;
  const int tid = threadIdx.x;
  const JTile& jt = a.tiles[t];
  int16_t* coef = a.coef + jt.coef0 * 64;
  const int64_t n = jt.data_len;
  const int R = jt.restart;
  const int nint = R > 0 ? (jt.mcux * jt.mcuy + R - 1) / R : 0;
  if (nint > max_iv || n > cap || n < 256) {
    if (tid == 0) jpeg_entropy(jt, a.tabsets[jt.tabset], a.data, coef);
    return;
  }
  int* ist = reinterpret_cast<int*>(smem + kTabBytes + cap + 16);
  // tables: dc[0], dc[1], ac[0], ac[1] as 32-bit words
  HuffTab* tabs = reinterpret_cast<HuffTab*>(smem);
  {
    const JTables& tb = a.tabsets[jt.tabset];
    const uint32_t* src0 = reinterpret_cast<const uint32_t*>(&tb.dc[0]);
    const uint32_t* src1 = reinterpret_cast<const uint32_t*>(&tb.ac[0]);
    uint32_t* dst = reinterpret_cast<uint32_t*>(smem);
    constexpr int half = 2 * (int)sizeof(HuffTab) / 4;
    for (int i = tid; i < half; i += kJLanes) {
      dst[i] = src0[i];
      dst[half + i] = src1[i];
    }
  }
  // destuff the segment into LDS, 16 bytes per thread per round; with
  // restart markers, also drop every RST marker (0xFF 0xD0..0xD7) and record
  // where the next interval starts
  uint8_t* ds = smem + kTabBytes;
  const uint8_t* g = a.data + jt.data_off;
  int out = 0, nmk = 0;
  auto is_rst = [](uint8_t v) { return v >= 0xD0 && v <= 0xD7; };
  for (int64_t base = 0; base < n; base += kJLanes * 16) {
    const int64_t i0 = base + tid * 16;
    uint8_t b[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) b[k] = i0 + k < n ? g[i0 + k] : 0;
    const uint8_t before = (i0 > 0 && i0 - 1 < n) ? g[i0 - 1] : 0;
    const uint8_t after = (R > 0 && i0 + 16 < n) ? g[i0 + 16] : 0;
    uint32_t keep = 0, mk = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint8_t pv = k ? b[k - 1] : before;
      const uint8_t nx = k < 15 ? b[k + 1] : after;
      const bool in = i0 + k < n;
      const bool rst_ff = R > 0 && b[k] == 0xFF && is_rst(nx);
      const bool rst_n = R > 0 && pv == 0xFF && is_rst(b[k]);
      if (in && !(pv == 0xFF && b[k] == 0x00) && !rst_ff && !rst_n) keep |= 1u << k;
      if (in && rst_ff && i0 + k + 1 < n) mk |= 1u << k;
    }
    // one scan for both counts: kept bytes (< 2^16 per round) and markers
    int total;
    const int ex = block_excl_scan(__popc(keep) | (__popc(mk) << 16), scan_sh, &total);
    int pos = out + (ex & 0xFFFF);
    int mi = nmk + (ex >> 16);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (mk & (1u << k)) {
        if (mi < max_iv) ist[mi] = pos;  // the next kept byte starts an interval
        ++mi;
      }
      if (keep & (1u << k)) ds[pos++] = b[k];
    }
    out += total & 0xFFFF;
    nmk += total >> 16;
  }
  __shared__ uint8_t zz[64];
  __shared__ int8_t ucomp[12], ubx[12], uby[12];
  if (tid < 64) zz[tid] = (uint8_t)kZigzag[tid];
  if (tid < 8) ds[out + tid] = 0;  // the refill's second word past the end
  const int ncomp = jt.ncomp;
  int U = 0;
  for (int c = 0; c < ncomp; ++c) U += jt.comp[c].h * jt.comp[c].v;
  if (tid == 0) {  // units of an MCU: components in frame order, h x v blocks each
    int u = 0;
    for (int c = 0; c < ncomp; ++c)
      for (int by = 0; by < jt.comp[c].v; ++by)
        for (int bx = 0; bx < jt.comp[c].h; ++bx, ++u) {
          ucomp[u] = (int8_t)c;
          ubx[u] = (int8_t)bx;
          uby[u] = (int8_t)by;
        }
  }
  __syncthreads();
  JSeg seg;
  seg.s = ds;
  seg.n = out;
  seg.tabs = tabs;
  seg.zz = zz;
  seg.ucomp = ucomp;
  seg.ubx = ubx;
  seg.uby = uby;
  seg.U = U;
  seg.mcux = jt.mcux;
  seg.total = jt.mcux * jt.mcuy * U;
  seg.coef = coef;
  for (int c = 0; c < 3; ++c) {
    const int cc = c < ncomp ? c : 0;
    seg.dtab[c] = jt.comp[cc].td;
    seg.atab[c] = 2 + jt.comp[cc].ta;
    seg.ch[c] = jt.comp[cc].h;
    seg.cv[c] = jt.comp[cc].v;
    seg.cbw[c] = jt.comp[cc].bw;
    seg.cb0[c] = jt.comp[cc].block0;
  }
  if (R > 0) {
    jseg_intervals(seg, ist, min(nmk, max_iv), nint, R, tid);
    return;
  }
  // each thread owns S bits (a multiple of 32) of the destuffed segment
  const int64_t nbits = (int64_t)out * 8;
  const int64_t S = ((nbits + kJLanes - 1) / kJLanes + 31) / 32 * 32;
  const int64_t b0 = std::min<int64_t>((int64_t)tid * S, nbits);
  const int64_t b1 = std::min<int64_t>(b0 + S, nbits);
  // A: speculative decode of the own part from an assumed MCU start
  JState st{b0, 0, 0};
  int gb = 0;
  int nblk = 0;
  JState en = jseg_run<false>(seg, st, b1, gb, nblk);
  // B: take the left neighbour's end state as the own start until no end
  // state changes (thread 0's start is exact, so thread k is exact after k
  // rounds; decoding self-synchronises, so a few rounds suffice)
  for (int round = 0; round < kJLanes; ++round) {
    end_bit[tid] = en.bit;
    end_u[tid] = en.u;
    end_z[tid] = en.z;
    __syncthreads();
    JState ps{0, 0, 0};
    if (tid > 0) ps = JState{end_bit[tid - 1], end_u[tid - 1], end_z[tid - 1]};
    bool changed = false;
    if (ps.bit != st.bit || ps.u != st.u || ps.z != st.z) {
      st = ps;
      gb = 0;
      nblk = 0;
      const JState e2 = jseg_run<false>(seg, st, b1, gb, nblk);
      changed = e2.bit != en.bit || e2.u != en.u || e2.z != en.z;
      en = e2;
    }
    if (!__syncthreads_or(changed)) break;
  }
  // C: first block of each part = blocks completed by the parts before it
  int tot;
  gb = block_excl_scan(nblk, scan_sh, &tot);
  // D: decode again, storing coefficients (DC as differences)
  nblk = 0;
  jseg_run<true>(seg, st, b1, gb, nblk);
  __threadfence_block();
  __syncthreads();
  // E: DC prediction = running sum of the differences per component, in
  // decode order
  const int nmcu = jt.mcux * jt.mcuy;
  for (int c = 0; c < ncomp; ++c) {
    const int hv = seg.ch[c] * seg.cv[c];
    const int N = nmcu * hv;
    const int per = (N + kJLanes - 1) / kJLanes;
    const int k0 = min(tid * per, N), k1 = min(k0 + per, N);
    int sum = 0;
    for (int k = k0; k < k1; ++k) sum += *jseg_dc(seg, c, hv, k);
    int run = block_excl_scan(sum, scan_sh, &tot);
    for (int k = k0; k < k1; ++k) {
      int16_t* p = jseg_dc(seg, c, hv, k);
      run += *p;
      *p = (int16_t)run;
    }
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    p[i] = v;
}

// One work-group per 8x8 block of a tile: the inverse transform into the
// tile's component plane.
__global__ __launch_bounds__(64) void k_jpeg_idct(const JDev a, int64_t nblk) {
  for (int64_t g = blockIdx.x * 64 + threadIdx.x; g < nblk; g += (int64_t)gridDim.x * 64) {
    // the tile holding block g: the last one whose first block is <= g
    int lo = 0, hi = a.ntiles - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.tiles[mid].coef0 <= g) lo = mid;
      else hi = mid - 1;
    }
    const JTile& jt = a.tiles[lo];
    const int64_t b = g - jt.coef0;  // block within the tile
    int c = 0;
    while (c + 1 < jt.ncomp && b >= jt.comp[c + 1].block0) ++c;
    const JComp& k = jt.comp[c];
    const int64_t lb = b - k.block0;
    const int by = (int)(lb / k.bw), bx = (int)(lb - (int64_t)by * k.bw);
    uint8_t* plane = a.planes + (jt.coef0 + k.block0) * 64;
    const int pw = k.bw * 8;
    idct_islow(a.coef + g * 64, a.tabsets[jt.tabset].qtab[k.tq],
               plane + (int64_t)by * 8 * pw + bx * 8, pw);
  }
}

// One thread per region pixel covered by a JPEG tile.
__global__ __launch_bounds__(256) void k_jpeg_color(const JDev a) {
  const int t = blockIdx.y;
  const JTile& jt = a.tiles[t];
  // the tile's overlap with the region (tile and region offsets fit 32 bits
  // relative to each other: a tile spans at most 65535 pixels)
  const int64_t ox0 = std::max<int64_t>(jt.x0, a.rx), oy0 = std::max<int64_t>(jt.y0, a.ry);
  const int64_t ox1 = std::min<int64_t>(jt.x0 + jt.vis_w, a.rx + a.rw);
  const int64_t oy1 = std::min<int64_t>(jt.y0 + jt.vis_h, a.ry + a.rh);
  if (ox1 <= ox0 || oy1 <= oy0) return;
  const uint32_t ow = (uint32_t)(ox1 - ox0), n = ow * (uint32_t)(oy1 - oy0);
  uint8_t* planes[3];
  for (int c = 0; c < 3; ++c)
    planes[c] = a.planes + (jt.coef0 + jt.comp[c < jt.ncomp ? c : 0].block0) * 64;
  // everything the loop reads of the tile, in registers (the byte stores
  // below could otherwise alias the tile table and force reloads)
  const PixCtx pc = pix_ctx(jt, planes);
  const int tx0 = (int)(ox0 - jt.x0), ty0 = (int)(oy0 - jt.y0);
  const int nch = a.nch;
  uint8_t* __restrict__ out = a.out + ((oy0 - a.ry) * a.rw + (ox0 - a.rx)) * nch;
  const int64_t orow = a.rw * nch;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t yy = i / ow, xx = i - yy * ow;
    jpeg_pixel(pc, tx0 + (int)xx, ty0 + (int)yy, out + yy * orow + xx * nch, nch);
  }
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    p[i] = v;
}

// ------------------------------------------------------ region read --
struct TileRef {
  int64_t tx, ty;  // tile column / row in the level
};

void tiles_of_region(const Ifd& f, int64_t x, int64_t y, int64_t w, int64_t h,
                     std::vector<TileRef>& out) {
  const int64_t x0 = std::max<int64_t>(x, 0), y0 = std::max<int64_t>(y, 0);
  const int64_t x1 = std::min<int64_t>(x + w, f.width), y1 = std::min<int64_t>(y + h, f.height);
  if (x1 <= x0 || y1 <= y0) return;
  for (int64_t ty = y0 / f.tile_h; ty <= (y1 - 1) / f.tile_h; ++ty)
    for (int64_t tx = x0 / f.tile_w; tx <= (x1 - 1) / f.tile_w; ++tx) out.push_back({tx, ty});
}

bool read_tile_bytes(const Wsi& w, const Ifd& f, const TileRef& t, std::vector<uint8_t>& buf) {
  const int64_t i = t.ty * f.tiles_x() + t.tx;
  const uint64_t off = f.offsets[i], n = f.counts[i];
  if (n > (1ull << 31)) return false;
  buf.resize(n);
  return n == 0 || read_at(w.fd, off, buf.data(), n);
}

// Copy the part of a decoded tile (tw x th pixels of nch samples, rows of
// tile_w pixels) that lies in the region.
void blit(const Ifd& f, const TileRef& t, const uint8_t* px, int nch, uint8_t* out, int64_t rx,
          int64_t ry, int64_t rw, int64_t rh) {
  const int64_t gx0 = t.tx * f.tile_w, gy0 = t.ty * f.tile_h;
  const int64_t ox0 = std::max(gx0, rx), oy0 = std::max(gy0, ry);
  const int64_t ox1 = std::min({gx0 + f.tile_w, rx + rw, f.width});
  const int64_t oy1 = std::min({gy0 + f.tile_h, ry + rh, f.height});
  for (int64_t gy = oy0; gy < oy1; ++gy)
    std::memcpy(out + ((gy - ry) * rw + (ox0 - rx)) * nch,
                px + ((gy - gy0) * f.tile_w + (ox0 - gx0)) * nch, (size_t)(ox1 - ox0) * nch);
}

// Decode one lossless tile into px (tile_w x tile_h x spp).
int decode_lossless(const Ifd& f, const std::vector<uint8_t>& src, std::vector<uint8_t>& px) {
  const size_t need = (size_t)(f.tile_w * f.tile_h * f.spp);
  px.assign(need, 0);
  if (f.compression == 1) {
    std::memcpy(px.data(), src.data(), std::min(need, src.size()));
  } else if (f.compression == 8 || f.compression == 32946) {
    uLongf outlen = (uLongf)need;
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return fail(HC_ERR_VALUE, "zlib init");
    zs.next_in = const_cast<Bytef*>(src.data());
    zs.avail_in = (uInt)src.size();
    zs.next_out = px.data();
    zs.avail_out = (uInt)outlen;
    const int r = inflate(&zs, Z_FINISH);
    inflateEnd(&zs);
    if (r != Z_STREAM_END && r != Z_OK && r != Z_BUF_ERROR)
      return fail(HC_ERR_VALUE, "corrupt deflate tile");
  } else if (f.compression == 5) {
    if (!lzw_decode(src.data(), src.size(), px.data(), need))
      return fail(HC_ERR_VALUE, "corrupt LZW tile");
  }
  if (f.predictor == 2) undo_predictor(px.data(), f.tile_w, f.tile_h, f.spp);
  return HC_OK;
}

// The colour space of three-component JPEG tiles follows the TIFF
// photometric tag, as the reference's decoder is told (ifd.cpp:90-106):
// RGB stays RGB, YCbCr is converted; otherwise the stream's own markers
// decide.
void set_color(JTile& jt, const Ifd& f) {
  if (jt.ncomp != 3) return;
  if (f.photometric == 2) jt.transform = 0;
  else if (f.photometric == 6) jt.transform = 1;
}

// Decode one JPEG tile on the host into tile pixels (tile_w x tile_h x nch).
int decode_jpeg_host(const Ifd& f, const std::vector<uint8_t>& src, int nch,
                     std::vector<uint8_t>& px) {
  JTile jt;
  JTables tb;
  int rc = parse_jpeg(f.jpeg_tables.data(), f.jpeg_tables.size(), src.data(), src.size(), jt, tb);
  if (rc != HC_OK) return rc;
  set_color(jt, f);
  if (jt.ncomp > nch) return fail(HC_ERR_VALUE, "JPEG tile has more components than the TIFF samples");
  std::vector<int16_t> coef((size_t)jt.nblocks * 64);
  jpeg_entropy(jt, tb, src.data(), coef.data());
  std::vector<uint8_t> planes((size_t)jt.nblocks * 64);
  for (int c = 0; c < jt.ncomp; ++c) {
    const JComp& k = jt.comp[c];
    uint8_t* plane = planes.data() + k.block0 * 64;
    for (int64_t b = 0; b < (int64_t)k.bw * k.bh; ++b) {
      const int by = (int)(b / k.bw), bx = (int)(b % k.bw);
      idct_islow(coef.data() + (k.block0 + b) * 64, tb.qtab[k.tq],
                 plane + (int64_t)by * 8 * k.bw * 8 + bx * 8, k.bw * 8);
    }
  }
  uint8_t* pl[3] = {nullptr, nullptr, nullptr};
  for (int c = 0; c < jt.ncomp; ++c) pl[c] = planes.data() + jt.comp[c].block0 * 64;
  const PixCtx pc = pix_ctx(jt, pl);
  px.assign((size_t)(f.tile_w * f.tile_h * nch), 0);
  const int64_t w = std::min<int64_t>(jt.width, f.tile_w), h = std::min<int64_t>(jt.height, f.tile_h);
  for (int64_t y = 0; y < h; ++y)
    for (int64_t x = 0; x < w; ++x)
      jpeg_pixel(pc, (int)x, (int)y, px.data() + (y * f.tile_w + x) * nch, nch);
  return HC_OK;
}

int read_region_host(const Wsi& w, const Ifd& f, int64_t x, int64_t y, int64_t rw, int64_t rh,
                     uint8_t* out, int workers) {
  int rc = check_level_codec(f);
  if (rc != HC_OK) return rc;
  const int nch = out_channels(f);
  std::memset(out, w.background, (size_t)(rw * rh * nch));
  std::vector<TileRef> tiles;
  tiles_of_region(f, x, y, rw, rh, tiles);
  std::atomic<size_t> next{0};
  std::atomic<int> err{HC_OK};
  std::mutex emu;
  std::string emsg;
  auto work = [&]() {
    std::vector<uint8_t> src, px;
    for (size_t i = next++; i < tiles.size(); i = next++) {
      if (err.load() != HC_OK) return;
      int r = read_tile_bytes(w, f, tiles[i], src) ? HC_OK : fail(HC_ERR_VALUE, "tile read failed");
      if (r == HC_OK)
        r = f.compression == 7 ? decode_jpeg_host(f, src, nch, px) : decode_lossless(f, src, px);
      if (r != HC_OK) {
        std::lock_guard<std::mutex> lk(emu);
        if (err.load() == HC_OK) {
          emsg = hc_last_error();
          err.store(r);
        }
        return;
      }
      blit(f, tiles[i], px.data(), nch, out, x, y, rw, rh);
    }
  };
  const int nt = std::max(1, std::min<int>(workers, (int)tiles.size()));
  if (nt <= 1) {
    work();
  } else {
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(work);
    for (auto& t : th) t.join();
  }
  if (err.load() != HC_OK) return fail(err.load(), emsg);
  return HC_OK;
}

}  // namespace
}  // namespace hc

using namespace hc;

struct hc_wsi {
  Wsi w;
  // staging (pinned host) and work (device) buffers of device reads, kept
  // between calls and grown as needed; device reads on one handle serialise
  mutable std::mutex mu;
  mutable uint8_t* pinned = nullptr;
  mutable size_t pinned_cap = 0;
  mutable uint8_t* dev = nullptr;
  mutable size_t dev_cap = 0;
  mutable int dev_id = -1;
  ~hc_wsi() {
    if (pinned) (void)hipHostFree(pinned);
    if (dev) (void)hipFree(dev);
  }
};

extern "C" {

int hc_wsi_open(const char* path, hc_wsi** out) {
  if (!path || !out) return fail(HC_ERR_VALUE, "null argument");
  *out = nullptr;
  std::unique_ptr<hc_wsi> h(new hc_wsi);
  h->w.fd = open(path, O_RDONLY);
  if (h->w.fd < 0) return fail(HC_ERR_VALUE, std::string("cannot open ") + path);
  const int rc = parse_tiff(h->w);
  if (rc != HC_OK) return rc;
  *out = h.release();
  return HC_OK;
}

void hc_wsi_close(hc_wsi* h) { delete h; }

int hc_wsi_level_count(const hc_wsi* h) { return h ? (int)h->w.levels.size() : -1; }

int hc_wsi_level_info(const hc_wsi* h, int level, int64_t* info) {
  if (!h || !info) return fail(HC_ERR_VALUE, "null argument");
  if (level < 0 || level >= (int)h->w.levels.size()) return fail(HC_ERR_VALUE, "level out of range");
  const Ifd& f = h->w.ifds[h->w.levels[level]];
  info[0] = f.width;
  info[1] = f.height;
  info[2] = f.tile_w;
  info[3] = f.tile_h;
  info[4] = out_channels(f);
  info[5] = f.compression;
  info[6] = f.photometric;
  info[7] = f.tiled ? 1 : 0;
  return HC_OK;
}

int hc_wsi_properties(const hc_wsi* h, double* res, int* background, char* desc, int desc_len,
                      char* vendor, int vendor_len) {
  if (!h) return fail(HC_ERR_VALUE, "null handle");
  const Ifd& f = h->w.ifds[0];
  if (res) {
    res[0] = f.xres;
    res[1] = f.yres;
    res[2] = f.res_unit;
  }
  if (background) *background = h->w.background;
  if (desc && desc_len > 0) {
    const size_t n = std::min<size_t>((size_t)desc_len - 1, f.description.size());
    std::memcpy(desc, f.description.data(), n);
    desc[n] = 0;
  }
  if (vendor && vendor_len > 0) {
    const size_t n = std::min<size_t>((size_t)vendor_len - 1, h->w.vendor.size());
    std::memcpy(vendor, h->w.vendor.data(), n);
    vendor[n] = 0;
  }
  return HC_OK;
}

int hc_wsi_read_region(const hc_wsi* h, int level, int64_t x, int64_t y, int64_t w, int64_t hgt,
                       uint8_t* out, int num_workers) {
  if (!h || !out) return fail(HC_ERR_VALUE, "null argument");
  if (level < 0 || level >= (int)h->w.levels.size()) return fail(HC_ERR_VALUE, "level out of range");
  if (w <= 0 || hgt <= 0) return fail(HC_ERR_VALUE, "region size must be positive");
  return read_region_host(h->w, h->w.ifds[h->w.levels[level]], x, y, w, hgt, out, num_workers);
}

int hc_wsi_associated_count(const hc_wsi* h) { return h ? (int)h->w.assoc.size() : -1; }

int hc_wsi_associated_info(const hc_wsi* h, int k, char* name, int name_len, int64_t* info) {
  if (!h || !info) return fail(HC_ERR_VALUE, "null argument");
  if (k < 0 || k >= (int)h->w.assoc.size()) return fail(HC_ERR_VALUE, "associated image out of range");
  const auto& a = h->w.assoc[k];
  const Ifd& f = h->w.ifds[a.second];
  if (name && name_len > 0) {
    const size_t n = std::min<size_t>((size_t)name_len - 1, a.first.size());
    std::memcpy(name, a.first.data(), n);
    name[n] = 0;
  }
  info[0] = f.width;
  info[1] = f.height;
  info[2] = out_channels(f);
  info[3] = f.compression;
  return HC_OK;
}

int hc_wsi_associated_read(const hc_wsi* h, int k, uint8_t* out, int num_workers) {
  if (!h || !out) return fail(HC_ERR_VALUE, "null argument");
  if (k < 0 || k >= (int)h->w.assoc.size()) return fail(HC_ERR_VALUE, "associated image out of range");
  const Ifd& f = h->w.ifds[h->w.assoc[k].second];
  return read_region_host(h->w, f, 0, 0, f.width, f.height, out, num_workers);
}

int hc_wsi_read_region_device(const hc_wsi* h, int level, int64_t x, int64_t y, int64_t w,
                              int64_t hgt, uint8_t* out_dev, int num_workers, void* stream) {
  if (!h || !out_dev) return fail(HC_ERR_VALUE, "null argument");
  if (level < 0 || level >= (int)h->w.levels.size()) return fail(HC_ERR_VALUE, "level out of range");
  if (w <= 0 || hgt <= 0) return fail(HC_ERR_VALUE, "region size must be positive");
  const Wsi& ws = h->w;
  const Ifd& f = ws.ifds[ws.levels[level]];
  int rc = check_level_codec(f);
  if (rc != HC_OK) return rc;
  const int nch = out_channels(f);
  hipStream_t st = (hipStream_t)stream;
  const int64_t nbytes = w * hgt * nch;
  if (f.compression != 7) {
    // lossless tiles: decoded on host threads into the handle's pinned
    // staging buffer, one upload of the region
    std::lock_guard<std::mutex> lock(h->mu);
    if (h->pinned_cap < (size_t)nbytes) {
      if (h->pinned) (void)hipHostFree(h->pinned);
      h->pinned = nullptr;
      h->pinned_cap = 0;
      HC_HIP(hipHostMalloc((void**)&h->pinned, (size_t)nbytes));
      h->pinned_cap = (size_t)nbytes;
    }
    rc = read_region_host(ws, f, x, y, w, hgt, h->pinned, num_workers);
    if (rc == HC_OK) {
      hipError_t e = hipMemcpyAsync(out_dev, h->pinned, (size_t)nbytes, hipMemcpyHostToDevice, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) rc = hip_check(e, "region upload");
    }
    return rc;
  }
  // JPEG tiles: byte ranges read and headers parsed on host threads, entropy
  // decoding and transforms on the GPU
  std::vector<TileRef> tiles;
  tiles_of_region(f, x, y, w, hgt, tiles);
  if (tiles.size() > 65535) return fail(HC_ERR_RESOURCE, "region spans more than 65535 JPEG tiles");
  const int nt = (int)tiles.size();
  std::lock_guard<std::mutex> lock(h->mu);
  // the level's shared tables, parsed once
  JTables base;
  std::memset(&base, 0, sizeof(base));
  JParser bp;
  bp.tb = &base;
  {
    JTile dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    if (!f.jpeg_tables.empty()) {
      rc = bp.segs(f.jpeg_tables.data(), f.jpeg_tables.size(), false, dummy);
      if (rc != HC_OK) return rc;
    }
  }
  // staging layout: tiles | table sets (at most one per tile + the shared
  // one) | compressed tiles back to back
  auto al = [](int64_t v) { return (v + 255) / 256 * 256; };
  std::vector<int64_t> boff(nt + 1, 0), foff(nt);
  for (int i = 0; i < nt; ++i) {
    const int64_t idx = tiles[i].ty * f.tiles_x() + tiles[i].tx;
    if (f.counts[idx] > (1ull << 31)) return fail(HC_ERR_VALUE, "implausible tile byte count");
    foff[i] = (int64_t)f.offsets[idx];
    boff[i + 1] = boff[i] + (int64_t)f.counts[idx];
  }
  const int64_t o_tabs = al((int64_t)sizeof(JTile) * std::max(nt, 1));
  const int64_t o_blob = al(o_tabs + (int64_t)sizeof(JTables) * (nt + 1));
  const size_t staging = (size_t)(o_blob + boff[nt] + 16);
  if (h->pinned_cap < staging) {
    if (h->pinned) (void)hipHostFree(h->pinned);
    h->pinned = nullptr;
    h->pinned_cap = 0;
    const size_t cap = staging + staging / 4;
    HC_HIP(hipHostMalloc((void**)&h->pinned, cap));
    h->pinned_cap = cap;
  }
  uint8_t* hp = h->pinned;
  JTile* T = reinterpret_cast<JTile*>(hp);
  JTables* TB = reinterpret_cast<JTables*>(hp + o_tabs);
  uint8_t* blob = hp + o_blob;
  TB[0] = base;
  // 1) reads: runs of tiles that lie back to back in the file, one pread each
  std::vector<std::pair<int, int>> runs;  // [first, last] tile of a run
  for (int i = 0; i < nt; ++i) {
    if (!runs.empty() && foff[i] == foff[runs.back().second] + (boff[runs.back().second + 1] -
                                                                 boff[runs.back().second]))
      runs.back().second = i;
    else
      runs.push_back({i, i});
  }
  const int workers = std::max(1, num_workers);
  std::atomic<int> err{HC_OK};
  std::mutex emu;
  std::string emsg;
  auto set_err = [&](int r) {
    std::lock_guard<std::mutex> lk(emu);
    if (err.load() == HC_OK) {
      emsg = hc_last_error();
      err.store(r);
    }
  };
  auto parallel = [&](int n, const std::function<void(int)>& fn) {
    std::atomic<int> next{0};
    auto body = [&]() {
      for (int i = next++; i < n; i = next++) {
        if (err.load() != HC_OK) return;
        fn(i);
      }
    };
    const int nth = std::min(workers, n);
    if (nth <= 1) {
      body();
      return;
    }
    std::vector<std::thread> th;
    for (int k = 0; k < nth; ++k) th.emplace_back(body);
    for (auto& t : th) t.join();
  };
  parallel((int)runs.size(), [&](int r) {
    const int a0 = runs[r].first, a1 = runs[r].second;
    if (!read_at(ws.fd, (uint64_t)foff[a0], blob + boff[a0], (size_t)(boff[a1 + 1] - boff[a0])))
      set_err(fail(HC_ERR_VALUE, "tile read failed"));
  });
  // 2) headers: each tile's stream on top of the shared tables
  std::atomic<int> nsets{1};
  parallel(nt, [&](int i) {
    JTile& j = T[i];
    std::memset(&j, 0, sizeof(j));
    JParser ps = bp;
    JTables unused;
    ps.tb = &unused;
    const uint8_t* src = blob + boff[i];
    const size_t n = (size_t)(boff[i + 1] - boff[i]);
    if (stream_has_tables(src, n)) {
      const int k = nsets++;
      TB[k] = base;
      ps.tb = &TB[k];
      j.tabset = k;
    }
    int r = ps.segs(src, n, true, j);
    if (r == HC_OK) r = ps.finish(j);
    if (r == HC_OK && j.ncomp > nch)
      r = fail(HC_ERR_VALUE, "JPEG tile has more components than the TIFF samples");
    if (r != HC_OK) {
      set_err(r);
      return;
    }
    set_color(j, f);
    j.data_off += boff[i];
    j.x0 = (int)(tiles[i].tx * f.tile_w);
    j.y0 = (int)(tiles[i].ty * f.tile_h);
    j.vis_w = (int)std::min<int64_t>(j.width, std::min(f.tile_w, f.width - j.x0));
    j.vis_h = (int)std::min<int64_t>(j.height, std::min(f.tile_h, f.height - j.y0));
  });
  if (err.load() != HC_OK) return fail(err.load(), emsg);
  int64_t nblk = 0;
  for (int i = 0; i < nt; ++i) {
    T[i].coef0 = nblk;
    nblk += T[i].nblocks;
  }
  // LDS staging capacity of the entropy kernel: the largest segment of a
  // tile without restart markers, rounded up, at most kMaxStage bytes (larger
  // ones are decoded from global memory)
  constexpr int64_t kMaxStage = 120 * 1024;
  int64_t need = 0;
  int max_iv = 0;  // restart intervals of the largest such tile (table in LDS)
  constexpr int kMaxIntervals = 4096;  // with a 120 KiB stage: < 160 KiB of LDS
  for (int i = 0; i < nt; ++i) {
    if (T[i].data_len > kMaxStage) continue;
    need = std::max(need, T[i].data_len);
    if (T[i].restart > 0) {
      const int nint = (T[i].mcux * T[i].mcuy + T[i].restart - 1) / T[i].restart;
      if (nint <= kMaxIntervals) max_iv = std::max(max_iv, nint);
    }
  }
  const int cap = (int)((need + 16 * kJLanes - 1) / (16 * kJLanes) * (16 * kJLanes));
  // device work buffer: the staged bytes at the same offsets, then the
  // coefficients and the component planes
  const int64_t o_coef = al(o_blob + boff[nt] + 16);
  const int64_t o_pl = al(o_coef + 128 * std::max<int64_t>(nblk, 1));
  const size_t total = (size_t)al(o_pl + 64 * std::max<int64_t>(nblk, 1));
  int dev_id = 0;
  HC_HIP(hipGetDevice(&dev_id));
  if (h->dev_cap < total || h->dev_id != dev_id) {
    if (h->dev) (void)hipFree(h->dev);
    h->dev = nullptr;
    h->dev_cap = 0;
    const size_t capb = total + total / 4;
    HC_HIP(hipMalloc((void**)&h->dev, capb));
    h->dev_cap = capb;
    h->dev_id = dev_id;
  }
  uint8_t* dev = h->dev;
  hipError_t e = hipSuccess;
  if (nt > 0) {
    e = hipMemcpyAsync(dev, hp, (size_t)(o_tabs + (int64_t)sizeof(JTables) * nsets.load()),
                       hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dev + o_blob, blob, (size_t)boff[nt], hipMemcpyHostToDevice, st);
  }
  if (e != hipSuccess) return hip_check(e, "JPEG tile upload");
  JDev a;
  a.tiles = reinterpret_cast<const JTile*>(dev);
  a.tabsets = reinterpret_cast<const JTables*>(dev + o_tabs);
  a.data = dev + o_blob;
  a.coef = reinterpret_cast<int16_t*>(dev + o_coef);
  a.planes = dev + o_pl;
  a.ntiles = nt;
  a.out = out_dev;
  a.rx = x;
  a.ry = y;
  a.rw = w;
  a.rh = hgt;
  a.nch = nch;
  rc = [&]() -> int {
    HC_LAUNCH("k_fill_u8", st, k_fill_u8, dim3((unsigned)std::min<int64_t>(4096, (nbytes + 255) / 256)),
              dim3(256), 0, out_dev, nbytes, ws.background);
    if (nt == 0) return HC_OK;
    // coefficients start at zero; the decoder stores the non-zero ones
    HC_LAUNCH("k_fill_u32", st, k_fill_u32,
              dim3((unsigned)std::min<int64_t>(4096, (nblk * 32 + 255) / 256)), dim3(256), 0,
              reinterpret_cast<uint32_t*>(a.coef), nblk * 32, 0u);
    const size_t smem = (size_t)(kTabBytes + cap + 16) + 4 * (size_t)max_iv;
    if (smem > 64 * 1024) {
      const hipError_t fe = hipFuncSetAttribute((const void*)k_jpeg_entropy_lds,
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)smem);
      if (fe != hipSuccess) return hip_check(fe, "hipFuncSetAttribute(k_jpeg_entropy_lds)");
    }
    HC_LAUNCH("k_jpeg_entropy_lds", st, k_jpeg_entropy_lds, dim3((unsigned)nt), dim3(kJLanes),
              smem, a, cap, max_iv);
    HC_LAUNCH("k_jpeg_idct", st, k_jpeg_idct,
              dim3((unsigned)std::min<int64_t>(65535, (nblk + 63) / 64)), dim3(64), 0, a, nblk);
    HC_LAUNCH("k_jpeg_color", st, k_jpeg_color, dim3(64, (unsigned)nt), dim3(256), 0, a);
    return HC_OK;
  }();
  // the staging and work buffers are reused by the next call: wait here
  if (rc == HC_OK) {
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_check(e, "JPEG decode");
  } else {
    (void)hipStreamSynchronize(st);
  }
  return rc;
}

}  // extern "C"
