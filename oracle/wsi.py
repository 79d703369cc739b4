"""Oracle: Python restatement of the reference's whole-slide region read.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference files (under ``cpp/``):
  src/cuimage.cpp:649                        CuImage::read_region (argument checks)
  plugins/cucim.kit.cuslide/src/cuslide/cuslide.cpp:227-234
                                             level_downsamples = mean of the
                                             width and height ratios (float32)
  plugins/cucim.kit.cuslide/src/cuslide/tiff/tiff.cpp:342-385, 670-707
                                             IFDs -> pyramid levels, vendor,
                                             background value, size checks and
                                             the level-0 -> level location
                                             division (:733-769)
  plugins/cucim.kit.cuslide/src/cuslide/tiff/ifd.cpp:576-1260
                                             region = the tiles it overlaps,
                                             pixels outside the level filled
                                             with the background value

Tile decoding uses the libraries the reference links, not a restatement:
baseline JPEG through Pillow's libjpeg-turbo (the reference decodes with
libjpeg-turbo's TurboJPEG API and its default flags: integer "islow" IDCT,
fancy upsampling, ``cuslide/jpeg/libjpeg_turbo.cpp:116-121``), deflate
through zlib.  TIFF LZW is restated here (:func:`lzw_decode`, the variable
9..12-bit, MSB-first, early-change code of the TIFF 6.0 specification,
section 13); it is pinned by files that libtiff wrote.
"""

import io
import struct
import zlib

import numpy as np

__all__ = ["TiffFile", "lzw_decode", "level_location"]

_TYPE_SIZE = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8,
              13: 4, 16: 8, 17: 8, 18: 8}
_TYPE_FMT = {1: "B", 2: "B", 3: "H", 4: "I", 6: "b", 7: "B", 8: "h", 9: "i", 11: "f", 12: "d",
             13: "I", 16: "Q", 17: "q", 18: "Q"}


def lzw_decode(src, nbytes):
    """TIFF LZW (compression 5): codes of 9..12 bits, most significant bit
    first; 256 clears the table, 257 ends the strip; the code width grows
    when the next free entry is one below the width's limit ("early
    change")."""
    out = bytearray()
    table = [bytes([i]) for i in range(256)] + [b"", b""]
    width, prev = 9, None
    bitbuf, nbits, pos = 0, 0, 0
    while len(out) < nbytes:
        while nbits < width and pos < len(src):
            bitbuf = (bitbuf << 8) | src[pos]
            pos += 1
            nbits += 8
        if nbits < width:
            break
        code = (bitbuf >> (nbits - width)) & ((1 << width) - 1)
        nbits -= width
        bitbuf &= (1 << nbits) - 1
        if code == 257:
            break
        if code == 256:
            table = table[:258]
            width, prev = 9, None
            continue
        if code < len(table):
            entry = table[code]
            if prev is not None:
                table.append(prev + entry[:1])
        elif code == len(table) and prev is not None:
            entry = prev + prev[:1]
            table.append(entry)
        else:
            raise ValueError("corrupt LZW stream")
        out += entry
        prev = entry
        if len(table) + 1 >= (1 << width) and width < 12:
            width += 1
    return bytes(out[:nbytes])


def level_location(loc, downsample):
    """Level-0 location -> level location: ``int64 /= float`` in C++, i.e.
    float32 division of the float32-converted value, truncated toward zero
    (tiff.cpp:762-768)."""
    q = np.float32(np.float32(loc) / np.float32(downsample))
    return int(q)


class _Ifd:
    pass


class TiffFile:
    """Structure of a (Big)TIFF file and region reads from it."""

    def __init__(self, path):
        with open(path, "rb") as fh:
            self.data = fh.read()
        d = self.data
        if d[:2] == b"II":
            self.bo = "<"
        elif d[:2] == b"MM":
            self.bo = ">"
        else:
            raise ValueError("not a TIFF file")
        magic = struct.unpack(self.bo + "H", d[2:4])[0]
        self.big = magic == 43
        nxt = (struct.unpack(self.bo + "Q", d[8:16])[0] if self.big
               else struct.unpack(self.bo + "I", d[4:8])[0])
        self.ifds = []
        while nxt:
            nxt = self._read_ifd(nxt)
        desc = self.ifds[0].tags.get(270, b"")
        self.vendor, self.background = "generic", 0
        if desc.startswith(b"Aperio"):
            self.vendor, self.background = "aperio", 255
        elif b"DPUfsImport" in desc:
            self.vendor, self.background = "philips", 255
        lv = [i for i, f in enumerate(self.ifds)
              if i == 0 or self.vendor == "generic" or f.tiled]
        lv.sort(key=lambda i: (-self.ifds[i].width, -self.ifds[i].height))
        self.levels = lv
        # associated images of Aperio files (tiff.cpp:670-700)
        self.associated = {}
        if self.vendor == "aperio":
            for i, f in enumerate(self.ifds):
                if i == 0 or f.tiled:
                    continue
                sub = f.tags.get(254, (0,))[0]
                name = ("thumbnail" if i == 1 and sub == 0 else
                        "label" if sub == 1 else "macro" if sub == 9 else None)
                if name:
                    self.associated[name] = i
        f0 = self.ifds[lv[0]]
        self.downsamples = []
        for i in lv:
            f = self.ifds[i]
            ds = (np.float32(f0.width) / np.float32(f.width)
                  + np.float32(f0.height) / np.float32(f.height)) / np.float32(2)
            self.downsamples.append(np.float32(ds))

    def _values(self, typ, count, raw_off_bytes):
        size = _TYPE_SIZE.get(typ)
        if size is None:
            return None
        n = size * count
        inline = 8 if self.big else 4
        if n <= inline:
            buf = raw_off_bytes[:n]
        else:
            off = struct.unpack(self.bo + ("Q" if self.big else "I"), raw_off_bytes[:inline])[0]
            buf = self.data[off:off + n]
        if typ in (2, 7):
            return bytes(buf)
        if typ in (5, 10):
            v = struct.unpack(self.bo + ("I" if typ == 5 else "i") * (2 * count), buf)
            return tuple(v)
        return struct.unpack(self.bo + _TYPE_FMT[typ] * count, buf)

    def _read_ifd(self, off):
        d, bo = self.data, self.bo
        if self.big:
            n = struct.unpack(bo + "Q", d[off:off + 8])[0]
            p, esz = off + 8, 20
        else:
            n = struct.unpack(bo + "H", d[off:off + 2])[0]
            p, esz = off + 2, 12
        tags = {}
        for e in range(n):
            ent = d[p + e * esz:p + (e + 1) * esz]
            tag, typ = struct.unpack(bo + "HH", ent[:4])
            if self.big:
                count = struct.unpack(bo + "Q", ent[4:12])[0]
                val = ent[12:20]
            else:
                count = struct.unpack(bo + "I", ent[4:8])[0]
                val = ent[8:12]
            v = self._values(typ, count, val)
            if v is not None:
                tags[tag] = v
        q = p + n * esz
        nxt = (struct.unpack(bo + "Q", d[q:q + 8])[0] if self.big
               else struct.unpack(bo + "I", d[q:q + 4])[0])
        f = _Ifd()
        f.tags = tags
        if 270 in tags:
            tags[270] = tags[270].rstrip(b"\0")
        f.width, f.height = tags[256][0], tags[257][0]
        f.spp = tags.get(277, (1,))[0]
        f.compression = tags.get(259, (1,))[0]
        f.photometric = tags.get(262, (2,))[0]
        f.predictor = tags.get(317, (1,))[0]
        f.tiled = 322 in tags
        if f.tiled:
            f.tw, f.th = tags[322][0], tags[323][0]
            f.offsets, f.counts = tags[324], tags[325]
        else:
            f.tw = f.width
            f.th = min(tags.get(278, (f.height,))[0], f.height)
            f.offsets, f.counts = tags[273], tags[279]
        f.tables = tags.get(347, b"")
        self.ifds.append(f)
        return nxt

    # ---------------------------------------------------------- decoding --
    def out_channels(self, f):
        if f.compression == 7:
            return 3 if f.spp >= 3 else 1
        return f.spp

    def decode_tile(self, f, index):
        """Pixels of one tile / strip as (th, tw, channels) uint8."""
        raw = self.data[f.offsets[index]:f.offsets[index] + f.counts[index]]
        nch = self.out_channels(f)
        if f.compression == 7:
            from PIL import Image

            s = raw
            if f.tables:
                # tables stream (SOI ... EOI) + tile stream without its SOI
                s = f.tables[:-2] + raw[2:]
            if nch == 3 and f.photometric in (2, 6):
                # the reference overrides the stream's colour space with the
                # TIFF photometric tag (ifd.cpp:98-106, libjpeg_turbo.cpp:309-313)
                s = _force_transform(s, 0 if f.photometric == 2 else 1)
            im = Image.open(io.BytesIO(s))
            a = np.asarray(im.convert("RGB") if nch == 3 else im.convert("L"))
            a = a.reshape(a.shape[0], a.shape[1], nch)
            out = np.zeros((f.th, f.tw, nch), np.uint8)
            h, w = min(a.shape[0], f.th), min(a.shape[1], f.tw)
            out[:h, :w] = a[:h, :w]
            return out
        need = f.tw * f.th * f.spp
        if f.compression == 1:
            buf = bytes(raw[:need])
        elif f.compression in (8, 32946):
            buf = zlib.decompressobj().decompress(bytes(raw), need)
        elif f.compression == 5:
            buf = lzw_decode(raw, need)
        else:
            raise NotImplementedError(f"compression {f.compression}")
        buf = buf + bytes(max(0, need - len(buf)))
        a = np.frombuffer(buf, np.uint8).reshape(f.th, f.tw, f.spp).copy()
        if f.predictor == 2:
            a = np.cumsum(a.astype(np.uint32), axis=1).astype(np.uint8)
        return a

    def level_image(self, level):
        """The whole level as (height, width, channels) uint8."""
        return self.ifd_image(self.levels[level])

    def ifd_image(self, index):
        """The whole image of IFD ``index`` as (height, width, channels) uint8."""
        f = self.ifds[index]
        nch = self.out_channels(f)
        img = np.zeros((f.height, f.width, nch), np.uint8)
        ntx = -(-f.width // f.tw)
        for i in range(len(f.offsets)):
            ty, tx = divmod(i, ntx)
            y0, x0 = ty * f.th, tx * f.tw
            if y0 >= f.height:
                break
            t = self.decode_tile(f, i)
            h, w = min(f.th, f.height - y0), min(f.tw, f.width - x0)
            img[y0:y0 + h, x0:x0 + w] = t[:h, :w]
        return img

    def read_region(self, location, size, level=0, _cache={}):
        """read_region of a level: location in level-0 coordinates (divided
        by the level's downsample), size in level pixels; pixels outside the
        level take the background value."""
        f0 = self.ifds[self.levels[0]]
        w, h = size
        if w <= 0 or h <= 0:
            raise ValueError("Invalid size")
        if w > f0.width or h > f0.height:
            raise ValueError("Invalid size (it exceeds the original image)")
        ds = self.downsamples[level]
        x, y = level_location(location[0], ds), level_location(location[1], ds)
        key = (id(self), level)
        if key not in _cache:
            _cache.clear()
            _cache[key] = self.level_image(level)
        img = _cache[key]
        H, W, C = img.shape
        out = np.full((h, w, C), self.background, np.uint8)
        xa, ya = max(x, 0), max(y, 0)
        xb, yb = min(x + w, W), min(y + h, H)
        if xb > xa and yb > ya:
            out[ya - y:yb - y, xa - x:xb - x] = img[ya:yb, xa:xb]
        return out


def _force_transform(stream, transform):
    """The stream with its APP0 (JFIF) and APP14 (Adobe) segments replaced
    by one Adobe segment of the given transform (0: components are RGB, 1:
    YCbCr), so that libjpeg decodes it in that colour space."""
    s = bytes(stream)
    out = bytearray(s[:2])
    p = 2
    while p + 4 <= len(s) and s[p] == 0xFF and s[p + 1] != 0xDA:
        L = (s[p + 2] << 8) | s[p + 3]
        if s[p + 1] not in (0xE0, 0xEE):
            out += s[p:p + 2 + L]
        p += 2 + L
    adobe = b"Adobe" + bytes([0, 100, 0, 0, 0, 0, transform])
    out = out[:2] + b"\xff\xee" + struct.pack(">H", len(adobe) + 2) + adobe + out[2:]
    return bytes(out) + s[p:]
