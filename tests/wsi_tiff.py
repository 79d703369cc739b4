"""Writer of small tiled / stripped (Big)TIFF pyramids for the region-read
tests: raw, deflate, LZW or JPEG tiles (JPEG through Pillow's libjpeg-turbo,
optionally with shared JPEGTables and restart markers), little- or big-endian,
classic or BigTIFF.  Test infrastructure; files go to a temporary folder.
"""

import io
import struct
import zlib

import numpy as np


def lzw_encode(data):
    """TIFF LZW: clear code first, codes widen when the next free entry
    reaches the width's limit (the decoder's early change is one entry
    behind the encoder), a clear code before the table fills, EOI last."""
    out = bytearray()
    acc, nacc = 0, 0

    def put(code, width):
        nonlocal acc, nacc
        acc = (acc << width) | code
        nacc += width
        while nacc >= 8:
            out.append((acc >> (nacc - 8)) & 0xFF)
            nacc -= 8
            acc &= (1 << nacc) - 1

    table = {bytes([i]): i for i in range(256)}
    nxt, width = 258, 9
    put(256, width)
    w = b""
    for byte in data:
        wc = w + bytes([byte])
        if wc in table:
            w = wc
            continue
        put(table[w], width)
        table[wc] = nxt
        nxt += 1
        if nxt >= (1 << width) and width < 12:
            width += 1
        if nxt >= 4094:
            put(256, width)
            table = {bytes([i]): i for i in range(256)}
            nxt, width = 258, 9
        w = bytes([byte])
    if w:
        put(table[w], width)
        if nxt + 1 >= (1 << width) and width < 12:
            width += 1
    put(257, width)
    if nacc:
        out.append((acc << (8 - nacc)) & 0xFF)
    return bytes(out)


def _split_jpeg(stream):
    """(tables, abbreviated) of a JPEG stream: DQT/DHT segments move to a
    tables-only stream (SOI ... EOI)."""
    s = stream
    tables, rest = bytearray(b"\xff\xd8"), bytearray(b"\xff\xd8")
    p = 2
    while s[p] == 0xFF and s[p + 1] != 0xDA:
        L = (s[p + 2] << 8) | s[p + 3]
        (tables if s[p + 1] in (0xDB, 0xC4) else rest).extend(s[p:p + 2 + L])
        p += 2 + L
    tables += b"\xff\xd9"
    rest += s[p:]
    return bytes(tables), bytes(rest)


def encode_tile(px, compression, quality=90, subsampling=2, keep_rgb=False, restart=0,
                predictor=1):
    """Compressed bytes of one tile (th, tw, spp) uint8."""
    if compression == "jpeg":
        from PIL import Image

        im = Image.fromarray(px[:, :, 0] if px.shape[2] == 1 else px)
        b = io.BytesIO()
        kw = dict(quality=quality, subsampling=subsampling)
        if keep_rgb:
            kw["keep_rgb"] = True
        if restart:
            kw["restart_marker_blocks"] = restart
        im.save(b, "JPEG", **kw)
        return b.getvalue()
    a = px
    if predictor == 2:
        a = px.copy()
        a[:, 1:] = (px[:, 1:].astype(np.int16) - px[:, :-1]).astype(np.uint8)
    raw = np.ascontiguousarray(a).tobytes()
    if compression == "raw":
        return raw
    if compression == "deflate":
        return zlib.compress(raw, 6)
    if compression == "lzw":
        return lzw_encode(raw)
    raise ValueError(compression)


_CODE = {"raw": 1, "lzw": 5, "jpeg": 7, "deflate": 8}


def write_tiff(path, levels, tile=(64, 64), compression="jpeg", big=False, byteorder="<",
               quality=90, subsampling=2, photometric=None, predictor=1, tables=True,
               restart=0, description=b"", strip_rows=None, associated=(),
               resolution=None):
    """Write ``levels`` (list of (h, w, spp) uint8 arrays, largest first) as
    one IFD each, tiled with ``tile`` = (tw, th) or, with ``strip_rows``, in
    strips.  ``associated``: (array, subfile_type) images appended as
    stripped IFDs (label, macro ...).  ``photometric``: 'rgb' / 'ycbcr' /
    None (rgb for lossless, ycbcr for JPEG, minisblack for one sample)."""
    bo = byteorder
    ifds = []
    blobs = bytearray()
    hdr_len = 16 if big else 8

    def add(arr, tiled, comp, desc, subfile, rows=None):
        h, w, spp = arr.shape
        if tiled:
            tw, th = tile
        else:
            tw, th = w, min(rows or h, h)
        ph = photometric
        if ph is None:
            ph = "gray" if spp == 1 else ("ycbcr" if comp == "jpeg" else "rgb")
        ph_code = {"gray": 1, "rgb": 2, "ycbcr": 6}[ph]
        offs, cnts = [], []
        jt = b""
        for ty in range(-(-h // th)):
            for tx in range(-(-w // tw)) if tiled else [0]:
                blk = np.zeros((th, tw, spp), np.uint8)
                part = arr[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw]
                blk[:part.shape[0], :part.shape[1]] = part
                if not tiled:
                    blk = blk[:part.shape[0]]  # the last strip holds only its rows
                data = encode_tile(blk, comp, quality, subsampling, ph == "rgb" and spp == 3,
                                   restart, predictor)
                if comp == "jpeg" and tables:
                    jt, data = _split_jpeg(data)
                offs.append(hdr_len + len(blobs))
                cnts.append(len(data))
                blobs.extend(data)
                if len(blobs) % 2:
                    blobs.append(0)
        tags = {256: (4, [w]), 257: (4, [h]), 258: (3, [8] * spp), 259: (3, [_CODE[comp]]),
                262: (3, [ph_code]), 277: (3, [spp]), 284: (3, [1])}
        if subfile:
            tags[254] = (4, [subfile])
        if desc:
            tags[270] = (2, desc + b"\0")
        if comp != "jpeg" and predictor != 1:
            tags[317] = (3, [predictor])
        if resolution is not None and not ifds:
            tags[282] = (5, [resolution[0], 1])
            tags[283] = (5, [resolution[1], 1])
            tags[296] = (3, [resolution[2]])
        otype = 16 if big else 4
        if tiled:
            tags[322] = (3, [tw])
            tags[323] = (3, [th])
            tags[324] = (otype, offs)
            tags[325] = (otype, cnts)
        else:
            tags[273] = (otype, offs)
            tags[278] = (4, [th])
            tags[279] = (otype, cnts)
        if comp == "jpeg" and tables and jt:
            tags[347] = (7, jt)
        if comp == "jpeg" and ph == "ycbcr":
            f = {0: (1, 1), 1: (2, 1), 2: (2, 2)}[subsampling]
            tags[530] = (3, list(f))
        ifds.append(tags)

    for i, arr in enumerate(levels):
        add(arr, strip_rows is None, compression, description if i == 0 else b"", 0, strip_rows)
    for arr, sub in associated:
        add(arr, False, "raw" if compression != "jpeg" else compression, b"", sub, 16)

    # IFDs after the data, each followed by its out-of-line values
    fmt = {1: "B", 2: "B", 3: "H", 4: "I", 5: "I", 7: "B", 16: "Q"}
    size = {1: 1, 2: 1, 3: 2, 4: 4, 5: 4, 7: 1, 16: 8}
    out = bytearray(blobs)
    pos = hdr_len + len(out)
    ifd_offsets = []
    for tags in ifds:
        if pos % 2:
            out.append(0)
            pos += 1
        ifd_offsets.append(pos)
        n = len(tags)
        esz, csz, nsz, inl = (20, 8, 8, 8) if big else (12, 2, 4, 4)
        ifd_len = csz + n * esz + nsz
        extra = bytearray()
        ents = bytearray()
        for tag in sorted(tags):
            typ, vals = tags[tag]
            if isinstance(vals, (bytes, bytearray)):
                payload = bytes(vals)
                count = len(payload)
            else:
                count = len(vals) // 2 if typ == 5 else len(vals)
                payload = struct.pack(bo + fmt[typ] * len(vals), *vals)
            ents += struct.pack(bo + "HH", tag, typ)
            ents += struct.pack(bo + ("Q" if big else "I"), count)
            if len(payload) <= inl:
                ents += payload + bytes(inl - len(payload))
            else:
                off = pos + ifd_len + len(extra)
                ents += struct.pack(bo + ("Q" if big else "I"), off)
                extra += payload
                if len(extra) % 2:
                    extra.append(0)
        blk = struct.pack(bo + ("Q" if big else "H"), n) + ents + bytes(nsz) + extra
        out += blk
        pos += len(blk)
    # link the IFDs
    for k, off in enumerate(ifd_offsets[:-1]):
        n = len(ifds[k])
        link = off + (8 if big else 2) + n * (20 if big else 12) - hdr_len
        out[link:link + (8 if big else 4)] = struct.pack(bo + ("Q" if big else "I"),
                                                          ifd_offsets[k + 1])
    if big:
        hdr = (b"II" if bo == "<" else b"MM") + struct.pack(bo + "HHHQ", 43, 8, 0, ifd_offsets[0])
    else:
        hdr = (b"II" if bo == "<" else b"MM") + struct.pack(bo + "HI", 42, ifd_offsets[0])
    with open(path, "wb") as fh:
        fh.write(hdr + bytes(out))
    return str(path)


def pyramid(h, w, nlevels, seed=0, spp=3):
    """Smooth synthetic RGB tissue-like levels (largest first), each half the
    size of the one before, so that JPEG tiles are realistic."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    base = np.zeros((h, w, spp))
    for c in range(spp):
        for _ in range(6):
            fy, fx, ph = rng.uniform(0.01, 0.08), rng.uniform(0.01, 0.08), rng.uniform(0, 6.3)
            base[..., c] += np.sin(yy * fy + xx * fx + ph)
    base = (base - base.min()) / (np.ptp(base) + 1e-9) * 200 + 20
    base += rng.normal(0, 6, base.shape)
    lv = [np.clip(base, 0, 255).astype(np.uint8)]
    for _ in range(nlevels - 1):
        a = lv[-1].astype(np.float64)
        hh, ww = a.shape[0] // 2, a.shape[1] // 2
        a = a[:2 * hh, :2 * ww].reshape(hh, 2, ww, 2, spp).mean(axis=(1, 3))
        lv.append(np.clip(np.round(a), 0, 255).astype(np.uint8))
    return lv
