"""Independent decoders for the codec tests: PNG (zlib + the five row
filters, non-interlaced, 8/16-bit) and TIFF (first IFD, chunky strips,
uncompressed or LZW, integer and float samples), written in Python from
the PNG specification and TIFF 6.0.  Test infrastructure only."""
import struct
import zlib

import numpy as np


def png_decode(data):
    """(h, w, channels) array of a non-interlaced PNG (no palette/tRNS)."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        t = data[pos + 4:pos + 8]
        d = data[pos + 8:pos + 8 + n]
        if t == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", d)
        elif t == b"IDAT":
            idat += d
        elif t == b"IEND":
            break
        pos += 12 + n
    w, h, depth, ctype, _, _, inter = hdr
    assert inter == 0 and depth in (8, 16) and ctype in (0, 2, 4, 6)
    ch = {0: 1, 2: 3, 4: 2, 6: 4}[ctype]
    bpp = ch * depth // 8
    rb = w * bpp
    raw = zlib.decompress(idat)
    out = bytearray()
    prev = bytearray(rb)
    for j in range(h):
        ft = raw[j * (rb + 1)]
        line = bytearray(raw[j * (rb + 1) + 1:(j + 1) * (rb + 1)])
        for i in range(rb):
            a = line[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if ft == 1:
                line[i] = (line[i] + a) & 255
            elif ft == 2:
                line[i] = (line[i] + b) & 255
            elif ft == 3:
                line[i] = (line[i] + (a + b) // 2) & 255
            elif ft == 4:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pr = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                line[i] = (line[i] + pr) & 255
        out += line
        prev = line
    dt = np.uint8 if depth == 8 else np.dtype(">u2")
    return np.frombuffer(bytes(out), dt).reshape(h, w, ch).astype(
        np.uint8 if depth == 8 else np.uint16)


def lzw_decode(src):
    """TIFF LZW (MSB-first codes, 9-12 bits, early change)."""
    table = [bytes([i]) for i in range(256)] + [b"", b""]
    out = bytearray()
    bit, width, old = 0, 9, None
    nbits = len(src) * 8
    while bit + width <= nbits:
        code = 0
        for _ in range(width):
            code = (code << 1) | ((src[bit >> 3] >> (7 - (bit & 7))) & 1)
            bit += 1
        if code == 257:
            break
        if code == 256:
            table = table[:258]
            width, old = 9, None
            continue
        if old is None:
            out += table[code]
            old = code
            continue
        if code < len(table):
            s = table[code]
            table.append(table[old] + s[:1])
        else:
            s = table[old] + table[old][:1]
            table.append(s)
        out += s
        if len(table) >= (1 << width) - 1 and width < 12:
            width += 1
        old = code
    return bytes(out)


def tiff_decode(data):
    """(h, w, spp) array of the first image of a chunky, striped TIFF."""
    be = data[:2] == b"MM"
    e = ">" if be else "<"
    assert struct.unpack(e + "H", data[2:4])[0] == 42
    ifd = struct.unpack(e + "I", data[4:8])[0]
    n = struct.unpack(e + "H", data[ifd:ifd + 2])[0]
    sizes = {1: 1, 3: 2, 4: 4}
    fmts = {1: "B", 3: "H", 4: "I"}
    tags = {}
    for k in range(n):
        t, typ, cnt = struct.unpack(e + "HHI", data[ifd + 2 + 12 * k:ifd + 10 + 12 * k])
        if typ not in sizes:
            continue
        nb = sizes[typ] * cnt
        off = ifd + 10 + 12 * k if nb <= 4 else struct.unpack(
            e + "I", data[ifd + 10 + 12 * k:ifd + 14 + 12 * k])[0]
        tags[t] = struct.unpack(e + fmts[typ] * cnt, data[off:off + nb])
    w, h = tags[256][0], tags[257][0]
    spp = tags.get(277, (1,))[0]
    bps = tags.get(258, (1,))[0]
    fmt = tags.get(339, (1,))[0]
    comp = tags.get(259, (1,))[0]
    assert tags.get(284, (1,))[0] == 1 and tags.get(317, (1,))[0] == 1
    raw = b""
    for o, c in zip(tags[273], tags[279]):
        chunk = data[o:o + c]
        raw += lzw_decode(chunk) if comp == 5 else chunk
    kind = {1: "u", 2: "i", 3: "f"}[fmt]
    dt = np.dtype(e + kind + str(bps // 8))
    return np.frombuffer(raw[:w * h * spp * bps // 8], dt).reshape(h, w, spp).astype(
        dt.newbyteorder("="))


# ------------------------------------------------------------ writers
# Test-file writers covering the variants of the formats the C decoders
# must read: PNG colour types, depths, palettes, tRNS, Adam7 and every row
# filter; TIFF byte orders, strips and tiles, planar layouts, compressions
# and predictors.

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4),
         (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(t, d):
    return struct.pack(">I", len(d)) + t + d + struct.pack(
        ">I", zlib.crc32(t + d) & 0xffffffff)


def _pack_row(vals, depth):
    if depth == 8:
        return bytes(int(v) for v in vals)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in vals)
    out, acc, nb = bytearray(), 0, 0
    for v in vals:
        acc = (acc << depth) | int(v)
        nb += depth
        if nb == 8:
            out.append(acc)
            acc, nb = 0, 0
    if nb:
        out.append(acc << (8 - nb))
    return bytes(out)


def _filter(line, prev, bpp, ft):
    out = bytearray(len(line))
    for i in range(len(line)):
        a = line[i - bpp] if i >= bpp else 0
        b = prev[i] if prev is not None else 0
        c = prev[i - bpp] if prev is not None and i >= bpp else 0
        if ft == 0:
            p = 0
        elif ft == 1:
            p = a
        elif ft == 2:
            p = b
        elif ft == 3:
            p = (a + b) // 2
        else:
            q = a + b - c
            pa, pb, pc = abs(q - a), abs(q - b), abs(q - c)
            p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
        out[i] = (line[i] - p) & 255
    return bytes([ft]) + bytes(out)


def png_encode(samples, ctype, depth, palette=None, trns=None, interlace=False):
    """samples: (h, w, ch) integers as stored (palette indices for ctype 3)."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    ft = 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = None
        for row in sub:
            line = _pack_row(row.reshape(-1), depth)
            raw += _filter(line, prev, bpp, ft % 5)
            ft += 1
            prev = line
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(
        ">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    z = zlib.compress(bytes(raw), 9)
    out += _chunk(b"IDAT", z[:len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])
    return out + _chunk(b"IEND", b"")


def lzw_encode(data):
    """TIFF LZW with early change, a clear code at the start and when the
    table fills."""
    out, acc, nacc = bytearray(), 0, 0

    def put(code, width):
        nonlocal acc, nacc
        acc = (acc << width) | code
        nacc += width
        while nacc >= 8:
            out.append((acc >> (nacc - 8)) & 255)
            nacc -= 8
        acc &= (1 << nacc) - 1

    width, table = 9, {}
    nxt = 258
    put(256, width)
    if data:
        ent = bytes(data[:1])
        for c in data[1:]:
            s = ent + bytes([c])
            if s in table or len(s) == 1:
                ent = s
                continue
            put(table[ent] if len(ent) > 1 else ent[0], width)
            table[s] = nxt
            nxt += 1
            ent = bytes([c])
            if nxt == 4094:
                put(256, width)
                table, nxt, width = {}, 258, 9
            elif nxt > (1 << width) - 1:
                width += 1
        put(table[ent] if len(ent) > 1 else ent[0], width)
        nxt += 1
        if nxt == 4094:
            put(256, width)
            width = 9
        elif nxt > (1 << width) - 1:
            width += 1
    put(257, width)
    if nacc:
        put(0, 8 - nacc)
    return bytes(out)


def packbits_encode(data):
    out, i = bytearray(), 0
    while i < len(data):
        j = i
        while j + 1 < len(data) and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j > i:
            out += bytes([(1 - (j - i + 1)) & 255, data[i]])
            i = j + 1
            continue
        j = i
        while j < len(data) and j - i < 128 and not (
                j + 1 < len(data) and data[j + 1] == data[j]):
            j += 1
        j = max(j, i + 1)
        out += bytes([j - i - 1]) + bytes(data[i:j])
        i = j
    return bytes(out)


def tiff_encode(arr, be=False, compression=1, predictor=1, tiles=None, planar=1,
                rps=None, bits=None, sampleformat=None):
    """arr: (h, w, spp) numpy array.  bits: 1/2/4 for packed unsigned."""
    h, w, spp = arr.shape
    e = ">" if be else "<"
    dt = arr.dtype
    fmt = sampleformat or {"u": 1, "i": 2, "f": 3}[dt.kind]
    bps = bits or dt.itemsize * 8

    def chunk_bytes(a):          # a: (rows, cols, s) of one chunk
        rows = []
        for r in a:
            if bits:
                rows.append(_pack_row(r.reshape(-1), bits))
                continue
            r = np.ascontiguousarray(r)
            if predictor == 2:
                x = r.reshape(-1).view(np.dtype("u%d" % dt.itemsize)).copy()
                s = r.shape[1]
                x[s:] = (x[s:] - x[:-s]).astype(x.dtype)
                rows.append(x.astype(np.dtype(e + "u%d" % dt.itemsize)).tobytes())
            elif predictor == 3:
                x = r.reshape(-1).astype(np.dtype(">" + dt.str[1:]))
                b = np.frombuffer(x.tobytes(), np.uint8).reshape(-1, dt.itemsize)
                planes = b.T.reshape(-1).copy()          # MSB plane first
                s = r.shape[1]
                planes[s:] = (planes[s:].astype(np.int16) - planes[:-s]).astype(np.uint8)
                rows.append(planes.tobytes())
            else:
                rows.append(r.astype(np.dtype(e + dt.str[1:])).tobytes())
        raw = b"".join(rows)
        if compression == 5:
            return lzw_encode(raw)
        if compression in (8, 32946):
            return zlib.compress(raw)
        if compression == 32773:
            return packbits_encode(raw)
        return raw

    planes = [arr] if planar == 1 else [arr[..., l:l + 1] for l in range(spp)]
    chunks = []
    if tiles:
        tw, tl = tiles
        for p in planes:
            for ty in range(0, h, tl):
                for tx in range(0, w, tw):
                    t = np.zeros((tl, tw, p.shape[2]), dt)
                    blk = p[ty:ty + tl, tx:tx + tw]
                    t[:blk.shape[0], :blk.shape[1]] = blk
                    chunks.append(chunk_bytes(t))
    else:
        rps = rps or h
        for p in planes:
            for y in range(0, h, rps):
                chunks.append(chunk_bytes(p[y:y + rps]))
    body = bytearray(b"MM\x00\x2a" if be else b"II\x2a\x00") + b"\0\0\0\0"
    offs = []
    for c in chunks:
        offs.append(len(body))
        body += c
        if len(body) & 1:
            body += b"\0"

    def arr_off(vals, code):
        nonlocal body
        o = len(body)
        body += struct.pack(e + code * len(vals), *vals)
        if len(body) & 1:
            body += b"\0"
        return o

    entries = []

    def add(tag, typ, vals):
        size = {3: 2, 4: 4}[typ] * len(vals)
        code = "H" if typ == 3 else "I"
        if size <= 4:
            v = struct.pack(e + code * len(vals), *vals).ljust(4, b"\0")
        else:
            v = struct.pack(e + "I", arr_off(vals, code))
        entries.append((tag, struct.pack(e + "HHI", tag, typ, len(vals)) + v))

    add(256, 4, [w])
    add(257, 4, [h])
    add(258, 3, [bps] * spp)
    add(259, 3, [compression])
    add(262, 3, [2 if spp == 3 else 1])
    add(277, 3, [spp])
    add(284, 3, [planar])
    if predictor != 1:
        add(317, 3, [predictor])
    add(339, 3, [fmt] * spp)
    if tiles:
        add(322, 4, [tiles[0]])
        add(323, 4, [tiles[1]])
        add(324, 4, offs)
        add(325, 4, [len(c) for c in chunks])
    else:
        add(273, 4, offs)
        add(278, 4, [rps])
        add(279, 4, [len(c) for c in chunks])
    entries.sort()
    ifd = len(body)
    body += struct.pack(e + "H", len(entries)) + b"".join(v for _, v in entries) + b"\0\0\0\0"
    body[4:8] = struct.pack(e + "I", ifd)
    return bytes(body)
