"""Writes small GeoTIFF files for the raster reader's tests: single band,
strips or tiles, no / Deflate / LZW compression, horizontal and
floating-point predictors, either byte order, classic TIFF or BigTIFF, a
geotransform (tiepoint + pixel scale, or a model transformation), GDAL's
NoData tag and its scale / offset metadata, PixelIsPoint or PixelIsArea.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

FORMATS = {np.dtype(np.uint8): 1, np.dtype(np.uint16): 1, np.dtype(np.uint32): 1, np.dtype(np.int8): 2,
           np.dtype(np.int16): 2, np.dtype(np.int32): 2, np.dtype(np.float32): 3, np.dtype(np.float64): 3}


def _width(nxt: int, width: int) -> int:
    """Code width once the table's next free entry is `nxt`."""
    return 12 if nxt >= 2048 else 11 if nxt >= 1024 else 10 if nxt >= 512 else width


def lzw_encode(data: bytes) -> bytes:
    """TIFF LZW: MSB-first codes, early change, a clear code first and when
    the table fills, end-of-information last."""
    out, acc, nbits = bytearray(), 0, 0

    def put(code, width):
        nonlocal acc, nbits
        acc = (acc << width) | code
        nbits += width
        while nbits >= 8:
            nbits -= 8
            out.append((acc >> nbits) & 0xFF)

    table = {bytes([i]): i for i in range(256)}
    nxt, width = 258, 9
    put(256, width)
    w = b""
    for b in data:
        wc = w + bytes([b])
        if wc in table:
            w = wc
            continue
        put(table[w], width)
        table[wc] = nxt
        nxt += 1
        if nxt == 4094:  # the table is full: clear and start again
            put(256, width)
            table = {bytes([i]): i for i in range(256)}
            nxt, width = 258, 9
        else:
            width = _width(nxt, width)
        w = bytes([b])
    if w:
        put(table[w], width)
        nxt += 1
        width = _width(nxt, width)
    put(257, width)
    if nbits:
        out.append((acc << (8 - nbits)) & 0xFF)
    return bytes(out)


def _predict(block: np.ndarray, predictor: int, endian: str) -> bytes:
    """block (rows, cols) in file byte order -> predicted bytes."""
    if predictor == 1:
        return block.astype(block.dtype.newbyteorder(endian)).tobytes()
    if predictor == 2:
        ints = block.astype(block.dtype)
        diff = ints.copy()
        diff[:, 1:] = ints[:, 1:] - ints[:, :-1]  # wraps in the sample width
        return diff.astype(block.dtype.newbyteorder(endian)).tobytes()
    # floating point: bytes of each row split into planes, most significant first, then byte differences
    rows, cols = block.shape
    bps = block.dtype.itemsize
    big = block.astype(block.dtype.newbyteorder(">")).view(np.uint8).reshape(rows, cols, bps)
    planes = big.transpose(0, 2, 1).reshape(rows, cols * bps).astype(np.int16)
    out = planes.copy()
    out[:, 1:] = planes[:, 1:] - planes[:, :-1]
    return (out & 0xFF).astype(np.uint8).tobytes()


def write(path, data, transform=(0.0, 1.0, 0.0, 0.0, 0.0, -1.0), nodata=None, scale=None, offset=None,
          compression="none", predictor=1, tile=None, rows_per_strip=None, endian="<", bigtiff=False,
          pixel_is_point=False, model_transformation=False, geokeys=(1024, 1, 1)):
    """Writes `data` (2-D numpy array) as band 1 of a GeoTIFF."""
    data = np.asarray(data)
    h, w = data.shape
    fmt = FORMATS[data.dtype]
    bps = data.dtype.itemsize
    comp = {"none": 1, "lzw": 5, "deflate": 8}[compression]

    def encode(block):
        raw = _predict(block, predictor, endian)
        if comp == 5:
            return lzw_encode(raw)
        if comp == 8:
            return zlib.compress(raw)
        return raw

    blocks = []
    if tile:
        tw, tl = tile
        for ty in range(0, h, tl):
            for tx in range(0, w, tw):
                b = np.zeros((tl, tw), data.dtype)
                part = data[ty:ty + tl, tx:tx + tw]
                b[:part.shape[0], :part.shape[1]] = part
                blocks.append(encode(b))
    else:
        rps = rows_per_strip or h
        for y in range(0, h, rps):
            blocks.append(encode(data[y:y + rps]))

    tags = {256: (4, [w]), 257: (4, [h]), 258: (3, [bps * 8]), 259: (3, [comp]), 262: (3, [1]), 277: (3, [1]),
            284: (3, [1]), 339: (3, [fmt])}
    if predictor != 1:
        tags[317] = (3, [predictor])
    if tile:
        tags[322] = (4, [tile[0]])
        tags[323] = (4, [tile[1]])
    else:
        tags[278] = (4, [rows_per_strip or h])
    gt = transform
    if model_transformation:
        tags[34264] = (12, [gt[1], gt[2], 0, gt[0], gt[4], gt[5], 0, gt[3], 0, 0, 0, 0, 0, 0, 0, 1])
    else:
        assert gt[2] == 0 and gt[4] == 0
        x0, y0 = gt[0], gt[3]
        if pixel_is_point:  # GDAL writes the tiepoint of the pixel centre
            x0, y0 = gt[0] + 0.5 * gt[1], gt[3] + 0.5 * gt[5]
        tags[33550] = (12, [gt[1], -gt[5], 0.0])
        tags[33922] = (12, [0.0, 0.0, 0.0, x0, y0, 0.0])
    keys = [1, 1, 0, 0]
    entries = [(1025, 0, 1, 2 if pixel_is_point else 1)]
    for k in range(0, len(geokeys), 3):
        entries.append((geokeys[k], 0, 1, geokeys[k + 1] * 1000 + geokeys[k + 2]))
    keys[3] = len(entries)
    for e in sorted(entries):
        keys.extend(e)
    tags[34735] = (3, keys)
    if nodata is not None:
        tags[42113] = (2, f"{nodata}\0".encode())
    if scale is not None or offset is not None:
        items = ""
        if offset is not None:
            items += f'<Item name="OFFSET" sample="0" role="offset">{offset}</Item>'
        if scale is not None:
            items += f'<Item name="SCALE" sample="0" role="scale">{scale}</Item>'
        tags[42112] = (2, f"<GDALMetadata>{items}</GDALMetadata>\0".encode())

    # layout: header, block data, out-of-line tag values, IFD
    e = endian
    head = 16 if bigtiff else 8
    body = bytearray()
    offsets = []
    for b in blocks:
        offsets.append(head + len(body))
        body += b
        if len(body) % 2:
            body += b"\0"
    key = 324 if tile else 273
    tags[key] = (16 if bigtiff else 4, offsets)
    tags[325 if tile else 279] = (16 if bigtiff else 4, [len(b) for b in blocks])
    tsize = {1: 1, 2: 1, 3: 2, 4: 4, 12: 8, 16: 8}
    tfmt = {1: "B", 3: "H", 4: "I", 12: "d", 16: "Q"}
    inline = 8 if bigtiff else 4
    extra = bytearray()
    ifd_entries = []
    n = len(tags)
    ifd_at = head + len(body)
    ifd_size = (8 + n * 20 + 8) if bigtiff else (2 + n * 12 + 4)
    extra_at = ifd_at + ifd_size
    for tag in sorted(tags):
        typ, vals = tags[tag]
        if typ == 2:
            raw = vals
            count = len(raw)
        else:
            raw = struct.pack(e + tfmt[typ] * len(vals), *vals)
            count = len(vals)
        if len(raw) <= inline:
            field = raw.ljust(inline, b"\0")
        else:
            field = struct.pack(e + ("Q" if bigtiff else "I"), extra_at + len(extra))
            extra += raw
            if len(extra) % 2:
                extra += b"\0"
        if bigtiff:
            ifd_entries.append(struct.pack(e + "HHQ", tag, typ, count) + field)
        else:
            ifd_entries.append(struct.pack(e + "HHI", tag, typ, count) + field)
    order = b"II" if e == "<" else b"MM"
    if bigtiff:
        header = order + struct.pack(e + "HHHQ", 43, 8, 0, ifd_at)
        ifd = struct.pack(e + "Q", n) + b"".join(ifd_entries) + struct.pack(e + "Q", 0)
    else:
        header = order + struct.pack(e + "HI", 42, ifd_at)
        ifd = struct.pack(e + "H", n) + b"".join(ifd_entries) + struct.pack(e + "I", 0)
    with open(path, "wb") as f:
        f.write(header + bytes(body) + ifd + bytes(extra))
