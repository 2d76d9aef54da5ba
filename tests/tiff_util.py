"""A TIFF reader for the tests (not collected), written from the TIFF 6.0
specification: header, one IFD, uncompressed strips of 16-bit grey samples."""
import struct

import numpy as np

TYPES = {3: ("H", 2), 4: ("I", 4)}      # SHORT, LONG


def read_tiff(path):
    """(tags {tag: [values]}, image uint16 [height][width])."""
    data = open(path, "rb").read()
    order = {b"II": "<", b"MM": ">"}[data[:2]]
    magic, ifd = struct.unpack_from(order + "HI", data, 2)
    assert magic == 42
    (n,) = struct.unpack_from(order + "H", data, ifd)
    tags = {}
    for k in range(n):
        tag, typ, count, raw = struct.unpack_from(order + "HHI4s", data, ifd + 2 + 12 * k)
        fmt, size = TYPES[typ]
        if size * count <= 4:
            vals = list(struct.unpack_from(order + fmt * count, raw, 0))
        else:
            (off,) = struct.unpack(order + "I", raw)
            vals = list(struct.unpack_from(order + fmt * count, data, off))
        tags[tag] = vals
    (next_ifd,) = struct.unpack_from(order + "I", data, ifd + 2 + 12 * n)
    assert next_ifd == 0
    w, h = tags[256][0], tags[257][0]
    assert tags[258] == [16] and tags[259] == [1] and tags[277] == [1]
    rows = bytearray()
    for off, cnt in zip(tags[273], tags[279]):
        rows += data[off:off + cnt]
    img = np.frombuffer(bytes(rows), dtype=order + "u2").reshape(h, w)
    return tags, img
