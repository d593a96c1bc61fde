"""Minimal OpenEXR reader for the tests: single-part scanline files with
uncompressed FLOAT channels, parsed from the OpenEXR file-layout description
(magic, version, attribute header, scanline offset table, scanline blocks).
Returns the attributes and a dict channel name -> [H][W] float32 array in
file row order (row 0 = top)."""
import struct

import numpy as np


def read_exr(path):
    b = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", b, 0)
    assert magic == 20000630, magic
    assert version & 0xFF == 2, version
    assert version >> 8 == 0, "tiled, long-name, deep or multi-part flags set"
    o = 8
    attrs = {}

    def cstr(o):
        e = b.index(b"\0", o)
        return b[o:e].decode(), e + 1
    while b[o] != 0:
        name, o = cstr(o)
        typ, o = cstr(o)
        (size,) = struct.unpack_from("<i", b, o)
        o += 4
        attrs[name] = (typ, b[o:o + size])
        o += size
    o += 1
    for req, typ in [("channels", "chlist"), ("compression", "compression"),
                     ("dataWindow", "box2i"), ("displayWindow", "box2i"),
                     ("lineOrder", "lineOrder"), ("pixelAspectRatio", "float"),
                     ("screenWindowCenter", "v2f"), ("screenWindowWidth", "float")]:
        assert attrs[req][0] == typ, (req, attrs.get(req))
    assert attrs["compression"][1] == b"\0", "compressed"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    chans = []
    cl = attrs["channels"][1]
    p = 0
    while cl[p] != 0:
        e = cl.index(b"\0", p)
        name = cl[p:e].decode()
        ptype, plin, xs, ys = struct.unpack_from("<iB3xii", cl, e + 1)
        assert ptype == 2 and xs == 1 and ys == 1, (name, ptype, xs, ys)
        chans.append(name)
        p = e + 1 + 16
    assert chans == sorted(chans), "channels must be stored in name order"
    offsets = struct.unpack_from(f"<{H}Q", b, o)
    out = {c: np.zeros((H, W), np.float32) for c in chans}
    last = -1
    for off in offsets:
        y, size = struct.unpack_from("<ii", b, off)
        assert size == W * 4 * len(chans) and y > last
        last = y
        for k, c in enumerate(chans):
            out[c][y - y0] = np.frombuffer(b, np.float32, W, off + 8 + k * W * 4)
    return attrs, out
