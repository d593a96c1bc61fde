""".mrspectra files: the reference's RGB->spectrum LUT format.

Writer layout (SpectraLUTGen/main.cpp:360-376, 590-609): "MR_SPECTRA",
u32 N = 64, u32 mode = 1, then nine float volumes, offset of coefficient c
of table l at voxel (x, y, z) = ((l * 3 + c) * N + z) * N^2 + y * N + x.
Reader checks (SpectrumContext.cu:317-351): character code (with a distinct
message for a byte-reversed one), N, mode, then a full read.  The tests
check the header bytes, the volume order against the in-memory tables,
the read-back, each rejection, and that a file placed with
mray_spectral_use_mrspectra drives the tables a renderer would use.
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from mray_amd import native
from oracle import refcpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 64
N3 = N ** 3


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    path = tmp_path_factory.mktemp("lut") / "REC_709.mrspectra"
    native.write_mrspectra(path)
    return path, native.spectral_tables()


def test_header_and_layout(written):
    path, t = written
    raw = path.read_bytes()
    assert len(raw) == 10 + 4 + 4 + 9 * N3 * 4
    assert raw[:10] == b"MR_SPECTRA"
    assert struct.unpack("<II", raw[10:18]) == (N, 1)
    vol = np.frombuffer(raw, np.float32, offset=18).reshape(3, 3, N, N, N)
    lut = t["lut"]                       # [l][z][y][x][c0, c1, c2, 0]
    for l in range(3):
        for c in range(3):
            assert np.array_equal(vol[l, c], lut[l, ..., c]), (l, c)
    # one voxel by the writer's own offset formula
    l, c, x, y, z = 2, 1, 5, 17, 40
    off = ((l * 3 + c) * N + z) * N * N + y * N + x
    assert vol.reshape(-1)[off] == lut[l, z, y, x, c]


def test_read_back(written):
    path, t = written
    assert np.array_equal(native.read_mrspectra(path), t["lut"])


def _corrupt(path, tmp_path, name, fn):
    raw = bytearray(path.read_bytes())
    fn(raw)
    out = tmp_path / name
    out.write_bytes(bytes(raw))
    return out


@pytest.mark.parametrize("name,fn,msg", [
    ("rev", lambda b: b.__setitem__(slice(0, 4), b"S_RM"), "Endianness"),
    ("cc", lambda b: b.__setitem__(slice(3, 5), b"XX"), "Wrong character code"),
    ("n", lambda b: b.__setitem__(slice(10, 14), struct.pack("<I", 32)), "Wrong size"),
    ("mode", lambda b: b.__setitem__(slice(14, 18), struct.pack("<I", 0)), "Wrong mode"),
    ("short", lambda b: b.__delitem__(slice(len(b) - 4, len(b))), "Unable to read"),
    ("nan", lambda b: b.__setitem__(slice(18, 22), struct.pack("<f", float("nan"))), "non-finite"),
])
def test_rejections(written, tmp_path, name, fn, msg):
    path, _ = written
    bad = _corrupt(path, tmp_path, name, fn)
    with pytest.raises(native.MRayError, match=msg):
        native.read_mrspectra(bad)


def test_file_tables_convert_like_the_built_ones(written):
    """The oracle's RGB->spectrum with tables read from the file equals the
    conversion with the built tables (same coefficients, same lookup)."""
    path, t = written
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0, 1, (64, 3)).astype(np.float32)
    waves = rng.uniform(380, 720, (64, 4)).astype(np.float32)
    refcpu.spectral_load(t)
    a = refcpu.rgb_to_spectrum(rgb, waves)
    refcpu.spectral_load(dict(t, lut=native.read_mrspectra(path)))
    b = refcpu.rgb_to_spectrum(rgb, waves)
    refcpu.spectral_load(t)
    assert np.array_equal(a, b)


def test_use_file_replaces_tables(written, tmp_path):
    """A file given to mray_spectral_use_mrspectra becomes the table set (in
    a fresh process, so this session's tables stay the built ones)."""
    path, _ = written
    raw = bytearray(path.read_bytes())
    vol = np.frombuffer(raw, np.float32, offset=18).copy()
    vol[:N3] += 1.0                      # table 0, c0 shifted
    mod = tmp_path / "mod.mrspectra"
    mod.write_bytes(bytes(raw[:18]) + vol.tobytes())
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "from mray_amd import native\n"
        f"native.check(native.lib().mray_spectral_use_mrspectra({str(mod)!r}.encode()))\n"
        "t = native.spectral_tables()\n"
        f"ref = native.read_mrspectra({str(mod)!r})\n"
        "print('same', np.array_equal(t['lut'], ref))\n"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "same True" in r.stdout
