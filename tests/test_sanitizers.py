"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

* The TracerI layer, the MRay JSON scene loader and its parser
  (mray_amd/host) take untrusted scene files: the sanitizer build of the
  TracerI front-end driver loads the reference's scene fixtures, every
  truncation of one, seeded byte mutations and a generated Cornell scene.
  Each run must end normally or with an MRayError, never with a sanitizer
  report.
* The oracle (oracle/refcpu.cpp) runs a small render, an LBVH build, a cast
  and a textured Skysphere under the same sanitizers, loaded into a child
  Python with the ASan runtime preloaded.
The builds are the `sanitize` targets of mray_amd/host/Makefile and
oracle/Makefile (g++; no GPU code).
"""
import fcntl
import os
import subprocess
import sys

import numpy as np
import pytest

from mray_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "_build", "san")
GOLDEN = os.path.join(ROOT, "tests", "golden", "scenes")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")
REPORT = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def _make(target_dir):
    """`make sanitize` under a file lock: parallel test workers share the
    build directory."""
    os.makedirs(SAN, exist_ok=True)
    with open(os.path.join(SAN, ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", target_dir, "sanitize"], check=True,
                       capture_output=True, timeout=600)


@pytest.fixture(scope="module")
def san_driver():
    _make(os.path.join(ROOT, "mray_amd", "host"))
    return os.path.join(SAN, "tracer_driver"), os.path.join(SAN, "libmray_tracer.so")


def _run(drv, path, tmp_path):
    exe, lib = drv
    r = subprocess.run([exe, lib, path, "flatten", str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120, env=ENV)
    text = r.stdout + r.stderr
    assert not any(k in text for k in REPORT), text[-3000:]
    # 0: loaded and flattened; 3: an MRayError reached the driver
    assert r.returncode in (0, 3), (r.returncode, text[-2000:])
    return r.returncode


def test_scene_fixtures_under_sanitizers(san_driver, tmp_path):
    assert _run(san_driver, os.path.join(GOLDEN, "basic.json"), tmp_path) == 0
    assert _run(san_driver, os.path.join(GOLDEN, "empty.json"), tmp_path) == 3
    _run(san_driver, os.path.join(GOLDEN, "minimal_valid.json"), tmp_path)
    src = str(tmp_path / "cornell.json")
    scenes.to_mray_json(scenes.cornell(), src)
    assert _run(san_driver, src, tmp_path) == 0


def test_truncated_scenes_under_sanitizers(san_driver, tmp_path):
    text = open(os.path.join(GOLDEN, "basic.json"), "rb").read()
    path = tmp_path / "cut.json"
    for n in list(range(0, len(text), 23)) + [text.rindex(b"}")]:
        path.write_bytes(text[:n])
        assert _run(san_driver, str(path), tmp_path) == 3


def test_mutated_scenes_under_sanitizers(san_driver, tmp_path):
    text = bytearray(open(os.path.join(GOLDEN, "basic.json"), "rb").read())
    rng = np.random.default_rng(12)
    alphabet = b'{}[]",:-+.eE0123456789 \n/*abcnrtu\\'
    path = tmp_path / "mut.json"
    for _ in range(120):
        m = bytearray(text)
        for _ in range(int(rng.integers(1, 4))):
            i = int(rng.integers(0, len(m)))
            op = int(rng.integers(0, 3))
            if op == 0:
                m[i] = alphabet[int(rng.integers(0, len(alphabet)))]
            elif op == 1:
                del m[i]
            else:
                m.insert(i, alphabet[int(rng.integers(0, len(alphabet)))])
        path.write_bytes(bytes(m))
        _run(san_driver, str(path), tmp_path)
    # numbers the loader must reject instead of casting: negative, fractional
    # and huge ids and indices, a non-finite-range coordinate
    bad = b'"primitive": 0'
    assert bad in text
    for rep in (b'"primitive": -1', b'"primitive": 0.5', b'"primitive": 1e300',
                b'"primitive": 4294967296'):
        path.write_bytes(bytes(text).replace(bad, rep))
        assert _run(san_driver, str(path), tmp_path) == 3
    path.write_bytes(bytes(text).replace(b"[   0,  0.5, 0]", b"[1e300, 0.5, 0]"))
    assert _run(san_driver, str(path), tmp_path) == 3


ORACLE_SCRIPT = r"""
import sys
sys.path.insert(0, ROOT)
import numpy as np
import oracle.refcpu as rc
rc._LIB = LIB
from mray_amd import abi, scenes
o = rc.OracleScene(scenes.cornell())
film, st = o.render(abi.render_params(16, 16, 4, rr=(3, 8), seed=1), threads=2)
assert st["dead_paths"] == 16 * 16 * 4 and np.isfinite(film).all()
s = scenes.terrain_spheres(grid=24, spheres=6, width=16, height=9)
ot = rc.OracleScene(s)
ot.lbvh()
rays, _ = ot.camera_rays(abi.render_params(16, 9, 1), 144)
ot.cast_closest(rays, brute=False)
ot.cast_visibility(rays)
rgb = np.random.default_rng(1).uniform(0.1, 2.0, size=(8, 16, 3)).astype(np.float32)
sc = scenes.cornell()
sc.boundary_type = abi.BOUNDARY_SKYSPHERE_COOCTA
sc.boundary_radiance = (1.0, 1.0, 1.0)
oe = rc.OracleScene(sc)
oe.set_environment(rgb)
oe.render(abi.render_params(8, 8, 2, rr=(2, 4), seed=3), threads=2)
print("oracle-ok")
"""


def test_oracle_under_sanitizers():
    _make(os.path.join(ROOT, "oracle"))
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    ubsan = subprocess.run(["g++", "-print-file-name=libubsan.so"], capture_output=True,
                           text=True).stdout.strip()
    env = dict(ENV, LD_PRELOAD=f"{asan}:{ubsan}")
    code = f"ROOT = {ROOT!r}\nLIB = {os.path.join(SAN, 'librefcpu.so')!r}\n" + ORACLE_SCRIPT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=600, env=env)
    text = r.stdout + r.stderr
    assert not any(k in text for k in REPORT), text[-4000:]
    assert r.returncode == 0 and "oracle-ok" in r.stdout, text[-4000:]
