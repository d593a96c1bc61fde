"""Host checks of the traversal trees (tests/cpp/bvh_check.cpp, built from the
library's own sah.cpp / bvh4.cpp): SAH tree validity, leaf coverage of the
4-wide and compressed 8-wide trees, conservative 8-bit box decoding, and a
host walk of the compressed tree against brute force."""
import os
import subprocess

import numpy as np
import pytest

from mray_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "bvh_check")


@pytest.mark.parametrize("name", ["soup", "terrain", "cornell", "flat_grid"])
def test_trees(tmp_path, name):
    if not os.path.exists(CHECK):
        pytest.fail("tests/cpp/bvh_check is not built; run __graft_entry__.build()")
    if name == "soup":
        s = scenes.random_triangles(6000, seed=3)
    elif name == "terrain":
        s = scenes.terrain_spheres(grid=96, spheres=60, width=64, height=36)
    elif name == "cornell":
        s = scenes.cornell()
    else:
        s = None
    if s is not None:
        tri = np.asarray(s.positions, np.float32)[np.asarray(s.indices)].reshape(-1, 9)
    else:
        # axis-aligned flat quads: zero-extent boxes on one axis everywhere
        q = []
        for i in range(20):
            for j in range(20):
                x0, y0 = i / 20, j / 20
                a, b, c, d = [x0, y0, 0], [x0 + .05, y0, 0], [x0 + .05, y0 + .05, 0], [x0, y0 + .05, 0]
                q += [a + b + c, a + c + d]
        tri = np.array(q, np.float32)
    path = str(tmp_path / "soup.bin")
    with open(path, "wb") as f:
        f.write(np.array([len(tri)], np.uint32).tobytes())
        f.write(tri.astype(np.float32).tobytes())
    r = subprocess.run([CHECK, path], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
