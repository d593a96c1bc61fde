"""Binary compatibility of the TracerI boundary with the reference's
(include/mray_tracer.hpp against Core/TracerI.h and the types it uses).

* tests/cpp/abi_layout: sizes, alignments and member offsets of every value
  type that crosses a TracerI call, how each by-value argument is passed,
  the vtable slot of the TracerI methods, the mangled names of the element
  types TransientData is tagged with, and the byte layout of TransientData
  and StaticVector at run time -- each against the figure the reference's
  declarations give on x86-64 (derivations in the file).  Its static checks
  fail the build, its run-time checks the exit status.
* libmray_tracer.so exports the factory pair and nothing of its own value
  types (they are built with hidden visibility, so their inline code cannot
  bind to a host's same-named symbols).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "abi_layout")
TRACER_LIB = os.path.join(ROOT, "mray_amd", "libmray_tracer.so")


def test_boundary_layout():
    if not os.path.exists(CHECK):
        pytest.fail("tests/cpp/abi_layout is not built; run __graft_entry__.build()")
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok: 0 failure(s)" in r.stdout


def test_exports():
    if not os.path.exists(TRACER_LIB):
        pytest.fail("libmray_tracer.so is not built; run __graft_entry__.build()")
    r = subprocess.run(["nm", "-D", "--defined-only", TRACER_LIB], capture_output=True,
                       text=True, timeout=60, check=True)
    names = [ln.split()[-1] for ln in r.stdout.splitlines() if ln.strip()]
    for sym in ("ConstructTracer", "DestroyTracer", "mray_tracer_flatten_to_file"):
        assert sym in names, sym
    ours = [n for n in names if "mray_amd" in n or "MRayError" in n or "StaticVector" in n
            or "TransientData" in n or n.startswith("_ZN6Vector")]
    # only the scene loader entry point among the library's own C++ symbols
    assert all("LoadSceneMRay" in n for n in ours if not n.startswith("_ZNSt")), ours
