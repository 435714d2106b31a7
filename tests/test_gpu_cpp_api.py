"""The C++ programs written against the header API (grid-calculus_amd/cpp/tests/*.cpp, built
in-tree by __graft_entry__.build() into grid-calculus_amd/cpp/bin/), run on the GPU: each
case prints [ OK ] or [ FAIL ] and the program exits non-zero on any failure."""
import glob
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "grid-calculus_amd", "cpp")
PROGRAMS = sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(CPP, "tests", "*.cpp")))


@pytest.mark.parametrize("name", PROGRAMS)
def test_header_api_program(name):
    exe = os.path.join(CPP, "bin", name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (run __graft_entry__.build())")
    print("running:", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    fails = [ln for ln in r.stdout.splitlines() if ln.startswith("[ FAIL")]
    assert r.returncode == 0 and not fails, r.stdout[-3000:] + r.stderr[-2000:]
    assert "0 failures" in r.stdout
