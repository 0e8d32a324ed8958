"""The graph-store reader under AddressSanitizer / UndefinedBehaviorSanitizer
(host code only, no GPU): tests/cpp/fuzz_store.cpp opens 3000 seeded,
randomly damaged copies of a store; every open must succeed or fail with a
status, and the sanitizers fail the run on any memory error or leak.

Two stores: the golden one, and an R-MAT scale-10 store written by the
library whose data files span many pages -- the reader maps the files, and
a read past the end of a truncated one-page file stays inside the mapped
page (zeros, no fault), so only multi-page files make such a read fault."""
import shutil
import subprocess
from pathlib import Path

import pytest

import vertexsurge_amd as vs
from conftest import GOLDEN

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "vertexsurge_amd" / "csrc"


@pytest.fixture(scope="module")
def fuzz_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path_factory.mktemp("fuzz") / "fuzz_store"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-fopenmp",
           f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "fuzz_store.cpp"),
           str(CSRC / "vs_host.cpp"), str(CSRC / "vs_store.cpp"), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    return exe


def _store_dir(which, tmp_path):
    if which == "golden":
        return GOLDEN / "store" / "knows"
    g = vs.HostCSR.rmat(10, 16, seed=3)
    d = tmp_path / "rmat10"
    d.mkdir()
    vs.store_write(d, g.offsets, g.adj, "Person", "knows")
    assert (d / "Person-knows-Person_csr_src_dst").stat().st_size > 16 * 4096
    return d


@pytest.mark.parametrize("which", ["golden", "rmat10"])
def test_store_reader_fuzz_under_sanitizers(fuzz_exe, tmp_path, which):
    exe = fuzz_exe
    iters = 3000
    r = subprocess.run([str(exe), str(_store_dir(which, tmp_path)), str(tmp_path / "scratch"),
                        str(iters)], capture_output=True, text=True, timeout=300,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
                            "UBSAN_OPTIONS": "print_stacktrace=1", "OMP_NUM_THREADS": "2",
                            "PATH": "/usr/bin:/bin"})
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    ok, rejected = (int(x) for x in r.stdout.split()[1::2])
    assert ok + rejected == iters and ok > 0 and rejected > iters // 2


def test_host_generators_fuzz_under_sanitizers(tmp_path):
    """tests/cpp/fuzz_host.cpp: random edge lists (some ids out of range) into
    vs_csr_from_edges, random and damaged CSRs into vs_sample_paths3, source
    samples of every size, and built CSRs written as stores with flat edge
    files, under ASan + UBSan; built CSRs must be ascending sets and sampled
    paths must follow them."""
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path / "fuzz_host"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-fopenmp",
           f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "fuzz_host.cpp"),
           str(CSRC / "vs_host.cpp"), str(CSRC / "vs_store.cpp"), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    iters = 1500
    scratch = tmp_path / "stores"
    scratch.mkdir()
    r = subprocess.run([str(exe), str(iters), str(scratch)], capture_output=True, text=True,
                       timeout=300,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
                            "UBSAN_OPTIONS": "print_stacktrace=1", "OMP_NUM_THREADS": "2",
                            "PATH": "/usr/bin:/bin"})
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    built, rejected = (int(x) for x in r.stdout.split()[1::2])
    assert built + rejected == iters and built > iters // 2 and rejected > 0
