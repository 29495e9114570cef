"""C and C++ clients compiled against include/ and linked with the
library, as a user of the reference's C ABI or C++ class would: the
reference's c++/examples/brunel_net.cpp network is built through the C++
API, and a plain C program exercises the C ABI and its error convention."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = os.path.join(ROOT, "tests", "cxx")
LIBDIR = os.path.join(ROOT, "neurongpu_amd")


def _build(tmp_path, src, compiler, std):
    if shutil.which(compiler) is None:
        pytest.skip(compiler + " not available")
    exe = str(tmp_path / os.path.splitext(src)[0])
    cmd = [compiler, std, "-I", os.path.join(ROOT, "include"),
           os.path.join(CXX, src), "-L", LIBDIR, "-lneurongpu",
           "-Wl,-rpath," + LIBDIR, "-o", exe]
    subprocess.check_call(cmd)
    return exe


def _run(exe, *args, timeout=120):
    r = subprocess.run([exe] + list(args), capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return dict(l.split(" ", 1) for l in r.stdout.strip().split("\n"))


def test_cxx_brunel_build(tmp_path):
    exe = _build(tmp_path, "brunel_build.cpp", "g++", "-std=c++17")
    out = _run(exe, "1000")
    assert int(out["exc"]) == 1000 * 800
    assert int(out["inh"]) == 1000 * 200
    assert int(out["pg"]) == 1000


def test_c_abi_client(tmp_path):
    exe = _build(tmp_path, "c_api.c", "gcc", "-std=c11")
    out = _run(exe)
    assert out["create"] == "0 err 0"
    assert out["bad"].startswith("0 err 2 msg ")
    assert out["connect"] == "err 0"
    assert out["conns"] == "100 ok"


@pytest.mark.gpu
def test_cxx_brunel_simulate(tmp_path):
    """The same C++ program simulates 100 ms on the GPU."""
    exe = _build(tmp_path, "brunel_build.cpp", "g++", "-std=c++17")
    out = _run(exe, "5000", "sim")
    print("spikes", out["spikes"], flush=True)
    assert int(out["spikes"]) > 0
