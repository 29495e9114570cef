"""CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

  * The FP32 oracle (oracle/liboracle_f32san.so, `make -C oracle
    sanitize`) runs networks of every kind the parity tests use (AdEx
    Brunel with Poisson drive, the legacy Poisson generator, the c2-shaped
    iaf_psc_exp network, a plastic network with dendritic delay, STDP
    reference cases, the ensemble of spike-generator-driven AdEx networks)
    in a child process with the sanitizer runtimes preloaded; any invalid
    access or undefined operation aborts it.  Its spike trains must equal
    those of the plain build.
  * The per-connection distributions (neurongpu_amd/csrc/conn_dist.h, host
    side) are sampled by tests/cxx/dist_check.cpp built with both
    sanitizers.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name],
                       capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


ASAN = _runtime("libasan.so")
UBSAN = _runtime("libubsan.so")
needs_san = pytest.mark.skipif(
    shutil.which("gcc") is None or not ASAN or not UBSAN,
    reason="gcc sanitizer runtimes not installed")

CHILD = r'''
import sys
sys.path.insert(0, %(tests)r)
sys.path.insert(0, %(root)r)
import numpy as np
import netbuild as nb
from neurongpu_amd import workloads as wl

def spikes(spec, steps, prec):
    net, _ = nb.build_oracle(spec, prec)
    net.run(steps, 4)
    st, nd, ct = net.spikes()
    return sorted(zip(st.tolist(), nd.tolist(), ct.tolist()))

specs = [
    ("brunel", nb.brunel_spec(order=200, w_ex=0.0625, w_in=0.375, w_p=0.375,
                              tol=1e-3), 300),
    ("legacy_poisson", nb.legacy_brunel_spec(order=200), 300),
    ("c2_small", wl.c2_spec(scale=0.02, integer_weights=True), 300),
    ("stdp_net", nb.stdp_net_spec(n_src=60, n_post=30, k_plastic=10), 600),
    ("ensemble", nb.ensemble_spec([nb.ensemble_instance_spec(80, 3),
                                   nb.ensemble_instance_spec(120, 4)]), 500),
]
cases = nb.load_stdp_cases()
for name in ("1", "2", "3"):
    specs.append(("stdp_case_" + name,
                  nb.stdp_case_spec(cases["cases"][name], cases["common"]),
                  int(round(cases["common"]["sim_time"] / 0.1))))
for name, spec, steps in specs:
    a = spikes(spec, steps, "f32san")
    b = spikes(spec, steps, "f32")
    assert a == b, name
    print(name, "spikes", len(a), flush=True)
# the plastic weights too
spec = nb.stdp_net_spec(n_src=60, n_post=30, k_plastic=10)
ws = []
for prec in ("f32san", "f32"):
    net, _ = nb.build_oracle(spec, prec)
    net.run(600, 4)
    ws.append(net.connections()[4])
assert np.array_equal(ws[0], ws[1])
print("OK")
'''


@needs_san
def test_oracle_under_sanitizers():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"),
                           "sanitize"])
    env = dict(os.environ)
    env["LD_PRELOAD"] = ASAN + ":" + UBSAN
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "4"
    code = CHILD % {"tests": HERE, "root": ROOT}
    r = subprocess.run([sys.executable, "-c", code], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "OK" in r.stdout, (
        r.stdout[-3000:] + r.stderr[-6000:])


@needs_san
def test_distributions_under_sanitizers(tmp_path):
    exe = str(tmp_path / "dist_check")
    subprocess.check_call([
        "g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
        "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
        "-fno-omit-frame-pointer",
        "-I" + os.path.join(ROOT, "neurongpu_amd", "csrc"),
        os.path.join(HERE, "cxx", "dist_check.cpp"), "-o", exe])
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1"
    r = subprocess.run([exe], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
