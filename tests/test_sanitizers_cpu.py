"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only; no kernel runs).

* The oracle (oracle/petl_oracle.cpp, built by `make -C oracle san` with g++ -fsanitize=address,undefined)
  runs its projectors in both back projector modes and both precisions, a ragged oblique geometry, the
  texture-fetch form, the priors, list-mode binning and an MLEM and an OSEM run, in a subprocess with
  gcc's libasan preloaded; any report aborts it.
* The C-ABI library's host code (petl_amd/build.py --sanitize-host: the host compilation with
  -fsanitize=address,undefined, device code as usual) runs the parameter API of the reference interface
  (planograms, volumes, models, set_default_volume) in a subprocess with clang's ASan runtime preloaded.

The sanitized builds are made on first use (the library takes a few minutes).
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SAN = os.path.join(ROOT, "oracle", "_san", "libpetl_oracle_san.so")
LIB_SAN = os.path.join(ROOT, "petl_amd", "_san", "libpetl_san.so")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", PYTHONPATH=ROOT, OMP_NUM_THREADS="2")

ORACLE_RUN = r"""
import numpy as np
from oracle import oracle as O
from petl_amd.workloads import ellipsoid_phantom, poisson_data
P = [O.planogram(0.0, 60.0, 32.0, 32.0, 2.0), O.planogram(90.0, 60.0, 32.0, 32.0, 2.0),
     O.planogram(30.0, 72.0, 30.0, 22.0, 2.0)]
V = O.volume(13, 11, 7, 2.0, 2.5, 1.0, -2.0, 3.0)
f = np.random.default_rng(0).random(V.shape).astype(np.float32)
for dt in (np.float32, np.float64):
    g = O.project(P, V, f, dtype=dt)
    for mode in ("petl", "matched"):
        O.backproject(P, V, g, mode=mode, dtype=dt)
O.backproject_petl_tex(P, V, O.project(P, V, f), "nearest")
O.rd_gradient(f, 2.0, 0.1); O.tv_gradient(f, 0.1, 0.1, 1.2); O.tv_cost(f, 0.1, 0.1); O.tv_quadform(f, f, 0.1, 0.1)
O.gaussian_filter(f.astype(np.float64), 3.0)
O.count_taps(P[0], V)
small = [O.planogram(0.0, 100.0, 64.0, 64.0, 2.0)]
Vs = O.volume(16, 16, 16, 4.0)
y = poisson_data(O.project(small, Vs, ellipsoid_phantom(16, 4.0, 3, seed=1)), 1e5, seed=2)
O.mlem(small, Vs, y, 2, beta=0.01)
O.osem(small, Vs, y, 1, 2)
rng = np.random.default_rng(3)
blue = np.c_[rng.uniform(-30, 30, 500), np.full(500, 50.0), rng.uniform(0, 60, 500)].astype(np.float32)
pink = np.c_[rng.uniform(-30, 30, 500), np.full(500, -50.0), rng.uniform(0, 60, 500)].astype(np.float32)
O.bin_events(small, blue, pink)
print("oracle clean")
"""

LIB_RUN = r"""
from petl_amd.petl_server import PETLserver
s = PETLserver()
s.print_warnings = False
s.add_planogram(0.0, 200.0, 256.0, 256.0, 2.0)
s.add_planogram(90.0, 101.0, 251.52, 156.72, 1.6)
s.add_planogram(30.0, 180.0, 192.0, 144.0, 2.0, 0.27, 0.27)
s.set_default_volume()
assert [tuple(int(x) for x in s.get_planogramSize(i)) for i in range(3)][0] == (101, 101, 128, 228)
s.set_volume(33, 17, 9, 2.0, 2.5, 1.0, 2.0, 3.0)
s.offsetX(-1.0); s.offsetY(2.0)
s.keep_only_planogram(2)
s.remove_planogram(0)
t = PETLserver()
t.add_planogram(0.0, 60.0, 32.0, 32.0, 2.0)
t.set_volume(8, 8, 8, 2.0)
assert t.libprojectors.copy_parameters(s.param_id)
print("library host code clean")
"""


def run(code, preload, env_extra):
    env = dict(ENV, LD_PRELOAD=preload, **env_extra)
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=ROOT,
                          timeout=900)


def test_oracle_under_sanitizers():
    if not os.path.exists(ORACLE_SAN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    libasan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    r = run(ORACLE_RUN, libasan, {"PETL_ORACLE_LIBRARY": ORACLE_SAN})
    assert r.returncode == 0 and "oracle clean" in r.stdout, r.stderr[-4000:]


def test_library_host_code_under_sanitizers():
    srcs = glob.glob(os.path.join(ROOT, "petl_amd", "csrc", "*"))
    if not os.path.exists(LIB_SAN) or os.path.getmtime(LIB_SAN) < max(os.path.getmtime(p) for p in srcs):
        subprocess.run([sys.executable, "-m", "petl_amd.build", "--sanitize-host"], check=True, cwd=ROOT)
    rt = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not rt:
        pytest.skip("clang's ASan runtime is not installed")
    r = run(LIB_RUN, rt[0], {"PETL_LIBRARY": LIB_SAN})
    assert r.returncode == 0 and "library host code clean" in r.stdout, r.stderr[-4000:]
