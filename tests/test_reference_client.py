"""The drop-in boundary as the reference's own Python client sees it.

tests/golden/reference_abi.json holds, as data, the reference's C declarations (src/petl_c_interface.h)
and every ctypes restype/argtypes its client declares (src/petl_server.py); make_reference_abi.py beside
it writes the file from the reference tree.  The reference client itself is not imported (it needs
imageio, which this image lacks), so these tests bind libpetl.so exactly as that client binds it -- the
same entry points, argument types and host-pointer conventions -- and drive it with the client's call
sequences:

* CPU: include/petl_c_interface.h declares each of the reference's entry points with the reference's
  types; every binding the client declares agrees with those types; every entry point the client calls
  is exported; the parameter calls (models, planograms, volume, getters, the host elementwise kernels)
  behave through those bindings as the reference's do.
* GPU: PETLserver.MLEM's host loop (src/petl_server.py:1534-1606: sensitivity by backproject of ones,
  project / rdivide / backproject per iteration, the relative-differences gradient, divide and multiply)
  issued call by call through those bindings on host arrays matches the fp32 oracle to the MLEM bar
  (1e-4 relative L2 after 5 iterations) and this build's device-resident MLEM to 1e-5.
"""
import ctypes
import json
import os
import re

import numpy as np
import pytest
from numpy.ctypeslib import ndpointer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ABI = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_abi.json")))
HEADER = os.path.join(ROOT, "include", "petl_c_interface.h")

CTYPE = {
    "bool": lambda: ctypes.c_bool, "int": lambda: ctypes.c_int, "float": lambda: ctypes.c_float,
    "char*": lambda: ctypes.c_char_p, "void*": lambda: ctypes.c_void_p,
    "float*": lambda: ctypes.POINTER(ctypes.c_float),
    "float[]": lambda: ndpointer(ctypes.c_float, flags="C_CONTIGUOUS"),
    "int[]": lambda: ndpointer(ctypes.c_int, flags="C_CONTIGUOUS"),
    "float**": lambda: ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
}


def our_declarations():
    text = re.sub(r"/\*.*?\*/", " ", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"PETL_EXTERN\s+([\w\s\*]+?)\s*\b(\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, args = m.group(1), m.group(2), " ".join(m.group(3).split())

        def norm(d):
            d = d.replace("const ", " ").strip()
            return d.replace("*", " ").split()[0] + "*" * d.count("*")
        argl = [] if args in ("", "void") else [norm(a) for a in args.split(",")]
        out[name] = {"ret": norm(ret), "args": argl}
    return out


def compatible(binding, decl):
    """A ctypes argument type and a C parameter type agree: an array 'T[]' passes as 'T*', c_void_p as any
    pointer."""
    return (binding == decl or binding.replace("[]", "*") == decl
            or (binding == "void*" and decl.endswith("*")))


def test_fixture_covers_the_reference_interface():
    assert len(ABI["c_declarations"]) == 64
    assert len(ABI["ctypes_bindings"]) >= 55
    for name in ("project", "backproject", "relativeDifferences_gradient", "rdivide4D", "set_volume"):
        assert name in ABI["ctypes_bindings"] and name in ABI["calls"]


def test_header_declares_the_reference_signatures():
    ours = our_declarations()
    for name, ref in ABI["c_declarations"].items():
        assert name in ours, name
        assert ours[name]["ret"] == ref["ret"], (name, ours[name], ref)
        assert ours[name]["args"] == ref["args"], (name, ours[name], ref)


def test_client_bindings_agree_with_the_declarations():
    decl = our_declarations()
    for name, b in ABI["ctypes_bindings"].items():
        d = decl[name]
        if b["restype"] is not None:
            assert b["restype"] == d["ret"], name
        for argl in b["argtypes"]:
            assert len(argl) == len(d["args"]), (name, argl, d["args"])
            assert all(compatible(x, y) for x, y in zip(argl, d["args"])), (name, argl, d["args"])


def test_every_called_entry_point_is_exported(petl_lib):
    for name in ABI["calls"]:
        assert hasattr(petl_lib, name), name


class RefClient:
    """libpetl.so bound with the reference client's own ctypes declarations (the host-pointer form of
    each), calling set_model before each call as src/petl_server.py does."""

    def __init__(self):
        from petl_amd.petl_server import library_path, load_library
        load_library()
        # a CDLL object of its own: the same loaded library, function objects bound as the client binds
        self.lib = ctypes.CDLL(library_path())
        for name, b in ABI["ctypes_bindings"].items():
            fn = getattr(self.lib, name)
            if b["restype"] is not None:
                fn.restype = CTYPE[b["restype"]]()
            if b["argtypes"]:
                host = [a for a in b["argtypes"] if "void*" not in a] or b["argtypes"]
                fn.argtypes = [CTYPE[t]() for t in host[-1]]
        self.id = self.lib.create_new_model()

    def __getattr__(self, name):
        fn = getattr(self.lib, name)

        def call(*args):
            assert self.lib.set_model(self.id)
            return fn(*args)
        return call

    @staticmethod
    def floatstarstar(g):
        p = ctypes.POINTER(ctypes.c_float)
        return (p * len(g))(*[x.ctypes.data_as(p) for x in g])

    def planograms(self):
        out = []
        for n in range(self.get_numPlanograms()):
            shape = np.zeros(4, dtype=np.int32)
            assert self.get_planogramSize(n, shape)
            out.append(np.zeros(tuple(shape), np.float32))
        return out

    def volume(self):
        return np.zeros((self.get_numZ(), self.get_numY(), self.get_numX()), np.float32)


def test_parameter_calls_through_the_reference_bindings():
    c = RefClient()
    assert not c.geometry_defined() and not c.volume_defined()
    # add_planogram(psi, D, L, H, v_m0, v_m1, T): the client's argument order, its default slopes
    assert c.add_planogram(0.0, 200.0, 256.0, 256.0, 1.0, 1.0, 2.0)
    assert c.add_planogram(90.0, 200.0, 256.0, 256.0, 1.0, 1.0, 2.0)
    assert not c.add_planogram(0.0, -1.0, 10.0, 10.0, 1.0, 1.0, 2.0)
    assert c.get_numPlanograms() == 2
    shape = np.zeros(4, dtype=np.int32)
    assert c.get_planogramSize(1, shape) and list(shape) == [101, 101, 128, 228]
    assert c.set_volume(128, 128, 96, 2.0, 2.5, 0.0, 0.0, 0.0)
    assert (c.get_numX(), c.get_numY(), c.get_numZ()) == (128, 128, 96)
    assert (c.get_voxelWidth(), c.get_voxelHeight()) == (2.0, 2.5)
    assert not c.set_offsetX(-1.0) and c.get_offsetX() == 0.0
    assert c.set_offsetZ(1.5) and c.get_offsetZ() == 1.5
    assert c.set_numZ(64) and c.get_numZ() == 64
    assert c.keep_only_planogram(1) and c.get_numPlanograms() == 1
    assert c.remove_planogram(0) and not c.geometry_defined()
    # a second client's model is its own
    d = RefClient()
    assert d.get_numPlanograms() == 0 and c.volume_defined()


def test_host_elementwise_calls_through_the_reference_bindings():
    c = RefClient()
    rng = np.random.default_rng(3)
    x = rng.random((3, 4, 5), dtype=np.float32) + 0.5
    y = rng.random((3, 4, 5), dtype=np.float32) + 0.5
    y[0, 0, 0] = 0.0
    want = x * y
    a = x.copy()
    assert c.multiply3D(a, y, 3, 4, 5)
    np.testing.assert_array_equal(a, want)
    a = x.copy()
    assert c.divide3D(a, y, 3, 4, 5)
    np.testing.assert_array_equal(a, np.where(y == 0, 0.0, x / np.where(y == 0, 1, y)).astype(np.float32))
    g = rng.random((2, 3, 4, 5), dtype=np.float32) + 0.5
    p = rng.random((2, 3, 4, 5), dtype=np.float32) + 0.5
    den = p.copy()
    assert c.rdivide4D(g, den, 2, 3, 4, 5)  # denom := num / denom
    np.testing.assert_array_equal(den, g / p)
    r = x.copy()
    r[1, 2, 3] = 0.0
    assert c.reciprocal3D(r, 3, 4, 5, 7.0)
    assert r[1, 2, 3] == 7.0 and r[0, 0, 1] == np.float32(1.0) / x[0, 0, 1]
    ip = c.inner_product3D(x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                           y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None, 3, 4, 5)
    assert abs(ip - float(np.dot(x.ravel().astype(np.float64), y.ravel()))) <= 1e-5 * abs(ip)


def client_mlem(c, g, num_iter, delta, beta):
    """src/petl_server.py:1534-1606 with f = None, no mask, no response, as C calls on host arrays."""
    f = c.volume()
    f[:] = 1.0
    ones = [np.ones_like(x) for x in g]
    sens = c.volume()
    assert c.backproject(c.floatstarstar(ones), sens)
    sens[sens == 0.0] = 1.0
    d = c.volume()
    Pd = c.planograms()
    nz, ny, nx = f.shape
    for _ in range(num_iter):
        assert c.project(c.floatstarstar(Pd), f)
        for n in range(len(g)):
            assert c.rdivide4D(g[n], Pd[n], *g[n].shape)
        assert c.backproject(c.floatstarstar(Pd), d)
        if beta <= 0.0:
            assert c.divide3D(d, sens, nz, ny, nx)
        else:
            tv = np.zeros_like(f)
            assert c.relativeDifferences_gradient(f, tv, nz, ny, nx, delta, beta)
            tv[:] += sens[:]
            np.maximum(0.1 * sens, tv, out=tv)
            assert c.divide3D(d, tv, nz, ny, nx)
        assert c.multiply3D(f, d, nz, ny, nx)
    return f


@pytest.mark.gpu
@pytest.mark.parametrize("beta", [0.0, 0.01])
def test_reference_client_mlem_loop(oracle, beta):
    from petl_amd.petl_server import PETLserver
    from petl_amd.workloads import ellipsoid_phantom
    psis, delta, iters = (0.0, 90.0), 2.0, 5
    c = RefClient()
    for psi in psis:
        assert c.add_planogram(psi, 100.0, 64.0, 64.0, 1.0, 1.0, 2.0)
    assert c.set_volume(32, 32, 32, 2.0, 2.0, 0.0, 0.0, 0.0)
    plans = [oracle.planogram(psi, 100.0, 64.0, 64.0, 2.0, 1.0, 1.0) for psi in psis]
    vol = oracle.volume(32, 32, 32, 2.0)
    assert [p.shape for p in plans] == [x.shape for x in c.planograms()]
    g = oracle.project(plans, vol, ellipsoid_phantom(32, 2.0, 6, seed=1))
    rng = np.random.default_rng(101)
    tot = sum(float(x.sum()) for x in g)
    y = [rng.poisson(x * (1e6 / tot)).astype(np.float32) for x in g]

    got = client_mlem(c, y, iters, delta, beta)
    ref = oracle.mlem(plans, vol, y, iters, delta=delta, beta=beta)
    rel = float(np.linalg.norm(got.astype(np.float64) - ref) / np.linalg.norm(ref.astype(np.float64)))
    print(f"reference client loop vs oracle: {rel:.2e}")
    assert rel <= 1e-4

    s = PETLserver()
    s.print_warnings = False
    for psi in psis:
        s.add_planogram(psi, 100.0, 64.0, 64.0, 2.0, 1.0, 1.0)
    s.set_volume(32, 32, 32, 2.0)
    dev = s.MLEM(y, None, iters, delta=delta, beta=beta)
    rel = float(np.linalg.norm(got.astype(np.float64) - dev) / np.linalg.norm(dev.astype(np.float64)))
    print(f"reference client loop vs device-resident MLEM: {rel:.2e}")
    assert rel <= 1e-5
