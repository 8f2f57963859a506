"""backproject_gpu_noalloc (the back projection entry point that allocates nothing: its workspace is a fixed
device array, the kernel tables come from an earlier prepare_gpu()) against backproject_gpu, bit for bit,
through the C ABI, on two calls in a row, the second with a smaller input; oversized and unprepared
geometries are refused with an error instead of running."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from petl_amd.petl_server import PETLserver, load_library  # noqa: E402
from petl_amd.workloads import setup_server  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
_fp = ctypes.POINTER(ctypes.c_float)


def abi():
    L = load_library()
    L.backproject_gpu_noalloc.restype = ctypes.c_bool
    L.backproject_gpu_noalloc.argtypes = [ctypes.POINTER(_fp), _fp, ctypes.c_void_p, ctypes.c_bool]
    L.prepare_gpu.restype = ctypes.c_bool
    L.prepare_gpu.argtypes = []
    return L


def noalloc(L, g, f):
    ptrs = (_fp * len(g))(*[ctypes.cast(ctypes.c_void_p(x.data_ptr()), _fp) for x in g])
    return L.backproject_gpu_noalloc(ptrs, ctypes.cast(ctypes.c_void_p(f.data_ptr()), _fp), None, True)


def server(name):
    s = PETLserver()
    s.print_warnings = False
    setup_server(s, name)
    shapes = [tuple(int(x) for x in s.get_planogramSize(i)) for i in range(s.get_numPlanograms())]
    return s, shapes, (s.numZ(), s.numY(), s.numX())


@pytest.mark.parametrize("mode", ["petl", "matched"])
def test_noalloc_equals_backproject_gpu_two_calls(mode):
    L = abi()
    gen = torch.Generator(device=DEV).manual_seed(5)
    for name in ("panel64", "panel32"):  # two planograms of 64 x 64 crystals, then one of 32 x 32: smaller
        s, shapes, vs = server(name)
        s.set_backprojector_mode(mode)
        s.set_model()
        assert L.prepare_gpu()
        g = [torch.rand(sh, generator=gen, device=DEV) for sh in shapes]
        want = torch.empty(vs, device=DEV)
        s.backproject_gpu(g, want)
        got = torch.full(vs, 7.0, device=DEV)  # the entry point overwrites its output
        s.set_model()
        assert noalloc(L, g, got)
        assert torch.equal(got, want), name


def test_noalloc_refuses_what_it_cannot_run():
    L = abi()
    s, shapes, vs = server("panel64")
    s.set_model()
    g = [torch.zeros(sh, device=DEV) for sh in shapes]
    f = torch.zeros(vs, device=DEV)
    assert not noalloc(L, g, f)        # kernel tables not built yet
    s, shapes, vs = server("panel128")
    s.set_model()
    assert L.prepare_gpu()
    g = [torch.zeros(sh, device=DEV) for sh in shapes]
    f = torch.zeros(vs, device=DEV)
    assert not noalloc(L, g, f)        # an 847 MB workspace does not fit the 128 MiB array
