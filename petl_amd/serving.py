"""Python interface of the reconstruction service (petl_service_* in libpetl.so).

    svc = ReconService(hbm_budget_bytes=64e9, max_slots=8)
    rid = svc.submit(psis=(0, 90), D=200, L=128, H=128, T=2, n=64, voxel=2.0, y=data, num_iter=10)
    image, times = svc.wait(rid)      # times = (queued s, on device s, total s)

A request is one MLEM/OSEM reconstruction of its own geometry and volume.  Requests are admitted while
others run, as soon as a slot is free and the request's device footprint fits the budget; the image is
the same as PETLserver.MLEM gives for the same inputs (ones as the initial estimate, no mask/response).
"""
import ctypes
import threading

import numpy as np

from .petl_server import load_library

MAX_PLANOGRAMS = 8
_c_float_p = ctypes.POINTER(ctypes.c_float)


class ReconRequest(ctypes.Structure):
    _fields_ = [("num_planograms", ctypes.c_int),
                ("psi", ctypes.c_float * MAX_PLANOGRAMS),
                ("D", ctypes.c_float), ("L", ctypes.c_float), ("H", ctypes.c_float), ("T", ctypes.c_float),
                ("v_m0", ctypes.c_float), ("v_m1", ctypes.c_float),
                ("numX", ctypes.c_int), ("numY", ctypes.c_int), ("numZ", ctypes.c_int),
                ("voxelWidth", ctypes.c_float), ("voxelHeight", ctypes.c_float),
                ("y", ctypes.POINTER(_c_float_p)),
                ("numIter", ctypes.c_int), ("numSubsets", ctypes.c_int), ("prior", ctypes.c_int),
                ("delta", ctypes.c_float), ("beta", ctypes.c_float), ("p", ctypes.c_float),
                ("storage", ctypes.c_int), ("image_storage", ctypes.c_int)]


_PRIORS = {None: 0, "none": 0, "rd": 1, "tv": 2}


def _declare(lib):
    lib.petl_service_create.restype = ctypes.c_void_p
    lib.petl_service_create.argtypes = [ctypes.c_double, ctypes.c_int]
    lib.petl_service_create_ex.restype = ctypes.c_void_p
    lib.petl_service_create_ex.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int]
    lib.petl_service_submit.restype = ctypes.c_int
    lib.petl_service_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(ReconRequest)]
    lib.petl_service_request_bytes.restype = ctypes.c_double
    lib.petl_service_request_bytes.argtypes = [ctypes.POINTER(ReconRequest)]
    lib.petl_service_wait.restype = ctypes.c_bool
    lib.petl_service_wait.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_float_p, ctypes.POINTER(ctypes.c_double)]
    lib.petl_service_stats.restype = ctypes.c_bool
    lib.petl_service_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    lib.petl_service_request_footprint.restype = ctypes.c_bool
    lib.petl_service_request_footprint.argtypes = [ctypes.POINTER(ReconRequest), ctypes.POINTER(ctypes.c_double)]
    lib.petl_service_destroy.restype = None
    lib.petl_service_destroy.argtypes = [ctypes.c_void_p]
    return lib


_STORAGE = {"f32": 0, "bf16": 1, "f16": 2}


def make_request(psis, D, L, H, T, n, voxel, y, num_iter, v_m0=None, v_m1=None, num_subsets=1, prior=None,
                 delta=2.0, beta=0.0, p=1.0, shape=None, storage="f32", image_storage=None):
    """Builds the C request and the array keeping the data pointers alive.  Defaults of v_m0/v_m1 follow
    PETLserver.add_planogram (src/petl_server.py:484-507).  storage / image_storage: device formats of the
    planograms and the image, as PETLserver.MLEM takes them."""
    psis = tuple(float(x) for x in psis)
    if not 1 <= len(psis) <= MAX_PLANOGRAMS:
        raise ValueError("1 to 8 planograms per request")
    if v_m0 is None:
        v_m0 = 1.0
    if v_m1 is None:
        v_m1 = min(1.0, (H - 1.5 * T) / D)
    nz, ny, nx = shape if shape is not None else (n, n, n)
    if len(y) != len(psis):
        raise ValueError("one data array per planogram")
    ys = [np.ascontiguousarray(a, dtype=np.float32) for a in y]
    ptrs = (_c_float_p * len(ys))(*[a.ctypes.data_as(_c_float_p) for a in ys])
    r = ReconRequest()
    r.num_planograms = len(psis)
    for i, v in enumerate(psis):
        r.psi[i] = v
    r.D, r.L, r.H, r.T = D, L, H, T
    r.v_m0, r.v_m1 = v_m0, v_m1
    r.numX, r.numY, r.numZ = nx, ny, nz
    r.voxelWidth, r.voxelHeight = voxel, voxel
    r.y = ptrs
    r.numIter, r.numSubsets = int(num_iter), int(num_subsets)
    r.prior = _PRIORS[prior]
    r.delta, r.beta, r.p = delta, beta, p
    r.storage = _STORAGE[storage]
    r.image_storage = _STORAGE[image_storage if image_storage is not None else storage]
    return r, (ys, ptrs)


class ReconService:
    def __init__(self, hbm_budget_bytes=0.0, max_slots=8, batched=False):
        """batched: requests that allow it (fp32 MLEM on the fast kernels, see petl_service_create_ex) run in
        batched iterations, one launch per phase for all of them."""
        self.lib = _declare(load_library())
        self.h = self.lib.petl_service_create_ex(float(hbm_budget_bytes), int(max_slots), 1 if batched else 0)
        if not self.h:
            raise RuntimeError("petl_service_create failed")
        self._keep = {}
        self._shapes = {}
        self._lock = threading.Lock()

    def request_bytes(self, **kw):
        r, _ = make_request(**kw)
        return self.lib.petl_service_request_bytes(ctypes.byref(r))

    def request_footprint(self, **kw):
        """(bytes the admission counts, bytes the request's buffers hold once allocated): allocates the
        request's buffers on the device and frees them again."""
        r, _ = make_request(**kw)
        out = (ctypes.c_double * 2)()
        if not self.lib.petl_service_request_footprint(ctypes.byref(r), out):
            raise RuntimeError("petl_service_request_footprint failed")
        return out[0], out[1]

    def submit(self, **kw):
        r, keep = make_request(**kw)
        rid = self.lib.petl_service_submit(self.h, ctypes.byref(r))
        if rid < 0:
            raise ValueError("request rejected (invalid geometry or larger than the HBM budget)")
        with self._lock:
            self._keep[rid] = keep
            self._shapes[rid] = (r.numZ, r.numY, r.numX)
        return rid

    def wait(self, rid):
        with self._lock:
            shape = self._shapes[rid]
        f = np.empty(shape, np.float32)
        t = (ctypes.c_double * 3)()
        ok = self.lib.petl_service_wait(self.h, int(rid), f.ctypes.data_as(_c_float_p), t)
        with self._lock:
            self._keep.pop(rid, None)
            self._shapes.pop(rid, None)
        if not ok:
            raise RuntimeError(f"request {rid} failed")
        return f, (t[0], t[1], t[2])

    def stats(self):
        s = (ctypes.c_double * 15)()
        self.lib.petl_service_stats(self.h, s)
        return dict(completed=int(s[0]), failed=int(s[1]), peak_active=int(s[2]), peak_reserved_bytes=s[3],
                    budget_bytes=s[4], queued=int(s[5]), admit_alloc_s=s[6], admit_setup_s=s[7],
                    admit_capture_s=s[8], release_s=s[9], pool_hits=int(s[10]), recaptures=int(s[11]),
                    batched_requests=int(s[12]), batched_rounds=int(s[13]),
                    batched_peak=int(s[14]))

    def close(self):
        if self.h:
            self.lib.petl_service_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
