"""PETLserver: ctypes bindings of the MI355X PETL library with the reference's Python API.

This mirrors the part of the reference's ``PETLserver`` class (src/petl_server.py) that sits on the
projector / MLEM path -- geometry, volume, allocation, project, backproject, sensitivity, elementwise
helpers, MLEM and the TV / relative-difference prior gradients -- with the same method names, argument
meanings, defaults and error behaviour (``TypeError('project failed')`` etc.), so PETL scripts that use
that path run unchanged.  The compute runs in libpetl.so (HIP kernels for gfx950); the library is
required and loading fails loudly without it.

Additions beyond the reference: ``set_backprojector_mode``, ``project_gpu`` / ``backproject_gpu`` on
torch tensors, a device-resident ``MLEM`` (with optional OSEM subsets and log-likelihood trace) and a
device-resident ``RWLS`` (and LS / WLS / RLS).
"""
import ctypes
import glob
import os

import numpy as np
from numpy.ctypeslib import ndpointer

_HERE = os.path.dirname(os.path.abspath(__file__))
_c_float_p = ctypes.POINTER(ctypes.c_float)
_f32 = ndpointer(ctypes.c_float, flags="C_CONTIGUOUS")

_SIGNATURES = {
    "set_model": (ctypes.c_bool, [ctypes.c_int]),
    "create_new_model": (ctypes.c_int, []),
    "about": (None, []),
    "copy_parameters": (ctypes.c_bool, [ctypes.c_int]),
    "clearAll": (None, []),
    "print_parameters": (None, []),
    "volume_defined": (ctypes.c_bool, []),
    "geometry_defined": (ctypes.c_bool, []),
    "get_numX": (ctypes.c_int, []), "set_numX": (ctypes.c_bool, [ctypes.c_int]),
    "get_numY": (ctypes.c_int, []), "set_numY": (ctypes.c_bool, [ctypes.c_int]),
    "get_numZ": (ctypes.c_int, []), "set_numZ": (ctypes.c_bool, [ctypes.c_int]),
    "get_voxelWidth": (ctypes.c_float, []), "set_voxelWidth": (ctypes.c_bool, [ctypes.c_float]),
    "get_voxelHeight": (ctypes.c_float, []), "set_voxelHeight": (ctypes.c_bool, [ctypes.c_float]),
    "get_offsetX": (ctypes.c_float, []), "set_offsetX": (ctypes.c_bool, [ctypes.c_float]),
    "get_offsetY": (ctypes.c_float, []), "set_offsetY": (ctypes.c_bool, [ctypes.c_float]),
    "get_offsetZ": (ctypes.c_float, []), "set_offsetZ": (ctypes.c_bool, [ctypes.c_float]),
    "get_numPlanograms": (ctypes.c_int, []),
    "get_planogramSize": (ctypes.c_bool, [ctypes.c_int, ndpointer(ctypes.c_int, flags="C_CONTIGUOUS")]),
    "add_planogram": (ctypes.c_bool, [ctypes.c_float] * 7),
    "remove_planogram": (ctypes.c_bool, [ctypes.c_int]),
    "keep_only_planogram": (ctypes.c_bool, [ctypes.c_int]),
    "set_default_volume": (ctypes.c_bool, [ctypes.c_float]),
    "set_volume": (ctypes.c_bool, [ctypes.c_int] * 3 + [ctypes.c_float] * 5),
    "project": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _f32]),
    "backproject": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _f32]),
    "multiply3D": (ctypes.c_bool, [_f32, _f32] + [ctypes.c_int] * 3),
    "multiply4D": (ctypes.c_bool, [_f32, _f32] + [ctypes.c_int] * 4),
    "divide3D": (ctypes.c_bool, [_f32, _f32] + [ctypes.c_int] * 3),
    "divide4D": (ctypes.c_bool, [_f32, _f32] + [ctypes.c_int] * 4),
    "rdivide3D": (ctypes.c_bool, [_f32, _f32] + [ctypes.c_int] * 3),
    "rdivide4D": (ctypes.c_bool, [_f32, _f32] + [ctypes.c_int] * 4),
    "reciprocal3D": (ctypes.c_bool, [_f32] + [ctypes.c_int] * 3 + [ctypes.c_float]),
    "reciprocal4D": (ctypes.c_bool, [_f32] + [ctypes.c_int] * 4 + [ctypes.c_float]),
    "inner_product3D": (ctypes.c_float, [_c_float_p, _c_float_p, _c_float_p] + [ctypes.c_int] * 3),
    "inner_product4D": (ctypes.c_float, [_c_float_p, _c_float_p, _c_float_p] + [ctypes.c_int] * 4),
    "TV_gradient": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 3 +
                    [ctypes.c_float] * 3 + [ctypes.c_bool]),
    "relativeDifferences_gradient": (ctypes.c_bool, [_f32, _f32] + [ctypes.c_int] * 3 + [ctypes.c_float] * 2),
    "TV_cost": (ctypes.c_float, [ctypes.c_void_p] + [ctypes.c_int] * 3 + [ctypes.c_float] * 3 + [ctypes.c_bool]),
    "TV_quadForm": (ctypes.c_float, [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 3 + [ctypes.c_float] * 3 +
                    [ctypes.c_bool]),
    "gaussian_filter": (ctypes.c_bool, [_f32] + [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int]),
    "FBP": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _f32]),
    "PFDR": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p)]),
    "add_object": (ctypes.c_bool, [ctypes.c_int, _f32, _f32, ctypes.c_float, _f32, _f32]),
    "add_module": (ctypes.c_bool, [ctypes.c_int, _f32, _f32, ctypes.c_float, _f32, _f32]),
    "clear_phantom": (ctypes.c_bool, []),
    "clear_modules": (ctypes.c_bool, []),
    "scale_phantom": (ctypes.c_bool, [ctypes.c_float] * 3),
    "ray_trace": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), ctypes.c_int]),
    "stopping_power": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), ctypes.c_int]),
    "set_solid_angle_correction": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), ctypes.c_bool]),
    "ray_trace_gpu": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_bool]),
    "RWLS_gpu": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _f32, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_int, ctypes.POINTER(_c_float_p), ctypes.c_float, ctypes.c_int, ctypes.c_bool,
                                 ctypes.POINTER(_c_float_p), ctypes.c_void_p, ctypes.c_void_p]),
    "set_backprojector_mode": (ctypes.c_bool, [ctypes.c_int]),
    "get_backprojector_mode": (ctypes.c_int, []),
    "set_projector_path": (ctypes.c_bool, [ctypes.c_int]),
    "get_projector_path": (ctypes.c_int, []),
    "planogram_uses_fast_path": (ctypes.c_int, [ctypes.c_int]),
    "project_gpu": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _c_float_p, ctypes.c_void_p, ctypes.c_bool]),
    "backproject_gpu": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _c_float_p, ctypes.c_void_p, ctypes.c_bool]),
    "MLEM_gpu": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _f32, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.POINTER(_c_float_p),
                                 ctypes.POINTER(_c_float_p), ctypes.c_void_p]),
    "bin": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), ctypes.c_char_p]),
    "bin_events": (ctypes.c_bool, [_f32, _f32, ctypes.c_longlong, ctypes.POINTER(_c_float_p), ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]),
    "bin_events_gpu": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                       ctypes.POINTER(_c_float_p), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_bool]),
    "project_gpu_ex": (ctypes.c_bool, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_bool]),
    "backproject_gpu_ex": (ctypes.c_bool, [ctypes.POINTER(ctypes.c_void_p), _c_float_p, ctypes.c_int, ctypes.c_float,
                                           ctypes.c_void_p, ctypes.c_bool]),
    "project_gpu_batch": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _c_float_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_bool]),
    "backproject_gpu_batch": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _c_float_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_bool]),
    "MLEM_gpu_storage": (ctypes.c_bool, [ctypes.POINTER(_c_float_p), _f32, ctypes.c_int, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]),
}

# storage formats of device buffers (include/petl_c_interface.h): fp32, bfloat16, binary16 with a scale
STORAGE = {"f32": 0, "bf16": 1, "f16": 2}


def _storage_of_tensor(t):
    name = str(t.dtype)
    for key, code in (("float32", 0), ("bfloat16", 1), ("float16", 2)):
        if name == "torch." + key:
            return code
    raise TypeError(f"unsupported device dtype {t.dtype} (float32, bfloat16 or float16)")

_lib = None


def library_path():
    """petl_amd/libpetl.so, or the library named by the PETL_LIBRARY environment variable."""
    return os.environ.get("PETL_LIBRARY") or os.path.join(_HERE, "libpetl.so")


def _share_torch_hip_runtime():
    """PyTorch's ROCm wheel carries its own HIP runtime (torch/lib/libamdhip64.so, soname libamdhip64.so.7).
    Loading it before libpetl.so makes libpetl.so bind to that same runtime, so a process that also uses
    torch holds one HIP/HSA runtime whichever of the two it loads first (two runtimes cannot both
    initialise the GPU).  Without torch installed, the system ROCm runtime is used."""
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    hip = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.isfile(hip):
        ctypes.CDLL(hip, mode=ctypes.RTLD_GLOBAL)


def load_library(path=None):
    """Loads libpetl.so (once per process) and declares every signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or library_path()
    if not os.path.isfile(path):
        raise RuntimeError(f"PETL library not found at {path}; build it with `python -m petl_amd.build`")
    _share_torch_hip_runtime()
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PETLserver:
    """Python interface of the PETL projector library (one instance = one PETL model)."""

    BACK_PETL = 0
    BACK_MATCHED = 1

    def __init__(self, param_id=None, lib_dir=""):
        if len(lib_dir) > 0:
            cands = glob.glob(os.path.join(lib_dir, "*petl*.so"))
            self.libprojectors = load_library(cands[0] if cands else None)
        else:
            self.libprojectors = load_library()
        if param_id is not None:
            self.param_id = param_id
        else:
            self.param_id = self.create_new_model()
        self.set_model()
        self.print_cost = False
        self.print_warnings = True
        self.response = None
        self.include_response_with_projectors = False

    # ------------------------------------------------------------------ model / geometry
    def about(self):
        self.set_model()
        self.libprojectors.about()

    def set_response(self, R):
        self.response = R

    def clear_response(self):
        self.response = None

    def create_new_model(self):
        return self.libprojectors.create_new_model()

    def set_model(self, i=None):
        return self.libprojectors.set_model(self.param_id if i is None else i)

    def volume_defined(self):
        self.set_model()
        return self.libprojectors.volume_defined()

    def geometry_defined(self):
        self.set_model()
        return self.libprojectors.geometry_defined()

    def all_defined(self):
        return bool(self.volume_defined() and self.geometry_defined())

    def print_parameters(self):
        self.set_model()
        self.libprojectors.print_parameters()

    def _get_set(self, name, value):
        self.set_model()
        if value is None:
            return getattr(self.libprojectors, "get_" + name)()
        if getattr(self.libprojectors, "set_" + name)(value):
            return getattr(self.libprojectors, "get_" + name)()
        return None

    def numX(self, N=None):
        return self._get_set("numX", N)

    def numY(self, N=None):
        return self._get_set("numY", N)

    def numZ(self, N=None):
        return self._get_set("numZ", N)

    def voxelWidth(self, w=None):
        return self._get_set("voxelWidth", w)

    def voxelHeight(self, h=None):
        return self._get_set("voxelHeight", h)

    def offsetX(self, x_0=None):
        return self._get_set("offsetX", x_0)

    def offsetY(self, y_0=None):
        return self._get_set("offsetY", y_0)

    def offsetZ(self, z_0=None):
        return self._get_set("offsetZ", z_0)

    def get_numPlanograms(self):
        self.set_model()
        return self.libprojectors.get_numPlanograms()

    def get_planogramSize(self, n):
        if self.geometry_defined():
            shape = np.zeros(4, dtype=np.int32)
            self.set_model()
            self.libprojectors.get_planogramSize(n, shape)
            return shape
        return None

    def allocate_projections(self):
        if self.geometry_defined():
            return [np.zeros(tuple(int(s) for s in self.get_planogramSize(n)), dtype=np.float32)
                    for n in range(self.get_numPlanograms())]
        return None

    def allocate_rebinned_projections(self):
        """Zeros of the rebinned planogram shapes [N_v0][N_u1][N_u0] (src/petl_server.py:432-448)."""
        if not self.geometry_defined():
            return None
        out = []
        for n in range(self.get_numPlanograms()):
            sh = self.get_planogramSize(n)
            out.append(np.zeros((int(sh[1]), int(sh[2]), int(sh[3])), dtype=np.float32))
        return out

    def FBP(self, g, f=None):
        """PFDR rebinning (when the data are not rebinned yet) + ramp filter + back projection
        (src/petl_server.py:757-784).  The geometry is kept; data that are already rebinned are filtered in place."""
        if not self.all_defined():
            raise ValueError('Error: volume and PET geometry must be defined')
        if f is None:
            f = self.allocate_volume()
        self.set_model()
        if not self.libprojectors.FBP(self.floatstarstar(g), f):
            raise TypeError('FBP failed')
        return f

    def PFDR(self, g, g_reb=None):
        """Planogram frequency-distance rebinning (src/petl_server.py:786-811).  As in the reference this
        changes the planogram specification to the rebinned one (N_v1 = 1)."""
        if not self.all_defined():
            raise ValueError('Error: volume and PET geometry must be defined')
        if g_reb is None:
            g_reb = self.allocate_rebinned_projections()
        self.set_model()
        if not self.libprojectors.PFDR(self.floatstarstar(g), self.floatstarstar(g_reb)):
            raise TypeError('PFDR failed')
        return g_reb

    def copy_data(self, x):
        if isinstance(x, list):
            return [a.copy() for a in x]
        return x.copy()

    def copyData(self, x):
        return self.copy_data(x)

    def allocate_volume(self):
        if self.volume_defined():
            return np.zeros((self.numZ(), self.numY(), self.numX()), dtype=np.float32)
        print('Error: volume is not defined')
        return None

    def clearAll(self):
        self.set_model()
        self.libprojectors.clearAll()

    def add_planogram(self, psi, D, L, H, T, v_m0=None, v_m1=None):
        """Adds a pair of parallel flat panels (src/petl_server.py:484-507 defaults)."""
        if v_m0 is None:
            v_m0 = 1.0
        if v_m1 is None:
            v_m1 = min(1.0, (H - 1.5 * T) / D)
        self.set_model()
        if not self.libprojectors.add_planogram(psi, D, L, H, v_m0, v_m1, T):
            raise ValueError('Error: add_planogram failed')

    def remove_planogram(self, n):
        self.set_model()
        if not self.libprojectors.remove_planogram(n):
            raise ValueError('Error: remove_planogram failed')

    def keep_only_planogram(self, n):
        self.set_model()
        if not self.libprojectors.keep_only_planogram(n):
            raise ValueError('Error: keep_only_planogram failed')

    def set_default_volume(self, scale=1.0):
        self.set_model()
        if not self.libprojectors.set_default_volume(scale):
            raise ValueError('Error: set_default_volume failed')

    def set_volume(self, numX, numY, numZ, voxelWidth, voxelHeight=0.0, offsetX=0.0, offsetY=0.0, offsetZ=0.0):
        if voxelHeight == 0.0:
            voxelHeight = voxelWidth
        self.set_model()
        if not self.libprojectors.set_volume(numX, numY, numZ, voxelWidth, voxelHeight, offsetX, offsetY, offsetZ):
            raise ValueError('Error: set_volume failed')

    # ------------------------------------------------------------------ projectors
    def floatstarstar(self, g):
        ptrs = [x.ctypes.data_as(_c_float_p) for x in g]
        return (_c_float_p * len(ptrs))(*ptrs)

    @staticmethod
    def _check_arrays(g, f):
        for x in list(g) + [f]:
            if x.dtype != np.float32 or not x.flags["C_CONTIGUOUS"]:
                raise TypeError("arrays must be C-contiguous float32")

    def project(self, g, f):
        if not self.all_defined():
            raise ValueError('Error: volume and PET geometry must be defined')
        self._check_arrays(g, f)
        self.set_model()
        if not self.libprojectors.project(self.floatstarstar(g), f):
            raise TypeError('project failed')
        if self.response is not None and self.include_response_with_projectors:
            self.multiply(g, self.response)
        return g

    def backproject(self, g, f=None):
        if not self.all_defined():
            raise ValueError('Error: volume and PET geometry must be defined')
        if f is None:
            f = self.allocate_volume()
        if self.response is not None and self.include_response_with_projectors:
            Rg = self.copyData(g)
            self.multiply(Rg, self.response)
        else:
            Rg = g
        self._check_arrays(Rg, f)
        self.set_model()
        if not self.libprojectors.backproject(self.floatstarstar(Rg), f):
            raise TypeError('backproject failed')
        return f

    def set_backprojector_mode(self, mode):
        """'petl' (reference semantics, default) or 'matched' (exact transpose of project)."""
        m = {"petl": 0, "matched": 1}.get(mode, mode)
        self.set_model()
        if not self.libprojectors.set_backprojector_mode(int(m)):
            raise ValueError('Error: unknown back projector mode')

    def get_backprojector_mode(self):
        self.set_model()
        return ["petl", "matched"][self.libprojectors.get_backprojector_mode()]

    def set_projector_path(self, path):
        """'auto' (fast separable kernels where the geometry allows) or 'generic'."""
        p = {"auto": 0, "generic": 1}.get(path, path)
        self.set_model()
        if not self.libprojectors.set_projector_path(int(p)):
            raise ValueError('Error: unknown projector path')

    def uses_fast_path(self, n):
        self.set_model()
        return bool(self.libprojectors.planogram_uses_fast_path(n))

    def _dev_ptrs(self, g):
        ptrs = [ctypes.cast(ctypes.c_void_p(x.data_ptr()), _c_float_p) for x in g]
        return (_c_float_p * len(ptrs))(*ptrs)

    def project_gpu(self, g, f, stream=None, sync=True, f_scale=1.0, g_scale=1.0):
        """Forward projection on device tensors (torch, contiguous) without host copies.  float32 tensors
        use the reference-compatible entry; bfloat16 / float16 tensors (the volume in the planograms' format
        or float32) are read and written in that format with fp32 arithmetic; for float16, value =
        stored * scale (f_scale for the volume, g_scale for the planograms)."""
        self.set_model()
        s = None if stream is None else ctypes.c_void_p(stream)
        sf, sg = _storage_of_tensor(f), _storage_of_tensor(g[0])
        if sf == 0 and sg == 0:
            ok = self.libprojectors.project_gpu(self._dev_ptrs(g), ctypes.cast(ctypes.c_void_p(f.data_ptr()),
                                                                                _c_float_p), s, sync)
        else:
            ptrs = (ctypes.c_void_p * len(g))(*[x.data_ptr() for x in g])
            ok = self.libprojectors.project_gpu_ex(ptrs, ctypes.c_void_p(f.data_ptr()), sf, sg, float(f_scale),
                                                   float(g_scale), s, sync)
        if not ok:
            raise TypeError('project failed')
        return g

    def backproject_gpu(self, g, f, stream=None, sync=True, g_scale=1.0):
        """Back projection of device planograms (float32, or bfloat16 / float16 read with fp32 arithmetic)
        into a float32 device volume."""
        self.set_model()
        s = None if stream is None else ctypes.c_void_p(stream)
        sg = _storage_of_tensor(g[0])
        fp = ctypes.cast(ctypes.c_void_p(f.data_ptr()), _c_float_p)
        if sg == 0:
            ok = self.libprojectors.backproject_gpu(self._dev_ptrs(g), fp, s, sync)
        else:
            ptrs = (ctypes.c_void_p * len(g))(*[x.data_ptr() for x in g])
            ok = self.libprojectors.backproject_gpu_ex(ptrs, fp, sg, float(g_scale), s, sync)
        if not ok:
            raise TypeError('backproject failed')
        return f

    def project_gpu_batch(self, g, f, stream=None, sync=True):
        """Forward projection of a batch: f is a (B, nz, ny, nx) float32 device tensor, g a list with one
        (B, N_v1, N_v0, N_u1, N_u0) float32 device tensor per planogram (all contiguous)."""
        self.set_model()
        s = None if stream is None else ctypes.c_void_p(stream)
        fp = ctypes.cast(ctypes.c_void_p(f.data_ptr()), _c_float_p)
        if not self.libprojectors.project_gpu_batch(self._dev_ptrs(g), fp, int(f.shape[0]), s, sync):
            raise TypeError('project failed')
        return g

    def backproject_gpu_batch(self, g, f, mode=None, stream=None, sync=True):
        """Back projection of a batch (shapes as project_gpu_batch) into f, overwritten; mode 'petl' or
        'matched' for this call only (None = the server's mode)."""
        self.set_model()
        s = None if stream is None else ctypes.c_void_p(stream)
        m = -1 if mode is None else {"petl": 0, "matched": 1}[mode]
        fp = ctypes.cast(ctypes.c_void_p(f.data_ptr()), _c_float_p)
        if not self.libprojectors.backproject_gpu_batch(self._dev_ptrs(g), fp, int(f.shape[0]), m, s, sync):
            raise TypeError('backproject failed')
        return f

    # ------------------------------------------------------------------ list-mode binning
    def bin(self, file_name, g=None):
        """Bins list mode data (src/petl_server.py:653-677): the planograms are zeroed, then the file's
        coincidences are binned (".lst" or the acquisition layout; millimetres, chest-wall origin,
        bilinear u bins).  Returns g."""
        if not self.all_defined():
            raise ValueError('Error: volume and PET geometry must be defined')
        if g is None:
            g = self.allocate_projections()
        file_name = bytes(str(file_name), 'ascii')
        self.set_model()
        if not self.libprojectors.bin(self.floatstarstar(g), file_name):
            pass  # the reference ignores a failed bin (src/petl_server.py:674-676)
        return g

    def bin_events(self, blue, pink, g=None, nearest=False, chest_wall=True):
        """Adds coincidences (blue, pink: (n, 3) float32 end points, mm) to the planograms g (zeros when
        None).  Returns (g, coincidences accepted per planogram)."""
        if not self.geometry_defined():
            raise ValueError('Error: PET geometry must be defined')
        blue = np.ascontiguousarray(blue, dtype=np.float32).reshape(-1, 3)
        pink = np.ascontiguousarray(pink, dtype=np.float32).reshape(-1, 3)
        if blue.shape != pink.shape:
            raise ValueError('blue and pink must have the same shape')
        if g is None:
            g = self.allocate_projections()
        acc = (ctypes.c_longlong * self.get_numPlanograms())()
        self.set_model()
        if not self.libprojectors.bin_events(blue, pink, len(blue), self.floatstarstar(g), int(bool(nearest)),
                                             int(bool(chest_wall)), acc):
            raise TypeError('bin_events failed')
        return g, [int(x) for x in acc]

    def bin_events_gpu(self, blue, pink, g, nearest=False, chest_wall=True, stream=None, sync=True):
        """bin_events on device tensors: blue, pink (n, 3) float32, g the device planograms (added to)."""
        self.set_model()
        s = None if stream is None else ctypes.c_void_p(stream)
        if not self.libprojectors.bin_events_gpu(ctypes.c_void_p(blue.data_ptr()), ctypes.c_void_p(pink.data_ptr()),
                                                 int(blue.shape[0]), self._dev_ptrs(g), int(bool(nearest)),
                                                 int(bool(chest_wall)), s, sync):
            raise TypeError('bin_events_gpu failed')
        return g

    # ------------------------------------------------------------------ response model
    _OBJECT_TYPES = {"ball": 0, "sphere": 0, "ellipsoid": 0, "box": 1, "parallelepiped": 1, "can_x": 2,
                     "can_y": 3, "can_z": 4, "can": 4, "cone_x": 5, "cone_y": 6, "cone_z": 7, "cone": 7}

    @staticmethod
    def _solid_args(c, r, A, clip):
        if A is None:
            A = np.eye(3, dtype=np.float32)
        if clip is None:
            clip = np.zeros(3, dtype=np.float32)
        if isinstance(c, (int, float)):
            c = [c, c, c]
        if isinstance(r, (int, float)):
            r = [r, r, r]
        return (np.ascontiguousarray(c, dtype=np.float32), np.ascontiguousarray(r, dtype=np.float32),
                np.ascontiguousarray(A, dtype=np.float32), np.ascontiguousarray(clip, dtype=np.float32))

    def add_object(self, typeOfObject, c, r, val, A=None, clip=None):
        """Adds a geometric object to the ray-tracing phantom (src/petl_server.py:813-875): type 0-7 or a name
        ('sphere', 'box', 'can_x', ..., 'cone'); later objects override earlier ones where they overlap."""
        if isinstance(typeOfObject, str):
            t = self._OBJECT_TYPES.get(typeOfObject.lower())
            if t is None:
                print('Error: unknown object type')
                return False
            typeOfObject = t
        c, r, A, clip = self._solid_args(c, r, A, clip)
        self.set_model()
        return self.libprojectors.add_object(int(typeOfObject), c, r, float(val), A, clip)

    def add_module(self, c, r, val=0.0815, A=None):
        """Adds a scintillator module (a box; val = attenuation coefficient at 511 keV, 1/mm) to the model the
        stopping power is computed from (src/petl_server.py:877-910)."""
        c, r, A, clip = self._solid_args(c, r, A, None)
        self.set_model()
        return self.libprojectors.add_module(1, c, r, float(val), A, clip)

    def clear_phantom(self):
        self.set_model()
        self.libprojectors.clear_phantom()

    def clear_modules(self):
        self.set_model()
        self.libprojectors.clear_modules()

    def scale_phantom(self, c):
        """Scales the ray-tracing phantom (src/petl_server.py:922-944)."""
        if isinstance(c, (int, float)):
            cx = cy = cz = c
        elif isinstance(c, np.ndarray) and c.size == 3:
            cx, cy, cz = (float(v) for v in c.ravel())
        else:
            return False
        self.set_model()
        return self.libprojectors.scale_phantom(cx, cy, cz)

    def ray_trace(self, g=None, oversampling=1):
        """Line integrals of the phantom along every bin's ray (src/petl_server.py:967-993)."""
        if not self.geometry_defined():
            raise ValueError("Error: PET scanner not specified")
        if g is None:
            g = self.allocate_projections()
        self.set_model()
        return g if self.libprojectors.ray_trace(self.floatstarstar(g), int(oversampling)) else None

    def stopping_power(self, g=None, oversampling=1):
        """Scintillator stopping power (1 - e^-a)(1 - e^-b) of every bin's two half rays through the modules
        (1 everywhere without modules), then smoothed with a FWHM 0.6571 Gaussian (src/petl_server.py:995-1022)."""
        if not self.geometry_defined():
            raise ValueError("Error: PET scanner not specified")
        if g is None:
            g = self.allocate_projections()
        self.set_model()
        if self.libprojectors.stopping_power(self.floatstarstar(g), int(oversampling)):
            self.gaussian_filter(g, 0.6571)
            return g
        return None

    def ray_trace_gpu(self, g, oversampling=1, stopping=False, stream=None, sync=True):
        """ray_trace (or, with stopping=True, the unsmoothed stopping power) into float32 device tensors."""
        self.set_model()
        if not self.libprojectors.ray_trace_gpu(self._dev_ptrs(g), int(oversampling), int(bool(stopping)), stream,
                                                bool(sync)):
            raise TypeError('ray_trace_gpu failed')
        return g

    def apply_solid_angle_correction(self, g=None, do_inverse=False):
        """Each view times 2 pi (1 + v0^2 + v1^2)^1.5 / (T_v1 T_v0), or its reciprocal (src/petl_server.py:1024-1043)."""
        if g is None:
            g = self.allocate_projections()
            self.setToOne(g)
        self.set_model()
        return g if self.libprojectors.set_solid_angle_correction(self.floatstarstar(g), bool(do_inverse)) else None

    def apply_solid_angle_model(self, g=None):
        return self.apply_solid_angle_correction(g, True)

    def calc_response(self, oversampling=1, mu=None):
        """response = stopping power (oversampling > 0) x solid-angle model / exp(P mu) (src/petl_server.py:578-605)."""
        if oversampling > 0:
            self.response = self.stopping_power(self.response, oversampling=oversampling)
        else:
            self.response = None
        self.response = self.apply_solid_angle_model(self.response)
        if mu is not None:
            Pmu = self.allocate_projections()
            self.project(Pmu, mu)
            for a in Pmu:
                np.exp(a, out=a)
            self.divide(self.response, Pmu)
        return self.response

    def sensitivity(self, Pstar1=None):
        """P*1, the back projection of all-ones planograms (src/petl_server.py:1257-1273)."""
        if Pstar1 is None:
            Pstar1 = self.allocate_volume()
        ones = self.allocate_projections()
        for n in range(len(ones)):
            ones[n][:] = 1.0
        self.backproject(ones, Pstar1)
        return Pstar1

    # ------------------------------------------------------------------ elementwise helpers
    def clip(self, x, min_val=0.0, max_val=None):
        if isinstance(x, list):
            for a in x:
                np.clip(a, min_val, max_val, out=a)
        else:
            np.clip(x, min_val, max_val, out=x)
        return x

    def scale(self, x, c=1.0):
        if isinstance(x, list):
            for a in x:
                a[:] *= c
        else:
            x[:] *= c
        return x

    def square(self, x):
        return self.multiply(x, x)

    def subtract(self, x, y):
        if isinstance(x, list):
            for a, b in zip(x, y):
                a[:] -= b[:]
        else:
            x[:] -= y[:]
        return x

    def scalarAdd(self, x, c, y):
        if isinstance(x, list):
            for a, b in zip(x, y):
                a[:] += c * b[:]
        else:
            x[:] += c * y[:]
        return x

    def _binary(self, op, x, y):
        self.set_model()
        xs = x if isinstance(x, list) else [x]
        ys = y if isinstance(y, list) else [y]
        for a, b in zip(xs, ys):
            if a.ndim == 4:
                ok = getattr(self.libprojectors, op + "4D")(a, b, *a.shape)
            elif a.ndim == 3:
                ok = getattr(self.libprojectors, op + "3D")(a, b, *a.shape)
            else:
                raise ValueError('inputs must be 3D or 4D numpy arrays')
            if not ok:
                print(op + ' failed!')
                return None
        return x

    def multiply(self, x, y):
        """x = x * y"""
        return self._binary("multiply", x, y)

    def divide(self, num, denom):
        """num = num / denom (0 where denom == 0)"""
        return self._binary("divide", num, denom)

    def rdivide(self, denom, num):
        """denom = num / denom (0 where denom == 0)"""
        self._binary("rdivide", num, denom)
        return denom

    def reciprocal(self, x, divide_by_zero_value=0.0):
        self.set_model()
        for a in (x if isinstance(x, list) else [x]):
            fn = self.libprojectors.reciprocal4D if a.ndim == 4 else self.libprojectors.reciprocal3D
            if not fn(a, *a.shape, divide_by_zero_value):
                print('reciprocal failed!')
                return None
        return x

    def as_float_ptr(self, arr):
        if arr is None:
            return None
        return np.asarray(arr, dtype=np.float32, order="C").ctypes.data_as(_c_float_p)

    def inner_product_helper(self, x, y, w=None):
        self.set_model()
        if x.ndim == 4:
            return self.libprojectors.inner_product4D(self.as_float_ptr(x), self.as_float_ptr(y),
                                                      self.as_float_ptr(w), *x.shape)
        if x.ndim == 3:
            return self.libprojectors.inner_product3D(self.as_float_ptr(x), self.as_float_ptr(y),
                                                      self.as_float_ptr(w), *x.shape)
        return np.sum(x * y * w)

    def inner_product(self, x, y, w=None):
        if isinstance(x, list):
            acc = 0.0
            for n in range(len(x)):
                if w is None:
                    acc += self.inner_product_helper(x[n], y[n])
                elif isinstance(w, (int, float)):
                    acc += self.inner_product_helper(x[n], y[n]) * w
                else:
                    acc += self.inner_product_helper(x[n], y[n], w[n])
            return acc
        return self.inner_product_helper(x, y, w)

    def innerProd(self, x, y, w=None):
        return self.inner_product(x, y, w)

    def isAllZeros(self, f):
        return not np.any(f)

    def setToConstant(self, x, c=0.0):
        if isinstance(x, list):
            for a in x:
                a[:] = c
        else:
            x[:] = c

    def setToZero(self, x):
        self.setToConstant(x, 0.0)

    def setToOne(self, x):
        self.setToConstant(x, 1.0)

    # ------------------------------------------------------------------ reconstruction
    def MLEM(self, g, f, numIter, delta=2.0, beta=0.0, mask=None, prior="rd", p=1.0, num_subsets=1,
             return_loglik=False, storage="f32", image_storage=None):
        r"""Maximum Likelihood-Expectation Maximization, f <- f / (P^T 1 + R'(f)) * P^T[g / Pf].

        Same arguments and result as the reference's PETLserver.MLEM (src/petl_server.py:1534-1606):
        f None -> ones, all-zero f -> ones, negatives clipped; beta > 0 adds the relative-difference
        prior gradient (delta) one-step-late; the response, when set, weights the projections.  The
        whole loop runs on the GPU.  Extras: prior='tv' uses the p-Huber TV gradient instead,
        num_subsets > 1 runs OSEM, return_loglik=True also returns sum(y log Pf - Pf) per iteration;
        storage='bf16' / 'f16' keeps the planograms (data, ratio) on the device in that 16-bit format and
        image_storage (default: the same) the image, with fp32 arithmetic (no mask, response or
        log-likelihood then).
        """
        if image_storage is None:
            image_storage = storage
        if storage != "f32" or image_storage != "f32":
            if mask is not None or return_loglik or self.response is not None:
                raise ValueError("16-bit storage takes no mask, response or log-likelihood")
            if f is None:
                f = self.allocate_volume()
                self.setToOne(f)
            self._check_arrays(g, f)
            kind = 0 if beta <= 0.0 else (2 if prior == "tv" else 1)
            self.set_model()
            if not self.libprojectors.MLEM_gpu_storage(self.floatstarstar(g), f, int(numIter), float(delta),
                                                       float(beta), kind, float(p), int(num_subsets),
                                                       STORAGE[storage], STORAGE[image_storage]):
                raise TypeError('MLEM failed')
            return f
        if mask is not None and (len(mask) != len(g) or any(m.shape != a.shape for m, a in zip(mask, g))):
            print('Error: mask must be the same shape as the projection data!')
            return None
        if f is None:
            f = self.allocate_volume()
            self.setToOne(f)
        self._check_arrays(g, f)
        if self.print_warnings:
            print('MLEM: ' + str(numIter) + ' iterations on the GPU')
        kind = 0 if beta <= 0.0 else (2 if prior == "tv" else 1)
        ll = np.zeros(max(numIter, 1), dtype=np.float64)
        resp = self.floatstarstar(self.response) if self.response is not None else None
        msk = self.floatstarstar(mask) if mask is not None else None
        self.set_model()
        ok = self.libprojectors.MLEM_gpu(self.floatstarstar(g), f, int(numIter), float(delta), float(beta), kind,
                                         float(p), int(num_subsets), msk, resp,
                                         ll.ctypes.data_as(ctypes.c_void_p) if return_loglik else None)
        if not ok:
            raise TypeError('MLEM failed')
        if return_loglik:
            return f, ll[:numIter]
        return f

    def OSEM(self, g, f, numIter, numSubsets, delta=2.0, beta=0.0, mask=None):
        return self.MLEM(g, f, numIter, delta=delta, beta=beta, mask=mask, num_subsets=numSubsets)

    def LS(self, g, f, numIter, preconditioner=None, nonnegativityConstraint=True):
        """Least squares, 0.5 ||Pf - g||^2, by preconditioned CG (src/petl_server.py:1608-1633)."""
        return self.RWLS(g, f, numIter, 1.0, 0.0, 1.0, preconditioner, nonnegativityConstraint)

    def WLS(self, g, f, numIter, W=None, preconditioner=None, nonnegativityConstraint=True):
        """Weighted least squares, 0.5 (Pf - g)^T W (Pf - g) (src/petl_server.py:1635-1661)."""
        return self.RWLS(g, f, numIter, 1.0, 0.0, W, preconditioner, nonnegativityConstraint)

    def RLS(self, g, f, numIter, delta=1.0, beta=0.0, preconditioner=None, nonnegativityConstraint=True):
        """Regularized least squares, 0.5 ||Pf - g||^2 + TV(f) (src/petl_server.py:1663-1689)."""
        return self.RWLS(g, f, numIter, delta, beta, 1.0, preconditioner, nonnegativityConstraint)

    def RWLS(self, g, f, numIter, delta=1.0, beta=0.0, W=None, preconditioner=None, nonnegativityConstraint=True,
             include_response_with_projectors=False, return_cost=False):
        r"""Regularized weighted least squares, 0.5 (Pf - g)^T W (Pf - g) + R(f), by preconditioned conjugate
        gradients (src/petl_server.py:1691-1845), the whole loop on the GPU.

        R is the anisotropic p-Huber TV (p = 1, delta, beta; none when beta <= 0).  W: None -> 1 /
        max(gaussian_filter(g, 3.0), 7.0); a list of planogram weights; or a scalar, which (as in the
        reference) leaves the gradient unweighted and scales <Pd, Pd> in the step.  preconditioner: 'SQS' or
        True for the separable quadratic surrogate (P* W P 1)^-1, anything else none.  f: the start (all zeros
        -> Pf = 0; negatives clipped when constrained), updated in place and returned.  With a response set,
        include_response_with_projectors=True makes the projectors include it; the default (False) needs the
        reference's apply_corrections (bad-pixel correction), which this library does not provide.
        Extra: return_cost=True also returns the cost 0.5 (Pf - g)^T W (Pf - g) + TV(f) after each iteration
        run (the values the reference prints with print_cost).
        """
        if self.response is not None and not include_response_with_projectors:
            raise ValueError("RWLS with a response needs include_response_with_projectors=True here "
                             "(apply_corrections is not available)")
        if f is None:
            f = self.allocate_volume()
        self._check_arrays(g, f)
        if W is None:
            mode, Wl, ws = 0, None, 1.0
        elif isinstance(W, list):
            if len(W) != len(g) or any(np.shape(a) != np.shape(b) for a, b in zip(W, g)):
                raise ValueError("W must be a list of arrays shaped like the projection data")
            mode, Wl, ws = 1, [np.ascontiguousarray(a, dtype=np.float32) for a in W], 1.0
        elif isinstance(W, (int, float)):
            mode, Wl, ws = 2, None, float(W)
        else:
            raise ValueError("W must be None, a list of planogram weights or a scalar")
        sqs = preconditioner is True or preconditioner == 'SQS'
        resp = self.floatstarstar(self.response) if (self.response is not None and include_response_with_projectors) \
            else None
        want_cost = return_cost or self.print_cost
        cost = np.zeros(max(int(numIter), 1), dtype=np.float64)
        done = ctypes.c_int(0)
        self.set_model()
        ok = self.libprojectors.RWLS_gpu(self.floatstarstar(g), f, int(numIter), float(delta), float(beta), mode,
                                         self.floatstarstar(Wl) if Wl is not None else None, ws, 1 if sqs else 0,
                                         bool(nonnegativityConstraint), resp,
                                         cost.ctypes.data_as(ctypes.c_void_p) if want_cost else None,
                                         ctypes.byref(done))
        if not ok:
            raise TypeError('RWLS failed')
        if self.print_warnings:
            for n in range(done.value):
                print('RWLS iteration ' + str(n + 1) + ' of ' + str(numIter))
            if done.value < numIter:
                print('invalid step size; quitting!')
        if self.print_cost:
            for n in range(done.value):
                print('\tcost = ' + str(cost[n]))
        if return_cost:
            return f, cost[:done.value].copy()
        return f

    def RWLSstepSize(self, f, grad, d, Pd, W, delta, beta, num=None):
        """<d, grad> / (<Pd, Pd>_W + <d, R''(f) d>), 0 when the denominator is ~0 (src/petl_server.py:1847-1877)."""
        if num is None:
            num = self.innerProd(d, grad)
        denomA = self.innerProd(Pd, Pd, W) if W is not None else self.innerProd(Pd, Pd)
        denomB = self.TVquadForm(f, d, delta, beta) if beta > 0.0 else 0.0
        denom = denomA + denomB
        stepSize = num / denom if np.abs(denom) > 1.0e-16 else 0.0
        if self.print_warnings:
            print('\tlambda = ' + str(stepSize))
        return stepSize

    # ------------------------------------------------------------------ priors
    @staticmethod
    def _dims3(f):
        if f.ndim == 3:
            return f.shape
        if f.ndim == 2:
            return (1,) + f.shape
        raise ValueError('input must be 2D or 3D')

    def TVgradient(self, f, delta, beta=0.0, p=1.0):
        """Gradient of the anisotropic p-Huber TV functional (src/petl_server.py:1924-1974)."""
        N_1, N_2, N_3 = self._dims3(f)
        self.set_model()
        if hasattr(f, "data_ptr"):
            Df = f.clone()
            self.libprojectors.TV_gradient(f.data_ptr(), Df.data_ptr(), N_1, N_2, N_3, delta, beta, p, not f.is_cuda)
            return Df
        Df = np.zeros(f.shape, dtype=np.float32)
        fc = np.ascontiguousarray(f, dtype=np.float32)
        self.libprojectors.TV_gradient(fc.ctypes.data, Df.ctypes.data, N_1, N_2, N_3, delta, beta, p, True)
        return Df

    def TVcost(self, f, delta, beta=0.0, p=1.0):
        """Anisotropic p-Huber TV functional of a 2D/3D array (src/petl_server.py:1879-1922)."""
        N_1, N_2, N_3 = self._dims3(f)
        self.set_model()
        if hasattr(f, "data_ptr"):
            return self.libprojectors.TV_cost(f.data_ptr(), N_1, N_2, N_3, delta, beta, p, not f.is_cuda)
        fc = np.ascontiguousarray(f, dtype=np.float32)
        return self.libprojectors.TV_cost(fc.ctypes.data, N_1, N_2, N_3, delta, beta, p, True)

    def TVquadForm(self, f, d, delta, beta=0.0, p=1.0):
        """<d, R''(f) d> of the p-Huber TV functional with h'(t)/t for h''(t) (src/petl_server.py:1976-2026)."""
        N_1, N_2, N_3 = self._dims3(f)
        self.set_model()
        if hasattr(f, "data_ptr"):
            return self.libprojectors.TV_quadForm(f.data_ptr(), d.data_ptr(), N_1, N_2, N_3, delta, beta, p,
                                                  not f.is_cuda)
        fc = np.ascontiguousarray(f, dtype=np.float32)
        dc = np.ascontiguousarray(d, dtype=np.float32)
        return self.libprojectors.TV_quadForm(fc.ctypes.data, dc.ctypes.data, N_1, N_2, N_3, delta, beta, p, True)

    def gaussian_filter(self, f, FWHM=2.0, numDims=3):
        """Gaussian smoothing in place (src/petl_server.py:2129-2155): lists elementwise, a 4D planogram over
        (u1, u0) within each (v1, v0) row, a 3D array along numDims axes, a 2D array as one slice."""
        if isinstance(f, list):
            for a in f:
                self.gaussian_filter(a, FWHM, numDims)
            return
        if f.ndim == 4:
            N_1, N_2, N_3, numDims = f.shape[0] * f.shape[1], f.shape[2], f.shape[3], 2
        elif f.ndim == 3:
            N_1, N_2, N_3 = f.shape
        else:
            N_1, N_2, N_3 = 1, f.shape[0], f.shape[1]
        self.set_model()
        self.libprojectors.gaussian_filter(f, N_1, N_2, N_3, FWHM, numDims)

    def relative_differences_grad(self, f, delta, beta):
        """Gradient of the relative-difference penalty (src/petl_server.py:2112-2127)."""
        N_1, N_2, N_3 = self._dims3(f)
        Df = np.zeros(f.shape, dtype=np.float32)
        self.set_model()
        self.libprojectors.relativeDifferences_gradient(np.ascontiguousarray(f, dtype=np.float32), Df,
                                                        N_1, N_2, N_3, delta, beta)
        return Df
