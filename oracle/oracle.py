"""CPU oracle loader -- TEST INFRASTRUCTURE ONLY.

ctypes bindings for ``libpetl_oracle.so`` (the C++ restatement of the reference hot path, see
petl_oracle.cpp) and for ``_ref/libpetl_ref_geom.so`` (the reference's own planogram/dataCube classes,
built from the reference sources when they are present).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module, to check or time results.  The product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.environ.get("PETL_ORACLE_LIBRARY") or os.path.join(_HERE, "libpetl_oracle.so")
_REF = os.path.join(_HERE, "_ref", "libpetl_ref_geom.so")


class Plan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("psi", "R", "L", "H", "v_m0", "v_m1")] + \
               [(n, ctypes.c_int) for n in ("N_u0", "N_u1", "N_v0", "N_v1")] + \
               [(n, ctypes.c_float) for n in ("T_u0", "T_u1", "T_v0", "T_v1", "u0_0", "u1_0", "v0_0",
                                               "v1_0", "cos_psi", "sin_psi")]

    @property
    def shape(self):
        return (self.N_v1, self.N_v0, self.N_u1, self.N_u0)


class Solid(ctypes.Structure):
    """A geometric solid as geometricObject::init builds it (src/phantom.cpp:535-640)."""
    _fields_ = [("type", ctypes.c_int), ("c", ctypes.c_float * 3), ("r", ctypes.c_float * 3),
                ("val", ctypes.c_float), ("A", ctypes.c_float * 9), ("clip", (ctypes.c_float * 4) * 6),
                ("rotated", ctypes.c_int), ("nclip", ctypes.c_int), ("cone", ctypes.c_float * 2)]


class Vol(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("numX", "numY", "numZ")] + \
               [(n, ctypes.c_float) for n in ("voxelWidth", "voxelHeight", "offsetX", "offsetY", "offsetZ",
                                               "x0", "y0", "z0")]

    @property
    def shape(self):
        return (self.numZ, self.numY, self.numX)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P, V = ctypes.POINTER(Plan), ctypes.POINTER(Vol)
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        c_int, c_float, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_double
        L.orc_plan_init.argtypes = [c_float] * 7 + [P]
        L.orc_plan_init.restype = c_int
        L.orc_vol_init.argtypes = [c_int] * 3 + [c_float] * 5 + [V]
        L.orc_bin_masked.argtypes = [P, c_int, c_int, c_int, c_int]
        L.orc_bin_masked.restype = c_int
        L.orc_project_f32.argtypes = [P, V, fp, fp, c_int]
        L.orc_project_f64.argtypes = [P, V, dp, dp, c_int]
        L.orc_backproject_f32.argtypes = [P, V, fp, fp, c_int, c_int]
        L.orc_backproject_f64.argtypes = [P, V, dp, dp, c_int, c_int]
        L.orc_bin_taps.argtypes = [P, V, c_int, c_int, c_int, c_int, i64p, dp, c_int,
                                   ctypes.POINTER(c_double)]
        L.orc_bin_taps.restype = c_int
        L.orc_count_taps.argtypes = [P, V, c_int, c_int, c_int, c_int]
        L.orc_count_taps.restype = ctypes.c_int64
        L.orc_rd_gradient_f32.argtypes = [fp, fp, c_int, c_int, c_int, c_float, c_float]
        L.orc_rd_gradient_f64.argtypes = [dp, dp, c_int, c_int, c_int, c_double, c_double]
        L.orc_rd_cost_f64.argtypes = [dp, c_int, c_int, c_int, c_double, c_double]
        L.orc_rd_cost_f64.restype = c_double
        L.orc_tv_gradient_f32.argtypes = [fp, fp, c_int, c_int, c_int, c_float, c_float, c_float]
        L.orc_tv_gradient_f64.argtypes = [dp, dp, c_int, c_int, c_int, c_float, c_float, c_float]
        L.orc_tv_cost_f64.argtypes = [dp, c_int, c_int, c_int, c_float, c_float, c_float]
        L.orc_tv_cost_f64.restype = c_double
        L.orc_tv_quadform_f64.argtypes = [dp, dp, c_int, c_int, c_int, c_float, c_float, c_float]
        L.orc_tv_quadform_f64.restype = c_double
        L.orc_gaussian_filter_f64.argtypes = [dp, dp, c_int, c_int, c_int, c_float, c_int]
        S = ctypes.POINTER(Solid)
        L.orc_solid_init.argtypes = [c_int, fp, fp, c_float, ctypes.c_void_p, ctypes.c_void_p, S]
        L.orc_solid_init.restype = c_int
        L.orc_intersect.argtypes = [S, fp, fp, fp]
        L.orc_intersect.restype = c_int
        L.orc_ray_value.argtypes = [S, c_int, fp, fp, c_int]
        L.orc_ray_value.restype = c_float
        L.orc_ray_trace.argtypes = [P, S, c_int, c_int, c_int, fp, c_int]
        L.orc_max_threads.restype = c_int
        L.orc_project_rows_f32.argtypes = [P, V, fp, fp, c_int, c_int, c_int]
        L.orc_backproject_petl_rows_f32.argtypes = [P, V, fp, fp, c_int, c_int, c_int]
        L.orc_project_subset_f32.argtypes = [P, V, fp, fp, c_int, c_int, c_int]
        L.orc_bin_events.argtypes = [P, fp, fp, ctypes.c_int64, c_int, c_int, fp]
        L.orc_bin_events.restype = ctypes.c_int64
        L.orc_event_coords.argtypes = [P, fp, fp, fp]
        L.orc_event_coords.restype = c_int
        L.orc_backproject_petl_subset_f32.argtypes = [P, V, fp, fp, c_int, c_int, c_int]
        L.orc_backproject_petl_tex_f32.argtypes = [P, V, fp, fp, c_int, c_int, c_int, c_int]
        _lib = L
    return _lib


def nthreads():
    return lib().orc_max_threads()


# --------------------------------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------------------------------
def planogram(psi, D, L, H, T, v_m0=None, v_m1=None):
    """Planogram sampling as produced by PETLserver.add_planogram (src/petl_server.py:484-507)."""
    if v_m0 is None:
        v_m0 = 1.0
    if v_m1 is None:
        v_m1 = min(1.0, (H - 1.5 * T) / D)
    p = Plan()
    if not lib().orc_plan_init(psi, D, L, H, v_m0, v_m1, T, ctypes.byref(p)):
        raise ValueError("invalid planogram")
    return p


def volume(numX, numY, numZ, voxelWidth, voxelHeight=0.0, offsetX=0.0, offsetY=0.0, offsetZ=0.0):
    if voxelHeight == 0.0:
        voxelHeight = voxelWidth
    v = Vol()
    lib().orc_vol_init(numX, numY, numZ, voxelWidth, voxelHeight, offsetX, offsetY, offsetZ, ctypes.byref(v))
    return v


# --------------------------------------------------------------------------------------------------
# projectors (lists of planograms, like PETLserver)
# --------------------------------------------------------------------------------------------------
def _nt(n):
    return nthreads() if n is None else n


def project(plans, vol, f, dtype=np.float32, nthreads=None):
    f = np.ascontiguousarray(f, dtype=dtype)
    fn = lib().orc_project_f32 if dtype == np.float32 else lib().orc_project_f64
    out = []
    for p in plans:
        g = np.zeros(p.shape, dtype=dtype)
        fn(ctypes.byref(p), ctypes.byref(vol), f, g, _nt(nthreads))
        out.append(g)
    return out


def backproject(plans, vol, g, mode="petl", dtype=np.float32, nthreads=None):
    """mode 'petl' = reference back projector semantics, 'matched' = exact transpose of project."""
    fn = lib().orc_backproject_f32 if dtype == np.float32 else lib().orc_backproject_f64
    f = np.zeros(vol.shape, dtype=dtype)
    m = 1 if mode == "matched" else 0
    for p, gi in zip(plans, g):
        fn(ctypes.byref(p), ctypes.byref(vol), np.ascontiguousarray(gi, dtype=dtype), f, m, _nt(nthreads))
    return f


def project_rows(plan, vol, f, g, v1_begin, v1_end, nthreads=None):
    """fp32 forward of v1 rows [v1_begin, v1_end) into g (full planogram array)."""
    lib().orc_project_rows_f32(ctypes.byref(plan), ctypes.byref(vol), f, g, v1_begin, v1_end, _nt(nthreads))


def backproject_petl_rows(plan, vol, g, f, v1_begin, v1_end, nthreads=None):
    """fp32 PETL-mode back projection of v1 rows [v1_begin, v1_end), accumulated into f."""
    lib().orc_backproject_petl_rows_f32(ctypes.byref(plan), ctypes.byref(vol), g, f, v1_begin, v1_end, _nt(nthreads))


TEX_QUANT = {"exact": 0, "nearest": 1, "truncate": 2}


def backproject_petl_tex(plans, vol, g, quant="nearest", v1_rows=None, nthreads=None):
    """PETL-mode back projection in the reference's texture-fetch form (bins merged in pairs by bilinear
    fetches, src/projectors.cu:351-370) with the interpolation fractions exact or quantised to 8
    fractional bits ('nearest' or 'truncate'); v1_rows: (begin, end) or None for all.  fp32."""
    f = np.zeros(vol.shape, np.float32)
    for p, gi in zip(plans, g):
        b, e = (0, p.N_v1) if v1_rows is None else v1_rows
        lib().orc_backproject_petl_tex_f32(ctypes.byref(p), ctypes.byref(vol), np.ascontiguousarray(gi, np.float32),
                                           f, TEX_QUANT[quant], b, e, _nt(nthreads))
    return f


def bin_taps(plan, vol, iv1, iv0, iu1, iu0, cap=1 << 16):
    idx = np.zeros(cap, np.int64)
    w = np.zeros(cap, np.float64)
    scale = ctypes.c_double()
    n = lib().orc_bin_taps(ctypes.byref(plan), ctypes.byref(vol), iv1, iv0, iu1, iu0, idx, w, cap,
                           ctypes.byref(scale))
    return idx[:n], w[:n], scale.value


def count_taps(plan, vol, v1_begin=0, v1_end=None, f64=False, nthreads=None):
    """Non-zero footprint taps of one projection over v1 rows [v1_begin, v1_end) (unmasked bins), weights
    formed in fp32 (default) or FP64."""
    v1_end = plan.N_v1 if v1_end is None else v1_end
    return int(lib().orc_count_taps(ctypes.byref(plan), ctypes.byref(vol), int(v1_begin), int(v1_end), int(f64),
                                    _nt(nthreads)))


def bin_masked(plan, iv1, iv0, iu1, iu0):
    return bool(lib().orc_bin_masked(ctypes.byref(plan), iv1, iv0, iu1, iu0))


# --------------------------------------------------------------------------------------------------
# priors
# --------------------------------------------------------------------------------------------------
def rd_gradient(f, delta, beta, dtype=np.float32):
    f = np.ascontiguousarray(f, dtype=dtype)
    D = np.zeros_like(f)
    N1, N2, N3 = f.shape
    (lib().orc_rd_gradient_f32 if dtype == np.float32 else lib().orc_rd_gradient_f64)(
        f, D, N1, N2, N3, delta, beta)
    return D


def rd_cost(f, delta, beta):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_rd_cost_f64(f, *f.shape, delta, beta)


def tv_gradient(f, delta, beta, p=1.0, dtype=np.float32):
    f = np.ascontiguousarray(f, dtype=dtype)
    D = np.zeros_like(f)
    N1, N2, N3 = f.shape
    (lib().orc_tv_gradient_f32 if dtype == np.float32 else lib().orc_tv_gradient_f64)(
        f, D, N1, N2, N3, delta, beta, p)
    return D


def _as3d(f):
    """2D arrays are taken as one slice (N_1 = 1), as PETLserver.TVcost does (src/petl_server.py:1906-1913)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    return f.reshape((1,) + f.shape) if f.ndim == 2 else f


def tv_cost(f, delta, beta, p=1.0):
    f = _as3d(f)
    return lib().orc_tv_cost_f64(f, *f.shape, delta, beta, p)


def tv_quadform(f, d, delta, beta, p=1.0):
    """<d, R''(f) d> of the p-Huber TV prior with the surrogate curvature h'(t)/t."""
    f, d = _as3d(f), _as3d(d)
    return lib().orc_tv_quadform_f64(f, d, *f.shape, delta, beta, p)


def gaussian_filter(f, FWHM=2.0, numDims=3):
    """PETLserver.gaussian_filter (src/petl_server.py:2129-2155): a 4D planogram is smoothed per
    (v1, v0) row over (u1, u0); a 3D array along numDims axes; a 2D array as one slice."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    if f.ndim == 4:
        shape3, numDims = (f.shape[0] * f.shape[1], f.shape[2], f.shape[3]), 2
    elif f.ndim == 3:
        shape3 = f.shape
    else:
        shape3 = (1,) + f.shape
    out = np.zeros(shape3, np.float64)
    lib().orc_gaussian_filter_f64(f.reshape(shape3), out, *shape3, FWHM, numDims)
    return out.reshape(f.shape)


# --------------------------------------------------------------------------------------------------
# elementwise semantics (src/data_cube.cpp:144-155) and MLEM (src/petl_server.py:1534-1606)
# --------------------------------------------------------------------------------------------------
def divide(x, y):
    """x / y with y == 0 -> 0 (dataCube DIV)."""
    out = np.zeros_like(x)
    nz = y != 0
    out[nz] = x[nz] / y[nz]
    return out


def rdivide(denom, num):
    """num / denom with denom == 0 -> 0 (dataCube RDIV)."""
    return divide(num, denom)


def _loglik_terms(q, y, rows=8):
    """FP64 sum of y log q - q over the bins with q > 0, by chunks of v1 rows (bounded temporaries)."""
    tot = 0.0
    for a in range(0, q.shape[0], rows):
        qa = np.asarray(q[a:a + rows], np.float64)
        ya = np.asarray(y[a:a + rows], np.float64)
        pos = qa > 0
        tot += float(np.sum(ya[pos] * np.log(qa[pos]) - qa[pos]))
    return tot


def sensitivity_image(plans, vol, mode="petl", dtype=np.float32, nthreads=None):
    """S = sum_i A_i^T 1, zeros replaced by ones (src/petl_server.py:1257-1273, 1580-1585)."""
    S = np.zeros(vol.shape, dtype=dtype)
    for pl in plans:
        S += backproject([pl], vol, [np.ones(pl.shape, dtype=dtype)], mode=mode, dtype=dtype, nthreads=nthreads)
    S[S == 0] = 1
    return S


def mlem(plans, vol, g, num_iter, f=None, delta=2.0, beta=0.0, prior="rd", p=1.0, dtype=np.float32,
         mode="petl", callback=None, nthreads=None, loglik=None, keep_projections=True, sens=None):
    """PETLserver.MLEM (src/petl_server.py:1534-1606) restated on the oracle projectors.  prior 'rd'
    follows the reference (relative differences, OSL); prior 'tv' uses the p-Huber TV gradient in the same
    OSL slot.  Planograms are processed one at a time (one planogram-sized temporary at a time), summing
    their back projections in planogram order, which is the order backproject() sums them in.  loglik: a
    list that receives, per iteration, the FP64 Poisson log-likelihood sum(y log Af - Af) over bins with
    Af > 0 of the image the iteration starts from.  callback(n, f, Pd) runs after every iteration; Pd is
    the list of that iteration's projections, or None when keep_projections is False.  sens: the
    sensitivity sensitivity_image() returns for the same planograms and mode, to skip forming it again
    (a long run can then be resumed one iteration at a time)."""
    if f is None or not np.any(f):
        f = np.ones(vol.shape, dtype=dtype)
    else:
        f = np.array(f, dtype=dtype)
        f[f < 0] = 0
    S = sensitivity_image(plans, vol, mode, dtype, nthreads) if sens is None else np.asarray(sens, dtype)
    for n in range(num_iter):
        d = np.zeros(vol.shape, dtype=dtype)
        Pd = [] if (callback is not None and keep_projections) else None
        ll = 0.0
        for pl, gi in zip(plans, g):
            q = project([pl], vol, f, dtype=dtype, nthreads=nthreads)[0]
            if loglik is not None:
                ll += _loglik_terms(q, gi)
            if Pd is not None:
                Pd.append(q.copy())
            # dataCube RDIV, in place: q <- y / q, 0 where q == 0 (src/data_cube.cpp:150-155)
            nz = q != 0
            q[nz] = np.asarray(np.asarray(gi)[nz], dtype=dtype) / q[nz]
            q[~nz] = 0
            del nz
            d += backproject([pl], vol, [q], mode=mode, dtype=dtype, nthreads=nthreads)
            del q
        if loglik is not None:
            loglik.append(ll)
        if beta <= 0.0:
            f = f * divide(d, S)
        else:
            if prior == "rd":
                G = rd_gradient(f, delta, beta, dtype=dtype)
            else:
                G = tv_gradient(f, delta, beta, p, dtype=dtype)
            G = G + S
            G = np.maximum(dtype(0.1) * S, G)
            f = f * divide(d, G)
        if callback is not None:
            callback(n, f, Pd)
    return f


def row_subset(g, first, stride):
    """A copy of planogram g that keeps the v1 rows first, first + stride, ... and zeroes the others."""
    out = np.zeros_like(g)
    out[first::stride] = g[first::stride]
    return out


def osem_composed(plans, vol, g, num_iter, num_subsets, f0=None, dtype=np.float32, mode="petl", nthreads=None,
                  callback=None):
    """OSEM built from the full-planogram oracle projectors: subset b's forward is the full forward restricted
    to its v1 rows, and its back projection is the back projection of a planogram that is zero outside its
    rows (the projectors are linear, and a zero bin adds nothing).  Independent of the row-range entry points
    osem() uses, and available in FP64 and in matched mode, where the EM identities hold exactly.
    callback(sub, f_before, f_after, S_sub, ratio_list) runs after every subset update."""
    f = np.ones(vol.shape, dtype) if f0 is None else np.array(f0, dtype=dtype)
    S = []
    for sub in range(num_subsets):
        Ss = backproject(plans, vol, [row_subset(np.ones(p.shape, dtype), sub, num_subsets) for p in plans],
                         mode=mode, dtype=dtype, nthreads=nthreads)
        Ss[Ss == 0] = 1
        S.append(Ss)
    for _ in range(num_iter):
        for sub in range(num_subsets):
            q = project(plans, vol, f, dtype=dtype, nthreads=nthreads)
            r = [row_subset(rdivide(qi, np.asarray(gi, dtype)), sub, num_subsets) for qi, gi in zip(q, g)]
            d = backproject(plans, vol, r, mode=mode, dtype=dtype, nthreads=nthreads)
            f_next = f * divide(d, S[sub])
            if callback is not None:
                callback(sub, f, f_next, S[sub], [row_subset(qi, sub, num_subsets) for qi in q])
            f = f_next
    return f


def rwls_default_weights(g):
    """W = 1 / max(gaussian_filter(g, FWHM 3), 7) per planogram (src/petl_server.py:1722-1726)."""
    return [1.0 / np.maximum(gaussian_filter(gi, 3.0), 7.0) for gi in g]


def rwls(plans, vol, g, num_iter, f=None, delta=1.0, beta=0.0, W=None, sqs=False, nonneg=True, mode="petl",
         response=None, callback=None, nthreads=None):
    """PETLserver.RWLS (src/petl_server.py:1691-1877) restated in FP64 on the oracle projectors:
    preconditioned conjugate gradients on 0.5 (Pf - g)^T W (Pf - g) + TV(f) with the p = 1 Huber TV of
    TVgradient / TVquadForm, Polak-Ribiere directions restarted every 50 iterations or when the direction
    is not a descent direction, the step <d, grad> / (<Pd, Pd>_W + <d, R'' d>), and the loop ending at a
    step <= 0.

    W: None (the default weights above), a list of planogram weights, or a scalar.  As in the reference a
    scalar weight leaves the gradient unweighted and only scales <Pd, Pd> in the step.  sqs: the separable
    quadratic surrogate preconditioner 1 / P*(W P 1) (0 -> 1).  response: planogram arrays that the
    projectors include (project -> R * Pf, backproject(x) -> P*(R x)).  Without the nonnegativity
    constraint the reference updates the residual as r <- W r - step Pd - g (its Pf_minus_g aliases Pf and
    was weighted in place); that update is restated as it is.  Returns (f, iterations run)."""
    dt = np.float64
    nP = len(plans)

    def P(x):
        q = project(plans, vol, x, dtype=dt, nthreads=nthreads)
        return [qi * ri for qi, ri in zip(q, response)] if response is not None else q

    def Pt(y):
        if response is not None:
            y = [yi * ri for yi, ri in zip(y, response)]
        return backproject(plans, vol, y, mode=mode, dtype=dt, nthreads=nthreads)

    def inner(x, y, w=None):
        if isinstance(x, list):
            if w is None:
                return sum(float(np.sum(a * b)) for a, b in zip(x, y))
            if np.isscalar(w):
                return sum(float(np.sum(a * b)) * w for a, b in zip(x, y))
            return sum(float(np.sum(a * b * c)) for a, b, c in zip(x, y, w))
        return float(np.sum(x * y))

    g = [np.asarray(gi, dt) for gi in g]
    if W is None:
        W = rwls_default_weights(g)
    elif not np.isscalar(W):
        W = [np.asarray(wi, dt) for wi in W]
    weighted = isinstance(W, list)
    f = np.zeros(vol.shape, dt) if f is None else np.array(f, dtype=dt)
    if np.any(f):
        if nonneg:
            f[f < 0] = 0
        r = P(f)
    else:
        r = [np.zeros_like(gi) for gi in g]
    r = [a - b for a, b in zip(r, g)]
    Q = None
    if sqs:
        Pd = P(np.ones(vol.shape, dt))
        if weighted:
            Pd = [a * b for a, b in zip(Pd, W)]
        Q = Pt(Pd)
        Q = np.where(Q == 0, 1.0, 1.0 / np.where(Q == 0, 1.0, Q))
    d = np.zeros(vol.shape, dt)
    grad_old = np.zeros(vol.shape, dt)
    gg_old = 0.0
    done = 0
    for n in range(num_iter):
        if weighted:
            r = [a * b for a, b in zip(r, W)]
        grad = Pt(r)
        if beta > 0.0:
            grad = grad + tv_gradient(f, delta, beta, 1.0, dtype=dt)
        u = Q * grad if Q is not None else grad.copy()
        if n % 50 == 0:
            d = u.copy()
        else:
            gamma = (inner(u, grad) - inner(u, grad_old)) / gg_old
            d = u + gamma * d
            if inner(d, grad) <= 0.0:
                d = u.copy()
        gg_old = inner(u, grad)
        grad_old = grad.copy()
        Pd = P(d)
        num = inner(d, grad)
        den = inner(Pd, Pd, W) + (tv_quadform(f, d, delta, beta, 1.0) if beta > 0.0 else 0.0)
        step = num / den if abs(den) > 1.0e-16 else 0.0
        if step <= 0.0:
            break
        f = f - step * d
        if nonneg:
            f[f < 0] = 0
            r = [a - b for a, b in zip(P(f), g)]
        else:
            r = [a - step * b - c for a, b, c in zip(r, Pd, g)]
        done = n + 1
        if callback is not None:
            callback(n, f)
    return f, done


# --------------------------------------------------------------------------------------------------
# response model: geometric solids, ray tracing, stopping power, solid angle
# --------------------------------------------------------------------------------------------------
def solid(type, c, r, val, A=None, clip=None):
    """type 0 ellipsoid, 1 parallelepiped, 2/3/4 cylinder x/y/z, 5/6/7 cone x/y/z; c, r: 3 floats; A: 3x3
    rotation (row-major); clip: 3 floats (non-zero entries clip the solid to coordinate <= value)."""
    out = Solid()
    cc = np.ascontiguousarray(c, np.float32)
    rr = np.ascontiguousarray(r, np.float32)
    Aa = None if A is None else np.ascontiguousarray(A, np.float32).reshape(9)
    cl = None if clip is None else np.ascontiguousarray(clip, np.float32)
    ok = lib().orc_solid_init(int(type), cc, rr, float(val), None if Aa is None else Aa.ctypes.data,
                              None if cl is None else cl.ctypes.data, ctypes.byref(out))
    if not ok:
        raise ValueError("invalid solid")
    return out


def _solid_array(solids):
    arr = (Solid * max(len(solids), 1))()
    for i, s in enumerate(solids):
        arr[i] = s
    return arr


def intersect(s, p, r):
    ts = np.zeros(2, np.float32)
    hit = lib().orc_intersect(ctypes.byref(s), np.asarray(p, np.float32), np.asarray(r, np.float32), ts)
    return bool(hit), float(ts[0]), float(ts[1])


def ray_value(solids, p, r, stopping=False):
    arr = _solid_array(solids)
    return float(lib().orc_ray_value(arr, len(solids), np.asarray(p, np.float32), np.asarray(r, np.float32),
                                     int(bool(stopping))))


def ray_trace(plan, solids, oversampling=1, stopping=False, nthreads=None):
    """Line integrals (stopping=False) or stopping powers (True) of every bin of one planogram."""
    g = np.zeros(plan.shape, np.float32)
    arr = _solid_array(solids)
    lib().orc_ray_trace(ctypes.byref(plan), arr, len(solids), int(oversampling), int(bool(stopping)), g,
                        _nt(nthreads))
    return g


def solid_angle_factor(plan, do_inverse=False):
    """planogram::solid_angle_correction_factor per (v1, v0) (src/planogram.cpp:262-265): 2 pi (1 + v0^2 +
    v1^2)^1.5 / (T_v1 T_v0), or its reciprocal.  Returns float32 (N_v1, N_v0)."""
    v1 = (np.arange(plan.N_v1, dtype=np.float32) * np.float32(plan.T_v1) + np.float32(plan.v1_0)).astype(np.float32)
    v0 = (np.arange(plan.N_v0, dtype=np.float32) * np.float32(plan.T_v0) + np.float32(plan.v0_0)).astype(np.float32)
    V1, V0 = np.meshgrid(v1, v0, indexing="ij")
    base = 1.0 + (V0 * V0).astype(np.float64) + (V1 * V1).astype(np.float64)
    c = (2.0 * 3.1415926535897932385 * base ** 1.5 / float(np.float32(plan.T_v1) * np.float32(plan.T_v0)))
    c = c.astype(np.float32)
    if do_inverse:
        c = (1.0 / c.astype(np.float64)).astype(np.float32)
    return c


# --------------------------------------------------------------------------------------------------
# analytic reconstruction: PFDR rebinning, ramp filter, FBP (src/pfdr.cu, src/ramp_filter.cu,
# src/petl.cpp:310-385), in FP64 with numpy's FFT
# --------------------------------------------------------------------------------------------------
def optimal_fft_size(N):
    """optimalFFTsize (src/vector_ops.cpp:481-505): the smallest 2^(n+1) 3^m >= N, from the reference's
    own search (which can return a larger value than the true minimum; its choice is what counts)."""
    import math
    if N <= 2:
        return 2
    c1 = math.log2(N / 2.0) / math.log2(3)
    c2 = 1.0 / math.log2(3)
    xbar = math.log2(N / 2.0)
    best = int(2 ** (int(math.ceil(xbar)) + 1))
    for x in range(int(math.ceil(xbar))):
        y = int(math.ceil(c1 - c2 * x))
        v = int(2 ** (x + 1) * 3 ** y)
        if v < best and y >= 0:
            best = v
    return best


def _fgrid(n, T0, start):
    return (np.arange(n, dtype=np.float32) * np.float32(T0) + np.float32(start)).astype(np.float64)


def pfdr(plan, g):
    """PFDR of one planogram (src/pfdr.cu:159-315): per v1 row the views are weighted by
    1/sqrt(1 + v1^2 + v0^2), Fourier transformed over (v0, u0) (zero-padded to optimal sizes), and added into
    the rebinned spectrum when |v1| <= T_v1/2 or |U| > |v1| / (4 T_u1 v_m0); the sum is divided by the number
    of rows added (except at U = 0) and by the transform size, transformed back and weighted by sqrt(1 + v0^2).
    As in the reference, a contributing oblique row is read at its own u1 (the kernel forms the shifted u1
    but indexes with the unshifted one), interpolated with the fractional part that float arithmetic leaves.
    Returns [N_v0][N_u1][N_u0]."""
    g = np.asarray(g, np.float64)
    Nv1, Nv0, Nu1, Nu0 = plan.shape
    NX, NY = optimal_fft_size(Nu0), optimal_fft_size(Nv0)
    NXH = NX // 2 + 1
    T_U = np.float32(1.0 / (NX * plan.T_u0))
    T_V = np.float32(1.0 / (NY * plan.T_v0))
    ac = np.float32(1.0 / (2.0 * plan.T_u1 * 2.0 * plan.v_m0))
    U = np.where(np.arange(NXH) < NX // 2, np.arange(NXH), np.arange(NXH) - NX).astype(np.float32) * T_U
    v0 = _fgrid(Nv0, plan.T_v0, plan.v0_0)
    u1 = (np.arange(Nu1, dtype=np.float32) * np.float32(plan.T_u1) + np.float32(plan.u1_0)).astype(np.float32)
    # the row the kernel interpolates at: (u1 - u1_0) / T_u1 in float
    iu1o = ((u1 - np.float32(plan.u1_0)) / np.float32(plan.T_u1)).astype(np.float32)
    valid = (iu1o >= 0) & (iu1o <= np.float32(Nu1 - 1))
    lo = np.clip(iu1o.astype(np.int64), 0, Nu1 - 1)
    hi = np.minimum(Nu1 - 1, lo + 1)
    d = (iu1o - lo.astype(np.float32)).astype(np.float64)
    Greb = np.zeros((Nu1, NY, NXH), np.complex128)
    wts = np.zeros((Nu1, NY, NXH), np.float64)
    T_v1 = np.float32(plan.T_v1)
    for iv1 in range(Nv1):
        v1 = float(np.float32(iv1) * np.float32(plan.T_v1) + np.float32(plan.v1_0))
        buf = np.zeros((Nu1, NY, NX))
        w = 1.0 / np.sqrt(1.0 + v1 * v1 + v0 * v0)
        buf[:, :Nv0, :Nu0] = np.transpose(g[iv1] * w[:, None, None], (1, 0, 2))
        G = np.fft.rfft2(buf)
        if abs(v1) <= 0.5 * T_v1:
            Greb += G
            wts += 1.0
        else:
            m = np.abs(U) > abs(v1) * float(ac)
            Gi = (1.0 - d)[:, None, None] * G[lo] + d[:, None, None] * G[hi]
            sel = valid[:, None, None] & m[None, None, :]
            Greb += np.where(sel, Gi, 0)
            wts += sel
    scale = np.full(wts.shape, 1.0 / (NY * NX))
    nz = (wts > 0) & (np.arange(NXH) != 0)[None, None, :]
    scale[nz] /= wts[nz]
    real = np.fft.irfft2(Greb * scale, s=(NY, NX)) * (NY * NX)
    out = np.transpose(real[:, :Nv0, :Nu0], (1, 0, 2)) * np.sqrt(1.0 + v0 * v0)[:, None, None]
    return out


def ramp_kernel(N, T):
    """rampImpulseResponse (src/ramp_filter.cu:334-349): h[i] = 1 / (T pi (1/4 - s^2)), s the signed index."""
    i = np.arange(N)
    s = np.where(i < N // 2, i, i - N).astype(np.float64)
    return 1.0 / (float(np.float32(T)) * np.pi * (0.25 - s * s))


def ramp_filter(g_reb, T_u0):
    """rampFilter1D (src/ramp_filter.cu:156-332) of rebinned data [N_v0][N_u1][N_u0] along u0: rows
    zero-padded to optimalFFTsize(2 N_u0), multiplied in frequency by Re(FFT(h)) / N_H."""
    g_reb = np.asarray(g_reb, np.float64)
    Nu0 = g_reb.shape[-1]
    NH = optimal_fft_size(2 * Nu0)
    H = np.real(np.fft.rfft(ramp_kernel(NH, T_u0))) / NH
    pad = np.zeros(g_reb.shape[:-1] + (NH,))
    pad[..., :Nu0] = g_reb
    return np.fft.irfft(np.fft.rfft(pad, axis=-1) * H, n=NH, axis=-1)[..., :Nu0] * NH


def rebinned_plan(plan):
    """The planogram PFDR leaves behind (src/petl.cpp:378-379): one v1 row at v1 = 0."""
    import copy
    p = copy.copy(plan)
    p.N_v1 = 1
    p.v1_0 = 0.0
    return p


def fbp(plans, vol, g, nthreads=None):
    """PETL::doFBP (src/petl.cpp:310-355): PFDR when N_v1 > 1, the ramp filter, the inverse planogram weight
    sqrt(1 + v0^2) (v1 = 0 after rebinning) and the PETL back projection of the rebinned planograms."""
    reb, rplans = [], []
    for p, gi in zip(plans, g):
        gr = pfdr(p, gi) if p.N_v1 > 1 else np.asarray(gi, np.float64).reshape(p.shape)[0]
        rp = rebinned_plan(p)
        v0 = _fgrid(p.N_v0, p.T_v0, p.v0_0)
        gr = ramp_filter(gr, p.T_u0) * np.sqrt(1.0 + v0 * v0)[:, None, None]
        reb.append(gr[None].astype(np.float64))
        rplans.append(rp)
    return backproject(rplans, vol, reb, mode="petl", dtype=np.float64, nthreads=nthreads)


def osem(plans, vol, g, num_iter, num_subsets, f0=None, nthreads=None):
    """OSEM with interleaved v1-row subsets (rows s, s+n, ...), fp32, PETL-mode back projection; one
    iteration = one pass over the subsets.  f0: start image (default ones).  With one subset and the
    default start this is mlem()."""
    L = lib()
    f = np.ones(vol.shape, np.float32) if f0 is None else np.array(f0, dtype=np.float32)
    S = []
    for sub in range(num_subsets):
        Ss = np.zeros(vol.shape, np.float32)
        for p in plans:
            L.orc_backproject_petl_subset_f32(ctypes.byref(p), ctypes.byref(vol), np.ones(p.shape, np.float32), Ss,
                                              sub, num_subsets, _nt(nthreads))
        Ss[Ss == 0] = 1
        S.append(Ss)
    for _ in range(num_iter):
        for sub in range(num_subsets):
            d = np.zeros(vol.shape, np.float32)
            for p, gi in zip(plans, g):
                q = np.zeros(p.shape, np.float32)
                L.orc_project_subset_f32(ctypes.byref(p), ctypes.byref(vol), f, q, sub, num_subsets, _nt(nthreads))
                r = rdivide(q, np.asarray(gi, np.float32))
                L.orc_backproject_petl_subset_f32(ctypes.byref(p), ctypes.byref(vol), r, d, sub, num_subsets,
                                                  _nt(nthreads))
            f = f * divide(d, S[sub])
    return f


# --------------------------------------------------------------------------------------------------
# list-mode binning (src/binning.cpp) and the two list-mode file layouts it reads
# --------------------------------------------------------------------------------------------------
def bin_events(plans, blue, pink, nearest=False, chest_wall=True, g=None):
    """Bins coincidences (blue, pink: (n, 3) float32 end points in mm) into the planograms, adding to g
    (zeros when None).  Returns (g, accepted counts per planogram)."""
    blue = np.ascontiguousarray(blue, dtype=np.float32)
    pink = np.ascontiguousarray(pink, dtype=np.float32)
    if g is None:
        g = [np.zeros(p.shape, np.float32) for p in plans]
    acc = []
    for p, gi in zip(plans, g):
        acc.append(int(lib().orc_bin_events(ctypes.byref(p), blue, pink, len(blue), int(bool(nearest)),
                                            int(bool(chest_wall)), gi)))
    return g, acc


def event_coords(plan, blue, pink):
    """(accepted, [v1, v0, u1, u0]) of one coincidence (u from the blue end point, no origin shift)."""
    out = np.zeros(4, np.float32)
    ok = lib().orc_event_coords(ctypes.byref(plan), np.asarray(blue, np.float32), np.asarray(pink, np.float32), out)
    return bool(ok), out


# .lst: records of two events {float pos[4] (x, y, z, energy); uint64 time} (src/binning.cpp:9-19, 105-114)
LST_EVENT = np.dtype([("pos", "<f4", (4,)), ("time", "<u8")])
LST_RECORD = np.dtype([("blue", LST_EVENT), ("pink", LST_EVENT)])


def write_lst(path, blue, pink, energy=511.0):
    rec = np.zeros(len(blue), LST_RECORD)
    rec["blue"]["pos"][:, :3] = blue
    rec["pink"]["pos"][:, :3] = pink
    rec["blue"]["pos"][:, 3] = energy
    rec["pink"]["pos"][:, 3] = energy
    rec["blue"]["time"] = np.arange(len(blue))
    rec["pink"]["time"] = np.arange(len(blue))
    rec.tofile(path)


def read_lst(path):
    rec = np.fromfile(path, LST_RECORD)
    return rec["blue"]["pos"][:, :3].copy(), rec["pink"]["pos"][:, :3].copy()


def write_acq(path, blue, pink, extra_interactions=None, energy=511.0, seed=0):
    """Acquisition layout of src/binning.cpp:598-617: int64 N, then per coincidence for blue and pink:
    int16 number of interactions, float pos[3], float energy, then 18 bytes per further interaction
    (skipped by the reader).  extra_interactions: (n, 2) counts of further interactions."""
    rng = np.random.default_rng(seed)
    n = len(blue)
    if extra_interactions is None:
        extra_interactions = np.zeros((n, 2), np.int64)
    parts = [np.array([n], "<i8").tobytes()]
    for i in range(n):
        for side, pos in ((0, blue[i]), (1, pink[i])):
            k = int(extra_interactions[i, side])
            parts.append(np.array([k + 1], "<i2").tobytes())
            parts.append(np.asarray(pos, "<f4").tobytes())
            parts.append(np.array([energy], "<f4").tobytes())
            parts.append(rng.integers(0, 255, size=18 * k, dtype=np.uint8).tobytes())
    with open(path, "wb") as fh:
        fh.write(b"".join(parts))


def read_acq(path, units_cm=False):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:8], "<i8")[0])
    off = 8
    blue = np.zeros((n, 3), np.float32)
    pink = np.zeros((n, 3), np.float32)
    for i in range(n):
        for out in (blue, pink):
            k = int(np.frombuffer(raw[off:off + 2], "<i2")[0])
            out[i] = np.frombuffer(raw[off + 2:off + 14], "<f4")
            off += 18 + 18 * max(k - 1, 0)
    if units_cm:  # bluePos[i] *= 10.0: a float times a double, stored as float
        blue = (blue.astype(np.float64) * 10.0).astype(np.float32)
        pink = (pink.astype(np.float64) * 10.0).astype(np.float32)
    return blue, pink


# --------------------------------------------------------------------------------------------------
# the reference's own geometry / dataCube (only where the reference sources could be compiled)
# --------------------------------------------------------------------------------------------------
_ref = None


def ref_lib():
    """Returns the reference-built helper library, or None when it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(_REF):
            return None
        R = ctypes.CDLL(_REF)
        c_int, c_float = ctypes.c_int, ctypes.c_float
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        R.ref_plan_init.argtypes = [c_float] * 7 + [np.ctypeslib.ndpointer(np.int32), fp]
        R.ref_plan_init.restype = c_int
        R.ref_binary3d.argtypes = [fp, fp, c_int, c_int, c_int, c_int]
        R.ref_binary3d.restype = c_int
        R.ref_reciprocal3d.argtypes = [fp, c_int, c_int, c_int, c_float]
        R.ref_inner3d.argtypes = [fp, fp, ctypes.c_void_p, c_int, c_int, c_int]
        R.ref_inner3d.restype = c_float
        R.ref_event_coords.argtypes = [c_float] * 7 + [fp, fp, fp]
        R.ref_event_coords.restype = c_int
        R.ref_solid_angle.argtypes = [c_float] * 7 + [c_int, fp]
        R.ref_solid_angle.restype = c_int
        R.ref_optimal_fft_size.argtypes = [c_int]
        R.ref_optimal_fft_size.restype = c_int
        _ref = R
    return _ref
