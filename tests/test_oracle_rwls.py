"""Pins of the oracle's RWLS pieces (CPU only): the Gaussian smoothing of the default weights, the TV cost
and quadratic form used in the step size, and the conjugate-gradient loop itself.

Each check is independent of the restatement it tests: textbook routines (scipy's Gaussian filter and
preconditioned CG), closed forms on tiny inputs, and identities the mathematics fixes (the TV gradient is
half the gradient of the TV cost; the quadratic form with h''(t) = h'(t)/t satisfies <f, R'' f> =
2 <grad R, f>; with p = 2 the prior is quadratic).
"""
import ctypes

import numpy as np
import pytest
from scipy import ndimage
from scipy.sparse.linalg import LinearOperator, cg

from oracle import oracle as orc

RNG = np.random.default_rng(1234)


def _sigma_radius(fwhm):
    sigma = np.float32(fwhm) / (np.float32(2.0) * np.sqrt(np.float32(2.0) * np.log(np.float32(2.0))))
    radius = int(np.ceil(np.sqrt(np.float32(2.0) * np.log(np.float32(10.0))) * sigma))
    return float(sigma), radius


def _scipy_sd(fwhm):
    """Standard deviation of the equivalent textbook Gaussian: exp(-t^2 / (2 sd^2)) with 1 / (2 sd^2)
    equal to the float exponent factor 1 / (2 sigma)."""
    sigma, _ = _sigma_radius(fwhm)
    den = float(np.float32(1.0) / (np.float32(2.0) * np.float32(sigma)))
    return np.sqrt(1.0 / (2.0 * den))


# ------------------------------------------------------------------------------------------------
# Gaussian smoothing (src/relative_differences.cu:314-420, 753-800)
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_gaussian_constant_is_unchanged(nd):
    f = np.full((5, 6, 7), 3.25)
    np.testing.assert_allclose(orc.gaussian_filter(f, 3.0, nd), f, rtol=1e-14)


@pytest.mark.parametrize("fwhm", [2.0, 3.0, 5.0])
def test_gaussian_interior_matches_scipy(fwhm):
    """Away from the borders the filter is a normalised Gaussian exp(-t^2 / (2 sigma)) truncated at the
    radius: scipy's with standard deviation sqrt(sigma) and the same radius."""
    sigma, R = _sigma_radius(fwhm)
    f = RNG.random((2 * R + 5, 2 * R + 6, 2 * R + 7))
    got = orc.gaussian_filter(f, fwhm, 3)
    want = ndimage.gaussian_filter(f, _scipy_sd(fwhm), radius=R, mode="nearest")
    inner = (slice(R, -R),) * 3
    np.testing.assert_allclose(got[inner], want[inner], rtol=1e-12, atol=1e-14)
    got2 = orc.gaussian_filter(f, fwhm, 2)
    want2 = ndimage.gaussian_filter(f, _scipy_sd(fwhm), radius=R, mode="nearest", axes=(1, 2))
    np.testing.assert_allclose(got2[:, R:-R, R:-R], want2[:, R:-R, R:-R], rtol=1e-12, atol=1e-14)


def test_gaussian_border_two_samples_closed_form():
    """Along an axis of length 2 with radius R, every tap past the border reads the border sample at
    distance 0 (weight 1): out[0] = ((R + 1) a + R e b) / (R + 1 + R e), e = exp(-1 / (2 sigma))."""
    fwhm = 3.0
    sigma, R = _sigma_radius(fwhm)
    e = np.exp(-float(np.float32(1.0) / (np.float32(2.0) * np.float32(sigma))))
    a, b = 2.0, 7.0
    f = np.array([a, b]).reshape(2, 1, 1)
    out = orc.gaussian_filter(f, fwhm, 1).ravel()
    np.testing.assert_allclose(out[0], ((R + 1) * a + R * e * b) / (R + 1 + R * e), rtol=1e-14)
    np.testing.assert_allclose(out[1], ((R + 1) * b + R * e * a) / (R + 1 + R * e), rtol=1e-14)


def test_gaussian_is_separable():
    """The 3D weight is the product of per-axis weights and the borders clamp per axis, so smoothing along
    the three axes one after the other equals the 3D filter (and the 2D filter is the last two axes)."""
    f = RNG.random((6, 9, 8))
    a = orc.gaussian_filter(f, 3.0, 1)
    b = np.moveaxis(orc.gaussian_filter(np.moveaxis(a, 1, 0).copy(), 3.0, 1), 0, 1)
    c = np.moveaxis(orc.gaussian_filter(np.moveaxis(b, 2, 0).copy(), 3.0, 1), 0, 2)
    np.testing.assert_allclose(orc.gaussian_filter(f, 3.0, 3), c, rtol=1e-12)
    bj = np.moveaxis(orc.gaussian_filter(np.moveaxis(f, 1, 0).copy(), 3.0, 1), 0, 1)
    ck = np.moveaxis(orc.gaussian_filter(np.moveaxis(bj, 2, 0).copy(), 3.0, 1), 0, 2)
    np.testing.assert_allclose(orc.gaussian_filter(f, 3.0, 2), ck, rtol=1e-12)


def test_gaussian_planogram_rows():
    """A planogram (4D) is smoothed over (u1, u0) within each (v1, v0) row."""
    g = RNG.random((3, 4, 9, 10))
    got = orc.gaussian_filter(g, 3.0)
    for a in range(3):
        for b in range(4):
            np.testing.assert_allclose(got[a, b], orc.gaussian_filter(g[a, b], 3.0)[...], rtol=1e-14)
    sigma, R = _sigma_radius(3.0)
    want = ndimage.gaussian_filter(g, _scipy_sd(3.0), radius=R, mode="nearest", axes=(2, 3))
    np.testing.assert_allclose(got[:, :, R:-R, R:-R], want[:, :, R:-R, R:-R], rtol=1e-12)


# ------------------------------------------------------------------------------------------------
# TV cost and quadratic form (src/total_variation.cu:80-417, 662-800)
# ------------------------------------------------------------------------------------------------
TV_CASES = [((6, 7, 8), 0.3, 1.0), ((6, 7, 8), 0.05, 1.5), ((1, 9, 8), 0.2, 1.0), ((5, 1, 7), 0.4, 1.2)]


@pytest.mark.parametrize("shape,delta,p", TV_CASES)
def test_tv_cost_gradient_is_twice_tv_gradient(shape, delta, p):
    """Every neighbour pair appears in both voxels' sums (border clamps pair a voxel with itself, h(0) =
    0), so d cost / d f_i = 2 * TVgradient_i."""
    beta = 0.7
    f = RNG.random(shape)
    G = orc.tv_gradient(f, delta, beta, p, dtype=np.float64)
    h = 1e-6
    for idx in [tuple(RNG.integers(0, s) for s in shape) for _ in range(12)]:
        e = np.zeros(shape)
        e[idx] = h
        fd = (orc.tv_cost(f + e, delta, beta, p) - orc.tv_cost(f - e, delta, beta, p)) / (2 * h)
        assert abs(fd - 2 * G[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, 2 * G[idx])


@pytest.mark.parametrize("shape,delta,p", TV_CASES)
def test_tv_quadform_of_f_is_twice_gradient_dot_f(shape, delta, p):
    """With the surrogate curvature h'(t)/t each pair term of <f, R'' f> is t h'(t), twice per pair, and
    <grad, f> collects t h'(t) once per pair."""
    f = RNG.random(shape)
    beta = 0.9
    q = orc.tv_quadform(f, f, delta, beta, p)
    G = orc.tv_gradient(f, delta, beta, p, dtype=np.float64)
    np.testing.assert_allclose(q, 2.0 * float(np.sum(G * f)), rtol=1e-10)


@pytest.mark.parametrize("shape", [(6, 7, 8), (1, 9, 8)])
def test_tv_quadform_quadratic_prior(shape):
    """p = 2: h(t) = t^2 / 2 on both branches, so <d, R'' d> = 2 cost(d), whatever f."""
    f, d = RNG.random(shape) * 5, RNG.standard_normal(shape)
    q = orc.tv_quadform(f, d, 0.1, 1.3, 2.0)
    np.testing.assert_allclose(q, 2.0 * orc.tv_cost(d, 0.1, 1.3, 2.0), rtol=1e-12)
    np.testing.assert_allclose(orc.tv_quadform(2 * f + 1, d, 0.1, 1.3, 2.0), q, rtol=1e-12)


def test_tv_parameter_rules():
    """beta <= 0 counts as 1 and delta is floored at 1e-8 (src/total_variation.cu:666-669)."""
    f, d = RNG.random((4, 5, 6)), RNG.random((4, 5, 6))
    assert orc.tv_cost(f, 0.3, 0.0) == pytest.approx(orc.tv_cost(f, 0.3, 1.0), rel=1e-15)
    assert orc.tv_quadform(f, d, 0.3, -2.0) == pytest.approx(orc.tv_quadform(f, d, 0.3, 1.0), rel=1e-15)
    assert orc.tv_cost(f, 0.0, 1.0) == pytest.approx(orc.tv_cost(f, 1e-8, 1.0), rel=1e-15)


# ------------------------------------------------------------------------------------------------
# the RWLS loop (src/petl_server.py:1691-1877) against textbook preconditioned CG
# ------------------------------------------------------------------------------------------------
def _small_system():
    plans = [orc.planogram(0.0, 40.0, 16.0, 16.0, 2.0), orc.planogram(90.0, 40.0, 16.0, 16.0, 2.0)]
    vol = orc.volume(8, 8, 6, 2.0)
    nvox = int(np.prod(vol.shape))
    cols = []
    for j in range(nvox):
        e = np.zeros(nvox)
        e[j] = 1.0
        cols.append(np.concatenate([q.ravel() for q in orc.project(plans, vol, e.reshape(vol.shape),
                                                                   dtype=np.float64)]))
    A = np.stack(cols, axis=1)
    return plans, vol, A


def _split(plans, y):
    out, o = [], 0
    for p in plans:
        n = int(np.prod(p.shape))
        out.append(y[o:o + n].reshape(p.shape))
        o += n
    return out


def _pcg(H, b, k, M=None):
    n = H.shape[0]
    Hop = LinearOperator((n, n), matvec=lambda x: H @ x)
    Mop = None if M is None else LinearOperator((n, n), matvec=lambda x: M * x)
    xs = []
    for it in range(1, k + 1):
        x, _ = cg(Hop, b, x0=np.zeros(n), rtol=0.0, atol=0.0, maxiter=it, M=Mop)
        xs.append(x)
    return xs


@pytest.fixture(scope="module")
def system():
    return _small_system()


@pytest.mark.parametrize("case", ["ls", "wls_sqs", "response_wls"])
def test_rwls_is_preconditioned_cg(system, case):
    """Without a prior the loop is CG on A^T W A f = A^T W g (Polak-Ribiere = Fletcher-Reeves with exact
    line search on a quadratic), started at 0; with SQS it is PCG with M = 1 / (A^T W A 1).  The data are
    consistent and positive enough that no iterate is clipped, so the nonnegativity step is idle."""
    plans, vol, A = system
    rng = np.random.default_rng(7)
    nvox = A.shape[1]
    seen = A.sum(axis=0) > 0
    x_true = np.where(seen, 4.0 + rng.random(nvox), 0.0)
    if case == "response_wls":
        resp = rng.uniform(0.5, 1.5, A.shape[0])
        Ause = resp[:, None] * A
    else:
        resp = None
        Ause = A
    y = Ause @ x_true
    if case == "ls":
        w, W = np.ones(A.shape[0]), 1.0
    else:
        w = rng.uniform(0.2, 2.0, A.shape[0])
        W = _split(plans, w)
    sqs = case != "ls"
    H = Ause.T @ (w[:, None] * Ause)
    b = Ause.T @ (w * y)
    M = None
    if sqs:
        q = Ause.T @ (w * (Ause @ np.ones(nvox)))
        M = np.where(q == 0, 1.0, 1.0 / np.where(q == 0, 1.0, q))
    k = 5
    want = _pcg(H, b, k, M)
    assert all(np.all(x >= 0) for x in want)
    g = _split(plans, y)
    got = []
    f, done = orc.rwls(plans, vol, g, k, W=W, sqs=sqs, nonneg=True, mode="matched",
                       response=None if resp is None else _split(plans, resp),
                       callback=lambda n, fi: got.append(fi.ravel().copy()))
    assert done == k
    for a, e in zip(got, want):
        np.testing.assert_allclose(a, e, rtol=1e-9, atol=1e-9 * np.abs(e).max())


def test_rwls_zero_data_stops_at_once(system):
    """g = 0 from f = 0: the gradient is 0, the step 0, and the loop ends before changing f."""
    plans, vol, _ = system
    g = [np.zeros(p.shape) for p in plans]
    f, done = orc.rwls(plans, vol, g, 4, W=1.0)
    assert done == 0 and not np.any(f)


def test_rwls_default_weights(system):
    plans, _, _ = system
    g = [RNG.poisson(20.0, p.shape).astype(np.float64) for p in plans]
    W = orc.rwls_default_weights(g)
    for wi, gi in zip(W, g):
        np.testing.assert_allclose(wi, 1.0 / np.maximum(orc.gaussian_filter(gi, 3.0), 7.0), rtol=0)
        assert np.all(wi <= 1.0 / 7.0)


# ------------------------------------------------------------------------------------------------
# the TV terms of the RWLS loop in the Huber quadratic regime
# ------------------------------------------------------------------------------------------------
def _shift_operators(shape):
    """The 26 neighbour shifts of a [N1][N2][N3] grid as dense matrices S_o (row c picks the voxel at
    c + o with every index clamped into the grid) with weights 1 / |o| (1, 1/sqrt 2, 1/sqrt 3)."""
    n = int(np.prod(shape))
    idx = np.arange(n).reshape(shape)
    ops = []
    for o in np.ndindex(3, 3, 3):
        o = np.array(o) - 1
        if not o.any():
            continue
        src = [np.clip(np.arange(shape[a]) + o[a], 0, shape[a] - 1) for a in range(3)]
        nb = idx[np.ix_(*src)].ravel()
        S = np.zeros((n, n))
        S[np.arange(n), nb] = 1.0
        ops.append((1.0 / np.sqrt(float(o @ o)), S))
    return ops


def test_tv_quadratic_regime_is_weighted_laplacian():
    """With delta above every neighbour difference h'(t) = t and h''(t) = 1, so the TV gradient is
    beta * sum_o w_o (I - S_o) f and the quadratic form is beta * sum_o w_o |(I - S_o) d|^2, whatever f."""
    shape = (3, 4, 5)
    f, d = RNG.random(shape), RNG.standard_normal(shape)
    beta, delta = 0.75, 10.0
    n = f.size
    L = sum(w * (np.eye(n) - S) for w, S in _shift_operators(shape))
    # the reference keeps the diagonal weights as float constants (relative rounding <= 6e-8)
    want = beta * L @ f.ravel()
    np.testing.assert_allclose(orc.tv_gradient(f, delta, beta, dtype=np.float64).ravel(), want,
                               rtol=2e-7, atol=2e-7 * np.abs(want).max())
    q = sum(w * float(np.sum(((np.eye(n) - S) @ d.ravel()) ** 2)) for w, S in _shift_operators(shape))
    assert orc.tv_quadform(f, d, delta, beta) == pytest.approx(beta * q, rel=2e-7)


def test_rwls_tv_quadratic_regime_iterates(system):
    """RWLS with the TV prior in the quadratic regime, against the same Polak-Ribiere iteration written
    with dense matrices: gradient A^T W (A f - g) + beta L f and step <d, grad> / (<Ad, Ad>_W + beta
    sum_o w_o |(I - S_o) d|^2), with L and S_o built from the neighbour definition above.  Without the prior
    the loop is pinned against scipy's CG; this pins how the loop adds the prior's gradient (at f, with a
    plus sign) and its curvature along d (the reference's step uses h'' with the pair counted from both
    voxels, twice the curvature of the gradient's own quadratic, so this is not CG on one quadratic)."""
    plans, vol, A = system
    rng = np.random.default_rng(8)
    nvox = A.shape[1]
    seen = A.sum(axis=0) > 0
    x_true = np.where(seen, 4.0 + rng.random(nvox), 0.0)
    y = A @ x_true
    w = rng.uniform(0.2, 2.0, A.shape[0])
    beta, delta = 0.25, 1.0e3
    ops = _shift_operators(vol.shape)
    L = sum(wo * (np.eye(nvox) - S) for wo, S in ops)
    Dq = sum(wo * (np.eye(nvox) - S).T @ (np.eye(nvox) - S) for wo, S in ops)
    q = A.T @ (w * (A @ np.ones(nvox)))
    M = np.where(q == 0, 1.0, 1.0 / np.where(q == 0, 1.0, q))
    # the loop of src/petl_server.py:1786-1830 on dense matrices (no iterate is clipped here)
    f = np.zeros(nvox)
    r = A @ f - y
    d = np.zeros(nvox)
    grad_old = np.zeros(nvox)
    gg_old = 0.0
    want = []
    for n in range(4):
        grad = A.T @ (w * r) + beta * L @ f
        u = M * grad
        if n == 0:
            d = u.copy()
        else:
            gamma = (u @ grad - u @ grad_old) / gg_old
            d = u + gamma * d
            if d @ grad <= 0.0:
                d = u.copy()
        gg_old = u @ grad
        grad_old = grad.copy()
        Ad = A @ d
        step = (d @ grad) / (Ad @ (w * Ad) + beta * d @ Dq @ d)
        f = f - step * d
        assert np.all(f >= 0)
        r = A @ f - y
        want.append(f.copy())
    got = []
    orc.rwls(plans, vol, _split(plans, y), 4, delta=delta, beta=beta, W=_split(plans, w), sqs=True, nonneg=True,
             mode="matched", callback=lambda n, fi: got.append(fi.ravel().copy()))
    assert len(got) == 4
    for a, e in zip(got, want):
        np.testing.assert_allclose(a, e, rtol=1e-6, atol=1e-6 * np.abs(e).max())
    # the prior changes the iterates (the check above is not the beta = 0 path)
    f0, _ = orc.rwls(plans, vol, _split(plans, y), 4, W=_split(plans, w), sqs=True, mode="matched")
    assert np.max(np.abs(f0.ravel() - want[-1])) > 1e-3 * np.abs(want[-1]).max()
