"""Pins the oracle's MLEM loop (restatement of src/petl_server.py:1534-1606) by invariants of the
algorithm, using the exactly-adjoint (matched) back projector where the invariant needs A^T:

* non-negativity;
* count preservation: sum_j S_j f_j^(n+1) = sum_i y_i over bins with (A f^n)_i > 0, S = A^T 1;
* the Poisson log-likelihood sum(y log Af - Af) is non-decreasing;
* noise-free consistent data y = A f is a fixed point when started at f;
* with the relative-difference prior (one-step-late): a constant image takes the plain step, a hot voxel
  is pulled down and its neighbours up, and the 0.1 S clamp of the denominator acts where S + grad R
  falls below it.
"""
import numpy as np
import pytest

GEO = dict(D=60.0, L=32.0, H=32.0, T=2.0)


def setup(oracle, psis=(0.0, 90.0)):
    plans = [oracle.planogram(psi, **GEO) for psi in psis]
    vol = oracle.volume(10, 10, 10, 2.0)
    rng = np.random.default_rng(30)
    f_true = rng.random(vol.shape) + 0.2
    g = oracle.project(plans, vol, f_true, dtype=np.float64)
    y = [rng.poisson(50.0 * gi).astype(np.float64) for gi in g]
    return plans, vol, f_true, y


def loglik(Af, y):
    return sum(float(np.sum(np.where(a > 0, yy * np.log(np.where(a > 0, a, 1.0)) - a, 0.0))) for a, yy in zip(Af, y))


def test_mlem_invariants(oracle):
    plans, vol, _, y = setup(oracle)
    S = oracle.backproject(plans, vol, [np.ones(p.shape) for p in plans], mode="matched", dtype=np.float64)
    S[S == 0] = 1
    lls, checks = [], []

    def cb(n, f, Pd):
        lls.append(loglik(Pd, y))

    f = np.ones(vol.shape)
    for it in range(6):
        Pd = oracle.project(plans, vol, f, dtype=np.float64)
        lls.append(loglik(Pd, y))
        f_next = oracle.mlem(plans, vol, y, 1, f=f, dtype=np.float64, mode="matched")
        assert np.all(f_next >= 0)
        counts = sum(float(yy[a > 0].sum()) for a, yy in zip(Pd, y))
        assert float(np.sum(S * f_next)) == pytest.approx(counts, rel=1e-10)
        f = f_next
    assert all(b >= a - 1e-9 * abs(a) for a, b in zip(lls, lls[1:]))


def test_consistent_data_fixed_point(oracle):
    plans, vol, f_true, _ = setup(oracle)
    y = oracle.project(plans, vol, f_true, dtype=np.float64)
    f = oracle.mlem(plans, vol, y, 3, f=f_true, dtype=np.float64, mode="matched")
    S = oracle.backproject(plans, vol, [np.ones(p.shape) for p in plans], mode="matched", dtype=np.float64)
    sel = S > 0
    np.testing.assert_allclose(f[sel], f_true[sel], rtol=1e-10)


def test_initial_estimate_rules(oracle):
    plans, vol, _, y = setup(oracle, psis=(0.0,))
    f0 = oracle.mlem(plans, vol, y, 1, f=np.zeros(vol.shape), dtype=np.float64)
    f1 = oracle.mlem(plans, vol, y, 1, f=None, dtype=np.float64)
    np.testing.assert_array_equal(f0, f1)
    neg = np.ones(vol.shape)
    neg[0, 0, 0] = -3.0
    fn = oracle.mlem(plans, vol, y, 1, f=neg, dtype=np.float64)
    assert fn[0, 0, 0] == 0.0


def test_osl_constant_image_is_the_plain_step(oracle):
    """A constant image has a zero relative-difference gradient, so the OSL step is the beta = 0 step."""
    plans, vol, _, y = setup(oracle, psis=(0.0,))
    f = np.full(vol.shape, 1.7)
    a = oracle.mlem(plans, vol, y, 1, f=f, delta=2.0, beta=0.5, dtype=np.float64)
    b = oracle.mlem(plans, vol, y, 1, f=f, beta=0.0, dtype=np.float64)
    np.testing.assert_array_equal(a, b)


def test_osl_hot_voxel_is_pulled_down(oracle):
    """The prior penalises differences: a hot voxel's update ends below its beta = 0 update and its face
    neighbours end above theirs (the sign of the gradient in the OSL denominator)."""
    plans, vol, _, y = setup(oracle, psis=(0.0,))
    f = np.ones(vol.shape)
    f[5, 5, 5] = 6.0
    a = oracle.mlem(plans, vol, y, 1, f=f, delta=2.0, beta=0.05, dtype=np.float64)
    b = oracle.mlem(plans, vol, y, 1, f=f, beta=0.0, dtype=np.float64)
    assert a[5, 5, 5] < b[5, 5, 5]
    for nb in ((4, 5, 5), (6, 5, 5), (5, 4, 5), (5, 6, 5), (5, 5, 4), (5, 5, 6)):
        assert a[nb] > b[nb], nb
    far = np.ones(vol.shape, bool)
    far[3:8, 3:8, 3:8] = False
    np.testing.assert_array_equal(a[far], b[far])  # the gradient is zero two voxels away


def test_osl_denominator_clamp(oracle):
    """With a beta large enough that S + grad R < 0.1 S next to a hot voxel, the denominator there is
    0.1 S: those voxels take exactly ten times the beta = 0 update factor."""
    plans, vol, _, y = setup(oracle, psis=(0.0,))
    f = np.ones(vol.shape)
    f[5, 5, 5] = 50.0
    beta = 1.0e4
    a = oracle.mlem(plans, vol, y, 1, f=f, delta=2.0, beta=beta, dtype=np.float64)
    b = oracle.mlem(plans, vol, y, 1, f=f, beta=0.0, dtype=np.float64)
    S = oracle.backproject(plans, vol, [np.ones(p.shape) for p in plans], dtype=np.float64)
    S[S == 0] = 1
    G = oracle.rd_gradient(f, 2.0, beta, dtype=np.float64)
    clamped = S + G < 0.1 * S
    assert clamped.sum() >= 6 and not clamped[5, 5, 5]
    np.testing.assert_allclose(a[clamped], 10.0 * b[clamped], rtol=1e-12)
    free = ~clamped & (G == 0)
    np.testing.assert_array_equal(a[free], b[free])
