"""Pins the oracle's MLEM loop (restatement of src/petl_server.py:1534-1606) by invariants of the
algorithm, using the exactly-adjoint (matched) back projector where the invariant needs A^T:

* non-negativity;
* count preservation: sum_j S_j f_j^(n+1) = sum_i y_i over bins with (A f^n)_i > 0, S = A^T 1;
* the Poisson log-likelihood sum(y log Af - Af) is non-decreasing;
* noise-free consistent data y = A f is a fixed point when started at f;
* with the relative-difference prior the OSL update equals the hand-written update formula.
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


def test_osl_prior_update_formula(oracle):
    plans, vol, _, y = setup(oracle, psis=(0.0,))
    f = np.random.default_rng(31).random(vol.shape) + 0.5
    beta, delta = 0.05, 2.0
    got = oracle.mlem(plans, vol, y, 1, f=f, delta=delta, beta=beta, dtype=np.float64)
    S = oracle.backproject(plans, vol, [np.ones(p.shape) for p in plans], dtype=np.float64)
    S[S == 0] = 1
    Pd = oracle.project(plans, vol, f, dtype=np.float64)
    d = oracle.backproject(plans, vol, [np.where(a == 0, 0, yy / np.where(a == 0, 1, a)) for a, yy in zip(Pd, y)],
                           dtype=np.float64)
    G = np.maximum(0.1 * S, oracle.rd_gradient(f, delta, beta, dtype=np.float64) + S)
    np.testing.assert_allclose(got, f * d / G, rtol=1e-12)
