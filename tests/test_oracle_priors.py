"""Pins the oracle's prior gradients (src/relative_differences.cu:21-28,107-173;
src/total_variation.cu:33-78,434-660).

The reference gradient kernels return half the gradient of their own cost sums (each unordered
neighbour pair appears twice in the cost, once in the gradient sum), so central finite differences are
taken of cost/2, with the cost written independently here (correct indexing, unlike the cost kernel at
src/relative_differences.cu:77).  A [1][1][2] array has a closed form: the only distinct neighbour of
voxel 0 is voxel 1, reached through 1 face, 4 edge and 4 corner offsets under border clamping.
"""
import numpy as np
import pytest


def cost_rd(f, delta, beta):
    f = np.where(f < 1e-8, 0.0, f)
    N1, N2, N3 = f.shape
    acc = 0.0
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                m = abs(di) + abs(dj) + abs(dk)
                if m == 0:
                    continue
                w = {1: 1.0, 2: float(np.float32(0.7071067811865475)), 3: float(np.float32(0.5773502691896258))}[m]
                ii = np.clip(np.arange(N1) + di, 0, N1 - 1)
                jj = np.clip(np.arange(N2) + dj, 0, N2 - 1)
                kk = np.clip(np.arange(N3) + dk, 0, N3 - 1)
                y = f[ii][:, jj][:, :, kk]
                x = f
                with np.errstate(invalid="ignore", divide="ignore"):
                    phi = np.where(x == y, 0.0, (x - y) ** 2 / ((x + y) + delta * np.abs(x - y)))
                acc += w * phi.sum()
    return beta * acc


def test_rd_cost_restatements_agree(oracle):
    f = np.random.default_rng(20).random((4, 5, 6)) + 0.1
    assert oracle.rd_cost(f, 2.0, 0.3) == pytest.approx(cost_rd(f, 2.0, 0.3), rel=1e-12)


@pytest.mark.parametrize("shape", [(5, 6, 7), (1, 4, 5), (3, 1, 4)])
def test_rd_gradient_is_half_cost_gradient(oracle, shape):
    rng = np.random.default_rng(21)
    f = rng.random(shape) + 0.2
    delta, beta = 2.0, 0.7
    G = oracle.rd_gradient(f, delta, beta, dtype=np.float64)
    h = 1e-6
    for idx in [tuple(rng.integers(0, s) for s in shape) for _ in range(12)]:
        fp, fm = f.copy(), f.copy()
        fp[idx] += h
        fm[idx] -= h
        fd = (oracle.rd_cost(fp, delta, beta) - oracle.rd_cost(fm, delta, beta)) / (2 * h)
        assert G[idx] == pytest.approx(0.5 * fd, rel=1e-6, abs=1e-9)


def test_rd_gradient_two_voxel_closed_form(oracle):
    a, b, delta, beta = 1.3, 0.4, 2.0, 0.5
    G = oracle.rd_gradient(np.array([[[a, b]]]), delta, beta, dtype=np.float64)
    D = lambda x, y: (x - y) * (delta * abs(x - y) + x + 3 * y) / (x + y + delta * abs(x - y)) ** 2
    w = 1.0 + 4 * float(np.float32(0.7071067811865475)) + 4 * float(np.float32(0.5773502691896258))
    assert G[0, 0, 0] == pytest.approx(beta * w * D(a, b), rel=1e-12)
    assert G[0, 0, 1] == pytest.approx(beta * w * D(b, a), rel=1e-12)


def test_rd_clip_and_zero_pairs(oracle):
    f = np.array([[[0.0, 5e-9, 2.0]]])
    G = oracle.rd_gradient(f, 2.0, 1.0, dtype=np.float64)
    # 5e-9 is clipped to 0: voxels 0 and 1 form a 0/0 pair that contributes nothing
    G2 = oracle.rd_gradient(np.array([[[0.0, 0.0, 2.0]]]), 2.0, 1.0, dtype=np.float64)
    np.testing.assert_array_equal(G, G2)
    assert G[0, 0, 0] == 0.0


@pytest.mark.parametrize("p", [1.0, 1.2, 2.0])
def test_tv_gradient_is_half_cost_gradient(oracle, p):
    rng = np.random.default_rng(22)
    f = rng.random((4, 5, 6))
    delta, beta = 0.05, 0.3
    G = oracle.tv_gradient(f, delta, beta, p, dtype=np.float64)
    h = 1e-7
    for idx in [tuple(rng.integers(0, s) for s in f.shape) for _ in range(12)]:
        fp, fm = f.copy(), f.copy()
        fp[idx] += h
        fm[idx] -= h
        fd = (oracle.tv_cost(fp, delta, beta, p) - oracle.tv_cost(fm, delta, beta, p)) / (2 * h)
        assert G[idx] == pytest.approx(0.5 * fd, rel=1e-5, abs=1e-8)


def test_tv_parameter_rules_and_p1_closed_form(oracle):
    """N1 == 1 arrays use the 8 in-plane neighbours (src/total_variation.cu:464-482).  For a [1][1][2]
    array voxel 0 reaches voxel 1 through one face and two clamped corner offsets."""
    a, b = 1.0, 0.2
    r2 = float(np.float32(0.7071067811865475))
    w = 1.0 + 2 * r2
    # beta <= 0 -> 1; p = 1 outside delta: h'(t) = delta * sign(t)
    G = oracle.tv_gradient(np.array([[[a, b]]]), 0.1, -5.0, 1.0, dtype=np.float64)
    assert G[0, 0, 0] == pytest.approx(w * 0.1, rel=1e-6)
    assert G[0, 0, 1] == pytest.approx(-w * 0.1, rel=1e-6)
    # delta floored at 1e-8
    G0 = oracle.tv_gradient(np.array([[[a, b]]]), 0.0, 1.0, 1.0, dtype=np.float64)
    assert G0[0, 0, 0] == pytest.approx(w * 1e-8, rel=1e-5)
    # [1][2][2], quadratic region (h'(t) = t): from (0,0) the face neighbours (1,0), (0,1) and, through
    # clamping, three corner offsets land on (1,1), (1,0), (0,1)
    f = np.array([[[1.0, 0.0], [0.0, 0.0]]])
    G2 = oracle.tv_gradient(f, 10.0, 1.0, 1.0, dtype=np.float64)
    assert G2[0, 0, 0] == pytest.approx(2.0 + 3 * r2, rel=1e-6)


def test_fp32_priors_match_fp64(oracle):
    f = np.random.default_rng(23).random((8, 9, 10)).astype(np.float32) + 0.05
    for fn, args in ((oracle.rd_gradient, (2.0, 0.01)), (oracle.tv_gradient, (0.05, 0.005, 1.0))):
        g32 = fn(f, *args)
        g64 = fn(f.astype(np.float64), *args, dtype=np.float64)
        assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-5
