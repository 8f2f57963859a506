"""Bounds the one detail of the reference back projector that nothing here can run: its hardware texture
filtering.  src/projectors.cu:351-370 merges the footprint's bins in pairs and reads each pair with one
bilinear texture fetch at i + w2 / (w1 + w2) (src/cuda_utils.cu:1037-1044 sets linear filtering); the
texture unit holds the interpolation fraction in 9-bit fixed point with 8 fractional bits (public CUDA C
Programming Guide, texture fetching appendix).  The guide does not state the rounding, so both rounding to
the nearest 1/256 and truncation are bounded.  The product and the oracle use exact weights; these CPU
tests measure, in the oracle only, how far the quantised reference form can lie from them:

* with exact fractions the pair-merged form equals the oracle's PETL back projection (pins the restated
  pair structure, border reads and weight sums);
* a single bin pair: the quantised fraction is within 1/512 (nearest) or 1/256 (truncated) of the exact one;
* relative L2 differences of one back projection at the small configuration (3e-5 to 8e-5 rounding to
  nearest, up to 3.2e-4 truncating) and of sampled rows at the full benchmark geometry (7e-6 / 3.7e-5), below
  1e-3; after 10 MLEM iterations at the small configuration the images differ by 6.4e-4 (nearest) and
  1.1e-3 (truncated), held to 2e-3 (the measured values are printed).
"""
import numpy as np
import pytest

from petl_amd.workloads import ellipsoid_phantom, poisson_data


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(np.asarray(b, np.float64)))


@pytest.fixture(scope="module")
def small(oracle):
    P = [oracle.planogram(0.0, 100.0, 64.0, 64.0, 2.0)]
    V = oracle.volume(32, 32, 32, 2.0)
    f = ellipsoid_phantom(32, 2.0, 6, seed=1)
    g = oracle.project(P, V, f)
    y = poisson_data(g, 1e6, seed=101)
    return P, V, g, y


@pytest.mark.parametrize("geo", ["small", "ragged"])
def test_exact_pair_form_equals_petl_back(oracle, small, geo):
    if geo == "small":
        P, V, g, _ = small
    else:
        P = [oracle.planogram(psi, 72.0, 30.0, 22.0, 2.0) for psi in (0.0, 90.0, 30.0)]
        V = oracle.volume(13, 11, 7, 2.0, 2.0, 1.0, -2.0, 3.0)
        rng = np.random.default_rng(4)
        g = [rng.random(p.shape).astype(np.float32) for p in P]
    exact = oracle.backproject(P, V, g, mode="petl")
    pair = oracle.backproject_petl_tex(P, V, g, "exact")
    assert rel(pair, exact) <= 2e-6


def test_single_pair_fraction_quantisation(oracle):
    """One bin pair lit: the fetch's fraction a = w2 / (w1 + w2) becomes a multiple of 1/256."""
    P = [oracle.planogram(0.0, 100.0, 64.0, 64.0, 2.0)]
    V = oracle.volume(16, 16, 16, 2.0)
    g = np.zeros(P[0].shape, np.float32)
    g[21, 37] = np.random.default_rng(5).random(g.shape[2:], dtype=np.float32)  # an oblique view
    exact = oracle.backproject(P, V, [g], mode="petl")
    lit = np.zeros_like(g)
    lit[21, 37] = 1.0
    wsum = oracle.backproject(P, V, [lit], mode="petl")  # per voxel: the sum of its fetches' weights
    for quant, bound in (("nearest", 1.0 / 512.0), ("truncate", 1.0 / 256.0)):
        q = oracle.backproject_petl_tex(P, V, [g], quant)
        # each fetch moves by at most (|da0| + |da1|) x (largest difference of neighbouring bins, < 1 here)
        # x its weights, so a voxel moves by at most 2 x bound x its weight sum
        assert np.all(np.abs(q - exact) <= 2.0 * bound * wsum + 1e-6)
        assert rel(q, exact) > 0.0


def test_one_back_projection_bound(oracle, small):
    P, V, g, y = small
    rng = np.random.default_rng(3)
    cases = {"Af": g, "random": [rng.random(p.shape).astype(np.float32) for p in P],
             "mlem_ratio": [np.where(a > 0, yy / np.where(a > 0, a, 1), 0).astype(np.float32) for a, yy in zip(g, y)]}
    for name, gg in cases.items():
        exact = oracle.backproject(P, V, gg, mode="petl")
        for quant in ("nearest", "truncate"):
            d = rel(oracle.backproject_petl_tex(P, V, gg, quant), exact)
            print(f"texture fractions ({quant}), back projection of {name}: rel-L2 {d:.2e}")
            assert d <= 1e-3


def test_mlem_bound(oracle, small):
    """10 MLEM iterations with the quantised texture form against exact weights (both fp32)."""
    P, V, g, y = small

    def run(back):
        f = np.ones(V.shape, np.float32)
        S = back([np.ones(p.shape, np.float32) for p in P])
        S[S == 0] = 1
        for _ in range(10):
            q = oracle.project(P, V, f)
            r = [oracle.rdivide(a, yy) for a, yy in zip(q, y)]
            f = f * oracle.divide(back(r), S)
        return f

    exact = run(lambda gg: oracle.backproject(P, V, gg, mode="petl"))
    for quant in ("nearest", "truncate"):
        d = rel(run(lambda gg: oracle.backproject_petl_tex(P, V, gg, quant)), exact)
        print(f"texture fractions ({quant}), 10 MLEM iterations: image rel-L2 {d:.2e}")
        # MLEM compounds the per-projection differences (6.4e-4 nearest, 1.1e-3 truncated measured here)
        assert d <= 2e-3


def test_full_size_sampled_rows_bound(oracle):
    """Rows 0 and 50 of the psi = 0 planogram at the benchmark geometry (128^3 volume), the projection of
    the seed-2 phantom as data."""
    P = oracle.planogram(0.0, 200.0, 256.0, 256.0, 2.0)
    V = oracle.volume(128, 128, 128, 2.0)
    f = ellipsoid_phantom(128, 2.0, 20, seed=2)
    g = np.zeros(P.shape, np.float32)
    for r in (0, 50):
        oracle.project_rows(P, V, f, g, r, r + 1)
    for r in (0, 50):
        exact = np.zeros(V.shape, np.float32)
        oracle.backproject_petl_rows(P, V, g, exact, r, r + 1)
        for quant in ("nearest", "truncate"):
            d = rel(oracle.backproject_petl_tex([P], V, [g], quant, v1_rows=(r, r + 1)), exact)
            print(f"texture fractions ({quant}), full-size row {r}: rel-L2 {d:.2e}")
            assert d <= 1e-3
