"""Pins of the oracle's analytic reconstruction (CPU only): FFT sizes, the ramp filter and PFDR rebinning.

  * optimal_fft_size equals the reference's compiled optimalFFTsize (src/vector_ops.cpp:481-505).
  * The ramp filter (zero padding to at least twice the row, multiplication by the transform of the kernel)
    equals the direct linear convolution of each row with the kernel.
  * PFDR of a planogram with a single v1 = 0 row is the identity (the two weights cancel and the transform
    pair is normalised); PFDR of data whose oblique rows all weight to the same views gives those views,
    times sqrt(1 + v0^2), whatever the number of rows each frequency collects (the normalisation divides by
    the count, except at U = 0 where only the direct row contributes).
"""
import numpy as np
import pytest

from oracle import oracle as orc


def test_optimal_fft_size_matches_reference():
    R = orc.ref_lib()
    if R is None:
        pytest.skip("reference sources absent")
    for n in range(1, 3000):
        assert orc.optimal_fft_size(n) == R.ref_optimal_fft_size(n), n


@pytest.mark.parametrize("Nu0,T", [(17, 2.0), (64, 1.6), (100, 2.0)])
def test_ramp_filter_is_linear_convolution(Nu0, T):
    rng = np.random.default_rng(3)
    g = rng.standard_normal((3, 4, Nu0))
    got = orc.ramp_filter(g, T)
    NH = orc.optimal_fft_size(2 * Nu0)
    h = orc.ramp_kernel(NH, T)
    want = np.zeros_like(g)
    for k in range(Nu0):
        for j in range(Nu0):
            want[..., k] += g[..., j] * h[(k - j) % NH]
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())


def test_pfdr_of_one_direct_row_is_identity():
    p = orc.rebinned_plan(orc.planogram(0.0, 100.0, 64.0, 64.0, 2.0))
    g = np.random.default_rng(4).random(p.shape)
    np.testing.assert_allclose(orc.pfdr(p, g), g[0], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("psi", [0.0, 90.0])
def test_pfdr_of_consistent_rows(psi):
    p = orc.planogram(psi, 100.0, 64.0, 64.0, 2.0)
    rng = np.random.default_rng(5)
    g0 = rng.random(p.shape[1:])
    v1 = (np.arange(p.N_v1, dtype=np.float32) * np.float32(p.T_v1) + np.float32(p.v1_0)).astype(np.float64)
    v0 = (np.arange(p.N_v0, dtype=np.float32) * np.float32(p.T_v0) + np.float32(p.v0_0)).astype(np.float64)
    g = g0[None] * np.sqrt(1.0 + v1[:, None] ** 2 + v0[None, :] ** 2)[:, :, None, None]
    got = orc.pfdr(p, g)
    want = g0 * np.sqrt(1.0 + v0 * v0)[:, None, None]
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-9 * np.abs(want).max())
