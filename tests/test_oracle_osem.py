"""Pins the oracle's OSEM (interleaved v1-row subsets; oracle.osem, built on the oracle's row-range
projector entry points) without restating it:

* one subset is MLEM, bit for bit (the reference loop of src/petl_server.py:1534-1606 with one pass);
* with k subsets it equals OSEM composed from the full-planogram projectors, where subset b's forward is
  the full forward restricted to its rows and its back projection is that of a planogram zeroed outside
  its rows (bit for bit: same per-bin and per-voxel arithmetic, zero bins add exactly nothing);
* the composed form, in FP64 with the exactly-adjoint back projector, satisfies the EM identities that
  fix the algorithm: per-subset count preservation sum_j S_b,j f+_j = sum_{i in b, (Af)_i > 0} y_i, and
  noise-free consistent data y = A f as a fixed point, for 2, 4 and 8 subsets.  A wrong subset (rows,
  stride, or sensitivity of another subset) breaks the fixed point; a wrong ratio breaks the count.
"""
import numpy as np
import pytest

GEO = dict(D=60.0, L=32.0, H=32.0, T=2.0)


def setup(oracle, psis=(0.0, 90.0), n=10):
    plans = [oracle.planogram(psi, **GEO) for psi in psis]
    vol = oracle.volume(n, n, n, 2.0)
    rng = np.random.default_rng(40)
    f_true = (rng.random(vol.shape) + 0.2).astype(np.float32)
    g = oracle.project(plans, vol, f_true)
    y = [rng.poisson(50.0 * gi).astype(np.float32) for gi in g]
    return plans, vol, f_true, y


def test_one_subset_is_mlem(oracle):
    plans, vol, _, y = setup(oracle)
    a = oracle.osem(plans, vol, y, 3, 1)
    b = oracle.mlem(plans, vol, y, 3)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("nsub", [2, 4])
def test_row_range_osem_equals_composed_osem(oracle, nsub):
    plans, vol, _, y = setup(oracle)
    a = oracle.osem(plans, vol, y, 2, nsub)
    b = oracle.osem_composed(plans, vol, y, 2, nsub)
    np.testing.assert_array_equal(a, b)
    # and it is not MLEM: the subsets change the iterate
    assert np.max(np.abs(a - oracle.mlem(plans, vol, y, 2))) > 1e-3


def test_row_range_osem_start_image(oracle):
    plans, vol, f_true, y = setup(oracle)
    f0 = np.where(f_true > 0.7, f_true, 0.0).astype(np.float32)
    a = oracle.osem(plans, vol, y, 1, 3, f0=f0)
    b = oracle.osem_composed(plans, vol, y, 1, 3, f0=f0)
    np.testing.assert_array_equal(a, b)
    assert np.all(a[f0 == 0] == 0)  # the multiplicative update keeps zeros


@pytest.mark.parametrize("nsub", [2, 4, 8])
def test_composed_osem_count_preservation(oracle, nsub):
    plans, vol, _, y = setup(oracle)
    y = [yi.astype(np.float64) for yi in y]
    checks = []

    def cb(sub, f_before, f_after, S_sub, q_sub):
        counts = sum(float(yi[sub::nsub][qi[sub::nsub] > 0].sum()) for yi, qi in zip(y, q_sub))
        checks.append((float(np.sum(S_sub * f_after)), counts))

    oracle.osem_composed(plans, vol, y, 2, nsub, dtype=np.float64, mode="matched", callback=cb)
    assert len(checks) == 2 * nsub
    for got, want in checks:
        assert got == pytest.approx(want, rel=1e-10)


@pytest.mark.parametrize("nsub", [2, 4, 8])
def test_composed_osem_fixed_point(oracle, nsub):
    plans, vol, f_true, _ = setup(oracle)
    f_true = f_true.astype(np.float64)
    y = oracle.project(plans, vol, f_true, dtype=np.float64)
    f = oracle.osem_composed(plans, vol, y, 2, nsub, f0=f_true, dtype=np.float64, mode="matched")
    # every voxel is seen by every subset here (central voxels of a 10^3 volume in two orientations)
    np.testing.assert_allclose(f, f_true, rtol=1e-10)


def test_composed_osem_one_subset_is_mlem_fp64(oracle):
    plans, vol, _, y = setup(oracle)
    a = oracle.osem_composed(plans, vol, y, 2, 1, dtype=np.float64, mode="matched")
    b = oracle.mlem(plans, vol, y, 2, dtype=np.float64, mode="matched")
    # the matched back projector sums per-thread partial volumes, so the planogram sums associate
    # differently in the two compositions
    np.testing.assert_allclose(a, b, rtol=1e-13)
