"""GPU parity of PFDR rebinning and FBP against the oracle (numpy FP64 FFTs on the oracle projectors).

Tolerances: hipFFT works in fp32, so each transform pair carries ~1e-7 relative error; PFDR sums up to
N_v1 rows per coefficient and FBP adds the fp32 ramp filter and back projection: 1e-5 relative L2 for
PFDR, 1e-4 for FBP.  The two FBP routes (PFDR inside FBP, or PFDR then FBP of the rebinned data) run the
same kernels on the same numbers and must agree bit for bit.
"""
import numpy as np
import pytest

from petl_amd.petl_server import PETLserver
from petl_amd.workloads import ellipsoid_phantom

pytestmark = pytest.mark.gpu


def rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def setup(oracle, psis=(0.0, 90.0), n=32, D=100.0, L=64.0, seed=80):
    s = PETLserver()
    s.print_warnings = False
    for psi in psis:
        s.add_planogram(psi, D, L, L, 2.0)
    s.set_volume(n, n, n, 2.0)
    plans = [oracle.planogram(psi, D, L, L, 2.0) for psi in psis]
    vol = oracle.volume(n, n, n, 2.0)
    f = ellipsoid_phantom(n, 2.0, 6, seed=seed)
    g = oracle.project(plans, vol, f)
    return s, plans, vol, g


@pytest.mark.parametrize("psis", [(0.0,), (90.0,), (30.0,)])
def test_pfdr_matches_oracle(oracle, psis):
    s, plans, vol, g = setup(oracle, psis)
    g_reb = s.PFDR([a.copy() for a in g])
    for p, a, b in zip(plans, g, g_reb):
        ref = oracle.pfdr(p, a)
        assert b.shape == ref.shape
        assert rel(b, ref) <= 1e-5, rel(b, ref)
    # the model now holds the rebinned geometry
    assert all(int(s.get_planogramSize(i)[0]) == 1 for i in range(len(psis)))


def test_pfdr_and_fbp_ragged_geometry(oracle):
    """Odd crystal counts (15 x 11, so L != H), odd view counts (37 v0, 9 v1) and an odd-sided volume off
    centre on every axis: PFDR and FBP against the oracle at the same tolerances."""
    psis, D, L, H, T = (0.0, 90.0), 72.0, 30.0, 22.0, 2.0
    offs = (1.0, -2.0, 3.0)

    def server():
        s = PETLserver()
        s.print_warnings = False
        for psi in psis:
            s.add_planogram(psi, D, L, H, T)
        s.set_volume(13, 11, 7, 2.0, 2.0, *offs)
        return s

    plans = [oracle.planogram(psi, D, L, H, T) for psi in psis]
    vol = oracle.volume(13, 11, 7, 2.0, 2.0, *offs)
    rng = np.random.default_rng(81)
    g = oracle.project(plans, vol, (0.5 + rng.random(vol.shape)).astype(np.float32))
    g_reb = server().PFDR([a.copy() for a in g])
    for p, a, b in zip(plans, g, g_reb):
        ref = oracle.pfdr(p, a)
        assert b.shape == ref.shape
        assert rel(b, ref) <= 1e-5, rel(b, ref)
    f = server().FBP([a.copy() for a in g])
    ref = oracle.fbp(plans, vol, g)
    assert rel(f, ref) <= 1e-4, rel(f, ref)


def test_fbp_matches_oracle(oracle):
    s, plans, vol, g = setup(oracle)
    f = s.FBP([a.copy() for a in g])
    ref = oracle.fbp(plans, vol, g)
    assert rel(f, ref) <= 1e-4, rel(f, ref)
    # the geometry is unchanged by FBP
    assert int(s.get_planogramSize(0)[0]) == plans[0].N_v1


def test_fbp_of_rebinned_data_is_the_same(oracle):
    s, plans, vol, g = setup(oracle)
    f1 = s.FBP([a.copy() for a in g])
    s2, _, _, _ = setup(oracle)
    g_reb = s2.PFDR([a.copy() for a in g])
    f2 = s2.FBP([a[None] for a in g_reb])
    np.testing.assert_array_equal(f2, f1)


def test_fbp_recovers_a_uniform_cylinder_shape(oracle):
    """Projections of a uniform cylinder: the FBP image is flat inside the cylinder (away from its edge) and
    near zero outside, whatever its absolute scale."""
    s, plans, vol, _ = setup(oracle, n=48, D=140.0, L=96.0)
    z, y, x = np.meshgrid(*(np.arange(48) - 23.5,) * 3, indexing="ij")
    cyl = ((x * x + y * y) <= 14.0 ** 2) & (np.abs(z) <= 14)
    g = oracle.project(plans, vol, cyl.astype(np.float32))
    f = s.FBP(g)
    inner = ((x * x + y * y) <= 8.0 ** 2) & (np.abs(z) <= 6)
    outer = ((x * x + y * y) >= 20.0 ** 2)
    vals = f[inner]
    assert float(vals.std()) < 0.05 * float(vals.mean())
    assert float(np.abs(f[outer]).mean()) < 0.05 * float(vals.mean())
