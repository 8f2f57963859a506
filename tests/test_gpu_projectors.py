"""GPU parity of the forward and back projectors against the CPU oracle (fp32 restatement of
src/projectors.cu), on small geometries that the oracle finishes in well under a second.

Tolerance: 1e-5 relative L2 (the parity bar for float32 projections) on both paths.  The fast separable
kernels compute the footprint edges once per (view, slice) in FP64, and the general kernels (any psi, any
pitch; path "generic" forces them on every geometry) take the planes' crossing coefficients from FP64
tables, so both differ from the fp32 oracle by rounding noise.
"""
import numpy as np
import pytest

from petl_amd.petl_server import PETLserver

pytestmark = pytest.mark.gpu

CASES = {
    # name: (psis, D, L, H, T, (nx, ny, nz), voxelWidth, voxelHeight, offsets)
    "psi0": ((0.0,), 60.0, 32.0, 32.0, 2.0, (12, 12, 12), 2.0, 2.0, (0.0, 0.0, 0.0)),
    "psi90": ((90.0,), 60.0, 32.0, 32.0, 2.0, (12, 12, 12), 2.0, 2.0, (0.0, 0.0, 0.0)),
    "pair_offsets": ((0.0, 90.0), 80.0, 48.0, 40.0, 2.0, (14, 10, 9), 2.0, 2.0, (2.0, -3.0, 1.0)),
    "panel32": ((0.0,), 100.0, 64.0, 64.0, 2.0, (32, 32, 32), 2.0, 2.0, (0.0, 0.0, 0.0)),
    # odd crystal counts (15 x 11), an odd number of views (37 v0, 9 v1), odd volume sides, offsets on every axis
    "ragged": ((0.0, 90.0), 72.0, 30.0, 22.0, 2.0, (13, 11, 7), 2.0, 2.0, (1.0, -2.0, 3.0)),
    # one v1 row, two u1 rows, fewer z slices than the back projector's shortest z chunk (4)
    "thin": ((0.0,), 20.0, 4.0, 6.0, 2.0, (5, 4, 3), 2.0, 2.0, (0.0, 0.0, 0.0)),
    "psi30_generic": ((30.0,), 60.0, 32.0, 32.0, 2.0, (12, 12, 12), 2.0, 2.0, (0.0, 0.0, 0.0)),
    "pitch_generic": ((0.0,), 60.0, 32.0, 32.0, 2.0, (16, 14, 12), 1.5, 2.5, (0.0, 0.0, 0.0)),
}
FAST_CASES = ["psi0", "psi90", "pair_offsets", "panel32", "ragged", "thin"]


def make_pair(oracle, case):
    psis, D, L, H, T, (nx, ny, nz), vw, vh, offs = CASES[case]
    s = PETLserver()
    s.print_warnings = False
    for psi in psis:
        s.add_planogram(psi, D, L, H, T)
    s.set_volume(nx, ny, nz, vw, vh, *offs)
    plans = [oracle.planogram(psi, D, L, H, T) for psi in psis]
    vol = oracle.volume(nx, ny, nz, vw, vh, *offs)
    return s, plans, vol


def rel(a, b):
    a = np.concatenate([x.ravel() for x in a]) if isinstance(a, list) else a.ravel()
    b = np.concatenate([x.ravel() for x in b]) if isinstance(b, list) else b.ravel()
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64)))


def tol(case, path):
    return 1e-5


@pytest.mark.parametrize("path", ["auto", "generic"])
@pytest.mark.parametrize("case", list(CASES))
def test_forward_matches_oracle(oracle, case, path):
    s, plans, vol = make_pair(oracle, case)
    s.set_projector_path(path)
    rng = np.random.default_rng(1)
    f = rng.random(vol.shape).astype(np.float32)
    g = s.allocate_projections()
    s.project(g, f)
    ref = oracle.project(plans, vol, f)
    assert rel(g, ref) <= tol(case, path)
    # bins the oracle projects to zero stay at round-off of a footprint edge; nothing is negative
    for gi, ri in zip(g, ref):
        stray = float(np.abs(gi[ri == 0.0]).max(initial=0.0))
        assert stray <= 1e-6 * float(np.abs(ri).max()), stray
        assert np.all(gi >= 0.0)


@pytest.mark.parametrize("case", FAST_CASES)
def test_forward_sparse_support_exact_zeros(oracle, case):
    """A volume that is zero outside a small block.  Bins beside the support must not pick up a round-off
    residual of neighbouring large sums: an MLEM ratio y / q turns one into a huge value."""
    s, plans, vol = make_pair(oracle, case)
    rng = np.random.default_rng(9)
    f = np.zeros(vol.shape, np.float32)
    nz, ny, nx = vol.shape
    block = f[nz // 3: nz // 3 + 3, ny // 4: ny // 4 + 4, nx // 2: nx // 2 + 3]  # clipped at small volumes
    block[...] = 1.0 + rng.random(block.shape)
    f[-1, -1, -1] = 5.0
    g = s.allocate_projections()
    s.project(g, f)
    ref = oracle.project(plans, vol, f)
    assert rel(g, ref) <= 1e-5
    for gi, ri in zip(g, ref):
        # a footprint edge can round to a zero weight on one side only: such bins are tiny on both sides
        stray = float(np.abs(gi[ri == 0.0]).max(initial=0.0))
        assert stray <= 1e-6 * float(ri.max()), stray
        assert np.all(np.abs(ri[gi == 0.0]) <= 1e-6 * ri.max())
        assert np.all(gi >= 0.0)


@pytest.mark.parametrize("case", FAST_CASES)
def test_fast_path_selected(oracle, case):
    s, plans, _ = make_pair(oracle, case)
    assert all(s.uses_fast_path(i) for i in range(len(plans)))


@pytest.mark.parametrize("case", ["psi30_generic", "pitch_generic"])
def test_generic_path_selected(oracle, case):
    s, plans, _ = make_pair(oracle, case)
    assert not any(s.uses_fast_path(i) for i in range(len(plans)))


@pytest.mark.parametrize("mode", ["petl", "matched"])
@pytest.mark.parametrize("path", ["auto", "generic"])
@pytest.mark.parametrize("case", list(CASES))
def test_back_matches_oracle(oracle, case, path, mode):
    s, plans, vol = make_pair(oracle, case)
    s.set_projector_path(path)
    s.set_backprojector_mode(mode)
    rng = np.random.default_rng(2)
    g = [rng.random(p.shape).astype(np.float32) for p in plans]
    f = s.backproject(g)
    ref = oracle.backproject(plans, vol, g, mode=mode)
    assert rel(f, ref) <= tol(case, path)


@pytest.mark.parametrize("path", ["auto", "generic"])
@pytest.mark.parametrize("case", list(CASES))
def test_adjoint(oracle, case, path):
    """<A x, y> = <x, A^T y> for the matched back projector, to 1e-6."""
    s, plans, vol = make_pair(oracle, case)
    s.set_projector_path(path)
    s.set_backprojector_mode("matched")
    rng = np.random.default_rng(3)
    x = rng.random(vol.shape).astype(np.float32)
    y = [rng.random(p.shape).astype(np.float32) for p in plans]
    Ax = s.project(s.allocate_projections(), x)
    ATy = s.backproject(y)
    lhs = sum(np.vdot(a.astype(np.float64), b.astype(np.float64)) for a, b in zip(Ax, y))
    rhs = np.vdot(x.astype(np.float64), ATy.astype(np.float64))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
