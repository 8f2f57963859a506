"""Pins the oracle's forward projector (restatement of src/projectors.cu:11-238) against what the
geometry fixes independently of that code:

* uniform volume: an interior bin's footprint overlaps sum to one in every slice, so the bin equals
  (slices x voxel width) / sqrt(1+v0^2+v1^2) -- checks l_phi, the obliquity factor and the weights;
* point source: per view, the footprint's first moment equals the projection of the voxel centre given
  by the reference's own binning geometry (src/planogram.cpp:130-141), and its mass is
  T_x / sqrt(1+v0^2+v1^2) -- catches a wrong sign, swapped axis or dropped shift;
* shift equivariance along z and along the row axis, linearity, exact zeros on masked bins;
* 90-degree rotation equivariance for a general panel angle (30 vs 120 degrees), which checks the
  x-slice branch against the y-slice branch;
* the fp32 restatement agrees with the fp64 one to float rounding.
"""
import numpy as np
import pytest

GEO = dict(D=60.0, L=32.0, H=32.0, T=2.0)


def views(p):
    v1 = np.arange(p.N_v1, dtype=np.float32) * np.float32(p.T_v1) + np.float32(p.v1_0)
    v0 = np.arange(p.N_v0, dtype=np.float32) * np.float32(p.T_v0) + np.float32(p.v0_0)
    u1 = np.arange(p.N_u1, dtype=np.float32) * np.float32(p.T_u1) + np.float32(p.u1_0)
    u0 = np.arange(p.N_u0, dtype=np.float32) * np.float32(p.T_u0) + np.float32(p.u0_0)
    return v1.astype(np.float64), v0.astype(np.float64), u1.astype(np.float64), u0.astype(np.float64)


@pytest.mark.parametrize("psi", [0.0, 90.0])
def test_uniform_volume_closed_form(oracle, psi):
    p = oracle.planogram(psi, **GEO)
    n = 20
    vol = oracle.volume(n, n, n, 2.0)
    g = oracle.project([p], vol, np.ones(vol.shape), dtype=np.float64)[0]
    v1, v0, u1, u0 = views(p)
    half = n * 2.0 / 2
    checked = 0
    for a in range(p.N_v1):
        for b in range(p.N_v0):
            obl = 1.0 / np.sqrt(1 + v0[b] ** 2 + v1[a] ** 2)
            # bins whose footprint stays inside the volume in every slice
            ok0 = (np.abs(u0) + half * abs(v0[b]) + 2.0 <= half - 1e-6) & \
                (np.abs(u0) - 1.0 <= p.L + p.R * abs(v0[b]) - 1e-3)
            ok1 = (np.abs(u1) + half * abs(v1[a]) + 2.0 <= half - 1e-6) & \
                (np.abs(u1) - 1.0 <= p.H - p.R * abs(v1[a]) - 1e-3)
            sel = np.outer(ok1, ok0)
            if sel.any():
                np.testing.assert_allclose(g[a, b][sel], n * 2.0 * obl, rtol=1e-7)
                checked += sel.sum()
    assert checked > 1000


@pytest.mark.parametrize("psi", [0.0, 90.0])
def test_point_source_moments(oracle, psi):
    """Footprint centroid = projection of the voxel centre (src/planogram.cpp:130-141)."""
    p = oracle.planogram(psi, **GEO)
    n = 16
    vol = oracle.volume(n, n, n, 2.0)
    ix, iy, iz = 9, 6, 7
    f = np.zeros(vol.shape)
    f[iz, iy, ix] = 1.0
    g = oracle.project([p], vol, f, dtype=np.float64)[0]
    x, y, z = vol.x0 + ix * 2.0, vol.y0 + iy * 2.0, vol.z0 + iz * 2.0
    cs, sn = float(p.cos_psi), float(p.sin_psi)
    t = y * cs - x * sn
    v1, v0, u1, u0 = views(p)
    checked = 0
    for a in range(p.N_v1):
        for b in range(p.N_v0):
            eu1 = z + v1[a] * t
            eu0 = x * cs + y * sn + v0[b] * t
            # skip views where the footprint touches a mask boundary or the array edge
            if abs(eu1) + 3 > p.H - p.R * abs(v1[a]) or abs(eu0) + 3 > p.L + p.R * abs(v0[b]):
                continue
            if abs(eu1) + 3 > abs(u1[0]) or abs(eu0) + 3 > abs(u0[0]):
                continue
            w = g[a, b]
            m = w.sum()
            obl = 1.0 / np.sqrt(1 + v0[b] ** 2 + v1[a] ** 2)
            assert m == pytest.approx(2.0 * obl, rel=1e-6)
            c1 = (w.sum(axis=1) * u1).sum() / m
            c0 = (w.sum(axis=0) * u0).sum() / m
            assert c1 == pytest.approx(eu1, abs=1e-5)
            assert c0 == pytest.approx(eu0, abs=1e-5)
            checked += 1
    assert checked > 50


@pytest.mark.parametrize("psi", [0.0, 90.0])
def test_shift_equivariance(oracle, psi):
    p = oracle.planogram(psi, **GEO)
    n = 14
    vol = oracle.volume(n, n, n, 2.0)
    rng = np.random.default_rng(5)
    f = np.zeros(vol.shape)
    f[3:11, 3:11, 3:11] = rng.random((8, 8, 8))
    g = oracle.project([p], vol, f, dtype=np.float64)[0]
    # z shift by one voxel -> u1 shift by one bin (T_z == T_u1)
    gz = oracle.project([p], vol, np.roll(f, 1, axis=0), dtype=np.float64)[0]
    mask = (g != 0) & (np.roll(g, 1, axis=2) != 0)
    both = np.zeros_like(mask)
    both[:, :, 1:, :] = True
    sel = mask & both & (gz != 0)
    assert sel.sum() > 1000
    np.testing.assert_allclose(gz[sel], np.roll(g, 1, axis=2)[sel], rtol=1e-9, atol=1e-12)
    # shift along the u0 row axis (x for psi=0, y for psi=90) -> u0 shift by one bin
    ax = 2 if psi == 0.0 else 1
    ga = oracle.project([p], vol, np.roll(f, 1, axis=ax), dtype=np.float64)[0]
    rolled = np.roll(g, 1, axis=3)
    sel = (ga != 0) & (rolled != 0)
    sel[..., 0] = False
    assert sel.sum() > 1000
    # at 90 degrees the float cos(pi/2) = -4.4e-8 couples u0 into the slice parameter at the 1e-7 level
    np.testing.assert_allclose(ga[sel], rolled[sel], rtol=1e-9 if psi == 0.0 else 1e-6, atol=1e-12)


def test_linearity_and_masks(oracle):
    p = oracle.planogram(0.0, **GEO)
    vol = oracle.volume(12, 12, 12, 2.0)
    rng = np.random.default_rng(6)
    a, b = rng.random(vol.shape), rng.random(vol.shape)
    ga, gb, gab = (oracle.project([p], vol, x, dtype=np.float64)[0] for x in (a, b, 2 * a - 3 * b))
    np.testing.assert_allclose(gab, 2 * ga - 3 * gb, rtol=1e-10, atol=1e-10)
    v1, v0, u1, u0 = views(p)
    for iv1 in (0, p.N_v1 // 2, p.N_v1 - 1):
        for iv0 in (0, p.N_v0 // 3, p.N_v0 - 1):
            for iu1 in range(p.N_u1):
                for iu0 in range(p.N_u0):
                    if oracle.bin_masked(p, iv1, iv0, iu1, iu0):
                        assert ga[iv1, iv0, iu1, iu0] == 0.0


def test_rotation_equivariance_general_angle(oracle):
    """planogram(psi + 90) of f == planogram(psi) of f rotated by 90 degrees about z."""
    n = 12
    vol = oracle.volume(n, n, n, 2.0)
    rng = np.random.default_rng(7)
    f = rng.random(vol.shape)
    h = np.ascontiguousarray(f[:, :, ::-1].transpose(0, 2, 1))
    g120 = oracle.project([oracle.planogram(120.0, **GEO)], vol, f, dtype=np.float64)[0]
    g30 = oracle.project([oracle.planogram(30.0, **GEO)], vol, h, dtype=np.float64)[0]
    assert np.linalg.norm(g120 - g30) / np.linalg.norm(g30) < 1e-5


@pytest.mark.parametrize("psi", [0.0, 90.0, 30.0])
def test_fp32_matches_fp64(oracle, psi):
    p = oracle.planogram(psi, D=100.0, L=64.0, H=64.0, T=2.0)
    vol = oracle.volume(24, 24, 24, 2.0)
    f = np.random.default_rng(8).random(vol.shape).astype(np.float32)
    g32 = oracle.project([p], vol, f)[0]
    g64 = oracle.project([p], vol, f.astype(np.float64), dtype=np.float64)[0]
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-6


def test_tap_count_matches_the_per_bin_listing(oracle):
    """count_taps (the work measure of one projection) in FP64 equals, row by row, the number of non-zero
    weights the per-bin tap listing (FP64) returns for every unmasked bin.  The fp32 count is the number of
    listed weights above 1e-6 of the row's largest: what FP64 leaves below that are rounding-level slivers
    (voxel faces that coincide with footprint edges) that fp32 forms as exact zeros.  Both stay under the
    4 N_v1 N_v0 V bound."""
    p = oracle.planogram(0.0, D=100.0, L=32.0, H=32.0, T=2.0)
    vol = oracle.volume(16, 16, 16, 2.0)
    for iv1 in (0, p.N_v1 // 2):
        ws = []
        for iv0 in range(p.N_v0):
            for iu1 in range(p.N_u1):
                for iu0 in range(p.N_u0):
                    ws.append(oracle.bin_taps(p, vol, iv1, iv0, iu1, iu0)[1])
        w = np.concatenate(ws)
        assert oracle.count_taps(p, vol, iv1, iv1 + 1, f64=True) == int(np.count_nonzero(w))
        assert oracle.count_taps(p, vol, iv1, iv1 + 1) == int(np.count_nonzero(w > 1e-6 * w.max()))
    bound = 4 * p.N_v1 * p.N_v0 * 16 ** 3
    assert oracle.count_taps(p, vol) <= oracle.count_taps(p, vol, f64=True) <= bound
