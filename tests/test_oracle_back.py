"""Pins the oracle's two back projectors.

* matched mode is the exact transpose of the forward: <Ax, y> = <x, A^T y> in fp64;
* PETL mode (an independent restatement of src/projectors.cu:240-377 with exact bilinear weights)
  equals the transpose of the forward wherever the reference's mask differences cannot act -- data
  that is zero on masked bins and on the two u1 rows inside the mask edge -- which ties the
  independently written gather formula to the forward's footprints;
* PETL-mode sensitivity (back projection of ones) of the centre voxel has the closed form
  T_x * sum over views whose voxel-centre u1 test passes of 1/sqrt(1+v0^2+v1^2).
"""
import numpy as np
import pytest

GEO = dict(D=60.0, L=32.0, H=32.0, T=2.0)


def coords(p):
    f32 = np.float32
    v1 = (np.arange(p.N_v1, dtype=f32) * f32(p.T_v1) + f32(p.v1_0)).astype(np.float64)
    v0 = (np.arange(p.N_v0, dtype=f32) * f32(p.T_v0) + f32(p.v0_0)).astype(np.float64)
    u1 = (np.arange(p.N_u1, dtype=f32) * f32(p.T_u1) + f32(p.u1_0)).astype(np.float64)
    u0 = (np.arange(p.N_u0, dtype=f32) * f32(p.T_u0) + f32(p.u0_0)).astype(np.float64)
    return v1, v0, u1, u0


@pytest.mark.parametrize("psi", [0.0, 90.0, 30.0, 135.0])
def test_matched_is_exact_adjoint(oracle, psi):
    p = oracle.planogram(psi, **GEO)
    vol = oracle.volume(12, 10, 11, 2.0, 2.0, 1.0, 0.0, -1.0)
    rng = np.random.default_rng(11)
    x = rng.random(vol.shape)
    y = rng.random(p.shape)
    Ax = oracle.project([p], vol, x, dtype=np.float64)[0]
    ATy = oracle.backproject([p], vol, [y], mode="matched", dtype=np.float64)
    lhs, rhs = np.vdot(Ax, y), np.vdot(x, ATy)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("psi", [0.0, 90.0])
def test_petl_mode_equals_transpose_away_from_masks(oracle, psi):
    p = oracle.planogram(psi, **GEO)
    vol = oracle.volume(12, 12, 12, 2.0)
    v1, v0, u1, u0 = coords(p)
    rng = np.random.default_rng(12)
    y = rng.random(p.shape)
    thr1 = p.H - p.R * np.abs(v1)                     # per v1
    thr0 = p.L + p.R * np.abs(v0)                     # per v0
    keep1 = (np.abs(u1)[None, :] - 0.5 * p.T_u1) <= (thr1[:, None] - 2.0 * p.T_u1)   # [v1][u1]
    keep0 = (np.abs(u0)[None, :] - 0.5 * p.T_u0) <= thr0[:, None]                     # [v0][u0]
    y *= keep1[:, None, :, None] * keep0[None, :, None, :]
    bm = oracle.backproject([p], vol, [y], mode="matched", dtype=np.float64)
    bp = oracle.backproject([p], vol, [y], mode="petl", dtype=np.float64)
    tol = 1e-12 if psi == 0.0 else 1e-6   # float cos(pi/2) = -4.4e-8 enters the PETL-mode Jacobian
    assert np.linalg.norm(bp - bm) <= tol * np.linalg.norm(bm)


def test_petl_mode_differs_only_near_masks(oracle):
    """On data that is zero on masked bins, the modes differ only on z slices whose voxel centres
    reach the u1 mask edge (the voxel-centre u1 test of src/projectors.cu:289)."""
    p = oracle.planogram(0.0, **GEO)
    vol = oracle.volume(12, 12, 12, 2.0)
    v1, v0, u1, u0 = coords(p)
    y = np.random.default_rng(13).random(p.shape)
    keep1 = (np.abs(u1)[None, :] - 0.5 * p.T_u1) <= (p.H - p.R * np.abs(v1))[:, None]
    keep0 = (np.abs(u0)[None, :] - 0.5 * p.T_u0) <= (p.L + p.R * np.abs(v0))[:, None]
    y *= keep1[:, None, :, None] * keep0[None, :, None, :]
    bm = oracle.backproject([p], vol, [y], mode="matched", dtype=np.float64)
    bp = oracle.backproject([p], vol, [y], mode="petl", dtype=np.float64)
    d = np.abs(bp - bm).max(axis=(1, 2))
    assert d[3:9].max() < 1e-9 * np.abs(bm).max()
    assert d.max() > 1e-3 * np.abs(bm).max()


def test_petl_sensitivity_closed_form(oracle):
    p = oracle.planogram(0.0, **GEO)
    n = 12
    vol = oracle.volume(n, n, n, 2.0)
    S = oracle.backproject([p], vol, [np.ones(p.shape)], mode="petl", dtype=np.float64)
    v1, v0, u1, u0 = coords(p)
    iz = iy = ix = n // 2
    x, y, z = vol.x0 + ix * 2.0, vol.y0 + iy * 2.0, vol.z0 + iz * 2.0
    t = y  # psi = 0
    expect = 0.0
    for a in range(p.N_v1):
        uc = np.float32(np.float32(z) + np.float32(np.float32(v1[a]) * np.float32(t)))
        if abs(uc) - 0.5 * p.T_u1 > p.H - p.R * abs(np.float32(v1[a])):
            continue
        for b in range(p.N_v0):
            expect += 1.0 / np.sqrt(1 + v0[b] ** 2 + v1[a] ** 2)
    assert S[iz, iy, ix] == pytest.approx(2.0 * expect, rel=1e-9)


def test_fp32_back_matches_fp64(oracle):
    p = oracle.planogram(0.0, D=100.0, L=64.0, H=64.0, T=2.0)
    vol = oracle.volume(20, 20, 20, 2.0)
    y = np.random.default_rng(14).random(p.shape).astype(np.float32)
    for mode in ("petl", "matched"):
        b32 = oracle.backproject([p], vol, [y], mode=mode)
        b64 = oracle.backproject([p], vol, [y.astype(np.float64)], mode=mode, dtype=np.float64)
        assert np.linalg.norm(b32 - b64) / np.linalg.norm(b64) < 1e-6
