"""Pins of the list-mode binning oracle (src/binning.cpp, src/planogram.cpp:113-141).

* event geometry: bit-exact against the reference's own planogram methods compiled from its sources;
* designed coincidences: LORs built from chosen planogram coordinates (p on the LOR at t = 0, direction
  r of src/projectors.cu:52-53) land in the chosen bins (nearest) or split with the chosen fractions
  (bilinear); clamping at the edges and the int(x + 0.5) truncation of negative positions;
* mass: every accepted, in-range coincidence adds 1 in total;
* the two file layouts round-trip through the readers.
The reference's binning.cpp itself cannot be built here: it needs the `parameters` class, whose
implementation requires the CUDA runtime.
"""
import numpy as np
import pytest

GEO = dict(D=100.0, L=64.0, H=64.0, T=2.0)


def plan_of(oracle, psi):
    return oracle.planogram(psi, GEO["D"], GEO["L"], GEO["H"], GEO["T"])


def lor_for(p, v1, v0, u1, u0, s=150.0, chest_wall=True):
    """Blue at the LOR point of parameter 0, pink s along -r: r = blue - pink = s * direction."""
    c, sn = float(p.cos_psi), float(p.sin_psi)
    blue = np.array([u0 * c, u0 * sn, u1 + (p.H if chest_wall else 0.0)], np.float64)
    d = np.array([-(v0 * c + sn), -(v0 * sn - c), -v1], np.float64)
    return blue.astype(np.float32), (blue - s * d).astype(np.float32)


def centre(p, iv1, iv0, iu1, iu0, f1=0.0, f0=0.0):
    return (p.v1_0 + iv1 * p.T_v1, p.v0_0 + iv0 * p.T_v0, p.u1_0 + (iu1 + f1) * p.T_u1, p.u0_0 + (iu0 + f0) * p.T_u0)


@pytest.mark.parametrize("psi", [0.0, 90.0, 30.0])
def test_event_geometry_matches_reference(oracle, psi):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference sources not present")
    p = plan_of(oracle, psi)
    rng = np.random.default_rng(3)
    n_acc = 0
    for _ in range(400):
        blue = rng.uniform(-60, 60, 3).astype(np.float32)
        pink = (blue + rng.uniform(-80, 80, 3)).astype(np.float32)
        ok, got = oracle.event_coords(p, blue, pink)
        ref = np.zeros(4, np.float32)
        rok = R.ref_event_coords(psi, GEO["D"], GEO["L"], GEO["H"], 1.0, p.v_m1, GEO["T"], blue, pink, ref)
        assert ok == bool(rok)
        if ok:
            n_acc += 1
            np.testing.assert_array_equal(got, ref)
    assert n_acc > 50


@pytest.mark.parametrize("psi", [0.0, 90.0])
def test_designed_events_nearest(oracle, psi):
    p = plan_of(oracle, psi)
    rng = np.random.default_rng(4)
    want = np.zeros(p.shape, np.float32)
    blue, pink = [], []
    for _ in range(300):
        idx = (int(rng.integers(1, p.N_v1 - 1)), int(rng.integers(1, p.N_v0 - 1)),
               int(rng.integers(1, p.N_u1 - 1)), int(rng.integers(1, p.N_u0 - 1)))
        ok, _ = oracle.event_coords(p, *lor_for(p, *centre(p, *idx)))
        if not ok:
            continue
        for _ in range(int(rng.integers(1, 4))):
            b, k = lor_for(p, *centre(p, *idx))
            blue.append(b)
            pink.append(k)
            want[idx] += 1
    g, acc = oracle.bin_events([p], np.array(blue), np.array(pink), nearest=True)
    assert acc[0] == len(blue)
    np.testing.assert_array_equal(g[0], want)


def test_bilinear_split_fractions(oracle):
    p = plan_of(oracle, 0.0)
    iv1, iv0, iu1, iu0 = p.N_v1 // 2, p.N_v0 // 2, 10, 20
    b, k = lor_for(p, *centre(p, iv1, iv0, iu1, iu0, f1=0.25, f0=0.5))
    g, _ = oracle.bin_events([p], b[None], k[None])
    row = g[0][iv1, iv0]
    np.testing.assert_allclose(row[iu1:iu1 + 2, iu0:iu0 + 2], [[0.75 * 0.5, 0.75 * 0.5], [0.25 * 0.5, 0.25 * 0.5]],
                               atol=2e-4)
    assert float(row.sum()) == pytest.approx(1.0, abs=1e-6)
    assert np.count_nonzero(row) == 4


def test_bilinear_clamps_at_edges(oracle):
    p = plan_of(oracle, 0.0)
    iv1, iv0 = p.N_v1 // 2, p.N_v0 // 2
    b1, k1 = lor_for(p, *centre(p, iv1, iv0, 5, -3))            # u0 below the first bin
    b2, k2 = lor_for(p, *centre(p, iv1, iv0, p.N_u1 + 2, 7))      # u1 beyond the last bin
    g, _ = oracle.bin_events([p], np.array([b1, b2]), np.array([k1, k2]))
    row = g[0][iv1, iv0]
    assert row[5, 0] == pytest.approx(1.0, abs=1e-6)
    assert row[p.N_u1 - 1, 7] == pytest.approx(1.0, abs=1e-6)
    assert float(row.sum()) == pytest.approx(2.0, abs=1e-6)


def test_nearest_truncation_of_negative_positions(oracle):
    """int(x + 0.5) truncates toward zero: u0 positions down to 1.5 bins below the first bin centre still
    land in bin 0 (src/binning.cpp:697-700); further out they are dropped."""
    p = plan_of(oracle, 0.0)
    iv1, iv0 = p.N_v1 // 2, p.N_v0 // 2
    ev = [lor_for(p, *centre(p, iv1, iv0, 6, -1.2)), lor_for(p, *centre(p, iv1, iv0, 6, -1.7))]
    g, acc = oracle.bin_events([p], np.array([e[0] for e in ev]), np.array([e[1] for e in ev]), nearest=True)
    assert acc[0] == 2
    assert g[0][iv1, iv0, 6, 0] == 1.0 and float(g[0].sum()) == 1.0


def test_mass_of_random_coincidences(oracle):
    """Bilinear binning adds exactly one (to rounding) per accepted coincidence whose v bins exist."""
    plans = [plan_of(oracle, 0.0), plan_of(oracle, 90.0)]
    rng = np.random.default_rng(5)
    n = 20000
    blue = np.column_stack([rng.uniform(-40, 40, n), np.full(n, 50.0), rng.uniform(0, 64, n)]).astype(np.float32)
    pink = np.column_stack([rng.uniform(-40, 40, n), np.full(n, -50.0), rng.uniform(0, 64, n)]).astype(np.float32)
    g, acc = oracle.bin_events(plans, blue, pink)
    assert acc[0] > 1000
    assert float(g[0].astype(np.float64).sum()) == pytest.approx(acc[0], rel=1e-6)
    # panels facing along y are seen edge-on by the psi = 90 planogram: the acceptance rejects them
    assert acc[1] < acc[0]


def test_chest_wall_origin_shifts_u1_by_H(oracle):
    p = plan_of(oracle, 0.0)
    b, k = lor_for(p, *centre(p, p.N_v1 // 2, p.N_v0 // 2, 9, 30), chest_wall=False)
    g_c, _ = oracle.bin_events([p], b[None], k[None], nearest=True, chest_wall=False)
    b2 = b.copy()
    b2[2] += np.float32(p.H)
    k2 = k.copy()
    k2[2] += np.float32(p.H)
    g_w, _ = oracle.bin_events([p], b2[None], k2[None], nearest=True, chest_wall=True)
    np.testing.assert_array_equal(g_c[0], g_w[0])
    assert g_c[0][p.N_v1 // 2, p.N_v0 // 2, 9, 30] == 1.0


def test_file_layouts_round_trip(oracle, tmp_path):
    rng = np.random.default_rng(6)
    blue = rng.uniform(-50, 50, (37, 3)).astype(np.float32)
    pink = rng.uniform(-50, 50, (37, 3)).astype(np.float32)
    oracle.write_lst(tmp_path / "a.lst", blue, pink)
    b, k = oracle.read_lst(tmp_path / "a.lst")
    np.testing.assert_array_equal(b, blue)
    np.testing.assert_array_equal(k, pink)
    extra = rng.integers(0, 3, (37, 2))
    oracle.write_acq(tmp_path / "a.acq", blue, pink, extra_interactions=extra)
    b, k = oracle.read_acq(tmp_path / "a.acq")
    np.testing.assert_array_equal(b, blue)
    np.testing.assert_array_equal(k, pink)
    assert (tmp_path / "a.lst").stat().st_size == 37 * 48
    assert (tmp_path / "a.acq").stat().st_size == 8 + 2 * 37 * 18 + 18 * int(extra.sum())
