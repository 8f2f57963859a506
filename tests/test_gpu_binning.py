"""List-mode binning on the GPU against the oracle (src/binning.cpp).  Every comparison is bit-exact: the
device adds the contributions to each bin in the reference's event order."""
import numpy as np
import pytest

from petl_amd.petl_server import PETLserver

pytestmark = pytest.mark.gpu

GEO = dict(D=100.0, L=64.0, H=64.0, T=2.0)


def server(psis=(0.0, 90.0)):
    s = PETLserver()
    s.print_warnings = False
    for psi in psis:
        s.add_planogram(psi, GEO["D"], GEO["L"], GEO["H"], GEO["T"])
    s.set_volume(32, 32, 32, 2.0)
    return s


def plans_of(oracle, psis=(0.0, 90.0)):
    return [oracle.planogram(psi, GEO["D"], GEO["L"], GEO["H"], GEO["T"]) for psi in psis]


def coincidences(n, seed):
    """End points on two panel pairs (y = +-50 mm and x = +-50 mm), z in the chest-wall frame [0, 64]."""
    rng = np.random.default_rng(seed)
    half = n // 2
    a = np.column_stack([rng.uniform(-40, 40, half), np.full(half, 50.0), rng.uniform(0, 64, half)])
    b = np.column_stack([rng.uniform(-40, 40, half), np.full(half, -50.0), rng.uniform(0, 64, half)])
    c = np.column_stack([np.full(n - half, 50.0), rng.uniform(-40, 40, n - half), rng.uniform(0, 64, n - half)])
    d = np.column_stack([np.full(n - half, -50.0), rng.uniform(-40, 40, n - half), rng.uniform(0, 64, n - half)])
    blue = np.vstack([a, c]).astype(np.float32)
    pink = np.vstack([b, d]).astype(np.float32)
    perm = rng.permutation(n)
    return blue[perm], pink[perm]


@pytest.mark.parametrize("nearest", [False, True])
@pytest.mark.parametrize("chest_wall", [True, False])
def test_bin_events_matches_oracle(oracle, nearest, chest_wall):
    blue, pink = coincidences(200000, 11)
    if not chest_wall:
        blue[:, 2] -= 32.0
        pink[:, 2] -= 32.0
    s = server()
    got, acc = s.bin_events(blue, pink, nearest=nearest, chest_wall=chest_wall)
    ref, racc = oracle.bin_events(plans_of(oracle), blue, pink, nearest=nearest, chest_wall=chest_wall)
    assert acc == racc and min(acc) > 1000
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_bin_events_adds_to_existing_values(oracle):
    blue, pink = coincidences(50000, 12)
    rng = np.random.default_rng(13)
    s = server()
    start = [rng.random(p.shape).astype(np.float32) * 3 for p in plans_of(oracle)]
    got, _ = s.bin_events(blue, pink, g=[x.copy() for x in start])
    ref, _ = oracle.bin_events(plans_of(oracle), blue, pink, g=[x.copy() for x in start])
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_hot_bin_keeps_event_order(oracle):
    """100000 coincidences with nearly the same LOR: the bilinear weights differ in their last bits, so the
    sums depend on the order in which they are added."""
    rng = np.random.default_rng(14)
    n = 100000
    blue = np.tile(np.array([[3.3, 50.0, 20.7]], np.float32), (n, 1))
    pink = np.tile(np.array([[-4.1, -50.0, 31.9]], np.float32), (n, 1))
    blue[:, 0] += rng.uniform(-1e-3, 1e-3, n).astype(np.float32)
    pink[:, 2] += rng.uniform(-1e-3, 1e-3, n).astype(np.float32)
    s = server((0.0,))
    got, _ = s.bin_events(blue, pink)
    ref, _ = oracle.bin_events(plans_of(oracle, (0.0,)), blue, pink)
    np.testing.assert_array_equal(got[0], ref[0])
    assert float(got[0].max()) > 1e4


def test_more_events_than_one_staging_chunk(oracle):
    blue, pink = coincidences((1 << 22) + 12345, 15)
    s = server((0.0,))
    got, acc = s.bin_events(blue, pink)
    ref, racc = oracle.bin_events(plans_of(oracle, (0.0,)), blue, pink)
    assert acc == racc
    np.testing.assert_array_equal(got[0], ref[0])


@pytest.mark.parametrize("layout", ["lst", "acq"])
def test_bin_file_matches_oracle_and_resets(oracle, tmp_path, layout):
    blue, pink = coincidences(20000, 16)
    path = tmp_path / f"events.{layout}"
    if layout == "lst":
        oracle.write_lst(path, blue, pink)
        rb, rp = oracle.read_lst(path)
    else:
        extra = np.random.default_rng(17).integers(0, 3, (len(blue), 2))
        oracle.write_acq(path, blue, pink, extra_interactions=extra)
        rb, rp = oracle.read_acq(path)
    s = server()
    g = [np.full(p.shape, 7.0, np.float32) for p in plans_of(oracle)]  # bin() starts from zero
    s.bin(str(path), g)
    ref, _ = oracle.bin_events(plans_of(oracle), rb, rp)
    for a, b in zip(g, ref):
        np.testing.assert_array_equal(a, b)


def test_bin_missing_file_leaves_data(oracle, tmp_path):
    s = server()
    g = [np.full(p.shape, 7.0, np.float32) for p in plans_of(oracle)]
    s.bin(str(tmp_path / "missing.lst"), g)
    assert all(np.all(x == 7.0) for x in g)


def test_bin_events_gpu_matches_host_path(oracle):
    torch = pytest.importorskip("torch")
    blue, pink = coincidences(100000, 18)
    s = server()
    host, _ = s.bin_events(blue, pink)
    g = [torch.zeros(p.shape, dtype=torch.float32, device="cuda:0") for p in plans_of(oracle)]
    s.bin_events_gpu(torch.from_numpy(blue).to("cuda:0"), torch.from_numpy(pink).to("cuda:0"), g)
    for a, b in zip(g, host):
        np.testing.assert_array_equal(a.cpu().numpy(), b)


@pytest.mark.parametrize("nearest", [False, True])
def test_bin_events_ragged_geometry(oracle, nearest):
    """Odd crystal counts (15 x 11, so L != H), odd view counts (37 v0, 9 v1), panels 72 mm apart:
    bit-exact against the oracle."""
    psis, D, L, H, T = (0.0, 90.0), 72.0, 30.0, 22.0, 2.0
    s = PETLserver()
    s.print_warnings = False
    for psi in psis:
        s.add_planogram(psi, D, L, H, T)
    s.set_volume(13, 11, 7, 2.0, 2.0, 1.0, -2.0, 3.0)
    plans = [oracle.planogram(psi, D, L, H, T) for psi in psis]
    rng = np.random.default_rng(19)
    n, R = 100000, D / 2
    half = n // 2
    a = np.column_stack([rng.uniform(-L / 2, L / 2, half), np.full(half, R), rng.uniform(0, H, half)])
    b = np.column_stack([rng.uniform(-L / 2, L / 2, half), np.full(half, -R), rng.uniform(0, H, half)])
    c = np.column_stack([np.full(n - half, R), rng.uniform(-L / 2, L / 2, n - half), rng.uniform(0, H, n - half)])
    d = np.column_stack([np.full(n - half, -R), rng.uniform(-L / 2, L / 2, n - half), rng.uniform(0, H, n - half)])
    blue = np.vstack([a, c]).astype(np.float32)
    pink = np.vstack([b, d]).astype(np.float32)
    got, acc = s.bin_events(blue, pink, nearest=nearest)
    ref, racc = oracle.bin_events(plans, blue, pink, nearest=nearest, chest_wall=True)
    assert acc == racc and min(acc) > 1000
    for x, y in zip(got, ref):
        np.testing.assert_array_equal(x, y)
