"""GPU parity of the response model (ray tracing, stopping power, solid angle, calc_response) against the
oracle, which restates the reference's kernels in the same float / double mix.

Tolerances: the two evaluate the same expressions in the same types with contraction off; they differ in
the device's expf / sqrt / division against the host's (an ulp or so) and in the order of nothing else, so
values agree to 1e-5 relative.  A ray that grazes a solid's face can flip an inside test on such an ulp;
the tests allow at most 1 bin in 20000 to differ by more, and report the count.
"""
import numpy as np
import pytest

from petl_amd.petl_server import PETLserver

pytestmark = pytest.mark.gpu

GEO = dict(psis=(0.0, 90.0), D=120.0, L=80.0, H=60.0, T=2.0)


def server(geo=GEO):
    s = PETLserver()
    s.print_warnings = False
    for psi in geo["psis"]:
        s.add_planogram(psi, geo["D"], geo["L"], geo["H"], geo["T"])
    s.set_volume(40, 40, 30, 2.0)
    return s


def plans_of(oracle, geo=GEO):
    return [oracle.planogram(psi, geo["D"], geo["L"], geo["H"], geo["T"]) for psi in geo["psis"]]


def modules(D, L, H, block=19.0, gap=1.0, depth=20.0):
    """Two opposed rings of box modules around a D-wide gap, as a PET/X-like scanner lays them out."""
    out = []
    nx = int(L // (block + gap))
    nz = int(H // (block + gap))
    for m in range(nz):
        cz = (m + 0.5) * (block + gap)
        for n in range(nx):
            cx = (n - (nx - 1) / 2) * (block + gap)
            out.append(([cx, D / 2 + depth / 2, cz], [block / 2, depth / 2, block / 2]))
            out.append(([cx, -D / 2 - depth / 2, cz], [block / 2, depth / 2, block / 2]))
            out.append(([D / 2 + depth / 2, cx, cz], [depth / 2, block / 2, block / 2]))
            out.append(([-D / 2 - depth / 2, cx, cz], [depth / 2, block / 2, block / 2]))
    return out


def compare(got, want, rtol=1e-5, atol=1e-7):
    bad = np.abs(got - want) > atol + rtol * np.abs(want)
    nbad = int(np.count_nonzero(bad))
    assert nbad <= max(1, got.size // 20000), (nbad, got.size, float(np.max(np.abs(got - want))))
    return nbad


@pytest.mark.parametrize("oversampling", [1, 3])
def test_stopping_power_matches_oracle(oracle, oversampling):
    s = server()
    mods = modules(GEO["D"], GEO["L"], GEO["H"])
    for c, r in mods:
        assert s.add_module(c, r, 0.0815)
    g = s.allocate_projections()
    assert s.libprojectors.stopping_power(s.floatstarstar(g), oversampling)
    solids = [oracle.solid(1, c, r, 0.0815) for c, r in mods]
    for p, gi in zip(plans_of(oracle), g):
        ref = oracle.ray_trace(p, solids, oversampling=oversampling, stopping=True)
        compare(gi, ref)
        assert float(gi.max()) > 0.1


def phantom_objects():
    rz = np.array([[np.cos(0.5), -np.sin(0.5), 0], [np.sin(0.5), np.cos(0.5), 0], [0, 0, 1.0]], np.float32)
    return [
        (0, [0, 0, 30.0], [30, 25, 28], 0.01, None, None),
        (1, [5, -3, 25.0], [8, 6, 10], 0.02, rz, None),
        (4, [-8, 6, 35.0], [5, 5, 12], 0.03, None, None),
        (2, [0, 10, 20.0], [15, 4, 3], 0.015, None, None),
        (7, [10, 0, 30.0], [2, 6, 8], 0.025, None, None),
        (6, [-10, 0, 30.0], [3, 7, 5], 0.02, None, None),
        (0, [0, 0, 30.0], [12, 12, 12], 0.005, None, [0, 0, 33.0]),
    ]


@pytest.mark.parametrize("oversampling", [1, 3])
def test_ray_trace_matches_oracle(oracle, oversampling):
    s = server()
    objs = phantom_objects()
    for t, c, r, v, A, clip in objs:
        assert s.add_object(t, c, r, v, A=A, clip=clip)
    g = s.ray_trace(oversampling=oversampling)
    solids = [oracle.solid(t, c, r, v, A=A, clip=clip) for t, c, r, v, A, clip in objs]
    for p, gi in zip(plans_of(oracle), g):
        ref = oracle.ray_trace(p, solids, oversampling=oversampling)
        compare(gi, ref)
        assert float(gi.max()) > 0.1


def test_many_overlapping_solids_take_the_streaming_path(oracle):
    """More solids on a ray than a thread keeps in registers: the end points are walked without storing
    them, and the result is the same."""
    s = server(dict(GEO, psis=(0.0,)))
    objs = [(0, [0, 0, 30.0], [4 + 2 * k, 4 + 2 * k, 4 + 2 * k], 0.01 * (k + 1)) for k in range(18)]
    objs += [(1, [0, 0, 30.0 + 2 * k], [30, 30, 0.5], 0.002) for k in range(-6, 7)]
    for t, c, r, v in objs:
        assert s.add_object(t, c, r, v)
    g = s.ray_trace()
    solids = [oracle.solid(t, c, r, v) for t, c, r, v in objs]
    ref = oracle.ray_trace(plans_of(oracle, dict(GEO, psis=(0.0,)))[0], solids)
    compare(g[0], ref)


def test_no_modules_gives_unit_stopping_power():
    s = server()
    g = s.stopping_power()
    for gi in g:
        np.testing.assert_array_equal(gi, np.ones_like(gi))


def test_solid_angle_matches_oracle(oracle):
    s = server()
    for inverse in (False, True):
        g = s.apply_solid_angle_correction(None, inverse)
        for p, gi in zip(plans_of(oracle), g):
            c = oracle.solid_angle_factor(p, inverse)
            np.testing.assert_array_equal(gi, np.broadcast_to(c[:, :, None, None], gi.shape))


def test_calc_response_matches_oracle_composition(oracle):
    s = server()
    mods = modules(GEO["D"], GEO["L"], GEO["H"])
    for c, r in mods:
        s.add_module(c, r)
    mu = np.zeros((30, 40, 40), np.float32)
    mu[5:25, 10:30, 10:30] = 0.0096
    resp = s.calc_response(3, mu=mu)
    solids = [oracle.solid(1, c, r, 0.0815) for c, r in mods]
    plans = plans_of(oracle)
    vol = oracle.volume(40, 40, 30, 2.0)
    Pmu = oracle.project(plans, vol, mu)
    for p, gi, pm in zip(plans, resp, Pmu):
        st = oracle.ray_trace(p, solids, oversampling=3, stopping=True)
        st = oracle.gaussian_filter(st, 0.6571).astype(np.float32)
        st = st * oracle.solid_angle_factor(p, True)[:, :, None, None]
        want = st / np.exp(pm.astype(np.float64))
        compare(gi, want, rtol=2e-5, atol=1e-9)


def test_device_entry_point_matches_host():
    torch = pytest.importorskip("torch")
    s = server()
    for c, r in modules(GEO["D"], GEO["L"], GEO["H"]):
        s.add_module(c, r)
    host = s.allocate_projections()
    assert s.libprojectors.stopping_power(s.floatstarstar(host), 3)
    dev = [torch.empty(a.shape, dtype=torch.float32, device="cuda") for a in host]
    s.ray_trace_gpu(dev, oversampling=3, stopping=True)
    for a, b in zip(host, dev):
        np.testing.assert_array_equal(b.cpu().numpy(), a)
