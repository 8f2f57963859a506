"""Reconstruction service (petl_amd.serving) on the GPU.

Every request's image must be bit-identical to PETLserver.MLEM on the same inputs run alone: the service
runs the same kernels, each request on its own buffers, and the kernels have a fixed reduction order (no
atomics), so neither the mix nor the replay of captured graphs may change a bit.
"""
import threading
import time

import numpy as np
import pytest

from petl_amd.petl_server import PETLserver
from petl_amd.serving import ReconService
from petl_amd.workloads import ellipsoid_phantom

pytestmark = pytest.mark.gpu

GEOS = {
    "p32": dict(psis=(0.0,), D=100.0, L=64.0, H=64.0, T=2.0, n=32, voxel=2.0),
    "p48": dict(psis=(0.0, 90.0), D=200.0, L=96.0, H=96.0, T=2.0, n=48, voxel=2.0),
    "p64": dict(psis=(0.0, 90.0), D=200.0, L=128.0, H=128.0, T=2.0, n=64, voxel=2.0),
    "p96": dict(psis=(0.0, 90.0), D=200.0, L=192.0, H=192.0, T=2.0, n=96, voxel=2.0),
    "p128": dict(psis=(0.0, 90.0), D=200.0, L=256.0, H=256.0, T=2.0, n=128, voxel=2.0),
    # oblique panels: both planograms take the general kernels
    "o48": dict(psis=(30.0, 120.0), D=200.0, L=96.0, H=96.0, T=2.0, n=48, voxel=2.0),
}


def server_for(geo):
    s = PETLserver()
    s.print_warnings = False
    for psi in geo["psis"]:
        s.add_planogram(psi, geo["D"], geo["L"], geo["H"], geo["T"])
    s.set_volume(geo["n"], geo["n"], geo["n"], geo["voxel"])
    return s


_data_cache = {}


def data_for(name, seed, counts=1e6):
    key = (name, seed, counts)
    if key not in _data_cache:
        geo = GEOS[name]
        s = server_for(geo)
        f = ellipsoid_phantom(geo["n"], geo["voxel"], 6, seed=seed)
        g = s.project(s.allocate_projections(), f)
        tot = sum(float(x.sum()) for x in g)
        rng = np.random.default_rng(seed + 100)
        _data_cache[key] = [rng.poisson(x * (counts / tot)).astype(np.float32) for x in g]
    return _data_cache[key]


def alone(name, y, num_iter, prior=None, beta=0.0, delta=2.0, p=1.0, num_subsets=1):
    s = server_for(GEOS[name])
    return s.MLEM(y, None, num_iter, delta=delta, beta=beta, prior=prior or "rd", p=p, num_subsets=num_subsets)


def kw(name, y, num_iter, **extra):
    return dict(GEOS[name], y=y, num_iter=num_iter, **extra)


MIX = [
    ("p32", 1, 10, {}),
    ("p48", 2, 5, dict(prior="rd", beta=0.01, delta=2.0)),
    ("p64", 3, 4, dict(num_subsets=4)),
    ("p32", 4, 6, dict(prior="tv", beta=0.005, delta=0.05, p=1.0)),
    ("p48", 5, 3, {}),
    ("p64", 6, 2, {}),
]


def expected(case):
    name, seed, it, extra = case
    return alone(name, data_for(name, seed), it, **extra)


def test_mix_matches_alone_bit_exact():
    want = [expected(c) for c in MIX]
    with ReconService(max_slots=4) as svc:
        ids = [svc.submit(**kw(n, data_for(n, sd), it, **ex)) for (n, sd, it, ex) in MIX]
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["completed"] == len(MIX) and st["failed"] == 0
    assert st["peak_active"] >= 2


@pytest.mark.parametrize("iters", [1, 2, 3, 4])
def test_each_graph_replay_matches_eager(iters):
    """k replays of the captured iteration equal k eager iterations, for k = 1..4."""
    y = data_for("p48", 2)
    want = alone("p48", y, iters, prior="rd", beta=0.01)
    with ReconService(max_slots=2) as svc:
        rid = svc.submit(**kw("p48", y, iters, prior="rd", beta=0.01))
        got, _ = svc.wait(rid)
    np.testing.assert_array_equal(got, want)


def test_hbm_budget_admits_one_at_a_time():
    cases = MIX[:3]
    with ReconService(max_slots=8) as probe:
        sizes = [probe.request_bytes(**kw(n, data_for(n, sd), it, **ex)) for (n, sd, it, ex) in cases]
    budget = 1.2 * max(sizes)
    assert budget < sum(sorted(sizes)[-2:])  # no two of the larger ones fit together
    want = [expected(c) for c in cases]
    with ReconService(hbm_budget_bytes=budget, max_slots=8) as svc:
        ids = [svc.submit(**kw(n, data_for(n, sd), it, **ex)) for (n, sd, it, ex) in cases]
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["peak_reserved_bytes"] <= budget


def test_requests_join_while_others_run():
    """A long request is running when short ones arrive; they are admitted beside it and finish first."""
    y_long = data_for("p64", 7)
    y_short = data_for("p32", 8)
    want_long = alone("p64", y_long, 300)   # ~1 ms per iteration: runs for a few hundred ms
    want_short = alone("p32", y_short, 3)
    with ReconService(max_slots=4) as svc:
        t0 = time.perf_counter()
        rid_long = svc.submit(**kw("p64", y_long, 300))
        done = {}

        def wait_long():
            done["long"] = svc.wait(rid_long)
            done["t_long"] = time.perf_counter() - t0

        th = threading.Thread(target=wait_long)
        th.start()
        time.sleep(0.02)
        short_ids = [svc.submit(**kw("p32", y_short, 3)) for _ in range(3)]
        shorts = [svc.wait(i) for i in short_ids]
        t_short = time.perf_counter() - t0
        th.join()
        st = svc.stats()
    np.testing.assert_array_equal(done["long"][0], want_long)
    for img, _ in shorts:
        np.testing.assert_array_equal(img, want_short)
    assert st["peak_active"] >= 2
    assert t_short < done["t_long"]


def test_oversized_request_is_rejected():
    y = data_for("p32", 1)
    with ReconService(max_slots=2) as probe:
        need = probe.request_bytes(**kw("p32", y, 2))
    with ReconService(hbm_budget_bytes=0.5 * need, max_slots=2) as svc:
        with pytest.raises(ValueError):
            svc.submit(**kw("p32", y, 2))


def test_pooled_resources_are_reused_bit_exact():
    """Requests with the same geometry and MLEM parameters reuse a finished request's model, buffers and
    captured iteration; new data and iteration counts still give the images MLEM gives alone."""
    cases = [("p48", 21, 3), ("p48", 22, 6), ("p48", 23, 2)]
    want = [alone(n, data_for(n, sd), it) for (n, sd, it) in cases]
    with ReconService(max_slots=4) as svc:
        got = []
        for (n, sd, it) in cases:
            rid = svc.submit(**kw(n, data_for(n, sd), it))
            got.append(svc.wait(rid)[0])
        st = svc.stats()
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["pool_hits"] == 2


def test_bf16_storage_requests_match_mlem_storage():
    y = data_for("p48", 31)
    want = server_for(GEOS["p48"]).MLEM(y, None, 4, storage="bf16", image_storage="f32")
    want16 = server_for(GEOS["p48"]).MLEM(y, None, 3, storage="f16")
    with ReconService(max_slots=4) as svc:
        a = svc.submit(**kw("p48", y, 4, storage="bf16", image_storage="f32"))
        b = svc.submit(**kw("p48", y, 3, storage="f16"))
        got, got16 = svc.wait(a)[0], svc.wait(b)[0]
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got16, want16)


def test_f16_pooled_request_with_other_scales():
    """Two binary16 requests with the same geometry and parameters but counts 100x apart: the second reuses
    the first one's pooled resources, picks its own data and image scales, and must still give the image
    MLEM gives alone (the captured iteration holds the scales, so it is captured again)."""
    y_lo = data_for("p48", 41, counts=1e5)
    y_hi = data_for("p48", 42, counts=1e7)
    want = [server_for(GEOS["p48"]).MLEM(y, None, 3, storage="f16") for y in (y_lo, y_hi)]
    with ReconService(max_slots=4) as svc:
        got = []
        for y in (y_lo, y_hi):
            rid = svc.submit(**kw("p48", y, 3, storage="f16"))
            got.append(svc.wait(rid)[0])
        st = svc.stats()
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["pool_hits"] == 1 and st["recaptures"] == 1


@pytest.mark.parametrize("storage,image", [("f32", "f32"), ("bf16", "f32"), ("bf16", "bf16"), ("f16", "f16")])
def test_admission_footprint_covers_the_buffers(storage, image):
    """The bytes admission counts for a request cover what its buffers hold once allocated, with no more
    than the fixed allowance for kernel tables and reduction scratch on top."""
    y = data_for("p64", 51)
    with ReconService(max_slots=1) as svc:
        counted, held = svc.request_footprint(**kw("p64", y, 2, storage=storage, image_storage=image))
    assert held <= counted <= held + (1 << 20) + 0.02 * held


def test_poisson_arrivals_over_all_panel_sizes():
    """A seeded Poisson-arrival stream (20 requests/s) of requests with 48, 64, 96 and 128-crystal panels and
    matched volumes, counts log-uniform in [1e5, 1e8], 2-5 iterations, some with the relative-difference
    prior: requests join while others run, and every image is bit-identical to MLEM run alone."""
    rng = np.random.default_rng(77)
    names = ["p48", "p64", "p96", "p128"]
    cases = []
    for k in range(12):
        name = names[k % 4] if k < 4 else names[int(rng.integers(4))]
        counts = float(10.0 ** rng.uniform(5.0, 8.0))
        extra = dict(prior="rd", beta=0.01, delta=2.0) if rng.random() < 0.3 else {}
        cases.append((name, 60 + k, int(rng.integers(2, 6)), counts, extra))
    gaps = rng.exponential(1.0 / 20.0, size=len(cases))
    want = [alone(n, data_for(n, sd, c), it, **ex) for (n, sd, it, c, ex) in cases]
    with ReconService(max_slots=8) as svc:
        ids = []
        for gap, (n, sd, it, c, ex) in zip(gaps, cases):
            time.sleep(gap)
            ids.append(svc.submit(**kw(n, data_for(n, sd, c), it, **ex)))
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
    for (n, *_), g, w in zip(cases, got, want):
        assert g.shape == (GEOS[n]["n"],) * 3
        np.testing.assert_array_equal(g, w)
    assert st["completed"] == len(cases) and st["failed"] == 0
    assert st["peak_active"] >= 2


# ---------------------------------------------------------------------------------------------------------
# batched iterations (ReconService(batched=True))
# ---------------------------------------------------------------------------------------------------------
BATCH_MIX = [
    ("p48", 81, 3, {}),
    ("p128", 82, 2, dict(prior="rd", beta=0.01, delta=2.0)),
    ("p64", 83, 5, dict(prior="rd", beta=0.01, delta=2.0)),
    ("p96", 84, 4, {}),
    ("p48", 85, 6, dict(prior="tv", beta=0.005, delta=0.05, p=1.0)),
    ("p64", 86, 1, {}),
    ("p128", 87, 3, {}),
]


def test_batched_iterations_match_alone_bit_exact():
    """Requests of four panel sizes, with and without priors and with odd and even iteration counts, advance
    together in one launch per phase; every image equals MLEM alone bit for bit."""
    want = [alone(n, data_for(n, sd), it, **ex) for (n, sd, it, ex) in BATCH_MIX]
    with ReconService(max_slots=8, batched=True) as svc:
        ids = [svc.submit(**kw(n, data_for(n, sd), it, **ex)) for (n, sd, it, ex) in BATCH_MIX]
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
    for (n, *_), g, w in zip(BATCH_MIX, got, want):
        assert g.shape == (GEOS[n]["n"],) * 3
        np.testing.assert_array_equal(g, w)
    assert st["batched_requests"] == len(BATCH_MIX) and st["failed"] == 0
    assert max(it for *_, it, _ in BATCH_MIX) <= st["batched_rounds"] <= sum(it for *_, it, _ in BATCH_MIX)


def test_batched_rounds_advance_several_requests():
    """Requests that are on the device at the same time share the batched iterations: one launch of the
    captured phases advances all of them (here a 48, a 64 and a 96-crystal request of 40 to 60 iterations,
    admitted within a few hundred milliseconds of each other)."""
    cases = [("p64", 111, 60), ("p48", 112, 50), ("p96", 113, 40)]
    want = [alone(n, data_for(n, sd), it) for (n, sd, it) in cases]
    with ReconService(max_slots=8, batched=True) as svc:
        ids = [svc.submit(**kw(n, data_for(n, sd), it)) for (n, sd, it) in cases]
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["batched_peak"] >= 2
    assert st["batched_rounds"] < sum(it for *_, it in cases)


def test_batched_service_runs_other_requests_on_their_own_graphs():
    """OSEM, binary16 storage and one-planogram geometries are outside the batched form: in a batched service
    they replay their own captured iteration beside the batched ones, with the images MLEM gives alone."""
    y48, y64, y32 = data_for("p48", 91), data_for("p64", 92), data_for("p32", 93)
    want = [alone("p48", y48, 3), alone("p64", y64, 2, num_subsets=4),
            server_for(GEOS["p48"]).MLEM(y48, None, 2, storage="bf16", image_storage="f32"), alone("p32", y32, 4)]
    with ReconService(max_slots=8, batched=True) as svc:
        ids = [svc.submit(**kw("p48", y48, 3)), svc.submit(**kw("p64", y64, 2, num_subsets=4)),
               svc.submit(**kw("p48", y48, 2, storage="bf16", image_storage="f32")), svc.submit(**kw("p32", y32, 4))]
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["batched_requests"] == 2 and st["failed"] == 0  # p48 fp32 and the one-planogram p32


def test_batched_poisson_arrivals_and_pool():
    """A Poisson stream into the batched service: requests join the running rounds as they are admitted,
    later ones reuse pooled resources, and every image equals MLEM alone."""
    rng = np.random.default_rng(95)
    names = ["p48", "p64", "p96", "p128"]
    cases = [(names[k % 4], 100 + (k % 6), int(rng.integers(2, 6))) for k in range(10)]
    gaps = rng.exponential(1.0 / 30.0, size=len(cases))
    want = [alone(n, data_for(n, sd), it) for (n, sd, it) in cases]
    with ReconService(max_slots=6, batched=True) as svc:
        ids = []
        for gap, (n, sd, it) in zip(gaps, cases):
            time.sleep(gap)
            ids.append(svc.submit(**kw(n, data_for(n, sd), it)))
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["batched_requests"] == len(cases) and st["pool_hits"] >= 1


def test_general_kernel_requests():
    """Requests on oblique panels (the general kernels, whose volume copies and partial volumes are allocated
    before the iteration is captured) beside fast-kernel requests: every image bit-identical to MLEM alone,
    with and without the RD prior, and the admission count covers the buffers held."""
    cases = [("o48", 81, 3, {}), ("p48", 82, 3, {}), ("o48", 83, 2, dict(prior="rd", beta=0.01, delta=2.0))]
    want = [expected(c) for c in cases]
    with ReconService(max_slots=4) as svc:
        ids = [svc.submit(**kw(n, data_for(n, sd), it, **ex)) for (n, sd, it, ex) in cases]
        got = [svc.wait(i)[0] for i in ids]
        st = svc.stats()
        counted, held = svc.request_footprint(**kw("o48", data_for("o48", 81), 3))
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    assert st["completed"] == len(cases) and st["failed"] == 0
    assert held <= counted
