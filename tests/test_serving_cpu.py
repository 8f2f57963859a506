"""Host-side checks of the reconstruction service that need no GPU: request validation, the admission
footprint, the service's thread lifecycle, and that a request that cannot run fails cleanly."""
import numpy as np
import pytest

from petl_amd.serving import ReconService, make_request

P32 = dict(psis=(0.0,), D=100.0, L=64.0, H=64.0, T=2.0, n=32, voxel=2.0)
P64 = dict(psis=(0.0, 90.0), D=200.0, L=128.0, H=128.0, T=2.0, n=64, voxel=2.0)


def zeros_for(shapes):
    return [np.zeros(s, np.float32) for s in shapes]


def test_make_request_checks_counts():
    with pytest.raises(ValueError):
        make_request(y=zeros_for([(31, 51, 32, 82)] * 2), num_iter=1, **P32)
    with pytest.raises(ValueError):
        make_request(y=[], num_iter=1, **dict(P32, psis=()))


def test_request_bytes_grow_with_size_and_subsets():
    with ReconService(hbm_budget_bytes=1e12, max_slots=2) as svc:
        b32 = svc.request_bytes(y=zeros_for([(31, 51, 32, 82)]), num_iter=1, **P32)
        b64 = svc.request_bytes(y=zeros_for([(63, 101, 64, 164)] * 2), num_iter=1, **P64)
        b64s = svc.request_bytes(y=zeros_for([(63, 101, 64, 164)] * 2), num_iter=1, num_subsets=4, **P64)
    planogram_bytes_64 = 2 * 63 * 101 * 64 * 164 * 4
    assert b64 > b32 > 0
    assert b64 >= 2 * planogram_bytes_64          # data + ratio planograms at least
    assert b64s - b64 == 3 * 64 ** 3 * 4           # one sensitivity volume per extra subset


def test_service_starts_and_stops_without_requests():
    svc = ReconService(hbm_budget_bytes=1e9, max_slots=3)
    assert svc.stats()["completed"] == 0 and svc.stats()["budget_bytes"] == 1e9
    svc.close()


def test_request_larger_than_budget_is_rejected():
    with ReconService(hbm_budget_bytes=1e6, max_slots=2) as svc:
        with pytest.raises(ValueError):
            svc.submit(y=zeros_for([(31, 51, 32, 82)]), num_iter=1, **P32)


def test_invalid_geometry_is_rejected():
    with ReconService(hbm_budget_bytes=1e12, max_slots=2) as svc:
        with pytest.raises(ValueError):
            svc.submit(y=zeros_for([(31, 51, 32, 82)]), num_iter=1, **dict(P32, T=-1.0))


def test_request_that_cannot_run_fails_cleanly():
    """Without a GPU, admission fails: wait() raises, the service counts the failure and stays usable."""
    torch = pytest.importorskip("torch")
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present; this checks the failure path on a host without one")
    y = zeros_for([(31, 51, 32, 82)])
    with ReconService(hbm_budget_bytes=1e12, max_slots=2) as svc:
        rid = svc.submit(y=y, num_iter=1, **P32)
        with pytest.raises(RuntimeError):
            svc.wait(rid)
        rid2 = svc.submit(y=y, num_iter=1, **P32)
        with pytest.raises(RuntimeError):
            svc.wait(rid2)
        st = svc.stats()
    assert st["failed"] == 2 and st["completed"] == 0
