"""Gradients through the projector (petl_amd.autograd) on the GPU.

Workload: a batch of 8 seeded 64^3 float32 volumes at 2 mm, two 64 x 64-crystal panels (psi 0 and 90 deg,
D = 200 mm) -> planograms (63, 101, 64, 164); Poisson data at 1e7 counts per volume; loss = the negative
Poisson log-likelihood sum(Af - y log Af).

* central finite differences on 64 seeded voxels: 1e-3 relative L2 over the 64 derivatives.  The loss
  difference is summed over bins (bins the voxel does not touch cancel exactly), so fp32 rounding of Af
  sets the error.  At 1e7 counts over 1.3e8 bins the counts are 0, 1 or 2 where Af ~ 0.08, so each
  derivative sums terms of +-12 to a value of order 10-100; with a step of 1e-3 x the mean voxel value
  that rounding reaches 5 % on such voxels.  The step is 2e-2 x the mean voxel value (truncation error
  of the central difference ~1e-6 relative);
* the gradient equals the oracle's matched back projection of the residual 1 - y / Af (oracle forward
  and back in fp32): 1e-4 relative L2;
* <A x, y> = <x, A^T y> through torch.autograd.grad, to 1e-6.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from petl_amd import autograd as pa  # noqa: E402
from petl_amd.petl_server import PETLserver  # noqa: E402
from petl_amd.workloads import ellipsoid_phantom, setup_server  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
B = 8
N = 64


@pytest.fixture(scope="module")
def setup():
    s = PETLserver()
    s.print_warnings = False
    setup_server(s, "panel64")
    shapes = [tuple(int(x) for x in s.get_planogramSize(i)) for i in range(2)]
    assert shapes == [(63, 101, 64, 164)] * 2
    assert all(s.uses_fast_path(i) for i in range(2))
    f_true = np.stack([ellipsoid_phantom(N, 2.0, 6, seed=3 + b) for b in range(B)])
    ft = torch.from_numpy(f_true).to(DEV)
    with torch.no_grad():
        g = pa.project(s, ft)
        tot = sum(x.double().sum(dim=(1, 2, 3, 4)) for x in g)
    scale = (1e7 / tot).float()                        # 1e7 expected counts per volume
    ft = ft * scale.view(B, 1, 1, 1)
    gen = torch.Generator(device=DEV).manual_seed(102)
    with torch.no_grad():
        y = [torch.poisson(x * scale.view(B, 1, 1, 1, 1), generator=gen) for x in g]
    # evaluation point: away from the data's optimum, positive everywhere in the field of view
    f0 = (0.8 * ft + 0.2 * ft.mean(dim=(1, 2, 3), keepdim=True)).contiguous()
    return s, ft, f0, y


def loss_of(s, f, y):
    q = pa.project(s, f)
    return sum(pa.poisson_nll(qi, yi) for qi, yi in zip(q, y))


def test_gradient_matches_finite_differences(setup):
    s, _, f0, y = setup
    f = f0.clone().requires_grad_(True)
    loss_of(s, f, y).backward()
    grad = f.grad.detach()
    rng = np.random.default_rng(7)
    # 64 voxels inside the central 60 % of the field of view, spread over the batch
    idx = [(int(rng.integers(B)),) + tuple(int(v) for v in rng.integers(int(0.2 * N), int(0.8 * N), 3))
           for _ in range(64)]
    h = 2e-2 * float(f0.mean())
    fd, ad = [], []
    with torch.no_grad():
        for (b, k, j, i) in idx:
            fb = f0[b:b + 1].clone()
            fb[0, k, j, i] += h
            qp = pa.project(s, fb)
            fb[0, k, j, i] -= 2 * h
            qm = pa.project(s, fb)
            diff = torch.zeros((), dtype=torch.float64, device=DEV)
            for a, c, yy in zip(qp, qm, y):
                lp = torch.where(a > 0, a.double() - yy[b:b + 1].double() * torch.log(a.double().clamp_min(1e-30)),
                                 torch.zeros_like(a, dtype=torch.float64))
                lm = torch.where(c > 0, c.double() - yy[b:b + 1].double() * torch.log(c.double().clamp_min(1e-30)),
                                 torch.zeros_like(c, dtype=torch.float64))
                diff += (lp - lm).sum()
            fd.append(float(diff) / (2 * h))
            ad.append(float(grad[b, k, j, i]))
    fd, ad = np.array(fd), np.array(ad)
    assert np.linalg.norm(fd - ad) / np.linalg.norm(fd) < 1e-3, (fd[:4], ad[:4])


def test_gradient_matches_oracle_residual_backprojection(setup, oracle):
    s, _, f0, y = setup
    b = 2
    f = f0[b:b + 1].clone().requires_grad_(True)
    loss = sum(pa.poisson_nll(qi, yi[b:b + 1]) for qi, yi in zip(pa.project(s, f), y))
    loss.backward()
    got = f.grad[0].cpu().numpy()
    plans = [oracle.planogram(psi, 200.0, 128.0, 128.0, 2.0) for psi in (0.0, 90.0)]
    vol = oracle.volume(N, N, N, 2.0)
    fh = f0[b].cpu().numpy()
    q = oracle.project(plans, vol, fh)
    r = []
    for qi, yi in zip(q, y):
        yh = yi[b].cpu().numpy()
        ri = np.zeros_like(qi)
        pos = qi > 0
        ri[pos] = 1.0 - yh[pos] / qi[pos]
        r.append(ri)
    ref = oracle.backproject(plans, vol, r, mode="matched")
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4


def test_adjoint_through_autograd(setup):
    s, _, _, _ = setup
    gen = torch.Generator(device=DEV).manual_seed(9)
    x = torch.rand((2, N, N, N), generator=gen, device=DEV).requires_grad_(True)
    Ax = pa.project(s, x)
    yv = [torch.rand(a.shape, generator=gen, device=DEV) for a in Ax]
    (ATy,) = torch.autograd.grad(Ax, x, grad_outputs=yv)
    lhs = sum(float(torch.dot(a.detach().reshape(-1).double(), c.reshape(-1).double())) for a, c in zip(Ax, yv))
    rhs = float(torch.dot(x.detach().reshape(-1).double(), ATy.reshape(-1).double()))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_unbatched_equals_batched(setup):
    s, _, f0, _ = setup
    with torch.no_grad():
        a = pa.project(s, f0[3])
        c = pa.project(s, f0)
    for x, z in zip(a, c):
        assert torch.equal(x, z[3])


def test_backward_restores_back_projector_mode(setup):
    s, _, f0, y = setup
    s.set_backprojector_mode("petl")
    f = f0[:1].clone().requires_grad_(True)
    sum(pa.poisson_nll(qi, yi[:1]) for qi, yi in zip(pa.project(s, f), y)).backward()
    assert s.get_backprojector_mode() == "petl"


def test_gradient_ragged_geometry_matches_oracle(oracle):
    """A batch of 3 volumes at odd crystal counts (15 x 11), odd view counts (37 v0, 9 v1) and an
    odd-sided volume off centre on every axis: each gradient of the Poisson loss equals the oracle's
    matched back projection of its residual 1 - y / Af, to 1e-4."""
    psis, D, L, H, T = (0.0, 90.0), 72.0, 30.0, 22.0, 2.0
    s = PETLserver()
    s.print_warnings = False
    for psi in psis:
        s.add_planogram(psi, D, L, H, T)
    s.set_volume(13, 11, 7, 2.0, 2.0, 1.0, -2.0, 3.0)
    assert all(s.uses_fast_path(i) for i in range(2))
    plans = [oracle.planogram(psi, D, L, H, T) for psi in psis]
    vol = oracle.volume(13, 11, 7, 2.0, 2.0, 1.0, -2.0, 3.0)
    rng = np.random.default_rng(103)
    fh = (0.5 + rng.random((3,) + vol.shape)).astype(np.float32)
    qh = [oracle.project(plans, vol, fh[b]) for b in range(3)]
    yh = [[rng.poisson(qi).astype(np.float32) for qi in q] for q in qh]
    f = torch.from_numpy(fh).to(DEV).requires_grad_(True)
    y = [torch.from_numpy(np.stack([yh[b][i] for b in range(3)])).to(DEV) for i in range(2)]
    loss = sum(pa.poisson_nll(qi, yi) for qi, yi in zip(pa.project(s, f), y))
    loss.backward()
    for b in range(3):
        r = []
        for qi, yi in zip(qh[b], yh[b]):
            ri = np.zeros_like(qi)
            pos = qi > 0
            ri[pos] = 1.0 - yi[pos] / qi[pos]
            r.append(ri)
        ref = oracle.backproject(plans, vol, r, mode="matched")
        got = f.grad[b].cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4


def test_batched_launch_equals_per_volume_calls(setup):
    """The batch axis of the fast kernels' grids computes every element as its own call would, bit for bit,
    for the forward and for both back projector modes; the backward's mode is an argument of the call and
    the server's own mode is left alone."""
    s, ft, _, _ = setup
    shapes = [tuple(int(x) for x in s.get_planogramSize(i)) for i in range(2)]
    gb = [torch.empty((B,) + sh, device=DEV) for sh in shapes]
    s.project_gpu_batch(gb, ft)
    gen = torch.Generator(device=DEV).manual_seed(71)
    r = [torch.rand((B,) + sh, generator=gen, device=DEV) for sh in shapes]
    s.set_backprojector_mode("petl")
    for mode in ("matched", "petl"):
        fb = torch.empty_like(ft)
        s.backproject_gpu_batch(r, fb, mode=mode)
        assert s.get_backprojector_mode() == "petl"
        s.set_backprojector_mode(mode)
        for b in range(B):
            one = [torch.empty(sh, device=DEV) for sh in shapes]
            s.project_gpu(one, ft[b].contiguous())
            for x, y in zip(one, gb):
                assert torch.equal(x, y[b])
            fo = torch.empty_like(ft[b])
            s.backproject_gpu([x[b].contiguous() for x in r], fo)
            assert torch.equal(fo, fb[b])
        s.set_backprojector_mode("petl")
