"""Host-side checks of petl_amd.autograd that need no GPU."""
import pytest

torch = pytest.importorskip("torch")

from petl_amd import autograd as pa  # noqa: E402
from petl_amd.petl_server import PETLserver  # noqa: E402


def server():
    s = PETLserver()
    s.print_warnings = False
    s.add_planogram(0.0, 100.0, 64.0, 64.0, 2.0)
    s.set_volume(32, 32, 32, 2.0)
    return s


def test_rejects_host_tensor():
    with pytest.raises(TypeError):
        pa.project(server(), torch.zeros((32, 32, 32)))


def test_rejects_non_server():
    with pytest.raises(TypeError):
        pa.project(object(), torch.zeros((32, 32, 32)))


def test_poisson_nll_closed_form():
    q = torch.tensor([1.0, 2.0, 0.0, 4.0])
    y = torch.tensor([1.0, 0.0, 3.0, 2.0])
    want = (1.0 - 0.0) + (2.0 - 0.0) + 0.0 + (4.0 - 2.0 * torch.log(torch.tensor(4.0, dtype=torch.float64)))
    assert float(pa.poisson_nll(q, y)) == pytest.approx(float(want), rel=1e-12)


def test_poisson_nll_gradient():
    q = torch.tensor([1.0, 2.0, 0.0, 4.0], dtype=torch.float64, requires_grad=True)
    y = torch.tensor([1.0, 0.0, 3.0, 2.0], dtype=torch.float64)
    pa.poisson_nll(q, y).backward()
    assert q.grad.tolist() == pytest.approx([0.0, 1.0, 0.0, 0.5])
