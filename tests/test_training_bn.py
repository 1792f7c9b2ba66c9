"""The training networks' batch norm (plain tensor operations, so that GPU
training avoids the MIOpen batch-norm kernels) is torch.nn.BatchNorm: same
outputs, gradients, running statistics and state-dict names."""
import numpy as np
import pytest


@pytest.mark.parametrize("dims", [2, 3])
def test_batchnorm_matches_torch(dims):
    import torch
    from deeprwcap_amd import training as T
    torch.manual_seed(dims)
    ref = (torch.nn.BatchNorm2d if dims == 2 else torch.nn.BatchNorm3d)(5).double()
    mine = T._BatchNorm(5).double()
    with torch.no_grad():
        ref.weight.uniform_(0.5, 1.5)
        ref.bias.uniform_(-0.5, 0.5)
    mine.load_state_dict(ref.state_dict())
    assert list(mine.state_dict()) == list(ref.state_dict())
    for _ in range(3):
        x = torch.randn(7, 5, *([6] * dims), dtype=torch.float64, requires_grad=True)
        x2 = x.detach().clone().requires_grad_(True)
        a, b = ref(x), mine(x2)
        g = torch.randn_like(a)
        a.backward(g)
        b.backward(g)
        assert torch.allclose(a, b, atol=1e-12) and torch.allclose(x.grad, x2.grad, atol=1e-12)
        assert torch.allclose(ref.weight.grad, mine.weight.grad, atol=1e-12)
    for k in ("running_mean", "running_var", "num_batches_tracked"):
        assert torch.allclose(getattr(ref, k).double(), getattr(mine, k).double(), atol=1e-12), k
    ref.eval()
    mine.eval()
    x = torch.randn(3, 5, *([6] * dims), dtype=torch.float64)
    assert np.allclose(ref(x).detach().numpy(), mine(x).detach().numpy(), atol=1e-12)
