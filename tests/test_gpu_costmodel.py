"""The MLP cost model's HIP forward / backward kernels against an FP64 numpy
restatement of the same network (tests/mlp_ref.py), and one training run.

Bar: predictions within 1e-5 relative, gradients within 1e-4 relative
(FP32 device arithmetic against FP64); training lowers the loss.
"""
import numpy as np
import pytest

from oracle.mlp import mlp_forward_backward

pytestmark = pytest.mark.gpu


def test_forward_and_gradients(ctx, torch_cuda):
    torch = torch_cuda
    from bbts_amd.device import CostModel
    cm = CostModel(ctx, hidden=32, seed=3)
    rng = np.random.default_rng(0)
    n, nf = 300, CostModel.num_features()
    X = rng.standard_normal((n, nf)).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    p = cm.get_params()
    tX, ty = torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda()
    pred = torch.zeros(n, dtype=torch.float32, device="cuda")
    cm.predict(tX, pred)
    grads = torch.zeros(cm.num_params(), dtype=torch.float32, device="cuda")
    loss = torch.zeros(4, dtype=torch.float32, device="cuda")
    cm.grad(tX, ty, grads, loss)
    yhat, L, g = mlp_forward_backward(p, X.astype(np.float64), y.astype(np.float64), 32, nf)
    assert np.allclose(pred.cpu().numpy(), yhat, rtol=1e-5, atol=1e-5)
    assert abs(loss[0].item() - L) <= 1e-5 * max(1.0, L)
    gd = grads.cpu().numpy()
    assert np.linalg.norm(gd - g) <= 1e-4 * np.linalg.norm(g)
    cm.close()


def test_training_reduces_loss(ctx, torch_cuda):
    torch = torch_cuda
    from bbts_amd.device import CostModel
    cm = CostModel(ctx, hidden=32, seed=4)
    rng = np.random.default_rng(1)
    n, nf = 2048, CostModel.num_features()
    X = rng.uniform(0, 1, (n, nf)).astype(np.float32)
    y = (2.0 * X[:, 0] + np.sin(3 * X[:, 1]) + 0.5).astype(np.float32)
    tX, ty = torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda()
    first = cm.train(tX, ty, batch=256, lr=3e-3, epochs=1)
    last = cm.train(tX, ty, batch=256, lr=3e-3, epochs=60)
    assert last < 0.1 * first
    cm.close()
