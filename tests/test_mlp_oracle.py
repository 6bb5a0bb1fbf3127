"""Pins the FP64 MLP oracle by central finite differences."""
import numpy as np

from oracle.mlp import loss_only, mlp_forward_backward


def test_mlp_oracle_gradient_by_finite_differences():
    rng = np.random.default_rng(0)
    H, nf, n = 6, 5, 40
    npar = H * nf + H + H * H + H + H + 1
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((n, nf))
    t = rng.standard_normal(n)
    _, L, g = mlp_forward_backward(p, X, t, H, nf)
    assert abs(L - loss_only(p, X, t, H, nf)) < 1e-15
    h = 1e-6
    fd = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (loss_only(p + e, X, t, H, nf) - loss_only(p - e, X, t, H, nf)) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-7)
