"""The torch restatement of the forward map (oracle/torch_ref.py, the autograd reference for the
adjoint) computes the C oracle's right-hand side and RK4 steps (CPU)."""
import ctypes

import numpy as np
import pytest

A, G0, OM = 0.6371229e7, 9.80616, 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)


def _state(G, N, seed):
    from cubedsphere_sbp_swm_amd import _capi
    n = G.n
    pitch = (n + 7) // 8 * 8
    H, U, V = (np.zeros((6, n, pitch)) for _ in range(3))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    assert _capi.lib().swmh_ic_sh(N, pitch, seed, 8, 1.0e4, 100.0, 40.0, dp(H), dp(U), dp(V), 4) == 0
    return [G.embed(X[:, :, :n]) for X in (H, U, V)]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("N,order", [(16, 42), (16, 43), (24, 63)])
def test_torch_restatement_matches_oracle(oracle_lib, N, order):
    """RHS to 5e-13 relative, 3 RK4 steps to 1e-11.  The torch side sums the SBP rows as matrix
    products in another order; the h tendency of a non-divergent flow is ~1e-3 of its terms and
    the d63 closure rows have entries of size ~10, so reordering shows at ~1e-13."""
    import torch
    from oracle.torch_ref import TorchSWM
    G = oracle_lib.Grid(N, order, order, A, OM, G0)
    T = TorchSWM(G, order)
    h, u, v = _state(G, N, 3)
    ref = G.rhs(h.copy(), u.copy(), v.copy())
    got = T.rhs(*(torch.tensor(G.interior(x).copy()) for x in (h, u, v)))
    for a, b in zip(got, ref):
        assert _rel(a.numpy(), G.interior(b)) < 5e-13
    y = tuple(torch.tensor(G.interior(x).copy()) for x in (h, u, v))
    G.rk4(h, u, v, 600.0, 3)
    for a, b in zip(T.rk4(y, 600.0, 3), (h, u, v)):
        assert _rel(a.numpy(), G.interior(b)) < 1e-11


def test_torch_restatement_gradient_is_a_transpose(oracle_lib):
    """Dot-product identity of the autograd reference: <J d, m> (forward mode) equals
    <d, J^T m> (reverse mode) to 1e-12."""
    import torch
    from oracle.torch_ref import TorchSWM
    N, order = 16, 42
    G = oracle_lib.Grid(N, order, order, A, OM, G0)
    T = TorchSWM(G, order)
    y = [torch.tensor(G.interior(x).copy(), requires_grad=True) for x in _state(G, N, 5)]
    rng = np.random.default_rng(0)
    d = [torch.tensor(rng.standard_normal(t.shape)) for t in y]
    m = [torch.tensor(rng.standard_normal(t.shape)) for t in y]
    _, jd = torch.autograd.functional.jvp(lambda *a: T.rhs(*a), tuple(t.detach() for t in y), tuple(d))
    _, jtm = torch.autograd.functional.vjp(lambda *a: T.rhs(*a), tuple(t.detach() for t in y), tuple(m))
    lhs = sum(float((a * b).sum()) for a, b in zip(jd, m))
    rhs = sum(float((a * b).sum()) for a, b in zip(d, jtm))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))
