"""Adjoint form on the device: the hand-written transposed right-hand side and the
checkpointed reverse RK4 sweep.

References:
  * FP64 reference gradient: torch autograd through a tensor restatement of the oracle's
    forward map (oracle/torch_ref.py, itself checked against the C oracle here), on the CPU;
  * central finite differences of the device forward model along 32 random directions at the
    C192 sbp42 adjoint configuration (200 RK4 steps, checkpoints every 20 steps).
Tolerances: 1e-12 relative (max abs difference / max abs reference, per field) against the
autograd reference for one transposed RHS; 1e-11 for multi-step gradients (the two sides round
differently and the errors accumulate over the steps); 1e-6 relative per direction for the
finite-difference check, normalised by the root-mean-square directional derivative (at step
1e-4 the central-difference truncation error is O(1e-8) of it and rounding O(1e-10)).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, G0, OM = 0.6371229e7, 9.80616, 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)


def _op(N, order):
    import cubedsphere_sbp_swm_amd as swm
    return swm, swm.create_swm_operator(N, **swm.operator_names(order), scale=A, omega=OM, grav=G0)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("tile_x", [16, 32, 0])
@pytest.mark.parametrize("N,order", [(16, 42), (16, 43), (24, 63), (33, 42)])
def test_rhs_vjp_matches_autograd(oracle_lib, N, order, tile_x):
    import torch
    from oracle.torch_ref import TorchSWM
    swm, op = _op(N, order)
    print(f"[vjp] N={N} order={order} tile_x={tile_x}", flush=True)
    op.set_adjoint_tiling(tile_x)
    G = oracle_lib.Grid(N, order, order, A, OM, G0)
    T = TorchSWM(G, order)
    h, u, v = op.initial_state(seed=7, lmax=8)
    rng = np.random.default_rng(N + order)
    mu = [rng.standard_normal(h.shape) for _ in range(3)]
    x, m, out = (swm.create_stvec_swm(op) for _ in range(3))
    x.upload(h, u, v)
    m.upload(*mu)
    op.rhs_vjp(out, x, m)
    got = out.download()
    y = [torch.tensor(a, requires_grad=True) for a in (h, u, v)]
    F = T.rhs(*y)
    sum((f * torch.tensor(w)).sum() for f, w in zip(F, mu)).backward()
    errs = [_rel(g, t.grad.numpy()) for g, t in zip(got, y)]
    print(f"  rel err h,u,v = {errs}")
    assert max(errs) < 1e-12


def _loss_grad_device(swm, op, tape, y0, tgt, w):
    import torch
    from cubedsphere_sbp_swm_amd.autograd import rk4_trajectory
    dev = torch.device("cuda")
    t0 = [torch.tensor(a, device=dev, requires_grad=True) for a in y0]
    hT, uT, vT = rk4_trajectory(tape, *t0)
    wt, tg = torch.tensor(w, device=dev), torch.tensor(tgt, device=dev)
    loss = 0.5 * (wt * (hT - tg) ** 2).sum()
    loss.backward()
    return float(loss), [t.grad.cpu().numpy() for t in t0]


@pytest.mark.parametrize("nsteps,every", [(5, 2), (4, 10), (6, 3)])
def test_trajectory_gradient_matches_autograd(oracle_lib, nsteps, every):
    """Loss 0.5*sum(w (h_T - h_target)^2) with SBP quadrature weights w: device adjoint gradient
    (checkpoints every `every` steps, including a partial last segment) = torch autograd
    gradient of the restated forward map."""
    import torch
    from oracle.torch_ref import TorchSWM
    from cubedsphere_sbp_swm_amd.autograd import AdjointTape
    N, order, dt = 16, 42, 600.0
    swm, op = _op(N, order)
    print(f"[traj] steps={nsteps} every={every}", flush=True)
    G = oracle_lib.Grid(N, order, order, A, OM, G0)
    T = TorchSWM(G, order)
    y0 = op.initial_state(seed=1, lmax=8)
    tgt = op.initial_state(seed=2, lmax=8)[0]
    w = op.quadrature_weights()
    tape = AdjointTape(op, nsteps, dt, every)
    loss, grads = _loss_grad_device(swm, op, tape, y0, tgt, w)
    yr = [torch.tensor(a, requires_grad=True) for a in y0]
    hT = T.rk4(tuple(yr), dt, nsteps)[0]
    ref_loss = 0.5 * (torch.tensor(w) * (hT - torch.tensor(tgt)) ** 2).sum()
    ref_loss.backward()
    errs = [_rel(g, t.grad.numpy()) for g, t in zip(grads, yr)]
    print(f"  loss {loss} vs {float(ref_loss)}; grad rel err h,u,v = {errs}")
    assert abs(loss - float(ref_loss)) <= 1e-12 * abs(float(ref_loss))
    assert max(errs) < 1e-11


def test_gradient_finite_differences_c192():
    """C192 sbp42 FP64, 200 RK4 steps at 150 s (gravity-wave Courant 0.9), checkpoints every 20
    steps; loss against the 200-step state of a second seed.  <grad, d> equals the central
    difference (L(y0 + eps d) - L(y0 - eps d)) / (2 eps) of the device forward model on 32
    random directions d."""
    import cubedsphere_sbp_swm_amd as swm_mod
    from cubedsphere_sbp_swm_amd.autograd import AdjointTape
    N, nsteps, dt = 192, 200, 150.0
    swm, op = _op(N, 42)
    y0 = op.initial_state(seed=0, lmax=16)
    x = swm.create_stvec_swm(op)
    x.upload(*op.initial_state(seed=1, lmax=16))
    ts = swm_mod.create_timescheme(op, "rk4")
    ts.steps(x, dt, nsteps)
    tgt = x.download()[0]
    w = op.quadrature_weights()
    tape = AdjointTape(op, nsteps, dt, 20)
    loss, grads = _loss_grad_device(swm, op, tape, y0, tgt, w)

    def L(y):
        x.upload(*y)
        ts.steps(x, dt, nsteps)
        hT = x.download()[0]
        return 0.5 * float(np.sum(w * (hT - tgt) ** 2))

    assert L(y0) == pytest.approx(loss, rel=1e-13)
    rng = np.random.default_rng(42)
    eps = 1e-4
    ads, fds = [], []
    for k in range(32):
        d = [rng.standard_normal(y0[0].shape) * s for s in (1.0, 0.1, 0.1)]
        ad = sum(float(np.sum(g * dd)) for g, dd in zip(grads, d))
        fd = (L([a + eps * dd for a, dd in zip(y0, d)]) - L([a - eps * dd for a, dd in zip(y0, d)])) / (2 * eps)
        ads.append(ad)
        fds.append(fd)
        print(f"  direction {k}: adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
    ads, fds = np.array(ads), np.array(fds)
    scale = np.sqrt(np.mean(fds ** 2))
    worst = float(np.max(np.abs(ads - fds)) / scale)
    print(f"  max |adjoint - fd| / rms(fd) = {worst:.2e}")
    assert worst < 1e-6


@pytest.mark.parametrize("nsteps,every", [(7, 3), (4, 10)])
def test_repeated_backward_is_bitwise_stable(nsteps, every):
    """The forward records the last segment's stage inputs, so the first reverse sweep after it
    recomputes one segment fewer than the next ones: all sweeps give the same bits, also with
    another seed in between and after a second forward from another initial state."""
    from cubedsphere_sbp_swm_amd.autograd import AdjointTape
    swm, op = _op(16, 42)
    print(f"[repeat] steps={nsteps} every={every}", flush=True)
    tape = AdjointTape(op, nsteps, 600.0, every)
    ya, yb = op.initial_state(seed=3, lmax=8), op.initial_state(seed=4, lmax=8)
    rng = np.random.default_rng(5)
    lam_a = [rng.standard_normal(ya[0].shape) for _ in range(3)]
    lam_b = [rng.standard_normal(ya[0].shape) for _ in range(3)]
    lt, l0 = swm.create_stvec_swm(op), swm.create_stvec_swm(op)

    def grad(lam):
        lt.upload(*lam)
        tape.backward(lt, l0)
        return l0.download()

    def forward(y):
        tape.x0.upload(*y)
        tape.forward(tape.x0, tape.xT)

    forward(ya)
    g1 = grad(lam_a)
    g2 = grad(lam_a)
    grad(lam_b)
    g3 = grad(lam_a)
    forward(yb)
    gb1 = grad(lam_a)
    forward(ya)
    g4 = grad(lam_a)
    forward(yb)
    grad(lam_b)
    gb2 = grad(lam_a)
    for g in (g2, g3, g4):
        assert all(np.array_equal(a, b) for a, b in zip(g1, g))
    assert all(np.array_equal(a, b) for a, b in zip(gb1, gb2))
    assert not all(np.array_equal(a, b) for a, b in zip(g1, gb1))


@pytest.mark.parametrize("nsteps", [9, 1])
def test_checkpoint_interval_does_not_change_bits(nsteps):
    """Checkpoints every 1, 2 and 4 steps, and one segment (every stage input of the trajectory
    stored, no recomputation in the reverse sweep): the same final state and the same gradient
    bit for bit, since a recomputed segment repeats the forward's arithmetic exactly."""
    from cubedsphere_sbp_swm_amd.autograd import AdjointTape
    swm, op = _op(16, 42)
    print(f"[intervals] steps={nsteps}", flush=True)
    y0 = op.initial_state(seed=6, lmax=8)
    rng = np.random.default_rng(7)
    lam = [rng.standard_normal(y0[0].shape) for _ in range(3)]
    lt, l0 = swm.create_stvec_swm(op), swm.create_stvec_swm(op)
    lt.upload(*lam)
    outs = []
    for every in (1, 2, 4, nsteps):
        tape = AdjointTape(op, nsteps, 600.0, every)
        tape.x0.upload(*y0)
        tape.forward(tape.x0, tape.xT)
        tape.backward(lt, l0)
        outs.append((tape.xT.download(), l0.download()))
        tape.close()
    for xT, g in outs[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(outs[0][0], xT))
        assert all(np.array_equal(a, b) for a, b in zip(outs[0][1], g))


@pytest.mark.parametrize("N,order", [(16, 42), (40, 43), (47, 63), (63, 42), (97, 63)])
def test_adjoint_tilings_agree_bitwise(N, order):
    """The fused kernel on 16 x 16 and 32 x 16 tiles and the two-launch path evaluate every
    point with the same operations: identical bits, on panels whose size is not a multiple of
    the tile (and is for n = 48, 64)."""
    swm, op = _op(N, order)
    print(f"[tilings] N={N} order={order}", flush=True)
    h, u, v = op.initial_state(seed=11, lmax=8)
    rng = np.random.default_rng(N)
    x, m, out = (swm.create_stvec_swm(op) for _ in range(3))
    x.upload(h, u, v)
    m.upload(*[rng.standard_normal(h.shape) for _ in range(3)])
    got = {}
    for tx in (0, 16, 32):
        op.set_adjoint_tiling(tx)
        op.rhs_vjp(out, x, m)
        got[tx] = out.download()
    for tx in (16, 32):
        for a, b in zip(got[0], got[tx]):
            assert np.array_equal(a, b), tx
