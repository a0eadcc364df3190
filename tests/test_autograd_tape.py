"""RK4Trajectory on a tape shared by several forward calls (CPU, with a stand-in tape).

The stand-in implements the tape interface of cubedsphere_sbp_swm_amd.autograd.AdjointTape
for the map x -> x*x per field: its forward keeps the last input as the "checkpoint" and its
backward uses that checkpoint, as the device tape's reverse sweep uses its stored states.  Two
trajectories on one tape in one loss must give each input its own gradient 2*x*g."""
import pytest
import torch

from cubedsphere_sbp_swm_amd.autograd import RK4Trajectory


class _Vec:
    def __init__(self):
        self.f = None

    def from_tensors(self, h, u, v):
        self.f = [h.clone(), u.clone(), v.clone()]

    def to_tensors(self, h, u, v):
        for dst, src in zip((h, u, v), self.f):
            dst.copy_(src)


class _SquareTape:
    def __init__(self):
        self.x0, self.xT, self.l0, self.lT = _Vec(), _Vec(), _Vec(), _Vec()
        self.generation = 0
        self.ckpt = None
        self.forwards = 0

    def forward(self, x0, xT):
        self.ckpt = [t.clone() for t in x0.f]
        xT.f = [t * t for t in x0.f]
        self.generation += 1
        self.forwards += 1

    def backward(self, lT, l0):
        l0.f = [2.0 * c * g for c, g in zip(self.ckpt, lT.f)]


def _inputs(seed):
    g = torch.Generator().manual_seed(seed)
    return [torch.randn(6, 5, 5, dtype=torch.float64, generator=g).requires_grad_() for _ in range(3)]


def test_two_trajectories_on_one_tape_get_their_own_gradients():
    tape = _SquareTape()
    a, b = _inputs(1), _inputs(2)
    ya = RK4Trajectory.apply(*a, tape)
    yb = RK4Trajectory.apply(*b, tape)
    wa, wb = _inputs(3), _inputs(4)
    loss = sum((y * w.detach()).sum() for y, w in zip(ya, wa)) + sum((y * w.detach()).sum() for y, w in zip(yb, wb))
    loss.backward()
    for x, w in zip(a, wa):
        assert torch.allclose(x.grad, 2.0 * x.detach() * w, rtol=0, atol=1e-14)
    for x, w in zip(b, wb):
        assert torch.allclose(x.grad, 2.0 * x.detach() * w, rtol=0, atol=1e-14)
    # yb's backward ran on the tape as left by its own forward; ya's forward ran again
    assert tape.forwards == 3


def test_latest_trajectory_needs_no_recompute():
    tape = _SquareTape()
    a = _inputs(5)
    y = RK4Trajectory.apply(*a, tape)
    (y[0].sum() + y[2].sum()).backward()  # u's output unused: its seed is zero
    assert tape.forwards == 1
    assert torch.allclose(a[0].grad, 2.0 * a[0].detach())
    assert torch.count_nonzero(a[1].grad) == 0
