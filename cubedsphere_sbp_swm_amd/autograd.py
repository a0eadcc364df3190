"""Adjoint form: RK4 trajectories of the device path as a torch.autograd.Function.

The forward pass runs `nsteps` fused RK4 steps on the device and keeps a checkpoint of the
state every `checkpoint_every` steps (swm_adj_forward); the backward pass runs the
hand-written adjoint kernels in reverse, recomputing each segment of states from its
checkpoint and each step's stage inputs (swm_adj_backward).  Gradients are with respect to
the initial h, u, v at every stored Ah point (the duplicated panel-edge points are
independent inputs, as in the forward model).  PyTorch is only the autograd wrapper: no
torch operation is on the differentiated path.

A tape holds the checkpoints of its most recent forward only.  Each forward bumps the tape's
generation; a backward whose forward is no longer the tape's latest (two trajectories on one
tape in one loss, or a checkpointed region run again) first runs its forward again from the
saved initial state, so that every backward sees its own trajectory.
"""
import ctypes

import torch

from . import _capi
from ._capi import check


class AdjointTape:
    """Checkpoint storage and work space of one trajectory length / time step."""

    def __init__(self, op, nsteps, dt, checkpoint_every=20):
        from . import create_stvec_swm
        self.op, self.nsteps, self.dt, self.checkpoint_every = op, nsteps, dt, checkpoint_every
        t = ctypes.c_void_p()
        check(_capi.lib().swm_adj_create(op.ctx, int(nsteps), float(dt), int(checkpoint_every), ctypes.byref(t)))
        self.tape = t
        self.x0, self.xT = create_stvec_swm(op), create_stvec_swm(op)
        self.l0, self.lT = create_stvec_swm(op), create_stvec_swm(op)
        self.generation = 0  # number of forward runs; the checkpoints belong to the latest

    def close(self):
        if getattr(self, "tape", None):
            _capi.lib().swm_adj_destroy(self.tape)
            self.tape = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, x0, xT):
        check(_capi.lib().swm_adj_forward(self.tape, ctypes.c_void_p(x0.ptr), ctypes.c_void_p(xT.ptr)))
        self.generation += 1

    def backward(self, lT, l0):
        check(_capi.lib().swm_adj_backward(self.tape, ctypes.c_void_p(lT.ptr), ctypes.c_void_p(l0.ptr)))


class RK4Trajectory(torch.autograd.Function):
    """(h0, u0, v0) -> (hT, uT, vT) after tape.nsteps RK4 steps; tensors (6, n, n) float64 on
    the device."""

    @staticmethod
    def forward(ctx, h0, u0, v0, tape):
        tape.x0.from_tensors(h0.contiguous(), u0.contiguous(), v0.contiguous())
        tape.forward(tape.x0, tape.xT)
        outs = [torch.empty_like(h0, memory_format=torch.contiguous_format) for _ in range(3)]
        tape.xT.to_tensors(*outs)
        ctx.tape = tape
        ctx.generation = tape.generation
        ctx.save_for_backward(h0, u0, v0)
        return tuple(outs)

    @staticmethod
    def backward(ctx, gh, gu, gv):
        tape = ctx.tape
        h0, u0, v0 = ctx.saved_tensors
        if tape.generation != ctx.generation:
            # another forward ran on this tape since: recompute this call's trajectory
            tape.x0.from_tensors(h0.contiguous(), u0.contiguous(), v0.contiguous())
            tape.forward(tape.x0, tape.xT)
        gs = [torch.zeros_like(h0, memory_format=torch.contiguous_format) if g is None else g.contiguous()
              for g in (gh, gu, gv)]
        tape.lT.from_tensors(*gs)
        tape.backward(tape.lT, tape.l0)
        outs = [torch.empty_like(gs[0]) for _ in range(3)]
        tape.l0.to_tensors(*outs)
        return outs[0], outs[1], outs[2], None


def rk4_trajectory(tape, h0, u0, v0):
    return RK4Trajectory.apply(h0, u0, v0, tape)
