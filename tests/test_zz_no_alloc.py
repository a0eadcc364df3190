"""swm_diagnostics_noalloc (partial sums in a library-owned device scratch, device output, no
allocation) against swm_diagnostics on the same states, through the C ABI: bit for bit, on two
calls in a row with a smaller cube the second time (so stale scratch contents would show), and
on an FP32 context."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, G0, OM = 0.6371229e7, 9.80616, 2.0 * np.pi / (23 * 60 * 60 + 56 * 60 + 4)
KEYS = ("hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy")


def _pair(N, seed, precision=64):
    import torch
    import cubedsphere_sbp_swm_amd as swm
    from cubedsphere_sbp_swm_amd import _capi
    op = swm.create_swm_operator(N, **swm.operator_names(42), scale=A, omega=OM, grav=G0, precision=precision)
    x = swm.create_stvec_swm(op)
    x.upload(*op.initial_state(seed=seed, lmax=8))
    d = op.get_diagnostics(x)
    ref = np.array([d[k] for k in KEYS])
    out = torch.full((7,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    print(f"[noalloc] N={N} precision={precision}", flush=True)
    _capi.check(_capi.lib().swm_diagnostics_noalloc(op.ctx, ctypes.c_void_p(x.ptr), ctypes.c_void_p(out.data_ptr())))
    op.synchronize()
    return ref, out.cpu().numpy()


def test_noalloc_diagnostics_equal_diagnostics():
    first = _pair(48, 1)
    second = _pair(20, 2)
    third = _pair(16, 3, precision=32)
    for ref, got in (first, second, third):
        print("  ref", ref.tolist(), "\n  got", got.tolist())
        assert np.array_equal(ref, got)
    assert not np.array_equal(first[1], second[1])
