"""Reverse-mode differentiation of the mixing path through torch.autograd.

The reference is differentiated by JAX's autodiff of its own code
(veros/core/utilities.py:51-59 and veros/core/isoneutral/diffusion.py:172-207
under ``jax.grad``); here each forward kernel has a hand-written adjoint
kernel in libvxmix.so behind a ``torch.autograd.Function``:

* ``solve_implicit_ad(a, b, c, d, kbot, b_edge=None, d_edge=None)`` — the
  kbot-masked tridiagonal solve, differentiable in a, b, c, d, b_edge and
  d_edge (vx_tdma_adjoint_f64: a transposed Thomas sweep that recomputes the
  elimination factors from a, b, c instead of storing them).
* ``isoneutral_step(state, T_tau, T_taup1, K_iso)`` — one isoneutral tracer
  step (iso=True, energy branch off) with the state's isoneutral tensor held
  fixed, differentiable in T[tau], T[taup1] and K_iso
  (vx_iso_diffusion_adjoint_f64: transposed K33 solve, then the transposed
  flux stencil scattered with float64 atomics).
* ``mixing_step(state, a, b, c, d, T_tau, K_iso, kbot)`` — the two composed:
  the solve's result becomes the tracer's taup1 level on the interior.

Forward passes run the same kernels as the non-differentiable API; the TDMA
forward may use the mixed float32 + float64-refinement solve
(``precision="mixed"``), its backward is always float64.
"""
from __future__ import annotations

import torch

from . import _native
from ._arrays import ptr, stream_handle, to_device
from .isoneutral import _check_array, grid_struct, params_struct
from .utilities import solve_implicit_kbot


def _contig(t):
    return None if t is None else t.contiguous()


class TdmaSolve(torch.autograd.Function):
    """x = solve_implicit(a, b, c, d) with gradients for a, b, c, d, b_edge, d_edge."""

    @staticmethod
    def forward(ctx, a, b, c, d, b_edge, d_edge, kbot, precision, refine_steps):
        x = solve_implicit_kbot(a, b, c, d, kbot, b_edge=b_edge, d_edge=d_edge, precision=precision,
                                refine_steps=refine_steps)
        ctx.save_for_backward(a, b, c, b_edge, d_edge, kbot, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        a, b, c, b_edge, d_edge, kbot, x = ctx.saved_tensors
        gx = gx.contiguous()
        nx, ny, nz = a.shape
        ga, gb, gc, gd = (torch.empty_like(a) for _ in range(4))
        gbe = torch.empty_like(a) if b_edge is not None else None
        gde = torch.empty_like(a) if d_edge is not None else None
        sx, sy, _ = a.stride()
        rc = _native.lib().vx_tdma_adjoint_f64(
            stream_handle(), ptr(a), ptr(b), ptr(c), ptr(b_edge), ptr(d_edge), ptr(kbot), kbot.stride(0),
            kbot.stride(1), ptr(x), ptr(gx), ptr(ga), ptr(gb), ptr(gc), ptr(gd), ptr(gbe), ptr(gde),
            nx, ny, nz, sx, sy)
        _native.check(rc, "vx_tdma_adjoint_f64")
        return ga, gb, gc, gd, gbe, gde, None, None, None


def solve_implicit_ad(a, b, c, d, kbot, b_edge=None, d_edge=None, *, precision="fp64", refine_steps=1):
    """Differentiable solve_implicit on kbot-masked columns (float64 tensors on the GPU)."""
    ts = [a, b, c, d, b_edge, d_edge]
    for t in ts:
        if t is not None and (not isinstance(t, torch.Tensor) or t.device.type != "cuda"):
            raise TypeError("solve_implicit_ad takes torch tensors on the GPU")
    a, b, c, d, b_edge, d_edge = (_contig(t) for t in ts)
    kb, _ = to_device(kbot, torch.int32)
    return TdmaSolve.apply(a, b, c, d, b_edge, d_edge, kb.contiguous(), precision, refine_steps)


class _IsoStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, T_tau, T_taup1, K_iso, state):
        vs, st = state.variables, state.settings
        X, Y, nz = K_iso.shape
        keep = []
        g = grid_struct(vs, keep)
        p = params_struct(st)
        tensors = [_check_array(n, getattr(vs, n), (X, Y, nz, 2, 2)) for n in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by")]
        tensors += [_check_array(n, getattr(vs, n), (X, Y, nz)) for n in ("K_11", "K_22", "K_33")]
        tau, taup1 = int(vs.tau), int(vs.taup1)
        if tau == taup1:
            raise ValueError("tau and taup1 must differ")
        tr = torch.zeros((X, Y, nz, 3), dtype=torch.float64, device=K_iso.device)
        tr[..., tau] = T_tau
        tr[..., taup1] = T_taup1
        dtr = torch.zeros((X, Y, nz), dtype=torch.float64, device=K_iso.device)
        L = _native.lib()
        ws = torch.empty(L.vx_iso_workspace_bytes(g, 1), dtype=torch.uint8, device=K_iso.device)
        rc = L.vx_iso_diffusion_f64(stream_handle(), g, p, ptr(tr), tau, taup1, ptr(K_iso),
                                    *(ptr(t) for t in tensors), ptr(dtr), None, None, ptr(ws))
        _native.check(rc, "isoneutral_step")
        ctx.state = state
        ctx.save_for_backward(tr, K_iso)
        return tr[..., taup1].contiguous()

    @staticmethod
    def backward(ctx, g_new):
        tr, K_iso = ctx.saved_tensors
        vs, st = ctx.state.variables, ctx.state.settings
        X, Y, nz = K_iso.shape
        keep = []
        g = grid_struct(vs, keep)
        p = params_struct(st)
        tensors = [getattr(vs, n) for n in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by", "K_11", "K_22", "K_33")]
        g_new = g_new.contiguous()
        g_tau, g_taup1, g_K = (torch.empty((X, Y, nz), dtype=torch.float64, device=K_iso.device) for _ in range(3))
        L = _native.lib()
        ws = torch.empty(L.vx_iso_adjoint_workspace_bytes(g), dtype=torch.uint8, device=K_iso.device)
        rc = L.vx_iso_diffusion_adjoint_f64(stream_handle(), g, p, ptr(tr), int(vs.tau), ptr(K_iso),
                                            *(ptr(t) for t in tensors), ptr(g_new), ptr(g_tau), ptr(g_taup1),
                                            ptr(g_K), ptr(ws))
        _native.check(rc, "isoneutral_step (adjoint)")
        return g_tau, g_taup1, g_K, None


def isoneutral_step(state, T_tau, T_taup1, K_iso):
    """T_new = tracer[..., taup1] after one isoneutral step of the tracer whose
    levels tau / taup1 are ``T_tau`` / ``T_taup1`` (each (X, Y, nz)), with
    ``K_iso`` in place of the state's and the state's Ai_*, K_11, K_22, K_33
    held fixed (veros/core/isoneutral/diffusion.py:172-207, iso=True, energy
    branch off). Differentiable in T_tau, T_taup1 and K_iso."""
    X, Y, nz = state.variables.K_iso.shape
    for name, t in (("T_tau", T_tau), ("T_taup1", T_taup1), ("K_iso", K_iso)):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float64:
            raise TypeError(f"{name} must be a float64 torch tensor on the GPU")
        if tuple(t.shape) != (X, Y, nz):
            raise ValueError(f"{name} must have shape {(X, Y, nz)}, got {tuple(t.shape)}")
    return _IsoStep.apply(T_tau, T_taup1, K_iso.contiguous(), state)


def mixing_step(state, a, b, c, d, T_tau, K_iso, b_edge=None, d_edge=None, *, precision="fp64",
                refine_steps=1):
    """solve_implicit of d on the interior columns (the state's kbot), whose
    result replaces the interior of the tracer's taup1 level (halo from
    ``state.variables.temp``), then one isoneutral step. Differentiable in
    a, b, c, d, b_edge, d_edge, T_tau and K_iso."""
    vs = state.variables
    kb = vs.kbot[2:-2, 2:-2]
    x = solve_implicit_ad(a, b, c, d, kb, b_edge, d_edge, precision=precision, refine_steps=refine_steps)
    base = vs.temp[..., vs.taup1]
    T_taup1 = torch.nn.functional.pad(x, (0, 0, 2, 2, 2, 2))
    halo = torch.ones_like(base, dtype=torch.bool)
    halo[2:-2, 2:-2, :] = False
    T_taup1 = torch.where(halo, base, T_taup1)
    return isoneutral_step(state, T_tau, T_taup1, K_iso)
