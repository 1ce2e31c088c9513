"""Differentiable electron run and its hand-written adjoint (libgasx
gasx_diff_*), exposed to torch.autograd.

The reference has no gradients; this wraps the frozen-event model of
include/gasx.h: the collisions of a recording pass at a reference density
profile n0 are kept pathwise (common random numbers), later passes at a
profile n replay the kicks and reweight each particle by the likelihood
ratio of its recorded decisions.  The objective is the electric energy of
the outer shell (i + 1/2) dr >= r_frac Rp at the last step, and the inputs
are the profile nodes and the source amplitude A (FN -> A FN).

    run = DiffRun(ctx, electrons, nodes0, n_steps=200, checkpoint_every=20)
    nodes = torch.tensor(nodes0, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    W = shell_energy(run, nodes, amp)
    W.backward()          # nodes.grad, amp.grad from gasx_diff_backward
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._abi import DiffSpec
from .core import Context, ParticleSet


class DiffRun:
    """A recorded differentiable run on a Context (one per context)."""

    def __init__(self, ctx: Context, electrons: ParticleSet, nodes0, n_steps: int,
                 checkpoint_every: int = 20, seed: int = 0, run_id: int = 0, r_frac: float = 0.9):
        self.ctx = ctx
        self.nodes0 = np.ascontiguousarray(np.asarray(nodes0, dtype=np.float64))
        if self.nodes0.ndim != 1:
            raise ValueError("nodes0 must be a vector")
        self.n_nodes = int(self.nodes0.size)
        self.n_steps = int(n_steps)
        self.n = electrons.n
        spec = DiffSpec(self.n_nodes, self.nodes0.ctypes.data_as(C.POINTER(C.c_double)), self.n_steps,
                        int(checkpoint_every), int(seed) & (2**64 - 1), int(run_id) & 0xFFFFFFFF,
                        float(r_frac))
        s = electrons.struct()
        ctx._check(ctx.lib.gasx_diff_record(ctx.ctx, C.byref(s), C.byref(spec)))
        self._forward_id = 0
        self._last = None

    def forward(self, nodes, amplitude: float = 1.0) -> float:
        """W at profile `nodes` and amplitude (replay of the recorded events)."""
        v = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64))
        if v.shape != (self.n_nodes,):
            raise ValueError(f"nodes must have shape ({self.n_nodes},)")
        out = C.c_double()
        self.ctx._check(self.ctx.lib.gasx_diff_forward(
            self.ctx.ctx, v.ctypes.data_as(C.POINTER(C.c_double)), float(amplitude), C.byref(out)))
        self._forward_id += 1
        self._last = (v.copy(), float(amplitude))
        return out.value

    def backward(self) -> tuple[np.ndarray, float]:
        """(dW/dnodes, dW/dA) of the last forward."""
        g = np.zeros(self.n_nodes, dtype=np.float64)
        ga = C.c_double()
        self.ctx._check(self.ctx.lib.gasx_diff_backward(
            self.ctx.ctx, g.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ga)))
        return g, ga.value

    def events(self, step: int) -> dict:
        """Recorded kick, p0 and collision flag of one step (host arrays)."""
        dev = self.ctx.device
        dv = torch.empty((3, self.n), dtype=torch.float32, device=f"cuda:{dev}")
        p0 = torch.empty(self.n, dtype=torch.float32, device=f"cuda:{dev}")
        col = torch.empty(self.n, dtype=torch.uint8, device=f"cuda:{dev}")
        self.ctx._check(self.ctx.lib.gasx_diff_events(
            self.ctx.ctx, int(step), dv[0].data_ptr(), dv[1].data_ptr(), dv[2].data_ptr(), p0.data_ptr(),
            col.data_ptr()))
        return {"dv": dv.cpu().numpy(), "p0": p0.cpu().numpy(), "col": col.cpu().numpy()}

    def release(self):
        if self.ctx.ctx:
            self.ctx._check(self.ctx.lib.gasx_diff_release(self.ctx.ctx))


class _ShellEnergy(torch.autograd.Function):
    @staticmethod
    def forward(fctx, run: DiffRun, nodes: torch.Tensor, amplitude: torch.Tensor):
        v = nodes.detach().to("cpu", torch.float64).numpy()
        a = float(amplitude.detach().cpu())
        w = run.forward(v, a)
        fctx.run = run
        fctx.point = (v.copy(), a)
        fctx.forward_id = run._forward_id
        return torch.tensor(w, dtype=torch.float64, device=nodes.device)

    @staticmethod
    def backward(fctx, gw: torch.Tensor):
        run = fctx.run
        if run._forward_id != fctx.forward_id:
            # another forward ran since: restore this one's trajectory first
            run.forward(*fctx.point)
        g, ga = run.backward()
        s = gw.detach().to("cpu", torch.float64)
        gn = (s * torch.from_numpy(g)).to(gw.device)
        gamp = (s * ga).to(gw.device)
        return None, gn, gamp


def shell_energy(run: DiffRun, nodes: torch.Tensor, amplitude: torch.Tensor) -> torch.Tensor:
    """W(nodes, amplitude) as a differentiable torch scalar (float64)."""
    return _ShellEnergy.apply(run, nodes, amplitude)
