"""Differentiable rollout of the timestep behind torch.autograd.Function.

forward(u0, f) runs `nsteps` full RK3 timesteps (mif_timestep) from the
initial velocity u0 (three FP64 device tensors in the reference layout,
shape (sz, sy, sx) per component) with a constant body force f (3,) and
returns u(T).  It keeps a checkpoint of (u, p) every `checkpoint_every`
steps, and records the three stencil sources of every step of the last
interval (u^n and the two stage buffers) as it runs them.  backward runs the
hand-written adjoint of each step (mif_timestep_adjoint) in reverse: the last
interval from the forward's records, every earlier one after replaying it
once from its checkpoint.  Memory: one (u, p) per checkpoint plus 3 velocity
triples per step of one interval.

No counterpart in the reference (it has no adjoint); the forward map is the
reference's timestep() (src/Timestep.cpp:97-156) with a body force added to
the momentum right-hand side.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib as L
from .core import Context, VelocityField, adjoint_force_take, function, timestep, timestep_adjoint


class Rollout:
    """nsteps of timestep() on one context, differentiable w.r.t. the initial
    velocity and a constant body force."""

    def __init__(self, ctx: Context, nsteps: int, checkpoint_every: int = 8, exact_velocity=None):
        if ctx.precision != L.FP64 and ctx.precision != L.MIXED:
            raise ValueError("unknown precision")
        self.ctx = ctx
        # the adjoint steps read the projected stage states the forward
        # leaves in its buffers
        ctx.set_stage_records(True)
        self.nsteps = int(nsteps)
        self.k = max(1, int(checkpoint_every))
        self.ev = exact_velocity if exact_velocity is not None else function(L.FN_ZERO)
        self.vel, self.buf, self.buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
        self.p, self.pb = ctx.pressure(), ctx.pressure()
        ncp = (self.nsteps + self.k - 1) // self.k
        self.cp_u = [ctx.velocity() for _ in range(ncp)]
        self.cp_p = [ctx.pressure() for _ in range(ncp)]
        self.rec = [(ctx.velocity(), ctx.velocity(), ctx.velocity()) for _ in range(min(self.k, self.nsteps))]
        self.ubar, self.pbar = ctx.velocity(), ctx.pressure()
        self.shapes = [f.shape for f in self.vel.components]
        # the records hold the last interval of the latest forward (until a
        # backward replays an earlier interval into them)
        self.tail_recorded = False
        # seconds spent in the last backward: replay of the forward steps and
        # the adjoint steps proper (measured only when `timed` is set, which
        # adds a device synchronisation per checkpoint interval)
        self.timed = False
        self.last_backward = {"recompute_s": 0.0, "adjoint_s": 0.0, "recompute_steps": 0, "adjoint_steps": 0}
        # forwards run so far: a graph node may only replay the latest one,
        # whose checkpoints and records the Rollout holds
        self.generation = 0
        self.force = (0.0, 0.0, 0.0)

    # -- device transfers with torch tensors -----------------------------------
    def _load(self, vf: VelocityField, tensors):
        # contiguous copies first (torch queues them on its current stream for
        # non-contiguous tensors, e.g. the expanded gradient of a sum), then
        # wait for torch's work, then the library reads the storage on its own
        # stream
        dense = []
        for f, t in zip(vf.components, tensors):
            t = t.detach().contiguous()
            if t.dtype != torch.float64 or not t.is_cuda or tuple(t.shape) != f.shape:
                raise ValueError(f"expected a float64 CUDA tensor of shape {f.shape}, got {t.dtype} "
                                 f"{'cuda' if t.is_cuda else 'cpu'} {tuple(t.shape)}")
            dense.append(t)
        torch.cuda.synchronize()
        for f, t in zip(vf.components, dense):
            f.upload_device(t.data_ptr())

    def _store(self, vf: VelocityField):
        out = []
        for f in vf.components:
            t = torch.empty(f.shape, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            f.download_device(t.data_ptr())
            out.append(t)
        return out

    def _zero(self, field):
        field.upload(np.zeros(field.shape))

    def _step(self, s: int):
        timestep(self.vel, self.buf, self.buf2, self.ev, s * self.ctx.dt, self.p, self.pb)

    # -- forward / backward -----------------------------------------------------
    def _record_step(self, s: int, s0: int):
        # timestep() leaves Y2 and Y3 (after their projections) in its two
        # buffers, so the recorded triples serve as the buffers; only u^n,
        # which the step overwrites, is copied
        un, y2, y3 = self.rec[s - s0]
        un.copy_from(self.vel)
        timestep(self.vel, y2, y3, self.ev, s * self.ctx.dt, self.p, self.pb)

    def run_forward(self, u0, f):
        self.generation += 1
        self.force = tuple(float(x) for x in f)
        self.ctx.set_body_force(self.force)
        self._load(self.vel, u0)
        self._zero(self.p)
        last = (len(self.cp_u) - 1) * self.k
        for s in range(self.nsteps):
            if s % self.k == 0:
                c = s // self.k
                self.cp_u[c].copy_from(self.vel)
                self.cp_p[c].copy_from(self.p)
            if s >= last:
                self._record_step(s, last)
            else:
                self._step(s)
        self.tail_recorded = True
        self.ctx.synchronize()
        return self._store(self.vel)

    def run_backward(self, grad_uT, generation=None):
        """Adjoint of the latest forward; `generation` (the forward's number)
        is checked against it when given."""
        if generation is not None and generation != self.generation:
            raise RuntimeError("Rollout: another forward ran on this Rollout after the one being differentiated; "
                               "its checkpoints are gone (use one Rollout per graph, or call backward first)")
        # the forward's body force (the context may serve other rollouts)
        self.ctx.set_body_force(self.force)
        self._load(self.ubar, grad_uT)
        self._zero(self.pbar)
        adjoint_force_take(self.ctx)  # start from a zero accumulator
        ncp = len(self.cp_u)
        t_rec = t_adj = 0.0
        n_rec = 0
        clock = time.perf_counter
        for c in reversed(range(ncp)):
            s0 = c * self.k
            s1 = min(self.nsteps, s0 + self.k)
            if self.timed:
                self.ctx.synchronize()
                t0 = clock()
            if c == ncp - 1 and self.tail_recorded:
                pass  # the forward recorded this interval
            else:
                self.vel.copy_from(self.cp_u[c])
                self.p.copy_from(self.cp_p[c])
                for s in range(s0, s1):
                    self._record_step(s, s0)
                n_rec += s1 - s0
            self.tail_recorded = False
            if self.timed:
                self.ctx.synchronize()
                t1 = clock()
                t_rec += t1 - t0
            for s in reversed(range(s0, s1)):
                un, y2, y3 = self.rec[s - s0]
                timestep_adjoint(un, y2, y3, self.ubar, self.pbar, None)
            if self.timed:
                self.ctx.synchronize()
                t_adj += clock() - t1
        fbar = adjoint_force_take(self.ctx)
        self.last_backward = {"recompute_s": t_rec, "adjoint_s": t_adj, "recompute_steps": n_rec,
                              "adjoint_steps": self.nsteps}
        return self._store(self.ubar), fbar

    def __call__(self, u0, f):
        return _RolloutFn.apply(self, f, *u0)


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx_, roll: Rollout, f, u, v, w):
        ctx_.roll = roll
        f = torch.as_tensor(f)
        ctx_.f_device, ctx_.f_dtype = f.device, f.dtype
        out = roll.run_forward((u, v, w), f.detach().cpu().numpy())
        ctx_.generation = roll.generation
        return tuple(out)

    @staticmethod
    def backward(ctx_, gu, gv, gw):
        roll = ctx_.roll
        grads = []
        for g, shp in zip((gu, gv, gw), roll.shapes):
            grads.append(torch.zeros(shp, dtype=torch.float64, device="cuda") if g is None else g)
        gu0, fbar = roll.run_backward(grads, generation=ctx_.generation)
        gf = torch.tensor(fbar, dtype=ctx_.f_dtype, device=ctx_.f_device)
        return (None, gf, *gu0)
