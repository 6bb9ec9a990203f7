"""Reverse-mode gradients through the 2D periodic SR step, with checkpoints.

``SRAdjoint`` runs an ``SRPICEngine`` forward for ``nsteps`` steps and
propagates the adjoint of the loss L = sum over active cells of |B|^2 / 2
at the final step back to the initial state, with the hand-written VJP
kernels of the library (``ntt_sr_push_vjp``, ``ntt_deposit_vjp``,
``ntt_yee_vjp``).  Memory is bounded by checkpointing the state every
``checkpoint`` steps and recomputing the steps in between.

``FieldEnergyLoss`` wraps it as a ``torch.autograd.Function`` of the two
parameters of a counter-streaming (Weibel) setup: the drift 4-velocity u_d
(species s gets u_x += sign_s u_d) and a uniform background B_z amplitude b.

Reverse of one step (forward: B_h = B - cf curl E; push with E, B_h;
deposit; filter; B' = B_h - cf curl E; E' = E + ca curl B' + k J_f):
    lam_B  += N^T lam_E           lam_Jf = k lam_E         (Ampere, current)
    lam_E  += M^T lam_B                                      (second Faraday half)
    lam_J   = F^T...F^T lam_Jf                               (filter passes)
    deposit VJP, push VJP (adds to lam_E, lam_B)             (particles)
    lam_E  += M^T lam_B                                      (first Faraday half)
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from .containers import field_abi
from .engine import SRPICEngine, _stream
from .kernels import communicate_fields, mink

_PRT_STATE = ("i1", "i2", "i1_prev", "i2_prev", "dx1", "dx2", "dx1_prev", "dx2_prev",
              "ux1", "ux2", "ux3", "weight", "tag")


def _snap(eng: SRPICEngine) -> dict:
    return {"em": eng.fields.em.clone(), "step": eng.step, "time": eng.time,
            "prt": [{k: getattr(sp, k)[: sp.npart].clone() for k in _PRT_STATE}
                    for sp in eng.species]}


def _restore(eng: SRPICEngine, s: dict) -> None:
    eng.fields.em.copy_(s["em"])
    eng.step, eng.time = s["step"], s["time"]
    for sp, st in zip(eng.species, s["prt"]):
        for k, v in st.items():
            getattr(sp, k)[: v.shape[0]].copy_(v)


class _PrtView:
    """Particle arrays of a snapshot, in the layout of the C ABI."""

    def __init__(self, st: dict, maxnpart: int):
        self.st = st
        self.n = st["tag"].shape[0]
        self.maxnpart = maxnpart

    def abi(self) -> _abi.NttParticles:
        s = _abi.NttParticles()
        g = self.st
        s.i1, s.i2 = g["i1"].data_ptr(), g["i2"].data_ptr()
        s.i1_prev, s.i2_prev = g["i1_prev"].data_ptr(), g["i2_prev"].data_ptr()
        s.dx1, s.dx2 = g["dx1"].data_ptr(), g["dx2"].data_ptr()
        s.dx1_prev, s.dx2_prev = g["dx1_prev"].data_ptr(), g["dx2_prev"].data_ptr()
        s.ux1, s.ux2, s.ux3 = (g[k].data_ptr() for k in ("ux1", "ux2", "ux3"))
        s.weight, s.tag = g["weight"].data_ptr(), g["tag"].data_ptr()
        s.npart, s.maxnpart = self.n, self.maxnpart
        return s


class SRAdjoint:
    def __init__(self, engine: SRPICEngine, nsteps: int, checkpoint: int = 8, precision: int = 32):
        p = engine.params
        if p.dim != 2:
            raise ValueError("the adjoint covers 2D boxes")
        if p.shape_order not in (1, 2):
            raise ValueError("the adjoint covers shape orders 1 and 2")
        if any(s.pusher != _abi.PUSHER_BORIS for s in engine.species):
            raise ValueError("the adjoint covers the Boris pusher")
        if p.sort_interval:
            raise ValueError("re-sorting reorders particles: use sort_interval = 0")
        if p.fd:
            raise ValueError("the adjoint covers the Yee stencil")
        self.eng, self.nsteps, self.K, self.prec = engine, nsteps, max(1, checkpoint), precision
        self.rdt = torch.float64 if precision == 64 else torch.float32
        n1, n2 = p.n
        self.nn = (C.c_int32 * 2)(n1, n2)
        # the step's field coefficients, in single precision as the step forms them
        f = np.float32
        dt, dx = f(p.dt), f(p.dx)
        dT = f(0.5) * f(p.correction) * dt
        self.far = (float(dT / (dx * dx)), float(dT))        # fieldsolvers.h:44-57
        dTa = f(p.correction) * dt
        self.amp = (float(dTa / (dx * dx)), float(dTa))      # fieldsolvers.h:108-117
        self.kJ = float(-dt * f(p.q0) / (f(p.B0) * f(p.V0)))  # fieldsolvers.h:157
        self.lib = engine.lib

    # -- helpers -------------------------------------------------------------------
    def _active_B(self) -> torch.Tensor:
        G = self.eng.shape.ngh
        n1, n2 = self.eng.params.n
        return self.eng.fields.em[3:6, G:G + n2, G:G + n1]

    def loss(self) -> float:
        return 0.5 * float((self._active_B().double() ** 2).sum())

    def _em_half(self) -> torch.Tensor:
        """E and B after the step's first Faraday half (what the push gathers)."""
        e = self.eng
        em = e.fields.em.clone()
        communicate_fields(em, e.shape, 0, 6, strict=e.strict)
        k = mink.Faraday_kernel(em, e.shape, *self.far)
        k.strict = e.strict
        k()
        communicate_fields(em, e.shape, 3, 6, strict=e.strict)
        return em

    def _vjp_yee(self, op, c1, c2, lin, lout):
        _abi.check(self.lib.ntt_yee_vjp(op, self.nn, c1, c2, self.prec, lin.data_ptr(),
                                        lout.data_ptr(), _stream()), "ntt_yee_vjp")

    def _padj(self, d) -> _abi.NttParticleAdjoint:
        return _abi.NttParticleAdjoint(*(d[k].data_ptr() for k in ("lx1", "lx2", "lu1", "lu2",
                                                                   "lu3")))

    # -- forward / backward ----------------------------------------------------------
    def forward(self) -> float:
        for _ in range(self.nsteps):
            self.eng.step_forward()
        return self.loss()

    def gradient(self):
        """Runs forward and backward from the engine's current (initial) state.
        Returns (loss, lam_em, lam_u) with lam_em = dL/d(E, B) on active cells
        ([6][n2][n1]) and lam_u[s] = (dL/du1, dL/du2, dL/du3) per species at the
        initial state.  The engine ends at the final state."""
        e = self.eng
        N, K = self.nsteps, self.K
        ckpts = {}
        for n in range(N):
            if n % K == 0:
                ckpts[n] = _snap(e)
            e.step_forward()
        torch.cuda.synchronize()
        loss = self.loss()
        final = _snap(e)
        n1, n2 = e.params.n
        dev = e.device
        lam = torch.zeros((6, n2, n1), dtype=self.rdt, device=dev)
        lam[3:6] = self._active_B().to(self.rdt)
        lamJ = torch.zeros((3, n2, n1), dtype=self.rdt, device=dev)
        lamJ2 = torch.zeros_like(lamJ)
        post = [{k: torch.zeros(sp.npart, dtype=self.rdt, device=dev)
                 for k in ("lx1", "lx2", "lu1", "lu2", "lu3")} for sp in e.species]
        pre = [{k: torch.zeros_like(v) for k, v in d.items()} for d in post]
        metric = _abi.NttMinkowski(2, e.params.dx)
        for s0 in reversed(range(0, N, K)):
            _restore(e, ckpts[s0])
            recs = []
            for n in range(s0, min(s0 + K, N)):
                before = {"prt": [{k: getattr(sp, k)[: sp.npart].clone() for k in _PRT_STATE}
                                  for sp in e.species]}
                emh = self._em_half()
                e.step_forward()
                after = {"prt": [{k: getattr(sp, k)[: sp.npart].clone() for k in _PRT_STATE}
                                 for sp in e.species]}
                recs.append((before, emh, after))
            for before, emh, after in reversed(recs):
                # Ampere + current source, second Faraday half
                lamJ.copy_(lam[0:3]).mul_(self.kJ)
                self._vjp_yee(_abi.VJP_AMPERE, *self.amp, lam[0:3], lam[3:6])
                self._vjp_yee(_abi.VJP_FARADAY, *self.far, lam[3:6], lam[0:3])
                # filter passes (symmetric stencil)
                src, dst = lamJ, lamJ2
                for _ in range(e.params.n_filters):
                    self._vjp_yee(_abi.VJP_FILTER, 0.0, 0.0, src, dst)
                    src, dst = dst, src
                # particles: deposit, then push (both per species)
                for s, sp in enumerate(e.species):
                    pv_post = _PrtView(after["prt"][s], sp.maxnpart).abi()
                    pa_post, pa_pre = self._padj(post[s]), self._padj(pre[s])
                    _abi.check(self.lib.ntt_deposit_vjp(
                        C.byref(pv_post), self.nn, sp.charge, e.params.dt, e.params.shape_order,
                        self.prec, src.data_ptr(), C.byref(pa_post), pre[s]["lx1"].data_ptr(),
                        pre[s]["lx2"].data_ptr(), _stream()), "ntt_deposit_vjp")
                    pv_pre = _PrtView(before["prt"][s], sp.maxnpart).abi()
                    ctx = _abi.NttSrPusherCtx(sp.mass, sp.charge, e.params.dt, e.params.omegaB0,
                                              n1, n2, 1, sp.pusher)
                    fe = field_abi(emh, e.shape)
                    _abi.check(self.lib.ntt_sr_push_vjp(
                        C.byref(ctx), C.byref(pv_pre), C.byref(fe), C.byref(metric),
                        e.params.shape_order, self.prec, C.byref(pa_post), C.byref(pa_pre),
                        lam.data_ptr(), _stream()), "ntt_sr_push_vjp")
                # first Faraday half
                self._vjp_yee(_abi.VJP_FARADAY, *self.far, lam[3:6], lam[0:3])
                post, pre = pre, post
        _restore(e, final)
        torch.cuda.synchronize()
        lam_u = [(d["lu1"], d["lu2"], d["lu3"]) for d in post]
        return loss, lam, lam_u


class FieldEnergyLoss(torch.autograd.Function):
    """L(u_d, b): final magnetic energy of a counter-streaming run.

    ``setup(u_d, b)`` returns a fresh SRPICEngine at its initial state in
    which species s drifts at sign[s] * u_d along x and B_z = b * profile on
    the active cells; ``profile`` is that [n2][n1] tensor (None: uniform).
    Call as FieldEnergyLoss.apply(u_d, b, setup, sign, profile, nsteps,
    checkpoint, precision)."""

    @staticmethod
    def forward(ctx, ud, b, setup, sign, profile, nsteps, checkpoint, precision):
        eng = setup(float(ud), float(b))
        adj = SRAdjoint(eng, nsteps, checkpoint, precision)
        loss, lam, lam_u = adj.gradient()
        g_ud = sum(sg * float(lu[0].double().sum()) for sg, lu in zip(sign, lam_u))
        lb3 = lam[5].double()
        g_b = float(lb3.sum() if profile is None else (lb3 * profile.to(lb3)).sum())
        ctx.save_for_backward(torch.tensor([g_ud, g_b], dtype=torch.float64))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        (g,) = ctx.saved_tensors
        return grad * g[0], grad * g[1], None, None, None, None, None, None
