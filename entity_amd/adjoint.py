"""Reverse-mode gradients through the 2D periodic SR step, with checkpoints.

``SRAdjoint`` runs an ``SRPICEngine`` forward for ``nsteps`` steps and
propagates the adjoint of the loss L = sum over active cells of |B|^2 / 2
at the final step back to the initial state, with the hand-written VJP
kernels of the library (``ntt_sr_particle_vjp``, the fused form of
``ntt_deposit_vjp`` and ``ntt_sr_push_vjp``, and ``ntt_yee_vjp``).  Memory is bounded by checkpointing the state every
``checkpoint`` steps and recomputing the steps in between.

Particle adjoints are indexed by particle ID when the engine carries an ID
payload (``SRPICEngine(..., with_payload=True)``): SRAdjoint numbers the
particles 0..npart-1 in ``pld_i`` at the start, and the tile re-sorts of the
forward then permute particles without disturbing the adjoint.  Without a
payload the adjoint is indexed by position and re-sorting must be off.

``FieldEnergyLoss`` wraps it as a ``torch.autograd.Function`` of the two
parameters of a counter-streaming (Weibel) setup: the drift 4-velocity u_d
(species s gets u_x += sign_s u_d) and the amplitude b of B_z = b * profile.

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

# arrays the push changes, in snapshot order (int32 or float32, npart each)
_MOVED = ("i1", "i2", "i1_prev", "i2_prev", "dx1", "dx2", "dx1_prev", "dx2_prev",
          "ux1", "ux2", "ux3")
# float arrays only a re-sort changes (permutes)
_SORTED = ("weight",)
_INT32 = ("i1", "i2", "i1_prev", "i2_prev")
_DX_PREV = ("dx1_prev", "dx2_prev")


class _Snap:
    """Particle arrays of one species at one instant, packed in one int32
    buffer (one launch to take), plus tag and IDs when re-sorts can move them.
    ``base``: an earlier snapshot with the sorted arrays (weight, tag, IDs) of
    the same particle order.  A push in the periodic box changes none of them,
    so a snapshot taken after pushes (and before the next re-sort) reads them
    there instead of copying them again."""

    def __init__(self, sp, with_sorted: bool, views: dict, base=None, pre=None):
        """pre: the snapshot of the same particles before the push that led
        here; the push copies dx into dx_prev bit for bit, so this snapshot
        reads its dx*_prev there instead of copying them."""
        n = sp.npart
        self.n = n
        self.pre = pre
        moved = _MOVED if pre is None else tuple(k for k in _MOVED if k not in _DX_PREV)
        self.names = moved + (_SORTED if with_sorted else ())
        self.has_pld = with_sorted and sp.pld_i is not None
        # the int16 tags ride in the int32 buffer as n/2 words when n is even
        self.tag_packed = with_sorted and n % 2 == 0
        # int32 views of the species' arrays, made once per (species, kind)
        key = (id(sp), sp.version, with_sorted, pre is None)
        parts = views.get(key)
        if parts is None:
            parts = self._views_of(sp)
            views[key] = parts
        self.buf = torch.cat(parts)
        self.tag = sp.tag[:n].clone() if with_sorted and not self.tag_packed else None
        # weight / tag / IDs when they are not in the snapshot
        self.shared = base if base is not None else sp
        self._abi = {}

    def _views_of(self, sp) -> list:
        """int32 views of the species' arrays in buffer order."""
        n = self.n
        parts = [getattr(sp, k)[:n].view(torch.int32) for k in self.names]
        if self.has_pld:
            parts.append(sp.pld_i[:n].view(torch.int32))
        if self.tag_packed:
            parts.append(sp.tag[:n].view(torch.int32))
        return parts

    def array(self, name: str):
        n = self.n
        if name in self.names:
            k = self.names.index(name)
            v = self.buf[k * n:(k + 1) * n]
            return v if name in _INT32 else v.view(torch.float32)
        if name in _DX_PREV and self.pre is not None:
            return self.pre.array(name[:-len("_prev")])
        k = len(self.names) * n
        if name == "pld_i" and self.has_pld:
            return self.buf[k:k + 2 * n].view(torch.int64)
        if name == "tag" and self.tag_packed:
            k += 2 * n if self.has_pld else 0
            return self.buf[k:k + n // 2].view(torch.int16)
        if name == "tag" and self.tag is not None:
            return self.tag
        if isinstance(self.shared, _Snap):
            return self.shared.array(name)
        a = getattr(self.shared, name)
        return a[:n] if a is not None else None

    def abi(self, maxnpart: int, ids: bool = True) -> _abi.NttParticles:
        """The C ABI view (built once; pointers into the packed buffer).
        ids=False leaves pld_i out, so that kernels index adjoints by position."""
        if ids not in self._abi:
            s = _abi.NttParticles()
            base, n = self.buf.data_ptr(), self.n
            for k, name in enumerate(self.names):
                setattr(s, name, base + 4 * k * n)
            for name in ("weight", "tag") + (_DX_PREV if self.pre is not None else ()):
                if name not in self.names:
                    setattr(s, name, self.array(name).data_ptr())
            pld = self.array("pld_i") if ids else None
            s.pld_i = pld.data_ptr() if pld is not None else None
            s.npart, s.maxnpart = n, maxnpart
            self._abi[ids] = s
        return self._abi[ids]

    def restore(self, sp) -> None:
        """Writes the snapshot back into the species (one multi-tensor copy for
        the packed buffer)."""
        dst, src, k = [], [], 0
        for d in self._views_of(sp):
            m = d.numel()
            s = self.buf[k:k + m]
            k += m
            if s.data_ptr() != d.data_ptr():
                dst.append(d)
                src.append(s)
        if dst:
            torch._foreach_copy_(dst, src)
        for name in ("weight", "tag", "pld_i"):
            if name in self.names or (name == "pld_i" and self.has_pld) or \
                    (name == "tag" and self.tag_packed):
                continue
            a = self.array(name)
            if a is None:
                continue
            d = getattr(sp, name)[: self.n]
            if a.data_ptr() != d.data_ptr():
                d.copy_(a)


class SRAdjoint:
    def __init__(self, engine: SRPICEngine, nsteps: int, checkpoint: int = 8, precision: int = 32):
        p = engine.params
        if p.dim != 2:
            raise ValueError("the adjoint covers 2D boxes")
        if p.shape_order not in (1, 2):
            raise ValueError("the adjoint covers shape orders 1 and 2")
        if any(s.pusher != _abi.PUSHER_BORIS for s in engine.species):
            raise ValueError("the adjoint covers the Boris pusher")
        self.ids = all(s.pld_i is not None for s in engine.species)
        if p.sort_interval and not self.ids:
            raise ValueError("re-sorting permutes particles: create the engine with "
                             "with_payload=True (particle IDs) or use sort_interval = 0")
        if p.fd:
            raise ValueError("the adjoint covers the Yee stencil")
        if precision not in (32, 64):
            raise ValueError("precision is 32 or 64")
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
        self._views = {}

    # -- helpers -------------------------------------------------------------------
    def _active_B(self) -> torch.Tensor:
        G = self.eng.shape.ngh
        n1, n2 = self.eng.params.n
        return self.eng.fields.em[3:6, G:G + n2, G:G + n1]

    def loss(self) -> float:
        return 0.5 * float((self._active_B().double() ** 2).sum())

    def _checkpoint(self) -> dict:
        e = self.eng
        return {"em": e.fields.em.clone(), "step": e.step, "time": e.time,
                "prt": [_Snap(sp, True, self._views) for sp in e.species],
                "tiles": [t.clone() if t is not None else None for t in e.tile_offsets]}

    def _restore(self, ck: dict) -> None:
        e = self.eng
        e.fields.em.copy_(ck["em"])
        e.step, e.time = ck["step"], ck["time"]
        for sp, snap in zip(e.species, ck["prt"]):
            snap.restore(sp)
        e.tile_offsets = [t.clone() if t is not None else None for t in ck["tiles"]]

    def _vjp_yee(self, op, c1, c2, lin, lout):
        _abi.check(self.lib.ntt_yee_vjp(op, self.nn, c1, c2, self.prec, lin.data_ptr(),
                                        lout.data_ptr(), _stream()), "ntt_yee_vjp")

    @staticmethod
    def _padj(d) -> _abi.NttParticleAdjoint:
        """d: [5][n] = (lx1, lx2, lu1, lu2, lu3)"""
        return _abi.NttParticleAdjoint(*(d[k].data_ptr() for k in range(5)), d.shape[1])

    def _prepare_reverse(self, post, pre):
        """ABI structs of the two particle-adjoint sets (built once)."""
        return ([self._padj(d) for d in post], [self._padj(d) for d in pre])

    def _record(self, n0: int, n1: int) -> list:
        """Runs steps n0..n1-1 from the engine's state, recording per step the
        particles before the push, E/B after the first Faraday half and the
        particles after the push."""
        e = self.eng
        recs = []
        before = [_Snap(sp, True, self._views) for sp in e.species]
        sorted_ = before   # the snapshots that hold weight, tag and IDs in this order
        for _ in range(n0, n1):
            if e.maybe_sort():
                before = sorted_ = [_Snap(sp, True, self._views) for sp in e.species]
            emh = torch.empty_like(e.fields.em)   # E, B as the push reads them
            e.advance(em_half_out=emh)
            after = [_Snap(sp, False, self._views, b, a)
                     for sp, b, a in zip(e.species, sorted_, before)]
            recs.append((before, emh, after))
            before = after
        return recs

    # -- forward / backward ----------------------------------------------------------
    def forward(self) -> float:
        for _ in range(self.nsteps):
            self.eng.step_forward()
        return self.loss()

    def gradient(self):
        """Runs forward and backward from the engine's current (initial) state.
        Returns (loss, lam_em, lam_u): lam_em = dL/d(E, B) on the active cells
        ([6][n2][n1]); lam_u[s] = (dL/du1, dL/du2, dL/du3) of species s at the
        initial state, indexed by particle ID (the initial position) when the
        engine carries IDs.  The engine ends at the final state; with IDs, its
        pld_i are overwritten with the initial positions 0..npart-1."""
        e = self.eng
        N, K = self.nsteps, self.K
        if self.ids:
            for sp in e.species:
                sp.pld_i[: sp.npart] = torch.arange(sp.npart, device=e.device)
        # checkpoints every K steps; the last segment is recorded on the way
        # (it is the first one the reverse pass needs)
        last = ((N - 1) // K) * K if N > 0 else 0
        ckpts = {}
        for n in range(last):
            if n % K == 0:
                ckpts[n] = self._checkpoint()
            e.step_forward()
        recs_last = self._record(last, N)
        loss = self.loss()
        final = self._checkpoint()
        n1, n2 = e.params.n
        dev = e.device
        lam = torch.zeros((6, n2, n1), dtype=self.rdt, device=dev)
        lam[3:6] = self._active_B().to(self.rdt)
        lamJ = torch.zeros((3, n2, n1), dtype=self.rdt, device=dev)
        lamJ2 = torch.zeros_like(lamJ)
        # particle adjoints [5][n] = (dx1, dx2, u1, u2, u3), held in the order of
        # the particles of the step at hand, so the kernels read them coalesced.
        # post_ids: the IDs of that order.  A re-sort, or a segment recomputed
        # from its checkpoint (whose re-sorts may order tiles differently),
        # changes the order: the adjoint then follows the particles by ID.
        post = [torch.zeros((5, sp.npart), dtype=self.rdt, device=dev) for sp in e.species]
        pre = [torch.zeros_like(d) for d in post]
        tmp = [torch.empty_like(d) for d in post]
        post_ids = [None] * len(e.species)

        def reorder(s, d, ids_from, ids_to):
            tmp[s].index_copy_(1, ids_from, d)                 # by ID
            torch.index_select(tmp[s], 1, ids_to, out=d)       # in the new order

        metric = _abi.NttMinkowski(2, e.params.dx)
        ctxs = [_abi.NttSrPusherCtx(sp.mass, sp.charge, e.params.dt, e.params.omegaB0, n1, n2, 1,
                                    sp.pusher) for sp in e.species]
        pa_post_all, pa_pre_all = self._prepare_reverse(post, pre)
        lam_ptr, lamE, lamB = lam.data_ptr(), lam[0:3], lam[3:6]
        order, prec, stream = e.params.shape_order, self.prec, _stream()
        for s0 in reversed(range(0, N, K)):
            if s0 == last:
                recs, recs_last = recs_last, None
            else:
                self._restore(ckpts.pop(s0))
                recs = self._record(s0, min(s0 + K, N))
            for before, emh, after in reversed(recs):
                # Ampere + current source, second Faraday half
                torch.mul(lamE, self.kJ, out=lamJ)
                self._vjp_yee(_abi.VJP_AMPERE, *self.amp, lamE, lamB)
                self._vjp_yee(_abi.VJP_FARADAY, *self.far, lamB, lamE)
                # filter passes (symmetric stencil)
                src, dst = lamJ, lamJ2
                for _ in range(e.params.n_filters):
                    self._vjp_yee(_abi.VJP_FILTER, 0.0, 0.0, src, dst)
                    src, dst = dst, src
                fe = field_abi(emh, e.shape)
                # particles: deposit and push VJPs of each species in one pass
                for s, sp in enumerate(e.species):
                    ids_q = after[s].array("pld_i")
                    if post_ids[s] is not None and post_ids[s].data_ptr() != ids_q.data_ptr():
                        reorder(s, post[s], post_ids[s], ids_q)
                    post_ids[s] = before[s].array("pld_i")   # the order pre is written in
                    _abi.check(self.lib.ntt_sr_particle_vjp(
                        C.byref(ctxs[s]), C.byref(before[s].abi(sp.maxnpart, False)),
                        C.byref(after[s].abi(sp.maxnpart, False)), C.byref(fe),
                        C.byref(metric), order, prec, src.data_ptr(),
                        C.byref(pa_post_all[s]), C.byref(pa_pre_all[s]), lam_ptr, stream),
                        "ntt_sr_particle_vjp")
                # first Faraday half
                self._vjp_yee(_abi.VJP_FARADAY, *self.far, lamB, lamE)
                post, pre = pre, post
                pa_post_all, pa_pre_all = pa_pre_all, pa_post_all
        self._restore(final)
        # to the ID order (pld_i = 0..npart-1 at the start)
        for s, d in enumerate(post):
            if post_ids[s] is not None:
                tmp[s].index_copy_(1, post_ids[s], d)
                post[s] = tmp[s]
        torch.cuda.synchronize()
        lam_u = [(d[2], d[3], d[4]) for d in post]
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
