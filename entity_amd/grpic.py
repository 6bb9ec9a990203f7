"""GRPIC pusher-only runs on one GPU (test particles in fixed Kerr-Schild fields).

Mirrors the reference's GRPIC engine in the configuration with the field
solver and the current deposit turned off (pgens/wald/wald.toml: the Wald
test problem), where one step is:
* at step 0, em0 <- em (grpic::CopyFields, src/engines/grpic/grpic.hpp:293-301
  and src/engines/grpic/utils.h:40-43);
* the push of every species, D gathered from em and B from em0
  (grpic::ParticlePush, src/engines/grpic/particle_pusher.h:26-89);
* Metadomain::SortParticles (src/framework/domain/metadomain_sort.cpp:15-34):
  every clear_interval steps (step > 0) the dead particles are removed
  (Particles::RemoveDead, src/framework/containers/particles_sort.cpp:104-195),
  here through ntt_sort_particles with one tile covering the grid.

``GRRun`` is one run (a 2D axisymmetric r-theta box with its own metric).
``GRBatch`` advances many runs with one batched launch per step
(ntt_gr_batch_push over every (run, species) entry); its step is capturable
in a HIP graph, and every run's particles equal those of GRRun.step_forward
on that run alone bit for bit.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _abi
from .containers import GridShape, Particles, field_abi
from .engine import _stream
from .kernels import PusherBoundaries, gr


@dataclass
class GRSpeciesSpec:
    label: str
    mass: float
    charge: float
    pusher: int = _abi.PUSHER_BORIS


class GRRun:
    """One pusher-only GRPIC run: fields em (D in 0..2, B in 3..5, contravariant
    code components with ghosts) and em0, and its species."""

    def __init__(self, metric: gr.KerrSchild, species: list[GRSpeciesSpec], maxnpart: int, *,
                 dt: float, omegaB0: float = 1.0, niter: int = 10, precision: int = 64,
                 shape_order: int = 1, clear_interval: int = 100, bc=None, device="cuda"):
        self.metric = metric
        self.shape = GridShape(2, tuple(metric.n))
        self.device = torch.device(device)
        self.em = torch.zeros(self.shape.tensor_shape(6), device=self.device)
        self.em0 = torch.zeros_like(self.em)
        self.species = [Particles(2, maxnpart, device=self.device, label=s.label, mass=s.mass,
                                  charge=s.charge, pusher=s.pusher, with_phi=True)
                        for s in species]
        self.dt, self.omegaB0, self.niter, self.precision = dt, omegaB0, niter, precision
        self.shape_order = shape_order
        self.clear_interval = clear_interval
        # absorbing at both r boundaries, axis at both theta boundaries (the
        # Wald problem's particle boundaries)
        self.bc = bc or PusherBoundaries([(_abi.PRTL_BC_ABSORB, _abi.PRTL_BC_ABSORB),
                                          (_abi.PRTL_BC_AXIS, _abi.PRTL_BC_AXIS), (0, 0)])
        self.step = 0
        self.lib = _abi.load(False)
        self._sort_ws = None

    def pusher_context(self, sp: Particles) -> gr.PusherContext:
        return gr.PusherContext(sp.pusher, sp.mass, sp.charge, self.dt, self.omegaB0,
                                self.shape.n[0], self.shape.n[1], niter=self.niter,
                                precision=self.precision)

    def entry(self, sp: Particles) -> _abi.NttGrBatchEntry:
        """The species' arguments of ntt_gr_push as one batch entry."""
        e = _abi.NttGrBatchEntry()
        e.ctx = self.pusher_context(sp).abi()
        e.bc = self.bc.abi()
        e.prtls = sp.abi()
        e.em = field_abi(self.em, self.shape)
        e.em0 = field_abi(self.em0, self.shape)
        e.metric = self.metric.abi()
        return e

    def begin_step(self) -> None:
        """grpic.hpp:293-301: at step 0 the pusher-only engine copies em into em0."""
        if self.step == 0:
            self.em0.copy_(self.em)

    def push(self, stream=None) -> None:
        for sp in self.species:
            if sp.npart == 0:
                continue
            gr.Pusher_kernel(self.pusher_context(sp), self.bc, sp, self.em, self.em0,
                             self.shape, self.metric, self.shape_order)(stream)

    def clear_due(self) -> bool:
        return bool(self.clear_interval) and self.step > 0 and \
            self.step % self.clear_interval == 0

    def remove_dead(self, stream=None) -> list:
        """RemoveDead of every species: live particles first (ntt_sort_particles
        with one tile over the grid), npart set to their count.  Returns the
        new counts."""
        n = (C.c_int32 * 3)(self.shape.n[0], self.shape.n[1], 1)
        tile = (C.c_int32 * 3)(self.shape.n[0], self.shape.n[1], 1)
        need = 0
        for sp in self.species:
            b = C.c_uint64(0)
            _abi.check(self.lib.ntt_sort_workspace_size(sp.maxnpart, 2, n, tile, C.byref(b)),
                       "ntt_sort_workspace_size")
            need = max(need, int(b.value))
        if self._sort_ws is None or self._sort_ws.numel() < need:
            self._sort_ws = torch.empty(max(need, 1), dtype=torch.uint8, device=self.device)
        counts = []
        for sp in self.species:
            if sp.npart == 0:
                counts.append(0)
                continue
            alive = C.c_int64(0)
            abi = sp.abi()
            _abi.check(self.lib.ntt_sort_particles(C.byref(abi), 2, n, tile,
                                                   self._sort_ws.data_ptr(),
                                                   self._sort_ws.numel(), None, C.byref(alive),
                                                   _stream(stream)), "ntt_sort_particles")
            sp.npart = int(alive.value)
            counts.append(sp.npart)
        return counts

    def step_forward(self, stream=None) -> None:
        """One step of the pusher-only engine (grpic.hpp:293-301, 402-445, 615-617)."""
        self.begin_step()
        self.push(stream)
        if self.clear_due():
            self.remove_dead(stream)
        self.step += 1


class GRBatch:
    """Many independent GRRun runs pushed by one launch per step."""

    def __init__(self, runs: list[GRRun]):
        if not runs:
            raise ValueError("empty batch")
        if len({r.step for r in runs}) != 1:
            raise ValueError("all runs must be at the same step")
        orders = {r.shape_order for r in runs}
        if len(orders) != 1:
            raise ValueError("all runs must share the shape order")
        self.runs = list(runs)
        self.order = orders.pop()
        self.step = runs[0].step
        self.lib = runs[0].lib
        self._graph = None
        self._plan = C.c_void_p(None)
        self._build()

    def _build(self) -> None:
        self._slots = [(r, sp) for r in self.runs for sp in r.species]
        arr = (_abi.NttGrBatchEntry * len(self._slots))()
        for k, (r, sp) in enumerate(self._slots):
            arr[k] = r.entry(sp)
        if self._plan.value:
            self.lib.ntt_gr_batch_destroy(self._plan)
        plan = C.c_void_p(None)
        _abi.check(self.lib.ntt_gr_batch_create(arr, len(self._slots), self.order,
                                                C.byref(plan)), "ntt_gr_batch_create")
        self._plan = plan
        self._arrays = [tuple(getattr(sp, n).data_ptr() for n in ("i1", "tag", "ux1"))
                        for _, sp in self._slots]
        self._graph = None

    def __del__(self):
        try:
            if self._plan.value:
                self.lib.ntt_gr_batch_destroy(self._plan)
        except Exception:
            pass

    def _check_plan(self) -> None:
        now = [tuple(getattr(sp, n).data_ptr() for n in ("i1", "tag", "ux1"))
               for _, sp in self._slots]
        if now != self._arrays:
            raise RuntimeError("a run's particle arrays were replaced since the batch was "
                               "planned: rebuild the GRBatch")

    def push(self, stream=None) -> None:
        """One batched push of every (run, species) entry."""
        self._check_plan()
        _abi.check(self.lib.ntt_gr_batch_push(self._plan, _stream(stream)),
                   "ntt_gr_batch_push")

    def _after_push(self, stream=None) -> None:
        if any(r.clear_due() for r in self.runs):
            for r in self.runs:
                if r.clear_due():
                    r.remove_dead(stream)
            for k, (r, sp) in enumerate(self._slots):
                _abi.check(self.lib.ntt_gr_batch_set_npart(self._plan, k, sp.npart,
                                                           _stream(stream)),
                           "ntt_gr_batch_set_npart")
        for r in self.runs:
            r.step += 1
        self.step += 1

    def step_forward(self, stream=None) -> None:
        """One step of every run: em0 <- em at step 0, the batched push, and
        RemoveDead of the runs whose clear interval is due."""
        for r in self.runs:
            r.begin_step()
        self.push(stream)
        self._after_push(stream)

    def capture(self) -> None:
        """Records the batched push (step > 0) in a HIP graph; step_graph()
        replays it, with the dead-particle removal outside the graph."""
        if self.step == 0:
            raise ValueError("capture after the first step (step 0 copies em into em0)")
        self._check_plan()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                _abi.check(self.lib.ntt_gr_batch_push(self._plan, s.cuda_stream),
                           "ntt_gr_batch_push")
        torch.cuda.current_stream().wait_stream(s)
        self._graph = g

    def step_graph(self) -> None:
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._check_plan()
        self._graph.replay()
        self._after_push()


def match_range(metric: gr.KerrSchild, upper: bool, ds: float, ngh: int = 2):
    """x1 cells [lo, hi) (ghost-extended) of a match layer of width ds at one r
    boundary, and its edge radius: the box grpic::MatchFieldsIn builds
    (src/engines/grpic/fields_bcs.h:41-94), ghosts excluded at its lower end
    and included at its upper end, through Mesh::ExtentToRange
    (src/framework/domain/mesh.h:133-203)."""
    rmin, rmax = metric.r_extent
    n1 = metric.n[0]

    def x1_of(r):   # convert<1, Ph, Cd>
        if metric.kind == _abi.METRIC_QKERR_SCHILD:
            c0 = math.log(rmin - metric.r0)
            return (math.log(r - metric.r0) - c0) / ((math.log(rmax - metric.r0) - c0) / n1)
        return (r - rmin) / ((rmax - rmin) / n1)

    lo_r, hi_r, edge = (rmax - ds, rmax, rmax) if upper else (rmin, rmin + ds, rmin)
    lo_r = min(max(rmin, lo_r), rmax)
    hi_r = max(min(rmax, hi_r), rmin)
    lo = max(math.floor(x1_of(lo_r)), 0) + ngh
    hi = min(math.ceil(x1_of(hi_r)) + 2 * ngh, n1 + 2 * ngh)
    return lo, hi, edge


class GRPICEngine(GRRun):
    """The full GRPIC step of one 2D (r, theta) box: field solver, pusher,
    deposit, current filter and boundaries (GRPICEngine::step_forward,
    src/engines/grpic/grpic.hpp:66-634).

    Fields: em, em0 and aux (D or E in 0..2, B or H in 3..5), cur, cur0, buff.
    flds_bc / prtl_bc: ((x1 min, x1 max), (x2 min, x2 max)) of _abi.FLDS_BC_* /
    PRTL_BC_* codes (grid.boundaries.fields / particles; theta sides AXIS).
    ds: match layer widths per side of x1 (grid.boundaries.match.ds).
    target: the fields the match layer relaxes to (the problem generator's
    field at every staggered point), comps: the components it defines.
    Single domain: the halo exchanges of the reference (CommunicateFields,
    SynchronizeFields) have no neighbour to exchange with on an r-theta box
    with these boundaries, so they do nothing."""

    def __init__(self, metric: gr.KerrSchild, species: list, maxnpart: int, *, dt: float,
                 flds_bc, prtl_bc, ds=((1.0, 1.0),), target=None, comps: int = 0x3F,
                 nfilter: int = 0, q0: float = 1.0, B0: float = 1.0, fieldsolver: bool = True,
                 deposit: bool = True, correction: float = 1.0, **kw):
        pbc = PusherBoundaries([tuple(prtl_bc[0]), tuple(prtl_bc[1]), (0, 0)])
        super().__init__(metric, species, maxnpart, dt=dt, bc=pbc, **kw)
        for side in (0, 1):
            if flds_bc[0][side] not in (_abi.FLDS_BC_MATCH, _abi.FLDS_BC_HORIZON):
                raise ValueError("r boundaries of the fields: MATCH or HORIZON")
            if prtl_bc[0][side] not in (_abi.PRTL_BC_ABSORB, _abi.PRTL_BC_HORIZON):
                raise ValueError("r boundaries of the particles: ABSORB or HORIZON")
            if flds_bc[1][side] != _abi.FLDS_BC_AXIS:
                raise ValueError("theta boundaries of the fields: AXIS")
        if deposit and prtl_bc[0][0] == _abi.PRTL_BC_ABSORB:
            # grpic/fields_bcs.h:58-60
            raise ValueError("absorption of currents only possible in +x1 (+r)")
        self.flds_bc = (tuple(flds_bc[0]), tuple(flds_bc[1]), (0, 0))
        self.prtl_bc = prtl_bc
        self.ds = ds
        self.aux = torch.zeros_like(self.em)
        self.cur = torch.zeros(self.shape.tensor_shape(3), device=self.device)
        self.cur0 = torch.zeros_like(self.cur)
        self.buff = torch.zeros_like(self.cur)
        self.target = target if target is not None else torch.zeros_like(self.em)
        self.comps, self.nfilter = comps, nfilter
        self.fieldsolver, self.deposit = fieldsolver, deposit
        self.dT = correction * dt
        self.coeff_j = -dt * q0 / B0                    # grpic/fieldsolvers.h:237-240
        self._layers = {side: match_range(metric, side == 1, ds[0][side]) for side in (0, 1)}

    def _bcs(self):
        return [self.flds_bc[0], self.flds_bc[1]]

    # ---- grpic/fields_bcs.h:234-266 ----
    def boundaries(self, tags: int, kind: str = "main", stream=None) -> None:
        if kind == "aux":       # the horizon acts on main fields only
            return
        if kind == "curr":
            if self.prtl_bc[0][1] == _abi.PRTL_BC_ABSORB:
                lo, hi, edge = self._layers[1]
                gr.AbsorbCurrents_kernel(self.cur0, self.shape, self.metric, edge,
                                         self.ds[0][1], (lo, hi))(stream)
            return
        for d, side in ((0, 0), (1, 0), (1, 1), (0, 1)):   # Directions<2D>::orth
            bc = self.flds_bc[d][side]
            for F in (self.em, self.em0):
                if bc == _abi.FLDS_BC_MATCH:
                    lo, hi, edge = self._layers[side]
                    gr.MatchBoundaries_kernel(F, self.target, self.shape, self.metric, edge,
                                              self.ds[0][side], tags, (lo, hi), self.comps,
                                              self._bcs())(stream)
                elif bc == _abi.FLDS_BC_AXIS:
                    gr.AxisBoundaries_kernel(F, self.shape, side == 1, tags)(stream)
                elif bc == _abi.FLDS_BC_HORIZON:
                    gr.HorizonBoundaries_kernel(F, self.shape, tags, self.nfilter)(stream)

    def _aux_e(self, D, B, st):
        gr.ComputeAuxE_kernel(D, B, self.aux, self.shape, self.metric, self._bcs())(st)

    def _aux_h(self, D, B, st):
        gr.ComputeAuxH_kernel(D, B, self.aux, self.shape, self.metric, self._bcs())(st)

    def _faraday(self, Bin, Bout, coeff, st):
        gr.Faraday_kernel(Bin, Bout, self.aux, self.shape, self.metric, coeff, self._bcs())(st)

    def _ampere(self, Din, Dout, coeff, st):
        gr.Ampere_kernel(Din, Dout, self.aux, self.shape, self.metric, coeff, self._bcs())(st)

    def _currents(self, J, st):
        gr.CurrentsAmpere_kernel(self.em0, J, self.shape, self.metric, self.coeff_j,
                                 self._bcs())(st)

    def _swap(self):
        self.em, self.em0 = self.em0, self.em
        self.cur, self.cur0 = self.cur0, self.cur

    def _init_fields(self, st) -> None:
        """Step 0 with the field solver on (grpic.hpp:74-280)."""
        B, D, dT = _abi.BC_B, _abi.BC_D, self.dT
        self.boundaries(B | D, stream=st)
        self.em0.copy_(self.em)
        self._aux_e(self.em, self.em0, st)
        self._aux_h(self.em, self.em0, st)
        self._faraday(self.em0, self.em0, 0.5 * dT, st)
        self.boundaries(B, stream=st)
        self._ampere(self.em, self.em, 0.5 * dT, st)
        self.boundaries(D, stream=st)
        self._aux_e(self.em, self.em0, st)
        self._aux_h(self.em, self.em0, st)
        self._faraday(self.em, self.em0, dT, st)
        self.boundaries(B, stream=st)
        self._ampere(self.em0, self.em0, dT, st)
        self.boundaries(D, stream=st)
        self._aux_h(self.em0, self.em0, st)
        self._ampere(self.em, self.em0, dT, st)
        self.boundaries(D, stream=st)
        self._swap()

    def deposit_currents(self, stream=None) -> None:
        """grpic::CurrentsDeposit (currents.h:34-77) into cur0, then the
        absorbing layer and the filter passes (currents.h:80-110)."""
        self.cur0.zero_()
        for sp in self.species:
            if sp.npart == 0 or sp.charge == 0.0:
                continue
            gr.DepositCurrents_kernel(self.cur0, self.shape, sp, self.metric, sp.charge,
                                      self.dt, self.precision)(stream)
        self.boundaries(0x1C0, "curr", stream)
        for _ in range(self.nfilter):
            self.buff.copy_(self.cur0)
            gr.DigitalFilter_kernel(self.cur0, self.buff, self.shape, self._bcs())(stream)

    def step_forward(self, stream=None) -> None:
        st, dT = stream, self.dT
        B, D = _abi.BC_B, _abi.BC_D
        if self.step == 0:
            if self.fieldsolver:
                self._init_fields(st)
            else:
                self.em0.copy_(self.em)
        if self.fieldsolver:
            gr.TimeAverageDB_kernel(self.em, self.em0, self.shape)(st)
            self._aux_e(self.em0, self.em, st)
            self._faraday(self.em0, self.em0, dT, st)
            self.boundaries(B, stream=st)
            self._aux_h(self.em, self.em0, st)
        self.push(st)
        if self.deposit:
            self.deposit_currents(st)
        if self.fieldsolver:
            if self.deposit:
                gr.TimeAverageJ_kernel(self.cur, self.cur0, self.shape)(st)
            self._aux_e(self.em, self.em0, st)
            self._faraday(self.em, self.em0, dT, st)
            self.boundaries(B, stream=st)
            self._ampere(self.em0, self.em0, dT, st)
            if self.deposit:
                self._currents(self.cur, st)
            self.boundaries(D, stream=st)
            self._aux_h(self.em0, self.em0, st)
            self._ampere(self.em, self.em0, dT, st)
            if self.deposit:
                self._currents(self.cur0, st)
            self._swap()
            self.boundaries(D, stream=st)
        if self.clear_due():
            self.remove_dead(st)
        self.step += 1
