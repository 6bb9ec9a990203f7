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
