"""Batched ensembles of independent periodic SR boxes on one GPU.

``SREnsemble`` advances many ``SRPICEngine`` boxes (any sizes, shape orders,
pushers, filter counts and particle loads) together: each phase of the step
is one launch over all boxes (``ntt_ensemble_step``), and so is the tile
re-sort of every box (``ntt_ensemble_sort``, run when any box's interval is
due).  A step, with or without the re-sort, is captured once in a HIP graph
and replayed.  Every box's fields equal those of ``SRPICEngine.step_forward``
on that box alone bit for bit, and its particles as a set (the order inside
a tile may differ).  The reference runs such parameter scans as separate
simulations.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .engine import SRPICEngine, _stream


class SREnsemble:
    def __init__(self, engines: list[SRPICEngine]):
        if not engines:
            raise ValueError("empty ensemble")
        steps = {e.step for e in engines}
        if len(steps) != 1:
            raise ValueError("all boxes must be at the same step")
        self.engines = list(engines)
        self.lib = engines[0].lib
        self.step = engines[0].step
        self._graphs = {}
        self._plan = C.c_void_p(None)
        self._build()

    def _build(self) -> None:
        boxes = (_abi.NttBox * len(self.engines))()
        keep = []
        for b, eng in zip(boxes, self.engines):
            if eng.params.sort_interval and any(t is None for t in eng.tile_offsets):
                eng.sort()      # tile offsets the plan can point at
            sp = eng.species_abi()
            keep.append(sp)
            b.em, b.cur, b.buff = (eng.fields.abi(x) for x in ("em", "cur", "buff"))
            b.species = C.cast(sp, C.POINTER(_abi.NttSpecies))
            b.nspecies = len(eng.species)
            b.prm = eng.step_params()
            b.metric = _abi.NttMinkowski(eng.params.dim, eng.params.dx)
            b.workspace = eng._step_ws.data_ptr()
            b.workspace_bytes = eng._step_ws.numel()
            if eng.params.sort_interval:
                t = list(eng.params.tile) + [1] * (3 - eng.params.dim)
                b.tile[0], b.tile[1], b.tile[2] = t
        if self._plan.value:
            self.lib.ntt_ensemble_destroy(self._plan)
        # the plan holds the device pointers of every box: remember the
        # containers' versions to catch arrays replaced behind its back
        self._versions = self._current_versions()
        plan = C.c_void_p(None)
        _abi.check(self.lib.ntt_ensemble_create(boxes, len(self.engines), C.byref(plan)),
                   "ntt_ensemble_create")
        self._plan = plan
        self._graphs = {}

    def _current_versions(self):
        return [(e.fields.version, tuple(sp.version for sp in e.species)) for e in self.engines]

    def _check_plan(self) -> None:
        if self._current_versions() != self._versions:
            raise RuntimeError("a box's arrays or particle counts changed since the ensemble was "
                               "planned (e.g. a box stepped or re-sorted on its own): rebuild "
                               "the SREnsemble")

    def __del__(self):
        try:
            if self._plan.value:
                self.lib.ntt_ensemble_destroy(self._plan)
        except Exception:
            pass

    def _sort_due(self) -> bool:
        """True when some box's re-sort interval is due at this step (then every
        box that re-sorts at all is re-sorted, in one batched pass)."""
        return any(e.params.sort_interval and self.step > 0 and
                   self.step % e.params.sort_interval == 0 for e in self.engines)

    def _enqueue(self, sort: bool, stream) -> None:
        if sort:
            _abi.check(self.lib.ntt_ensemble_sort(self._plan, stream), "ntt_ensemble_sort")
        _abi.check(self.lib.ntt_ensemble_step(self._plan, self.step, stream), "ntt_ensemble_step")

    def _advance_counters(self) -> None:
        self.step += 1
        for eng in self.engines:
            eng.step += 1
            eng.time += eng.params.dt

    def step_forward(self, stream=None) -> None:
        """One step of every box (after the batched re-sort when one is due)."""
        self._check_plan()
        self._enqueue(self._sort_due(), _stream(stream))
        self._advance_counters()

    def capture(self) -> None:
        """Records the ensemble step (step > 0) in two HIP graphs, without and
        with the batched re-sort in front; step_graph() replays the one the
        step needs."""
        if self.step == 0:
            raise ValueError("capture after the first step (step 0 also fills ghosts)")
        for sort in (False, True):
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(s):
                with torch.cuda.graph(g, stream=s):
                    self._enqueue(sort, s.cuda_stream)
            torch.cuda.current_stream().wait_stream(s)
            self._graphs[sort] = g

    def step_graph(self) -> None:
        """One step by replaying a captured graph."""
        if not self._graphs:
            raise RuntimeError("capture() first")
        self._check_plan()
        self._graphs[self._sort_due()].replay()
        self._advance_counters()
