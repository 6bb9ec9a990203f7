"""Autotuner for the launch choices of the SR step: the tile shape of the
spatial sort (which is also the chunk of the fused push + deposit kernel), the
re-sort cadence, and the particles per thread of the unsorted layout.

None of these choices changes results: the current is summed in integers, so
fields are bit-identical and particles identical as a set whatever the
chunking.  The tuner therefore only measures: every candidate runs a fresh
box for a few steps and the fastest per step wins.  Results can be kept in a
JSON file keyed by the box's signature (dimension, cells, particles per cell,
shape order, pushers, filters), so a parameter scan tunes each kind of box
once.

    best = tune(engine_runner(factory), candidates_for(dim))
    factory(Candidate) -> a fresh SRPICEngine at its initial state
"""
from __future__ import annotations

import json
import time
from dataclasses import asdict, dataclass
from pathlib import Path


@dataclass(frozen=True)
class Candidate:
    tile: tuple
    sort_interval: int           # 0: never re-sort (chunks of 256 * ppt particles)
    particles_per_thread: int = 4


@dataclass
class TuneResult:
    best: Candidate
    ms_per_step: dict            # Candidate -> measured ms per step

    def table(self) -> list:
        return sorted(({"candidate": asdict(c), "ms_per_step": t}
                       for c, t in self.ms_per_step.items()), key=lambda r: r["ms_per_step"])


def fused_window(dim: int) -> tuple:
    """Cells per dimension of the fused kernel's LDS window (ntt_fused_window)."""
    import ctypes as C

    from . import _abi
    w = (C.c_int32 * 3)()
    _abi.check(_abi.load(False).ntt_fused_window(dim, w), "ntt_fused_window")
    return tuple(int(x) for x in w[:dim])


def tile_fits_window(tile, window) -> bool:
    """A freshly sorted tile of t cells spans t - 1 cells; the chunk keeps the
    LDS path while that span + 4 (the stencil's [-1, +2] reach and one cell of
    drift) fits the window, i.e. while t + 3 <= w in every dimension, with
    at least one more cell of room for the drift of a few steps."""
    return all(t + 4 <= w for t, w in zip(tile, window))


def candidates_for(dim: int, window=None) -> list:
    """The default search space: sorted layouts over a few tile shapes and
    cadences, and the unsorted layout over a few chunk sizes.  Tiles that do
    not fit the fused kernel's LDS window (`window`, cells per dimension; by
    default the built library's) are left out: every chunk of theirs would
    take the HBM path."""
    if dim == 2:
        tiles = [(8, 8), (16, 8), (8, 16), (16, 16)]
        cadences = [2, 4, 8]
        unsorted = (8, 8)
    elif dim == 3:
        tiles = [(4, 4, 4), (8, 4, 4), (4, 4, 2)]
        cadences = [2, 4]
        unsorted = (4, 4, 4)
    else:
        raise ValueError("dim must be 2 or 3")
    if window is None:
        window = fused_window(dim)
    tiles = [t for t in tiles if tile_fits_window(t, window)]
    out = [Candidate(t, c) for t in tiles for c in cadences]
    out += [Candidate(unsorted, 0, ppt) for ppt in (1, 2, 4, 8)]
    return out


def engine_runner(factory, steps: int = 16, warmup: int = 16):
    """A runner for tune(): builds the candidate's box with factory(candidate)
    and returns its wall-clock milliseconds per step over `steps` steps (after
    `warmup` steps), with the device synchronised around the timed steps.
    The warm-up also ages the layout: a freshly generated (sorted) plasma
    favours never re-sorting over a short window, so time a window in which
    the particles have moved a few cells."""
    import torch

    def run(c: Candidate) -> float:
        eng = factory(c)
        for _ in range(warmup):
            eng.step_forward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.step_forward()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        del eng
        return ms

    return run


def signature(params, ppc: int, pushers) -> str:
    """Key of a box kind in the tuning cache."""
    return json.dumps({"dim": params.dim, "n": list(params.n), "ppc": ppc,
                       "order": params.shape_order, "filters": params.n_filters,
                       "pushers": sorted(int(p) for p in pushers)}, sort_keys=True)


def tune(run, candidates, cache: str | Path | None = None, key: str | None = None) -> TuneResult:
    """Measures every candidate with run(candidate) -> ms per step and returns
    the fastest.  With cache and key, a stored result for key is returned
    without measuring, and a new result is stored."""
    if not candidates:
        raise ValueError("no candidates")
    path = Path(cache) if cache is not None else None
    if path is not None and key is not None and path.exists():
        stored = json.loads(path.read_text()).get(key)
        if stored is not None:
            best = Candidate(tuple(stored["best"]["tile"]), stored["best"]["sort_interval"],
                             stored["best"]["particles_per_thread"])
            return TuneResult(best, {best: stored["ms_per_step"]})
    times = {c: float(run(c)) for c in candidates}
    best = min(times, key=times.get)
    if path is not None and key is not None:
        data = json.loads(path.read_text()) if path.exists() else {}
        data[key] = {"best": asdict(best), "ms_per_step": times[best]}
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    return TuneResult(best, times)
