"""SRPIC engine for one periodic domain on one GPU.

Mirrors the reference's engine layer for the hot path:
* ``SimulationParams``-style scalars derived as the reference derives them
  (src/framework/parameters/parameters.cpp:47-78, algorithms.cpp:18-35,
  grid.cpp:39-51): dt = CFL * dx / sqrt(D), B0 = omegaB0 = 1 / larmor0,
  V0 = dx^D, n0 = ppc0 / V0, q0 = V0 / (ppc0 * skindepth0^2).
* ``SRPICEngine.step_forward`` -- SRPICEngine::step_forward
  (src/engines/srpic/srpic.hpp:65-188) through ``ntt_sr_step``.
* ``SRPICEngine.sort`` -- Metadomain::SortParticles
  (src/framework/domain/metadomain_sort.cpp:15-34) with a tile key, which also
  compacts dead particles (RemoveDead).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import torch

from . import _abi
from .containers import Fields, GridShape, Particles


@dataclass
class SpeciesSpec:
    label: str
    mass: float
    charge: float
    pusher: int = _abi.PUSHER_BORIS


@dataclass
class SRParams:
    """Hot-path keys of the reference's input (input.example.toml:300-420)."""
    dim: int
    n: tuple
    extent: tuple = None            # ((x1min, x1max), ...); default [0, n_d]
    cfl: float = 0.5
    correction: float = 1.0
    larmor0: float = 1.0
    skindepth0: float = 1.0
    ppc0: float = 16.0
    shape_order: int = 1
    n_filters: int = 0
    fieldsolver_enable: bool = True
    deposit_enable: bool = True
    fd: dict = field(default_factory=dict)
    sort_interval: int = 0          # 0: never re-sort
    tile: tuple = None              # sort / generation tile
    particles_per_thread: int = 0   # chunk of the unsorted layout (0: the library's 4)

    def __post_init__(self):
        if self.extent is None:
            self.extent = tuple((0.0, float(x)) for x in self.n)
        if self.tile is None:
            # 128-cell tiles (2048 particles at 16 ppc, inside the fused kernel's
            # LDS windows of 32 x 16 and 14 x 10 x 10) halve the workgroups of
            # 64-cell tiles and each pays its latency chain once per more
            # particles; a box with fewer than 512 of them (two workgroups per
            # CU of the 256) keeps 64-cell tiles, or the launches leave CUs idle
            big = (16, 8) if self.dim == 2 else (8, 4, 4)
            small = (8, 8) if self.dim == 2 else (4, 4, 4)
            ncell = math.prod(self.n[: self.dim])
            fits = all(x % t == 0 for x, t in zip(self.n, big))
            self.tile = big if fits and ncell // math.prod(big) >= 512 else small

    # derived scalars (single precision, as the FP32 reference build holds them)
    @property
    def dx(self) -> float:
        return (self.extent[0][1] - self.extent[0][0]) / self.n[0]

    @property
    def dt(self) -> float:
        # dt = CFL * dxMin, dxMin = dx / sqrt(D), in single precision
        # (algorithms.cpp:21-22, minkowski.h:82-84)
        import numpy as np
        dx_min = np.float32(self.dx) / np.sqrt(np.float32(self.dim))
        return float(np.float32(self.cfl) * np.float32(dx_min))

    @property
    def omegaB0(self) -> float:
        return 1.0 / self.larmor0

    @property
    def B0(self) -> float:
        return 1.0 / self.larmor0

    @property
    def V0(self) -> float:
        return self.dx ** self.dim

    @property
    def q0(self) -> float:
        return self.V0 / (self.ppc0 * self.skindepth0 ** 2)


class SRPICEngine:
    def __init__(self, params: SRParams, species: list[SpeciesSpec], maxnpart: int | list,
                 *, device="cuda", with_payload=False, strict=False, double_buffer=False):
        """double_buffer: re-sort out of place (ntt_sort_particles_to) into a
        second set of particle arrays, then swap the sets -- about half the
        memory traffic of the in-place sort, for twice the particle memory.
        The species' array attributes then change at every re-sort."""
        self.params = params
        self.device = torch.device(device)
        self.shape = GridShape(params.dim, tuple(params.n))
        self.fields = Fields(self.shape, device=self.device)
        if not isinstance(maxnpart, (list, tuple)):
            maxnpart = [maxnpart] * len(species)
        self.species = [
            Particles(params.dim, m, device=self.device, label=s.label, mass=s.mass,
                      charge=s.charge, pusher=s.pusher, with_payload=with_payload)
            for s, m in zip(species, maxnpart)]
        self.step = 0
        self.time = 0.0
        self.double_buffer = double_buffer
        self._alt = None        # the second set of particle arrays (double_buffer)
        # bound on |weight| of every particle (sets the current's fixed-point scale)
        self.weight_max = 1.0
        self._sort_ws = None
        self.strict = strict
        self.lib = _abi.load(strict)
        # tile offsets of every species (device uint32, ntiles + 1), or None
        self.tile_offsets: list = [None] * len(self.species)
        self.ntiles = _ntiles(self.lib, params)
        # workspace of ntt_sr_step: the fixed-point current accumulator, and in 3D
        # the split launches' flags and one box-record set per species, kept from
        # step to step (zeroed: no valid record before the first push)
        ws = C.c_uint64(0)
        cur = self.fields.abi("cur")
        _abi.check(self.lib.ntt_sr_step_workspace_size_species(C.byref(cur), len(self.species),
                                                               C.byref(ws)),
                   "ntt_sr_step_workspace_size_species")
        self._step_ws = torch.zeros(int(ws.value), dtype=torch.uint8, device=self.device)

    # -- initial conditions ---------------------------------------------------
    def inject_uniform(self, ppc: int, *, temperature: float, drifts=None, seed=1,
                       position_seed=7, weight=1.0):
        """Fills every species with ppc particles per cell (co-located across
        species); drifts[s] is a 3-vector of drift 4-velocity for species s."""
        n = (C.c_int32 * 3)(*(list(self.params.n) + [1] * (3 - self.params.dim)))
        for s, sp in enumerate(self.species):
            init = _abi.NttPlasmaInit()
            init.seed = seed
            init.position_seed = position_seed
            init.species_id = s
            init.ppc = ppc
            init.temperature = temperature
            d = (drifts[s] if drifts is not None else (0.0, 0.0, 0.0))
            init.drift_u[0], init.drift_u[1], init.drift_u[2] = d
            init.weight = weight
            t = list(self.params.tile) + [1] * (3 - self.params.dim)
            init.tile[0], init.tile[1], init.tile[2] = t
            abi = sp.abi()
            _abi.check(self.lib.ntt_init_uniform_plasma(C.byref(abi), self.params.dim, n,
                                                         C.byref(init), _stream()),
                       "ntt_init_uniform_plasma")
            sp.npart = abi.npart
            self.weight_max = max(self.weight_max, abs(weight))
            # the generator lays particles out tile by tile: ppc * cells_per_tile each
            per_tile = ppc * math.prod(self.params.tile[: self.params.dim])
            self.tile_offsets[s] = (torch.arange(self.ntiles + 1, device=self.device,
                                                 dtype=torch.int64) * per_tile
                                    ).to(torch.int32)

    # -- one step -----------------------------------------------------------------
    def step_params(self) -> _abi.NttSrStepParams:
        p = self.params
        s = _abi.NttSrStepParams()
        s.dt, s.correction, s.omegaB0 = p.dt, p.correction, p.omegaB0
        s.q0, s.B0, s.V0, s.ppc0 = p.q0, p.B0, p.V0, p.ppc0
        s.shape_order, s.n_filters = p.shape_order, p.n_filters
        s.fieldsolver_enable, s.deposit_enable = int(p.fieldsolver_enable), int(p.deposit_enable)
        s.step = self.step
        s.weight_max = self.weight_max
        s.particles_per_thread = p.particles_per_thread
        for k, v in p.fd.items():
            setattr(s.fd, k, v)
        for d in range(3):
            s.prtl_bc.bc[d][0] = s.prtl_bc.bc[d][1] = _abi.PRTL_BC_PERIODIC if d < p.dim else 0
        return s

    def species_abi(self):
        arr = (_abi.NttSpecies * len(self.species))()
        for k, sp in enumerate(self.species):
            arr[k].prtls = sp.abi()
            arr[k].mass, arr[k].charge, arr[k].pusher_flags = sp.mass, sp.charge, sp.pusher
            off = self.tile_offsets[k]
            arr[k].tile_offsets = off.data_ptr() if off is not None else None
            arr[k].ntiles = self.ntiles if off is not None else 0
        return arr

    def step_forward(self, stream=None) -> None:
        self.maybe_sort(stream)
        self.advance(stream)

    def maybe_sort(self, stream=None) -> bool:
        """The tile re-sort that opens a step every sort_interval steps; True if it ran."""
        p = self.params
        if p.sort_interval and self.step > 0 and self.step % p.sort_interval == 0:
            self.sort(stream=stream, keep_prev=False)   # the push that follows rewrites them
            return True
        return False

    def _step_args(self):
        """The ABI structures of ntt_sr_step, rebuilt only when a particle array
        or count (the containers' versions), a tile-offset array, the weight
        bound or a parameter has changed."""
        key = (tuple(sp.version for sp in self.species), self.fields.version,
               tuple(t.data_ptr() if t is not None else 0 for t in self.tile_offsets),
               self.weight_max, tuple(self.params.__dict__.values()))
        if getattr(self, "_args_key", None) != key:
            p = self.params
            self._args = (tuple(self.fields.abi(x) for x in ("em", "cur", "buff")),
                          self.species_abi(), self.step_params(),
                          _abi.NttMinkowski(p.dim, p.dx))
            self._args_key = key
        return self._args

    def advance(self, stream=None, em_half_out: torch.Tensor | None = None) -> None:
        """One ntt_sr_step (the step without its opening re-sort).  em_half_out (a
        tensor shaped like fields.em) receives E and B as the push reads them."""
        (em, cur, buff), sp, prm, metric = self._step_args()
        prm.step = self.step
        if em_half_out is not None and (em_half_out.shape != self.fields.em.shape or
                                        em_half_out.dtype != self.fields.em.dtype or
                                        not em_half_out.is_contiguous()):
            raise ValueError("em_half_out must be a contiguous tensor shaped like fields.em")
        prm.em_half_out = em_half_out.data_ptr() if em_half_out is not None else None
        _abi.check(self.lib.ntt_sr_step(C.byref(em), C.byref(cur), C.byref(buff), sp,
                                        len(self.species), C.byref(prm), C.byref(metric),
                                        self._step_ws.data_ptr(), self._step_ws.numel(),
                                        _stream(stream)), "ntt_sr_step")
        self.step += 1
        self.time += self.params.dt

    # -- spatial sort ----------------------------------------------------------------
    def sort(self, stream=None, compact: bool = False, keep_prev: bool = True) -> None:
        """Tile re-sort of every species (ntt_sort_particles[_to]).  keep_prev=False
        leaves the i*_prev / dx*_prev arrays of pushed species unpermuted, for a
        sort the next push follows (the push rewrites them first)."""
        p = self.params
        n = (C.c_int32 * 3)(*(list(p.n) + [1] * (3 - p.dim)))
        tile = (C.c_int32 * 3)(*(list(p.tile) + [1] * (3 - p.dim)))
        nt = _ntiles(self.lib, p)
        if nt != self.ntiles:   # the tile shape changed: offsets of the new size
            self.ntiles = nt
            self.tile_offsets = [None] * len(self.species)
        need = max(_sort_ws_bytes(self.lib, sp.maxnpart, p.dim, n, tile, self.double_buffer)
                   for sp in self.species)
        if self._sort_ws is None or self._sort_ws.numel() < need:
            self._sort_ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        if self.double_buffer and self._alt is None:
            self._alt = [Particles(p.dim, sp.maxnpart, device=self.device, label=sp.label,
                                   mass=sp.mass, charge=sp.charge, pusher=sp.pusher,
                                   with_phi=sp.phi is not None, with_payload=sp.pld_i is not None)
                         for sp in self.species]
        for k, sp in enumerate(self.species):
            skip = not keep_prev and sp.pusher != _abi.PUSHER_NONE
            abi = _without_prev(sp.abi()) if skip else sp.abi()
            alive = C.c_int64(0)
            if self.tile_offsets[k] is None:
                self.tile_offsets[k] = torch.zeros(self.ntiles + 1, dtype=torch.int32,
                                                   device=self.device)
            common = (p.dim, n, tile, self._sort_ws.data_ptr(), self._sort_ws.numel(),
                      self.tile_offsets[k].data_ptr(), C.byref(alive) if compact else None,
                      _stream(stream))
            if self.double_buffer:
                alt = self._alt[k]
                dst = _without_prev(alt.abi()) if skip else alt.abi()
                _abi.check(self.lib.ntt_sort_particles_to(C.byref(abi), C.byref(dst), *common),
                           "ntt_sort_particles_to")
                alt.npart = sp.npart      # dst now holds the sorted species
                sp.swap_arrays(alt)
            else:
                _abi.check(self.lib.ntt_sort_particles(C.byref(abi), *common),
                           "ntt_sort_particles")
            if compact:
                sp.npart = int(alive.value)


_PREV = ("i1_prev", "i2_prev", "i3_prev", "dx1_prev", "dx2_prev", "dx3_prev")


def _without_prev(abi: _abi.NttParticles) -> _abi.NttParticles:
    """A copy of the ABI view with the prev arrays left out (not permuted)."""
    out = _abi.NttParticles.from_buffer_copy(abi)
    for f in _PREV:
        setattr(out, f, None)
    return out


def _ntiles(lib, p: SRParams) -> int:
    n = (C.c_int32 * 3)(*(list(p.n) + [1] * (3 - p.dim)))
    tile = (C.c_int32 * 3)(*(list(p.tile) + [1] * (3 - p.dim)))
    v = C.c_int64(0)
    _abi.check(lib.ntt_sort_ntiles(p.dim, n, tile, C.byref(v)), "ntt_sort_ntiles")
    return int(v.value)


def _sort_ws_bytes(lib, maxnpart, dim, n, tile, out_of_place=False) -> int:
    b = C.c_uint64(0)
    fn = lib.ntt_sort_workspace_size_to if out_of_place else lib.ntt_sort_workspace_size
    _abi.check(fn(maxnpart, dim, n, tile, C.byref(b)), "ntt_sort_workspace_size")
    return int(b.value)


def _stream(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
