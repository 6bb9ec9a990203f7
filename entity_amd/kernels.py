"""Python mirrors of Entity's kernel functors, backed by libntt_hip.

Each class is constructed with the arguments of the reference functor and
launched by calling it; a launch covers the range the reference's driver
passes to ``Kokkos::parallel_for`` (active particles, or active cells).
Launches are asynchronous on the current torch stream; failures raise
``NttError``.  There is no CPU fallback: without the built HIP library every
launch raises.

Reference functors (paths relative to the Entity source tree):
* sr.Pusher_kernel           src/kernels/pushers/sr.hpp:63-332
* DepositCurrents_kernel     src/kernels/currents_deposit.hpp:33-762
* mink.Faraday_kernel        src/kernels/faraday_mink.hpp:25-167
* mink.Ampere_kernel         src/kernels/ampere_mink.hpp:31-89
* mink.CurrentsAmpere_kernel src/kernels/ampere_mink.hpp:96-215
* DigitalFilter_kernel       src/kernels/digital_filter.hpp:58-389
* gr.Pusher_kernel           src/kernels/pushers/gr.hpp:71-865 (metrics
                             src/metrics/kerr_schild*.h, qkerr_schild.h)
* gr.Faraday_kernel          src/kernels/faraday_gr.hpp:32-106
* gr.Ampere_kernel, gr.CurrentsAmpere_kernel   src/kernels/ampere_gr.hpp:33-181
* gr.ComputeAuxE_kernel, gr.ComputeAuxH_kernel, gr.TimeAverageDB_kernel,
  gr.TimeAverageJ_kernel     src/kernels/aux_fields_gr.hpp:31-306
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import torch

from . import _abi
from .containers import Fields, GridShape, Particles, field_abi


def _stream(stream=None) -> int | None:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


@dataclass
class PusherContext:
    """kernel::sr::PusherContext (src/kernels/pushers/context.h:74-122)."""
    pusher_flags: int
    mass: float
    charge: float
    dt: float
    omegaB0: float
    ni1: int
    ni2: int
    ni3: int = 1

    def abi(self) -> _abi.NttSrPusherCtx:
        return _abi.NttSrPusherCtx(self.mass, self.charge, self.dt, self.omegaB0,
                                   self.ni1, self.ni2, self.ni3, self.pusher_flags)


@dataclass
class PusherBoundaries:
    """kernel::sr::PusherBoundaries<D> (context.h:124-178): PrtlBC per face."""
    bc: list = field(default_factory=lambda: [(_abi.PRTL_BC_PERIODIC,) * 2] * 3)

    def abi(self) -> _abi.NttPrtlBC:
        s = _abi.NttPrtlBC()
        for d in range(3):
            lo, hi = self.bc[d] if d < len(self.bc) else (0, 0)
            s.bc[d][0], s.bc[d][1] = lo, hi
        return s


@dataclass
class Minkowski:
    """metric::Minkowski<D> (src/metrics/minkowski.h:30-84): uniform cell size."""
    dim: int
    dx: float = 1.0

    def abi(self) -> _abi.NttMinkowski:
        return _abi.NttMinkowski(self.dim, self.dx)


class _Launch:
    strict = False

    @property
    def lib(self):
        return _abi.load(self.strict)


class sr:  # namespace kernel::sr
    class Pusher_kernel(_Launch):
        def __init__(self, ctx: PusherContext, bc: PusherBoundaries, particles: Particles,
                     EB: torch.Tensor, shape: GridShape, metric: Minkowski,
                     shape_order: int = 0, strict: bool = False):
            self.ctx, self.bc, self.particles = ctx, bc, particles
            self.EB, self.shape, self.metric = EB, shape, metric
            self.shape_order, self.strict = shape_order, strict

        def __call__(self, stream=None) -> None:
            p = self.particles.abi()
            f = field_abi(self.EB, self.shape)
            st = self.lib.ntt_sr_push(C.byref(self.ctx.abi()), C.byref(self.bc.abi()),
                                      C.byref(p), C.byref(f), C.byref(self.metric.abi()),
                                      self.shape_order, _stream(stream))
            _abi.check(st, "ntt_sr_push")


class gr:  # namespace kernel::gr
    @dataclass
    class PusherContext:
        """kernel::gr::PusherContext (context.h:182-211).  precision 64: FP64
        metric and pusher on FP32 particle storage; 32: FP32 arithmetic."""
        pusher_flags: int
        mass: float
        charge: float
        dt: float
        omegaB0: float
        ni1: int
        ni2: int
        niter: int = 10
        epsilon: float = 1e-6
        precision: int = 64

        def abi(self) -> _abi.NttGrPusherCtx:
            return _abi.NttGrPusherCtx(self.mass, self.charge, self.dt, self.omegaB0,
                                       self.epsilon, self.niter, self.ni1, self.ni2, 1,
                                       self.pusher_flags, self.precision)

    @dataclass
    class KerrSchild:
        """metric::KerrSchild / QKerrSchild / KerrSchild0 of an (r, theta) grid."""
        kind: int
        n: tuple
        r_extent: tuple
        theta_extent: tuple = (0.0, 3.141592653589793)
        a: float = 0.0
        r0: float = 0.0
        h: float = 0.0

        def abi(self) -> _abi.NttKsMetric:
            return _abi.NttKsMetric(self.kind, self.n[0], self.n[1], self.r_extent[0],
                                    self.r_extent[1], self.theta_extent[0], self.theta_extent[1],
                                    self.a, self.r0, self.h)

    class Pusher_kernel(_Launch):
        """kernel::gr::Pusher_kernel<M>(ctx, bc, particles, DB, DB0, metric):
        D is gathered from DB (em), B from DB0 (em0)."""

        def __init__(self, ctx, bc: PusherBoundaries, particles: Particles, DB, DB0,
                     shape: GridShape, metric, shape_order: int = 1):
            self.ctx, self.bc, self.particles = ctx, bc, particles
            self.DB, self.DB0, self.shape, self.metric = DB, DB0, shape, metric
            self.shape_order = shape_order

        def __call__(self, stream=None) -> None:
            p = self.particles.abi()
            fe = field_abi(self.DB, self.shape) if self.DB is not None else None
            f0 = field_abi(self.DB0, self.shape) if self.DB0 is not None else None
            st = self.lib.ntt_gr_push(C.byref(self.ctx.abi()), C.byref(self.bc.abi()),
                                      C.byref(p), C.byref(fe) if fe is not None else None,
                                      C.byref(f0) if f0 is not None else None,
                                      C.byref(self.metric.abi()), self.shape_order,
                                      _stream(stream))
            _abi.check(st, "ntt_gr_push")


    # ---- field solver (src/kernels/faraday_gr.hpp, ampere_gr.hpp,
    # aux_fields_gr.hpp): 2D fields on the metric's (r, theta) grid,
    # boundaries a list of FldsBC pairs per direction (AXIS at theta marks
    # the polar rows)
    class _Field(_Launch):
        @staticmethod
        def _bc(boundaries):
            if boundaries is None:
                return None
            b = _abi.NttFldsBC()
            for d, (lo, hi) in enumerate(boundaries):
                b.bc[d][0], b.bc[d][1] = lo, hi
            return b

        def _call(self, name, *fields, coeff=None, stream=None, with_bc=True):
            args = [C.byref(field_abi(f, self.shape)) for f in fields]
            args.append(C.byref(self.metric.abi()))
            if coeff is not None:
                args.append(C.c_float(coeff))
            if with_bc:
                bc = self._bc(self.boundaries)
                args.append(C.byref(bc) if bc is not None else None)
            args.append(_stream(stream))
            _abi.check(getattr(self.lib, name)(*args), name)

    class ComputeAuxE_kernel(_Field):
        """kernel::gr::ComputeAuxE_kernel(Df, Bf, Ef, metric)."""

        def __init__(self, Df, Bf, Ef, shape: GridShape, metric, boundaries=None):
            self.D, self.B, self.E, self.shape = Df, Bf, Ef, shape
            self.metric, self.boundaries = metric, boundaries

        def __call__(self, stream=None) -> None:
            self._call("ntt_gr_aux_e", self.D, self.B, self.E, stream=stream)

    class ComputeAuxH_kernel(_Field):
        """kernel::gr::ComputeAuxH_kernel(Df, Bf, Hf, metric)."""

        def __init__(self, Df, Bf, Hf, shape: GridShape, metric, boundaries=None):
            self.D, self.B, self.H, self.shape = Df, Bf, Hf, shape
            self.metric, self.boundaries = metric, boundaries

        def __call__(self, stream=None) -> None:
            self._call("ntt_gr_aux_h", self.D, self.B, self.H, stream=stream)

    class Faraday_kernel(_Field):
        """kernel::gr::Faraday_kernel(Bin, Bout, E, metric, coeff, ni2, boundaries)."""

        def __init__(self, Bin, Bout, E, shape: GridShape, metric, coeff: float,
                     boundaries=None):
            self.Bin, self.Bout, self.E, self.shape = Bin, Bout, E, shape
            self.metric, self.coeff, self.boundaries = metric, coeff, boundaries

        def __call__(self, stream=None) -> None:
            f = [field_abi(t, self.shape) for t in (self.Bin, self.Bout, self.E)]
            bc = self._bc(self.boundaries)
            _abi.check(self.lib.ntt_gr_faraday(C.byref(f[0]), C.byref(f[1]), C.byref(f[2]),
                                               C.byref(self.metric.abi()), self.coeff,
                                               C.byref(bc) if bc is not None else None,
                                               _stream(stream)), "ntt_gr_faraday")

    class Ampere_kernel(_Field):
        """kernel::gr::Ampere_kernel(Din, Dout, H, metric, coeff, ni2, boundaries)."""

        def __init__(self, Din, Dout, H, shape: GridShape, metric, coeff: float,
                     boundaries=None):
            self.Din, self.Dout, self.H, self.shape = Din, Dout, H, shape
            self.metric, self.coeff, self.boundaries = metric, coeff, boundaries

        def __call__(self, stream=None) -> None:
            f = [field_abi(t, self.shape) for t in (self.Din, self.Dout, self.H)]
            bc = self._bc(self.boundaries)
            _abi.check(self.lib.ntt_gr_ampere(C.byref(f[0]), C.byref(f[1]), C.byref(f[2]),
                                              C.byref(self.metric.abi()), self.coeff,
                                              C.byref(bc) if bc is not None else None,
                                              _stream(stream)), "ntt_gr_ampere")

    class CurrentsAmpere_kernel(_Field):
        """kernel::gr::CurrentsAmpere_kernel(Df, J, metric, coeff, ni2, boundaries)."""

        def __init__(self, Df, J, shape: GridShape, metric, coeff: float, boundaries=None):
            self.D, self.J, self.shape = Df, J, shape
            self.metric, self.coeff, self.boundaries = metric, coeff, boundaries

        def __call__(self, stream=None) -> None:
            f = [field_abi(t, self.shape) for t in (self.D, self.J)]
            bc = self._bc(self.boundaries)
            _abi.check(self.lib.ntt_gr_currents_ampere(C.byref(f[0]), C.byref(f[1]),
                                                       C.byref(self.metric.abi()), self.coeff,
                                                       C.byref(bc) if bc is not None else None,
                                                       _stream(stream)),
                       "ntt_gr_currents_ampere")

    class TimeAverageDB_kernel(_Launch):
        """kernel::gr::TimeAverageDB_kernel(BDf, BDf0): BDf0 = (BDf0 + BDf) / 2."""

        def __init__(self, BDf, BDf0, shape: GridShape):
            self.BD, self.BD0, self.shape = BDf, BDf0, shape

        def __call__(self, stream=None) -> None:
            a, b = field_abi(self.BD, self.shape), field_abi(self.BD0, self.shape)
            _abi.check(self.lib.ntt_gr_time_average_db(C.byref(a), C.byref(b), _stream(stream)),
                       "ntt_gr_time_average_db")

    class TimeAverageJ_kernel(_Launch):
        """kernel::gr::TimeAverageJ_kernel(Jf, Jf0): Jf = (Jf0 + Jf) / 2."""

        def __init__(self, Jf, Jf0, shape: GridShape):
            self.J, self.J0, self.shape = Jf, Jf0, shape

        def __call__(self, stream=None) -> None:
            a, b = field_abi(self.J, self.shape), field_abi(self.J0, self.shape)
            _abi.check(self.lib.ntt_gr_time_average_j(C.byref(a), C.byref(b), _stream(stream)),
                       "ntt_gr_time_average_j")


    # ---- boundaries, current filter and deposit of the GRPIC step
    # (src/kernels/fields_bcs.hpp, digital_filter.hpp, currents_deposit.hpp)
    class MatchBoundaries_kernel(_Field):
        """kernel::bc::MatchBoundaries_kernel<GRPIC, M, FS, in::x1>(Fld, fset,
        metric, xg_edge, ds, tags, boundaries) over x1 cells i1_range: fset is
        `target`, the field setter's values at every staggered point, comps the
        components it defines (bit c for em component c)."""

        def __init__(self, Fld, target, shape: GridShape, metric, r_edge: float, ds: float,
                     tags: int, i1_range, comps: int = 0x3F, boundaries=None):
            self.F, self.T, self.shape, self.metric = Fld, target, shape, metric
            self.m = _abi.NttGrMatch(r_edge, ds, int(i1_range[0]), int(i1_range[1]), tags, comps)
            self.boundaries = boundaries

        def __call__(self, stream=None) -> None:
            f, t = field_abi(self.F, self.shape), field_abi(self.T, self.shape)
            bc = self._bc(self.boundaries)
            _abi.check(self.lib.ntt_gr_match_boundaries(
                C.byref(f), C.byref(t), C.byref(self.metric.abi()),
                C.byref(bc) if bc is not None else None, C.byref(self.m), _stream(stream)),
                "ntt_gr_match_boundaries")

    class HorizonBoundaries_kernel(_Launch):
        """kernel::bc::gr::HorizonBoundaries_kernel(Fld, i1_min, tags, nfilter)."""

        def __init__(self, Fld, shape: GridShape, tags: int, nfilter: int):
            self.F, self.shape, self.tags, self.nfilter = Fld, shape, tags, nfilter

        def __call__(self, stream=None) -> None:
            f = field_abi(self.F, self.shape)
            _abi.check(self.lib.ntt_gr_horizon_boundaries(C.byref(f), self.tags, self.nfilter,
                                                          _stream(stream)),
                       "ntt_gr_horizon_boundaries")

    class AxisBoundaries_kernel(_Launch):
        """kernel::bc::AxisBoundaries_kernel<2D, P>(Fld, i_edge, tags); upper = P."""

        def __init__(self, Fld, shape: GridShape, upper: bool, tags: int):
            self.F, self.shape, self.upper, self.tags = Fld, shape, upper, tags

        def __call__(self, stream=None) -> None:
            f = field_abi(self.F, self.shape)
            _abi.check(self.lib.ntt_gr_axis_boundaries(C.byref(f), int(self.upper), self.tags,
                                                       _stream(stream)),
                       "ntt_gr_axis_boundaries")

    class AbsorbCurrents_kernel(_Launch):
        """kernel::bc::gr::AbsorbCurrents_kernel<M, 1>(J, metric, xg_edge, ds)
        over x1 cells i1_range."""

        def __init__(self, J, shape: GridShape, metric, r_edge: float, ds: float, i1_range):
            self.J, self.shape, self.metric = J, shape, metric
            self.r_edge, self.ds, self.range = r_edge, ds, (int(i1_range[0]), int(i1_range[1]))

        def __call__(self, stream=None) -> None:
            j = field_abi(self.J, self.shape)
            _abi.check(self.lib.ntt_gr_absorb_currents(C.byref(j), C.byref(self.metric.abi()),
                                                       self.r_edge, self.ds, self.range[0],
                                                       self.range[1], _stream(stream)),
                       "ntt_gr_absorb_currents")

    class DigitalFilter_kernel(_Field):
        """kernel::DigitalFilter_kernel<2D, spherical>(array, buffer, size, boundaries)."""

        def __init__(self, array, buffer, shape: GridShape, boundaries=None):
            self.A, self.B, self.shape, self.boundaries = array, buffer, shape, boundaries

        def __call__(self, stream=None) -> None:
            a, b = field_abi(self.A, self.shape), field_abi(self.B, self.shape)
            bc = self._bc(self.boundaries)
            _abi.check(self.lib.ntt_gr_digital_filter(C.byref(a), C.byref(b),
                                                      C.byref(bc) if bc is not None else None,
                                                      _stream(stream)), "ntt_gr_digital_filter")

    class DepositCurrents_kernel(_Launch):
        """kernel::DepositCurrents_kernel<GRPIC, M>(J, particles..., metric,
        charge, dt): order 1; precision 64 evaluates the metric in FP64."""

        def __init__(self, J, shape: GridShape, particles: Particles, metric, charge: float,
                     dt: float, precision: int = 64):
            self.J, self.shape, self.particles, self.metric = J, shape, particles, metric
            self.charge, self.dt, self.precision = charge, dt, precision

        def __call__(self, stream=None) -> None:
            p = self.particles.abi()
            j = field_abi(self.J, self.shape)
            _abi.check(self.lib.ntt_gr_deposit_currents(C.byref(p), C.byref(j),
                                                        C.byref(self.metric.abi()), self.charge,
                                                        self.dt, self.precision,
                                                        _stream(stream)),
                       "ntt_gr_deposit_currents")


class DepositCurrents_kernel(_Launch):
    def __init__(self, J: torch.Tensor, shape: GridShape, particles: Particles,
                 metric: Minkowski, charge: float, dt: float, shape_order: int = 0,
                 strict: bool = False):
        self.J, self.shape, self.particles, self.metric = J, shape, particles, metric
        self.charge, self.dt, self.shape_order, self.strict = charge, dt, shape_order, strict

    def __call__(self, stream=None) -> None:
        p = self.particles.abi()
        f = field_abi(self.J, self.shape)
        st = self.lib.ntt_deposit_currents(C.byref(p), C.byref(f), C.byref(self.metric.abi()),
                                           self.charge, self.dt, self.shape_order,
                                           _stream(stream))
        _abi.check(st, "ntt_deposit_currents")


class mink:  # namespace kernel::mink
    class Faraday_kernel(_Launch):
        def __init__(self, EB: torch.Tensor, shape: GridShape, coeff1: float, coeff2: float,
                     deltax=0.0, deltay=0.0, betaxy=0.0, betayx=0.0, deltaz=0.0, betaxz=0.0,
                     betazx=0.0, betayz=0.0, betazy=0.0, strict: bool = False):
            self.EB, self.shape, self.c1, self.c2, self.strict = EB, shape, coeff1, coeff2, strict
            self.fd = _abi.NttFdCoeffs(deltax, deltay, deltaz, betaxy, betayx, betaxz, betazx,
                                       betayz, betazy)

        def __call__(self, stream=None) -> None:
            f = field_abi(self.EB, self.shape)
            _abi.check(self.lib.ntt_faraday_mink(C.byref(f), self.c1, self.c2, C.byref(self.fd),
                                                 _stream(stream)), "ntt_faraday_mink")

    class Ampere_kernel(_Launch):
        def __init__(self, EB: torch.Tensor, shape: GridShape, coeff1: float, coeff2: float,
                     strict: bool = False):
            self.EB, self.shape, self.c1, self.c2, self.strict = EB, shape, coeff1, coeff2, strict

        def __call__(self, stream=None) -> None:
            f = field_abi(self.EB, self.shape)
            _abi.check(self.lib.ntt_ampere_mink(C.byref(f), self.c1, self.c2, _stream(stream)),
                       "ntt_ampere_mink")

    class CurrentsAmpere_kernel(_Launch):
        def __init__(self, E: torch.Tensor, J: torch.Tensor, shape: GridShape, coeff: float,
                     ppc0: float, strict: bool = False):
            self.E, self.J, self.shape = E, J, shape
            self.coeff, self.ppc0, self.strict = coeff, ppc0, strict

        def __call__(self, stream=None) -> None:
            e, j = field_abi(self.E, self.shape), field_abi(self.J, self.shape)
            _abi.check(self.lib.ntt_currents_ampere_mink(C.byref(e), C.byref(j), self.coeff,
                                                         self.ppc0, _stream(stream)),
                       "ntt_currents_ampere_mink")


class DigitalFilter_kernel(_Launch):
    def __init__(self, array: torch.Tensor, buffer: torch.Tensor, shape: GridShape,
                 boundaries=None, strict: bool = False):
        self.array, self.buffer, self.shape, self.strict = array, buffer, shape, strict
        self.bc = None
        if boundaries is not None:
            self.bc = _abi.NttFldsBC()
            for d, (lo, hi) in enumerate(boundaries):
                self.bc.bc[d][0], self.bc.bc[d][1] = lo, hi

    def __call__(self, stream=None) -> None:
        a, b = field_abi(self.array, self.shape), field_abi(self.buffer, self.shape)
        bc = C.byref(self.bc) if self.bc is not None else None
        _abi.check(self.lib.ntt_digital_filter(C.byref(a), C.byref(b), bc, _stream(stream)),
                   "ntt_digital_filter")


def communicate_fields(t: torch.Tensor, shape: GridShape, comp_begin: int, comp_end: int,
                       stream=None, strict: bool = False) -> None:
    """Periodic ghost fill (Metadomain::CommunicateFields, single periodic domain)."""
    f = field_abi(t, shape)
    _abi.check(_abi.load(strict).ntt_communicate_fields(C.byref(f), comp_begin, comp_end,
                                                        _stream(stream)),
               "ntt_communicate_fields")


def synchronize_currents(cur: torch.Tensor, buff: torch.Tensor, shape: GridShape,
                         stream=None, strict: bool = False) -> None:
    """Periodic current fold (Metadomain::SynchronizeFields(Comm::J))."""
    j, b = field_abi(cur, shape), field_abi(buff, shape)
    _abi.check(_abi.load(strict).ntt_synchronize_currents(C.byref(j), C.byref(b),
                                                          _stream(stream)),
               "ntt_synchronize_currents")


def device_errors(strict: bool = False) -> int:
    """Synchronize and return/clear the library's device error word."""
    v = C.c_uint32(0)
    _abi.check(_abi.load(strict).ntt_get_device_errors(C.byref(v)), "ntt_get_device_errors")
    return int(v.value)


__all__ = ["PusherContext", "PusherBoundaries", "Minkowski", "sr", "gr", "mink",
           "DepositCurrents_kernel", "DigitalFilter_kernel", "communicate_fields",
           "synchronize_currents", "device_errors", "Fields", "GridShape", "Particles"]
