"""ctypes binding of the C ABI declared in include/ntt.h (libntt_hip.so).

The structures below mirror include/ntt.h field by field.  The library is
loaded from the package's lib/ directory; importing torch first makes the
process share torch's HIP runtime with the library.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime before libntt_hip)

LIBDIR = Path(__file__).resolve().parent / "lib"

NTT_OK, NTT_ERR_ARG, NTT_ERR_HIP, NTT_ERR_DEVICE, NTT_ERR_UNSUPPORTED, NTT_ERR_CAPACITY = range(6)

DEVERR_INVALID_TAG = 0x1
DEVERR_OUT_OF_BOUNDS = 0x2
DEVERR_PUSHER_FLAGS = 0x4
DEVERR_CURRENT_RANGE = 0x8

TAG_DEAD, TAG_ALIVE = 0, 1

PUSHER_NONE = 0
PUSHER_PHOTON = 1 << 0
PUSHER_BORIS = 1 << 1
PUSHER_VAY = 1 << 2
PUSHER_GCA = 1 << 3

PRTL_BC_PERIODIC = 1
PRTL_BC_ABSORB = 2
PRTL_BC_ATMOSPHERE = 3
PRTL_BC_REFLECT = 5
PRTL_BC_HORIZON = 6
PRTL_BC_AXIS = 7

FLDS_BC_PERIODIC = 1
FLDS_BC_MATCH = 2
FLDS_BC_FIXED = 3
FLDS_BC_HORIZON = 6
FLDS_BC_AXIS = 7
FLDS_BC_CONDUCTOR = 8

_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_i16p = C.POINTER(C.c_int16)
_i64p = C.POINTER(C.c_int64)


class NttParticles(C.Structure):
    _fields_ = [
        ("i1", C.c_void_p), ("i2", C.c_void_p), ("i3", C.c_void_p),
        ("i1_prev", C.c_void_p), ("i2_prev", C.c_void_p), ("i3_prev", C.c_void_p),
        ("dx1", C.c_void_p), ("dx2", C.c_void_p), ("dx3", C.c_void_p),
        ("dx1_prev", C.c_void_p), ("dx2_prev", C.c_void_p), ("dx3_prev", C.c_void_p),
        ("ux1", C.c_void_p), ("ux2", C.c_void_p), ("ux3", C.c_void_p),
        ("weight", C.c_void_p), ("phi", C.c_void_p), ("tag", C.c_void_p),
        ("pld_i", C.c_void_p),
        ("npart", C.c_int64), ("maxnpart", C.c_int64),
    ]


class NttField(C.Structure):
    _fields_ = [("data", C.c_void_p), ("dim", C.c_int32), ("n", C.c_int32 * 3),
                ("ngh", C.c_int32), ("ncomp", C.c_int32)]


class NttSrPusherCtx(C.Structure):
    _fields_ = [("mass", C.c_float), ("charge", C.c_float), ("dt", C.c_float),
                ("omegaB0", C.c_float), ("ni1", C.c_int32), ("ni2", C.c_int32),
                ("ni3", C.c_int32), ("pusher_flags", C.c_uint8)]


class NttPrtlBC(C.Structure):
    _fields_ = [("bc", (C.c_uint8 * 2) * 3)]


class NttFldsBC(C.Structure):
    _fields_ = [("bc", (C.c_uint8 * 2) * 3)]


class NttMinkowski(C.Structure):
    _fields_ = [("dim", C.c_int32), ("dx", C.c_float)]


class NttFdCoeffs(C.Structure):
    _fields_ = [("delta_x", C.c_float), ("delta_y", C.c_float), ("delta_z", C.c_float),
                ("beta_xy", C.c_float), ("beta_yx", C.c_float), ("beta_xz", C.c_float),
                ("beta_zx", C.c_float), ("beta_yz", C.c_float), ("beta_zy", C.c_float)]


class NttSpecies(C.Structure):
    _fields_ = [("prtls", NttParticles), ("mass", C.c_float), ("charge", C.c_float),
                ("pusher_flags", C.c_uint8), ("tile_offsets", C.c_void_p), ("ntiles", C.c_int64)]


class NttSrStepParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("correction", C.c_float), ("omegaB0", C.c_float),
                ("q0", C.c_float), ("B0", C.c_float), ("V0", C.c_float), ("ppc0", C.c_float),
                ("shape_order", C.c_int32), ("n_filters", C.c_int32),
                ("fieldsolver_enable", C.c_int32), ("deposit_enable", C.c_int32),
                ("step", C.c_int64), ("fd", NttFdCoeffs), ("prtl_bc", NttPrtlBC),
                ("particles_per_thread", C.c_int32), ("weight_max", C.c_float),
                ("em_half_out", C.c_void_p)]


class NttBox(C.Structure):
    _fields_ = [("em", NttField), ("cur", NttField), ("buff", NttField),
                ("species", C.POINTER(NttSpecies)), ("nspecies", C.c_int32),
                ("prm", NttSrStepParams), ("metric", NttMinkowski), ("workspace", C.c_void_p),
                ("workspace_bytes", C.c_uint64), ("tile", C.c_int32 * 3)]


class NttPlasmaInit(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("position_seed", C.c_uint64), ("species_id", C.c_int32),
                ("ppc", C.c_int32), ("temperature", C.c_float), ("drift_u", C.c_float * 3),
                ("weight", C.c_float), ("tile", C.c_int32 * 3)]


METRIC_KERR_SCHILD, METRIC_QKERR_SCHILD, METRIC_KERR_SCHILD_0 = 0, 1, 2


class NttKsMetric(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nx1", C.c_int32), ("nx2", C.c_int32),
                ("x1_min", C.c_double), ("x1_max", C.c_double), ("x2_min", C.c_double),
                ("x2_max", C.c_double), ("a", C.c_double), ("r0", C.c_double), ("h", C.c_double)]


class NttGrPusherCtx(C.Structure):
    _fields_ = [("mass", C.c_float), ("charge", C.c_float), ("dt", C.c_float),
                ("omegaB0", C.c_float), ("epsilon", C.c_float), ("niter", C.c_int32),
                ("ni1", C.c_int32), ("ni2", C.c_int32), ("ni3", C.c_int32),
                ("pusher_flags", C.c_uint8), ("precision", C.c_uint8)]


BC_D, BC_B = 0x7, 0x38   # NTT_BC_D, NTT_BC_B


class NttGrMatch(C.Structure):
    _fields_ = [("r_edge", C.c_double), ("ds", C.c_double), ("i1_lo", C.c_int32),
                ("i1_hi", C.c_int32), ("tags", C.c_uint32), ("comps", C.c_uint32)]


class NttGrBatchEntry(C.Structure):
    _fields_ = [("ctx", NttGrPusherCtx), ("bc", NttPrtlBC), ("prtls", NttParticles),
                ("em", NttField), ("em0", NttField), ("metric", NttKsMetric)]


class NttParticleAdjoint(C.Structure):
    _fields_ = [("lx1", C.c_void_p), ("lx2", C.c_void_p), ("lu1", C.c_void_p),
                ("lu2", C.c_void_p), ("lu3", C.c_void_p), ("size", C.c_int64)]


VJP_FARADAY, VJP_AMPERE, VJP_FILTER = 0, 1, 2


class NttError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: status {status} ({_STATUS_NAMES.get(status, '?')})")


_STATUS_NAMES = {0: "ok", 1: "invalid argument", 2: "HIP runtime error",
                 3: "device error flag set", 4: "unsupported configuration",
                 5: "capacity or workspace too small"}

_P = C.POINTER

# name -> (restype, argtypes); every symbol declared in include/ntt.h
SIGNATURES: dict[str, tuple] = {
    "ntt_abi_version": (C.c_int32, []),
    "ntt_status_string": (C.c_char_p, [C.c_int]),
    "ntt_get_device_errors": (C.c_int, [_P(C.c_uint32)]),
    "ntt_sr_push": (C.c_int, [_P(NttSrPusherCtx), _P(NttPrtlBC), _P(NttParticles), _P(NttField),
                              _P(NttMinkowski), C.c_int32, C.c_void_p]),
    "ntt_gr_push": (C.c_int, [_P(NttGrPusherCtx), _P(NttPrtlBC), _P(NttParticles), _P(NttField),
                              _P(NttField), _P(NttKsMetric), C.c_int32, C.c_void_p]),
    "ntt_sr_push_vjp": (C.c_int, [_P(NttSrPusherCtx), _P(NttParticles), _P(NttField),
                                  _P(NttMinkowski), C.c_int32, C.c_int32,
                                  _P(NttParticleAdjoint), _P(NttParticleAdjoint), C.c_void_p,
                                  C.c_void_p]),
    "ntt_deposit_vjp": (C.c_int, [_P(NttParticles), _P(C.c_int32), C.c_float, C.c_float,
                                  C.c_int32, C.c_int32, C.c_void_p, _P(NttParticleAdjoint),
                                  _P(NttParticleAdjoint), C.c_void_p]),
    "ntt_sr_particle_vjp": (C.c_int, [_P(NttSrPusherCtx), _P(NttParticles), _P(NttParticles),
                                      _P(NttField), _P(NttMinkowski), C.c_int32, C.c_int32,
                                      C.c_void_p, _P(NttParticleAdjoint),
                                      _P(NttParticleAdjoint), C.c_void_p, C.c_void_p]),
    "ntt_yee_vjp": (C.c_int, [C.c_int32, _P(C.c_int32), C.c_float, C.c_float, C.c_int32,
                              C.c_void_p, C.c_void_p, C.c_void_p]),
    "ntt_gr_batch_create": (C.c_int, [_P(NttGrBatchEntry), C.c_int32, C.c_int32,
                                      _P(C.c_void_p)]),
    "ntt_gr_batch_push": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ntt_gr_batch_set_npart": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p]),
    "ntt_gr_batch_destroy": (C.c_int, [C.c_void_p]),
    "ntt_gr_aux_e": (C.c_int, [_P(NttField), _P(NttField), _P(NttField), _P(NttKsMetric),
                               _P(NttFldsBC), C.c_void_p]),
    "ntt_gr_aux_h": (C.c_int, [_P(NttField), _P(NttField), _P(NttField), _P(NttKsMetric),
                               _P(NttFldsBC), C.c_void_p]),
    "ntt_gr_faraday": (C.c_int, [_P(NttField), _P(NttField), _P(NttField), _P(NttKsMetric),
                                 C.c_float, _P(NttFldsBC), C.c_void_p]),
    "ntt_gr_ampere": (C.c_int, [_P(NttField), _P(NttField), _P(NttField), _P(NttKsMetric),
                                C.c_float, _P(NttFldsBC), C.c_void_p]),
    "ntt_gr_currents_ampere": (C.c_int, [_P(NttField), _P(NttField), _P(NttKsMetric), C.c_float,
                                         _P(NttFldsBC), C.c_void_p]),
    "ntt_gr_time_average_db": (C.c_int, [_P(NttField), _P(NttField), C.c_void_p]),
    "ntt_gr_time_average_j": (C.c_int, [_P(NttField), _P(NttField), C.c_void_p]),
    "ntt_gr_match_boundaries": (C.c_int, [_P(NttField), _P(NttField), _P(NttKsMetric),
                                          _P(NttFldsBC), _P(NttGrMatch), C.c_void_p]),
    "ntt_gr_horizon_boundaries": (C.c_int, [_P(NttField), C.c_uint32, C.c_int32, C.c_void_p]),
    "ntt_gr_axis_boundaries": (C.c_int, [_P(NttField), C.c_int32, C.c_uint32, C.c_void_p]),
    "ntt_gr_absorb_currents": (C.c_int, [_P(NttField), _P(NttKsMetric), C.c_double, C.c_double,
                                         C.c_int32, C.c_int32, C.c_void_p]),
    "ntt_gr_digital_filter": (C.c_int, [_P(NttField), _P(NttField), _P(NttFldsBC), C.c_void_p]),
    "ntt_gr_deposit_currents": (C.c_int, [_P(NttParticles), _P(NttField), _P(NttKsMetric),
                                          C.c_float, C.c_float, C.c_int32, C.c_void_p]),
    "ntt_deposit_currents": (C.c_int, [_P(NttParticles), _P(NttField), _P(NttMinkowski),
                                       C.c_float, C.c_float, C.c_int32, C.c_void_p]),
    "ntt_faraday_mink": (C.c_int, [_P(NttField), C.c_float, C.c_float, _P(NttFdCoeffs),
                                   C.c_void_p]),
    "ntt_ampere_mink": (C.c_int, [_P(NttField), C.c_float, C.c_float, C.c_void_p]),
    "ntt_currents_ampere_mink": (C.c_int, [_P(NttField), _P(NttField), C.c_float, C.c_float,
                                           C.c_void_p]),
    "ntt_digital_filter": (C.c_int, [_P(NttField), _P(NttField), _P(NttFldsBC), C.c_void_p]),
    "ntt_communicate_fields": (C.c_int, [_P(NttField), C.c_int32, C.c_int32, C.c_void_p]),
    "ntt_synchronize_currents": (C.c_int, [_P(NttField), _P(NttField), C.c_void_p]),
    "ntt_sr_step": (C.c_int, [_P(NttField), _P(NttField), _P(NttField), _P(NttSpecies), C.c_int32,
                              _P(NttSrStepParams), _P(NttMinkowski), C.c_void_p, C.c_uint64,
                              C.c_void_p]),
    "ntt_sr_step_workspace_size": (C.c_int, [_P(NttField), _P(C.c_uint64)]),
    "ntt_sr_step_workspace_size_species": (C.c_int, [_P(NttField), C.c_int32, _P(C.c_uint64)]),
    "ntt_sr_step_static": (C.c_int, [_P(NttField), _P(NttField), _P(NttField), _P(NttSpecies),
                                     C.c_int32, _P(NttSrStepParams), _P(NttMinkowski),
                                     C.c_void_p]),
    "ntt_fused_window": (C.c_int, [C.c_int32, _P(C.c_int32)]),
    "ntt_ensemble_create": (C.c_int, [_P(NttBox), C.c_int32, _P(C.c_void_p)]),
    "ntt_ensemble_step": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "ntt_ensemble_sort": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ntt_ensemble_destroy": (C.c_int, [C.c_void_p]),
    "ntt_sort_workspace_size": (C.c_int, [C.c_int64, C.c_int32, _P(C.c_int32), _P(C.c_int32),
                                          _P(C.c_uint64)]),
    "ntt_sort_particles": (C.c_int, [_P(NttParticles), C.c_int32, _P(C.c_int32), _P(C.c_int32),
                                     C.c_void_p, C.c_uint64, C.c_void_p, _P(C.c_int64),
                                     C.c_void_p]),
    "ntt_sort_particles_to": (C.c_int, [_P(NttParticles), _P(NttParticles), C.c_int32,
                                        _P(C.c_int32), _P(C.c_int32), C.c_void_p, C.c_uint64,
                                        C.c_void_p, _P(C.c_int64), C.c_void_p]),
    "ntt_sort_workspace_size_to": (C.c_int, [C.c_int64, C.c_int32, _P(C.c_int32), _P(C.c_int32),
                                             _P(C.c_uint64)]),
    "ntt_sort_ntiles": (C.c_int, [C.c_int32, _P(C.c_int32), _P(C.c_int32), _P(C.c_int64)]),
    "ntt_profile_enable": (C.c_int, [C.c_int32]),
    "ntt_profile_read": (C.c_int, [_P(C.c_double), _P(C.c_int64), C.c_int32]),
    "ntt_init_uniform_plasma": (C.c_int, [_P(NttParticles), C.c_int32, _P(C.c_int32),
                                          _P(NttPlasmaInit), C.c_void_p]),
}

_LIBS: dict[str, C.CDLL] = {}


def library_path(strict: bool = False) -> Path:
    return LIBDIR / ("libntt_hip_strict.so" if strict else "libntt_hip.so")


def load(strict: bool = False) -> C.CDLL:
    """Load (once) and return libntt_hip.so, or the strict-FP build."""
    key = "strict" if strict else "fast"
    if key in _LIBS:
        return _LIBS[key]
    path = library_path(strict)
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: the HIP extension is not built "
            "(run `python -c 'import __graft_entry__ as g; g.build()'`)")
    lib = C.CDLL(str(path), mode=C.RTLD_LOCAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _LIBS[key] = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != NTT_OK:
        raise NttError(status, what)
