"""ctypes binding of the CPU oracle (libntt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or the CPU
baseline, never as the thing measured or shipped.

Arrays are numpy; particle dicts use the reference's SoA names
(i1, i1_prev, dx1, dx1_prev, ux1, weight, tag, ...), fields are arrays of shape
(ncomp, [N3,] N2, N1) (component slowest, x fastest).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libntt_oracle.so"

_lib = None


class OrtParticles(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "i1", "i2", "i3", "i1_prev", "i2_prev", "i3_prev", "dx1", "dx2", "dx3",
        "dx1_prev", "dx2_prev", "dx3_prev", "ux1", "ux2", "ux3", "weight", "phi", "tag")] + [
        ("npart", C.c_int64)]


class OrtField(C.Structure):
    _fields_ = [("data", C.c_void_p), ("dim", C.c_int32), ("n", C.c_int32 * 3),
                ("ngh", C.c_int32), ("ncomp", C.c_int32)]


class OrtSrParams(C.Structure):
    _fields_ = [("mass", C.c_float), ("charge", C.c_float), ("dt", C.c_double),
                ("omegaB0", C.c_double), ("dx", C.c_double), ("ni1", C.c_int32),
                ("ni2", C.c_int32), ("ni3", C.c_int32), ("pusher_flags", C.c_uint8),
                ("prtl_bc", (C.c_uint8 * 2) * 3), ("shape_order", C.c_int32)]


class OrtFdCoeffs(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("delta_x", "delta_y", "delta_z", "beta_xy", "beta_yx",
                                          "beta_xz", "beta_zx", "beta_yz", "beta_zy")]


class OrtSpecies(C.Structure):
    _fields_ = [("prtls", OrtParticles), ("mass", C.c_float), ("charge", C.c_float),
                ("pusher_flags", C.c_uint8)]


class OrtStepParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("correction", C.c_double), ("omegaB0", C.c_double),
                ("q0", C.c_double), ("B0", C.c_double), ("V0", C.c_double), ("ppc0", C.c_double),
                ("dx", C.c_double), ("shape_order", C.c_int32), ("n_filters", C.c_int32),
                ("fieldsolver_enable", C.c_int32), ("deposit_enable", C.c_int32),
                ("step", C.c_int64), ("fd", OrtFdCoeffs)]


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: build the oracle (make -C oracle)")
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


REAL = {"f32": np.float32, "f64": np.float64}
PRT_INT = ("i1", "i2", "i3", "i1_prev", "i2_prev", "i3_prev")
PRT_DX = ("dx1", "dx2", "dx3", "dx1_prev", "dx2_prev", "dx3_prev")
PRT_REAL = ("ux1", "ux2", "ux3", "weight", "phi")


def _ptr(a):
    return None if a is None else a.ctypes.data


def normalize_particles(p: dict, prec: str) -> dict:
    """Return a dict of contiguous arrays with the dtypes the oracle expects
    (copies, so the oracle may update them in place)."""
    out = {}
    n = None
    for k, v in p.items():
        if k in PRT_INT:
            a = np.ascontiguousarray(v, dtype=np.int32).copy()
        elif k in PRT_DX:
            a = np.ascontiguousarray(v, dtype=np.float32).copy()
        elif k in PRT_REAL:
            a = np.ascontiguousarray(v, dtype=REAL[prec]).copy()
        elif k == "tag":
            a = np.ascontiguousarray(v, dtype=np.int16).copy()
        else:
            a = np.array(v, copy=True)
        out[k] = a
        if k != "pld_i":
            n = a.shape[0] if n is None else n
    out["_n"] = n or 0
    return out


def particles_struct(p: dict) -> OrtParticles:
    s = OrtParticles()
    for name in PRT_INT + PRT_DX + PRT_REAL + ("tag",):
        setattr(s, name, _ptr(p.get(name)))
    s.npart = p["_n"]
    return s


def field_struct(arr: np.ndarray, dim: int, ngh: int = 2) -> OrtField:
    assert arr.flags.c_contiguous
    f = OrtField()
    f.data = arr.ctypes.data
    f.dim = dim
    ext = list(reversed(arr.shape[1:]))
    n = [e - 2 * ngh for e in ext] + [1] * (3 - dim)
    f.n[0], f.n[1], f.n[2] = n
    f.ngh = ngh
    f.ncomp = arr.shape[0]
    return f


def _call(name, prec, *args):
    fn = getattr(lib(), f"ntt_oracle_{name}_{prec}")
    rc = fn(*args)
    return rc


def sr_push(p: dict, em: np.ndarray, *, dim, mass, charge, dt, omegaB0, pusher, order,
            bc=((1, 1), (1, 1), (1, 1)), dx=1.0, prec="f32"):
    """Runs the pusher over all particles in place; returns the error code."""
    q = p if "_n" in p else normalize_particles(p, prec)
    prm = OrtSrParams()
    prm.mass, prm.charge, prm.dt, prm.omegaB0, prm.dx = mass, charge, dt, omegaB0, dx
    ext = list(reversed(em.shape[1:]))
    ni = [e - 4 for e in ext] + [0] * (3 - dim)
    prm.ni1, prm.ni2, prm.ni3 = ni
    prm.pusher_flags = pusher
    for d in range(3):
        prm.prtl_bc[d][0], prm.prtl_bc[d][1] = bc[d]
    prm.shape_order = order
    s = particles_struct(q)
    f = field_struct(em, dim)
    rc = _call("sr_push", prec, C.byref(prm), C.byref(s), C.byref(f))
    return q, rc


def interpolate(p: dict, em: np.ndarray, *, dim, order, prec="f32"):
    q = normalize_particles(p, prec)
    prm = OrtSrParams()
    prm.shape_order = order
    n = q["_n"]
    e = np.zeros((n, 3), dtype=REAL[prec])
    b = np.zeros((n, 3), dtype=REAL[prec])
    s = particles_struct(q)
    f = field_struct(em, dim)
    _call("interp", prec, C.byref(prm), C.byref(s), C.byref(f), C.c_void_p(e.ctypes.data),
          C.c_void_p(b.ctypes.data))
    return e, b


def deposit(p: dict, J: np.ndarray, *, dim, charge, dt, order, dx=1.0, prec="f32"):
    q = p if "_n" in p else normalize_particles(p, prec)
    s = particles_struct(q)
    f = field_struct(J, dim)
    return _call("deposit", prec, C.byref(s), C.byref(f), C.c_double(charge), C.c_double(dt),
                 C.c_double(dx), C.c_int32(order))


def faraday(em, *, dim, coeff1, coeff2, fd=None, prec="f32"):
    f = field_struct(em, dim)
    c = OrtFdCoeffs(**(fd or {}))
    return _call("faraday", prec, C.byref(f), C.c_double(coeff1), C.c_double(coeff2), C.byref(c))


def ampere(em, *, dim, coeff1, coeff2, prec="f32"):
    f = field_struct(em, dim)
    return _call("ampere", prec, C.byref(f), C.c_double(coeff1), C.c_double(coeff2))


def currents_ampere(em, J, *, dim, coeff, ppc0, prec="f32"):
    return _call("currents_ampere", prec, C.byref(field_struct(em, dim)),
                 C.byref(field_struct(J, dim)), C.c_double(coeff), C.c_double(ppc0))


def filter_pass(arr, buf, *, dim, prec="f32"):
    return _call("filter", prec, C.byref(field_struct(arr, dim)), C.byref(field_struct(buf, dim)))


def communicate(fld, c0, c1, *, dim, prec="f32"):
    return _call("communicate", prec, C.byref(field_struct(fld, dim)), C.c_int32(c0),
                 C.c_int32(c1))


def synchronize(cur, buff, *, dim, prec="f32"):
    return _call("synchronize", prec, C.byref(field_struct(cur, dim)),
                 C.byref(field_struct(buff, dim)))


def sr_step(em, cur, buff, species: list[dict], *, dim, prm: dict, bc, prec="f32",
            threads=False):
    """One SRPIC step (srpic.hpp:65-188) over normalized species dicts.
    species entries: {"p": normalized particle dict, "mass", "charge", "pusher"}.
    threads=True (FP32 only) runs the step's loops on OpenMP threads with
    per-thread current copies (the CPU baseline); the default is the
    sequential checker."""
    arr = (OrtSpecies * len(species))()
    for k, sp in enumerate(species):
        arr[k].prtls = particles_struct(sp["p"])
        arr[k].mass, arr[k].charge, arr[k].pusher_flags = sp["mass"], sp["charge"], sp["pusher"]
    sp_ = OrtStepParams()
    for key, val in prm.items():
        if key == "fd":
            sp_.fd = OrtFdCoeffs(**val)
        else:
            setattr(sp_, key, val)
    bcs = ((C.c_uint8 * 2) * 3)()
    for d in range(3):
        bcs[d][0], bcs[d][1] = bc[d]
    args = (C.byref(field_struct(em, dim)), C.byref(field_struct(cur, dim)),
            C.byref(field_struct(buff, dim)), arr, C.c_int32(len(species)), C.byref(sp_), bcs)
    if threads:
        if prec != "f32":
            raise ValueError("the threaded step is FP32")
        return lib().ntt_oracle_sr_step_omp_f32(*args)
    return _call("sr_step", prec, *args)


def omp_threads() -> int:
    """Threads the threaded step uses (omp_get_max_threads)."""
    return int(lib().ntt_oracle_omp_threads())


# ---- GR pusher over the Kerr-Schild metrics ---------------------------------------
METRIC_KIND = {"kerr_schild": 0, "qkerr_schild": 1, "kerr_schild_0": 2}

# order of the quantities returned by gr_metric, per point
GR_METRIC_QUANTITIES = (
    "h_11", "h_22", "h_33", "h_13", "h11", "h22", "h33", "h13", "alpha", "beta1",
    "dr_alpha", "dt_alpha", "dr_beta1", "dt_beta1", "dr_h11", "dr_h22", "dr_h33", "dr_h13",
    "dt_h11", "dt_h22", "dt_h33", "dt_h13", "r", "theta", "sqrt_det_h", "sqrt_det_h_tilde",
    "polar_area")

GR_TRANSFORMS = {"D->U": 0, "U->D": 1, "D->T": 2, "T->D": 3, "U->T": 4, "T->U": 5}


class OrtGrParams(C.Structure):
    _fields_ = [("mass", C.c_float), ("charge", C.c_float), ("dt", C.c_double),
                ("omegaB0", C.c_double), ("niter", C.c_int32), ("ni1", C.c_int32),
                ("ni2", C.c_int32), ("massless", C.c_int32), ("absorb_i1min", C.c_int32),
                ("absorb_i1max", C.c_int32), ("axis_i2min", C.c_int32),
                ("axis_i2max", C.c_int32), ("metric_kind", C.c_int32), ("a", C.c_double),
                ("r0", C.c_double), ("h", C.c_double), ("x1_min", C.c_double),
                ("x1_max", C.c_double), ("x2_min", C.c_double), ("x2_max", C.c_double),
                ("nx1", C.c_int32), ("nx2", C.c_int32)]


def gr_params(metric: str, n, extent, *, a=0.0, r0=0.0, h=0.0, dt=0.0, mass=1.0, charge=1.0,
              omegaB0=1.0, niter=10, massless=False, absorb=(True, True), axis=(True, True)):
    p = OrtGrParams()
    p.mass, p.charge, p.dt, p.omegaB0, p.niter = mass, charge, dt, omegaB0, niter
    p.ni1, p.ni2 = n
    p.nx1, p.nx2 = n
    p.massless = int(massless)
    p.absorb_i1min, p.absorb_i1max = (int(x) for x in absorb)
    p.axis_i2min, p.axis_i2max = (int(x) for x in axis)
    p.metric_kind = METRIC_KIND[metric]
    p.a, p.r0, p.h = a, r0, h
    (p.x1_min, p.x1_max), (p.x2_min, p.x2_max) = extent
    return p


def gr_push(prm: OrtGrParams, p: dict, em=None, em0=None, *, order=1, prec="f64") -> int:
    """kernel::gr::Pusher_kernel over every particle (D from em, B from em0)."""
    s = particles_struct(p)
    fe = field_struct(em, 2) if em is not None else None
    f0 = field_struct(em0, 2) if em0 is not None else None
    fn = getattr(lib(), f"ntt_oracle_gr_push_{prec}")
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
    return fn(C.addressof(prm), C.addressof(s), C.addressof(fe) if fe is not None else None,
              C.addressof(f0) if f0 is not None else None, int(order))


def gr_metric(prm: OrtGrParams, x1, x2, *, prec="f64") -> dict:
    """Metric quantities at code coordinates (arrays); keys GR_METRIC_QUANTITIES."""
    x1 = np.ascontiguousarray(x1, dtype=np.float64).ravel()
    x2 = np.ascontiguousarray(x2, dtype=np.float64).ravel()
    out = np.zeros((x1.size, len(GR_METRIC_QUANTITIES)))
    fn = getattr(lib(), f"ntt_oracle_gr_metric_{prec}")
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    assert fn(C.addressof(prm), x1.size, _ptr(x1), _ptr(x2), _ptr(out)) == 0
    return {k: out[:, q] for q, k in enumerate(GR_METRIC_QUANTITIES)}


def gr_transform(prm: OrtGrParams, which: str, x1, x2, v, *, prec="f64") -> np.ndarray:
    """Vector transformation ("D->U", "U->D", "D->T", "T->D", "U->T", "T->U")."""
    x1 = np.ascontiguousarray(x1, dtype=np.float64).ravel()
    x2 = np.ascontiguousarray(x2, dtype=np.float64).ravel()
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(v)
    fn = getattr(lib(), f"ntt_oracle_gr_transform_{prec}")
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_void_p]
    assert fn(C.addressof(prm), GR_TRANSFORMS[which], x1.size, _ptr(x1), _ptr(x2),
              _ptr(v), _ptr(out)) == 0
    return out


GR_FIELD_OPS = {"aux_e": 0, "aux_h": 1, "faraday": 2, "ampere": 3, "currents_ampere": 4,
                "time_average_db": 5, "time_average_j": 6}


def gr_field(op: str, prm: OrtGrParams, a, b, c=None, *, coeff=0.0, axis=(True, True),
             mutate=0, prec="f64") -> int:
    """One GR field-solver operation on 2D fields (arrays (ncomp, N2, N1)):
    aux_e(D=a, B=b, E=c), aux_h(D=a, B=b, H=c), faraday(Bin=a, Bout=b, E=c),
    ampere(Din=a, Dout=b, H=c), currents_ampere(D=a, J=b), time_average_db
    (BD=a, BD0=b: b = (a + b) / 2), time_average_j(J0=a, J=b)."""
    fa = field_struct(a, 2)
    fb = field_struct(b, 2)
    fc = field_struct(c, 2) if c is not None else None
    fn = getattr(lib(), f"ntt_oracle_gr_field_{prec}")
    fn.restype = C.c_int
    fn.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                   C.c_int32, C.c_int32, C.c_int32]
    return fn(GR_FIELD_OPS[op], C.addressof(prm), C.addressof(fa), C.addressof(fb),
              C.addressof(fc) if fc is not None else None, float(coeff), int(axis[0]),
              int(axis[1]), int(mutate))
