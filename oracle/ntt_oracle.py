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
import math
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
    threads=True runs the step's loops on OpenMP threads with per-thread
    current copies (the CPU baseline in FP32, finite-difference references in
    FP64); the default is the sequential checker."""
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
        return getattr(lib(), f"ntt_oracle_sr_step_omp_{prec}")(*args)
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


# ---- GR field boundaries, the spherical filter, the GR deposit and the GRPIC
# step (src/engines/grpic/grpic.hpp:66-634) -----------------------------------
GR_BC_OPS = {"match": 0, "horizon": 1, "axis": 2, "absorb_currents": 3, "filter": 4}
# the reference's BCTags bits (src/global/global.h:306-329): D shares the bits
# of E, H those of B
BC_D, BC_B = 0x7, 0x38
FLDS_MATCH, FLDS_HORIZON, FLDS_AXIS = 2, 6, 7
PRTL_ABSORB, PRTL_HORIZON, PRTL_AXIS = 2, 6, 7


class OrtGrBcArgs(C.Structure):
    _fields_ = [("xg_edge", C.c_double), ("ds", C.c_double), ("i1_lo", C.c_int32),
                ("i1_hi", C.c_int32), ("tags", C.c_uint32), ("comps", C.c_uint32),
                ("nfilter", C.c_int32), ("upper", C.c_int32), ("axis_min", C.c_int32),
                ("axis_max", C.c_int32)]


def gr_bc(op: str, prm: OrtGrParams, a, b=None, *, xg_edge=0.0, ds=1.0, i1_range=(0, 0),
          tags=BC_D | BC_B, comps=0x3F, nfilter=0, upper=False, axis=(True, True), mutate=0,
          prec="f64") -> int:
    """One GR boundary operation on 2D fields (arrays (ncomp, N2, N1)):
    match(F=a, target=b), horizon(F=a), axis(F=a), absorb_currents(J=a),
    filter(array=a, buffer=b)."""
    g = OrtGrBcArgs(float(xg_edge), float(ds), int(i1_range[0]), int(i1_range[1]), int(tags),
                    int(comps), int(nfilter), int(upper), int(axis[0]), int(axis[1]))
    fa = field_struct(a, 2)
    fb = field_struct(b, 2) if b is not None else None
    fn = getattr(lib(), f"ntt_oracle_gr_bc_{prec}")
    fn.restype = C.c_int
    fn.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
    return fn(GR_BC_OPS[op], C.addressof(prm), C.addressof(fa),
              C.addressof(fb) if fb is not None else None, C.addressof(g), int(mutate))


def gr_x1_to_code(prm: OrtGrParams, r: float) -> float:
    """convert<1, Ph, Cd> (kerr_schild.h, qkerr_schild.h): radius to code x1."""
    if prm.metric_kind == METRIC_KIND["qkerr_schild"]:
        chi_min = math.log(prm.x1_min - prm.r0)
        dchi = (math.log(prm.x1_max - prm.r0) - chi_min) / prm.nx1
        return (math.log(r - prm.r0) - chi_min) / dchi
    return (r - prm.x1_min) / ((prm.x1_max - prm.x1_min) / prm.nx1)


def gr_match_range(prm: OrtGrParams, upper: bool, ds: float, G: int = 2):
    """The x1 cell range of a match layer of width ds at one r boundary:
    Mesh::ExtentToRange (src/framework/domain/mesh.h:133-203) of the box that
    MatchFieldsIn builds (src/engines/grpic/fields_bcs.h:41-94), with ghosts
    excluded below and included above.  Returns (lo, hi), ghost-extended
    indices, and the edge radius."""
    n1 = prm.nx1
    if upper:
        lo_r, hi_r, edge = prm.x1_max - ds, prm.x1_max, prm.x1_max
    else:
        lo_r, hi_r, edge = prm.x1_min, prm.x1_min + ds, prm.x1_min
    lo_r = min(max(prm.x1_min, lo_r), prm.x1_max)
    hi_r = max(min(prm.x1_max, hi_r), prm.x1_min)
    lo = max(math.floor(gr_x1_to_code(prm, lo_r)), 0)
    hi = math.ceil(gr_x1_to_code(prm, hi_r))
    return lo + G, min(hi + 2 * G, n1 + 2 * G), edge


def gr_deposit(prm: OrtGrParams, p: dict, J, *, charge, dt, prec="f64") -> int:
    """DepositCurrents_kernel<GRPIC, M> (always order 1,
    src/engines/grpic/currents.h:34-77): the 2D Esirkepov deposit with the
    out-of-plane velocity of the metric."""
    q = p if "_n" in p else normalize_particles(p, prec)
    s = particles_struct(q)
    vz = np.zeros(max(q["_n"], 1))
    fn = getattr(lib(), f"ntt_oracle_gr_deposit_velocity_{prec}")
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    assert fn(C.addressof(prm), C.addressof(s), _ptr(vz)) == 0
    f = field_struct(J, 2)
    return _call("deposit_vz", prec, C.byref(s), C.byref(f), C.c_double(charge),
                 C.c_double(dt), C.c_int32(1), C.c_void_p(_ptr(vz)))


class GRStep:
    """The GRPIC step of one 2D (r, theta) domain in FP64
    (GRPICEngine::step_forward, src/engines/grpic/grpic.hpp:66-634): fields
    em, em0, aux (D/E in 0..2, B/H in 3..5), cur, cur0, buff; species as
    (prm, particle dict) pairs; boundaries as in grpic/fields_bcs.h:234-266.
    flds_bc / prtl_bc: ((x1 min, x1 max), (x2 min, x2 max)) codes; ds: match
    layer widths ((x1 min, x1 max), ...); target: the fields the match layer
    relaxes to (the problem's initial fields), comps: which it defines."""

    def __init__(self, prm: OrtGrParams, em, species, *, flds_bc, prtl_bc, ds, target,
                 comps=0x3F, nfilter=0, q0=1.0, B0=1.0, order=1, fieldsolver=True, deposit=True,
                 clear_interval=100, correction=1.0):
        self.prm = prm
        self.em = np.array(em, dtype=np.float64)
        self.em0 = np.zeros_like(self.em)
        self.aux = np.zeros_like(self.em)
        shp = (3,) + self.em.shape[1:]
        self.cur, self.cur0, self.buff = (np.zeros(shp) for _ in range(3))
        self.species = [(sp, normalize_particles(p, "f64")) for sp, p in species]
        self.flds_bc, self.prtl_bc, self.ds = flds_bc, prtl_bc, ds
        self.target = np.array(target, dtype=np.float64)
        self.comps, self.nfilter, self.order = comps, nfilter, order
        self.fieldsolver, self.deposit = fieldsolver, deposit
        self.clear_interval = clear_interval
        self.dt = prm.dt
        self.dT = correction * prm.dt
        self.coeff_j = -prm.dt * q0 / B0
        self.axis = (flds_bc[1][0] == FLDS_AXIS, flds_bc[1][1] == FLDS_AXIS)
        self.step = 0

    # ---- grpic/fields_bcs.h:234-266 ----
    def boundaries(self, tags, kind="main"):
        if kind == "aux":      # only horizon sides act on aux, and only for main
            return
        prm = self.prm
        if kind == "curr":
            for d, side in ((0, 0), (1, 0), (1, 1), (0, 1)):
                if self.prtl_bc[d][side] == PRTL_ABSORB:
                    if d != 0 or side == 0:
                        raise ValueError("absorption of currents only in +x1")
                    lo, hi, edge = gr_match_range(prm, True, self.ds[0][1])
                    assert gr_bc("absorb_currents", prm, self.cur0, xg_edge=edge,
                                 ds=self.ds[0][1], i1_range=(lo, hi)) == 0
            return
        for d, side in ((0, 0), (1, 0), (1, 1), (0, 1)):   # Directions<2D>::orth
            bc = self.flds_bc[d][side]
            for F in (self.em, self.em0):
                if bc == FLDS_MATCH:
                    if d != 0:
                        raise ValueError("match boundaries only in x1")
                    lo, hi, edge = gr_match_range(prm, side == 1, self.ds[0][side])
                    assert gr_bc("match", prm, F, self.target, xg_edge=edge, ds=self.ds[0][side],
                                 i1_range=(lo, hi), tags=tags, comps=self.comps,
                                 axis=self.axis) == 0
                elif bc == FLDS_AXIS:
                    assert gr_bc("axis", prm, F, tags=tags, upper=side == 1) == 0
                elif bc == FLDS_HORIZON:
                    assert gr_bc("horizon", prm, F, tags=tags, nfilter=self.nfilter) == 0

    def _f(self, op, a, b, c=None, coeff=0.0):
        assert gr_field(op, self.prm, a, b, c, coeff=coeff, axis=self.axis) == 0

    def _swap(self):
        self.em, self.em0 = self.em0, self.em
        self.cur, self.cur0 = self.cur0, self.cur

    def _init_fields(self):
        """Step 0 with the field solver on (grpic.hpp:74-280)."""
        dT = self.dT
        self.boundaries(BC_B | BC_D)
        self.em0[...] = self.em
        self._f("aux_e", self.em, self.em0, self.aux)
        self._f("aux_h", self.em, self.em0, self.aux)
        self._f("faraday", self.em0, self.em0, self.aux, 0.5 * dT)
        self.boundaries(BC_B)
        self._f("ampere", self.em, self.em, self.aux, 0.5 * dT)
        self.boundaries(BC_D)
        self._f("aux_e", self.em, self.em0, self.aux)
        self._f("aux_h", self.em, self.em0, self.aux)
        self._f("faraday", self.em, self.em0, self.aux, dT)
        self.boundaries(BC_B)
        self._f("ampere", self.em0, self.em0, self.aux, dT)
        self.boundaries(BC_D)
        self._f("aux_h", self.em0, self.em0, self.aux)
        self._f("ampere", self.em, self.em0, self.aux, dT)
        self.boundaries(BC_D)
        self._swap()

    def step_forward(self):
        dT = self.dT
        if self.step == 0:
            if self.fieldsolver:
                self._init_fields()
            else:
                self.em0[...] = self.em
        if self.fieldsolver:
            self._f("time_average_db", self.em, self.em0)
            self._f("aux_e", self.em0, self.em, self.aux)
            self._f("faraday", self.em0, self.em0, self.aux, dT)
            self.boundaries(BC_B)
            self._f("aux_h", self.em, self.em0, self.aux)
        for sp, p in self.species:
            if p["_n"]:
                assert gr_push(sp, p, self.em, self.em0, order=self.order, prec="f64") == 0
        if self.deposit:
            self.cur0[...] = 0.0
            for sp, p in self.species:
                if p["_n"] and sp.charge != 0.0:
                    assert gr_deposit(sp, p, self.cur0, charge=sp.charge, dt=self.dt) == 0
            self.boundaries(0x1C0, "curr")
            for _ in range(self.nfilter):
                self.buff[...] = self.cur0
                assert gr_bc("filter", self.prm, self.cur0, self.buff, axis=self.axis) == 0
        if self.fieldsolver:
            if self.deposit:
                self._f("time_average_j", self.cur0, self.cur)
            self._f("aux_e", self.em, self.em0, self.aux)
            self._f("faraday", self.em, self.em0, self.aux, dT)
            self.boundaries(BC_B)
            self._f("ampere", self.em0, self.em0, self.aux, dT)
            if self.deposit:
                self._f("currents_ampere", self.em0, self.cur, coeff=self.coeff_j)
            self.boundaries(BC_D)
            self._f("aux_h", self.em0, self.em0, self.aux)
            self._f("ampere", self.em, self.em0, self.aux, dT)
            if self.deposit:
                self._f("currents_ampere", self.em0, self.cur0, coeff=self.coeff_j)
            self._swap()
            self.boundaries(BC_D)
        if self.clear_interval and self.step > 0 and self.step % self.clear_interval == 0:
            for k, (sp, p) in enumerate(self.species):
                live = p["tag"][: p["_n"]] == 1
                q = {key: (v[: p["_n"]][live].copy() if key != "_n" else v)
                     for key, v in p.items()}
                q["_n"] = int(live.sum())
                self.species[k] = (sp, q)
        self.step += 1
