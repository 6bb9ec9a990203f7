#!/usr/bin/env python3
"""Throughput of the SR PIC step (push + deposit + fields) on one MI355X.

Workload (default): 3D periodic 512 x 256 x 256 cells, e-/e+ pair plasma at
16 particles per cell per species (1,073,741,824 particles), theta = 0.3
thermal spread, background B0 along z plus a 16-mode random-phase
perturbation (|k| <= 4 * 2 pi / L, dB/B0 = 0.5), shape order 2, Vay pusher,
2 filter passes, CFL 0.5, tile re-sort every 4 steps.  Synthetic inputs are
generated on the GPU.

One step = ntt_sr_step (Faraday 1/2, fused push+deposit per species,
current fold, filters, Faraday 1/2, Ampere + current source) plus the tile
re-sort every --sort-interval steps, all inside the timed region.

Prints one JSON line.  Also reports the fused push+deposit kernel against the
HBM roofline (HIP events around every launch in the timed region) and a CPU
baseline: the C++ restatement of the step (not the reference binary) on
OpenMP threads, on a sub-box of the same workload, on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s (push+deposit+fields) 3D SR; batched & adjoint vs plain"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=str, default="512,256,256")
    ap.add_argument("--ppc", type=int, default=16, help="particles per cell per species")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--pusher", choices=["vay", "boris"], default="vay")
    ap.add_argument("--filters", type=int, default=2)
    ap.add_argument("--sort-interval", type=int, default=4)
    ap.add_argument("--tile", type=str, default="8,4,4")
    ap.add_argument("--double-buffer", type=int, default=1,
                    help="re-sort out of place into a second set of particle arrays")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-cells", type=str, default="128,128,64")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--batched", type=int, default=1, help="also time the 64-box ensemble")
    ap.add_argument("--batched-steps", type=int, default=12)
    ap.add_argument("--gr", type=int, default=1, help="also time the GR pusher runs")
    ap.add_argument("--gr-steps", type=int, default=300)
    ap.add_argument("--gr-full", type=int, default=1,
                    help="also time the full GRPIC step (field solver and deposit) on one run")
    ap.add_argument("--gr-full-steps", type=int, default=20)
    ap.add_argument("--adjoint", type=int, default=1, help="also time the adjoint (gradient) run")
    ap.add_argument("--autotune", type=int, default=1, help="also run the launch autotuner")
    ap.add_argument("--beams2d", type=int, default=1,
                    help="also run the 2D 256^2 CPU-runnable case")
    ap.add_argument("--parity", type=int, default=1,
                    help="also count the one-step differences from the oracle on a small box")
    ap.add_argument("--functor", type=int, default=1,
                    help="also time the kernel-level push and deposit entry points")
    return ap.parse_args()


def init_fields(eng, seed=12345, nmodes=16, kmax=4, dB=0.5, B0=1.0):
    """B = B0 z + curl A with A a sum of random-phase modes (discrete curl on
    the Yee grid, so div B = 0 to round-off); E = 0."""
    import numpy as np
    import torch
    from entity_amd.kernels import communicate_fields, mink

    shape = eng.shape
    n = shape.n
    L = [float(x) for x in n]
    G = shape.ngh
    rng = np.random.default_rng(seed)
    modes = []
    while len(modes) < nmodes:
        m = rng.integers(-kmax, kmax + 1, 3)
        if 0 < np.sqrt((m ** 2).sum()) <= kmax:
            modes.append(m)
    em = eng.fields.em
    em.zero_()
    dev = em.device
    axes = [torch.arange(e, device=dev, dtype=torch.float32) - G for e in shape.extents]
    shifts = {0: (0.5, 0, 0), 1: (0, 0.5, 0), 2: (0, 0, 0.5)}
    for c in range(3):
        z, y, x = torch.meshgrid(axes[2] + shifts[c][2], axes[1] + shifts[c][1],
                                 axes[0] + shifts[c][0], indexing="ij")
        acc = torch.zeros_like(x)
        for m in modes:
            k = [2 * math.pi * m[d] / L[d] for d in range(3)]
            amp = rng.standard_normal(3)
            kk = np.array(k)
            amp -= kk * (amp @ kk) / (kk @ kk)   # transverse polarization
            ph = rng.uniform(0, 2 * math.pi)
            acc += float(amp[c] / max(np.linalg.norm(kk), 1e-9)) * torch.cos(
                k[0] * x + k[1] * y + k[2] * z + ph)
        em[c].copy_(acc)
        del x, y, z, acc
    communicate_fields(em, shape, 0, 3)
    mink.Faraday_kernel(em, shape, -1.0, 0.0)()   # B = curl A
    em[:3].zero_()
    act = (slice(None),) + tuple(slice(G, G + s) for s in reversed(n))
    B = em[3:6][act]
    rms = torch.sqrt((B.double() ** 2).sum(0).mean()).item()
    em[3:6] *= dB * B0 / rms
    em[5] += B0
    communicate_fields(em, shape, 0, 6)
    torch.cuda.synchronize()


# the 3D step's push and deposit of one species (shape orders 1, 2, Boris or
# Vay, sorted species): two launches, k_push_window then k_deposit_window,
# timed together (HIP events around the species' launches)
SPLIT_KERNELS = "k_push_window + k_deposit_window"
TRAFFIC_PROFILE = "k_push_window_k_deposit_window_3d_o2_vay_512x256x256_traffic.md"
COUNTERS_PROFILE = "k_push_window_k_deposit_window_3d_o2_vay_128x64x64_counters.md"


def measured_traffic(cells, ppc, tile):
    """HBM bytes of one species' push and deposit launches from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (None for
    another workload or tile)."""
    path = ROOT / "profiles" / TRAFFIC_PROFILE
    want = (f"workload: 3D periodic {'x'.join(map(str, cells))} cells, e-/e+ {ppc} ppc per "
            f"species, {'x'.join(map(str, tile))} tiles")
    try:
        text = path.read_text()
    except OSError:
        return None
    if want not in text:
        return None
    for line in text.splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def measured_valu_per_64(order, pusher):
    """VALU wavefront instructions per 64 particles of one species' push and
    deposit launches in 3D, from the committed rocprofv3 SQ_INSTS_VALU
    summary (order 2, Vay)."""
    if order != 2 or pusher != "vay":
        return None
    path = ROOT / "profiles" / COUNTERS_PROFILE
    try:
        for line in path.read_text().splitlines():
            if line.startswith("valu_per_64_particles:"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


# CU count and peak engine clock of the MI355X: one 64-wide FP32 VALU
# wavefront instruction per CU-cycle at most (four 16-lane SIMDs)
MI355X_CUS, MI355X_CLOCK_HZ = 256, 2.4e9


def parity_leg(args):
    """The 3D benchmark step's physics (order 2, Vay, 2 filter passes, B0 z plus the
    random-phase modes, theta = 0.3) on a 32 x 16 x 16 box, one ntt_sr_step
    against one step of the oracle from the same state: the particles whose
    cell index or tag differs (all within 4 float ulps of a cell face, which
    is required), the largest deviations of the others, and of E, B and J."""
    import numpy as np
    import torch
    from entity_amd import _abi
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    from entity_amd.kernels import device_errors
    sys.path.insert(0, str(ROOT / "oracle"))
    import ntt_oracle
    n, ppc = (32, 16, 16), args.ppc
    pusher = _abi.PUSHER_VAY if args.pusher == "vay" else _abi.PUSHER_BORIS
    prm = SRParams(dim=3, n=n, cfl=0.5, shape_order=args.order, n_filters=args.filters,
                   ppc0=2.0 * ppc, larmor0=8.0, skindepth0=8.0, sort_interval=4, tile=(8, 4, 4))
    eng = SRPICEngine(prm, [SpeciesSpec("e-", 1.0, -1.0, pusher),
                            SpeciesSpec("e+", 1.0, 1.0, pusher)], int(np.prod(n)) * ppc)
    eng.inject_uniform(ppc, temperature=0.3, seed=2024, position_seed=99)
    init_fields(eng)
    parts0 = [sp.to_host() for sp in eng.species]
    em = eng.fields.em.cpu().numpy().copy()
    cur = eng.fields.cur.cpu().numpy().copy()
    eng.step_forward()
    torch.cuda.synchronize()
    assert device_errors() == 0
    p = eng.params
    sp = [{"p": ntt_oracle.normalize_particles(h, "f32"), "mass": s.mass, "charge": s.charge,
           "pusher": s.pusher} for s, h in zip(eng.species, parts0)]
    step = dict(dt=p.dt, correction=p.correction, omegaB0=p.omegaB0, q0=p.q0, B0=p.B0, V0=p.V0,
                ppc0=p.ppc0, dx=p.dx, shape_order=p.shape_order, n_filters=p.n_filters,
                fieldsolver_enable=1, deposit_enable=1, step=0)
    assert ntt_oracle.sr_step(em, cur, np.zeros_like(cur), sp, dim=3, prm=step,
                              bc=((1, 1),) * 3) == 0
    eps = np.float32(4) * np.finfo(np.float32).eps
    mism = off_face = 0
    dev = {"dx": 0.0, "u": 0.0}
    for s, ref in zip(eng.species, sp):
        got, q = s.to_host(), ref["p"]
        near = np.zeros(len(q["tag"]), bool)
        for d in (1, 2, 3):
            for x in (got[f"dx{d}"], q[f"dx{d}"]):
                near |= (x < eps) | (x > np.float32(1) - eps)
        diff = np.zeros_like(near)
        for k in ("i1", "i2", "i3", "tag"):
            diff |= got[k] != q[k]
        mism += int(diff.sum())
        off_face += int((diff & ~near).sum())
        ok = ~diff
        for k in ("dx1", "dx2", "dx3"):
            dev["dx"] = max(dev["dx"], float(np.abs(got[k][ok] - q[k][ok]).max()))
        for k in ("ux1", "ux2", "ux3"):
            dev["u"] = max(dev["u"], float((np.abs(got[k][ok] - q[k][ok]) /
                                            np.maximum(np.abs(q[k][ok]), 1e-3)).max()))
    rel = {}
    for name, ref in (("em", em), ("cur", cur)):
        got = getattr(eng.fields, name).cpu().numpy()
        rel[name] = float(np.abs(got - ref).max() / np.abs(ref).max())
    return {"workload": f"3D periodic 32x16x16, e-/e+ {ppc} ppc, the benchmark's physics, one "
                        "step against the oracle's step (product FP32 build)",
            "particles": 2 * int(np.prod(n)) * ppc,
            "cell_or_tag_mismatches_near_face": mism, "cell_or_tag_mismatches_off_face": off_face,
            "max_abs_dx_deviation": dev["dx"], "max_rel_u_deviation": dev["u"],
            "max_rel_em_deviation": rel["em"], "max_rel_current_deviation": rel["cur"]}


def functor_leg(eng, frac=16):
    """The kernel-level entry points a driver calls one by one
    (ntt_sr_push, then ntt_deposit_currents into the current array), timed
    with HIP events on the first 1/frac of species 0 of the benchmark's box
    after the timed steps: the functor-level integration path."""
    import ctypes as C
    import torch
    from entity_amd import _abi
    from entity_amd.kernels import Minkowski, PusherBoundaries, PusherContext
    sp = eng.species[0]
    p = eng.params
    abi = sp.abi()
    abi.npart = sp.npart // frac
    n = list(p.n) + [1] * (3 - p.dim)
    ctx = PusherContext(sp.pusher, sp.mass, sp.charge, p.dt, p.omegaB0, *n).abi()
    bc = PusherBoundaries().abi()
    em = eng.fields.abi("em")
    cur = eng.fields.abi("cur")
    met = Minkowski(p.dim, p.dx).abi()
    st = torch.cuda.current_stream().cuda_stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    eng.fields.cur.zero_()
    torch.cuda.synchronize()
    ev[0].record()
    _abi.check(eng.lib.ntt_sr_push(C.byref(ctx), C.byref(bc), C.byref(abi), C.byref(em),
                                   C.byref(met), p.shape_order, st), "ntt_sr_push")
    ev[1].record()
    _abi.check(eng.lib.ntt_deposit_currents(C.byref(abi), C.byref(cur), C.byref(met), sp.charge,
                                            p.dt, p.shape_order, st), "ntt_deposit_currents")
    ev[2].record()
    torch.cuda.synchronize()
    t_push, t_dep = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    npart = int(abi.npart)
    return {"workload": f"ntt_sr_push + ntt_deposit_currents (fields "
                        f"from HBM, J per workgroup in LDS, then atomics into cur) on {npart} particles (1/{frac} of "
                        f"one species of the benchmark's box)",
            "particles": npart, "push_ms": t_push, "deposit_ms": t_dep,
            "value": npart / ((t_push + t_dep) * 1e-3), "unit": "particle-steps/s (push + deposit)"}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, cells, order, pusher_flag, nfilt):
    """The oracle's C++ restatement (FP32) on a sub-box of the workload, its
    step's loops on OpenMP threads (OMP_NUM_THREADS of the environment) with
    per-thread current copies."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import ntt_oracle
    dim = 3
    n = cells
    ppc = args.ppc
    rng = np.random.default_rng(5)
    ncell = int(np.prod(n))
    npart = ncell * ppc
    em = np.zeros((6,) + tuple(x + 4 for x in reversed(n)), np.float32)
    em[5] = 1.0
    em[3:5] += (rng.standard_normal(em[3:5].shape) * 0.3).astype(np.float32)
    cur = np.zeros((3,) + em.shape[1:], np.float32)
    buff = np.zeros_like(cur)
    species = []
    for s, q in enumerate((-1.0, 1.0)):
        cell = np.repeat(np.arange(ncell), ppc)
        p = {"i1": (cell % n[0]).astype(np.int32), "i2": ((cell // n[0]) % n[1]).astype(np.int32),
             "i3": (cell // (n[0] * n[1])).astype(np.int32)}
        for d in range(3):
            p[f"dx{d+1}"] = rng.random(npart, dtype=np.float32)
            p[f"i{d+1}_prev"] = p[f"i{d+1}"].copy()
            p[f"dx{d+1}_prev"] = p[f"dx{d+1}"].copy()
            p[f"ux{d+1}"] = (rng.standard_normal(npart) * math.sqrt(0.3)).astype(np.float32)
        p["weight"] = np.ones(npart, np.float32)
        p["tag"] = np.ones(npart, np.int16)
        species.append({"p": ntt_oracle.normalize_particles(p, "f32"), "mass": 1.0,
                        "charge": q, "pusher": pusher_flag})
    dt = float(np.float32(0.5) * (np.float32(1.0) / np.sqrt(np.float32(3))))
    prm = dict(dt=dt, correction=1.0, omegaB0=1.0 / 8, q0=1.0 / (2 * ppc * 64.0), B0=1.0 / 8,
               V0=1.0, ppc0=2.0 * ppc, dx=1.0, shape_order=order, n_filters=nfilt,
               fieldsolver_enable=1, deposit_enable=1)
    threads = ntt_oracle.omp_threads()
    t0 = time.perf_counter()
    for k in range(args.cpu_steps):
        prm["step"] = k
        rc = ntt_oracle.sr_step(em, cur, buff, species, dim=3, prm=prm, bc=((1, 1),) * 3,
                                threads=True)
        assert rc == 0, rc
    el = time.perf_counter() - t0
    return {"value": 2 * npart * args.cpu_steps / el, "unit": "particle-steps/s",
            "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{args.cpu_steps} full steps of a {n[0]}x{n[1]}x{n[2]} sub-box of the same "
                      f"workload ({2 * npart} particles, order {order}, {nfilt} filter passes), "
                      f"C++ restatement in FP32 on {threads} OpenMP threads (per-thread current "
                      f"copies), {el:.1f} s"}


def ensemble_specs(sizes=(64, 128, 256, 512), nboxes=64, seed=777):
    """The seeded parameter scan of the batched workload: per box a size n
    (n x n cells) from `sizes`, ppc per species in [8, 32], theta log-uniform
    in [0.01, 1], counter-drift beta uniform in [0, 0.9] (drift 4-velocity
    u_d = beta gamma), shape order 1 or 2, and the position/velocity seed."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for b in range(nboxes):
        n = int(rng.choice(list(sizes)))
        ppc = int(rng.integers(8, 33))
        theta = float(10 ** rng.uniform(-2, 0))
        beta = float(rng.uniform(0, 0.9))
        out.append({"n": n, "ppc": ppc, "theta": theta, "ud": beta / np.sqrt(1 - beta * beta),
                    "order": int(rng.choice([1, 2])), "seed": 1000 + b})
    return out


def ensemble_box(spec, **engine_kw):
    """One box of the scan as an SRPICEngine at its initial state: e-/e+ at
    +-u_d along x, Boris, 1 filter pass, re-sort every 4 steps, the engine's
    default tile for the box size (16 x 8 from 256^2 up, 8 x 8 below)."""
    from entity_amd import _abi
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    n, ppc, ud = spec["n"], spec["ppc"], spec["ud"]
    prm = SRParams(dim=2, n=(n, n), shape_order=spec["order"], n_filters=1, ppc0=2.0 * ppc,
                   larmor0=4.0, skindepth0=4.0, sort_interval=4)
    eng = SRPICEngine(prm, [SpeciesSpec("e-", 1.0, -1.0, _abi.PUSHER_BORIS),
                            SpeciesSpec("e+", 1.0, 1.0, _abi.PUSHER_BORIS)], n * n * ppc,
                      **engine_kw)
    eng.inject_uniform(ppc, temperature=spec["theta"], drifts=[(ud, 0, 0), (-ud, 0, 0)],
                       seed=spec["seed"])
    return eng


def batched_leg(args, sizes=(64, 128, 256, 512), label=None):
    """Ensemble of 64 independent 2D periodic boxes (parameter scan): sizes
    from {64^2, 128^2, 256^2, 512^2}, 8-32 ppc per species, theta log-uniform
    in [0.01, 1], counter-drift beta in [0, 0.9], shape order 1 or 2, Boris,
    1 filter pass, re-sort every 4 steps.  Times the plain per-box loop
    (SRPICEngine.step_forward box after box) and the batched step replayed
    from a HIP graph, on the same boxes."""
    import torch
    from entity_amd.ensemble import SREnsemble
    specs = ensemble_specs(sizes)
    engines = [ensemble_box(s) for s in specs]
    total = sum(2 * s["n"] ** 2 * s["ppc"] for s in specs)
    for eng in engines:
        eng.step_forward()
    torch.cuda.synchronize()
    K = args.batched_steps
    t0 = time.perf_counter()
    for _ in range(K):
        for eng in engines:
            eng.step_forward()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / K
    ens = SREnsemble(engines)
    ens.step_forward()
    ens.capture()
    ens.step_graph()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        ens.step_graph()
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t0) / K
    sz = "..".join(f"{x}^2" for x in (min(sizes), max(sizes))) if len(sizes) > 1 \
        else f"{sizes[0]}^2"
    return {"workload": label or (f"64 independent 2D periodic boxes ({sz}, 8-32 ppc/species, "
                                  "theta 0.01-1, drift beta 0-0.9, order 1/2, Boris, 1 filter "
                                  "pass, re-sort every 4 steps)"),
            "particles": total, "steps": K,
            "value": total / t_batched, "unit": "particle-steps/s",
            "plain_loop_value": total / t_plain,
            "ms_per_step": 1e3 * t_batched, "plain_loop_ms_per_step": 1e3 * t_plain,
            "speedup_vs_plain_loop": t_plain / t_batched}


def _qks(N, rmin, rmax, a):
    """Quasi-spherical Kerr-Schild grid with h = 0, r0 = 0 (r = e^chi, chi uniform
    in x1; theta uniform in x2), in float64 numpy: functions of code coordinates
    (x1, x2) returning r, theta, dr/dx1, dth/dx2, h_11, h_22, h_33, h_13,
    beta^1 (kerr_schild.h:143-330 and qkerr_schild.h:84-93 as formulas)."""
    import numpy as np
    chi0, chi1 = np.log(rmin), np.log(rmax)
    dchi, dth = (chi1 - chi0) / N[0], np.pi / N[1]

    def m(x1, x2):
        r = np.exp(chi0 + x1 * dchi)
        th = x2 * dth
        drdx = r * dchi
        S = r * r + (a * np.cos(th)) ** 2
        z = 2 * r / S
        D = r * r - 2 * r + a * a
        A = (r * r + a * a) ** 2 - (a * np.sin(th)) ** 2 * D
        s2 = np.sin(th) ** 2
        return dict(r=r, th=th, drdx=drdx, dthdx=dth, h11=(1 + z) * drdx ** 2, h22=S * dth ** 2,
                    h33=A * s2 / S, h13=-a * (1 + z) * s2 * drdx, beta1=z / (1 + z) / drdx)
    return m


def wald_vertical_b(N, rmin, rmax, a, B0=1.0):
    """Contravariant code components of Wald's uniform field aligned with the
    spin (pgens/wald/pgen.hpp:41-110): A_phi = B0/2 (h_33 + 2 a h_13 beta^1),
    A_r = B0/2 (h_13 + 2 a h_11 beta^1), B^i = curl A / sqrt(det h), each on
    its staggered position with ghost cells (2), as float32 arrays [3][N2+4][N1+4]."""
    import numpy as np
    m = _qks(N, rmin, rmax, a)
    G = 2
    i = np.arange(N[0] + 2 * G, dtype=np.float64)[None, :] - G
    j = np.arange(N[1] + 2 * G, dtype=np.float64)[:, None] - G

    def A3(x1, x2):
        q = m(x1, x2)
        return 0.5 * B0 * (q["h33"] + 2 * a * q["h13"] * q["beta1"])

    def A1(x1, x2):
        q = m(x1, x2)
        return 0.5 * B0 * (q["h13"] + 2 * a * q["h11"] * q["beta1"])

    def sqrt_det(x1, x2):
        q = m(x1, x2)
        return np.sqrt(q["h22"] * (q["h11"] * q["h33"] - q["h13"] ** 2))

    with np.errstate(divide="ignore", invalid="ignore"):
        b1 = (A3(i, j + 1.0) - A3(i, j)) / sqrt_det(i, j + 0.5)            # at (i, j + 1/2)
        b2 = -(A3(i + 1.0, j) - A3(i, j)) / sqrt_det(i + 0.5, j)          # at (i + 1/2, j)
        b3 = -(A1(i + 0.5, j + 1.0) - A1(i + 0.5, j)) / sqrt_det(i + 0.5, j + 0.5)
    b2 = np.where(np.isfinite(b2), b2, 0.0)  # on the axis rows
    b1 = np.where(np.isfinite(b1), b1, 0.0)
    b3 = np.where(np.isfinite(b3), b3, 0.0)
    return np.stack([b1 + 0 * i, b2 + 0 * j, b3]).astype(np.float32), m


def gr_run(N, a, *, ppc=8, rmax=50.0, theta=0.1, seed=4242, clear_interval=100, B0=1.0,
           cfl=0.5):
    """One GR test-particle run of the batched GR workload: a quasi-spherical
    Kerr-Schild grid of N cells (h = 0, r0 = 0: log-r in [0.9 r_h, rmax],
    theta in [0, pi]) with spin a, Wald's uniform field aligned with the spin
    (B0, no electric field), e-/e+ at ppc per species with thermal momenta of
    temperature theta in the local frame (covariant code components scaled by
    sqrt(h_ii)), uniform positions, CFL `cfl` on the smallest cell, FP64 metric
    and pusher on FP32 storage, first order shape, 10 iterations, absorbing r
    boundaries, axis reflection, dead particles removed every clear_interval
    steps.  Returns the GRRun at step 0."""
    import numpy as np
    import torch
    from entity_amd import _abi
    from entity_amd.grpic import GRRun, GRSpeciesSpec
    from entity_amd.kernels import gr
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rh = 1 + np.sqrt(1 - a * a)
    met = gr.KerrSchild(_abi.METRIC_QKERR_SCHILD, N, (0.9 * rh, rmax), (0.0, np.pi), a=float(a))
    bw, mfun = wald_vertical_b(N, 0.9 * rh, rmax, float(a), B0)
    ic = np.arange(N[0]) + 0.5
    jc = np.arange(N[1]) + 0.5
    q = mfun(ic[None, :], jc[:, None])
    dt = float(cfl * min(q["drdx"].min(), (q["r"] * q["dthdx"]).min()) / np.sqrt(2))
    npart = N[0] * N[1] * ppc
    run = GRRun(met, [GRSpeciesSpec("e-", 1.0, -1.0), GRSpeciesSpec("e+", 1.0, 1.0)], npart,
                dt=dt, clear_interval=clear_interval)
    run.em[3:6] = torch.from_numpy(bw).to(dev)
    sc = [torch.from_numpy(np.sqrt(q[k]).astype(np.float32).ravel()).to(dev)
          for k in ("h11", "h22", "h33")]
    sig = float(np.sqrt(theta))      # non-relativistic thermal spread
    cell = torch.arange(npart, device=dev) // ppc
    for P in run.species:
        P.load(i1=(cell % N[0]).to(torch.int32), i2=(cell // N[0]).to(torch.int32),
               dx1=torch.rand(npart, device=dev, generator=gen),
               dx2=torch.rand(npart, device=dev, generator=gen),
               ux1=torch.randn(npart, device=dev, generator=gen) * sig * sc[0][cell],
               ux2=torch.randn(npart, device=dev, generator=gen) * sig * sc[1][cell],
               ux3=torch.randn(npart, device=dev, generator=gen) * sig * sc[2][cell],
               weight=torch.ones(npart, device=dev),
               tag=torch.ones(npart, dtype=torch.int16, device=dev))
    return run


# FP64 flop per particle-step of the GR push on the batched workload, from the
# SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 counters (FMA counted twice)
GR_FP64_FLOP_PER_PARTICLE = 3506.0
FP64_VECTOR_PEAK_TFLOPS = 78.6


def gr_leg(args):
    """GR pusher-only runs, batched: 16 runs on 768 x 768 quasi-spherical
    Kerr-Schild grids (r in [0.9 r_h, 50], theta in [0, pi], h = 0), spins
    linspace(0, 0.99, 16), Wald's uniform magnetic field aligned with the spin
    (B0 = 1, no electric field), e-/e+ at 8 ppc per species with theta = 0.1
    thermal momenta in the local frame, FP32 storage with FP64 metric and
    pusher, first order shape, 10 iterations, absorbing r boundaries and axis
    reflection, dead particles removed every 100 steps.  Step 0 (em0 <- em and
    the first push) runs eagerly; steps 1..gr_steps-1 replay one HIP graph
    holding the single batched launch over the 32 (run, species) entries."""
    import numpy as np
    import torch
    from entity_amd.grpic import GRBatch
    N = (768, 768)
    runs = [gr_run(N, float(a), seed=4242 + k) for k, a in enumerate(np.linspace(0.0, 0.99, 16))]
    total0 = sum(sp.npart for r in runs for sp in r.species)
    batch = GRBatch(runs)
    batch.step_forward()
    batch.capture()
    torch.cuda.synchronize()
    K = args.gr_steps - 1
    pushed = 0
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s0.record()
    for _ in range(K):
        pushed += sum(sp.npart for r in runs for sp in r.species)
        batch.step_graph()
    s1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = s0.elapsed_time(s1) / K
    rate = pushed / (ms * 1e-3 * K)
    alive = sum(sp.npart for r in runs for sp in r.species)
    tflops = rate * GR_FP64_FLOP_PER_PARTICLE / 1e12
    return {"workload": "16 QKS runs 768x768 (r in [0.9 r_h, 50]), spins 0..0.99, Wald field "
                        "B0 = 1, e-/e+ 8 ppc/species theta 0.1 (151M particles), CFL 0.5, "
                        "FP32 storage / FP64 metric, 10 iterations, absorbing r / axis theta, "
                        "dead removed every 100 steps; one batched launch per step (graph)",
            "particles": total0, "particles_at_end": alive, "steps": K + 1, "timed_steps": K,
            "value": rate, "unit": "particle-steps/s (particles held at each step)",
            "ms_per_step": ms, "wall_s": wall,
            "hbm_gbs": rate * 78 / 1e9,
            "roofline": {"bound": "valu_fp64", "achieved": tflops,
                         "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tflops / FP64_VECTOR_PEAK_TFLOPS,
                         "flop_per_particle_step": GR_FP64_FLOP_PER_PARTICLE,
                         "flop_source": "profiles/k_gr_push_qks_768x768_counters.md"}}


def gr_full_leg(args):
    """The full GRPIC step (field solver, push, deposit, one filter pass,
    boundaries) on one run of the GR workload: 768 x 768 QKS grid, a = 0.9,
    Wald's field, e-/e+ at 8 ppc per species, horizon fields / particles at
    r_min, match layer (ds = 5) and absorbing current layer at r_max = 50,
    axis rows, q0 = 1e-3; the engine's launches run eagerly (the step swaps
    em / em0 and cur / cur0)."""
    import torch
    from entity_amd import _abi
    from entity_amd.grpic import GRPICEngine, GRSpeciesSpec
    N, a = (768, 768), 0.9
    run = gr_run(N, a, seed=4343, cfl=0.5)
    flds = ((_abi.FLDS_BC_HORIZON, _abi.FLDS_BC_MATCH), (_abi.FLDS_BC_AXIS, _abi.FLDS_BC_AXIS))
    prtl = ((_abi.PRTL_BC_HORIZON, _abi.PRTL_BC_ABSORB), (_abi.PRTL_BC_AXIS, _abi.PRTL_BC_AXIS))
    specs = [GRSpeciesSpec(sp.label, sp.mass, sp.charge) for sp in run.species]
    eng = GRPICEngine(run.metric, specs, 1, dt=run.dt, flds_bc=flds, prtl_bc=prtl,
                      ds=((0.5, 5.0),), target=run.em.clone(), nfilter=1, q0=1e-3, B0=1.0)
    eng.species = run.species      # the run's particles, in place
    eng.em.copy_(run.em)
    del run
    npart = sum(sp.npart for sp in eng.species)
    for _ in range(2):
        eng.step_forward()
    torch.cuda.synchronize()
    K = args.gr_full_steps
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    pushed = 0
    s0.record()
    for _ in range(K):
        pushed += sum(sp.npart for sp in eng.species)
        eng.step_forward()
    s1.record()
    torch.cuda.synchronize()
    ms = s0.elapsed_time(s1) / K
    return {"workload": "full GRPIC step, QKS 768x768 (r in [0.9 r_h, 50]), a = 0.9, Wald field, "
                        "e-/e+ 8 ppc/species (9.4M particles), deposit + 1 filter pass, "
                        "horizon / match (ds 5) / axis boundaries, q0 = 1e-3, FP64 metric",
            "particles": npart, "timed_steps": K, "ms_per_step": ms,
            "value": pushed / (ms * 1e-3 * K), "unit": "particle-steps/s",
            "cell_updates_per_s": N[0] * N[1] / (ms * 1e-3)}


# the adjoint workload: 128 x 128 cells, 4 species x 16 ppc, 40 steps,
# checkpoint every 8, drift +-u_d on the species with ADJ_SIGN, uniform B_z = b
ADJ_N, ADJ_PPC, ADJ_STEPS, ADJ_CHECKPOINT = (128, 128), 16, 40, 8
ADJ_SIGN = (1.0, 1.0, -1.0, -1.0)
ADJ_UD, ADJ_B = 0.57735, 0.1


def adjoint_box(ud, b):
    """Initial state of the adjoint workload with drift u_d and B_z amplitude b:
    e-/e+ beams (species charges -1, +1, -1, +1; u_x shifted by ADJ_SIGN * u_d)
    over theta = 0.1, shape order 2, 1 filter pass, re-sort every 4 steps
    into 8 x 8 tiles, particle IDs carried for the adjoint."""
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    n, ppc = ADJ_N, ADJ_PPC
    prm = SRParams(dim=2, n=n, shape_order=2, n_filters=1, ppc0=4.0 * ppc, larmor0=1.0,
                   skindepth0=4.0, sort_interval=4, tile=(8, 8))   # 256 tiles per species
    sp = [SpeciesSpec(f"s{k}", 1.0, q) for k, q in enumerate((-1.0, 1.0, -1.0, 1.0))]
    eng = SRPICEngine(prm, sp, n[0] * n[1] * ppc, with_payload=True)
    eng.inject_uniform(ppc, temperature=0.1, seed=31)
    for sg, s in zip(ADJ_SIGN, eng.species):
        s.ux1[: s.npart] += sg * ud
    G = eng.shape.ngh
    eng.fields.em[5, G:G + n[1], G:G + n[0]] = b
    return eng


def adjoint_leg(args):
    """Adjoint of a 2D counter-streaming run: periodic 128 x 128 cells, e-/e+
    beams at u_x +- u_d (u_d = 0.57735, beta = 0.5) over theta = 0.1 thermal
    momenta, 16 ppc per species x 4 species (64 ppc), uniform B_z = b = 0.1,
    shape order 2, Boris, 1 filter pass, tile re-sort every 4 steps (particle
    IDs carry the adjoint through the re-sorts), 40 steps.  Loss
    L = sum |B|^2 / 2 at the last step; gradients dL/du_d and dL/db from the
    checkpointed reverse pass (checkpoint every 8 steps).  Reports the forward
    and the float gradient wall time (medians of three runs) and the double
    gradient's, float vs double adjoint agreement, and central
    finite differences (h = 1e-3) of the forward."""
    import torch
    from entity_amd.adjoint import SRAdjoint
    n, ppc, nsteps, K = ADJ_N, ADJ_PPC, ADJ_STEPS, ADJ_CHECKPOINT
    sign = ADJ_SIGN
    ud0, b0, h = ADJ_UD, ADJ_B, 1e-3
    setup = adjoint_box

    def grads(prec):
        eng = setup(ud0, b0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss, lam, lam_u = SRAdjoint(eng, nsteps, K, prec).gradient()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        g_ud = sum(sg * float(lu[0].double().sum()) for sg, lu in zip(sign, lam_u))
        return loss, g_ud, float(lam[5].double().sum()), t

    def loss_of(ud, b):
        eng = setup(ud, b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = SRAdjoint(eng, nsteps, K).forward()
        torch.cuda.synchronize()
        return L, time.perf_counter() - t0

    loss_of(ud0, b0)            # warm-up
    # medians of three timed runs each (single runs vary by a few %)
    fwd = [loss_of(ud0, b0) for _ in range(3)]
    L0, t_fwd = fwd[0][0], sorted(t for _, t in fwd)[1]
    grads(32)                   # warm-up
    g32 = [grads(32) for _ in range(3)]
    L32, gu32, gb32 = g32[0][:3]
    t32 = sorted(g[3] for g in g32)[1]
    L64, gu64, gb64, t64 = grads(64)
    fd_u = (loss_of(ud0 + h, b0)[0] - loss_of(ud0 - h, b0)[0]) / (2 * h)
    fd_b = (loss_of(ud0, b0 + h)[0] - loss_of(ud0, b0 - h)[0]) / (2 * h)
    # L(+h) - L(-h) at h = 1e-3 is about 1e-6 of L, a few float roundings of
    # the FP32 forward, so those quotients carry ~1e-3 of rounding noise; the
    # same differences at h = 1e-2 carry a tenth of it (the test suite holds
    # the 1e-3 bar against an FP64 forward's differences at h = 1e-3)
    h2 = 1e-2
    fd2_u = (loss_of(ud0 + h2, b0)[0] - loss_of(ud0 - h2, b0)[0]) / (2 * h2)
    fd2_b = (loss_of(ud0, b0 + h2)[0] - loss_of(ud0, b0 - h2)[0]) / (2 * h2)
    npart = 4 * n[0] * n[1] * ppc
    return {"workload": "2D periodic 128x128, e-/e+ beams u_x +- 0.57735 over theta 0.1, "
                        "4 species x 16 ppc (1,048,576 particles), uniform B_z = 0.1, order 2, "
                        "Boris, 1 filter pass, re-sort every 4 steps, 40 steps; L = sum B^2/2; "
                        "checkpoint every 8",
            "particles": npart, "steps": nsteps,
            "forward_ms": 1e3 * t_fwd, "forward_value": npart * nsteps / t_fwd,
            "gradient_ms_f32": 1e3 * t32, "gradient_ms_f64": 1e3 * t64,
            "gradient_over_forward_f32": t32 / t_fwd,
            "value": npart * nsteps / t32, "unit": "particle-steps/s (forward + reverse)",
            "loss": L32, "dL_dud": {"f32": gu32, "f64": gu64, "fd": fd_u, "fd_h1e-2": fd2_u},
            "dL_db": {"f32": gb32, "f64": gb64, "fd": fd_b, "fd_h1e-2": fd2_b},
            "rel_err_vs_fd_f64": {"ud": abs(gu64 - fd_u) / abs(fd_u),
                                  "b": abs(gb64 - fd_b) / abs(fd_b)},
            "rel_err_vs_fd_h1e-2_f64": {"ud": abs(gu64 - fd2_u) / abs(fd2_u),
                                        "b": abs(gb64 - fd2_b) / abs(fd2_b)},
            "fd_note": "FP32 forward; at h = 1e-3 the differences carry ~1e-3 of rounding "
                       "noise, at h = 1e-2 a tenth of it; the FP64 adjoint's 1e-3 bar against "
                       "an FP64 forward's differences at h = 1e-3 is checked in the GPU tests"}


BEAMS2D_N, BEAMS2D_PPC = (256, 256), 16
BEAMS2D_DRIFTS = [(0.57735, 0, 0), (0.57735, 0, 0), (-0.57735, 0, 0), (-0.57735, 0, 0)]
BEAMS2D_CHARGES = (-1.0, 1.0, -1.0, 1.0)


def beams2d_params(tile=None, sort_interval=4, particles_per_thread=0):
    from entity_amd.engine import SRParams
    return SRParams(dim=2, n=BEAMS2D_N, cfl=0.5, shape_order=1, n_filters=1,
                    ppc0=4.0 * BEAMS2D_PPC,
                    larmor0=8.0, skindepth0=8.0, sort_interval=sort_interval, tile=tile,
                    particles_per_thread=particles_per_thread)


def autotune_leg(args):
    """Autotuner on the 2D beam box (2D periodic 256 x 256, e-/e+ beams u_x =
    +-0.57735 over theta 0.1, 4 species x 16 ppc, order 1, Boris, 1 filter
    pass): every candidate (tile shape x re-sort cadence, and the unsorted
    layout x chunk size) runs 16 timed steps after 32 steps that age the
    layout; reports the fastest and its gain over 8x8 tiles re-sorted every 4
    steps.  The 2D beam leg then runs with the pick."""
    import torch
    from entity_amd.autotune import Candidate, candidates_for, engine_runner, tune
    from entity_amd.engine import SpeciesSpec, SRPICEngine
    n, ppc = BEAMS2D_N, BEAMS2D_PPC
    sp = [SpeciesSpec(f"s{k}", 1.0, q) for k, q in enumerate(BEAMS2D_CHARGES)]
    ref = SRPICEngine(beams2d_params(tile=(8, 8)), sp, n[0] * n[1] * ppc)
    ref.inject_uniform(ppc, temperature=0.1, drifts=BEAMS2D_DRIFTS, seed=101)
    prts = [s_.to_host() for s_ in ref.species]
    del ref

    def make(c):
        eng = SRPICEngine(beams2d_params(c.tile, c.sort_interval, c.particles_per_thread), sp,
                          n[0] * n[1] * ppc)
        for s_, h in zip(eng.species, prts):
            s_.load(**h)
        if c.sort_interval:
            eng.sort()
        return eng

    res = tune(engine_runner(make, steps=16, warmup=32), candidates_for(2))
    default = Candidate((8, 8), 4)
    t_def = res.ms_per_step.get(default)
    torch.cuda.synchronize()
    return {"workload": "the 2D beam box: 2D periodic 256x256, e-/e+ beams, 4 species x 16 ppc "
                        "(4,194,304 particles), order 1, Boris, 1 filter pass",
            "best": {"tile": list(res.best.tile), "sort_interval": res.best.sort_interval,
                     "particles_per_thread": res.best.particles_per_thread},
            "best_ms_per_step": res.ms_per_step[res.best], "default_ms_per_step": t_def,
            "gain_over_default": (t_def / res.ms_per_step[res.best]) if t_def else None,
            "candidates": len(res.ms_per_step)}


def beams2d_leg(args, tuned=None):
    """The reference's CPU-runnable case on the GPU: 2D periodic 256 x 256,
    dx = 1, CFL 0.5, e-/e+ beams at u_x = +-0.57735 (beta 0.5) over
    theta = 0.1, 16 ppc per species x 4 species (4,194,304 particles), zero
    initial fields, shape order 1, Boris, 1 filter pass, 100 steps, with the
    autotuner's layout when the autotune leg ran (else the default tile,
    re-sorted every 4 steps).  Reports particle-steps/s, cell-updates/s and
    the 2D fused kernel against the HBM roofline (78 B per particle + 36 B
    per cell per launch), with the oracle timed on the host for 20 steps of
    the same box."""
    import ctypes as C
    import numpy as np
    import torch
    from entity_amd import _abi
    from entity_amd.engine import SpeciesSpec, SRPICEngine
    n, ppc, nsteps = BEAMS2D_N, BEAMS2D_PPC, 100
    drifts = BEAMS2D_DRIFTS
    if tuned is not None:
        b = tuned["best"]
        prm = beams2d_params(tuple(b["tile"]) if b["tile"] else None, b["sort_interval"],
                        b["particles_per_thread"])
        layout = dict(b, source="autotune leg")
    else:
        prm = beams2d_params()      # default tile: 16 x 8 at 256^2, re-sort every 4
        layout = {"tile": "default", "sort_interval": 4, "particles_per_thread": 1,
                  "source": "defaults"}
    sp = [SpeciesSpec(f"s{k}", 1.0, q) for k, q in enumerate(BEAMS2D_CHARGES)]
    eng = SRPICEngine(prm, sp, n[0] * n[1] * ppc)
    eng.inject_uniform(ppc, temperature=0.1, drifts=drifts, seed=101)
    for _ in range(4):
        eng.step_forward()
    torch.cuda.synchronize()
    lib = eng.lib
    lib.ntt_profile_read(None, None, 1)
    lib.ntt_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        eng.step_forward()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.ntt_profile_enable(0)
    ms, nl = C.c_double(0), C.c_int64(0)
    _abi.check(lib.ntt_profile_read(C.byref(ms), C.byref(nl), 1), "ntt_profile_read")
    npart = 4 * n[0] * n[1] * ppc
    per = n[0] * n[1] * ppc
    avg = ms.value / max(nl.value, 1)
    achieved = (per * 78 + n[0] * n[1] * 36) / (avg * 1e-3) / 1e9
    out = {"workload": "2D periodic 256x256, e-/e+ beams u_x +-0.57735 over theta 0.1, "
                       "4 species x 16 ppc (4,194,304 particles), order 1, Boris, 1 filter pass, "
                       "100 steps",
           "layout": layout,
           "particles": npart, "steps": nsteps, "value": npart * nsteps / el,
           "unit": "particle-steps/s", "ms_per_step": 1e3 * el / nsteps,
           "cell_updates_per_s": n[0] * n[1] * nsteps / el,
           "fused_kernel": {"avg_launch_ms": avg, "achieved_gbs": achieved,
                            "frac": achieved / HBM_PEAK_GBS}}
    if args.cpu_baseline:
        out["cpu_baseline"] = beams2d_cpu_baseline(eng, n, ppc, drifts)
    return out


def beams2d_cpu_baseline(eng, n, ppc, drifts):
    """The oracle's C++ restatement (FP32, on OpenMP threads) on 20 steps of
    the same box, from the GPU engine's state copied to the host."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import ntt_oracle
    p = eng.params
    em = eng.fields.em.cpu().numpy().astype(np.float32)
    cur = np.zeros((3,) + em.shape[1:], np.float32)
    buff = np.zeros_like(cur)
    species = [{"p": ntt_oracle.normalize_particles(sp.to_host(), "f32"), "mass": sp.mass,
                "charge": sp.charge, "pusher": sp.pusher} for sp in eng.species]
    prm = dict(dt=p.dt, correction=p.correction, omegaB0=p.omegaB0, q0=p.q0, B0=p.B0, V0=p.V0,
               ppc0=p.ppc0, dx=p.dx, shape_order=p.shape_order, n_filters=p.n_filters,
               fieldsolver_enable=1, deposit_enable=1)
    nst = 20
    threads = ntt_oracle.omp_threads()
    t0 = time.perf_counter()
    for k in range(nst):
        prm["step"] = eng.step + k
        rc = ntt_oracle.sr_step(em, cur, buff, species, dim=2, prm=prm,
                                bc=((1, 1), (1, 1), (0, 0)), threads=True)
        assert rc == 0, rc
    el = time.perf_counter() - t0
    npart = sum(int(sp["p"]["_n"]) for sp in species)
    return {"value": npart * nst / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{nst} full steps of the same 256x256 box ({npart} particles), C++ "
                      f"restatement in FP32 on {threads} OpenMP threads, {el:.1f} s"}

def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                          "(multi-GPU decomposition is a design note, not a measured path)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from entity_amd import _abi
    from entity_amd.engine import SRParams, SpeciesSpec, SRPICEngine
    from entity_amd.kernels import device_errors

    cells = tuple(int(x) for x in args.cells.split(","))
    tile = tuple(int(x) for x in args.tile.split(","))
    dim = len(cells)
    pusher = _abi.PUSHER_VAY if args.pusher == "vay" else _abi.PUSHER_BORIS
    prm = SRParams(dim=dim, n=cells, cfl=0.5, shape_order=args.order, n_filters=args.filters,
                   ppc0=2.0 * args.ppc, larmor0=8.0, skindepth0=8.0,
                   sort_interval=args.sort_interval, tile=tile)
    ncell = int(np.prod(cells))
    npart = ncell * args.ppc
    t_init = time.perf_counter()
    eng = SRPICEngine(prm, [SpeciesSpec("e-", 1.0, -1.0, pusher),
                            SpeciesSpec("e+", 1.0, 1.0, pusher)], npart,
                      double_buffer=bool(args.double_buffer))
    eng.inject_uniform(args.ppc, temperature=0.3, seed=2024, position_seed=99)
    init_fields(eng)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t_init
    print(f"# init: {2 * npart} particles on {cells} in {t_init:.1f} s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        eng.step_forward()
    torch.cuda.synchronize()
    assert device_errors() == 0
    lib = eng.lib
    lib.ntt_profile_read(None, None, 1)
    lib.ntt_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        eng.step_forward()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.ntt_profile_enable(0)
    import ctypes as C
    ms_pd, nl = C.c_double(0), C.c_int64(0)
    _abi.check(lib.ntt_profile_read(C.byref(ms_pd), C.byref(nl), 1), "ntt_profile_read")
    errs = device_errors()
    ms_step = 1000.0 * el / args.steps
    total_particles = 2 * npart
    value = total_particles * args.steps / el
    # algorithmic bytes of one species' push and deposit: per particle 102 B
    # (3D) / 78 B (2D), per cell 36 B (E/B read + J write)
    b_p = 102 if dim == 3 else 78
    bytes_launch = npart * b_p + ncell * 36
    avg_ms = ms_pd.value / max(nl.value, 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    v64 = measured_valu_per_64(args.order, args.pusher) if dim == 3 else None
    valu_frac = (npart / 64 * v64 / (MI355X_CUS * MI355X_CLOCK_HZ * avg_ms * 1e-3)
                 if v64 else None)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{dim}D periodic {'x'.join(map(str, cells))} cells, e-/e+ "
                        f"{args.ppc} ppc per species ({total_particles} particles), theta=0.3, "
                        f"B0 z + 16-mode dB/B0=0.5, shape order {args.order}, {args.pusher} pusher, "
                        f"{args.filters} filter passes, CFL 0.5, re-sort every "
                        f"{args.sort_interval} steps"
                        + (" (out of place, double-buffered particles)" if args.double_buffer
                           else ""),
            "cells": list(cells), "particles": total_particles, "shape_order": args.order,
            "pusher": args.pusher, "filters": args.filters, "sort_interval": args.sort_interval,
            "tile": list(tile), "double_buffer": bool(args.double_buffer),
        },
        "cell_updates_per_s": ncell * args.steps / el,
        "roofline": {
            "bound": "hbm",
            "kernel": (SPLIT_KERNELS + " (push with E/B from LDS, then Esirkepov deposit into "
                       "LDS windows; one species per pair, timed together)")
                      if dim == 3 and args.order in (1, 2) else
                      "k_push_deposit (fused push + Esirkepov deposit, one species per launch)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(cells, args.ppc, tile) if args.order == 2 and
                       args.pusher == "vay" else None,
            "traffic_source": f"profiles/{TRAFFIC_PROFILE} (rocprofv3 FETCH_SIZE and "
                              "WRITE_SIZE of these kernels on this workload, copy-kernel "
                              "calibrated)",
            "avg_launch_ms": avg_ms,
            "launches": int(nl.value),
            "algorithmic_bytes_per_launch": bytes_launch,
            "valu_frac": valu_frac,
            "valu_source": f"profiles/{COUNTERS_PROFILE} (SQ_INSTS_VALU per 64 particles, "
                           "both launches) x particles per species / (256 CUs x 2.4 GHz x the "
                           "time of the pair measured here)",
        },
        "device_errors": errs,
    }
    if args.functor:
        out["functor_path"] = functor_leg(eng)
    del eng
    torch.cuda.empty_cache()
    if args.parity:
        out["parity"] = parity_leg(args)
    if args.batched:
        out["batched"] = batched_leg(args)
        # the same scan restricted to the smallest box size, where one box
        # cannot fill the GPU and per-launch overhead dominates
        out["batched_small_boxes"] = batched_leg(args, sizes=(64,))
    if args.gr:
        out["gr"] = gr_leg(args)
        torch.cuda.empty_cache()
    if args.gr_full:
        out["gr_full_step"] = gr_full_leg(args)
        torch.cuda.empty_cache()
    if args.adjoint:
        out["adjoint"] = adjoint_leg(args)
    if args.autotune:
        out["autotune"] = autotune_leg(args)
    if args.beams2d:
        out["beams_2d"] = beams2d_leg(args, out.get("autotune"))
    if args.cpu_baseline:
        ccells = tuple(int(x) for x in args.cpu_cells.split(","))
        out["cpu_baseline"] = cpu_baseline(args, ccells, args.order, pusher, args.filters)
        out["cpu_baseline"]["cores_available"] = os.cpu_count()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
