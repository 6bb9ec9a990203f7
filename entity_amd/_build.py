"""Build recipes for libntt_hip.so (gfx950).

Two builds of the same HIP sources are produced:

* ``lib/libntt_hip.so`` -- the product build.  In the SR sources each source
  expression's multiply-adds become FMAs (``-ffp-contract=on``): contraction
  is decided per expression by the front end, not per basic block by the back
  end, so an expression rounds the same in every kernel that inlines it (the
  fused, split and batched push kernels give bit-identical particles).  The GR
  sources, bound by FP64 issue, keep the back end's contraction across
  statements (``-ffp-contract=fast``: 7% fewer instructions in the pusher).
* ``lib/libntt_hip_strict.so`` -- identical sources compiled with
  ``-ffp-contract=off``.  Its single-precision arithmetic follows the
  reference's operation order exactly, which is what lets the parity tests
  demand bit-identical particle state against the CPU oracle.

Every HIP translation unit is compiled with packed-FP32 instructions disabled
(``-target-feature -packed-fp32-ops``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["common.hip", "fields.hip", "sr_kernels.hip", "push_deposit.hip",
               "sort.hip", "init.hip", "engine.hip", "gr_pusher.hip", "gr_fields.hip", "gr_bcs.hip",
               "adjoint.hip"]

ARCH = os.environ.get("NTT_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libntt_hip)")


# sources whose product build contracts across statements (see above)
FAST_CONTRACT_SOURCES = {"gr_pusher.hip", "gr_fields.hip", "gr_bcs.hip"}


def _common_flags(strict: bool, source: str = "") -> list[str]:
    contract = "off" if strict else ("fast" if source in FAST_CONTRACT_SOURCES else "on")
    return [
        "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
        f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-ffp-contract={contract}",
        "-Wall", "-Wno-unused-function", "-Wno-unused-result",
        f"-I{INCLUDE}", f"-I{CSRC}",
    ] + (["-DNTT_STRICT_FP=1"] if strict else [])


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


def build_hip(strict: bool = False, jobs: int | None = None) -> Path:
    """Compile the HIP sources into lib/libntt_hip[_strict].so and return its path."""
    hipcc = _hipcc()
    tag = "strict" if strict else "fast"
    objdir = PKG / "_build" / tag
    objdir.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()]
    objs = [objdir / (s.stem + ".o") for s in srcs]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [hipcc, *_common_flags(strict, s.name), "-c", str(s), "-o",
                                 str(o)])
                for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    out = LIBDIR / ("libntt_hip_strict.so" if strict else "libntt_hip.so")
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-Wl,-Bsymbolic",
          "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        futs = [ex.submit(build_hip, False, 4), ex.submit(build_hip, True, 4)]
        for f in futs:
            f.result()


if __name__ == "__main__":
    build_all()
    print("built:", LIBDIR / "libntt_hip.so", LIBDIR / "libntt_hip_strict.so")
