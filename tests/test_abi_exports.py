"""The C-ABI libraries load and export every symbol declared in include/ntt.h
(no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    txt = (ROOT / "include" / "ntt.h").read_text()
    return sorted(set(re.findall(r"^NTT_API\s+[\w\s\*]+?\b(ntt_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ntt_sr_push", "ntt_deposit_currents", "ntt_faraday_mink", "ntt_ampere_mink",
              "ntt_currents_ampere_mink", "ntt_digital_filter", "ntt_get_device_errors"):
        assert s in syms


@pytest.mark.parametrize("strict", [False, True])
def test_library_exports_all_declared_symbols(strict):
    import torch  # noqa: F401  (same load order as the package)
    from entity_amd import _abi
    path = _abi.library_path(strict)
    assert path.exists(), f"{path} not built"
    lib = ctypes.CDLL(str(path), mode=ctypes.RTLD_LOCAL)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    lib.ntt_abi_version.restype = ctypes.c_int32
    assert lib.ntt_abi_version() == 1


def test_python_signatures_cover_header():
    from entity_amd import _abi
    missing = [s for s in declared_symbols() if s not in _abi.SIGNATURES]
    assert not missing, missing


def test_product_path_has_no_oracle_dependency():
    # the package never imports, links or loads the oracle
    for f in (ROOT / "entity_amd").rglob("*"):
        if f.suffix in (".py", ".hip", ".hpp", ".cpp", ".h"):
            txt = f.read_text()
            assert "ntt_oracle" not in txt and "libntt_oracle" not in txt, f


def test_container_versions_track_assignments():
    """The engine rebuilds its cached ABI view of the particles when an array
    or the particle count of a container is assigned."""
    from entity_amd.containers import Fields, GridShape, Particles
    p = Particles(2, 8, device="cpu")
    v0 = p.version
    p.npart = 3
    assert p.version > v0
    v1 = p.version
    p.pld_i = None
    assert p.version > v1
    f = Fields(GridShape(2, (4, 4)), device="cpu")
    w = f.version
    f.em = f.em.clone()
    assert f.version > w


def test_empty_out_of_place_sort_sets_dst_count():
    """ntt_sort_particles_to of an empty species (no tile offsets asked for,
    so no device work) leaves dst holding the sorted, empty species."""
    import ctypes as C

    from entity_amd import _abi
    lib = _abi.load(False)
    src, dst = _abi.NttParticles(), _abi.NttParticles()
    src.npart, src.maxnpart = 0, 16
    dst.npart, dst.maxnpart = 7, 16      # stale count from an earlier use
    n = (C.c_int32 * 3)(8, 8, 1)
    tile = (C.c_int32 * 3)(4, 4, 1)
    alive = C.c_int64(-1)
    st = lib.ntt_sort_particles_to(C.byref(src), C.byref(dst), 2, n, tile, None, 0, None,
                                   C.byref(alive), None)
    assert st == _abi.NTT_OK
    assert dst.npart == 0 and alive.value == 0
