"""Particle and field containers: device memory held in torch tensors, handed to
libntt_hip through the C ABI structures.

Mirrors the reference containers:
* ``Particles`` -- ntt::ParticleArrays / Particles<D, C>
  (src/framework/containers/particles.h:47-71, 78-320; allocation
  particles.cpp:47-86): i1..3, i1..3_prev (int32), dx1..3, dx1..3_prev
  (float), ux1..3, weight (real), tag (int16), optional phi and an integer
  payload; active prefix [0, npart) of maxnpart.
* ``Fields`` -- ntt::Fields<D, S> (src/framework/containers/fields.h:54-82;
  fields.cpp:10-47): em (6 components), cur and buff (3 components), each with
  ``n + 2 * ngh`` extents, stored x-fastest and component-slowest.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _abi

N_GHOSTS = 2  # (SHAPE_ORDER + 1) / 2 + 1 for orders <= 2 (src/global/global.h:130-136)

_INT_NAMES = ("i1", "i2", "i3", "i1_prev", "i2_prev", "i3_prev")
_DX_NAMES = ("dx1", "dx2", "dx3", "dx1_prev", "dx2_prev", "dx3_prev")
_U_NAMES = ("ux1", "ux2", "ux3")


class Particles:
    """Structure-of-arrays particle storage of one species on one device.

    ``version`` changes whenever an attribute is assigned (an array replaced,
    npart changed), so holders of the C ABI view know when to rebuild it."""

    def __setattr__(self, name, value):
        object.__setattr__(self, "version", self.__dict__.get("version", 0) + 1)
        object.__setattr__(self, name, value)

    def __init__(self, dim: int, maxnpart: int, *, device="cuda", label: str = "",
                 mass: float = 1.0, charge: float = 1.0, pusher: int = _abi.PUSHER_BORIS,
                 with_phi: bool = False, with_payload: bool = False):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.maxnpart = int(maxnpart)
        self.npart = 0
        self.label = label
        self.mass = float(mass)
        self.charge = float(charge)
        self.pusher = int(pusher)
        self.device = torch.device(device)
        n = self.maxnpart
        dev = self.device
        for k, name in enumerate(_INT_NAMES):
            used = (k % 3) < dim
            setattr(self, name, torch.zeros(n if used else 0, dtype=torch.int32, device=dev))
        for k, name in enumerate(_DX_NAMES):
            used = (k % 3) < dim
            setattr(self, name, torch.zeros(n if used else 0, dtype=torch.float32, device=dev))
        for name in _U_NAMES:
            setattr(self, name, torch.zeros(n, dtype=torch.float32, device=dev))
        self.weight = torch.ones(n, dtype=torch.float32, device=dev)
        self.tag = torch.zeros(n, dtype=torch.int16, device=dev)
        self.phi = torch.zeros(n, dtype=torch.float32, device=dev) if with_phi else None
        self.pld_i = torch.zeros(n, dtype=torch.int64, device=dev) if with_payload else None

    # -- ABI ---------------------------------------------------------------
    def abi(self) -> _abi.NttParticles:
        s = _abi.NttParticles()
        for name in _INT_NAMES + _DX_NAMES + _U_NAMES + ("weight", "tag"):
            t = getattr(self, name)
            setattr(s, name, t.data_ptr() if t.numel() else None)
        s.phi = self.phi.data_ptr() if self.phi is not None else None
        s.pld_i = self.pld_i.data_ptr() if self.pld_i is not None else None
        s.npart = self.npart
        s.maxnpart = self.maxnpart
        return s

    def swap_arrays(self, other: "Particles") -> None:
        """Exchange every particle array with `other` (same capacity and
        optional arrays) together with the particle counts."""
        names = _INT_NAMES + _DX_NAMES + _U_NAMES + ("weight", "tag", "phi", "pld_i")
        for name in names:
            a, b = getattr(self, name), getattr(other, name)
            setattr(self, name, b)
            setattr(other, name, a)
        self.npart, other.npart = other.npart, self.npart

    # -- host transfer (tests, init) ----------------------------------------
    def arrays(self) -> dict[str, torch.Tensor]:
        names = list(_INT_NAMES[: self.dim]) + list(_INT_NAMES[3:3 + self.dim])
        names += list(_DX_NAMES[: self.dim]) + list(_DX_NAMES[3:3 + self.dim])
        names += list(_U_NAMES) + ["weight", "tag"]
        out = {n: getattr(self, n)[: self.npart] for n in names}
        if self.phi is not None:
            out["phi"] = self.phi[: self.npart]
        if self.pld_i is not None:
            out["pld_i"] = self.pld_i[: self.npart]
        return out

    def load(self, **arrays) -> "Particles":
        """Copy host/device arrays into the active prefix (sets npart)."""
        n = None
        for name, val in arrays.items():
            t = torch.as_tensor(val)
            if n is None:
                n = t.shape[0]
            elif t.shape[0] != n:
                raise ValueError(f"{name}: length {t.shape[0]} != {n}")
            if n > self.maxnpart:
                raise ValueError("npart exceeds maxnpart")
            dst = getattr(self, name)
            dst[:n].copy_(t.to(dtype=dst.dtype))
        self.npart = int(n or 0)
        return self

    def to_host(self) -> dict[str, "object"]:
        return {k: v.detach().cpu().numpy() for k, v in self.arrays().items()}


@dataclass
class GridShape:
    dim: int
    n: tuple[int, ...]
    ngh: int = N_GHOSTS

    @property
    def extents(self) -> tuple[int, ...]:
        return tuple(x + 2 * self.ngh for x in self.n)

    def tensor_shape(self, ncomp: int) -> tuple[int, ...]:
        # component slowest, x fastest: torch (C-order) shape is reversed
        return (ncomp,) + tuple(reversed(self.extents))


class Fields:
    """em, cur and buff arrays of one domain (Yee staggering as in fields.h:40-52).
    ``version`` changes whenever an array is replaced."""

    def __setattr__(self, name, value):
        object.__setattr__(self, "version", self.__dict__.get("version", 0) + 1)
        object.__setattr__(self, name, value)

    def __init__(self, shape: GridShape, *, device="cuda"):
        self.shape = shape
        dev = torch.device(device)
        self.em = torch.zeros(shape.tensor_shape(6), dtype=torch.float32, device=dev)
        self.cur = torch.zeros(shape.tensor_shape(3), dtype=torch.float32, device=dev)
        self.buff = torch.zeros(shape.tensor_shape(3), dtype=torch.float32, device=dev)

    def abi(self, name: str) -> _abi.NttField:
        return field_abi(getattr(self, name), self.shape)


def field_abi(t: torch.Tensor, shape: GridShape) -> _abi.NttField:
    if not t.is_contiguous():
        raise ValueError("field tensors must be contiguous")
    if t.dtype != torch.float32:
        raise ValueError("field tensors must be float32")
    ncomp = t.shape[0]
    if tuple(t.shape) != shape.tensor_shape(ncomp):
        raise ValueError(f"field shape {tuple(t.shape)} != {shape.tensor_shape(ncomp)}")
    f = _abi.NttField()
    f.data = t.data_ptr()
    f.dim = shape.dim
    n = list(shape.n) + [1] * (3 - shape.dim)
    f.n[0], f.n[1], f.n[2] = n
    f.ngh = shape.ngh
    f.ncomp = ncomp
    return f
