"""Periodic boxes as a property of the pair geometry (reference ``scene/periodic.py``): the stored particle positions
are never wrapped; a body sees a particle at its nearest image relative to the body centre, and a disk bundle that
spans the box carries the tiled copies of its disks (``DiskArrayRep.tiled``).  The minimum image of a displacement d
on a periodic axis of length L is d - L round(d / L) (support below L / 2)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

import torch

_LENGTHS = {}


@dataclass(frozen=True)
class Periodic:
    """an axis-aligned box [lo, hi] with ``flags`` selecting its periodic axes."""
    lo: Tuple[float, float]
    hi: Tuple[float, float]
    flags: Tuple[bool, bool] = (True, True)

    def length(self, ref: torch.Tensor) -> torch.Tensor:
        """[2] box lengths of the periodic axes (0 on the others), like ``ref``; built once per (box, device, dtype),
        so later calls make no host-to-device copy (legal while a graph is captured)."""
        key = (self, ref.device, ref.dtype)
        t = _LENGTHS.get(key)
        if t is None:
            vals = [float(h) - float(l) if f else 0.0 for l, h, f in zip(self.lo, self.hi, self.flags)]
            t = _LENGTHS[key] = torch.tensor(vals, dtype=ref.dtype, device=ref.device)
        return t

    def checkSupport(self, H: float):
        """ValueError unless every periodic length exceeds twice the support (one image within reach)."""
        for a, f in enumerate(self.flags):
            if f and 2.0 * float(H) >= float(self.hi[a]) - float(self.lo[a]):
                raise ValueError(f"periodic axis {a}: the box length {float(self.hi[a]) - float(self.lo[a])} must "
                                 f"exceed twice the support {float(H)}")


def min_image(d: torch.Tensor, periodic: Optional[Periodic]) -> torch.Tensor:
    """the displacements ``d`` [..., 2] with whole box lengths removed on the periodic axes (``d`` itself without a
    box)."""
    if periodic is None:
        return d
    L = periodic.length(d)
    safe = torch.where(L > 0, L, torch.ones_like(L))
    return d - L * torch.round(d / safe)
