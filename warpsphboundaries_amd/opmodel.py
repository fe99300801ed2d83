"""Deterministic operation model of the 2-D twelve-channel closed form (algorithmic flops, not measured ones).

Counting rules: fused multiply-add = 2, add / sub / mul / min / max = 1, each sqrt / log / atan2 / division = 4
flop-equivalents (quarter-rate transcendental issue).  The count follows the shared-primitive closed form of
``csrc/edge2d_device.h`` term by term, with the number of nonzero profile coefficients taken from the exact plan
tables, so it is a property of the algorithm and the kernel, not of the compiled code.

Algorithmic bytes per particle: positions (2 x 8 B) + support (8 B) read, twelve float64 outputs written per kernel.
"""
from __future__ import annotations

from .plan import block_profiles

BYTES_PER_PARTICLE_2D = 2 * 8 + 8 + 12 * 8          # 120 B


def bytes_per_particle_2d(n_kernels: int = 1) -> int:
    """positions and support read once, twelve float64 outputs written per kernel (one pass over several kernels)."""
    return 2 * 8 + 8 + n_kernels * 12 * 8
TRANS = 4


def pair_geometry_flops() -> int:
    """relative coordinates, edge frame, predicate: once per (particle, edge) pair."""
    rel = 4 * 1 + 4 * TRANS                  # (v - x) / h for both endpoints
    frame = 2 + 3 + TRANS + TRANS + 2        # d, ell^2 (mul + fma), sqrt, 1/ell, t
    z = 3 + 1                                # separately rounded cross product, * 1/ell
    s = 2 * (3 + 1)                          # d.p and d.q (mul + fma), * 1/ell
    return rel + frame + z + s + 2           # |z|, z^2


def block_nm(b) -> int:
    """the largest power of r a block's profiles carry."""
    return max(len(b.P), len(b.Phi)) - 1


def chord_flops(nm: int) -> int:
    """the primitives of one active (pair, radius) chord, whatever kernel uses them: chord, endpoint radii, one log,
    one angle, the dI recurrence up to power nm and the powers of the endpoint radii."""
    chord = 3 + TRANS + 2
    radii = 2 * (2 + TRANS)
    asinh = 2 * (1 + TRANS) + TRANS + TRANS          # F(hi), F(lo), ratio, log
    dI01 = 1 + (3 + 2 + 1)
    rec = max(nm - 1, 0) * (2 + 3 + 1 + 2 + 1)        # powers, d[s r^m], m z^2, fma, * 1/(m+1)
    powers = 2 * 2 + (nm + 1)                          # the two extra powers and d[r^m]
    angle = 2 + 2 + TRANS
    return chord + radii + asinh + dI01 + rec + powers + angle


def kernel_block_flops(b) -> int:
    """one kernel block over an active chord: the four profile sums over dI, the two over d[r^(n+2)], and the
    twelve channel updates."""
    nz = lambda arr: sum(1 for c in arr if c != 0)
    sums = 2 * (2 * nz(b.P) + 2 * nz(b.Phi))         # A and V of P and Phi over dI
    bsums = 2 * (nz(b.P) + nz(b.Phi))                  # B of P and Phi over d[r^(n+2)]
    channels = 2 * 3 + 4 * 3 + 24                      # lam, valF, Y terms, twelve updates
    return sums + bsums + channels


def block_flops(b) -> int:
    """one active (pair, block) chord of a single kernel: its primitives and its sums and channels."""
    return chord_flops(block_nm(b)) + kernel_block_flops(b)


def multi_flops(knames, npairs: int, active_by_radius: dict, n_rows: int) -> int:
    """flop-equivalents of several kernels in one pass: the pair geometry once per pair, each radius's chord
    primitives once per active (pair, radius) up to the largest power any kernel needs there, then per kernel its
    block sums and channels, the 1/pi scaling and the row work.  ``active_by_radius`` {R: active chords}."""
    total = npairs * pair_geometry_flops()
    by_r = {}
    for k in knames:
        for b in block_profiles(k):
            by_r.setdefault(b.R, []).append(b)
    for R, blocks in by_r.items():
        a = active_by_radius.get(R, 0)
        total += a * (chord_flops(max(block_nm(b) for b in blocks)) + sum(kernel_block_flops(b) for b in blocks))
    return total + len(knames) * (npairs * 12 + n_rows * row_flops())


def pair_flops(kname: str) -> int:
    """flop-equivalents of one (pair) evaluation when every block's chord is active, plus the final 1/pi scaling."""
    return pair_geometry_flops() + sum(block_flops(b) for b in block_profiles(kname)) + 12


def row_flops() -> int:
    """per particle: indicator x plane vector, physical scaling and world rotation of twelve channels."""
    return 12 * 2 + 10 + 30
