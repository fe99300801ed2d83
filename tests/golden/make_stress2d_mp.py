"""Writes tests/golden/stress2d_mp.npz: mpmath (50 digits) values of the twelve channels at a fixed stratified
subsample of the 2-D precision stress workload (``workloads.config4_2d()``, 262,144 particles over 256 wedges):
every 64th particle, 16 per wedge (8 anchored on edge interiors, 8 on vertices), d/h = 10^U(-7, 0).

The inputs are stored with the values (wedge vertices, positions, h), so the test does not depend on regenerating
them.  Run from the repository root:  python tests/golden/make_stress2d_mp.py  (about ten minutes on one core).
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

from tests.stress_cases import polygon_oracle                  # noqa: E402
from warpsphboundaries_amd import workloads as WL              # noqa: E402

STRIDE = 64


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "stress2d_mp.npz"), limit=None):
    w = WL.config4_2d()
    idx = np.arange(0, len(w.positions), STRIDE)
    if limit:
        idx = idx[:limit]
    kernels = list(w.kernels)
    values = np.zeros((len(kernels), len(idx), 12))
    t0 = time.time()
    for i in np.unique(w.poly[idx]):
        sel = np.nonzero(w.poly[idx] == i)[0]
        ref = polygon_oracle(w.polygons[i], w.positions[idx[sel]], w.h, kernels, mode="mp", dps=50)
        for ki, k in enumerate(kernels):
            values[ki, sel] = ref[k]
        print(f"wedge {i}: {len(sel)} points, {time.time() - t0:.0f} s", flush=True)
    used = np.unique(w.poly[idx])
    remap = {int(p): j for j, p in enumerate(used)}
    verts = [w.polygons[p] for p in used]
    np.savez_compressed(out, h=w.h, kernels=np.array(kernels), positions=w.positions[idx],
                        poly=np.array([remap[int(p)] for p in w.poly[idx]], dtype=np.int32),
                        d_over_h=w.d_over_h[idx], anchor=w.anchor[idx],
                        angles_deg=w.angles_deg[used],
                        verts=np.concatenate(verts), verts_offsets=np.cumsum([0] + [len(v) for v in verts]),
                        values=values)
    print("wrote", out, values.shape)


if __name__ == "__main__":
    main(*sys.argv[1:2], limit=int(sys.argv[2]) if len(sys.argv) > 2 else None)
