"""CPU oracle of the wall-integral hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import anything
here, and only as the checker (or the timed CPU baseline), never as the product path: the product
(``warpsphboundaries_amd``) runs on the HIP library and fails loudly when it is missing.

Contents (each function cites the reference file:line it restates):

* ``edge2d``    numpy restatement of the stage-1 edge reductions (``edge/core.py``, ``edge/primitives.py``,
                ``edge/geometry.py``) per (query, edge) pair: the twelve edge-local channels.
* ``adjacency`` FP64 numpy restatement of the static cell list, the fixed-capacity slot fill and the winding
                indicator (``scene/scene.py:100-138,273-371``, ``scene/fixedadj.py:67-147``).
* ``mesh3d``    numpy restatement of the 3-D face->edge reductions of ``mesh3d/reference.py`` (closed surfaces,
                tetrahedra, open triangles), the element adjacency predicates and the face solid angle.
* ``kernels``   the kernel blocks re-derived with sympy from ``curvbound/kernels.py``.
* ``disk2d``    the disk element (circle identity of ``edge/disk.py`` with second moments) and the disk cell list /
                slot fill (``scene/scene.py:869-888``, ``scene/fixedadj.py:199-222``).
* ``cone2d``, ``triangle2d``  the cone area / direction extreme and the conditioned triangle quadrature.
* ``contour2d`` the marching-squares zero level set of a sampled distance (``scene/scene.py:762-806``).
* ``sdf2d``     the bilinear signed distance of a sampled field with node-blended jets (``scene/scene.py:665-706``).
* ``sceneop``   the boundary operations over per-query moments (``scene/scene.py:1640-1707``), pinned by quadrature.
* ``torch2d``   a torch-CPU restatement of the stage-1 reductions, the adjacency and the indicator (the timed CPU
                baseline of ``bench.py``).
* ``arith``     the arithmetic the restatements run in: float64 (default), x86 long double, or mpmath.

How each part is pinned (tests/test_oracle_*.py, tests/test_mesh3d_cpu.py): the reference's 40-digit 2-D golden
fixtures (``tests/fixtures/edge2d_golden.json`` of the reference, copied as data under ``tests/golden/``) and its
70-digit 3-D fixtures (``fluxbridge-v1.json``, ``fluxbridge-nearplane-v1.json``), exact rational whole-support
moments, the half-plane closed form, the cavity identity, and brute force for the adjacency; the extended-precision
modes against the same fixtures at 1e-15 (long double) and 1e-33 (mpmath).
"""
