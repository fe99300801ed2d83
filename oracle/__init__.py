"""CPU oracle of the wall-integral hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import anything
here, and only as the checker (or the timed CPU baseline), never as the product path: the product
(``warpsphboundaries_amd``) runs on the HIP library and fails loudly when it is missing.

Contents (each function cites the reference file:line it restates):

* ``edge2d``    FP64 numpy restatement of the stage-1 edge reductions (``edge/core.py``, ``edge/primitives.py``,
                ``edge/geometry.py``) per (query, edge) pair: the twelve edge-local channels.
* ``adjacency`` FP64 numpy restatement of the static cell list, the fixed-capacity slot fill and the winding
                indicator (``scene/scene.py:100-138,273-371``, ``scene/fixedadj.py:67-147``).
* ``scene2d``   the per-particle twelve moments of a surface representation (slots + channels + indicator x plane
                vector + physical scaling; ``scene/scene.py:384-416``) and the signed distance (``:1312-1372``).
* ``mp_edge``   small-input mpmath restatement used to pin single open-edge values at high precision.

How each part is pinned (tests/test_oracle_*.py): the reference's 40-digit golden fixtures
(``tests/fixtures/edge2d_golden.json`` of the reference, copied as data under ``tests/golden/``), exact rational
whole-support moments, the half-plane closed form, edge-subdivision invariance, an independent fan-triangle area
quadrature for the open-edge value channel, and brute force for the adjacency.
"""
