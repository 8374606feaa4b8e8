"""FP64 CPU oracle of the per-view transport hot path — TEST INFRASTRUCTURE ONLY.

This package is a from-scratch restatement, in numpy float64, of the algorithms of the reference's scatter solver
(``citrus/scatter_solver``). It exists to check the HIP product path; it is never imported by ``citrus_amd``. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker (or, for ``cpu_baseline``, as the CPU timing reference).

Each function cites the reference file:line it follows, and is pinned by checks that do not retype its own formula
(closed forms, invariants, brute force, library routines, or the reference's own recorded outputs). The pins, by
module (tests/test_oracle.py unless named):

* ``quadrature`` — real spherical harmonics against ``scipy.special.sph_harm_y`` and the closed form of degree 1,
  the addition theorem, orthonormality on product sets; level-symmetric sets by exact integration of even monomials,
  octant symmetry, weight sum and N(N+2) counts, and S12's exact integration of the P5 basis products.
* ``ld.cell_solve`` — ``numpy.linalg.solve`` of the dense 4x4 cell matrix.
* ``sweep.sweep`` — the 1-D slab recursion, exact particle balance, the dense global upwind assembly solved by
  ``numpy.linalg.solve``, Beer's-law refinement order, reflection symmetry.
* ``scattering`` — brute-force loops over groups; the source iteration by its fixed-point identity.
* ``rays`` — voxel traversal and line integrals against midpoint sampling on a random non-uniform field and slab
  chords; the panel frame and distances through the primary's vacuum and slab closed forms.
* ``uncollided`` — group transmission and the inverse-square and Beer's-law closed forms
  (tests/test_uncollided_hip.py); panel hits, off-focal weights and bilinear profile/filter lookups against hand-
  computed cases (inside, on the margin, at and between off-focal nodes, past the last node).
* ``first_collision`` — per-element interpolation by ``numpy.interp``, the synthetic kernel's closed form, exact
  Legendre moments against a sphere quadrature, adaptive-table node positions, exact restriction of linear fields,
  the conservation identities of the sampling correction.
* ``image`` — box-matrix properties, single-voxel bookkeeping, convergence of the collapsed-cone image to a brute-force
  scatter integral, detector-table lookups on the nodes, the direction-context weights (flat and curved) restated
  from the cone set alone and compared with the product's context.
* ``fields.resample_by_lattice`` — exact block means on nested grids and mass conservation
  (tests/test_resample_hip.py).
* ``grid`` — the direct transmission by midpoint brute force over the entry offset and the track, and its special
  cases; the lamella track sampling (tests/test_grid.py).
* ``density`` — the model's identities (the attenuation reproduced, nominal values map to one material, X <= 0 to
  none, at most two adjacent materials mix).
* ``state_store`` — the float16 format's exact points (the maximum maps to 2^14, zeros, values on the float16 grid)
  and numpy's IEEE rounding.
* ``surrogate`` — convolutions against brute force, all gradients against central finite differences, pooling and
  upsampling adjoints, Adam's first step in closed form, bfloat16 rounding against known bit patterns.
* The whole view: the reference's recorded regression (tests/golden/pre_quadrature_reference_synthetic.json, cases
  synthetic_cpu and synthetic_cuda of its test_quadrature_grid.py) reproduced by ``ScatterSolver.solve_view``
  (tests/test_golden_view.py).
"""
