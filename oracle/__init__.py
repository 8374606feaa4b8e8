"""FP64 CPU oracle of the per-view transport hot path — TEST INFRASTRUCTURE ONLY.

This package is a from-scratch restatement, in numpy float64, of the algorithms of the reference's scatter solver
(``citrus/scatter_solver``). It exists to check the HIP product path; it is never imported by ``citrus_amd``. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker (or, for ``cpu_baseline``, as the CPU timing reference).

Each function cites the reference file:line it follows. Every function is pinned by at least one check that does not
retype its own formula (see ``tests/test_oracle.py``):

* ``quadrature.real_spherical_harmonics`` — ``scipy.special.sph_harm_y`` (different algorithm) and the closed form
  ``Y_1 = sqrt(3/4pi) (y, z, x)``; orthonormality on product sets.
* ``quadrature.level_symmetric_quadrature`` — exact integration of even monomials up to the set's order, octant
  symmetry, 4 pi weight sum, published direction counts N(N+2).
* ``ld.cell_solve`` — ``numpy.linalg.solve`` of the dense 4x4 cell matrix.
* ``sweep.sweep`` — the standalone 1-D slab 2x2 recursion, exact particle balance, the dense global upwind
  assembly solved by ``numpy.linalg.solve``, Beer's-law refinement order, reflection symmetry.
* ``scattering.scatter_moments`` — brute-force double loop over groups on tiny inputs and the synthetic library's
  closed-form kernel ``a (1 + b eta) / 4 pi`` (exact Legendre moments).
* ``scattering.source_iteration`` — the fixed-point identity ``sweep(S(phi_tot)) + sweep(S_FC) = phi_tot``.
"""
