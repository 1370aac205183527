"""FP64 NumPy oracle for the skyrmion_solver hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, site-for-site and in the
reference's own floating-point evaluation order, the arithmetic of the
reference's Numba-CUDA kernels:

* ``core/derivatives.py``  -> :mod:`oracle.stencils`
* ``core/utils.py`` reductions -> :mod:`oracle.reductions`
* ``theories/{nuclear,omega,coulomb,rho}_skyrme/kernels.py`` point functions,
  energy and baryon densities -> :mod:`oracle.theories`
* ``core/integrator.py`` RK4 / arrested Newton flow -> :mod:`oracle.integrator`
  (and :mod:`oracle.parallel`, the same step split into x-slabs over a thread
  pool, bit-identical to the single-threaded driver)
* ``theories/nuclear_skyrme/initial_config.py`` rational-map / hedgehog /
  product initial data -> :mod:`oracle.initial_config`

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it, and only as the checker (or the timed CPU baseline), never
as part of the engine.  The engine (``skyrmion_solver_amd``) never imports it.

How each part is pinned (the reference ships no tests, fixtures or golden
vectors for this path):

* stencils: closed forms (exact on low-degree polynomials), O(h^4)
  convergence, and the exact zeroed-band / asymmetric-guard index pattern;
* point gradients: the Euler-Lagrange identity against ``torch.autograd`` of
  the coded energy densities (pion, Coulomb, rho; omega/Coulomb coupling terms
  against the restored commented-out energy term), vacuum fixed point;
* integrator: closed-form RK4 amplification on a linear problem, S^3 and
  tangency invariants, the arrest rule;
* energy / baryon number: rotation/translation invariance, B -> integer on
  smooth hedgehogs, Derrick-type scaling identities;
* reductions: exact small-integer data, signed-max semantics;
* initial data: profile-fit and map constants against values extracted from
  the reference's source text, lattice baryon number of every rational map
  equal to its degree, Hamilton product and Euler matrices in textbook form.
"""

from . import stencils, reductions, theories, integrator  # noqa: F401
