"""FP64 NumPy oracle for the skyrmion_solver hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, site-for-site and in the
reference's own floating-point evaluation order, the arithmetic of the
reference's Numba-CUDA kernels:

* ``core/derivatives.py``  -> :mod:`oracle.stencils`
* ``core/utils.py`` reductions -> :mod:`oracle.reductions`
* ``theories/{nuclear,omega,coulomb,rho}_skyrme/kernels.py`` point functions,
  energy and baryon densities -> :mod:`oracle.theories`
* ``core/integrator.py`` RK4 / arrested Newton flow -> :mod:`oracle.integrator`

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it, and only as the checker (or the timed CPU baseline), never
as part of the engine.  The engine (``skyrmion_solver_amd``) never imports it.

How each part is pinned (there are no reference tests, fixtures or golden
vectors to check against; the reference cannot be imported here):

* stencils: closed forms (exact on low-degree polynomials), O(h^4)
  convergence, and the exact zeroed-band / asymmetric-guard index pattern;
* point gradients: the Euler-Lagrange identity against ``torch.autograd`` of
  the coded energy densities (pion, Coulomb, rho; omega/Coulomb coupling terms
  against the restored commented-out energy term), vacuum fixed point;
* integrator: closed-form RK4 amplification on a linear problem, S^3 and
  tangency invariants, the arrest rule;
* energy / baryon number: rotation/translation invariance, B -> integer on
  smooth hedgehogs, Derrick-type scaling identities;
* reductions: exact small-integer data, signed-max semantics.
"""

from . import stencils, reductions, theories, integrator  # noqa: F401
