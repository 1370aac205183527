"""Engine binding of the Berger Skyrme model kernels.

The reference's Numba kernels of this theory (gradient step, RK4 finalise,
energy and baryon-density kernels) are replaced by the engine's HIP stage
kernels for theory ``THEORY_BERGER``; ``Simulation`` dispatches on ``THEORY_ID``.
"""

from skyrmion_solver_amd import _engine

THEORY_ID = _engine.THEORY_BERGER
NUMBER_TOTAL_FIELDS = 4
N_SIGMA_FIELDS = 4
