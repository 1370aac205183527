"""MI355X (gfx950) engine for the skyrmion_solver hot path.

The reference's Python API (``Simulation``, theory plugins, packed parameter
arrays) over hand-written HIP kernels reached through a ctypes C ABI
(``include/skx.h``, ``libskx.so``).
"""

__version__ = "0.1.0"
