"""``Simulation``: the reference's driver API on top of the gfx950 engine.

Mirrors ``core/simulation.py`` of the reference (``Simulation.__init__``
``:46-122``, ``initialize`` ``:124-173``, ``set_params`` ``:175-197``,
``observables`` ``:199-217``, the density helpers ``:219-285``, ``step``
``:287-308``, ``save_output`` ``:310-370`` and the post-convergence
observables ``:390-493``) with the same method names, arguments and return
values.  The per-step work runs in the engine's HIP kernels through the C ABI
(``include/skx.h``); there is no CPU path.

Device buffers the reference exposes as Numba ``DeviceNDArray`` objects
(``Field``, ``Velocity``, ``en``) are exposed as :class:`DeviceBuffer`, which
keeps the ``copy_to_host`` / ``copy_to_device`` methods the reference's
callers use.
"""

from __future__ import annotations

import numpy as np

from skyrmion_solver_amd import _engine
from skyrmion_solver_amd import initial_data
from skyrmion_solver_amd.core.params import Params

PRECISIONS = {"fp64": _engine.FP64, "float64": _engine.FP64, "fp32": _engine.FP32, "float32": _engine.FP32}


class DeviceBuffer:
    """A field-sized engine buffer with the DeviceNDArray copy methods."""

    def __init__(self, sim: "Simulation", which: int):
        self._sim = sim
        self._which = which

    @property
    def size(self) -> int:
        return int(self._sim.rp.dim_fields)

    @property
    def shape(self):
        return (self.size,)

    dtype = np.dtype(np.float64)

    def copy_to_host(self) -> np.ndarray:
        return self._sim.engine.download(self._which)

    def copy_to_device(self, host) -> None:
        self._sim.engine.upload(self._which, np.asarray(host, dtype=np.float64))


class HostDensity:
    """The last density computed by a ``compute_*_density`` call (the reference's ``en``)."""

    def __init__(self, values: np.ndarray | None = None):
        self.values = values

    def copy_to_host(self) -> np.ndarray:
        if self.values is None:
            raise RuntimeError("no density has been computed yet")
        return self.values.copy()


class Simulation:
    """Bind a theory to the engine (``core/simulation.py:28-122``).

    Parameters
    ----------
    params : Params
        Theory parameter set (resolved with ``params.resolved()``).
    theory : module
        A theory package from :func:`skyrmion_solver_amd.theories.load_theory`.
    precision : str
        ``"fp64"`` (plain form, reference arithmetic) or ``"fp32"`` (FP32
        storage and stencils, FP64 reductions and renormalisation).
    device : int
        HIP device ordinal.
    engine_options : dict
        ``flags``, ``tile_z``, ``x_chunk`` passed to ``skx_create``.
    """

    def __init__(self, params, theory, precision: str = "fp64", device: int = 0, engine_options: dict | None = None):
        self.theory = theory
        self.kernels = theory.kernels
        self.initial_config = theory.initial_config
        self.observables_mod = theory.observables
        self.threads3d = getattr(theory, "threads3d", (8, 8, 4))
        if getattr(self.kernels, "THEORY_ID", None) is None:
            raise AttributeError("Theory kernels must provide THEORY_ID (the engine theory this plugin runs on).")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.device = device
        self.engine_options = dict(engine_options or {})
        self.engine = None
        self.set_params(params)
        self.en = HostDensity()
        self.entmp = HostDensity()

    # ------------------------------------------------------------ parameters
    def set_params(self, p):
        """Update the parameters (``core/simulation.py:175-197``).

        The packed arrays are handed to a fresh engine handle; an existing
        field and velocity are carried over when the lattice size is unchanged.
        """
        old = self.engine
        carry = None
        self.p = p
        self.rp = p if not isinstance(p, Params) else p.resolved()
        self.p_i_h, self.p_f_h = self.theory.params.pack_device_params(self.rp)
        if old is not None and int(old.p_i[7]) == int(self.p_i_h[7]):
            carry = (old.download(_engine.FIELD), old.download(_engine.VELOCITY))
        if old is not None:
            old.close()
        self.engine = _engine.Engine(self.kernels.THEORY_ID, PRECISIONS[self.precision], self.p_i_h, self.p_f_h,
                                     device=self.device, **self.engine_options)
        self.Field = DeviceBuffer(self, _engine.FIELD)
        self.Velocity = DeviceBuffer(self, _engine.VELOCITY)
        if carry is not None:
            self.engine.upload(_engine.FIELD, carry[0])
            self.engine.upload(_engine.VELOCITY, carry[1])

    # ------------------------------------------------------------ initial data
    def initialize(self, init_config: dict | None = None):
        """Set the initial field and zero the velocity (``core/simulation.py:124-173``)."""
        cfg = dict(init_config or {})
        F = self.initial_config.initialize(sim=self, config=cfg)
        self.engine.upload(_engine.FIELD, F)
        self.engine.upload(_engine.VELOCITY, np.zeros_like(F))
        post = getattr(self.initial_config, "post_initialize", None)
        if post is not None:
            post(sim=self, config=cfg)

    # ------------------------------------------------------------ observables
    def observables(self) -> dict:
        """Energy and baryon number (``core/simulation.py:199-217``)."""
        obs = {"energy": self.observables_mod.compute_energy(self)}
        if hasattr(self.observables_mod, "compute_skyrmion_number"):
            obs["skyrmion_number"] = self.observables_mod.compute_skyrmion_number(self)
        return obs

    def compute_energy_density(self):
        """Per-site energy (times cell volume) into ``en`` (``core/simulation.py:219-234``)."""
        self.en = HostDensity(self.engine.density(_engine.DENSITY_ENERGY))

    def compute_baryon_density(self):
        """Per-site baryon density (times cell volume) into ``en`` (``:236-251``)."""
        self.en = HostDensity(self.engine.density(_engine.DENSITY_BARYON))

    def compute_potential_density(self):
        """The omega / Coulomb potential component into ``en`` (``:253-268``)."""
        self.en = HostDensity(self.engine.density(_engine.DENSITY_POTENTIAL))

    def compute_rho_meson_density(self):
        """Sum of squared rho components into ``en`` (``:270-285``)."""
        self.en = HostDensity(self.engine.density(_engine.DENSITY_RHO))

    # ------------------------------------------------ post-convergence observables
    # core/simulation.py:390-493: each delegates to the theory's observables
    # module, which raises AttributeError where the theory has no such observable
    # (the omega / Coulomb / rho theories have the centre of mass and RMS radius).
    def compute_center_of_mass(self) -> np.ndarray:
        """Baryon-density centre of mass, shape ``(3,)`` (``core/simulation.py:390-403``)."""
        return self.observables_mod.compute_center_of_mass(self)

    def compute_rms_radius(self) -> float:
        """RMS radius about the centre of mass (``:405-418``)."""
        return self.observables_mod.compute_rms_radius(self)

    def compute_U_tensor(self) -> np.ndarray:
        """Isospin inertia tensor ``U``, shape ``(3, 3)`` (``:420-433``)."""
        return self.observables_mod.compute_U_tensor(self)

    def compute_V_tensor(self) -> np.ndarray:
        """Spin inertia tensor ``V``, shape ``(3, 3)`` (``:435-448``)."""
        return self.observables_mod.compute_V_tensor(self)

    def compute_W_tensor(self) -> np.ndarray:
        """Mixed inertia tensor ``W``, shape ``(3, 3)`` (``:450-463``)."""
        return self.observables_mod.compute_W_tensor(self)

    def compute_quadrupole_tensor(self) -> np.ndarray:
        """Isospin-0 electric quadrupole tensor ``Q``, shape ``(3, 3)`` (``:465-478``)."""
        return self.observables_mod.compute_quadrupole_tensor(self)

    def compute_D_term(self) -> float:
        """Monopole D-term (``:480-493``)."""
        return self.observables_mod.compute_D_term(self)

    # ------------------------------------------------------------ the hot path
    def step(self, prev_energy: float) -> tuple[float, float]:
        """One arrested-Newton-flow RK4 step (``core/simulation.py:287-308``).

        Returns ``(new_energy, err)`` with the reference's ``err`` (the signed
        maximum of the stage-4 energy gradient).  The true ``max |g|`` of the
        same gradient is kept in ``self.last_max_force``.
        """
        e, err, err_abs = self.engine.anf_step(prev_energy)
        self.last_max_force = err_abs
        return e, err

    def minimize(self, tol: float = 1.0e-4, max_steps: int = 100000, check_every: int = 0,
                 refine_fp64: bool = False, refine_factor: float = 3.0) -> dict:
        """Run ANF on the device until the stage-4 ``max |g| < tol``.

        The reference documents ``minimize(tol=...)`` (``core/simulation.py:8,43``)
        but has no such method; this is the device-resident loop of ``step``.

        ``refine_fp64`` (FP32 simulations): the FP32 form's force has a rounding
        floor near 1e-4 on stiff solutions, so the run goes in FP32 until
        ``max |g| < refine_factor * tol`` and then finishes in FP64 from the same
        field and velocity; the simulation stays in FP64 afterwards.  The result
        then carries ``steps_fp32`` and ``refined``.
        """
        if not refine_fp64 or self.precision != "fp32":
            return self.engine.anf_run(max_steps, tol, check_every).as_dict()
        first = self.engine.anf_run(max_steps, refine_factor * tol, check_every).as_dict()
        if first["err_abs"] < tol or first["steps"] >= max_steps:
            return dict(first, steps_fp32=first["steps"], refined=False)
        self.precision = "fp64"
        self.set_params(self.p)  # a fresh FP64 handle; field and velocity carried over
        rest = self.engine.anf_run(max_steps - first["steps"], tol, check_every).as_dict()
        return dict(rest, steps=first["steps"] + rest["steps"], steps_fp32=first["steps"], refined=True,
                    seconds=first["seconds"] + rest["seconds"])

    def save_output(self, output_dir: str | None = None, precision: int = 32, binary: bool = False):
        """Write the theory's output bundle (``core/simulation.py:310-370``).

        The field, the coordinate grid and the per-site energy and baryon
        densities (device FP64 passes) go to ``theory.io.output_data_bundle``
        in the reference's .dat text format and file names; ``output_dir``
        defaults to ``results/`` beside the theory package, as in the
        reference.  ``binary=True`` writes ``.npy`` arrays instead of the grid
        text files (the fast path for large lattices).
        """
        import inspect
        from pathlib import Path

        from skyrmion_solver_amd.core.io import lattice_grid

        if output_dir is None:
            output_dir = str(Path(self.theory.io.__file__).resolve().parent / "results")
        Path(output_dir).mkdir(parents=True, exist_ok=True)
        X, Y, Z = (int(v) for v in self.p_i_h[:3])
        kwargs = dict(output_dir=output_dir,
                      h_Field=self.engine.download(_engine.FIELD),
                      h_EnergyDensity=self.engine.density(_engine.DENSITY_ENERGY),
                      h_BaryonDensity=self.engine.density(_engine.DENSITY_BARYON),
                      h_grid=lattice_grid(X, Y, Z, *(float(v) for v in self.p_f_h[3:6])),
                      xlen=X, ylen=Y, zlen=Z,
                      number_coordinates=int(self.p_i_h[4]),
                      number_total_fields=int(self.p_i_h[5]),
                      precision=precision,
                      binary=binary)
        accepted = set(inspect.signature(self.theory.io.output_data_bundle).parameters)
        self.theory.io.output_data_bundle(**{k: v for k, v in kwargs.items() if k in accepted})

    def load_output(self, output_dir: str, field_dump_name: str = "Field.dat"):
        """Restore the field from a bundle written by ``save_output`` and zero the velocity.

        The reference writes no loader (and no velocity), so a restart
        continues from the field with V = 0.
        """
        from skyrmion_solver_amd.core.io import read_field

        X, Y, Z = (int(v) for v in self.p_i_h[:3])
        F = read_field(output_dir, X, Y, Z, int(self.p_i_h[5]), field_dump_name)
        self.engine.upload(_engine.FIELD, F)
        self.engine.upload(_engine.VELOCITY, np.zeros_like(F))

    def save_checkpoint(self, path: str):
        """Field and velocity, exactly, as ``.npz`` (with ``p_i`` / ``p_f``) for a bit-exact resume."""
        np.savez(path, Field=self.engine.download(_engine.FIELD), Velocity=self.engine.download(_engine.VELOCITY),
                 p_i=self.p_i_h, p_f=self.p_f_h)

    def load_checkpoint(self, path: str):
        """Restore a ``save_checkpoint`` file written for the same lattice."""
        with np.load(path, allow_pickle=False) as d:
            if not np.array_equal(d["p_i"], self.p_i_h):
                raise ValueError("checkpoint was written for different integer parameters (lattice / theory)")
            self.engine.upload(_engine.FIELD, d["Field"])
            self.engine.upload(_engine.VELOCITY, d["Velocity"])

    def print_instructions(self):
        if hasattr(self.theory, "print_instructions"):
            self.theory.print_instructions()
        else:
            print("No specific instructions available for this theory.")

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None
