"""ctypes binding of the engine's C ABI (``include/skx.h``).

The shared library ``libskx.so`` is built in-tree by
:func:`skyrmion_solver_amd._build.build`.  There is no fallback: if the
library is missing or cannot be loaded, every engine call raises.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libskx.so"

SKX_OK = 0
ERRORS = {
    -1: "SKX_EINVAL",
    -2: "SKX_ENOMEM",
    -3: "SKX_EHIP",
    -4: "SKX_ENONFINITE",
    -5: "SKX_EBUSY",
    -6: "SKX_ENOTSUP",
}

THEORY_PION, THEORY_OMEGA, THEORY_COULOMB, THEORY_RHO = 0, 1, 2, 3
THEORY_BERGER, THEORY_LIGHTLY_BOUND = 4, 5
FP64, FP32 = 0, 1
FIELD, VELOCITY = 0, 1
DENSITY_ENERGY, DENSITY_BARYON, DENSITY_POTENTIAL, DENSITY_RHO = 0, 1, 2, 3
FLAG_NO_GRAPH, FLAG_TRACE = 1, 2
# FP64 pion: RK4 stage kernels with one thread per site (the default is two wherever the lattice allows)
FLAG_ONE_LANE = 4
LUMP_HEDGEHOG, LUMP_RATIONAL_MAP = 0, 1
PROFILE_EXP, PROFILE_GAUSS, PROFILE_FIT = 0, 1, 2
MAX_LUMPS = 16
OBS_COM, OBS_RMS, OBS_QUADRUPOLE, OBS_U, OBS_V, OBS_W, OBS_D = 0, 1, 2, 3, 4, 5, 6
OBS_COUNT = (4, 1, 9, 9, 9, 9, 1)

EXPORTED = (
    "skx_create", "skx_destroy", "skx_footprint_bytes", "skx_upload", "skx_download", "skx_anf_step",
    "skx_anf_run", "skx_anf_prepare", "skx_rk4_steps", "skx_energy", "skx_baryon", "skx_density",
    "skx_eval_gradient",
    "skx_history_attach", "skx_history_read", "skx_trace_read", "skx_trace_reset", "skx_stage_shape", "skx_sync",
    "skx_batch_create", "skx_batch_destroy", "skx_batch_add", "skx_batch_run", "skx_batch_launch",
    "skx_batch_wait", "skx_batch_poll",
    "skx_batch_remove", "skx_diag_stream_copy", "skx_device_info", "skx_last_error", "skx_version",
    "skx_init_field", "skx_batch_init", "skx_observable", "skx_observable_noalloc", "skx_write_dat",
)


class SkxError(RuntimeError):
    """An engine call returned a negative status code."""

    def __init__(self, code: int, where: str, message: str):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {message}")


class Desc(C.Structure):
    _fields_ = [
        ("theory", C.c_int32), ("precision", C.c_int32),
        ("p_i", C.POINTER(C.c_int32)), ("n_i", C.c_int32),
        ("p_f", C.POINTER(C.c_double)), ("n_f", C.c_int32),
        ("device", C.c_int32), ("flags", C.c_uint32),
        ("tile_z", C.c_int32), ("x_chunk", C.c_int32),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int32), ("converged", C.c_int32), ("energy", C.c_double),
        ("err_signed", C.c_double), ("err_abs", C.c_double), ("baryon", C.c_double),
        ("seconds", C.c_double), ("issued", C.c_int32), ("graph_launches", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchDesc(C.Structure):
    _fields_ = [("theory", C.c_int32), ("precision", C.c_int32), ("device", C.c_int32), ("flags", C.c_uint32),
                ("tile_z", C.c_int32), ("max_slots", C.c_int32), ("max_blocks", C.c_int32)]


class Lump(C.Structure):
    _fields_ = [("kind", C.c_int32), ("baryon_number", C.c_int32), ("profile", C.c_int32), ("pad_", C.c_int32),
                ("profile_param", C.c_double), ("center", C.c_double * 3), ("space_rot", C.c_double * 9),
                ("iso_rot", C.c_double * 9)]


class InitDesc(C.Structure):
    _fields_ = [("n_lumps", C.c_int32), ("vacuum_band", C.c_int32), ("noise", C.c_double),
                ("noise_seed", C.c_uint64), ("omega_cs", C.c_double), ("lumps", Lump * MAX_LUMPS)]


class SlotStatus(C.Structure):
    _fields_ = [("slot", C.c_int32), ("active", C.c_int32), ("steps", C.c_int32), ("done", C.c_int32),
                ("converged", C.c_int32), ("nonfinite", C.c_int32), ("energy", C.c_double),
                ("err_signed", C.c_double), ("err_abs", C.c_double), ("baryon", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepRecord(C.Structure):
    _fields_ = [("energy", C.c_double), ("err_signed", C.c_double), ("err_abs", C.c_double),
                ("baryon", C.c_double)]


_lib = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load ``libskx.so`` (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"engine library {p} is missing: build it with skyrmion_solver_amd._build.build()")
    lib = C.CDLL(str(p))
    dp = C.POINTER(C.c_double)
    H = C.c_void_p
    sig = {
        "skx_create": ([C.POINTER(Desc), C.POINTER(H)], C.c_int),
        "skx_destroy": ([H], C.c_int),
        "skx_footprint_bytes": ([C.POINTER(Desc), C.POINTER(C.c_size_t)], C.c_int),
        "skx_upload": ([H, C.c_int32, dp, C.c_size_t], C.c_int),
        "skx_download": ([H, C.c_int32, dp, C.c_size_t], C.c_int),
        "skx_anf_step": ([H, C.c_double, dp, dp, dp], C.c_int),
        "skx_anf_run": ([H, C.c_int32, C.c_double, C.c_int32, C.POINTER(RunStats)], C.c_int),
        "skx_anf_prepare": ([H, C.c_int32], C.c_int),
        "skx_rk4_steps": ([H, C.c_int32, C.POINTER(RunStats)], C.c_int),
        "skx_energy": ([H, dp], C.c_int),
        "skx_baryon": ([H, dp], C.c_int),
        "skx_density": ([H, C.c_int32, dp, C.c_size_t], C.c_int),
        "skx_observable": ([H, C.c_int32, dp, dp, C.c_int32], C.c_int),
        "skx_observable_noalloc": ([H, C.c_int32, dp, dp, C.c_int32], C.c_int),
        "skx_write_dat": ([C.c_char_p, dp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32],
                          C.c_int),
        "skx_eval_gradient": ([H, dp, dp, dp, C.c_size_t], C.c_int),
        "skx_history_attach": ([H, C.c_int32], C.c_int),
        "skx_history_read": ([H, C.POINTER(StepRecord), C.c_int32, C.POINTER(C.c_int32)], C.c_int),
        "skx_trace_read": ([H, dp, C.POINTER(C.c_int64)], C.c_int),
        "skx_trace_reset": ([H], C.c_int),
        "skx_stage_shape": ([H, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
        "skx_sync": ([H], C.c_int),
        "skx_batch_create": ([C.POINTER(BatchDesc), C.POINTER(H)], C.c_int),
        "skx_batch_destroy": ([H], C.c_int),
        "skx_batch_add": ([H, C.POINTER(Desc), dp, dp, C.c_int32, C.c_double, C.POINTER(C.c_int32)], C.c_int),
        "skx_batch_run": ([H, C.c_int32], C.c_int),
        "skx_batch_launch": ([H, C.c_int32], C.c_int),
        "skx_batch_wait": ([H], C.c_int),
        "skx_batch_poll": ([H, C.POINTER(SlotStatus), C.c_int32, C.POINTER(C.c_int32)], C.c_int),
        "skx_batch_remove": ([H, C.c_int32, dp, dp, C.POINTER(SlotStatus)], C.c_int),
        "skx_diag_stream_copy": ([C.c_int32, C.c_size_t, C.c_int32, dp, dp], C.c_int),
        "skx_device_info": ([C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
        "skx_init_field": ([H, C.POINTER(InitDesc)], C.c_int),
        "skx_batch_init": ([H, C.c_int32, C.POINTER(InitDesc)], C.c_int),
        "skx_last_error": ([], C.c_char_p),
        "skx_version": ([], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def _check(rc: int, where: str):
    if rc != SKX_OK:
        msg = load_library().skx_last_error().decode(errors="replace")
        raise SkxError(rc, where, msg)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_desc(theory: int, precision: int, p_i, p_f, device: int = 0, flags: int = 0, tile_z: int = 0,
              x_chunk: int = 0):
    """Build an ``skx_desc``; returns ``(desc, keepalive)``."""
    pi = np.ascontiguousarray(p_i, dtype=np.int32)
    pf = np.ascontiguousarray(p_f, dtype=np.float64)
    d = Desc(theory=theory, precision=precision,
             p_i=pi.ctypes.data_as(C.POINTER(C.c_int32)), n_i=pi.size,
             p_f=pf.ctypes.data_as(C.POINTER(C.c_double)), n_f=pf.size,
             device=device, flags=flags, tile_z=tile_z, x_chunk=x_chunk)
    return d, (pi, pf)


def footprint_bytes(theory: int, precision: int, p_i, p_f, tile_z: int = 0, x_chunk: int = 0) -> int:
    lib = load_library()
    d, keep = make_desc(theory, precision, p_i, p_f, tile_z=tile_z, x_chunk=x_chunk)
    out = C.c_size_t(0)
    _check(lib.skx_footprint_bytes(C.byref(d), C.byref(out)), "skx_footprint_bytes")
    return int(out.value)


def stream_copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20, device: int = 0):
    """Measured copy bandwidth (GB/s, read + write) and ms per copy of ``nbytes``."""
    lib = load_library()
    g, ms = C.c_double(), C.c_double()
    _check(lib.skx_diag_stream_copy(device, nbytes, reps, C.byref(g), C.byref(ms)), "skx_diag_stream_copy")
    return g.value, ms.value


def device_info(device: int = 0):
    """(architecture name, compute-unit count) of `device`."""
    buf = C.create_string_buffer(64)
    cu = C.c_int32()
    _check(load_library().skx_device_info(int(device), buf, 64, C.byref(cu)), "skx_device_info")
    return buf.value.decode(), int(cu.value)


class Engine:
    """One lattice resident on one GPU (an ``skx_handle``)."""

    def __init__(self, theory: int, precision: int, p_i, p_f, device: int = 0, flags: int = 0,
                 tile_z: int = 0, x_chunk: int = 0):
        self._lib = load_library()
        d, keep = make_desc(theory, precision, p_i, p_f, device, flags, tile_z, x_chunk)
        h = C.c_void_p()
        _check(self._lib.skx_create(C.byref(d), C.byref(h)), "skx_create")
        self._h = h
        self.theory = theory
        self.precision = precision
        self.p_i = np.array(p_i, dtype=np.int32)
        self.p_f = np.array(p_f, dtype=np.float64)
        self.nf = int(self.p_i[5])
        self.N = int(self.p_i[0]) * int(self.p_i[1]) * int(self.p_i[2])

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.skx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- data movement
    def upload(self, which: int, host: np.ndarray):
        a = np.ascontiguousarray(host, dtype=np.float64).ravel()
        _check(self._lib.skx_upload(self._h, which, _dptr(a), a.size), "skx_upload")

    def init_field(self, desc: "InitDesc"):
        """Set the field on the device from an ``InitDesc`` (see ``initial_data.device_init_desc``); V = 0."""
        _check(self._lib.skx_init_field(self._h, C.byref(desc)), "skx_init_field")

    def download(self, which: int) -> np.ndarray:
        out = np.empty(self.nf * self.N, dtype=np.float64)
        _check(self._lib.skx_download(self._h, which, _dptr(out), out.size), "skx_download")
        return out

    # -- the hot path
    def anf_step(self, prev_energy: float):
        e, s, a = C.c_double(), C.c_double(), C.c_double()
        _check(self._lib.skx_anf_step(self._h, float(prev_energy), C.byref(e), C.byref(s), C.byref(a)),
               "skx_anf_step")
        return e.value, s.value, a.value

    def anf_run(self, max_steps: int, tol: float = -1.0, check_every: int = 0) -> RunStats:
        st = RunStats()
        _check(self._lib.skx_anf_run(self._h, int(max_steps), float(tol), int(check_every), C.byref(st)),
               "skx_anf_run")
        return st

    def prepare(self, check_every: int = 0):
        """Capture the step graphs of ``anf_run(..., check_every)`` for both ping-pong parities now, so that
        a later run replays them without capturing."""
        _check(self._lib.skx_anf_prepare(self._h, int(check_every)), "skx_anf_prepare")

    def rk4_steps(self, n: int) -> RunStats:
        st = RunStats()
        _check(self._lib.skx_rk4_steps(self._h, int(n), C.byref(st)), "skx_rk4_steps")
        return st

    # -- observables
    def energy(self) -> float:
        e = C.c_double()
        _check(self._lib.skx_energy(self._h, C.byref(e)), "skx_energy")
        return e.value

    def baryon(self) -> float:
        b = C.c_double()
        _check(self._lib.skx_baryon(self._h, C.byref(b)), "skx_baryon")
        return b.value

    def density(self, kind: int) -> np.ndarray:
        out = np.empty(self.N, dtype=np.float64)
        _check(self._lib.skx_density(self._h, int(kind), _dptr(out), out.size), "skx_density")
        return out

    def observable(self, kind: int, com=None) -> np.ndarray:
        """Lattice sums of one observable's per-site integrand (``OBS_*``; FP64)."""
        out = np.empty(OBS_COUNT[kind], dtype=np.float64)
        c = np.ascontiguousarray(np.zeros(3) if com is None else com, dtype=np.float64)
        if c.size != 3:
            raise ValueError("com must hold 3 values")
        _check(self._lib.skx_observable(self._h, int(kind), _dptr(c), _dptr(out), out.size), "skx_observable")
        return out

    def eval_gradient(self, temp: np.ndarray):
        t = np.ascontiguousarray(temp, dtype=np.float64).ravel()
        eg = np.empty_like(t)
        l = np.empty_like(t)
        _check(self._lib.skx_eval_gradient(self._h, _dptr(t), _dptr(eg), _dptr(l), t.size), "skx_eval_gradient")
        return eg, l

    # -- diagnostics
    def history_attach(self, capacity: int):
        _check(self._lib.skx_history_attach(self._h, int(capacity)), "skx_history_attach")

    def history_read(self, n: int):
        buf = (StepRecord * max(int(n), 1))()
        got = C.c_int32()
        _check(self._lib.skx_history_read(self._h, buf, int(n), C.byref(got)), "skx_history_read")
        return [(r.energy, r.err_signed, r.err_abs, r.baryon) for r in buf[: got.value]]

    def trace_read(self):
        ms = (C.c_double * 5)()
        cnt = (C.c_int64 * 5)()
        _check(self._lib.skx_trace_read(self._h, ms, cnt), "skx_trace_read")
        return list(ms), list(cnt)

    def trace_reset(self):
        _check(self._lib.skx_trace_reset(self._h), "skx_trace_reset")

    def stage_shape(self, stage: int) -> dict:
        """Launch shape of stage kind `stage` (0 = energy pass, 1..4 = RK4 stages)."""
        buf = (C.c_int32 * 4)()
        _check(self._lib.skx_stage_shape(self._h, int(stage), buf), "skx_stage_shape")
        return dict(tile_z=buf[0], tile_y=buf[1], x_chunk=buf[2], lanes=buf[3])

    def sync(self):
        _check(self._lib.skx_sync(self._h), "skx_sync")


class Batch:
    """Many lattices of one theory / precision advanced by shared launches (``skx_batch_*``).

    ``tile_z`` 16, 32 or 64 runs every stage on that tile; 0 takes the single-lattice default of a
    tile per stage (FP32: 32 x 16 where that is faster, 32 x 8 elsewhere)."""

    def __init__(self, theory: int, precision: int, max_slots: int = 64, max_blocks: int = 16384,
                 tile_z: int = 32, device: int = 0, flags: int = 0):
        self._lib = load_library()
        d = BatchDesc(theory=theory, precision=precision, device=device, flags=flags, tile_z=tile_z,
                      max_slots=max_slots, max_blocks=max_blocks)
        h = C.c_void_p()
        _check(self._lib.skx_batch_create(C.byref(d), C.byref(h)), "skx_batch_create")
        self._h = h
        self.theory, self.precision, self.max_slots = theory, precision, max_slots
        self._shape = {}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.skx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, p_i, p_f, F0=None, V0=None, max_steps: int = 100000, tol: float = -1.0, device: int = 0,
            x_chunk: int = 0, init: "InitDesc | None" = None) -> int:
        """Add a lattice; raises ``SkxError`` (code -5, SKX_EBUSY) when the batch is full.

        The field is `F0` (host FP64), or built on the device from `init`
        (an ``InitDesc``), or the vacuum when both are None.  ``x_chunk`` =
        planes per block along x (0: the whole lattice)."""
        d, keep = make_desc(self.theory, self.precision, p_i, p_f, device, 0, 0, x_chunk)
        pi = np.asarray(p_i)
        n = int(pi[5]) * int(pi[0]) * int(pi[1]) * int(pi[2])
        f = None if F0 is None else np.ascontiguousarray(F0, dtype=np.float64).ravel()
        v = None if V0 is None else np.ascontiguousarray(V0, dtype=np.float64).ravel()
        slot = C.c_int32()
        _check(self._lib.skx_batch_add(self._h, C.byref(d), _dptr(f) if f is not None else None,
                                       _dptr(v) if v is not None else None, int(max_steps), float(tol),
                                       C.byref(slot)), "skx_batch_add")
        self._shape[slot.value] = n
        if init is not None:
            try:
                _check(self._lib.skx_batch_init(self._h, slot.value, C.byref(init)), "skx_batch_init")
            except SkxError:
                self.remove(slot.value, want_fields=False)
                raise
        return slot.value

    def run(self, steps: int):
        _check(self._lib.skx_batch_run(self._h, int(steps)), "skx_batch_run")

    def launch(self, steps: int):
        """Enqueue `steps` steps and return at once (pair with :meth:`wait`)."""
        _check(self._lib.skx_batch_launch(self._h, int(steps)), "skx_batch_launch")

    def wait(self):
        _check(self._lib.skx_batch_wait(self._h), "skx_batch_wait")

    def poll(self) -> list:
        buf = (SlotStatus * self.max_slots)()
        n = C.c_int32()
        _check(self._lib.skx_batch_poll(self._h, buf, self.max_slots, C.byref(n)), "skx_batch_poll")
        return [buf[i].as_dict() for i in range(n.value)]

    def remove(self, slot: int, want_fields: bool = True):
        n = self._shape.pop(slot)
        F = np.empty(n) if want_fields else None
        V = np.empty(n) if want_fields else None
        st = SlotStatus()
        _check(self._lib.skx_batch_remove(self._h, int(slot), _dptr(F) if want_fields else None,
                                          _dptr(V) if want_fields else None, C.byref(st)), "skx_batch_remove")
        return st.as_dict(), F, V
