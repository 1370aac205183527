"""Text output in the reference's .dat format, plus a binary fast path.

Same functions, arguments and files as ``core/io.py`` of the reference
(``output_field_dat`` ``:52-92``, ``output_iteration_data_dat`` ``:94-121``,
``output_density_data_dat`` ``:123-160``, ``output_data_bundle_core``
``:162-247``).  The grids are formatted by the native writer
``skx_write_dat`` (``include/skx.h``), byte-identical to the reference's
``format(v, ".{p}g")`` loops; the one-line scalar files stay in Python.

``binary=True`` on the bundle writes each array as ``.npy`` next to the same
file names instead (``Field.npy`` holds ``(nf, X, Y, Z)``; the grids
``(X, Y, Z)``), and ``read_field`` reads either form back.
"""

from __future__ import annotations

import os

import numpy as np

from skyrmion_solver_amd import _engine


def _write_grids(flat, path, xlen, ylen, zlen, ncomp, precision, component_break, first=0):
    a = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    n = xlen * ylen * zlen
    if a.size < (first + ncomp) * n:
        raise ValueError(f"array holds {a.size} values, need {(first + ncomp) * n}")
    view = a[first * n:(first + ncomp) * n]
    lib = _engine.load_library()
    rc = lib.skx_write_dat(os.fsencode(path), _engine._dptr(view), int(xlen), int(ylen), int(zlen), int(ncomp),
                           int(precision), 1 if component_break else 0)
    _engine._check(rc, "skx_write_dat")


def output_field_dat(field_flat, path, xlen, ylen, zlen, nfields, precision=32):
    """All ``nfields`` components, each followed by a blank line (``core/io.py:52-92``)."""
    _write_grids(field_flat, path, xlen, ylen, zlen, nfields, precision, True)


def output_iteration_data_dat(values, path, precision=32):
    """One tab-separated line of scalars (``core/io.py:94-121``)."""
    with open(path, "w", newline="\n") as f:
        f.write("".join(format(float(v), f".{precision}g") + "\t" for v in values))


def output_density_data_dat(density_flat, a, path, xlen, ylen, zlen, precision=32):
    """Component ``a`` of a flat grid array (``core/io.py:123-160``)."""
    _write_grids(density_flat, path, xlen, ylen, zlen, 1, precision, False, first=int(a))


def lattice_grid(xlen, ylen, zlen, lsx, lsy, lsz) -> np.ndarray:
    """The reference's coordinate grid, flat ``(3, X, Y, Z)`` with ``x = lsx * i`` (``create_grid_kernel``)."""
    g = np.empty((3, xlen, ylen, zlen))
    g[0] = (lsx * np.arange(xlen, dtype=np.float64))[:, None, None]
    g[1] = (lsy * np.arange(ylen, dtype=np.float64))[None, :, None]
    g[2] = (lsz * np.arange(zlen, dtype=np.float64))[None, None, :]
    return g.ravel()


def output_data_bundle_core(output_dir, *, h_Field, h_grid, xlen, ylen, zlen, number_coordinates,
                            number_total_fields, precision, bundle_spec, arrays,
                            lattice_points_name="LatticePoints.dat", lattice_vectors_name="LatticeVectors.dat",
                            field_dump_name="d_Field.dat", binary=False):
    """Lattice metadata, the full field dump and the named component grids (``core/io.py:162-247``)."""
    os.makedirs(output_dir, exist_ok=True)
    nc = int(number_coordinates)
    n = xlen * ylen * zlen
    last = (xlen - 1) * ylen * zlen + (ylen - 1) * zlen + (zlen - 1)
    ext = [float(h_grid[c * n + last]) if nc > c else 0.0 for c in range(3)]
    output_iteration_data_dat(np.array([xlen, ylen, zlen], dtype=np.float64)[:nc],
                              os.path.join(output_dir, lattice_points_name), precision)
    output_iteration_data_dat(np.diag(ext).ravel()[:nc * nc], os.path.join(output_dir, lattice_vectors_name),
                              precision)
    if binary:
        F = np.asarray(h_Field, dtype=np.float64).ravel()[:number_total_fields * n]
        np.save(os.path.join(output_dir, _npy(field_dump_name)), F.reshape(number_total_fields, xlen, ylen, zlen))
    else:
        output_field_dat(h_Field, os.path.join(output_dir, field_dump_name), xlen, ylen, zlen, number_total_fields,
                         precision)
    for key, a, filename in bundle_spec:
        arr = arrays.get(key)
        if arr is None:
            continue
        path = os.path.join(output_dir, filename)
        if binary:
            comp = np.asarray(arr, dtype=np.float64).ravel()[int(a) * n:(int(a) + 1) * n]
            np.save(os.path.join(output_dir, _npy(filename)), comp.reshape(xlen, ylen, zlen))
        else:
            output_density_data_dat(arr, int(a), path, xlen, ylen, zlen, precision)


def _npy(name: str) -> str:
    return os.path.splitext(name)[0] + ".npy"


def read_field(output_dir, xlen, ylen, zlen, nfields, field_dump_name="Field.dat") -> np.ndarray:
    """Read a field dump back as flat ``(nf * N,)`` doubles from ``.npy`` if present, else from the text file.

    The reference writes no loader; this is what a restart needs.
    """
    npy = os.path.join(output_dir, _npy(field_dump_name))
    if os.path.exists(npy):
        F = np.load(npy, allow_pickle=False)
    else:
        with open(os.path.join(output_dir, field_dump_name)) as f:
            F = np.array(f.read().split(), dtype=np.float64)
    F = np.asarray(F, dtype=np.float64).ravel()
    if F.size != nfields * xlen * ylen * zlen:
        raise ValueError(f"field dump holds {F.size} values, expected {nfields * xlen * ylen * zlen}")
    return F
