"""The reference's text outputs, for figure reproduction.

Numbers are written as C++ `ostream << setprecision(16) << v` writes them in
the default float format, i.e. C's "%.16g".

* velocity dumps of the two-surface solve (highreg_stokes_2surf/main.cpp:258-270):
  one line "u v w" per quadrature point, one file per surface;
* mid-plane slices y = 1.5 of the grid program
  (highreg_stokes_extend_to_grid/main.cpp:115-118, 294-303, 543-549, 919-925):
  values of the slice j (j h = 1.5) for k = 0..N, each followed by two spaces.
  The pressure files (p_midY, err_p_midY) end a line after every (i, j) pair,
  so all lines but the slice's are empty; the velocity files (err_u_midY,
  err_gradu_midY) end a line after every i.
"""
from __future__ import annotations

import numpy as np

__all__ = ["fmt", "write_velocity", "read_velocity", "write_mid_y_slice", "read_mid_y_slice",
           "write_grid_files"]


def fmt(v):
    """%.16g, the reference's setprecision(16) default-format number."""
    return "%.16g" % float(v)


def write_velocity(path, u):
    """One "u v w" line per point, u = 3n interleaved (main.cpp:258-270)."""
    u = np.asarray(u, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        for a, b, c in u:
            f.write(f"{fmt(a)} {fmt(b)} {fmt(c)}\n")


def read_velocity(path):
    return np.loadtxt(path, dtype=np.float64, ndmin=2).ravel()


def _mid_index(N, h, mid):
    js = [j for j in range(N + 1) if abs(j * h - mid) < 1.e-8]
    return js


def write_mid_y_slice(path, field, h, newline_per="i", mid=1.5):
    """field: (N+1)^3 grid; the values at j h = mid (|j h - mid| < 1e-8).
    newline_per "j": a line break after every (i, j) (the pressure files),
    "i": after every i (the velocity files)."""
    field = np.asarray(field, dtype=np.float64)
    N = field.shape[0] - 1
    js = set(_mid_index(N, h, mid))
    with open(path, "w") as f:
        for i in range(N + 1):
            for j in range(N + 1):
                if j in js:
                    f.write("".join(fmt(field[i, j, k]) + "  " for k in range(N + 1)))
                if newline_per == "j":
                    f.write("\n")
            if newline_per == "i":
                f.write("\n")


def read_mid_y_slice(path):
    """The non-empty lines of a slice file as an array (rows x (N+1))."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([float(v) for v in line.split()])
    return np.array(rows)


def write_grid_files(res, directory, tag="h32"):
    """The four slice files of the grid program from hor.grid.grid_program's
    result: p_midY (exact pressure), err_p_midY, err_u_midY, err_gradu_midY
    (main.cpp:115-118), names suffixed with tag as the reference's "_h32"."""
    import os
    h = res["h"]
    outside_on = (res["phi"] > 0) | (np.abs(res["phi"]) < 1.e-12)
    err_p = np.where(outside_on, np.abs(res["p"] - res["p_exact"]), 0.0)
    err_u = np.where(outside_on, np.sqrt(((res["u"] - res["u_exact"]) ** 2).sum(0)), 0.0)
    from .grid import gradient_error
    paths = {}
    for name, field, per in (("p_midY", res["p_exact"], "j"), ("err_p_midY", err_p, "j"),
                             ("err_u_midY", err_u, "i"),
                             ("err_gradu_midY", gradient_error(res["u"], res["u_exact"], h), "i")):
        p = os.path.join(directory, f"{name}_{tag}")
        write_mid_y_slice(p, field, h, newline_per=per)
        paths[name] = p
    return paths
