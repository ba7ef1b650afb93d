"""Stratigraphic horizons (FastScape_Strati, src/FastScape_api.f90:745-758; Activate_Strati and run_Strati,
src/FastScape_ctx.f90:781-839; Strati, src/Strati.f90:1-82; VTK, VTK_filled and VTK_CUBE,
src/VTK.f90:1-256): the fields of every reflector at an output step and the files the reference writes.

The reflectors themselves live on the GPU and are updated inside every step (uplifted, cut down to h when
buried, set to h from the active one up, with the erosion below each accumulated); at an output step the
host reads them back with their slopes and distances to shore (computed on the GPU), completes the ten
fields here and writes, into ``VTK/`` as the reference does:

* ``Horizon<iii>-<step>.vtk`` per reflector: binary structured grid (x, y, depth * vex) with the scalars
  ``H`` and the ten fields, floats in the machine's byte order (as the reference's unformatted direct
  access writes them);
* ``Strati-<step>.vtk``: the filled stratigraphy as an ASCII unstructured grid of hexahedra between the
  basement and successive reflectors, with Fortran's F16.4 / E10.4 / I10 edit descriptors;
* ``CUBE.vtk`` once the last reflector is active: every field on an nx x ny x nreflector block (the
  header counts nreflector + 1 layers and the record keeps the reference's length, zero-filled).
"""
from __future__ import annotations

import os

import numpy as np

from .vtk import _i, _points, _sections

NFIELD = 10
NAMES = ["1.CurrentDepth(m)", "2.CurrentSlope(Deg)", "3.ThicknessToNextReflector(m)", "4.ThicknessToBasement(m)",
         "5.DepositionalBathymetry(m)", "6.DepositionalSlope(Deg)", "7.DistanceToShore(m)", "8.Sand/ShaleRatio",
         "9.ReflectorAge(yr)", "A.ThicknessErodedBelow(m)"]


def update_fields(fields, refl, slopes, dists, b, fmix, iref, sealevel, dtref):
    """Strati's fields 1..9 (src/Strati.f90:31-50) in place: fields (nref, 10, nn), refl / slopes / dists
    (nref, nn), iref the active reflector (1-based).  Fields 5..8 of a reflector below the active one
    keep the values of the last output at which it was active; field 10 is the GPU's running sum."""
    nref = refl.shape[0]
    for i in range(1, nref + 1):
        r = refl[i - 1]
        f = fields[i - 1]
        f[0] = r
        f[1] = slopes[i - 1]
        f[2] = r - refl[i - 2] if i > 1 else r - b
        f[3] = r - b
        if i >= iref:
            f[4] = r + sealevel
            f[5] = slopes[i - 1]
            f[6] = dists[i - 1]
            f[7] = fmix
        f[8] = max(dtref * float(nref - i), 0.0)


def _cstep(step):
    return f"{int(step):07d}"


def _e10_4(v):
    """Fortran's E10.4 edit descriptor: 0.dddd E+ee (exponent beyond two digits without the E), the
    optional leading zero dropped for negative values."""
    v = float(v)
    if v != v:
        return "       NaN"
    if v in (float("inf"), float("-inf")):
        return "  Infinity" if v > 0 else " -Infinity"
    if v == 0.0:
        return "-.0000E+00" if np.signbit(v) else "0.0000E+00"
    m, e = f"{abs(v):.3E}".split("E")
    digits = m.replace(".", "")
    e = int(e) + 1
    ex = f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"
    return ("-." if v < 0 else "0.") + digits + ex


def write_horizons(step, nx, ny, xl, yl, refl, fields, vex, directory="VTK"):
    """Horizon<iii>-<step>.vtk of every reflector (VTK, src/VTK.f90:1-74, as Strati calls it :64-69)."""
    nn = nx * ny
    dx, dy = xl / (nx - 1), yl / (ny - 1)
    header, footer, part1, part2 = _sections(nx, ny)
    os.makedirs(directory, exist_ok=True)
    paths = []
    for i in range(1, refl.shape[0] + 1):
        h = np.asarray(refl[i - 1], np.float64)
        p = os.path.join(directory, f"Horizon{i:03d}-{_cstep(step)}.vtk")
        with open(p, "wb") as fh:
            fh.write(header.encode())
            fh.write(_points(nx, ny, dx, dy, h * vex))
            fh.write(footer.encode())
            fh.write((part1 + "H" + part2).encode() + h.astype("=f4").tobytes())
            for k in range(NFIELD):
                fh.write((part1 + NAMES[k] + part2).encode() + np.asarray(fields[i - 1, k, :nn]).astype("=f4").tobytes())
        paths.append(p)
    return paths


def write_filled(step, nx, ny, xl, yl, basement, refl, fields, distb, vex, directory="VTK"):
    """Strati-<step>.vtk (VTK_filled, src/VTK.f90:132-256): ASCII unstructured grid of the basement and
    every reflector, hexahedra between successive surfaces, the reflector index and the ten fields
    (on the basement layer: depth and slope of the first reflector, distance to shore of the basement,
    age extrapolated from the first two reflectors, 0 otherwise)."""
    nn = nx * ny
    nref = refl.shape[0]
    dx, dy = xl / (nx - 1), yl / (ny - 1)
    nnode = nn * (nref + 1)
    nelem = (nx - 1) * (ny - 1) * nref
    xs = dx * np.arange(nx, dtype=np.float64)
    ys = dy * np.arange(ny, dtype=np.float64)
    out = ["# vtk DataFile Version 3.0", "FilledStratigraphy", "ASCII", "DATASET UNSTRUCTURED_GRID",
           "POINTS " + _i(nnode, 10) + " float"]
    for k in range(nref + 1):
        z = (basement if k == 0 else refl[k - 1]) * vex
        for j in range(ny):
            for i in range(nx):
                out.append(f"{xs[i]:16.4f}{ys[j]:16.4f}{z[j * nx + i]:16.4f}")
    out.append("CELLS " + _i(nelem, 10) + _i(9 * nelem, 10))
    for k in range(1, nref + 1):
        for j in range(1, ny):
            for i in range(1, nx):
                ij = (k - 1) * nn + (j - 1) * nx + i - 1
                out.append("".join(_i(v, 10) for v in (8, ij, ij + 1, ij + 1 + nx, ij + nx, ij + nn, ij + 1 + nn,
                                                        ij + 1 + nx + nn, ij + nx + nn)))
    out.append("CELL_TYPES " + _i(nelem, 10))
    out.extend(["12"] * nelem)
    out.append("POINT_DATA " + _i(nnode, 10))
    out.append("SCALARS 0.Reflector float 1")
    out.append("LOOKUP_TABLE default")
    for k in range(nref + 1):
        out.extend([_e10_4(np.float32(k))] * nn)
    for l in range(1, NFIELD + 1):
        out.append("SCALARS " + f"{NAMES[l - 1]:<30s}" + " float 1")
        out.append("LOOKUP_TABLE default")
        for k in range(nref + 1):
            if k == 0:
                if l in (1, 2):
                    vals = fields[0, l - 1]
                elif l == 7:
                    vals = distb
                elif l == 9:  # the reference reads a second reflector's age (out of bounds with only one)
                    vals = 2 * fields[0, 8] - (fields[1, 8] if nref > 1 else 0.0)
                else:
                    vals = np.zeros(nn)
            else:
                vals = fields[k - 1, l - 1]
            out.extend(_e10_4(v) for v in vals)
    os.makedirs(directory, exist_ok=True)
    p = os.path.join(directory, f"Strati-{_cstep(step)}.vtk")
    with open(p, "w") as fh:
        fh.write("\n".join(out) + "\n")
    return p


def write_cube(nx, ny, xl, yl, fields, directory="VTK"):
    """CUBE.vtk (VTK_CUBE, src/VTK.f90:78-128): the fields of every reflector on a block of nx x ny x
    nreflector points spaced dz = (dx + dy) / 2 vertically.  The header counts nx*ny*(nreflector + 1)
    points and the record has the length those would take; the bytes past the written data are zero."""
    nref = fields.shape[0]
    nn1 = nx * ny
    nn = nn1 * (nref + 1)
    dx, dy = xl / (nx - 1), yl / (ny - 1)
    dz = (dx + dy) / 2.0
    header = ("# vtk DataFile Version 3.0\nFastScape\nBINARY\nDATASET STRUCTURED_GRID\n"
              f"DIMENSIONS {_i(nx, 6)} {_i(ny, 6)} {_i(nref, 3)}\nPOINTS{_i(nn, 12)} float\n")
    footer = f"POINT_DATA{_i(nn, 12)}\n"
    part1, part2 = "SCALARS ", " float 1\nLOOKUP_TABLE default\n"
    x = (dx * np.arange(nx, dtype=np.float64)).astype(np.float32)
    y = (dy * np.arange(ny, dtype=np.float64)).astype(np.float32)
    pts = np.empty((nref, ny, nx, 3), np.float32)
    pts[..., 0] = x[None, None, :]
    pts[..., 1] = y[None, :, None]
    pts[..., 2] = (dz * np.arange(nref, dtype=np.float64)).astype(np.float32)[:, None, None]
    body = header.encode() + pts.astype("=f4").tobytes() + footer.encode()
    for k in range(NFIELD):
        body += (part1 + NAMES[k] + part2).encode() + np.asarray(fields[:, k, :nn1]).astype("=f4").tobytes()
    recl = len(header) + 12 * nn + len(footer) + NFIELD * (len(part1) + len(part2) + 4 * nn) + sum(map(len, NAMES))
    body += b"\0" * max(0, recl - len(body))
    os.makedirs(directory, exist_ok=True)
    p = os.path.join(directory, "CUBE.vtk")
    with open(p, "wb") as fh:
        fh.write(body)
    return p


def write_output(step, nx, ny, xl, yl, refl, fields, basement, distb, vex, iref, directory="VTK"):
    """Everything Strati writes at an output step (src/Strati.f90:64-76); returns the paths."""
    paths = write_horizons(step, nx, ny, xl, yl, refl, fields, vex, directory)
    paths.append(write_filled(step, nx, ny, xl, yl, basement, refl, fields, distb, vex, directory))
    if iref == refl.shape[0]:
        paths.append(write_cube(nx, ny, xl, yl, fields, directory))
    return paths
