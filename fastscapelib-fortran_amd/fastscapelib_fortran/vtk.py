"""Legacy-VTK snapshots of a landscape (FastScape_VTK, src/FastScape_api.f90:728-741; Make_VTK,
src/FastScape_ctx.f90:692-777): host-side file output from the fields copied off the GPU.

The files are the reference's byte for byte: a structured grid of nx x ny points (x, y, h |vex|) in
single precision, then the scalars "H" (h) and "HHHHH" (the caller's field f); with vex < 0 also a
basement file (b |vex|, scalars "B" and "HHHHH") and a sea-level file (scalar "SL").  The reference
writes the floats with Fortran unformatted direct access, i.e. in the machine's byte order (little
endian here), although the legacy format's BINARY section is read as big endian by most readers; that
quirk is kept so that files match.
"""
from __future__ import annotations

import os

import numpy as np


def _i(v, w):
    """Fortran's (iw) edit descriptor: right-justified in w characters."""
    return f"{int(v):{w}d}"


def _sections(nx, ny):
    nn = nx * ny
    header = ("# vtk DataFile Version 3.0\nFastScape\nBINARY\nDATASET STRUCTURED_GRID\n"
              f"DIMENSIONS {_i(nx, 6)} {_i(ny, 6)} 1\nPOINTS{_i(nn, 12)} float\n")
    footer = f"POINT_DATA{_i(nn, 12)}\n"
    part1 = "SCALARS "
    part2 = " float 1\nLOOKUP_TABLE default\n"
    return header, footer, part1, part2


def _points(nx, ny, dx, dy, z):
    """(x, y, z) per node, i fastest: sngl(dx*(i-1)), sngl(dy*(j-1)), sngl(z) -- products in double."""
    x = (dx * np.arange(nx, dtype=np.float64)).astype(np.float32)
    y = (dy * np.arange(ny, dtype=np.float64)).astype(np.float32)
    pts = np.empty((ny, nx, 3), np.float32)
    pts[:, :, 0] = x[None, :]
    pts[:, :, 1] = y[:, None]
    pts[:, :, 2] = np.asarray(z, np.float64).reshape(ny, nx).astype(np.float32)
    return pts.astype("=f4").tobytes()


def write_vtk(step, nx, ny, xl, yl, h, b, f, vex, sealevel=0.0, directory="VTK"):
    """Write VTK/Topography<step>.vtk (and, for vex < 0, Basement and SeaLevel) as Make_VTK does.
    Returns the paths written."""
    nn = nx * ny
    dx, dy = xl / (nx - 1), yl / (ny - 1)
    h = np.asarray(h, np.float64).reshape(-1)[:nn]
    f = np.asarray(f, np.float64).reshape(-1)[:nn]
    header, footer, part1, part2 = _sections(nx, ny)
    os.makedirs(directory, exist_ok=True)
    cstep = f"{int(step):07d}"
    av = abs(vex)

    def scalars(name, vals):
        return (part1 + name + part2).encode() + np.asarray(vals, np.float64).astype("=f4").tobytes()

    paths = []

    def emit(stem, z, blocks):
        p = os.path.join(directory, f"{stem}{cstep}.vtk")
        with open(p, "wb") as fh:
            fh.write(header.encode())
            fh.write(_points(nx, ny, dx, dy, z))
            fh.write(footer.encode())
            for blk in blocks:
                fh.write(blk)
        paths.append(p)

    emit("Topography", h * av, [scalars("H", h), scalars("HHHHH", f)])
    if vex < 0.0:
        b = np.asarray(b, np.float64).reshape(-1)[:nn]
        emit("Basement", b * av, [scalars("B", b), scalars("HHHHH", f)])
        emit("SeaLevel", np.full(nn, sealevel * av), [scalars("SL", np.full(nn, sealevel))])
    return paths
