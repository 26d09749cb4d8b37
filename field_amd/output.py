"""Profile writers: the VTK structured-points layout of MPS.cpp:326-335."""
from __future__ import annotations

import numpy as np


def vtk_text(phi: np.ndarray, dims) -> str:
    """VTK text of an interior profile. `phi` is the halo-padded lattice vector
    ((MX*MY*MZ), last gradient contiguous); values are written x-major with z
    fastest, as MPS.cpp:329-335 loops i, j, l over x, y, z."""
    mx, my, mz = dims
    a = np.asarray(phi, dtype=np.float64).reshape(mx + 2, my + 2, mz + 2)[1:-1, 1:-1, 1:-1]
    head = (
        "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\n"
        f"DIMENSIONS {mx} {my} {mz}\n"
        f"SPACING 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {mx * my * mz}\n"
        "SCALARS SFBox_profile double\nLOOKUP_TABLE default\n"
    )
    return head + "".join(f"{v:e}\n" for v in a.reshape(-1))


def write_vtk(path: str, phi: np.ndarray, dims) -> None:
    with open(path, "w") as f:
        f.write(vtk_text(phi, dims))


def read_vtk(path: str):
    """Returns (dims, values) of a VTK file written in the layout above."""
    with open(path) as f:
        lines = f.read().splitlines()
    dims = None
    for i, l in enumerate(lines):
        if l.startswith("DIMENSIONS"):
            dims = tuple(int(v) for v in l.split()[1:4])
        if l.startswith("LOOKUP_TABLE"):
            vals = np.array([float(v) for v in lines[i + 1:] if v.strip()])
            return dims, vals
    raise ValueError("not a structured-points VTK file")
