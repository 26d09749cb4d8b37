"""Writers: the VTK structured-points layout of MPS.cpp:326-335, and the
"kal" table of Output.cpp:640-718 (one header line of `class : name :
property` keys, one tab-separated row of values per calculation).
"""
from __future__ import annotations

import os
import re
from typing import Dict, List, Optional, Sequence, Tuple

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


def _wild(pattern: str, text: str) -> bool:
    """Input::WildCompare (Input.cpp:977-1030): '*' matches any run of
    characters, everything else itself."""
    rx = "^" + ".*".join(re.escape(part) for part in pattern.split("*")) + "$"
    return re.match(rx, text) is not None


def _fmt(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return f"{float(v):.15g}"


def kal_values(result: dict, names: Dict[str, object]) -> List[Tuple[Tuple[str, str, str], object]]:
    """The scalar output lines of one calculation (a `sfbox_input.run`
    result) under the reference's keys: sys free energy and grand potential
    (SF_System.cpp), per molecule phibulk, ln(G(N|1)) and FH-MU
    (SF_Molecule.cpp:355-398, SF_MolList1stO.cpp output), newton
    iterations, accuracy and error occurred (SF_Solve.cpp:191-211)."""
    sysn, newn = str(names.get("sys", "noname")), str(names.get("newton", "noname"))
    out = [(("sys", sysn, "free energy"), result["free_energy"]),
           (("sys", sysn, "grand potential"), result["grand_potential"])]
    for m, mn in enumerate(names.get("mol", [])):
        out += [(("mol", mn, "phibulk"), result["phibulk"][m]),
                (("mol", mn, "ln(G(N|1))"), result["lngn"][m]),
                (("mol", mn, "FH-MU"), result["chemical_potential"][m])]
    out += [(("newton", newn, "iterations"), int(result["iterations"])),
            (("newton", newn, "accuracy"), result["residual_norm"]),
            (("newton", newn, "error occurred"), not result["converged"])]
    return out


def kal_text(results: Sequence[dict], names: Dict[str, object], template: Optional[str] = None,
             header: bool = True) -> str:
    """Header (unless `header` is false, as when appending to a file that
    exists) and one row per calculation; with a template (SFbox input
    syntax), only lines whose class, name and property match one of its
    lines, '*' allowed (Output::ValueInTemplate)."""
    keep = None
    if template is not None:
        from .sfbox_input import tokenize
        keep = [l[:3] for b in tokenize(template) for l in b]
    rows = []
    for r in results:
        vals = kal_values(r, names)
        if keep is not None:
            vals = [(k, v) for k, v in vals if any(all(_wild(w, x) for w, x in zip(t, k)) for t in keep)]
        rows.append(vals)
    lines = []
    if header and rows:
        lines.append("\t".join(" : ".join(k) for k, _ in rows[0]))
    for vals in rows:
        lines.append("\t".join(_fmt(v) for _, v in vals))
    return "".join(l + "\n" for l in lines)


def write_kal(path: str, results: Sequence[dict], names: Dict[str, object], template: Optional[str] = None,
              append: bool = False) -> None:
    """OutputKal (Output.cpp:640-718): a new file starts with the header; an
    appended one gets rows only when it already exists."""
    header = not (append and os.path.exists(path))
    with open(path, "a" if append else "w") as f:
        f.write(kal_text(results, names, template, header))
