"""SFbox input files for the part of SFbox this package runs.

Grammar (Input.cpp:6-130): a line holds four fields separated by ':',
`class : name : property : value`, each stripped of blanks; '#' or '//' ends a
line; a line `start` ends a calculation; blank lines are skipped. A
calculation sees every value set in it and, for what it does not set, the
value of the latest earlier calculation that does (Input.cpp:424-485).

Supported subset, mapped onto LatticeSpec / SystemSpec:
  lat : L : gradients (1, 2, 3), geometry (flat / cylindrical / spherical),
        n_layers (1-D) or n_layers_x / _y / _z, lowerbound / upperbound (1-D)
        or lowerbound_x ... upperbound_z (surface, mirror1, mirror2,
        periodic; default mirror1, Lat1DFlat.cpp:26-35), lambda (1-D and
        2-D), latticetype (standard, FCC, HEX), layer_adjustment,
        offset_first_layer, lambda_chi
  mon : X : freedom (free, frozen), frozen_range (Lat1DFlat.cpp:301-380,
        Lat3D.cpp:617-700: one or two points 'x[,y[,z]]' separated by ';',
        each coordinate an integer 0..n+1 or lowerbound / upperbound /
        lastlayer), chi - Y (set on either side, SF_SegmentList.cpp:145-200),
        Pf - Y (the straight-bond weight of second order, one value)
  mol : P : composition (linear blocks '(A)12(B)28'), freedom (free,
        restricted, solvent, range_restricted), phibulk, theta,
        restricted_range (a range as above, inside the lattice), phi_range
  sys : S : matrix_approach (first_order, second_order; SF_System.cpp:24-60)
  newton : N : method (DIIS runs DIIS; pseudohessian and hessian run
        SFNewton's pseudo-Hessian Newton; any other method, or none, runs
        Newton-Krylov, except with a range_restricted molecule, which runs
        SFNewton), tolerance, iterationlimit, m
  var : V : step, end_value, scale (linear, exponential), type (integer,
        real), and one `class : instance` entry naming the varied
        parameter (Vary.cpp:3-230)
Anything else that changes the physics (pinned or grafted segments, branched
compositions, charges, states, 'bulk' bounds, stencil lattices, ...) raises
InputError naming it. Output, Monte Carlo and solver-tuning properties the
GPU solvers do not use are ignored.
"""
from __future__ import annotations

import math
import os
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

from .spec import LatticeSpec, MoleculeSpec, SystemSpec

Key = Tuple[str, str, str]


class InputError(ValueError):
    pass


def tokenize(text: str) -> List[List[Tuple[str, str, str, str]]]:
    """The calculations of an input text, each a list of (class, name,
    property, value) lines (Input.cpp:6-130). A trailing `start` with no
    lines after it ends the last calculation."""
    blocks: List[List[Tuple[str, str, str, str]]] = [[]]
    for n, raw in enumerate(text.splitlines(), 1):
        line = raw
        for mark in ("#", "//"):
            i = line.find(mark)
            if i >= 0:
                line = line[:i]
        fields = [f.strip() for f in line.split(":")] if line.strip() else []
        if not fields:
            continue
        if len(fields) == 1 and fields[0] == "start":
            if blocks[-1]:
                blocks.append([])
            continue
        if len(fields) > 4:
            raise InputError(f"line {n}: too many elements")
        if len(fields) < 4:
            raise InputError(f"line {n}: too few elements")
        if any(f == "" for f in fields):
            raise InputError(f"line {n}: statement missing")
        blocks[-1].append(tuple(fields))
    if not blocks[-1] and len(blocks) > 1:
        blocks.pop()
    return blocks


class Input:
    """An input text read as calculations (`start` blocks)."""

    def __init__(self, text: str):
        self.blocks = tokenize(text)

    @classmethod
    def from_file(cls, path: str) -> "Input":
        with open(path) as f:
            return cls(f.read())

    def __len__(self) -> int:
        return len(self.blocks)

    def values(self, calc: int) -> Dict[Key, str]:
        """Every (class, name, property) -> value that calculation `calc`
        (0-based) sees: its own, else the latest earlier one."""
        out: Dict[Key, str] = {}
        for b in self.blocks[: calc + 1]:
            for c, nm, p, v in b:
                out[(c, nm, p)] = v
        return out

    def names(self, calc: int, cls: str) -> List[str]:
        """Names of a class in order of first appearance up to `calc`."""
        seen: List[str] = []
        for b in self.blocks[: calc + 1]:
            for c, nm, _, _ in b:
                if c == cls and nm not in seen:
                    seen.append(nm)
        return seen


_BOUNDS = ("surface", "mirror1", "mirror2", "periodic")


def _real(v: str, what: str) -> float:
    try:
        return float(v)
    except ValueError:
        raise InputError(f"{what}: '{v}' is not a number") from None


def _int(v: str, what: str) -> int:
    if not re.fullmatch(r"[+-]?\d+", v):
        raise InputError(f"{what}: '{v}' is not an integer")
    return int(v)


def _lattice(vals: Dict[Key, str], name: str) -> LatticeSpec:
    g = lambda p, d=None: vals.get(("lat", name, p), d)
    grad = _int(g("gradients", "1"), "lat gradients")
    if grad not in (1, 2, 3):
        raise InputError("lat gradients must be 1, 2 or 3")
    geometry = g("geometry")
    if geometry is None:
        raise InputError("lat geometry is required (flat, cylindrical or spherical)")
    if geometry not in ("flat", "cylindrical", "spherical"):
        raise InputError(f"lat geometry '{geometry}' is not supported")
    lt = g("latticetype", "standard")
    if lt not in ("standard", "FCC", "HEX"):
        raise InputError(f"lat latticetype '{lt}' is not supported")
    axes = ("",) if grad == 1 else ("_x", "_y", "_z")[:grad]
    n, lo, hi = [], [], []
    for a in axes:
        v = g("n_layers" + a)
        if v is None:
            raise InputError(f"lat n_layers{a} is required")
        n.append(_int(v, f"lat n_layers{a}"))
        for side, dst in (("lowerbound", lo), ("upperbound", hi)):
            b = g(side + a, "mirror1")
            if b == "bulk":
                raise InputError(f"lat {side}{a}: bulk bounds are not supported")
            if b not in _BOUNDS:
                raise InputError(f"lat {side}{a}: unknown bound '{b}'")
            dst.append(b)
    lam = g("lambda")
    if grad == 1 and lam is None:
        raise InputError("lat lambda is required on a 1-D lattice (Lat1DFlat.cpp:109)")
    return LatticeSpec(
        grad, tuple(n), geometry=geometry, lattice_type=lt,
        lowerbound=tuple(lo), upperbound=tuple(hi),
        lambda_=_real(lam, "lat lambda") if lam is not None else 1.0 / 6.0,
        lambda_chi=_real(g("lambda_chi", "0"), "lat lambda_chi"),
        layer_adjustment=_real(g("layer_adjustment", "1"), "lat layer_adjustment"),
        offset_first_layer=_real(g("offset_first_layer", "0"), "lat offset_first_layer"),
    )


def _range_mask(lat: LatticeSpec, rng: str) -> np.ndarray:
    """Sites of a range 'p1[;p2]' as a mask over the halo-padded lattice:
    coordinate c is site c (0 and n+1 are the halo layers); lowerbound = 0,
    upperbound = n+1, lastlayer = n (Lat1DFlat.cpp:315-380, Lat3D.cpp:617-700)."""
    dims = lat.dims
    n = [int(lat.n_layers[k]) for k in range(lat.gradients)]
    pts = rng.split(";")
    if len(pts) > 2 or rng.startswith("=") or rng.startswith("<"):
        raise InputError(f"range '{rng}': only one or two points are supported")

    def point(p: str) -> List[int]:
        cs = [c.strip() for c in p.split(",")]
        if len(cs) != lat.gradients:
            raise InputError(f"range '{rng}': expected {lat.gradients} coordinates per point")
        out = []
        for k, c in enumerate(cs):
            if c == "lowerbound":
                out.append(0)
            elif c == "upperbound":
                out.append(n[k] + 1)
            elif c == "lastlayer":
                out.append(n[k])
            else:
                v = _int(c, f"range '{rng}'")
                if v < 0 or v > n[k] + 1:
                    raise InputError(f"range '{rng}': coordinate {v} outside 0..{n[k] + 1}")
                out.append(v)
        return out

    a = point(pts[0])
    b = point(pts[-1])
    m = np.zeros(dims, dtype=bool)
    sl = [slice(min(a[k], b[k]), max(a[k], b[k]) + 1) for k in range(lat.gradients)]
    sl += [slice(0, 1)] * (3 - lat.gradients)
    m[tuple(sl)] = True
    return m.reshape(-1)


_SEGMENT_UNSUPPORTED = ("pinned_range", "grafted_range", "valence", "epsilon", "state1", "state2",
                        "alpha_bulk", "external_potential")
_MOL_UNSUPPORTED = ("phibulk_range", "theta_range", "n", "n_excess")


def _range_box(lat: LatticeSpec, rng: str):
    """(lo, hi) interior coordinates (1-based, inclusive, three entries) of a
    range: the sites of _range_mask, which must all be interior."""
    m = _range_mask(lat, rng).reshape(lat.dims)
    idx = np.argwhere(m)
    lo = [int(v) for v in idx.min(axis=0)][:lat.gradients]
    hi = [int(v) for v in idx.max(axis=0)][:lat.gradients]
    for k in range(lat.gradients):
        if lo[k] < 1 or hi[k] > int(lat.n_layers[k]):
            raise InputError(f"range '{rng}': a restricted range must lie inside the lattice")
    lo += [1] * (3 - len(lo))
    hi += [1] * (3 - len(hi))
    return tuple(lo[:3]), tuple(hi[:3])


def build(inp: Input, calc: int = 0):
    """(LatticeSpec, SystemSpec, solver options) of calculation `calc`."""
    vals = inp.values(calc)
    for (c, nm, p) in vals:
        if c in ("state", "reaction"):
            raise InputError(f"'{c} : {nm}': segment states and reactions are not supported")
    lats = inp.names(calc, "lat")
    if len(lats) != 1:
        raise InputError("exactly one 'lat' is required")
    lat = _lattice(vals, lats[0])
    sysn = inp.names(calc, "sys")
    approach = vals.get(("sys", sysn[0], "matrix_approach"), "first_order") if sysn else "first_order"
    if approach not in ("first_order", "second_order"):
        raise InputError(f"sys matrix_approach '{approach}' is not supported")
    lat.order = 2 if approach == "second_order" else 1

    states = inp.names(calc, "mon")
    if not states:
        raise InputError("no 'mon' segments")
    ns = len(states)
    chi = np.zeros((ns, ns))
    frozen: Dict[str, np.ndarray] = {}
    pf: List[float] = []
    for i, st in enumerate(states):
        for (c, nm, p), v in vals.items():
            if c != "mon" or nm != st:
                continue
            if p in _SEGMENT_UNSUPPORTED:
                raise InputError(f"mon {st} {p}: not supported")
            if p.startswith("chi - ") or p.startswith("chi -"):
                other = p.split("-", 1)[1].strip()
                if other not in states:
                    raise InputError(f"mon {st} {p}: unknown segment '{other}'")
                j = states.index(other)
                x = _real(v, f"mon {st} {p}")
                if chi[i, j] != 0 and chi[i, j] != x:
                    raise InputError(f"chi {st}-{other} set to two values")
                chi[i, j] = chi[j, i] = x
            elif p.startswith("Pf -"):
                pf.append(_real(v, f"mon {st} {p}"))
        fr = vals.get(("mon", st, "freedom"), "free")
        if fr == "frozen":
            rng = vals.get(("mon", st, "frozen_range"))
            if rng is None:
                raise InputError(f"mon {st}: a frozen segment needs frozen_range")
            frozen[st] = _range_mask(lat, rng).astype(float)
        elif fr != "free":
            raise InputError(f"mon {st} freedom '{fr}' is not supported")
    stiffness = 0.0
    if lat.order == 2:
        if not pf or any(abs(p - pf[0]) > 0 for p in pf):
            raise InputError("second order needs one straight-bond weight (mon : X : Pf - Y)")
        stiffness = pf[0]

    mols = []
    for m in inp.names(calc, "mol"):
        g = lambda p, d=None: vals.get(("mol", m, p), d)
        for (c, nm, p) in vals:
            if c == "mol" and nm == m and p in _MOL_UNSUPPORTED:
                raise InputError(f"mol {m} {p}: not supported")
        comp = g("composition")
        if comp is None:
            raise InputError(f"mol {m}: composition is required")
        comp = comp.replace(" ", "")
        if not re.fullmatch(r"(\([A-Za-z0-9_]+\)\d+)+", comp):
            raise InputError(f"mol {m} composition '{comp}': only linear blocks '(A)n(B)m...' are supported")
        for seg in re.findall(r"\(([A-Za-z0-9_]+)\)", comp):
            if seg not in states:
                raise InputError(f"mol {m}: segment '{seg}' has no 'mon' entry")
        fr = g("freedom", "free")
        if fr == "free":
            mols.append(MoleculeSpec(m, comp, "free", phibulk=_real(g("phibulk", "0"), f"mol {m} phibulk")))
        elif fr == "restricted":
            if g("theta") is None:
                raise InputError(f"mol {m}: a restricted molecule needs theta")
            mols.append(MoleculeSpec(m, comp, "restricted", theta=_real(g("theta"), f"mol {m} theta")))
        elif fr == "solvent":
            mols.append(MoleculeSpec(m, comp, "solvent"))
        elif fr == "range_restricted":
            # SF_Molecule.cpp:169-180
            if g("restricted_range") is None or g("phi_range") is None:
                raise InputError(f"mol {m}: a range_restricted molecule needs restricted_range and phi_range")
            lo, hi = _range_box(lat, g("restricted_range"))
            mols.append(MoleculeSpec(m, comp, "range_restricted", range_lo=lo, range_hi=hi,
                                     phi_range=_real(g("phi_range"), f"mol {m} phi_range")))
        else:
            raise InputError(f"mol {m} freedom '{fr}' is not supported")
    if not mols:
        raise InputError("no 'mol' molecules")
    sysd = SystemSpec(tuple(states), chi, mols, frozen=frozen, stiffness=stiffness)

    solver = {"method": "nk", "tolerance": 1e-7, "max_iter": 1000}
    nn = inp.names(calc, "newton")
    if nn:
        g = lambda p, d=None: vals.get(("newton", nn[0], p), d)
        meth = g("method", "").lower()
        solver["method"] = {"diis": "diis", "pseudohessian": "newton", "hessian": "newton"}.get(meth, "nk")
        solver["tolerance"] = _real(g("tolerance", "1e-7"), "newton tolerance")
        solver["max_iter"] = _int(g("iterationlimit", "1000"), "newton iterationlimit")
        if g("m") is not None and solver["method"] == "diis":
            solver["diis_m"] = _int(g("m"), "newton m")
    if solver["method"] == "nk" and any(m.freedom == "range_restricted" for m in mols):
        solver["method"] = "newton"  # no Jacobian products for range_restricted molecules
    return lat, sysd, solver


def vary(inp: Input, calc: int = 0) -> Optional[Tuple[Key, List[float], str]]:
    """The parameter a `var` block varies, its values and their type
    ('integer' or 'real') (Vary.cpp:3-230), or None. Values are stored with 15 significant digits, as Vary writes
    them back into the input (Vary.cpp:161-169)."""
    vals = inp.values(calc)
    names = inp.names(calc, "var")
    if not names:
        return None
    cls = names[0]
    own = {p: v for (c, nm, p), v in vals.items() if c == "var" and nm == cls}
    keys = [p for p in own if p not in ("type", "scale", "end_value", "output_name", "step")]
    if len(keys) != 1:
        raise InputError("var: name exactly one 'class : instance' entry to vary")
    inst, param = keys[0], own[keys[0]]
    key = (cls, inst, param)
    if key not in vals:
        raise InputError("var: only a value set in the input can be varied")
    step = _real(own.get("step", "0"), "var step")
    scale = own.get("scale", "linear")
    typ = own.get("type")
    if typ not in ("integer", "real"):
        raise InputError("var type must be integer or real")
    if scale not in ("linear", "exponential"):
        raise InputError("var scale must be linear or exponential")
    start = _real(vals[key], "var start value")
    end = _real(own.get("end_value", "nan"), "var end_value")
    if typ == "integer":
        end = float(int(end))
    maxn = 10000000
    if scale == "linear":
        if end - start < 0 and step > 0:
            step = -step
        if end - start > 0 and step < 0:
            step = -step
        if abs(end - start) < abs(step) or step == 0:
            raise InputError("var: invalid end_value / step combination")
    else:
        if abs(math.log10(abs(end)) - math.log10(abs(start))) <= 0:
            raise InputError("var: invalid end_value / step combination")
        if math.log(end) - math.log(start) < 0 and step > 0:
            step = -step
        if math.log(end) - math.log(start) > 0 and step < 0:
            step = -step
    fmt = (lambda v: float(int(v))) if typ == "integer" else (lambda v: float(f"{v:.15g}"))
    out = [start]
    cur = start
    while True:
        if typ == "integer":
            if (int(cur) - int(end) >= 0 and step > 0) or (int(cur) - int(end) <= 0 and step < 0):
                break
        elif scale == "linear" and abs(cur - end) < abs(step / maxn):
            break
        elif scale == "exponential" and abs(cur - end) < abs(cur / maxn):
            break
        old = cur
        f = (lambda c: c + step) if scale == "linear" else (lambda c: c * math.exp(math.log(10.0) / step))
        cur = f(cur)
        if typ == "integer":
            while int(cur) - int(old) == 0:
                cur = f(cur)
        if (cur > end and step > 0) or (cur < end and step < 0):
            cur = end
        cur = fmt(cur)
        out.append(cur)
        if len(out) > maxn:
            raise InputError("var: too many calculations")
    return key, out, typ


def calculations(inp: Input):
    """Every calculation of the input in order, var steps expanded: yields
    (calc, var value or None, LatticeSpec, SystemSpec, solver options)."""
    for c in range(len(inp)):
        v = vary(inp, c)
        if v is None:
            yield (c, None) + build(inp, c)
            continue
        key, values, typ = v
        for x in values:
            sub = Input("")
            sub.blocks = [list(b) for b in inp.blocks[: c + 1]]
            text = str(int(x)) if typ == "integer" else repr(x)
            sub.blocks[c] = [l for l in sub.blocks[c] if l[:3] != key] + [key + (text,)]
            yield (c, x) + build(sub, c)


def output_names(inp: Input, calc: int = 0) -> Dict[str, object]:
    """Names the output lines of a calculation carry: its sys and newton
    names ('noname' when absent) and its molecules."""
    sysn = inp.names(calc, "sys")
    newn = inp.names(calc, "newton")
    return {"sys": sysn[0] if sysn else "noname", "newton": newn[0] if newn else "noname",
            "mol": inp.names(calc, "mol")}


def run(path_or_text: str, kal: Optional[str] = None, kal_template: Optional[str] = None, **solver_overrides):
    """Solve every calculation of an SFbox input on the GPU. A calculation
    starts from the previous solution when the iteration vector keeps its
    size, as the reference's default initial guess does, and from u = 0
    otherwise. Returns one dict per calculation: var value, convergence,
    iterations, residual norm, free energy, grand potential, per molecule
    ln G_N, phi_b and the chemical potential, and the solution x with the
    lattice and system it belongs to. With `kal`, the scalar results also go
    to that file as a kal table (field_amd.output.write_kal), one row per
    calculation, filtered by `kal_template` (input syntax) when given."""
    from . import sfx
    text = open(path_or_text).read() if os.path.exists(path_or_text) else path_or_text
    inp = Input(text)
    if kal is not None:
        from .output import write_kal
    results, x = [], None
    for c, v, lat, sysd, solver in calculations(inp):
        opts = {**solver, **solver_overrides}
        method = opts.pop("method")
        G = sfx.System(lat, sysd)
        x0 = x if x is not None and len(x) == G.nvar else np.zeros(G.nvar)
        x, st = G.solve(x0, method=method, check_converged=False, **opts)
        G.residual(x, precision="fp64")
        nm = len(sysd.molecules)
        results.append({
            "calculation": c, "var": v, "x": np.asarray(x).copy(), "lattice": lat, "system": sysd,
            "converged": bool(st["converged"]), "iterations": st["iterations"],
            "residual_norm": st["residual_norm"], "free_energy": G.free_energy(),
            "grand_potential": G.grand_potential(),
            "lngn": [G.lngn(m) for m in range(nm)], "phibulk": [G.phibulk(m) for m in range(nm)],
            "chemical_potential": [G.chemical_potential(m) for m in range(nm)],
        })
        G.close()
        if kal is not None:  # one row per calculation, appended as it finishes
            write_kal(kal, results[-1:], output_names(inp, c), kal_template, append=len(results) > 1)
    return results
