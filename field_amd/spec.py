"""Problem descriptions: lattice, molecules, system -> C ABI descriptors.

The names follow the reference's input file vocabulary
(`lat : name : n_layers`, `mol : name : composition : (A)12(B)28`,
`mon : A : chi - B : 0.4`, freedoms free / restricted / solvent, Input.h:12-132).
"""
from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _abi


@dataclass
class LatticeSpec:
    gradients: int
    n_layers: Sequence[int]
    geometry: str = "flat"
    lattice_type: str = "standard"
    order: int = 1
    lowerbound: Sequence[str] = ("mirror1", "mirror1", "mirror1")
    upperbound: Sequence[str] = ("mirror1", "mirror1", "mirror1")
    lambda_: float = 1.0 / 6.0
    lambda_chi: float = 0.0
    layer_adjustment: float = 1.0
    offset_first_layer: float = 0.0

    @property
    def dims(self):
        n = list(self.n_layers) + [0, 0, 0]
        MX = n[0] + 2
        MY = n[1] + 2 if self.gradients >= 2 else 1
        MZ = n[2] + 2 if self.gradients >= 3 else 1
        return MX, MY, MZ

    @property
    def M(self) -> int:
        MX, MY, MZ = self.dims
        return MX * MY * MZ

    def interior_mask(self) -> np.ndarray:
        MX, MY, MZ = self.dims
        m = np.zeros((MX, MY, MZ), dtype=bool)
        sx = slice(1, MX - 1)
        sy = slice(1, MY - 1) if self.gradients >= 2 else slice(0, 1)
        sz = slice(1, MZ - 1) if self.gradients >= 3 else slice(0, 1)
        m[sx, sy, sz] = True
        return m.reshape(-1)

    def desc(self) -> _abi.LatticeDesc:
        d = _abi.LatticeDesc()
        d.gradients = self.gradients
        d.geometry = _abi._GEOM[self.geometry]
        d.lattice_type = _abi._LATTYPE[self.lattice_type.lower()]
        d.order = self.order
        for k in range(3):
            d.n_layers[k] = int(self.n_layers[k]) if k < len(self.n_layers) else 0
            d.lowerbound[k] = _abi._BOUNDS[self.lowerbound[k]] if k < len(self.lowerbound) else 0
            d.upperbound[k] = _abi._BOUNDS[self.upperbound[k]] if k < len(self.upperbound) else 0
        d.lambda_ = self.lambda_
        d.lambda_chi = self.lambda_chi
        d.layer_adjustment = self.layer_adjustment
        d.offset_first_layer = self.offset_first_layer
        return d


def parse_composition(comp: str, states: Sequence[str]) -> List[int]:
    """'(A)12(B)28' -> state indices per rank (SF_SegmentBlock syntax subset)."""
    seq: List[int] = []
    for name, count in re.findall(r"\(([^()]+)\)(\d+)", comp.replace(" ", "")):
        seq += [list(states).index(name)] * int(count)
    if not seq:
        raise ValueError(f"cannot parse composition {comp!r}")
    return seq


@dataclass
class MoleculeSpec:
    name: str
    composition: str
    freedom: str = "free"          # free / restricted / solvent / range_restricted
    phibulk: float = 0.0
    theta: float = 0.0
    # range_restricted: sum of phi over the box lo..hi (1-based, inclusive) is phi_range
    range_lo: Sequence[int] = (1, 1, 1)
    range_hi: Sequence[int] = (1, 1, 1)
    phi_range: float = 0.0


@dataclass
class SystemSpec:
    states: Sequence[str]
    chi: np.ndarray
    molecules: List[MoleculeSpec]
    frozen: Dict[str, np.ndarray] = field(default_factory=dict)   # state -> profile (M)
    form: str = "sfbox"            # sfbox / melt
    melt_phibulk: Optional[Sequence[float]] = None
    stiffness: float = 0.0

    def sequences(self) -> List[List[int]]:
        return [parse_composition(m.composition, self.states) for m in self.molecules]

    def desc(self, lat: LatticeSpec):
        """Returns (SystemDesc, keepalive)."""
        ns = len(self.states)
        chi = np.ascontiguousarray(np.asarray(self.chi, dtype=np.float64).reshape(ns, ns))
        frozen = np.zeros(ns, dtype=np.int32)
        prof = np.zeros((ns, lat.M), dtype=np.float64)
        for name, p in self.frozen.items():
            i = list(self.states).index(name)
            frozen[i] = 1
            prof[i] = np.asarray(p, dtype=np.float64).reshape(-1)
        seqs = [np.ascontiguousarray(np.asarray(s, dtype=np.int32)) for s in self.sequences()]
        mols = (_abi.MoleculeDesc * len(self.molecules))()
        fmap = {"free": _abi.FREE, "restricted": _abi.RESTRICTED, "solvent": _abi.SOLVENT,
                "range_restricted": _abi.RANGE_RESTRICTED}
        for k, (m, s) in enumerate(zip(self.molecules, seqs)):
            mols[k].length = len(s)
            mols[k].sequence = s.ctypes.data_as(C.POINTER(C.c_int))
            mols[k].freedom = fmap[m.freedom]
            mols[k].phibulk = m.phibulk
            mols[k].theta = m.theta
            for a in range(3):
                mols[k].range_lo[a] = int(m.range_lo[a]) if a < len(m.range_lo) else 1
                mols[k].range_hi[a] = int(m.range_hi[a]) if a < len(m.range_hi) else 1
            mols[k].phi_range = m.phi_range
        d = _abi.SystemDesc()
        d.n_states = ns
        d.chi = chi.ctypes.data_as(C.POINTER(C.c_double))
        d.frozen = frozen.ctypes.data_as(C.POINTER(C.c_int))
        d.frozen_profile = prof.ctypes.data_as(C.POINTER(C.c_double)) if self.frozen else None
        d.n_molecules = len(self.molecules)
        d.molecules = C.cast(mols, C.POINTER(_abi.MoleculeDesc))
        d.stiffness = self.stiffness
        d.form = _abi.FORM_MELT if self.form == "melt" else _abi.FORM_SFBOX
        mpb = np.zeros(ns) if self.melt_phibulk is None else np.asarray(self.melt_phibulk, dtype=np.float64)
        mpb = np.ascontiguousarray(mpb)
        d.melt_phibulk = mpb.ctypes.data_as(C.POINTER(C.c_double))
        keep = (chi, frozen, prof, seqs, mols, mpb)
        return d, keep

    # ---- iteration-variable bookkeeping (SF_Solve.cpp:39-163) ----
    def itvar_groups(self) -> List[int]:
        """state -> itvar index (-1 frozen): liquid states with equal chi rows share one."""
        ns = len(self.states)
        chi = np.asarray(self.chi, dtype=np.float64).reshape(ns, ns)
        out, reps = [], []
        for i, s in enumerate(self.states):
            if s in self.frozen:
                out.append(-1)
                continue
            hit = -1
            for j in reps:
                if np.array_equal(chi[i], chi[j]):
                    hit = out[j]
                    break
            if hit < 0:
                hit = len(reps)
                reps.append(i)
            out.append(hit)
        return out

    def n_itvar(self) -> int:
        return max(self.itvar_groups()) + 1

    def itvar_sites(self, lat: LatticeSpec) -> np.ndarray:
        m = lat.interior_mask().copy()
        for p in self.frozen.values():
            m &= ~(np.asarray(p).reshape(-1) > 0.999)
        return np.nonzero(m)[0]

    def nvar(self, lat: LatticeSpec) -> int:
        if self.form == "melt":
            return self.n_itvar() * lat.M
        return len(self.itvar_sites(lat)) * self.n_itvar()
