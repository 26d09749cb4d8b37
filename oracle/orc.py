"""ORACLE binding (test infrastructure only).

ctypes access to liborc.so, the FP64 CPU restatement of the reference's SCF
hot path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module: as the checker, or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")
_lib = None

D = C.c_double
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)
PF = C.POINTER(C.c_float)
RESID_FN = C.CFUNCTYPE(None, C.c_void_p, PD, PD)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_sys_new.restype = C.c_void_p
        L.orc_sys_new.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_sys_free.argtypes = [C.c_void_p]
        L.orc_sys_nvar.restype = C.c_long
        L.orc_sys_nvar.argtypes = [C.c_void_p]
        L.orc_sys_M.restype = C.c_long
        L.orc_sys_M.argtypes = [C.c_void_p]
        L.orc_sys_residual.argtypes = [C.c_void_p, PD, PD]
        L.orc_sys_compute_phi.argtypes = [C.c_void_p, PD]
        L.orc_sys_get_phi.argtypes = [C.c_void_p, C.c_int, PD]
        L.orc_sys_get_swf.argtypes = [C.c_void_p, C.c_int, PD]
        L.orc_sys_lngn.restype = D
        L.orc_sys_lngn.argtypes = [C.c_void_p, C.c_int]
        L.orc_sys_phibulk.restype = D
        L.orc_sys_phibulk.argtypes = [C.c_void_p, C.c_int]
        L.orc_sys_free_energy.restype = D
        L.orc_sys_free_energy.argtypes = [C.c_void_p]
        L.orc_sys_chemical_potential.restype = D
        L.orc_sys_chemical_potential.argtypes = [C.c_void_p, C.c_int]
        L.orc_sys_grand_potential.restype = D
        L.orc_sys_grand_potential.argtypes = [C.c_void_p]
        L.orc_sys_solve_diis.argtypes = [C.c_void_p, PD, D, C.c_int, C.c_int, D, PI, PD]
        L.orc_mps_new.restype = C.c_void_p
        L.orc_mps_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D, D, D, D]
        L.orc_mps_free.argtypes = [C.c_void_p]
        L.orc_mps_M.restype = C.c_long
        L.orc_mps_M.argtypes = [C.c_void_p]
        L.orc_mps_compute_g.argtypes = [C.c_void_p, PD, PD]
        L.orc_mps_get_phi.argtypes = [C.c_void_p, PD]
        L.orc_mps_gn.restype = D
        L.orc_mps_gn.argtypes = [C.c_void_p]
        L.orc_mps_helmholtz.restype = D
        L.orc_mps_helmholtz.argtypes = [C.c_void_p, PD]
        L.orc_mps_scf.argtypes = [C.c_void_p, PD, D, C.c_int, C.c_int, D, PI, PD]
        L.orc_lat_new.restype = C.c_void_p
        L.orc_lat_new.argtypes = [C.c_void_p]
        L.orc_lat_free.argtypes = [C.c_void_p]
        L.orc_lat_M.restype = C.c_long
        L.orc_lat_M.argtypes = [C.c_void_p]
        L.orc_setb.argtypes = [C.c_void_p, PD]
        L.orc_propagate.argtypes = [C.c_void_p, PD, PD, PD]
        L.orc_lngn.restype = D
        L.orc_lngn.argtypes = [C.c_void_p, PD]
        L.orc_side.argtypes = [C.c_void_p, PD, PD]
        L.orc_subtract_boundaries.argtypes = [C.c_void_p, PD]
        L.orc_diis_coeffs.argtypes = [PD, C.c_int, PD]
        L.orc_propagate2.argtypes = [C.c_void_p, PD, PD, PD, D, C.c_int]
        L.orc_ncomp2.argtypes = [C.c_void_p]
        L.orc_set_cyl2d_split_outward.argtypes = [C.c_int]
        L.orc_lngn2.restype = D
        L.orc_lngn2.argtypes = [C.c_void_p, PD]
        L.orc_connect2.argtypes = [C.c_void_p, PD, PD, PD, PD]
        L.orc_correct2.argtypes = [C.c_void_p, PD, PD, PD]
        L.orc_propagate_force.argtypes = [C.c_void_p, PD, PD, PD, D]
        L.orc_norm_phi_free.argtypes = [C.c_void_p, PD, PD, D]
        L.orc_norm_phi_restr.argtypes = [C.c_void_p, PD, PD, PD, D]
        L.orc_newton_defaults.argtypes = [C.c_void_p]
        L.orc_newton.argtypes = [RESID_FN, C.c_void_p, PD, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_newton_sys.argtypes = [C.c_void_p, PD, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_ldu_multiply.argtypes = [PF, C.c_int, PD, D, PD]
        L.orc_ldu_sign.argtypes = [PF, C.c_int]
        L.orc_ldu_update.argtypes = [PF, C.c_int, PD, PD, D]
        L.orc_ldu_solve.argtypes = [PF, C.c_int, PD, PD]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(PD)


class NewtonOpts(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("iterationlimit", C.c_int), ("deltamax", C.c_double),
                ("deltamin", C.c_double), ("linesearchlimit", C.c_int),
                ("reset_hessian_criterion", C.c_double), ("small_alpha", C.c_double),
                ("max_n_small_alpha", C.c_int)]


class NewtonStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("functioncalls", C.c_int), ("hessian_resets", C.c_int),
                ("reversed_directions", C.c_int), ("time_dependent", C.c_int), ("converged", C.c_int),
                ("accuracy", C.c_double), ("residual_norm", C.c_double)]


def _newton_opts(**kw):
    o = NewtonOpts()
    lib().orc_newton_defaults(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _stats_dict(st):
    return {k: getattr(st, k) for k, _ in NewtonStats._fields_}


def newton(fun, x, **kw):
    """SFNewton (pseudo-Hessian Newton, Scheutjens line search) on g = fun(x);
    returns (x, stats)."""
    x = _f64(x).copy()
    n = x.size

    def cb(_ctx, xp, gp):
        xv = np.ctypeslib.as_array(xp, shape=(n,))
        np.ctypeslib.as_array(gp, shape=(n,))[:] = fun(xv.copy())

    fn = RESID_FN(cb)
    st = NewtonStats()
    rc = lib().orc_newton(fn, None, _p(x), n, C.byref(_newton_opts(**kw)), C.byref(st))
    if rc:
        raise ValueError(f"orc_newton returned {rc}")
    return x, _stats_dict(st)


def ldu_identity(n, diag=1.0):
    h = np.zeros((n, n), dtype=np.float32)
    np.fill_diagonal(h, diag)
    return h


def ldu_dense(h):
    """The matrix L D U that a factorisation array h stands for (h[r + n c] is
    element (r, c): numpy's column-major view, h.T in row-major terms)."""
    a = np.asarray(h, dtype=np.float64).T
    L = np.tril(a, -1) + np.eye(a.shape[0])
    U = np.triu(a, 1) + np.eye(a.shape[0])
    return L @ np.diag(np.diag(a)) @ U


def _pf(a):
    return a.ctypes.data_as(PF)


def ldu_update(h, w, v, alpha):
    w, v = _f64(w).copy(), _f64(v).copy()
    lib().orc_ldu_update(_pf(h), h.shape[0], _p(w), _p(v), alpha)


def ldu_solve(h, g):
    p = np.zeros(h.shape[0])
    lib().orc_ldu_solve(_pf(h), h.shape[0], _p(_f64(g)), _p(p))
    return p


def ldu_multiply(h, w, alpha=1.0):
    v = np.zeros(h.shape[0])
    lib().orc_ldu_multiply(_pf(h), h.shape[0], _p(_f64(w)), alpha, _p(v))
    return v


def ldu_sign(h):
    return lib().orc_ldu_sign(_pf(h), h.shape[0])


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Lattice:
    def __init__(self, spec):
        self.spec = spec
        self._d = spec.desc()
        self.h = lib().orc_lat_new(C.byref(self._d))
        self.M = lib().orc_lat_M(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_lat_free(self.h)

    def propagate(self, Gin, g):
        Gin = _f64(Gin).copy()
        out = np.zeros(self.M)
        lib().orc_propagate(self.h, _p(Gin), _p(_f64(g)), _p(out))
        return out

    def set_boundaries(self, P):
        P = _f64(P).copy()
        lib().orc_setb(self.h, _p(P))
        return P

    def lngn(self, G):
        return lib().orc_lngn(self.h, _p(_f64(G)))

    def propagate2(self, Gin, g, S, backward=False):
        """One second-order step over the lattice's lumped components
        (ncomp2 blocks of M); halos of the input are set first."""
        Gin = _f64(Gin).copy()
        out = np.zeros_like(Gin)
        lib().orc_propagate2(self.h, _p(Gin), _p(_f64(g)), _p(out), S, int(backward))
        return out

    def lngn2(self, G):
        return lib().orc_lngn2(self.h, _p(_f64(G)))

    def propagate_force(self, Gin, g, f):
        Gin = _f64(Gin).copy()
        out = np.zeros(self.M)
        lib().orc_propagate_force(self.h, _p(Gin), _p(_f64(g)), _p(out), float(f))
        return out

    def connect2(self, A, B, out, OUT=None):
        """(out, OUT) after ConnectG with PHI; OUT None: no per-axis parts."""
        out = _f64(out).copy()
        OUT = None if OUT is None else _f64(OUT).copy()
        lib().orc_connect2(self.h, _p(_f64(A)), _p(_f64(B)), _p(out), None if OUT is None else _p(OUT))
        return out, OUT

    def correct2(self, v, V, G):
        v = _f64(v).copy()
        V = None if V is None else _f64(V).copy()
        lib().orc_correct2(self.h, _p(v), None if V is None else _p(V), _p(_f64(G)))
        return v, V

    def norm_phi_free(self, phi, PHI, C):
        phi = _f64(phi).copy()
        PHI = None if PHI is None else _f64(PHI).copy()
        lib().orc_norm_phi_free(self.h, _p(phi), None if PHI is None else _p(PHI), float(C))
        return phi, PHI

    def norm_phi_restr(self, phi, PHI, Gi, C):
        phi = _f64(phi).copy()
        PHI = None if PHI is None else _f64(PHI).copy()
        lib().orc_norm_phi_restr(self.h, _p(phi), None if PHI is None else _p(PHI), _p(_f64(Gi)), float(C))
        return phi, PHI

    def ncomp2(self):
        return lib().orc_ncomp2(self.h)

    def sides(self, P):
        P = _f64(P).copy()
        out = np.zeros(self.M)
        lib().orc_side(self.h, _p(P), _p(out))
        return out


class System:
    """Oracle SCF system built from field_amd.spec descriptions."""

    def __init__(self, lat_spec, sys_spec):
        self.lat, self.sys = lat_spec, sys_spec
        self._ld = lat_spec.desc()
        self._sd, self._keep = sys_spec.desc(lat_spec)
        self.h = lib().orc_sys_new(C.byref(self._ld), C.byref(self._sd))
        if not self.h:
            raise ValueError("oracle rejected the system description")
        self.nvar = lib().orc_sys_nvar(self.h)
        self.M = lib().orc_sys_M(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sys_free(self.h)

    def residual(self, x):
        x = _f64(x)
        f = np.zeros(self.nvar)
        rc = lib().orc_sys_residual(self.h, _p(x), _p(f))
        if rc:
            raise FloatingPointError(f"oracle residual returned {rc}")
        return f

    def compute_phi(self, x):
        lib().orc_sys_compute_phi(self.h, _p(_f64(x)))

    def phi(self, state):
        out = np.zeros(self.M)
        lib().orc_sys_get_phi(self.h, state, _p(out))
        return out

    def swf(self, state):
        out = np.zeros(self.M)
        lib().orc_sys_get_swf(self.h, state, _p(out))
        return out

    def lngn(self, mol):
        return lib().orc_sys_lngn(self.h, mol)

    def phibulk(self, mol):
        return lib().orc_sys_phibulk(self.h, mol)

    def free_energy(self):
        return lib().orc_sys_free_energy(self.h)

    def chemical_potential(self, mol):
        return lib().orc_sys_chemical_potential(self.h, mol)

    def grand_potential(self):
        return lib().orc_sys_grand_potential(self.h)

    def solve_newton(self, x, **kw):
        """SFNewton's default iteration from x; returns (x, stats)."""
        x = _f64(x).copy()
        st = NewtonStats()
        rc = lib().orc_newton_sys(self.h, _p(x), x.size, C.byref(_newton_opts(**kw)), C.byref(st))
        if rc:
            raise ValueError(f"orc_newton_sys returned {rc}")
        return x, _stats_dict(st)

    def solve_diis(self, x, tol=1e-7, max_iter=1000, m=10, eta=0.1):
        x = _f64(x).copy()
        it = C.c_int(0)
        err = C.c_double(0)
        rc = lib().orc_sys_solve_diis(self.h, _p(x), tol, max_iter, m, eta, C.byref(it), C.byref(err))
        return x, it.value, err.value, rc


class MPS:
    """Literal restatement of MPS.cpp (3-D periodic AB melt)."""

    def __init__(self, n, n1, n2, chi, n_chains, phib_a=0.0, phib_b=0.0):
        self.h = lib().orc_mps_new(n, n, n, n1, n2, chi, n_chains, phib_a, phib_b)
        self.M = lib().orc_mps_M(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mps_free(self.h)

    def residual(self, x):
        g = np.zeros(2 * self.M)
        lib().orc_mps_compute_g(self.h, _p(_f64(x)), _p(g))
        return g

    def phi(self):
        out = np.zeros(2 * self.M)
        lib().orc_mps_get_phi(self.h, _p(out))
        return out

    def gn(self):
        return lib().orc_mps_gn(self.h)

    def helmholtz(self, x):
        return lib().orc_mps_helmholtz(self.h, _p(_f64(x)))

    def scf(self, x, tol=1e-7, max_iter=1000, m=10, eta=0.1):
        x = _f64(x).copy()
        it = C.c_int(0)
        err = C.c_double(0)
        rc = lib().orc_mps_scf(self.h, _p(x), tol, max_iter, m, eta, C.byref(it), C.byref(err))
        return x, it.value, err.value, rc


def set_cyl2d_split_outward(on):
    """Tests only: make the 2-D cylinder's second-order backward sweep split
    its outward radial step as the forward sweep does."""
    lib().orc_set_cyl2d_split_outward(int(bool(on)))


def diis_coeffs(A):
    A = _f64(A)
    n = A.shape[0]
    X = np.zeros(n)
    lib().orc_diis_coeffs(_p(A), n, _p(X))
    return X
