"""ctypes binding of libsfx.so (the C ABI in include/sfx.h).

The product path: every call goes to the HIP library. If the library is not
built this module raises at load time; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from .spec import LatticeSpec, SystemSpec

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsfx.so")
_lib = None

PD = C.POINTER(C.c_double)

# every entry point of include/sfx.h with its ctypes signature
_SIGS = {
    "sfx_abi_version": (C.c_int, []),
    "sfx_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "sfx_set_device": (C.c_int, [C.c_int]),
    "sfx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "sfx_system_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sfx_system_destroy": (C.c_int, [C.c_void_p]),
    "sfx_system_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "sfx_compute_phi": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    "sfx_residual": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    "sfx_jvp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    "sfx_vjp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    "sfx_get_phi": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "sfx_get_lngn": (C.c_int, [C.c_void_p, C.c_int, PD]),
    "sfx_get_phibulk": (C.c_int, [C.c_void_p, C.c_int, PD]),
    "sfx_free_energy": (C.c_int, [C.c_void_p, PD]),
    "sfx_grand_potential": (C.c_int, [C.c_void_p, PD]),
    "sfx_chemical_potential": (C.c_int, [C.c_void_p, C.c_int, PD]),
    "sfx_set_chi": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
    "sfx_stream_set": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sfx_solver_defaults": (C.c_int, [C.c_void_p]),
    "sfx_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "sfx_newton_iterate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]),
    "sfx_precondition": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "sfx_profile_start": (C.c_int, [C.c_void_p]),
    "sfx_profile_start_select": (C.c_int, [C.c_void_p, C.c_uint]),
    "sfx_profile_stop": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sfx_batch_create": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sfx_batch_destroy": (C.c_int, [C.c_void_p]),
    "sfx_batch_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.c_void_p]),
    "sfx_batch_tuning": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_void_p]),
    "sfx_batch_residual": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "sfx_batch_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "sfx_batch_get_lngn": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sfx_batch_get_phi": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "sfx_lattice_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "sfx_lattice_destroy": (C.c_int, [C.c_void_p]),
    "sfx_lattice_sites": (C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    "sfx_lattice_propagate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfx_lattice_connect": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfx_lattice_correct_double_count": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfx_lattice_lngn": (C.c_int, [C.c_void_p, C.c_void_p, PD, C.c_void_p]),
    "sfx_lattice_set_boundaries": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfx_lattice_lngn_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "sfx_lattice_propagate_force": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                              C.c_void_p]),
    "sfx_lattice_norm_phi_free": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "sfx_lattice_norm_phi_restr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                             C.c_void_p]),
    "sfx_lattice_propagate2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                         C.c_void_p]),
    "sfx_lattice_connect2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfx_lattice_correct_double_count2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                    C.c_void_p]),
    "sfx_lattice_sides": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
}


class SFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sfx error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first (python -m field_amd.build)")
        # torch ships its own HIP runtime; loading it first makes the library
        # bind to that same runtime, so device pointers and streams are shared
        import torch  # noqa: F401

        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().sfx_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int):
    if rc != 0:
        raise SFError(rc, last_error())


def device_count() -> int:
    n = C.c_int(0)
    check(lib().sfx_device_count(C.byref(n)))
    return n.value


def _ptr(a):
    """(pointer, where) for a numpy array or a device tensor."""
    if isinstance(a, np.ndarray):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return C.c_void_p(a.ctypes.data), _abi.HOST
    # torch tensor: the library runs on its own stream, so work torch queued
    # on the tensor's stream must be finished before the library reads it
    assert a.dtype.__str__() == "torch.float64" and a.is_contiguous()
    if a.is_cuda:
        import torch

        torch.cuda.current_stream(a.device).synchronize()
    return C.c_void_p(a.data_ptr()), (_abi.DEVICE if a.is_cuda else _abi.HOST)


_PREC = {"fp64": _abi.FP64, "fp32": _abi.FP32, None: None}


RESIDUAL_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def newton_iterate(fun, x0, check_converged=True, **kw):
    """SFNewton's iteration (sfx_newton_iterate) on g = fun(x), host only:
    returns (x, stats). kw sets sfx_solver_opts fields (tolerance, max_iter,
    newton_deltamax, newton_linesearch_limit, newton_reset_criterion)."""
    o = _abi.SolverOpts()
    check(lib().sfx_solver_defaults(C.byref(o)))
    o.method = _abi.SOLVER_NEWTON
    for k, v in kw.items():
        setattr(o, k, v)
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    n = x.size

    def cb(_ctx, xp, gp):
        xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        np.ctypeslib.as_array(gp, shape=(n,))[:] = fun(xv)

    fn = RESIDUAL_FN(cb)
    st = _abi.SolveStats()
    rc = lib().sfx_newton_iterate(fn, None, x.ctypes.data_as(C.c_void_p), n, C.byref(o), C.byref(st))
    stats = {k: getattr(st, k) for k, _ in _abi.SolveStats._fields_}
    if rc != 0 and not (rc == _abi.SFX_E_NOTCONVERGED and not check_converged):
        raise SFError(rc, last_error())
    return x, stats


class System:
    """One SCF system on the GPU (sfx_system_create)."""

    _SWEEP = {"auto": _abi.SWEEP_AUTO, "per_rank": _abi.SWEEP_PER_RANK, "blocked": _abi.SWEEP_BLOCKED}
    _SCHEDULE = {"auto": _abi.SCHEDULE_AUTO, "sequential": _abi.SCHEDULE_SEQUENTIAL}
    _TRANSFER = {"auto": _abi.TRANSFER_AUTO, "rows": _abi.TRANSFER_ROWS, "mfma": _abi.TRANSFER_MFMA}

    def __init__(self, lat: LatticeSpec, sysd: SystemSpec, precision: str = "fp64",
                 memory_budget: int = 0, sweep_kernel: str = "auto", chain_schedule: str = "auto",
                 transfer_kernel: str = "auto"):
        self.lat, self.sys = lat, sysd
        self._ld = lat.desc()
        self._sd, self._keep = sysd.desc(lat)
        opt = _abi.Options()
        opt.precision = _PREC[precision]
        opt.memory_budget = memory_budget
        opt.sweep_kernel = self._SWEEP[sweep_kernel]
        opt.chain_schedule = self._SCHEDULE[chain_schedule]
        opt.transfer_kernel = self._TRANSFER[transfer_kernel]
        self.precision = opt.precision
        h = C.c_void_p()
        check(lib().sfx_system_create(C.byref(self._ld), C.byref(self._sd), C.byref(opt), C.byref(h)))
        self.h = h
        nv, M = C.c_long(), C.c_long()
        check(lib().sfx_system_sizes(self.h, C.byref(nv), C.byref(M)))
        self.nvar, self.M = nv.value, M.value

    def close(self):
        # at interpreter exit the module's library handle may already be gone
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.sfx_system_destroy(h)
        self.h = None

    __del__ = close

    def _p(self, precision):
        return self.precision if precision is None else _PREC[precision]

    def _vec(self, x):
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float64)
            assert x.size == self.nvar, (x.size, self.nvar)
        return x

    def residual(self, x, precision=None, out=None):
        x = self._vec(x)
        if out is None:
            out = np.zeros(self.nvar) if isinstance(x, np.ndarray) else x.new_zeros(self.nvar)
        px, where = _ptr(x)
        po, _ = _ptr(out)
        check(lib().sfx_residual(self.h, px, po, where, self._p(precision)))
        return out

    def precondition(self, r):
        """The Newton-Krylov preconditioner applied to r (sfx_precondition)."""
        r = self._vec(r)
        out = np.zeros(self.nvar) if isinstance(r, np.ndarray) else r.new_zeros(self.nvar)
        pr, where = _ptr(r)
        check(lib().sfx_precondition(self.h, pr, _ptr(out)[0], where))
        return out

    def compute_phi(self, x, precision=None):
        x = self._vec(x)
        px, where = _ptr(x)
        check(lib().sfx_compute_phi(self.h, px, where, self._p(precision)))

    def jvp(self, x, v, precision=None):
        x, v = self._vec(x), self._vec(v)
        out = np.zeros(self.nvar) if isinstance(x, np.ndarray) else x.new_zeros(self.nvar)
        px, where = _ptr(x)
        check(lib().sfx_jvp(self.h, px, _ptr(v)[0], _ptr(out)[0], where, self._p(precision)))
        return out

    def vjp(self, x, w, precision=None):
        x, w = self._vec(x), self._vec(w)
        out = np.zeros(self.nvar) if isinstance(x, np.ndarray) else x.new_zeros(self.nvar)
        px, where = _ptr(x)
        check(lib().sfx_vjp(self.h, px, _ptr(w)[0], _ptr(out)[0], where, self._p(precision)))
        return out

    def phi(self, state: int) -> np.ndarray:
        out = np.zeros(self.M)
        check(lib().sfx_get_phi(self.h, state, C.c_void_p(out.ctypes.data), _abi.HOST))
        return out

    def lngn(self, mol: int) -> float:
        v = C.c_double()
        check(lib().sfx_get_lngn(self.h, mol, C.byref(v)))
        return v.value

    def phibulk(self, mol: int) -> float:
        v = C.c_double()
        check(lib().sfx_get_phibulk(self.h, mol, C.byref(v)))
        return v.value

    def free_energy(self) -> float:
        v = C.c_double()
        check(lib().sfx_free_energy(self.h, C.byref(v)))
        return v.value

    def grand_potential(self) -> float:
        v = C.c_double()
        check(lib().sfx_grand_potential(self.h, C.byref(v)))
        return v.value

    def chemical_potential(self, mol: int) -> float:
        v = C.c_double()
        check(lib().sfx_chemical_potential(self.h, mol, C.byref(v)))
        return v.value

    def set_chi(self, a: int, b: int, value: float):
        """chi(a,b) = chi(b,a) = value from here on (one step of a Vary.cpp
        sweep over a chi parameter; the caller re-solves from its last x)."""
        check(lib().sfx_set_chi(self.h, int(a), int(b), float(value)))

    def set_stream(self, stream=None):
        """Run the following calls on `stream` (a torch.cuda.Stream or a raw
        hipStream_t as int; None: the system's own stream), after what this
        system queued before (sfx_stream_set). Keep the stream alive while
        the system uses it."""
        if stream is None:
            ptr = None
        elif isinstance(stream, int):
            ptr = stream
        else:
            ptr = stream.cuda_stream
        self._stream = stream
        check(lib().sfx_stream_set(self.h, C.c_void_p(ptr)))

    def profile_start(self, only=None):
        """Time propagator launches with HIP events until profile_stop; `only`:
        a list of (kernel kind, "fp32"/"fp64") pairs to time, the rest untimed."""
        if only is None:
            check(lib().sfx_profile_start(self.h))
            return
        sel = 0
        for kind, prec in only:
            sel |= 1 << (2 * kind + (1 if prec == "fp32" else 0))
        check(lib().sfx_profile_start_select(self.h, C.c_uint(sel)))

    def profile_stop(self):
        """Per kernel kind and precision: launches, summed HIP-event ms, bytes per launch."""
        rep = _abi.ProfileReport()
        check(lib().sfx_profile_stop(self.h, C.byref(rep)))
        out = []
        for e in rep.entries[: rep.count]:
            if e.launches:
                out.append({"kernel": _abi.KERNEL_NAMES[e.kind], "kind": e.kind,
                            "precision": "fp32" if e.precision == _abi.FP32 else "fp64",
                            "launches": e.launches, "ranks": e.ranks, "total_ms": e.total_ms,
                            "avg_us": 1e3 * e.total_ms / e.launches, "bytes_per_launch": e.bytes_per_launch})
        return out

    def solve(self, x0, method="diis", check_converged=True, **kw):
        o = _abi.SolverOpts()
        check(lib().sfx_solver_defaults(C.byref(o)))
        o.method = {"diis": _abi.SOLVER_DIIS, "nk": _abi.SOLVER_NK, "newton": _abi.SOLVER_NEWTON}[method]
        for k, v in kw.items():
            if k == "jvp_precision":
                v = _PREC[v]
            setattr(o, k, v)
        x = self._vec(x0)
        x = x.copy() if isinstance(x, np.ndarray) else x.clone()
        st = _abi.SolveStats()
        px, where = _ptr(x)
        rc = lib().sfx_solve(self.h, C.byref(o), px, where, C.byref(st))
        stats = {k: getattr(st, k) for k, _ in _abi.SolveStats._fields_}
        if rc != 0 and not (rc == _abi.SFX_E_NOTCONVERGED and not check_converged):
            raise SFError(rc, last_error())
        return x, stats

    def solve_chi_sweep(self, x0, a, b, values, method="nk", check_converged=True, continuation=None, **kw):
        """A `var` sweep over chi(a,b) (Vary.cpp:3-80, SF_Box.cpp:145-176):
        for each value in turn set chi and solve from the previous step's
        solution; the system keeps the last value. `continuation` (a dict of
        solver options) overrides kw for the warm-started steps after the
        first, e.g. plain Newton (ptc_dt0=0) from a neighbouring solution.
        Returns (x, [stats])."""
        x, out = x0, []
        for i, v in enumerate(values):
            self.set_chi(a, b, v)
            last = i == len(values) - 1
            opts = {**kw, **(continuation or {})} if i > 0 else kw
            x, st = self.solve(x, method=method, check_converged=check_converged and last, **opts)
            st["chi"] = float(v)
            out.append(st)
            if not st["converged"]:
                break
        return x, out


class Batch:
    """Many 1-D adsorption problems solved at once (sfx_batch_create).

    `problems` is a sequence of dicts with keys n_layers, N, chi_as, chi_s,
    phibulk and optionally lambda_ (default 1/6)."""

    def __init__(self, problems):
        self.problems = [dict(p) for p in problems]
        arr = (_abi.AdsorptionDesc * len(self.problems))()
        for a, p in zip(arr, self.problems):
            a.n_layers, a.N = int(p["n_layers"]), int(p["N"])
            a.lambda_ = float(p.get("lambda_", 1.0 / 6.0))
            a.chi_as, a.chi_s, a.phibulk = float(p["chi_as"]), float(p["chi_s"]), float(p["phibulk"])
        self._arr = arr
        opt = _abi.Options()
        opt.precision = _abi.FP64
        h = C.c_void_p()
        check(lib().sfx_batch_create(arr, len(self.problems), C.byref(opt), C.byref(h)))
        self.h = h
        nv = C.c_longlong()
        offs = np.zeros(len(self.problems) + 1, dtype=np.int64)
        check(lib().sfx_batch_sizes(self.h, C.byref(nv), C.c_void_p(offs.ctypes.data)))
        self.nvar, self.offsets = nv.value, offs

    def close(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.sfx_batch_destroy(h)
        self.h = None

    __del__ = close

    def tuning(self):
        th = C.c_int()
        ms = (C.c_float * 4)()
        check(lib().sfx_batch_tuning(self.h, C.byref(th), ms))
        return th.value, list(ms)

    def split(self, v):
        return [v[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.problems))]

    def residual(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64) if isinstance(x, np.ndarray) else x
        f = np.empty(self.nvar) if isinstance(x, np.ndarray) else x.new_empty(self.nvar)
        px, where = _ptr(x)
        pf, _ = _ptr(f)
        check(lib().sfx_batch_residual(self.h, px, pf, where))
        return f

    def solve(self, x0=None, check_converged=False, **kw):
        """DIIS on every problem; returns (x, iterations, norms, stats)."""
        o = _abi.SolverOpts()
        check(lib().sfx_solver_defaults(C.byref(o)))
        o.method = _abi.SOLVER_DIIS
        for k, v in kw.items():
            setattr(o, k, v)
        if x0 is None:
            x = np.zeros(self.nvar)
        elif isinstance(x0, np.ndarray):
            x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        else:
            x = x0.clone()
        its = np.zeros(len(self.problems), dtype=np.int32)
        norms = np.zeros(len(self.problems))
        st = _abi.SolveStats()
        px, where = _ptr(x)
        rc = lib().sfx_batch_solve(self.h, C.byref(o), px, where, C.c_void_p(its.ctypes.data),
                                   C.c_void_p(norms.ctypes.data), C.byref(st))
        stats = {k: getattr(st, k) for k, _ in _abi.SolveStats._fields_}
        if rc != 0 and not (rc == _abi.SFX_E_NOTCONVERGED and not check_converged):
            raise SFError(rc, last_error())
        return x, its, norms, stats

    def lngn(self):
        out = np.zeros(len(self.problems))
        check(lib().sfx_batch_get_lngn(self.h, C.c_void_p(out.ctypes.data)))
        return out

    def phi(self, problem, state):
        out = np.zeros(self.problems[problem]["n_layers"] + 2)
        check(lib().sfx_batch_get_phi(self.h, problem, state, C.c_void_p(out.ctypes.data)))
        return out
