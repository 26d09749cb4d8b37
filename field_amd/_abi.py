"""ctypes mirror of the C ABI in include/sfx.h (structs and constants)."""
import ctypes as C

SFX_OK = 0
SFX_E_INVALID = -1
SFX_E_OOM = -2
SFX_E_UNSUPPORTED = -3
SFX_E_NONFINITE = -4
SFX_E_HIP = -5
SFX_E_NOTCONVERGED = -6

FLAT, CYLINDER, SPHERE = 0, 1, 2
LAT_STANDARD, LAT_FCC, LAT_HEX = 0, 1, 2
SURFACE, MIRROR1, MIRROR2, PERIODIC = 1, 2, 3, 4
FREE, RESTRICTED, SOLVENT, RANGE_RESTRICTED = 0, 1, 2, 3
FORM_SFBOX, FORM_MELT = 0, 1
FP64, FP32 = 0, 1
HOST, DEVICE = 0, 1
SOLVER_DIIS, SOLVER_NK, SOLVER_NEWTON = 0, 1, 2

_BOUNDS = {"surface": SURFACE, "mirror1": MIRROR1, "mirror2": MIRROR2, "periodic": PERIODIC}
_GEOM = {"flat": FLAT, "cylindrical": CYLINDER, "cylinder": CYLINDER, "spherical": SPHERE,
         "sphere": SPHERE}
_LATTYPE = {"standard": LAT_STANDARD, "sc": LAT_STANDARD, "fcc": LAT_FCC, "hex": LAT_HEX}


class LatticeDesc(C.Structure):
    _fields_ = [
        ("gradients", C.c_int),
        ("geometry", C.c_int),
        ("lattice_type", C.c_int),
        ("order", C.c_int),
        ("n_layers", C.c_int * 3),
        ("lowerbound", C.c_int * 3),
        ("upperbound", C.c_int * 3),
        ("lambda_", C.c_double),
        ("lambda_chi", C.c_double),
        ("layer_adjustment", C.c_double),
        ("offset_first_layer", C.c_double),
    ]


class MoleculeDesc(C.Structure):
    _fields_ = [
        ("length", C.c_int),
        ("sequence", C.POINTER(C.c_int)),
        ("freedom", C.c_int),
        ("phibulk", C.c_double),
        ("theta", C.c_double),
        ("range_lo", C.c_int * 3),
        ("range_hi", C.c_int * 3),
        ("phi_range", C.c_double),
    ]


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_states", C.c_int),
        ("chi", C.POINTER(C.c_double)),
        ("frozen", C.POINTER(C.c_int)),
        ("frozen_profile", C.POINTER(C.c_double)),
        ("n_molecules", C.c_int),
        ("molecules", C.POINTER(MoleculeDesc)),
        ("stiffness", C.c_double),
        ("form", C.c_int),
        ("melt_phibulk", C.POINTER(C.c_double)),
    ]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("memory_budget", C.c_size_t), ("sweep_kernel", C.c_int),
                ("chain_schedule", C.c_int), ("transfer_kernel", C.c_int)]


SWEEP_AUTO, SWEEP_PER_RANK, SWEEP_BLOCKED = 0, 1, 2
SCHEDULE_AUTO, SCHEDULE_SEQUENTIAL = 0, 1
TRANSFER_AUTO, TRANSFER_ROWS, TRANSFER_MFMA = 0, 1, 2


class SolverOpts(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("tolerance", C.c_double),
        ("max_iter", C.c_int),
        ("diis_m", C.c_int),
        ("diis_eta", C.c_double),
        ("gmres_restart", C.c_int),
        ("gmres_max_iter", C.c_int),
        ("gmres_rtol", C.c_double),
        ("jvp_precision", C.c_int),
        ("fp64_switch", C.c_double),
        ("precondition", C.c_int),
        ("ptc_dt0", C.c_double),
        ("verbose", C.c_int),
        ("diis_backtrack", C.c_int),
        ("newton_deltamax", C.c_double),
        ("newton_linesearch_limit", C.c_int),
        ("newton_reset_criterion", C.c_double),
    ]


class AdsorptionDesc(C.Structure):
    _fields_ = [
        ("n_layers", C.c_int),
        ("N", C.c_int),
        ("lambda_", C.c_double),
        ("chi_as", C.c_double),
        ("chi_s", C.c_double),
        ("phibulk", C.c_double),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("residual_evals", C.c_int),
        ("residual_evals_fp32", C.c_int),
        ("jvp_evals", C.c_int),
        ("residual_norm", C.c_double),
        ("seconds", C.c_double),
        ("converged", C.c_int),
        ("accuracy", C.c_double),
        ("graph_replays", C.c_int),
    ]


K_FWD, K_BWD, K_TFWD, K_TBWD = 0, 1, 2, 3
KERNEL_NAMES = {K_FWD: "propagate", K_BWD: "propagate+compose", K_TFWD: "tangent propagate",
                K_TBWD: "tangent propagate+compose"}


class ProfileEntry(C.Structure):
    _fields_ = [("kind", C.c_int), ("precision", C.c_int), ("launches", C.c_long), ("ranks", C.c_long),
                ("total_ms", C.c_double), ("bytes_per_launch", C.c_double)]


class ProfileReport(C.Structure):
    _fields_ = [("count", C.c_int), ("entries", ProfileEntry * 8)]
