"""The C ABI library loads and exports every entry point include/sfx.h declares.
No compute calls: this runs without a GPU."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    import glob
    text = "".join(open(p).read() for p in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_hot_path():
    syms = declared_symbols()
    for s in ("sfx_system_create", "sfx_residual", "sfx_jvp", "sfx_compute_phi", "sfx_free_energy",
              "sfx_solve", "sfx_lattice_propagate"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from field_amd import sfx
    lib = ctypes.CDLL(sfx.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from field_amd import sfx
    assert set(declared_symbols()) == set(sfx._SIGS), set(declared_symbols()) ^ set(sfx._SIGS)


def test_abi_version_and_struct_sizes():
    from field_amd import _abi, sfx
    assert sfx.lib().sfx_abi_version() == 3
    # lattice desc: 4 ints + 3*3 ints + 4 doubles
    assert ctypes.sizeof(_abi.LatticeDesc) == 4 * 4 + 9 * 4 + 4 + 4 * 8  # 4 bytes padding before doubles
    assert ctypes.sizeof(_abi.MoleculeDesc) == 72  # + range box (6 ints) and phi_range


def test_ctypes_layouts_match_the_c_headers(tmp_path):
    """Every ctypes mirror of an ABI struct has the size and field offsets the
    C compiler gives the declaration in include/*.h."""
    import subprocess

    from field_amd import _abi

    mirrors = {"sfx_lattice_desc": _abi.LatticeDesc, "sfx_molecule_desc": _abi.MoleculeDesc,
               "sfx_system_desc": _abi.SystemDesc, "sfx_options": _abi.Options,
               "sfx_solver_opts": _abi.SolverOpts, "sfx_solve_stats": _abi.SolveStats,
               "sfx_adsorption_desc": _abi.AdsorptionDesc, "sfx_profile_entry": _abi.ProfileEntry,
               "sfx_profile_report": _abi.ProfileReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sfx.h"', '#include "sfx_profile.h"',
             "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            cf = "lambda" if fname == "lambda_" else fname
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {cf}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_errors_come_back_as_codes():
    """Invalid descriptors are rejected before any device work."""
    import pytest
    from field_amd import sfx
    from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec
    import numpy as np
    lat = LatticeSpec(1, (10,), lowerbound=("mirror1",), upperbound=("mirror1",), lambda_=0.0)
    sysd = SystemSpec(states=("A", "S"), chi=np.zeros((2, 2)),
                      molecules=[MoleculeSpec("p", "(A)5", "free", phibulk=0.1),
                                 MoleculeSpec("s", "(S)1", "solvent")])
    with pytest.raises(sfx.SFError) as e:
        sfx.System(lat, sysd)
    assert e.value.code == -1 and "lambda" in str(e.value)
