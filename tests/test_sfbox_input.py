"""SFbox input files (field_amd.sfbox_input): the reference's grammar
(Input.cpp), its lattice ranges (Lat1DFlat.cpp, Lat3D.cpp) and its var
sweeps (Vary.cpp), checked on inputs written here, on the reference's own
sample inputs (as text fixtures of their syntax), and on the 1-D wall
workload, which an input file must reproduce exactly."""
import math
import os

import numpy as np
import pytest

from field_amd import configs
from field_amd.sfbox_input import Input, InputError, build, calculations, tokenize, vary

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

WALL = """
// the 1-D wall workload as an SFbox input
lat : flat : gradients : 1
lat : flat : geometry : flat
lat : flat : n_layers : 100
lat : flat : lowerbound : surface
lat : flat : upperbound : mirror1
lat : flat : lambda : 0.16666666666666666
mon : A : freedom : free
mon : S : freedom : free
mon : W : freedom : frozen
mon : W : frozen_range : lowerbound
mon : A : chi - S : 0.5
mon : A : chi - W : -1.0   # wall attraction
mol : pol : composition : (A)100
mol : pol : freedom : free
mol : pol : phibulk : 0.01
mol : sol : composition : (S)1
mol : sol : freedom : solvent
newton : isaac : method : DIIS
newton : isaac : tolerance : 1e-8
start
"""


def test_grammar():
    blocks = tokenize("a : b : c : d\n\n# note\nstart\ne:f:g:h // tail\nstart\nstart\n")
    assert blocks == [[("a", "b", "c", "d")], [("e", "f", "g", "h")]]
    for bad in ("a : b : c", "a : b : c : d : e", "a :  : c : d"):
        with pytest.raises(InputError):
            tokenize(bad)


def test_values_carry_over_between_calculations():
    inp = Input("x : a : p : 1\nx : a : q : 2\nstart\nx : a : p : 3\nstart\n")
    assert len(inp) == 2
    assert inp.values(0)[("x", "a", "p")] == "1"
    assert inp.values(1)[("x", "a", "p")] == "3" and inp.values(1)[("x", "a", "q")] == "2"


def test_wall_input_is_the_wall_workload(orc):
    """The input above builds the system configs.adsorption_1d describes: the
    same lattice, states, chi, frozen wall and molecules, so the oracle gives
    the same residual for both at a seeded field."""
    lat, sysd, solver = build(Input(WALL))
    lat0, sys0, _ = configs.adsorption_1d()
    assert (lat.gradients, tuple(lat.n_layers), lat.lowerbound, lat.upperbound) == (1, (100,), ("surface",), ("mirror1",))
    assert lat.lambda_ == pytest.approx(1 / 6, abs=1e-16)
    assert tuple(sysd.states) == ("A", "S", "W") and np.array_equal(sysd.chi, sys0.chi)
    assert np.array_equal(sysd.frozen["W"], sys0.frozen["W"])
    assert solver == {"method": "diis", "tolerance": 1e-8, "max_iter": 1000}
    x = np.random.default_rng(5).normal(0, 0.1, sysd.nvar(lat))
    assert np.array_equal(orc.System(lat, sysd).residual(x), orc.System(lat0, sys0).residual(x))


def test_ranges_on_3d_and_2d():
    text = """lat : L : gradients : 3
lat : L : geometry : flat
lat : L : n_layers_x : 4
lat : L : n_layers_y : 5
lat : L : n_layers_z : 6
mon : W : freedom : frozen
mon : W : frozen_range : 1,1,lastlayer;4,upperbound,6
mon : A : freedom : free
mol : p : composition : (A)3
mol : p : freedom : restricted
mol : p : theta : 10
"""
    lat, sysd, _ = build(Input(text))
    m = sysd.frozen["W"].reshape(lat.dims).astype(bool)
    want = np.zeros(lat.dims, dtype=bool)
    want[1:5, 1:7, 6:7] = True  # x 1..4, y 1..upperbound (6), z lastlayer (6)..6
    assert np.array_equal(m, want)
    with pytest.raises(InputError):
        build(Input(text.replace("1,1,lastlayer;4,upperbound,6", "1,1;4,6")))


def test_unsupported_features_are_named():
    for text, word in [(WALL.replace("freedom : frozen", "freedom : pinned"), "pinned"),
                       (WALL.replace("(A)100", "(A)50[(A)10](A)50"), "linear"),
                       (WALL.replace("lowerbound : surface", "lowerbound : bulk"), "bulk")]:
        with pytest.raises(InputError, match=word):
            build(Input(text))


@pytest.mark.parametrize("fname, ncalc", [("myinput.in", 1), ("mykal.tmp", 7)])  # 6 `start` lines, data after the last
def test_reference_sample_inputs_tokenize(fname, ncalc):
    """The reference's sample inputs (syntax fixtures, their systems use
    pinned segments and branched chains): they read as `ncalc` calculations,
    and building one names an unsupported feature instead of failing later."""
    inp = Input.from_file(os.path.join(GOLDEN, fname))
    assert len(inp) == ncalc
    with pytest.raises(InputError):
        build(inp, 0)


def test_var_linear_and_exponential():
    """Vary.cpp's sequences: linear steps clamped to end_value; exponential
    steps of 10^(1/step); integers advance to the next integer."""
    base = WALL.replace("start", "")
    lin = base + "var : mol : pol : phibulk\nvar : mol : step : 0.004\nvar : mol : end_value : 0.02\nvar : mol : type : real\n"
    key, vals, typ = vary(Input(lin))
    assert key == ("mol", "pol", "phibulk") and typ == "real"
    assert vals == pytest.approx([0.01, 0.014, 0.018, 0.02])
    ex = base + ("var : mol : pol : phibulk\nvar : mol : step : 2\nvar : mol : end_value : 0.1\n"
                 "var : mol : scale : exponential\nvar : mol : type : real\n")
    _, vals, _ = vary(Input(ex))
    assert vals == pytest.approx([0.01, 0.01 * math.sqrt(10), 0.1])
    it = base.replace("n_layers : 100", "n_layers : 10") + (
        "var : lat : flat : n_layers\nvar : lat : step : 5\nvar : lat : end_value : 22\nvar : lat : type : integer\n")
    _, vals, typ = vary(Input(it))
    assert typ == "integer" and vals == [10, 15, 20, 22]
    sizes = [tuple(c[2].n_layers) for c in calculations(Input(it))]
    assert sizes == [(10,), (15,), (20,), (22,)]


def test_kal_table_and_template():
    """The kal table (Output.cpp:640-718): a header of `class : name :
    property` keys and one tab-separated row per calculation; a template in
    input syntax, '*' allowed, keeps the matching lines (the reference's
    pro.tmp template style)."""
    from field_amd.output import kal_text
    from field_amd.sfbox_input import output_names
    names = output_names(Input(WALL))
    assert names == {"sys": "noname", "newton": "isaac", "mol": ["pol", "sol"]}
    res = [{"free_energy": -1.5, "grand_potential": 0.25, "phibulk": [0.01, 0.99], "lngn": [3.0, 0.0],
            "chemical_potential": [-2.0, -0.5], "iterations": 12, "residual_norm": 1e-9, "converged": True},
           {"free_energy": -1.25, "grand_potential": 0.5, "phibulk": [0.02, 0.98], "lngn": [3.5, 0.0],
            "chemical_potential": [-1.5, -0.25], "iterations": 7, "residual_norm": 2e-9, "converged": False}]
    tmpl = "sys : * : free energy : 1\nmol : pol : FH-MU : 1\nnewton : * : error occurred : false\n"
    text = kal_text(res, names, tmpl)
    assert text.splitlines() == [
        "sys : noname : free energy\tmol : pol : FH-MU\tnewton : isaac : error occurred",
        "-1.5\t-2\tfalse",
        "-1.25\t-1.5\ttrue"]
    full = kal_text(res, names).splitlines()
    assert full[0].split("\t")[:3] == ["sys : noname : free energy", "sys : noname : grand potential",
                                       "mol : pol : phibulk"]
    assert len(full) == 3 and len(full[1].split("\t")) == 2 + 3 * 2 + 3
