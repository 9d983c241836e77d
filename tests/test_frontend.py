"""Input-file front end and OpenDX output (CPU): the parser on the reference's
own example inputs, the refusals, PRINT id rules, the DX layout, and the
printed energy / force lines against the patterns the reference's regression
scripts read (tests/apbs_tester.py:130, apbs_check_intermediate_energies.py:19,
apbs_check_forces.py:20-27 in the reference)."""
import glob
import io
import json
import os
import re

import numpy as np
import pytest

from apbs_amd import dx, nosh

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "examples")
EXPECTED = json.load(open(os.path.join(EX, "expected.json")))

# the reference tester's patterns
FLOAT = r"([-+]?\d*\.\d+(?:[eE][-+]?\d+)?)"
GLOBAL = re.compile(r"Global net (?:ELEC|APOL) energy \= " + FLOAT)
INTERMEDIATE = re.compile(r"\s(?P<label>[a-zA-Z]+ [a-zA-Z]+ = )[\-]?(?P<x>[0-9]+(\.[0-9]+)?"
                          r"(E[\+,\-][0-9]+)?)")
FORCE = re.compile(r"\s+(?P<label>[a-zA-Z]+)\s+(?P<x>[+-]?\d\.\d+E[+-]\d+)\s+"
                   r"(?P<y>[+-]?\d\.\d+E[+-]\d+)\s+(?P<z>[+-]?\d\.\d+E[+-]\d+)")


def inputs():
    return sorted(k for k in EXPECTED)


@pytest.mark.parametrize("key", inputs())
def test_example_inputs_parse_and_are_runnable(key):
    inp = nosh.parse_file(os.path.join(EX, key + ".in"))
    assert inp.elecs, key
    assert nosh.check_runnable(inp) == [], key
    for m in inp.molecules:
        assert os.path.exists(inp.path(m.path)), m.path


def test_born_mol_auto_fields():
    inp = nosh.parse_file(os.path.join(EX, "born", "apbs-mol-auto.in"))
    e0, e1 = inp.elecs
    assert (e0.name, e0.method, e0.dime, e0.cglen, e0.fglen) == (
        "solvated", "mg-auto", (65, 65, 65), (50.0,) * 3, (12.0,) * 3)
    assert e0.cgcent.mol == 1 and e0.fgcent.mol == 1 and e0.mol == 1
    assert (e0.pbe, e0.bcfl, e0.pdie, e0.sdie, e0.chgm, e0.srfm) == (
        "lpbe", "mdh", 1.0, 78.54, "spl2", "mol")
    assert (e0.srad, e0.swin, e0.sdens, e0.temp, e0.calcenergy, e0.calcforce) == (
        1.4, 0.3, 10.0, 298.15, "total", "no")
    assert e0.writes == [("pot", "gz", "potential")]
    assert e1.sdie == 1.0 and e1.writes == []
    assert [(p.what, p.ids, p.ops) for p in inp.prints] == [
        ("elecenergy", ["solvated", "reference"], ["-"])]


def test_ion_keyword_forms_and_comments():
    text = """read mol pqr a.pqr end   # comment: end quit
    elec name x mg-manual dime 33 33 33 grid 0.5 0.5 0.5 gcent 1.0 -2.0 3e0 mol 1 npbe
      bcfl sdh ion charge 1 conc 0.1 radius 2.0 ion radius 1.5 charge -2 conc 0.05
      ion 1.0 0.2 2.5
      pdie 2 sdie 78.5 chgm spl0 srfm smol srad 1.4 sdens 10 swin 0.3 temp 300
      calcenergy comps calcforce total end
    print energy 1 end
    quit"""
    inp = nosh.parse(text)
    e = inp.elecs[0]
    assert e.ions == [(1.0, 0.1, 2.0), (-2.0, 0.05, 1.5), (1.0, 0.2, 2.5)]
    assert e.gcent.xyz == (1.0, -2.0, 3.0)
    assert inp.prints[0].what == "elecenergy"
    assert nosh.check_runnable(inp) == []


def test_bcfl_forms():
    base = ("read mol pqr a.pqr end elec mg-manual bcfl {} end quit")
    assert nosh.parse(base.format("2")).elecs[0].bcfl == "mdh"
    assert nosh.parse(base.format("focus")).elecs[0].bcfl == "focus"
    assert nosh.parse(base.format("MAP")).elecs[0].bcfl == "map"
    with pytest.raises(nosh.InputError):
        nosh.parse(base.format("3"))


def test_print_ids_digits_before_names():
    """An all-digit PRINT id is a section number even if a section has that
    name (NOsh_parsePRINT); other ids are names."""
    elecs = [nosh.Elec(name="2"), nosh.Elec(name="b")]
    assert nosh.resolve_id("2", elecs, "ELEC") == 1
    assert nosh.resolve_id("1", elecs, "ELEC") == 0
    assert nosh.resolve_id("b", elecs, "ELEC") == 1
    with pytest.raises(nosh.InputError):
        nosh.resolve_id("3", elecs, "ELEC")
    with pytest.raises(nosh.InputError):
        nosh.resolve_id("c", elecs, "ELEC")


@pytest.mark.parametrize("snippet,reason", [
    ("read mol pdb a.pdb end elec mg-manual end quit", "only PQR"),
    ("read mol pqr a.pqr end elec mg-dummy end quit", "not set: dime"),
    ("read mol pqr a.pqr end elec mg-para dime 9 9 9 cglen 9 9 9 fglen 5 5 5 cgcent mol 1 "
     "fgcent mol 1 mol 1 lpbe bcfl mdh pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 "
     "end quit", "pdime"),
    ("read mol pqr a.pqr end elec mg-para pdime 2 1 1 ofrac 0.1 async 2 dime 9 9 9 cglen 9 9 9 "
     "fglen 5 5 5 cgcent mol 1 fgcent mol 1 mol 1 lpbe bcfl mdh pdie 1 sdie 1 chgm spl2 "
     "srfm spl2 swin 0.3 temp 300 end quit", "outside the pdime"),
    ("read mol pqr a.pqr end elec fe-manual end quit", "fe-manual"),
    ("read mol pqr a.pqr end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 mol 1 lpbe "
     "bcfl map pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 end quit", "bcfl map"),
    ("read mol pqr a.pqr end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 mol 1 lpbe "
     "bcfl mdh pdie 1 sdie 1 chgm spl6 srfm spl2 swin 0.3 temp 300 end quit", "chgm spl6"),
    ("read mol pqr a.pqr end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 mol 1 lpbe "
     "bcfl mdh pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 usemap diel 1 end quit",
     "1 is not a valid diel map ID"),
    ("read mol pqr a.pqr diel dxbin x y z end elec mg-manual end quit", "read as OpenDX"),
    ("read mol pqr a.pqr mesh mcsf m end elec mg-manual end quit", "finite-element"),
    ("read mol pqr a.pqr kappa dx k.dx end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 "
     "mol 1 lpbe bcfl mdh ion 1 0.1 2 ion -1 0.1 2 pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 "
     "temp 300 usemap kappa 1 end quit", "only with dielectric maps"),
    ("read mol pqr a.pqr diel dx x y z end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 "
     "mol 1 lpbe bcfl mdh ion 1 0.1 2 ion -1 0.1 2 pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 "
     "temp 300 usemap diel 1 end quit", "need a kappa map"),
    ("read mol pqr a.pqr pot dx p.dx end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 "
     "mol 1 lpbe bcfl map pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 usemap pot 2 "
     "end quit", "2 is not a valid pot map ID"),
    ("read mol pqr a.pqr end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 mol 1 lpbe "
     "bcfl mdh pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 usemap phi 1 end quit",
     "usemap phi"),
    ("read mol pqr a.pqr end elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 mol 1 lpbe "
     "bcfl focus pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 end quit",
     "can't focus the first"),
    ("read mol pqr a.pqr end elec mg-manual dime 9 9 9 gcent mol 1 mol 1 lpbe "
     "bcfl mdh pdie 1 sdie 1 chgm spl2 srfm mol temp 300 end quit", "srad"),
])
def test_refusals_name_the_keyword(snippet, reason):
    why = nosh.check_runnable(nosh.parse(snippet))
    assert any(reason in w for w in why), why


def test_map_inputs_are_runnable():
    """READ diel/kappa/charge/pot maps with usemap, bcfl map and mg-dummy."""
    text = ("read mol pqr a.pqr diel dx x.dx y.dx z.dx kappa gz k.dx.gz charge dx q.dx "
            "pot dx p.dx end "
            "elec name d mg-dummy dime 9 9 9 glen 5 5 5 gcent mol 1 mol 1 lpbe bcfl mdh "
            "pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 write dielx dx dx end "
            "elec mg-manual dime 9 9 9 glen 5 5 5 gcent mol 1 mol 1 lpbe bcfl map "
            "ion 1 0.1 2 ion -1 0.1 2 pdie 1 sdie 1 chgm spl2 srfm spl2 swin 0.3 temp 300 "
            "usemap diel 1 usemap kappa 1 usemap charge 1 usemap pot 1 end quit")
    inp = nosh.parse(text)
    assert nosh.check_runnable(inp) == []
    assert [m[0] for m in inp.maps] == ["diel", "kappa", "charge", "pot"]
    assert inp.maps[0][2] == ["x.dx", "y.dx", "z.dx"]
    assert inp.elecs[1].usemaps == [("diel", 1), ("kappa", 1), ("charge", 1), ("pot", 1)]


@pytest.mark.parametrize("bad", ["read mol pqr a.pqr end", "elec mg-manual dime 1 2 end quit",
                                 "print elecEnergy 1 * 2 end quit", "bogus quit",
                                 "elec mg-manual ion charge 1 conc 0.1 depth 2 end quit"])
def test_malformed_inputs_raise(bad):
    with pytest.raises(nosh.InputError):
        nosh.parse(bad)


def test_dx_layout_and_round_trip(tmp_path):
    """Values three to a line with z fastest; header origin / deltas /
    counts; the reader inverts the writer."""
    nx, ny, nz = 4, 3, 2
    a = np.empty((nz, ny, nx))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                a[k, j, i] = 100 * i + 10 * j + k
    p = tmp_path / "g.dx"
    dx.write_dx(str(p), a, (-1.5, 0.0, 2.25), (0.5, 0.25, 1.0), "POTENTIAL (kT/e)")
    lines = p.read_text().splitlines()
    assert lines[2] == "# POTENTIAL (kT/e)"
    assert lines[4] == "object 1 class gridpositions counts 4 3 2"
    assert lines[5].split() == ["origin", "-1.500000e+00", "0.000000e+00", "2.250000e+00"]
    assert [ln.split()[1:] for ln in lines[6:9]] == [
        ["5.000000e-01", "0.000000e+00", "0.000000e+00"],
        ["0.000000e+00", "2.500000e-01", "0.000000e+00"],
        ["0.000000e+00", "0.000000e+00", "1.000000e+00"]]
    assert lines[10].startswith("object 3 class array type double rank 0 items 24")
    body = " ".join(lines[11:19]).split()
    order = [100 * i + 10 * j + k for i in range(nx) for j in range(ny) for k in range(nz)]
    assert [float(v) for v in body] == order
    assert all(len(ln.split()) == 3 for ln in lines[11:19])
    assert lines[19] == 'attribute "dep" string "positions"'
    b, origin, h = dx.read_dx(str(p))
    assert np.array_equal(a, b) and origin == (-1.5, 0.0, 2.25) and h == (0.5, 0.25, 1.0)
    pg = tmp_path / "g.dx.gz"
    dx.write_dx(str(pg), a, (0, 0, 0), (1, 1, 1), gz=True)
    assert np.array_equal(dx.read_dx(str(pg))[0], a)


def test_dx_partial_last_line(tmp_path):
    a = np.arange(5.0).reshape(5, 1, 1)
    p = tmp_path / "h.dx"
    dx.write_dx(str(p), a, (0, 0, 0), (1, 1, 1))
    assert np.array_equal(dx.read_dx(str(p))[0], a)


def _runner_with(results, inp):
    from apbs_amd import run
    out = io.StringIO()
    r = run.Runner(inp, out=out)
    r.per_rank = [(results, list(range(len(results))))]
    r.ranks = [0]
    return r, out


def test_print_energy_lines_match_tester_patterns():
    """printElecEnergy: kb*1e-3*Na*T*E per term, signed sum; the global
    line is the one the reference's tester reads."""
    from apbs_amd import run
    inp = nosh.parse("elec name a mg-manual end elec name b mg-manual end "
                     "print elecEnergy a - b end quit")
    ra = run.CalcResult(298.15, "total", "no")
    rb = run.CalcResult(298.15, "total", "no")
    ra.energy, rb.energy = 100.0, 40.0
    r, out = _runner_with([ra, rb], inp)
    r._print_energy(inp.prints[0])
    text = out.getvalue()
    want = run.VUNIT_KB * 1e-3 * run.VUNIT_NA * 298.15 * 100.0
    want += -1.0 * run.VUNIT_KB * 1e-3 * run.VUNIT_NA * 298.15 * 40.0
    got = [float(v) for v in GLOBAL.findall(text)]
    assert got == [float("%1.12E" % want)]
    assert "print energy 1 (a) - 2 (b) end" in text


def test_print_force_lines_match_tester_pattern():
    from apbs_amd import run
    inp = nosh.parse("elec name a mg-manual end elec name b mg-manual end "
                     "print elecForce a - b end quit")
    ra = run.CalcResult(300.0, "no", "total")
    rb = run.CalcResult(300.0, "no", "total")
    ra.forces = np.arange(9.0).reshape(1, 3, 3) * 1e-3
    rb.forces = np.ones((1, 3, 3)) * 1e-3
    r, out = _runner_with([ra, rb], inp)
    r._print_force(inp.prints[0])
    lines = out.getvalue().splitlines()
    start = next(i for i, ln in enumerate(lines) if ln.startswith("print force"))
    got = {}
    for ln in lines[start:]:
        m = FORCE.search(ln)
        if m:
            got[m.group("label")] = [float(m.group(c)) for c in "xyz"]
    c = run.VUNIT_KB * 1e-3 * run.VUNIT_NA * 300.0
    f = c * ra.forces[0] - c * rb.forces[0]
    assert np.allclose(got["qf"], f[0], rtol=1e-11)
    assert np.allclose(got["ib"], f[1], rtol=1e-11)
    assert np.allclose(got["db"], f[2], rtol=1e-11)
    assert np.allclose(got["tot"], f[0] + f[1] + f[2], rtol=1e-11)


def test_all_reference_example_inputs_parse_or_are_refused():
    """Every example input the reference ships either parses to a runnable
    input or is refused with a reason (nothing crashes the parser)."""
    ref = "/root/reference/examples"
    if not os.path.isdir(ref):
        pytest.skip("reference examples not present")
    for f in glob.glob(os.path.join(ref, "*", "*.in")):
        if "TEMPLATE" in f:
            continue
        inp = nosh.parse_file(f)
        nosh.check_runnable(inp)
