"""Full Vpmg calculations on the GPU (through the C ABI) against the oracle
and against the reference's own golden energies.

Parity protocol.  The GPU solver orders its coarse-level Gauss-Seidel sweeps
by 8 colours and coarsens below the reference's level count, solving the
coarsest grid exactly; the reference sweeps coarse levels lexicographically
and stops at nlev with a banded / CG coarse solve.  Both converge to the same
discrete solution, so:
  * at a tight tolerance (etol 1e-12) both solves are converged and the GPU
    potential and energy match the oracle's reference schedule to 1e-9
    relative;
  * against the oracle run with the GPU's own schedule (coarse_smoother 1,
    deep_coarse 1) at the default etol, the iteration counts are equal and
    energies agree to 1e-11;
  * the reference's golden energies (computed at etol 1e-6) are met to 1e-6.
"""
import json
import os

import numpy as np
import pytest

import orc_driver as od
from apbs_amd import driver, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))


def rel(a, b):
    return abs(a - b) / abs(b)


def pot_err(ug, uo):
    """(relative L2, relative Linf) of ug against uo."""
    d = ug - uo
    return (np.linalg.norm(d) / np.linalg.norm(uo), np.abs(d).max() / np.abs(uo).max())


def small_mol(n, radius, seed, net=None):
    xyz, q, r, _ = synth.globule(n, (radius,) * 3, seed, net_charge=net)
    return xyz, q, r


SALT = [(1, 0.15, 2.0), (-1, 0.15, 2.0)]


def pbe_of(srfm, chgm, **kw):
    p = dict(pdie=2.0, sdie=78.54, srfm=srfm, chgm=chgm, srad=1.4, swin=0.3, sdens=10.0,
             temp=298.15, bcfl="mdh", ions=SALT)
    p.update(kw)
    return p


# ------------------------------------------------------------ coefficient maps

@pytest.mark.parametrize("srfm,chgm", [("mol", "spl0"), ("smol", "spl2"), ("spl2", "spl2"),
                                       ("smol", "spl0"), ("spl3", "spl2"), ("spl4", "spl4"),
                                       ("spl4", "spl0"), ("mol", "spl4")])
def test_fillco_maps(srfm, chgm):
    """Vpmg_fillco: dielectric face maps, kappa and charge maps equal the
    oracle's bitwise; MDH boundary values to 32 ulp of the largest value (the
    device evaluates each term as c_a exp(-kappa d) / d with a refined
    reciprocal square root, a few ulp from the reference's expression, and
    the sum over atoms accumulates those differences)."""
    atoms = small_mol(80, 6.0, 11)
    c = od.mol_center(atoms[0])
    pbe = pbe_of(srfm, chgm)
    mol = driver.Molecule(*atoms)
    g = driver.Calc(mol, driver.manual_mgparm([41] * 3, c, grid=[0.45] * 3, chgm=chgm), pbe)
    g.fillco()
    o = od.OrcCalc(od.make_params([41] * 3, [0.45] * 3, c, 4, pbe), atoms)
    o.fillco()
    for name in ("epsx", "epsy", "epsz", "kappa", "charge"):
        a, b = g.array(name), o.array(name)
        assert np.array_equal(a, b), (name, np.abs(a - b).max())
    for a, b in zip(g.boundary(), o.boundary()):
        err = np.abs(a - b).max() / (np.finfo(float).eps * np.abs(b).max())
        print(f"{srfm}/{chgm}: boundary within {err:.1f} ulp of the largest value")
        assert err <= 32


# -------------------------------------------------------------- golden energies

@pytest.mark.parametrize("srfm,key", [("smol", "born_smol_auto"), ("mol", "born_mol_auto")])
def test_born_focusing_golden(srfm, key):
    """examples/born/apbs-{s}mol-auto.in on the GPU: every focusing level's
    energy against tests/test_cases.cfg."""
    atoms = od.read_pqr(os.path.join(HERE, "golden", "examples", "born", "ion.pqr"))
    c = od.mol_center(atoms[0])
    mol = driver.Molecule(*atoms)
    last = {}
    for state, sdie in (("solvated_levels", 78.54), ("reference_levels", 1.0)):
        pbe = dict(pdie=1.0, sdie=sdie, srfm=srfm, chgm="spl2", srad=1.4, swin=0.3,
                   sdens=10.0, temp=298.15, bcfl="mdh")
        e, _ = driver.run_auto(mol, [65] * 3, [50] * 3, [12] * 3, c, c, pbe)
        for got, want in zip(e, G[key][state]):
            assert rel(got, want) < 1e-6, (state, got, want)
        last[state] = e[-1]
    dg = last["solvated_levels"] - last["reference_levels"]
    assert rel(dg, G[key]["difference"]) < 1e-4


@pytest.mark.parametrize("srfm", ["smol", "mol"])
def test_solv_manual_golden(srfm):
    """examples/solv/apbs-{s}mol.in on the GPU (chgm spl0, no salt)."""
    out = []
    for name in ("methanol", "methoxide"):
        atoms = od.read_pqr(os.path.join(HERE, "golden", "examples", "solv", f"{name}.pqr"))
        c = od.mol_center(atoms[0])
        mol = driver.Molecule(*atoms)
        for sdie in (78.0, 1.0):
            pbe = dict(pdie=2.0, sdie=sdie, srfm=srfm, chgm="spl0", srad=0.0, swin=0.3,
                       sdens=10.0, temp=300.0, bcfl="mdh", ions=[(1, 0.0, 2.0), (-1, 0.0, 2.0)])
            e, _ = driver.run_manual(mol, [65] * 3, c, pbe, grid=[0.25] * 3)
            out.append(e)
    for got, want in zip(out, G["solv_" + srfm]["energies"]):
        assert rel(got, want) < 1e-6, (got, want)


# ------------------------------------------------------------ oracle parity

def _pair(atoms, dime, h, pbe, orc_extra=None):
    c = od.mol_center(atoms[0])
    mol = driver.Molecule(*atoms)
    eg, g = driver.run_manual(mol, [dime] * 3, c, pbe, grid=[h] * 3)
    po = dict(pbe)
    po.update(orc_extra or {})
    eo, o = od.run_manual(atoms, [dime] * 3, c, po, grid=[h] * 3)
    return eg, g, eo, o


@pytest.mark.parametrize("npbe", [False, True])
def test_tight_parity_reference_schedule(npbe):
    """The 65^3 lpbe workload (500-atom globule, 0.5 A, spl2/spl2, 0.15 M) and
    its npbe form, solved to etol 1e-12 on both sides: potential and energy
    within 1e-9 relative of the oracle's reference-schedule solve."""
    atoms = synth.make("lpbe65")
    pbe = dict(synth.CONFIGS["lpbe65"]["pbe"], etol=1e-12, npbe=npbe)
    eg, g, eo, o = _pair(atoms, 65, 0.5, pbe)
    l2, linf = pot_err(g.array("u"), o.array("u"))
    print(f"npbe={npbe}: energy gpu {eg:.12f} oracle {eo:.12f} rel {rel(eg, eo):.2e}; "
          f"potential rel L2 {l2:.2e} Linf {linf:.2e}; stats {g.stats()}")
    assert rel(eg, eo) < 1e-9
    assert l2 < 1e-9 and linf < 1e-9


@pytest.mark.parametrize("npbe", [False, True])
def test_same_schedule_parity(npbe):
    """Default etol; oracle with the GPU's schedule (8-colour coarse sweeps,
    deep coarsening, exact coarsest solve): same iteration counts, energies
    to 1e-11."""
    atoms = small_mol(300, 9.0, 5, net=-3)
    pbe = pbe_of("smol", "spl2", npbe=npbe)
    eg, g, eo, o = _pair(atoms, 49, 0.45, pbe, dict(coarse_smoother=1, deep_coarse=1))
    sg, so = g.stats(), o.stats()
    print(f"npbe={npbe}: gpu {eg:.12f} ({sg}) oracle {eo:.12f} ({so})")
    assert sg["iters"] == so["iters"] and sg["inner"] == so["inner"]
    assert sg["nlev"] == so["nlev"]
    assert rel(eg, eo) < 1e-11
    l2, linf = pot_err(g.array("u"), o.array("u"))
    assert linf < 1e-10


# ----------------------------------------------------------------- forces

@pytest.mark.parametrize("srfm,chgm", [("spl2", "spl2"), ("spl4", "spl4"), ("spl3", "spl0")])
def test_forces_against_oracle(srfm, chgm):
    """Vpmg_qfForce / ibForce / dbForce on a tightly converged solve, per
    atom, against the oracle; spl3 / spl4 select the polynomial-window
    gradients and chgm spl4 the 5th-order B-spline field."""
    atoms = small_mol(40, 5.0, 17, net=2)
    pbe = pbe_of(srfm, chgm, etol=1e-12)
    eg, g, eo, o = _pair(atoms, 41, 0.4, pbe)
    assert abs(eg - eo) <= 1e-9 * abs(eo), (eg, eo)
    fmax = 0.0
    diffs = []
    nrc0 = 0
    for i in range(len(atoms[1])):
        rg, *fg = g.force(i)
        ro, *fo = o.force(i)
        # spl3 / spl4 take the solvent radius from srad (routines.c:1371-1378),
        # so near-edge atoms' dielectric-boundary boxes leave the grid and
        # Vpmg_dbForce returns 0 for them on both sides
        assert bool(rg) == bool(ro), i
        nrc0 += not rg
        for a, b in zip(fg, fo):
            fmax = max(fmax, np.abs(b).max())
            diffs.append(np.abs(a - b).max())
    print(f"{srfm}/{chgm}: force max |F| {fmax:.4e}, max diff {max(diffs):.2e}, "
          f"{nrc0} atoms with return code 0")
    assert max(diffs) <= 1e-8 * fmax


# ---------------------------------------------------------- mixed precision

@pytest.mark.parametrize("npbe", [False, True])
def test_mixed_precision_energy(npbe):
    """FP32 V-cycles inside FP64 defect correction: energy within 1e-6 of the
    FP64 GPU solve, potential errors stated."""
    atoms = synth.make("lpbe65")
    pbe = dict(synth.CONFIGS["lpbe65"]["pbe"], npbe=npbe)
    c = od.mol_center(atoms[0])
    mol = driver.Molecule(*atoms)
    ref_pbe = dict(pbe, etol=1e-12)
    e64, g64 = driver.run_manual(mol, [65] * 3, c, ref_pbe, grid=[0.5] * 3)
    emx, gmx = driver.run_manual(mol, [65] * 3, c, pbe, grid=[0.5] * 3, precision=1,
                                 mixed_tol=1e-10)
    l2, linf = pot_err(gmx.array("u"), g64.array("u"))
    print(f"npbe={npbe}: fp64 {e64:.10f} mixed {emx:.10f} rel {rel(emx, e64):.2e} "
          f"L2 {l2:.2e} Linf {linf:.2e} stats {gmx.stats()}")
    assert rel(emx, e64) < 1e-6
    assert linf < 1e-6


@pytest.mark.parametrize("npbe", [False, True])
def test_mixed_precision_against_oracle(npbe):
    """Mixed precision held to the oracle directly: the 65^3 workload (lpbe and
    its npbe form) against the oracle's reference-schedule solve at etol
    1e-12.  Converged (FP64 defect correction to 1e-11, or Newton etol 1e-12
    with FP32 inner cycles): energy and potential L2 / Linf within 1e-9.  At
    the default etol (1e-6): energy within 1e-6, potential L2 / Linf within
    1e-4."""
    atoms = synth.make("lpbe65")
    pbe = dict(synth.CONFIGS["lpbe65"]["pbe"], npbe=npbe)
    c = od.mol_center(atoms[0])
    eo, o = od.run_manual(atoms, [65] * 3, c, dict(pbe, etol=1e-12), grid=[0.5] * 3)
    uo = o.array("u")
    mol = driver.Molecule(*atoms)
    tight = dict(pbe, etol=1e-12) if npbe else pbe
    et, gt = driver.run_manual(mol, [65] * 3, c, tight, grid=[0.5] * 3, precision=1,
                               mixed_tol=1e-11)
    l2t, linft = pot_err(gt.array("u"), uo)
    ed, gd = driver.run_manual(mol, [65] * 3, c, pbe, grid=[0.5] * 3, precision=1)
    l2d, linfd = pot_err(gd.array("u"), uo)
    print(f"npbe={npbe}: oracle {eo:.12f}; mixed converged {et:.12f} rel {rel(et, eo):.2e} "
          f"L2 {l2t:.2e} Linf {linft:.2e} ({gt.stats()}); mixed etol 1e-6 {ed:.12f} "
          f"rel {rel(ed, eo):.2e} L2 {l2d:.2e} Linf {linfd:.2e} ({gd.stats()})")
    assert rel(et, eo) < 1e-9 and l2t < 1e-9 and linft < 1e-9
    assert rel(ed, eo) < 1e-6 and l2d < 1e-4 and linfd < 1e-4


def test_c_caller_matches_python(tmp_path):
    """tests/c_caller.c runs the routines.c sequence through the C header
    alone; its energy and force equal the Python driver's bitwise."""
    import subprocess
    from test_abi import build_c_caller
    exe = build_c_caller(tmp_path)
    pqr = os.path.join(HERE, "golden", "examples", "born", "ion.pqr")
    out = subprocess.run([exe, "solve", pqr, "33", "0.5"], capture_output=True, text=True,
                         check=True, timeout=120).stdout
    vals = {l.split()[0]: [float(x) for x in l.split()[1:]] for l in out.splitlines()}
    atoms = od.read_pqr(pqr)
    mol = driver.Molecule(*atoms)
    pbe = dict(pdie=1.0, sdie=78.54, srfm="spl2", chgm="spl2", srad=0.3, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=SALT)
    e, calc = driver.run_manual(mol, [33] * 3, mol.center, pbe, grid=[0.5] * 3)
    rc, qf, ib, db = calc.force(0)
    assert vals["energy"][0] == e
    assert np.array_equal(np.array(vals["force"]), qf + db + ib)  # vpmg.c:5434-5436 order
    assert vals["umax"][0] == calc.array("u").max()


@pytest.mark.parametrize("npbe,precision", [(False, 0), (True, 0), (True, 1)])
def test_graph_replay_equals_eager(npbe, precision, monkeypatch):
    """V-cycles replayed from captured HIP graphs give bitwise the results of
    eager launches, over three solves on one object (each solve replays the
    captured cycles many times, with eager Newton work in between)."""
    atoms = small_mol(300, 9.0, 21, net=-2)
    pbe = pbe_of("smol", "spl2", npbe=npbe)
    mol = driver.Molecule(*atoms)
    c = od.mol_center(atoms[0])
    res = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("APBS_PMG_GRAPHS", graphs)
        calc = driver.Calc(mol, driver.manual_mgparm([49] * 3, c, grid=[0.45] * 3), pbe,
                           precision=precision)
        runs = []
        for _ in range(3):
            calc.fillco()
            calc.solve()
            calc.set_part_domain()
            runs.append((calc.energy(1), calc.array("u").copy()))
        res[graphs] = runs
    for (eg, ug), (ee, ue) in zip(res["1"], res["0"]):
        assert eg == ee
        assert np.array_equal(ug, ue)
    assert res["1"][0][0] == res["1"][2][0]


@pytest.mark.parametrize("env", ["APBS_PMG_SMOOTHER", "APBS_PMG_SMOOTHER27"])
@pytest.mark.parametrize("precision", [0, 1])
def test_smoother_forms_give_the_same_bits(env, precision, monkeypatch):
    """The one-launch-per-colour smoother forms (fine level, or every 27-point
    level) give bitwise the default forms' npbe solve, in FP64 and mixed
    precision (whose cycles start from zero: the per-colour fine form clears
    the correction, the fused form's first sweep does not read it)."""
    atoms = small_mol(300, 9.0, 23, net=-3)
    pbe = pbe_of("smol", "spl2", npbe=True)
    mol = driver.Molecule(*atoms)
    c = od.mol_center(atoms[0])
    res = {}
    for form in ("colour", "default"):
        if form == "colour":
            monkeypatch.setenv(env, "colour")
        else:
            monkeypatch.delenv(env, raising=False)
        calc = driver.Calc(mol, driver.manual_mgparm([49] * 3, c, grid=[0.45] * 3), pbe,
                           precision=precision)
        calc.fillco()
        calc.solve()
        res[form] = (calc.energy(1), calc.array("u").copy(), calc.stats()["iters"])
        calc.release()
    assert res["colour"][0] == res["default"][0]
    assert np.array_equal(res["colour"][1], res["default"][1])
    assert res["colour"][2] == res["default"][2]


@pytest.mark.parametrize("srfm,key", [("mol", "actin_mol_auto"), ("smol", "actin_smol_auto")])
def test_actin_dimer_golden(srfm, key):
    """examples/actin-dimer/apbs-{s}mol-auto.in on the GPU: mg-auto at 161^3,
    npbe, bcfl sdh on the coarse level, chgm spl0, 11,754-atom complex;
    every focusing level's energy against tests/test_cases.cfg.  The binding
    energy is a difference of numbers 5000x larger, so the reference's own
    convergence error (etol 1e-6) bounds how closely it can be met."""
    c = od.mol_center(od.read_pqr(os.path.join(HERE, "golden", "examples", "actin-dimer", "complex.pqr"))[0])
    pbe = dict(npbe=True, pdie=2.0, sdie=78.4, srfm=srfm, chgm="spl0", srad=1.4, swin=0.3,
               sdens=10.0, temp=298.15, bcfl="sdh", ions=[(1, 0.050, 2.0), (-1, 0.050, 2.0)])
    fine = {}
    for name in ("mol1", "mol2", "complex"):
        atoms = od.read_pqr(os.path.join(HERE, "golden", "examples", "actin-dimer", f"{name}.pqr"))
        mol = driver.Molecule(*atoms)
        e, _ = driver.run_auto(mol, [161] * 3, [156, 121, 162], [112, 91, 116], c, c, pbe)
        want = G[key]["levels"][name]
        assert len(e) == len(want)
        for got, w in zip(e, want):
            assert rel(got, w) < 1e-6, (name, got, w)
        fine[name] = e[-1]
    dg = fine["complex"] - fine["mol1"] - fine["mol2"]
    print(f"{key}: binding {dg:.6f} kJ/mol, golden {G[key]['difference']:.6f}")
    assert rel(dg, G[key]["difference"]) < 1e-3


@pytest.mark.parametrize("salt", [0.0, 0.15])
def test_sdh_boundary_against_oracle(salt):
    """bcfl sdh boundary values (vpmg.c:3479-3645) against the oracle."""
    atoms = small_mol(60, 6.0, 23, net=3)
    c = od.mol_center(atoms[0])
    pbe = pbe_of("mol", "spl0", bcfl="sdh", ions=[(1, salt, 2.0), (-1, salt, 2.0)])
    mol = driver.Molecule(*atoms)
    g = driver.Calc(mol, driver.manual_mgparm([41] * 3, c, grid=[0.5] * 3, chgm="spl0"), pbe)
    g.fillco()
    o = od.OrcCalc(od.make_params([41] * 3, [0.5] * 3, c, 4, pbe), atoms)
    o.fillco()
    for a, b in zip(g.boundary(), o.boundary()):
        assert np.abs(a - b).max() <= 8 * np.finfo(float).eps * np.abs(b).max()
