"""External maps on the GPU (usemap diel/kappa/charge/pot, bcfl map) against
the oracle, and through the input front end.

fillcoCoefMap, fillcoChargeMap and bcfl_map (vpmg.c:4098-4220, 4938-4985,
3953-3990) sample a Vgrid (vgrid.c:179-260) at every grid point; the device
evaluates the same arithmetic, so the coefficient and charge arrays equal the
oracle's bitwise.  Maps here live on a foreign grid (other spacing and origin
than the calculation's), with kappa values above 1 (scaled by the maximum) and
below VPMGSMALL (dropped).
"""
import io
import os

import numpy as np
import pytest

import orc_driver as od
from apbs_amd import driver, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SALT = [(1, 0.15, 2.0), (-1, 0.15, 2.0)]
DIME, H = 33, 0.45


def foreign_maps(center, seed, kappa_max=2.5):
    """Smooth random maps over [center - 9, center + 9.3] on their own grid."""
    rng = np.random.default_rng(seed)
    origin = np.asarray(center) - 9.0
    out = {}
    shape = (29, 33, 31)             # nz, ny, nx
    spacing = (0.61, 0.57, 0.66)

    def smooth():
        a = rng.random(shape)
        for ax in range(3):
            a = 0.5 * a + 0.25 * (np.roll(a, 1, ax) + np.roll(a, -1, ax))
        return (a - a.min()) / (a.max() - a.min())

    for k in ("dielx", "diely", "dielz"):
        out[k] = (2.0 + 76.54 * smooth(), origin, spacing)
    kap = kappa_max * smooth()
    kap[kap < 0.05] = 0.0
    kap[3, 4, 5] = 3e-13                 # below VPMGSMALL: dropped
    out["kappa"] = (kap, origin, spacing)
    out["charge"] = (0.02 * (smooth() - 0.5), origin, spacing)
    out["pot"] = (3.0 * (smooth() - 0.5), origin, spacing)
    return out


def gpu_calc(atoms, center, pbe, maps):
    mol = driver.Molecule(*atoms)
    dm = {k: driver.Map(*v) for k, v in maps.items()}
    pb = dict(pbe, maps=dm)
    c = driver.Calc(mol, driver.manual_mgparm([DIME] * 3, center, grid=[H] * 3,
                                              chgm=pbe["chgm"]), pb)
    c._keep = (mol, dm)
    return c


def orc_calc(atoms, center, pbe, maps):
    o = od.OrcCalc(od.make_params([DIME] * 3, [H] * 3, center, 4, pbe), atoms)
    for k, v in maps.items():
        o.set_map(k, *v)
    return o


CASES = [
    # which maps, salt, bcfl
    (("dielx", "diely", "dielz", "kappa", "charge"), True, "mdh"),
    (("dielx", "diely", "dielz"), False, "zero"),
    (("charge",), True, "mdh"),
    (("dielx", "diely", "dielz", "kappa", "pot"), True, "map"),
]


@pytest.mark.parametrize("kinds,salt,bcfl", CASES)
def test_maps_match_oracle(kinds, salt, bcfl):
    atoms = synth.globule(40, (5.0,) * 3, 17, net_charge=2)[:3]
    c = od.mol_center(atoms[0])
    allmaps = foreign_maps(c, 29)
    maps = {k: allmaps[k] for k in kinds}
    pbe = dict(pdie=2.0, sdie=78.54, srfm="mol", chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl=bcfl, ions=SALT if salt else [], etol=1e-12)
    print(f"maps {kinds} salt {salt} bcfl {bcfl}", flush=True)
    g = gpu_calc(atoms, c, pbe, maps)
    g.fillco()
    o = orc_calc(atoms, c, pbe, maps)
    o.fillco()
    for name in ("epsx", "epsy", "epsz", "kappa", "charge"):
        a, b = g.array(name), o.array(name)
        assert np.array_equal(a, b), (name, np.abs(a - b).max())
    if "kappa" in kinds:
        # the kappa map's maximum exceeds 1: scaled to 1
        assert g.array("kappa").max() == 1.0
    for a, b in zip(g.boundary(), o.boundary()):
        if bcfl == "map":
            assert np.abs(b).max() > 0
            assert np.array_equal(a, b)
        elif bcfl == "mdh":
            assert np.abs(a - b).max() <= 8 * np.finfo(float).eps * np.abs(b).max()
    g.solve()
    o.solve()
    g.set_part_domain()
    o.set_part_domain()
    eg, eo = g.energy(1), o.energy(1)
    assert abs(eg - eo) <= 1e-9 * abs(eo), (eg, eo)


def test_map_errors_are_reported():
    """Vpmg_fillco fails (and the kernels stay in bounds) for a charge map
    short of the grid, a negative kappa map, dielectric maps with salt but no
    kappa map, and bcfl map without a potential map or off it."""
    atoms = synth.globule(20, (4.0,) * 3, 3)[:3]
    c = od.mol_center(atoms[0])
    allmaps = foreign_maps(c, 5)
    pbe = dict(pdie=2.0, sdie=78.54, srfm="mol", chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=SALT)
    q, lo, h = allmaps["charge"]
    bad = [
        {"charge": (q[:, :, :10], lo, h)},
        {k: allmaps[k] for k in ("dielx", "diely", "dielz")},
        {k: (allmaps[k] if k != "kappa" else (-allmaps[k][0], lo, h))
         for k in ("dielx", "diely", "dielz", "kappa")},
    ]
    for maps in bad:
        g = gpu_calc(atoms, c, pbe, maps)
        with pytest.raises(RuntimeError):
            g.fillco()
        g.release()
    g = gpu_calc(atoms, c, dict(pbe, bcfl="map"), {})
    with pytest.raises(RuntimeError):
        g.fillco()
    g.release()
    p, _, h = allmaps["pot"]
    g = gpu_calc(atoms, c, dict(pbe, bcfl="map"), {"pot": (p, lo + 3.0, h)})
    with pytest.raises(RuntimeError):
        g.fillco()
    g.release()
    # and a good map still works afterwards
    g = gpu_calc(atoms, c, pbe, {"charge": allmaps["charge"]})
    g.fillco()
    g.release()


# ------------------------------------------------------------ input files

MOL = os.path.join(HERE, "golden", "examples", "solv", "methanol.pqr")

ELEC = """elec name {name}
    {method}
    dime 33 33 33
    glen 12 12 12
    gcent mol 1
    mol 1
    lpbe
    bcfl mdh
    ion charge 1 conc 0.1 radius 2.0
    ion charge -1 conc 0.1 radius 2.0
    pdie 2.0
    sdie 78.54
    chgm spl2
    srfm mol
    srad 1.4
    swin 0.3
    sdens 10.0
    temp 298.15
    calcenergy total
    calcforce no
{extra}end
"""


def _run(text, base, outdir):
    from apbs_amd import nosh, run
    buf = io.StringIO()
    run.Runner(nosh.parse(text, base=base), out=buf, outdir=outdir).run()
    return buf.getvalue()


def _global(text):
    import re
    return [float(v) for v in re.findall(r"Global net ELEC energy = (\S+)", text)]


def test_dummy_written_maps_reproduce_the_solve(tmp_path):
    """The helix example's pattern (examples/helix/Apbs_dummy-TEMPLATE.in,
    Apbs_solv-TEMPLATE.in): an mg-dummy calculation writes the dielectric,
    kappa and charge maps; a second input reads them (READ diel/kappa/charge
    dx, usemap) and must give the energy of the direct solve, to the 7
    significant digits the DX files carry."""
    writes = "".join(f"    write {w} dx {w}\n" for w in ("dielx", "diely", "dielz", "kappa",
                                                           "charge"))
    first = (f"read mol pqr {MOL} end\n" + ELEC.format(name="dummy", method="mg-dummy",
                                                       extra=writes)
             + ELEC.format(name="direct", method="mg-manual", extra="")
             + "print elecEnergy dummy end\nprint elecEnergy direct end\nquit\n")
    out1 = _run(first, str(tmp_path), str(tmp_path))
    assert "Skipping solve for mg-dummy run" in out1
    e_dummy, e_direct = _global(out1)
    assert e_dummy == 0.0 and e_direct != 0.0
    second = (f"read mol pqr {MOL}\n    diel dx dielx.dx diely.dx dielz.dx\n"
              "    kappa dx kappa.dx\n    charge dx charge.dx\nend\n"
              + ELEC.format(name="mapped", method="mg-manual",
                            extra="    usemap diel 1\n    usemap kappa 1\n    usemap charge 1\n")
              + "print elecEnergy mapped end\nquit\n")
    out2 = _run(second, str(tmp_path), str(tmp_path))
    (e_mapped,) = _global(out2)
    assert abs(e_mapped - e_direct) <= 1e-5 * abs(e_direct), (e_mapped, e_direct)


def test_bcfl_map_input_equals_focusing(tmp_path):
    """`bcfl map` with `usemap pot` on the coarse level's written potential
    gives the focused level's boundary, so its energy without the focusing
    correction matches a `bcfl focus` calculation to the DX precision."""
    coarse = ELEC.format(name="coarse", method="mg-manual", extra="    write pot dx coarse\n")
    coarse = coarse.replace("glen 12 12 12", "glen 24 24 24")
    fine = ELEC.format(name="fine", method="mg-manual", extra="").replace("bcfl mdh",
                                                                          "bcfl focus")
    first = (f"read mol pqr {MOL} end\n" + coarse + fine
             + "print elecEnergy fine end\nquit\n")
    out1 = _run(first, str(tmp_path), str(tmp_path))
    (e_focus,) = _global(out1)
    mapped = ELEC.format(name="mapped", method="mg-manual",
                         extra="    usemap pot 1\n").replace("bcfl mdh", "bcfl map")
    second = (f"read mol pqr {MOL}\n    pot dx coarse.dx\nend\n" + mapped
              + "print elecEnergy mapped end\nquit\n")
    out2 = _run(second, str(tmp_path), str(tmp_path))
    (e_map,) = _global(out2)
    assert abs(e_map - e_focus) <= 1e-5 * abs(e_focus), (e_map, e_focus)
