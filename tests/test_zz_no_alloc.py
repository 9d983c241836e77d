"""Vpmg_energyNoAlloc against Vpmg_energy, through the C ABI.

The allocation-free entry point runs the same reduction kernels as
Vpmg_energy, in the library's static device buffer instead of the object's
buffers, so the two energies are bitwise equal.  Checked on calls in a row
whose inputs shrink: a 65^3 npbe calculation (charge-volume, dielectric and
ionic energies), then a 33^3 lpbe one with point charges (chgm spl0, one
partial per atom in the buffer) and fewer atoms, then the first again (the
buffer keeps the previous call's contents, so the chop counter must be
cleared by the call itself).  Over-size input is refused before any device
work: the point-charge energy of more atoms than the buffer holds.
"""
import ctypes

import numpy as np
import pytest

from apbs_amd import driver, lib, synth

pytestmark = pytest.mark.gpu

SALT = [(1, 0.15, 2.0), (-1, 0.15, 2.0)]
SCRATCH_ATOMS = 1 << 17


def calc(n_atoms, radius, seed, dime, h, npbe, chgm):
    xyz, q, r, _ = synth.globule(n_atoms, (radius,) * 3, seed, net_charge=-2)
    mol = driver.Molecule(xyz, q, r)
    pbe = dict(pdie=2.0, sdie=78.54, srfm="spl2", chgm=chgm, srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=SALT, npbe=npbe)
    e, c = driver.run_manual(mol, [dime] * 3, mol.center, pbe, grid=[h] * 3, mirror=False)
    c._keep = mol
    return c


def no_alloc_energy(c, ext=1):
    out = ctypes.c_double(np.nan)
    rc = lib().Vpmg_energyNoAlloc(c.pmg, ext, ctypes.byref(out))
    return rc, out.value


def test_no_alloc_energy_matches_energy_on_shrinking_inputs():
    big = calc(400, 10.0, 11, 65, 0.5, True, "spl2")
    small = calc(60, 5.0, 12, 33, 0.5, False, "spl0")
    for c, label in ((big, "npbe 65^3 spl2"), (small, "lpbe 33^3 spl0"),
                     (big, "npbe 65^3 spl2 again")):
        want = c.energy(1)
        rc, got = no_alloc_energy(c)
        print(f"{label}: Vpmg_energy {want!r}, Vpmg_energyNoAlloc {got!r} (rc {rc})", flush=True)
        assert rc == 1
        assert got == want, (label, got, want)
        assert np.isfinite(got) and got != 0.0


def test_no_alloc_energy_refuses_oversize_point_charge_input(capfd):
    n = SCRATCH_ATOMS + 1
    rng = np.random.default_rng(5)
    xyz = rng.uniform(-20.0, 20.0, (n, 3))
    mol = driver.Molecule(xyz, np.zeros(n), np.full(n, 1.5))
    pbe = dict(pdie=2.0, sdie=78.54, srfm="spl2", chgm="spl0", srad=1.4, swin=0.3,
               sdens=10.0, temp=298.15, bcfl="mdh", ions=SALT)
    c = driver.Calc(mol, driver.manual_mgparm([17] * 3, [0.0] * 3, grid=[3.0] * 3,
                                              chgm="spl0"), pbe, mirror=False)
    rc, _ = no_alloc_energy(c)
    err = capfd.readouterr().err
    assert rc == 0
    assert "static buffer" in err, err
    c.release()
