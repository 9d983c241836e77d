"""The oracle's srfm spl3 / spl4 windows and chgm spl4 charges, pinned by
closed forms (CPU).

- The polynomial windows of fillcoCoefSpline4 / fillcoCoefSpline3
  (vpmg.c:7052, 7525) are the smoothsteps t^4 (35 - 84 t + 70 t^2 - 20 t^3) and
  t^3 (10 - 15 t + 6 t^2) of t = (s - b) / (e - b) over [b, e] = [r - w, r + w],
  expanded in the distance s: the dielectric maps of one atom must be
  eps_p + (eps_s - eps_p) chi(s) at every face point.
- spl4's kappa window is the same over [r + r_ion - w, r + r_ion + w].  spl3's
  kappa coefficients are the dielectric ones divided once more by (2w)^5
  (vpmg.c:7632-7643), so its window is chi_d(s) / (2w)^5 clamped to [0, 1]:
  over the ion window that is 1, and kappa is a step at r + r_ion - w.
- The 5th-order B-spline of chgm spl4 (bspline4 at 2.5 + i - f) spreads a
  charge with unit total weight, first moment at the atom and variance 5/12 h^2
  per axis, whatever the atom's offset from the grid.
- Forces: the spl3 / spl4 ib and db terms against the energy's central
  difference (test_oracle_golden.test_forces_match_energy_gradient holds the
  method), and the spl4 fixed-charge force against the spl2 one.
"""
import numpy as np
import pytest

import orc_driver as od


def smoothstep(order, t):
    t = np.clip(t, 0.0, 1.0)
    if order == 4:
        return t ** 4 * (35 - 84 * t + 70 * t ** 2 - 20 * t ** 3)
    return t ** 3 * (10 - 15 * t + 6 * t ** 2)


def one_atom_maps(srfm, salt, radius=2.0, dime=65, h=0.15, chgm="spl2"):
    atoms = (np.zeros((1, 3)), np.array([1.0]), np.array([radius]))
    pbe = dict(pdie=2.0, sdie=78.54, srfm=srfm, chgm=chgm, srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=[(1, salt, 2.0), (-1, salt, 2.0)])
    o = od.OrcCalc(od.make_params([dime] * 3, [h] * 3, [0.0, 0.0, 0.0], 4, pbe), atoms)
    o.fillco()
    x = (np.arange(dime) - (dime - 1) / 2) * h
    return o, x


def monomial_bound(order, b, e, s):
    """Rounding bound of the reference's evaluation, which expands the window
    in powers of the distance s (not of t): eps * sum |c_n s^n| for the
    coefficients of the smoothstep written in s."""
    t_poly = (np.poly1d([-20, 70, -84, 35, 0, 0, 0, 0]) if order == 4
              else np.poly1d([6, -15, 10, 0, 0, 0]))
    in_s = t_poly(np.poly1d([1.0 / (e - b), -b / (e - b)]))   # t = (s - b)/(e - b)
    terms = sum(abs(cn) * s ** n for n, cn in enumerate(in_s.coeffs[::-1]))
    return 16 * np.finfo(float).eps * terms


@pytest.mark.parametrize("order", [3, 4])
def test_dielectric_window_is_the_smoothstep(order):
    o, x = one_atom_maps(f"spl{order}", 0.0)
    b, e = 2.0 - 0.3, 2.0 + 0.3
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    h = x[1] - x[0]
    for name, (dx, dy, dz) in (("epsx", (h / 2, 0, 0)), ("epsy", (0, h / 2, 0)),
                               ("epsz", (0, 0, h / 2))):
        s = np.sqrt((X + dx) ** 2 + (Y + dy) ** 2 + (Z + dz) ** 2)
        chi_want = smoothstep(order, (s - b) / (e - b))
        chi_got = (o.array(name) - 2.0) / (78.54 - 2.0)
        inside = (s > b) & (s < e)
        # the window is used: some face points fall inside it
        assert np.count_nonzero(inside) > 100
        bound = np.where(inside, monomial_bound(order, b, e, s), 0.0) + 1e-13
        err = np.abs(chi_got - chi_want)
        assert np.all(err <= bound), (name, (err / bound).max())


def test_spl4_kappa_window_is_the_smoothstep():
    o, x = one_atom_maps("spl4", 0.15)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    s = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    b, e = 2.0 + 2.0 - 0.3, 2.0 + 2.0 + 0.3
    want = smoothstep(4, (s - b) / (e - b))
    inside = (s > b) & (s < e)
    assert np.count_nonzero(inside) > 100
    bound = np.where(inside, monomial_bound(4, b, e, s), 0.0) + 1e-13
    assert np.all(np.abs(o.array("kappa") - want) <= bound)


def test_spl3_kappa_is_a_step_at_the_inner_ion_radius():
    o, x = one_atom_maps("spl3", 0.15)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    s2 = X ** 2 + Y ** 2 + Z ** 2
    want = (s2 > (2.0 + 2.0 - 0.3) ** 2).astype(float)
    assert np.array_equal(o.array("kappa"), want)


@pytest.mark.parametrize("offset", [(0.0, 0.0, 0.0), (0.3, 0.7, 0.45), (0.5, 0.5, 0.5)])
def test_spl4_charge_moments(offset):
    h = 0.5
    pos = np.array(offset) * h + np.array([0.25, -0.5, 1.0])
    atoms = (pos[None, :], np.array([1.0]), np.array([1.5]))
    maps = {}
    for chgm in ("spl4", "spl2"):
        pbe = dict(pdie=2.0, sdie=78.54, srfm="spl2", chgm=chgm, srad=1.4, swin=0.3,
                   sdens=10.0, temp=298.15, bcfl="mdh", ions=[])
        o = od.OrcCalc(od.make_params([33] * 3, [h] * 3, [0.0, 0.0, 0.0], 4, pbe), atoms)
        o.fillco()
        maps[chgm] = o.array("charge")
    q = maps["spl4"]
    x = (np.arange(33) - 16) * h
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    W = q.sum()
    assert W == pytest.approx(maps["spl2"].sum(), rel=1e-13)
    for coord, c in ((X, pos[0]), (Y, pos[1]), (Z, pos[2])):
        assert (q * coord).sum() / W == pytest.approx(c, abs=1e-12)
        assert (q * (coord - c) ** 2).sum() / W == pytest.approx(5.0 / 12.0 * h * h, rel=1e-12)
    # at most 6 nodes per axis carry weight
    assert np.count_nonzero(q) <= 6 ** 3


@pytest.mark.parametrize("srfm,salt", [("spl4", 0.15), ("spl3", 0.0)])
def test_window_forces_match_energy_gradient(srfm, salt):
    """ib + db + qf of one ion of a pair against -dE/dx with a dielectric
    jump.  spl4 also with salt (both boundary terms non-zero).  spl3 without:
    its kappa map uses the reference's mis-scaled coefficients (a step) while
    its ion-boundary force differentiates the proper window, so in the
    reference too that force is not the gradient of the energy."""
    kT = od.kT_to_kJmol(1.0, 298.15)
    q, r = np.array([1.0, -1.0]), np.array([2.0, 1.8])
    pbe = dict(pdie=2.0, sdie=78.54, srfm=srfm, chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl="mdh", ions=[(1, salt, 2.0), (-1, salt, 2.0)], etol=1e-12)

    def run(xyz):
        return od.run_manual((xyz, q, r), [65] * 3, [0.0, 0.0, 0.0], pbe, grid=[0.3] * 3)

    x0 = np.array([[-3.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    _, c = run(x0)
    rc, qf, ib, db = c.force(0)
    assert rc == 1
    assert np.abs(db).max() > 0
    assert (np.abs(ib).max() > 0) == (salt > 0)
    f = qf + ib + db
    d = 0.02
    xp, xm = x0.copy(), x0.copy()
    xp[0, 0] += d
    xm[0, 0] -= d
    fd = -(run(xp)[0] - run(xm)[0]) / (2 * d) / kT
    assert abs(f[0] - fd) < 0.01 * abs(fd), (srfm, f[0], fd)


def test_spl4_fixed_charge_force_matches_spl2():
    """Two ions on grid points in a uniform dielectric without salt: only the
    fixed-charge force acts, the self-force vanishes by symmetry, and the 5th-
    and 3rd-order B-spline fields must agree to 1 %."""
    q, r = np.array([1.0, -1.0]), np.array([2.0, 1.8])
    x0 = np.array([[-3.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    out = {}
    for chgm in ("spl4", "spl2"):
        pbe = dict(pdie=78.54, sdie=78.54, srfm="spl2", chgm=chgm, srad=1.4, swin=0.3,
                   sdens=10.0, temp=298.15, bcfl="mdh", ions=[], etol=1e-12)
        _, c = od.run_manual((x0, q, r), [65] * 3, [0.0, 0.0, 0.0], pbe, grid=[0.3] * 3)
        rc, qf, ib, db = c.force(0)
        assert rc == 1 and np.all(ib == 0) and np.all(db == 0)
        out[chgm] = qf
    assert out["spl4"][0] > 0          # attracted towards the other ion, at +x
    assert abs(out["spl4"][0] - out["spl2"][0]) < 0.01 * abs(out["spl2"][0])
    assert abs(out["spl4"][1]) < 1e-9 and abs(out["spl4"][2]) < 1e-9
