"""The oracle's multigrid pieces against what the mathematics fixes, and its
fill/solve against closed forms.  A restatement that drops a term, flips a
sign, mixes up a stencil direction or transposes an operator fails these.

  * Galerkin coarse operator == P^T A P assembled densely (brute force);
  * restriction == P^T and interpolation == P (dense trilinear P);
  * residual == f - (A + diag(cc)) x with A assembled from the stencil arrays;
  * one red-black sweep == the colour-by-colour simultaneous update;
  * uniform dielectric, no salt: the potential of a unit charge is Coulomb's;
  * uniform dielectric with salt: the Debye-Hueckel sphere.
"""
import ctypes

import numpy as np
import pytest

import orc_driver as od

# stored stencil directions of src/pmgc/gsd.c:74-87 (0-based x, y, z offsets)
FWD = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, 1),
       (0, 1, 1), (0, -1, 1), (1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1)]
dp = ctypes.POINTER(ctypes.c_double)


def P_(a):
    return a.ctypes.data_as(dp)


@pytest.fixture(scope="module")
def O():
    return od.lib()


def interior_index(n3):
    nx, ny, nz = n3
    idx = -np.ones((nz, ny, nx), dtype=int)
    pts = [(i, j, k) for k in range(1, nz - 1) for j in range(1, ny - 1) for i in range(1, nx - 1)]
    for m, (i, j, k) in enumerate(pts):
        idx[k, j, i] = m
    return idx, pts


def dense_operator(a, n3, st):
    """The matrix the stencil arrays stand for, on interior unknowns."""
    nx, ny, nz = n3
    n = nx * ny * nz
    arr = a.reshape(-1, n)
    idx, pts = interior_index(n3)
    A = np.zeros((len(pts), len(pts)))
    ns = 4 if st == 7 else 14
    for m, (i, j, k) in enumerate(pts):
        p = (k * ny + j) * nx + i
        A[m, m] = arr[0, p]
        for s in range(1, ns):
            ex, ey, ez = FWD[s]
            i2, j2, k2 = i + ex, j + ey, k + ez
            if 0 < i2 < nx - 1 and 0 < j2 < ny - 1 and 0 < k2 < nz - 1:
                m2 = idx[k2, j2, i2]
                A[m, m2] = -arr[s, p]
                A[m2, m] = -arr[s, p]
    return A


def dense_prolongation(f3, c3):
    _, fpts = interior_index(f3)
    _, cpts = interior_index(c3)
    P = np.zeros((len(fpts), len(cpts)))
    for a, f in enumerate(fpts):
        for b, c in enumerate(cpts):
            w = 1.0
            for d in range(3):
                off = abs(f[d] - 2 * c[d])
                w *= 1.0 if off == 0 else (0.5 if off == 1 else 0.0)
            P[a, b] = w
    return P


def random_operator(rng, n3, st):
    nx, ny, nz = n3
    n = nx * ny * nz
    na = 4 if st == 7 else 14
    a = np.zeros((na, nz, ny, nx))
    inner = (slice(1, -1),) * 3
    for s in range(1, na):
        a[(s,) + inner] = rng.uniform(0.1, 1.0, (nz - 2, ny - 2, nx - 2))
    a[(0,) + inner] = rng.uniform(6.0, 12.0, (nz - 2, ny - 2, nx - 2)) * (1 if st == 7 else 3)
    return np.ascontiguousarray(a.reshape(na * n))


def vec_interior(x, n3):
    nx, ny, nz = n3
    return x.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1].ravel()


@pytest.mark.parametrize("st", [7, 27])
def test_galerkin_is_PtAP(O, st):
    rng = np.random.default_rng(1 + st)
    f3, c3 = (9, 11, 9), (5, 6, 5)
    af = random_operator(rng, f3, st)
    ac = np.zeros(14 * np.prod(c3))
    O.orc_galerkin(*f3, st, P_(af), *c3, P_(ac))
    Af = dense_operator(af, f3, st)
    P = dense_prolongation(f3, c3)
    want = P.T @ Af @ P
    got = dense_operator(ac, c3, 27)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    assert np.allclose(want, want.T)


def test_transfers_are_P_and_Pt(O):
    rng = np.random.default_rng(4)
    f3, c3 = (9, 11, 13), (5, 6, 7)
    P = dense_prolongation(f3, c3)
    xf = rng.normal(size=np.prod(f3))
    xc = np.zeros(np.prod(c3))
    O.orc_restrict(*f3, *c3, P_(xf.copy()), P_(xc))
    assert np.allclose(vec_interior(xc, c3), P.T @ vec_interior(xf, f3), rtol=1e-14, atol=1e-14)
    yc = np.zeros(np.prod(c3))
    yc.reshape(c3[2], c3[1], c3[0])[1:-1, 1:-1, 1:-1] = rng.normal(size=(c3[2] - 2, c3[1] - 2,
                                                                         c3[0] - 2))
    yf = np.zeros(np.prod(f3))
    O.orc_interp(*c3, *f3, P_(yc.copy()), P_(yf))
    assert np.allclose(vec_interior(yf, f3), P @ vec_interior(yc, c3), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("st", [7, 27])
def test_residual_is_f_minus_Ax(O, st):
    rng = np.random.default_rng(9 + st)
    n3 = (8, 9, 10)
    n = np.prod(n3)
    a = random_operator(rng, n3, st)
    cc = rng.uniform(0, 1, n)
    fc = rng.normal(size=n)
    x = np.zeros(n)
    x.reshape(n3[2], n3[1], n3[0])[1:-1, 1:-1, 1:-1] = rng.normal(size=(n3[2] - 2, n3[1] - 2,
                                                                       n3[0] - 2))
    r = np.zeros(n)
    O.orc_resid(*n3, st, P_(a), P_(cc), P_(fc), P_(x), P_(r))
    A = dense_operator(a, n3, st) + np.diag(vec_interior(cc, n3))
    want = vec_interior(fc, n3) - A @ vec_interior(x, n3)
    assert np.allclose(vec_interior(r, n3), want, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("iadjoint", [0, 1])
def test_red_black_sweep_is_colourwise_update(O, iadjoint):
    """gsd.c:95-165: one sweep updates all points of one colour at once from
    the current values, then the other colour; red = 1-based i+j+k even."""
    rng = np.random.default_rng(5 + iadjoint)
    n3 = (9, 10, 11)
    nx, ny, nz = n3
    n = np.prod(n3)
    a = random_operator(rng, n3, 7)
    cc, fc = rng.uniform(0, 1, n), rng.normal(size=n)
    x0 = np.zeros(n)
    x0.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1] = rng.normal(size=(nz - 2, ny - 2, nx - 2))
    x = x0.copy()
    O.orc_gsrb(*n3, 7, P_(a), P_(cc), P_(fc), P_(x), 1, iadjoint, 0)
    A = dense_operator(a, n3, 7)
    D = np.diag(A) + vec_interior(cc, n3)
    off = A - np.diag(np.diag(A))
    _, pts = interior_index(n3)
    par = np.array([(i + j + k) % 2 for (i, j, k) in pts])   # 0-based
    red = par == 1                                          # 1-based even
    y = vec_interior(x0, n3).copy()
    f = vec_interior(fc, n3)
    for colour in ([red, ~red] if iadjoint == 0 else [~red, red]):
        upd = (f - off @ y) / D
        y[colour] = upd[colour]
    assert np.allclose(vec_interior(x, n3), y, rtol=1e-13, atol=1e-13)


def pot_along_axis(calc, center, h, dime):
    u = calc.array("u")
    c = dime // 2
    r = np.arange(1, c) * h
    return r, u[c, c, c + 1:c + c]


def test_uniform_dielectric_point_charge_is_coulomb():
    """pdie = sdie, no salt, bcfl mdh: the boundary is exact and the interior
    solution tends to q / (4 pi eps0 eps r) (kT/e) away from the charge."""
    eps, temp, h, dime = 2.0, 298.15, 0.5, 65
    pbe = dict(pdie=eps, sdie=eps, srfm="spl2", chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=temp, bcfl="mdh", ions=[(1, 0.0, 2.0), (-1, 0.0, 2.0)], etol=1e-12)
    atoms = (np.zeros((1, 3)), np.array([1.0]), np.array([1.0]))
    _, calc = od.run_manual(atoms, [dime] * 3, [0.0, 0.0, 0.0], pbe, grid=[h] * 3)
    r, u = pot_along_axis(calc, 0.0, h, dime)
    ec = 1.6021773e-19
    pre1 = ec * ec / (4 * np.pi * 8.8541878e-12 * eps * 1.3806581e-23 * temp) * 1e10
    sel = (r >= 3.0) & (r <= 12.0)
    rel = np.abs(u[sel] - pre1 / r[sel]) / (pre1 / r[sel])
    assert rel.max() < 0.01, rel.max()


def test_uniform_dielectric_salt_is_debye_hueckel():
    """pdie = sdie with 0.15 M salt: outside the ion-exclusion sphere of
    radius a the potential is q e^{-kappa (r - a)} / (4 pi eps0 eps r
    (1 + kappa a)) (kT/e).  The ion radius is 0 so that the exclusion radius
    (atom + ion radius, vpmg.c:4631-4854) equals the radius the MDH boundary
    uses (the atom radius, vpmg.c:3289-3367) and both describe one sphere."""
    eps, temp, h, dime = 78.54, 298.15, 0.4, 65
    a = 3.0
    pbe = dict(pdie=eps, sdie=eps, srfm="spl2", chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=temp, bcfl="mdh", ions=[(1, 0.15, 0.0), (-1, 0.15, 0.0)], etol=1e-12)
    atoms = (np.zeros((1, 3)), np.array([1.0]), np.array([a]))
    _, calc = od.run_manual(atoms, [dime] * 3, [0.0, 0.0, 0.0], pbe, grid=[h] * 3)
    kappa = calc.stats()["xkappa"]
    r, u = pot_along_axis(calc, 0.0, h, dime)
    ec = 1.6021773e-19
    pre1 = ec * ec / (4 * np.pi * 8.8541878e-12 * eps * 1.3806581e-23 * temp) * 1e10
    want = pre1 * np.exp(-kappa * (r - a)) / (r * (1 + kappa * a))
    sel = r >= 4.5
    rel = np.abs(u[sel] - want[sel]) / want[sel]
    assert rel.max() < 0.005, rel.max()


def _boundary(atoms, bcfl, pdie, sdie, salt, dime=33, h=1.0):
    pbe = dict(pdie=pdie, sdie=sdie, srfm="spl2", chgm="spl2", srad=1.4, swin=0.3, sdens=10.0,
               temp=298.15, bcfl=bcfl, ions=[(1, salt, 2.0), (-1, salt, 2.0)])
    c = od.mol_center(atoms[0])
    o = od.OrcCalc(od.make_params([dime] * 3, [h] * 3, c, 4, pbe), atoms)
    o.fillco()
    return np.concatenate(o.boundary())


@pytest.mark.parametrize("salt", [0.0, 0.15])
def test_sdh_of_one_centred_ion_is_mdh(salt):
    """bcfl sdh (vpmg.c:3479-3645) for a single ion: the sphere is the ion
    (solute centre and radius are the atom's), the moments reduce to the
    charge, and the boundary equals the MDH sum (vpmg.c:3289-3367)."""
    atoms = (np.array([[0.3, -0.2, 0.1]]), np.array([1.5]), np.array([2.0]))
    sdh = _boundary(atoms, "sdh", 2.0, 78.54, salt)
    mdh = _boundary(atoms, "mdh", 2.0, 78.54, salt)
    assert np.allclose(sdh, mdh, rtol=1e-13, atol=0)


def test_sdh_far_field_is_the_multipole_expansion():
    """pdie = sdie, no salt: the SDH factors are the vacuum multipole terms,
    so at a boundary ~30 A away from a ~2 A molecule the SDH values equal the
    exact Coulomb sum (MDH) up to the octupole remainder."""
    rng = np.random.default_rng(8)
    xyz = rng.uniform(-1.0, 1.0, (6, 3))
    q = np.array([0.8, -0.5, 0.3, -0.9, 0.6, -0.3])
    atoms = (xyz, q, np.full(6, 1.2))
    sdh = _boundary(atoms, "sdh", 78.54, 78.54, 0.0, dime=65, h=1.0)
    mdh = _boundary(atoms, "mdh", 78.54, 78.54, 0.0, dime=65, h=1.0)
    err = np.abs(sdh - mdh).max() / np.abs(mdh).max()
    assert err < 2e-3, err
    # the dipole and quadrupole terms matter: a monopole-only boundary is far off
    assert np.abs(mdh).max() > 0
