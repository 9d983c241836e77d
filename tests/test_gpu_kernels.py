"""Kernel-by-kernel parity of the HIP path against the oracle.

Fine-level kernels (the fused 7-point red-black sweep, residuals, restriction,
prolongation, VbuildA_fv, Galerkin) evaluate the reference's expressions in
the reference's order without FMA contraction, so the bar is bitwise
equality.  The 8-colour 27-point sweep is compared with the oracle's
8-colour variant, also bitwise.  The nonlinear term goes through exp(),
whose device and glibc implementations may differ in the last ulp: 4 ulp.
"""
import ctypes

import numpy as np
import pytest

import orc_driver as od

pytestmark = pytest.mark.gpu

dp = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(dp)


@pytest.fixture(scope="module")
def L():
    import apbs_amd
    return ctypes.CDLL(apbs_amd.LIB_PATH)


@pytest.fixture(scope="module")
def O():
    return od.lib()


def rand_op(rng, n3, st):
    nx, ny, nz = n3
    n = nx * ny * nz
    na = 4 if st == 7 else 14
    a = np.zeros((na, nz, ny, nx))
    inner = (slice(1, -1),) * 3
    for s in range(1, na):
        a[(s,) + inner] = rng.uniform(0.1, 1.0, (nz - 2, ny - 2, nx - 2))
    a[(0,) + inner] = rng.uniform(6.0, 12.0, (nz - 2, ny - 2, nx - 2)) * (1 if st == 7 else 3)
    return np.ascontiguousarray(a.reshape(na * n))


def interior(x, n3):
    nx, ny, nz = n3
    return x.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("st,n3", [(7, (17, 19, 21)), (7, (33, 33, 33)), (7, (67, 13, 20)),
                                   (7, (130, 6, 11)), (7, (5, 70, 3)), (27, (17, 17, 17)),
                                   (27, (9, 11, 13)), (27, (45, 43, 41))])
@pytest.mark.parametrize("adj", [0, 1])
def test_gsrb_bitwise(L, O, st, n3, adj):
    rng = np.random.default_rng(7 + st + adj)
    nx, ny, nz = n3
    n = nx * ny * nz
    a = rand_op(rng, n3, st)
    cc = rng.uniform(0, 1, n)
    fc = rng.normal(size=n)
    x0 = np.zeros(n)
    x0.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1] = rng.normal(size=(nz - 2, ny - 2, nx - 2))
    xg = x0.copy()
    xo = x0.copy()
    L.pmgtest_gsrb(nx, ny, nz, st, P(a), P(cc), P(fc), P(xg), 2, adj)
    O.orc_gsrb(nx, ny, nz, st, P(a), P(cc), P(fc), P(xo), 2, adj, 1)
    assert np.array_equal(xg, xo)


@pytest.mark.parametrize("n3", [(17, 17, 17), (9, 11, 13), (45, 43, 41), (5, 5, 4),
                                (70, 21, 6), (81, 81, 9), (3, 3, 3), (130, 7, 6), (129, 6, 7)])
@pytest.mark.parametrize("mode", [1, 2, 3, 8, 16, 32])
def test_gsrb27_plane_sweeps_bitwise(L, O, n3, mode):
    """The plane-parity 27-point sweep (one launch per plane parity, tiles
    with recomputed margins) and the colour-pair sweep (mode 2 the vector
    kernel, 16-byte runs per lane; mode 3 one wavefront per row) give the
    oracle's 8-colour sweep bit for bit: every
    tile size, partial tiles, forward and adjoint, even and odd sweep
    counts, rows of 1 to 128 interior points."""
    rng = np.random.default_rng(sum(n3) + mode)
    nx, ny, nz = n3
    n = nx * ny * nz
    a = rand_op(rng, n3, 27)
    cc, fc = rng.uniform(0, 1, n), rng.normal(size=n)
    x0 = np.zeros(n)
    x0.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1] = rng.normal(size=(nz - 2, ny - 2, nx - 2))
    for adj in (0, 1):
        for nu in (2, 3):
            xg, xo = x0.copy(), x0.copy()
            L.pmgtest_gsrb27_mode(nx, ny, nz, P(a), P(cc), P(fc), P(xg), nu, adj, mode)
            O.orc_gsrb(nx, ny, nz, 27, P(a), P(cc), P(fc), P(xo), nu, adj, 1)
            assert np.array_equal(xg, xo), (n3, mode, adj, nu)


@pytest.mark.parametrize("tile", [416, 48, 816, 88, 84, 1616, 832, 44, 68, 640048, 800048,
                                  960048, 800044, 800068, 960416, 800816, 5088, 645084,
                                  805816, 7048, 7088, 647048, 807088, 7410, 807412, 967416,
                                  647410, 807816, 807412, 967816])
@pytest.mark.parametrize("n3", [(133, 21, 38), (133, 21, 41)])
def test_fused_sweep_tiles_bitwise(L, O, tile, n3):
    """Every tile shape of the fused red-black sweep gives the oracle's
    sweep bit for bit (partial tiles in every direction; 41 planes leave an
    odd number of planes in the last chunk)."""
    rng = np.random.default_rng(tile)
    nx, ny, nz = n3
    n = nx * ny * nz
    a = rand_op(rng, n3, 7)
    cc, fc = rng.uniform(0, 1, n), rng.normal(size=n)
    x0 = np.zeros(n)
    x0.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1] = rng.normal(size=(nz - 2, ny - 2, nx - 2))
    for adj in (0, 1):
        xg, xo = x0.copy(), x0.copy()
        L.pmgtest_gsrb_tile(nx, ny, nz, P(a), P(cc), P(fc), P(xg), 2, adj, tile)
        O.orc_gsrb(nx, ny, nz, 7, P(a), P(cc), P(fc), P(xo), 2, adj, 1)
        assert np.array_equal(xg, xo), (tile, adj)


@pytest.mark.parametrize("n3", [(17, 19, 21), (161, 23, 29), (133, 21, 38)])
@pytest.mark.parametrize("tile", [0, 5088, 800048, 416, 7048, 807088])
def test_fused_zero_start_bitwise(L, O, n3, tile):
    """The zero-start form of the fused sweep pair (the first smoothing of a
    mixed-precision correction cycle) ignores the interior of x (NaN here)
    and gives the oracle's two sweeps from x = 0, in the default shapes and,
    by clearing x first, in a shape without that form (416)."""
    rng = np.random.default_rng(sum(n3) + tile)
    nx, ny, nz = n3
    n = nx * ny * nz
    a = rand_op(rng, n3, 7)
    cc, fc = rng.uniform(0, 1, n), rng.normal(size=n)
    for adj in (0, 1):
        xg = np.zeros(n)
        xg.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1] = np.nan
        xo = np.zeros(n)
        assert L.pmgtest_gsrb_zero(nx, ny, nz, P(a), P(cc), P(fc), P(xg), adj, tile) == 1
        O.orc_gsrb(nx, ny, nz, 7, P(a), P(cc), P(fc), P(xo), 2, adj, 1)
        assert np.array_equal(xg, xo), (n3, tile, adj)


@pytest.mark.parametrize("n3", [(161, 23, 9), (81, 81, 9), (17, 19, 21), (7, 9, 11), (5, 8, 9),
                                (70, 6, 5)])
def test_resid7_fp32_bitwise(L, n3):
    """The FP32 7-point residual (vector form from 6 points a row, one point
    per thread below) against numpy FP32 in acc7's order, bit for bit; r's
    boundary untouched."""
    rng = np.random.default_rng(sum(n3))
    nx, ny, nz = n3
    sh = (nz, ny, nx)
    f32 = np.float32
    oE, oN, uC = (rng.uniform(0.1, 1.0, sh).astype(f32) for _ in range(3))
    D = rng.uniform(6.0, 12.0, sh).astype(f32)
    fc, x = rng.normal(size=sh).astype(f32), rng.normal(size=sh).astype(f32)
    r = np.full(sh, 7.0, f32)
    assert L.pmgtest_resid32(nx, ny, nz, F32(oE), F32(oN), F32(uC), F32(D), F32(fc), F32(x),
                             F32(r)) == 1

    def s_(a, dk, dj, di):
        return a[1 + dk:nz - 1 + dk, 1 + dj:ny - 1 + dj, 1 + di:nx - 1 + di]

    acc = s_(fc, 0, 0, 0).copy()
    acc = acc + s_(oN, 0, 0, 0) * s_(x, 0, 1, 0)
    acc = acc + s_(oN, 0, -1, 0) * s_(x, 0, -1, 0)
    acc = acc + s_(oE, 0, 0, 0) * s_(x, 0, 0, 1)
    acc = acc + s_(oE, 0, 0, -1) * s_(x, 0, 0, -1)
    acc = acc + s_(uC, -1, 0, 0) * s_(x, -1, 0, 0)
    acc = acc + s_(uC, 0, 0, 0) * s_(x, 1, 0, 0)
    ref = np.full(sh, 7.0, f32)
    ref[1:-1, 1:-1, 1:-1] = acc - s_(D, 0, 0, 0) * s_(x, 0, 0, 0)
    assert np.array_equal(r, ref), n3


@pytest.mark.parametrize("n3", [(81, 81, 9), (45, 43, 41), (17, 17, 17), (9, 11, 13),
                                (7, 6, 5), (130, 7, 6), (129, 6, 7)])
@pytest.mark.parametrize("mode", [2, 3, 4, 1])
def test_gsrb27_fp32_forms_bitwise(L, n3, mode):
    """FP32 27-point sweeps: the colour-pair kernels (vector and one point
    per lane) and the plane-parity sweep give the per-colour kernel's bits
    (that kernel is part27 term for term, and in FP64 it is the oracle's
    8-colour sweep bit for bit), with a non-zero boundary."""
    rng = np.random.default_rng(sum(n3) * 3 + mode)
    nx, ny, nz = n3
    sh = (nz, ny, nx)
    f32 = np.float32
    a = rng.uniform(0.05, 0.5, (14,) + sh).astype(f32)
    D = rng.uniform(8.0, 16.0, sh).astype(f32)
    fc = rng.normal(size=sh).astype(f32)
    x0 = rng.normal(size=sh).astype(f32)
    for adj in (0, 1):
        for nu in (2, 3):
            ref, got = x0.copy(), x0.copy()
            assert L.pmgtest_gsrb27_32(nx, ny, nz, F32(a), F32(D), F32(fc), F32(ref), nu, adj,
                                       0) == 1
            assert L.pmgtest_gsrb27_32(nx, ny, nz, F32(a), F32(D), F32(fc), F32(got), nu, adj,
                                       mode) == 1
            assert np.array_equal(got, ref), (n3, mode, adj, nu)


def rb_sweeps(oE, oN, uC, D, fc, x, first, nu):
    """In-place red-black sweeps in the arrays' own precision, colour
    `first` ((i + j + k) & 1) first, acc7's terms in its order with every
    product and sum rounded (numpy does not contract)."""
    x = x.copy()
    nz, ny, nx = x.shape
    kk, jj, ii = np.meshgrid(np.arange(1, nz - 1), np.arange(1, ny - 1), np.arange(1, nx - 1),
                             indexing="ij")
    par = (ii + jj + kk) & 1

    def sh(a, dk, dj, di):
        return a[1 + dk:nz - 1 + dk, 1 + dj:ny - 1 + dj, 1 + di:nx - 1 + di]

    for _ in range(nu):
        for colour in (first, 1 - first):
            acc = sh(fc, 0, 0, 0).copy()
            acc = acc + sh(oN, 0, 0, 0) * sh(x, 0, 1, 0)
            acc = acc + sh(oN, 0, -1, 0) * sh(x, 0, -1, 0)
            acc = acc + sh(oE, 0, 0, 0) * sh(x, 0, 0, 1)
            acc = acc + sh(oE, 0, 0, -1) * sh(x, 0, 0, -1)
            acc = acc + sh(uC, -1, 0, 0) * sh(x, -1, 0, 0)
            acc = acc + sh(uC, 0, 0, 0) * sh(x, 1, 0, 0)
            new = acc / sh(D, 0, 0, 0)
            xi = x[1:-1, 1:-1, 1:-1]
            xi[par == colour] = new[par == colour]
    return x


def test_rb_sweeps_reference_is_the_oracle(O):
    """The numpy sweep used for the FP32 kernels, run in FP64, gives the
    oracle's sweeps bit for bit (so it follows the reference's colour order
    and term order)."""
    rng = np.random.default_rng(5)
    n3 = (13, 11, 9)
    nx, ny, nz = n3
    n = nx * ny * nz
    a = rand_op(rng, n3, 7)
    cc, fc = rng.uniform(0, 1, n), rng.normal(size=n)
    x0 = rng.normal(size=n)
    for adj in (0, 1):
        xo = x0.copy()
        O.orc_gsrb(nx, ny, nz, 7, P(a), P(cc), P(fc), P(xo), 2, adj, 1)
        A = a.reshape(4, nz, ny, nx)
        D = A[0] + cc.reshape(nz, ny, nx)
        xr = rb_sweeps(A[1], A[2], A[3], D, fc.reshape(nz, ny, nx), x0.reshape(nz, ny, nx),
                       0 if adj else 1, 2)
        assert np.array_equal(xr.ravel(), xo), adj


fp = ctypes.POINTER(ctypes.c_float)


def F32(a):
    return a.ctypes.data_as(fp)


@pytest.mark.parametrize("n3", [(133, 21, 38), (133, 21, 41), (17, 19, 21), (161, 23, 29),
                                (7, 9, 11), (5, 8, 9)])
@pytest.mark.parametrize("tile", [0, 48, 5088, 7048, 7088, 7410, 7816, 647412, 967416, 967816])
def test_fused_fp32_bitwise(L, n3, tile):
    """The FP32 fused sweeps (scalar and vector forms, full and zero-start)
    against the numpy FP32 sweeps, bit for bit, with a non-zero boundary (a
    vector form needs rows of 6 points; narrower grids take the scalar
    form)."""
    rng = np.random.default_rng(sum(n3) * 7 + tile)
    nx, ny, nz = n3
    sh = (nz, ny, nx)
    f32 = np.float32
    oE, oN, uC = (rng.uniform(0.1, 1.0, sh).astype(f32) for _ in range(3))
    D = rng.uniform(6.0, 12.0, sh).astype(f32)
    fc = rng.normal(size=sh).astype(f32)
    x0 = rng.normal(size=sh).astype(f32)
    for adj in (0, 1):
        for nu in (1, 2, 3):
            xg = x0.copy()
            assert L.pmgtest_fused32(nx, ny, nz, F32(oE), F32(oN), F32(uC), F32(D), F32(fc),
                                     F32(xg), nu, adj, tile, 0) == 1
            xr = rb_sweeps(oE, oN, uC, D, fc, x0, 0 if adj else 1, nu)
            assert np.array_equal(xg, xr), (tile, adj, nu)
        # the zero-start pair: zero boundary, interior not read
        xz = np.zeros(sh, f32)
        xg = xz.copy()
        xg[1:-1, 1:-1, 1:-1] = np.nan
        assert L.pmgtest_fused32(nx, ny, nz, F32(oE), F32(oN), F32(uC), F32(D), F32(fc),
                                 F32(xg), 2, adj, tile, 1) == 1
        xr = rb_sweeps(oE, oN, uC, D, fc, xz, 0 if adj else 1, 2)
        assert np.array_equal(xg, xr), (tile, adj, "zero start")


@pytest.mark.parametrize("st,n3", [(7, (21, 17, 19)), (27, (21, 17, 19)), (7, (161, 9, 7)),
                                  (7, (5, 6, 5)), (7, (70, 5, 4))])
def test_resid_bitwise(L, O, st, n3):
    """The residual (for st 7 the vector form: 2-point runs) against the
    oracle, bit for bit, on the interior."""
    rng = np.random.default_rng(11)
    nx, ny, nz = n3
    n = nx * ny * nz
    a = rand_op(rng, n3, st)
    cc, fc, x = rng.uniform(0, 1, n), rng.normal(size=n), rng.normal(size=n)
    rg = np.zeros(n)
    ro = np.zeros(n)
    L.pmgtest_resid(nx, ny, nz, st, P(a), P(cc), P(fc), P(x), P(rg))
    O.orc_resid(nx, ny, nz, st, P(a), P(cc), P(fc), P(x), P(ro))
    assert np.array_equal(interior(rg, n3), interior(ro, n3))


def test_transfers_bitwise(L, O):
    rng = np.random.default_rng(3)
    f3, c3 = (33, 17, 25), (17, 9, 13)
    nf, nc = np.prod(f3), np.prod(c3)
    xin = rng.normal(size=nf)
    og, oo = np.zeros(nc), np.zeros(nc)
    L.pmgtest_restrict(*f3, *c3, P(xin), P(og))
    O.orc_restrict(*f3, *c3, P(xin.copy()), P(oo))
    assert np.array_equal(interior(og, c3), interior(oo, c3))
    xc = rng.normal(size=nc)
    xc.reshape(c3[2], c3[1], c3[0])[[0, -1], :, :] = 0
    xc.reshape(c3[2], c3[1], c3[0])[:, [0, -1], :] = 0
    xc.reshape(c3[2], c3[1], c3[0])[:, :, [0, -1]] = 0
    fg, fo = np.zeros(nf), np.zeros(nf)
    L.pmgtest_prolong(*c3, *f3, P(xc), P(fg), ctypes.c_double(1.0))
    O.orc_interp(*c3, *f3, P(xc.copy()), P(fo))
    assert np.array_equal(interior(fg, f3), interior(fo, f3))


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("st,f3,c3", [(7, (17, 17, 17), (9, 9, 9)),
                                      (27, (17, 17, 17), (9, 9, 9)),
                                      (7, (65, 57, 61), (33, 29, 31)),
                                      (27, (65, 57, 61), (33, 29, 31))])
def test_galerkin_bitwise(L, O, st, f3, c3, form):
    """Every kernel form (one thread, one 32-lane group, one thread with the
    loops unrolled at compile time per coefficient) and the size-based
    choice; the 33x29x31 level has more coarse points than the group form's
    cut-off, so form 0 takes a per-thread kernel there."""
    rng = np.random.default_rng(5 + st)
    nf, nc = np.prod(f3), np.prod(c3)
    af = rand_op(rng, f3, st)
    ag, ao = np.zeros(14 * nc), np.zeros(14 * nc)
    assert L.pmgtest_galerkin_form(*f3, st, P(af), *c3, P(ag), form) == 1
    O.orc_galerkin(*f3, st, P(af), *c3, P(ao))
    for s in range(14):
        assert np.array_equal(interior(ag[s * nc:(s + 1) * nc], c3),
                              interior(ao[s * nc:(s + 1) * nc], c3)), s


def test_buildA_bitwise(L, O):
    rng = np.random.default_rng(9)
    n3 = (19, 17, 21)
    nx, ny, nz = n3
    n = nx * ny * nz
    xf = -3.7 + np.arange(nx) * 0.41
    yf = 1.3 + np.arange(ny) * 0.37
    zf = -0.2 + np.arange(nz) * 0.43
    gx, gy, gz = (rng.normal(size=2 * ny * nz), rng.normal(size=2 * nx * nz),
                  rng.normal(size=2 * nx * ny))
    a1, a2, a3 = (rng.uniform(2, 80, n) for _ in range(3))
    ccf, fcf = rng.uniform(0, 1, n), rng.normal(size=n)
    outs_g = [np.zeros(n) for _ in range(6)]
    outs_o = [np.zeros(n) for _ in range(6)]
    L.pmgtest_buildA(nx, ny, nz, P(xf), P(yf), P(zf), P(gx), P(gy), P(gz), P(a1), P(a2), P(a3),
                     P(ccf), P(fcf), *[P(o) for o in outs_g])
    O.orc_buildA_fv(nx, ny, nz, P(xf), P(yf), P(zf), P(gx), P(gy), P(gz), P(a1), P(a2), P(a3),
                    P(ccf), P(fcf), *[P(o) for o in outs_o])
    for g, o in zip(outs_g, outs_o):
        assert np.array_equal(interior(g, n3), interior(o, n3))


def test_nonlinear_terms(L, O):
    rng = np.random.default_rng(13)
    n3 = (17, 17, 17)
    nx, ny, nz = n3
    n = nx * ny * nz
    a = rand_op(rng, n3, 7)
    cc = rng.uniform(0, 1, n)
    cc[::7] = 0.0
    fc, x = rng.normal(size=n), rng.normal(scale=30, size=n)
    q = np.array([1.0, -1.0, 2.0, -2.0])
    sc = np.array([0.2, 0.2, 0.05, 0.05])
    rg, dg = np.zeros(n), np.zeros(n)
    L.pmgtest_nonlinear(nx, ny, nz, P(a), P(cc), P(fc), P(x), 4, P(q), P(sc), P(rg), P(dg))
    ro, w, do = np.zeros(n), np.zeros(n), np.zeros(n)
    zero = np.zeros(n)
    O.orc_resid(nx, ny, nz, 7, P(a), P(zero), P(fc), P(x), P(ro))
    O.orc_cvec(n, 4, P(q), P(sc), P(cc), P(x), P(w))
    O.orc_dcvec(n, 4, P(q), P(sc), P(cc), P(x), P(do))
    ro = ro - w
    ri, rgi = interior(ro, n3), interior(rg, n3)
    scale = np.abs(ri).max()
    assert np.max(np.abs(rgi - ri)) <= 8 * np.finfo(float).eps * scale
    assert np.allclose(dg, do, rtol=4 * np.finfo(float).eps, atol=0)


@pytest.mark.parametrize("N", [64, 27, 8, 125])
def test_dense_inverse_forms(L, N):
    """The coarsest level's dense inverse: the register form (N <= 64) and
    the LDS / global form agree to 1e-13 relative and invert an SPD matrix
    like the coarsest 27-point operator (diagonally dominant, symmetric)."""
    rng = np.random.default_rng(N)
    A = rng.uniform(-1, 0, (N, N))
    A = 0.5 * (A + A.T)
    np.fill_diagonal(A, 0.0)
    A[np.diag_indices(N)] = -A.sum(axis=1) + rng.uniform(0.5, 2.0, N)
    A = np.ascontiguousarray(A)
    outs = []
    for form in (0, 1):
        out = np.zeros((N, N))
        assert L.pmgtest_dense_invert(N, P(A), P(out), form) == 1
        assert np.abs(A @ out - np.eye(N)).max() < 1e-12
        outs.append(out)
    assert np.abs(outs[0] - outs[1]).max() <= 1e-13 * np.abs(outs[1]).max()

