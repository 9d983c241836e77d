"""The C-ABI library loads here (no GPU) and exports every symbol that
include/*.h declares; no compute calls."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "apbs_amd", "libapbs_pmg.so")


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name in ("if", "sizeof", "return", "defined"):
                continue
            # declarations: identifier followed by '(' at top level of a prototype
            names.add(name)
    # keep only names that look like API functions
    return sorted(n for n in names
                  if re.match(r"^(Vpmg|Vpmgp|Vpbe|Valist|Vgrid|MGparm|apbs_pmg|pmgtest|PmgServer)_", n))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    return ctypes.CDLL(LIB)


def test_library_loads_and_exports_all_declared(lib):
    decl = declared_symbols()
    assert len(decl) >= 60
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [d for d in decl if d not in exported]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    import apbs_amd
    decl = {d for d in declared_symbols() if not d.startswith("pmgtest_")}
    assert decl <= set(apbs_amd.EXPORTED), sorted(decl - set(apbs_amd.EXPORTED))


def test_host_only_entry_points(lib):
    """Entry points that touch no device memory can run without a GPU."""
    import apbs_amd
    L = apbs_amd.lib()
    assert L.apbs_pmg_version().startswith(b"apbs_pmg")
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.Vpmgp_makeCoarse(2, 161, 97, 65, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    assert (a.value, b.value, c.value) == (41, 25, 17)
    m = L.MGparm_ctor(apbs_amd.MCT_MANUAL)
    m.contents.dime[:] = [161, 161, 161]
    m.contents.setdime = 1
    m.contents.grid[:] = [0.4, 0.4, 0.4]
    m.contents.setgrid = 1
    m.contents.setgcent = 1
    m.contents.setchgm = 1
    assert L.MGparm_check(m) == 1
    # mgparm.c:265-320: 161 = 5*2^5+1 -> nlev 4 (not the 5 a user may ask for)
    assert m.contents.nlev == 4
    L.MGparm_dtor(ctypes.byref(m))


def test_mgauto_chain_matches_restatement():
    """MGparm_setupAuto (library) against the test restatement of
    NOsh_setupCalcMGAUTO for the born example (cglen 50, fglen 12, 65^3)."""
    import numpy as np
    from apbs_amd import driver
    import orc_driver as od
    chain = driver.auto_chain([65] * 3, [50] * 3, [12] * 3, [0, 0, 0], [0.5, -0.25, 1.0])
    ref = od.auto_chain([65] * 3, [50] * 3, [12] * 3, [0, 0, 0], [0.5, -0.25, 1.0])
    assert len(chain) == len(ref) == 3
    for m, (grid, glen, center) in zip(chain, ref):
        assert np.allclose(m.contents.grid[:], grid, rtol=0, atol=0)
        assert np.allclose(m.contents.glen[:], glen, rtol=0, atol=0)
        assert np.allclose(m.contents.center[:], center, rtol=0, atol=0)


def build_c_caller(tmpdir):
    """tests/c_caller.c built with gcc against include/apbs_pmg.h and the
    library (a C binding needs nothing else)."""
    exe = os.path.join(str(tmpdir), "c_caller")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", exe, os.path.join(ROOT, "tests", "c_caller.c"),
                    "-L", os.path.join(ROOT, "apbs_amd"), "-lapbs_pmg",
                    "-Wl,-rpath," + os.path.join(ROOT, "apbs_amd")], check=True)
    return exe


def test_c_header_layout_matches_binding(lib, tmp_path):
    """The struct layouts a C caller compiles against equal the ones the
    Python binding uses (same sizes and field offsets)."""
    import apbs_amd
    exe = build_c_caller(tmp_path)
    out = subprocess.run([exe, "layout"], capture_output=True, text=True, check=True).stdout
    vals = dict(l.split() for l in out.splitlines())
    types = {"Vatom": apbs_amd.Vatom, "Valist": apbs_amd.Valist, "MGparm": apbs_amd.MGparm,
             "Vpmgp": apbs_amd.Vpmgp, "Vpmg": apbs_amd.Vpmg}
    checked = 0
    for key, v in vals.items():
        kind, field = key.split(".")
        if kind == "sizeof" and field in types:
            assert ctypes.sizeof(types[field]) == int(v), key
            checked += 1
        elif kind in types:
            assert getattr(types[kind], field).offset == int(v), key
            checked += 1
    assert checked >= 25
    # reference layout facts: sVatom carries two 64-byte name fields after
    # id (vatom.h:84-100); sMGparm puts chgm right after setdime
    # (mgparm.h:112-120)
    assert int(vals["sizeof.Vatom"]) == 192 and int(vals["MGparm.chgm"]) == 24


def test_footprint_estimate_host_only(lib):
    """PmgServer_footprint needs no GPU: it grows with the grid, with Newton
    and with the FP32 copies, and 161^3 npbe mixed sits between the bytes of
    its arrays (~30 fine FP64-equivalent arrays) and twice that."""
    import numpy as np
    from apbs_amd import serve
    xyz = np.zeros((1, 3))
    one = (xyz, np.array([1.0]), np.array([2.0]))

    def fp(dime, pbe, prec):
        return serve.Request(*one, (dime,) * 3, [0.4] * 3, [0, 0, 0], pbe, precision=prec).footprint()

    big = fp(161, serve.NPBE, 1)
    assert fp(161, serve.NPBE, 0) < big
    assert fp(161, serve.LPBE, 0) < fp(161, serve.NPBE, 0)
    assert fp(97, serve.NPBE, 1) < big
    words = 161 ** 3 * 8
    assert 25 * words < big < 60 * words
