"""apbs_amd -- Python binding of libapbs_pmg.so, the MI355X PMG
Poisson-Boltzmann solver with the APBS Vpmg C API.

The binding mirrors the C ABI in include/apbs_pmg.h one to one (ctypes), and
``apbs_amd.driver`` restates the mg-manual / mg-auto calculation sequence of
APBS's src/routines.c (initMG -> solveMG -> setPartMG -> energyMG ->
forceMG) on top of it.

Loading fails loudly when the library has not been built: there is no CPU
fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libapbs_pmg.so")

# enums of include/apbs_pmg.h (values of src/generic/vhal.h)
VSOL_Newton, VSOL_MG = 1, 2
VSM_MOL, VSM_MOLSMOOTH, VSM_SPLINE, VSM_SPLINE3, VSM_SPLINE4 = 0, 1, 2, 3, 4
VCM_TRIL, VCM_BSPL2, VCM_BSPL4 = 0, 1, 2
BCFL_ZERO, BCFL_SDH, BCFL_MDH, BCFL_FOCUS, BCFL_MAP = 0, 1, 2, 4, 6
NONLIN_LPBE, NONLIN_NPBE = 0, 1
MCT_MANUAL, MCT_AUTO, MCT_PARALLEL, MCT_DUMMY = 0, 1, 2, 3
PCE_NO, PCE_TOTAL = 0, 1
VDT_CHARGE, VDT_POT, VDT_DIELX, VDT_DIELY, VDT_DIELZ, VDT_KAPPA = 0, 1, 11, 12, 13, 14

_dp = ctypes.POINTER(ctypes.c_double)


class Vatom(ctypes.Structure):
    """struct sVatom (src/generic/vatom.h:84-100)"""
    _fields_ = [("position", ctypes.c_double * 3), ("radius", ctypes.c_double),
                ("charge", ctypes.c_double), ("partID", ctypes.c_double),
                ("epsilon", ctypes.c_double), ("id", ctypes.c_int),
                ("resName", ctypes.c_char * 64), ("atomName", ctypes.c_char * 64)]


class Valist(ctypes.Structure):
    """struct sValist (src/generic/valist.h:78-89)"""
    _fields_ = [("number", ctypes.c_int), ("center", ctypes.c_double * 3),
                ("mincrd", ctypes.c_double * 3), ("maxcrd", ctypes.c_double * 3),
                ("maxrad", ctypes.c_double), ("charge", ctypes.c_double),
                ("atoms", ctypes.POINTER(Vatom)), ("vmem", ctypes.c_void_p)]


class MGparm(ctypes.Structure):
    """struct sMGparm (src/generic/mgparm.h:112-183)"""
    _fields_ = [
        ("type", ctypes.c_int), ("parsed", ctypes.c_int),
        ("dime", ctypes.c_int * 3), ("setdime", ctypes.c_int),
        ("chgm", ctypes.c_int), ("setchgm", ctypes.c_int),
        ("chgs", ctypes.c_int),
        ("nlev", ctypes.c_int), ("setnlev", ctypes.c_int),
        ("etol", ctypes.c_double), ("setetol", ctypes.c_int),
        ("grid", ctypes.c_double * 3), ("setgrid", ctypes.c_int),
        ("glen", ctypes.c_double * 3), ("setglen", ctypes.c_int),
        ("cmeth", ctypes.c_int), ("center", ctypes.c_double * 3), ("centmol", ctypes.c_int),
        ("setgcent", ctypes.c_int),
        ("cglen", ctypes.c_double * 3), ("setcglen", ctypes.c_int),
        ("fglen", ctypes.c_double * 3), ("setfglen", ctypes.c_int),
        ("ccmeth", ctypes.c_int), ("ccenter", ctypes.c_double * 3), ("ccentmol", ctypes.c_int),
        ("setcgcent", ctypes.c_int),
        ("fcmeth", ctypes.c_int), ("fcenter", ctypes.c_double * 3), ("fcentmol", ctypes.c_int),
        ("setfgcent", ctypes.c_int),
        ("partDisjCenter", ctypes.c_double * 3), ("partDisjLength", ctypes.c_double * 3),
        ("partDisjOwnSide", ctypes.c_int * 6),
        ("pdime", ctypes.c_int * 3), ("setpdime", ctypes.c_int),
        ("proc_rank", ctypes.c_int), ("setrank", ctypes.c_int),
        ("proc_size", ctypes.c_int), ("setsize", ctypes.c_int),
        ("ofrac", ctypes.c_double), ("setofrac", ctypes.c_int),
        ("async", ctypes.c_int), ("setasync", ctypes.c_int),
        ("nonlintype", ctypes.c_int), ("setnonlintype", ctypes.c_int),
        ("method", ctypes.c_int), ("setmethod", ctypes.c_int),
        ("useAqua", ctypes.c_int), ("setUseAqua", ctypes.c_int),
    ]


class Vpmgp(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("nlev", ctypes.c_int),
        ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("hzed", ctypes.c_double),
        ("nonlin", ctypes.c_int),
        ("nxc", ctypes.c_int), ("nyc", ctypes.c_int), ("nzc", ctypes.c_int), ("nf", ctypes.c_int),
        ("nc", ctypes.c_int), ("narrc", ctypes.c_int), ("n_rpc", ctypes.c_int),
        ("n_iz", ctypes.c_int), ("n_ipc", ctypes.c_int),
        ("nrwk", ctypes.c_size_t), ("niwk", ctypes.c_int), ("narr", ctypes.c_int),
        ("ipkey", ctypes.c_int),
        ("xcent", ctypes.c_double), ("ycent", ctypes.c_double), ("zcent", ctypes.c_double),
        ("errtol", ctypes.c_double),
        ("itmax", ctypes.c_int), ("istop", ctypes.c_int), ("iinfo", ctypes.c_int),
        ("bcfl", ctypes.c_int),
        ("key", ctypes.c_int), ("iperf", ctypes.c_int), ("meth", ctypes.c_int),
        ("mgkey", ctypes.c_int), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
        ("mgsmoo", ctypes.c_int), ("mgprol", ctypes.c_int), ("mgcoar", ctypes.c_int),
        ("mgsolv", ctypes.c_int), ("mgdisc", ctypes.c_int),
        ("omegal", ctypes.c_double), ("omegan", ctypes.c_double),
        ("irite", ctypes.c_int), ("ipcon", ctypes.c_int),
        ("xlen", ctypes.c_double), ("ylen", ctypes.c_double), ("zlen", ctypes.c_double),
        ("xmin", ctypes.c_double), ("ymin", ctypes.c_double), ("zmin", ctypes.c_double),
        ("xmax", ctypes.c_double), ("ymax", ctypes.c_double), ("zmax", ctypes.c_double),
    ]


class Vgrid(ctypes.Structure):
    """struct sVgrid (src/mg/vgrid.h:85-104); data x fastest."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("hzed", ctypes.c_double),
        ("xmin", ctypes.c_double), ("ymin", ctypes.c_double), ("zmin", ctypes.c_double),
        ("xmax", ctypes.c_double), ("ymax", ctypes.c_double), ("zmax", ctypes.c_double),
        ("data", _dp), ("readdata", ctypes.c_int), ("ctordata", ctypes.c_int),
        ("mem", ctypes.c_void_p),
    ]


class Vpmg(ctypes.Structure):
    _fields_ = [
        ("vmem", ctypes.c_void_p), ("pmgp", ctypes.POINTER(Vpmgp)), ("pbe", ctypes.c_void_p),
        ("epsx", _dp), ("epsy", _dp), ("epsz", _dp), ("kappa", _dp), ("pot", _dp),
        ("charge", _dp), ("iparm", ctypes.c_void_p), ("rparm", _dp), ("iwork", ctypes.c_void_p),
        ("rwork", _dp), ("a1cf", _dp), ("a2cf", _dp), ("a3cf", _dp), ("ccf", _dp), ("fcf", _dp),
        ("tcf", _dp), ("u", _dp), ("xf", _dp), ("yf", _dp), ("zf", _dp), ("gxcf", _dp),
        ("gycf", _dp), ("gzcf", _dp), ("pvec", _dp),
        ("extDiEnergy", ctypes.c_double), ("extQmEnergy", ctypes.c_double),
        ("extQfEnergy", ctypes.c_double), ("extNpEnergy", ctypes.c_double),
        ("surfMeth", ctypes.c_int), ("splineWin", ctypes.c_double),
        ("chargeMeth", ctypes.c_int), ("chargeSrc", ctypes.c_int), ("filled", ctypes.c_int),
        ("useDielXMap", ctypes.c_int), ("dielXMap", ctypes.POINTER(Vgrid)),
        ("useDielYMap", ctypes.c_int), ("dielYMap", ctypes.POINTER(Vgrid)),
        ("useDielZMap", ctypes.c_int), ("dielZMap", ctypes.POINTER(Vgrid)),
        ("useKappaMap", ctypes.c_int), ("kappaMap", ctypes.POINTER(Vgrid)),
        ("usePotMap", ctypes.c_int), ("potMap", ctypes.POINTER(Vgrid)),
        ("useChargeMap", ctypes.c_int), ("chargeMap", ctypes.POINTER(Vgrid)),
        ("dev", ctypes.c_void_p),
    ]


MAXION = 10


class PmgRequest(ctypes.Structure):
    """include/apbs_pmg_server.h PmgRequest"""
    _fields_ = [
        ("natoms", ctypes.c_int), ("xyz", _dp), ("radius", _dp), ("charge", _dp),
        ("dime", ctypes.c_int * 3), ("grid", ctypes.c_double * 3),
        ("center", ctypes.c_double * 3), ("npbe", ctypes.c_int), ("nion", ctypes.c_int),
        ("ionq", ctypes.c_double * MAXION), ("ionc", ctypes.c_double * MAXION),
        ("ionr", ctypes.c_double * MAXION),
        ("pdie", ctypes.c_double), ("sdie", ctypes.c_double), ("srad", ctypes.c_double),
        ("swin", ctypes.c_double), ("sdens", ctypes.c_double), ("temp", ctypes.c_double),
        ("srfm", ctypes.c_int), ("chgm", ctypes.c_int), ("bcfl", ctypes.c_int),
        ("etol", ctypes.c_double), ("precision", ctypes.c_int), ("mixed_tol", ctypes.c_double),
        ("potential", _dp), ("tag", ctypes.c_long),
    ]


class PmgResult(ctypes.Structure):
    """include/apbs_pmg_server.h PmgResult"""
    _fields_ = [
        ("id", ctypes.c_long), ("tag", ctypes.c_long), ("status", ctypes.c_int),
        ("energy_kT", ctypes.c_double), ("iters", ctypes.c_int), ("inner", ctypes.c_int),
        ("relres", ctypes.c_double), ("t_submit", ctypes.c_double),
        ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
        ("footprint", ctypes.c_size_t),
    ]


# (name, restype, argtypes) of every exported entry point
_SIGS = [
    ("Valist_ctor", ctypes.POINTER(Valist), []),
    ("Valist_dtor", None, [ctypes.POINTER(ctypes.POINTER(Valist))]),
    ("Valist_readPQRFile", ctypes.c_int, [ctypes.POINTER(Valist), ctypes.c_char_p]),
    ("Valist_setAtoms", ctypes.c_int, [ctypes.POINTER(Valist), ctypes.c_int, _dp, _dp, _dp]),
    ("Valist_getStatistics", ctypes.c_int, [ctypes.POINTER(Valist)]),
    ("Valist_getNumberAtoms", ctypes.c_int, [ctypes.POINTER(Valist)]),
    ("Valist_getAtom", ctypes.POINTER(Vatom), [ctypes.POINTER(Valist), ctypes.c_int]),
    ("Vpbe_ctor", ctypes.c_void_p,
     [ctypes.POINTER(Valist), ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("Vpbe_dtor", None, [ctypes.POINTER(ctypes.c_void_p)]),
    ("Vpbe_getBulkIonicStrength", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getMaxIonRadius", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getTemperature", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getSoluteDiel", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getSolventDiel", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getSolventRadius", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getXkappa", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getZkappa2", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getZmagic", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getSoluteRadius", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getSoluteCharge", ctypes.c_double, [ctypes.c_void_p]),
    ("Vpbe_getSoluteCenter", _dp, [ctypes.c_void_p]),
    ("Vpbe_getValist", ctypes.POINTER(Valist), [ctypes.c_void_p]),
    ("Vpbe_getIons", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp]),
    ("MGparm_ctor", ctypes.POINTER(MGparm), [ctypes.c_int]),
    ("MGparm_dtor", None, [ctypes.POINTER(ctypes.POINTER(MGparm))]),
    ("MGparm_check", ctypes.c_int, [ctypes.POINTER(MGparm)]),
    ("MGparm_setupAuto", ctypes.c_int,
     [ctypes.POINTER(MGparm), ctypes.POINTER(ctypes.POINTER(MGparm)), ctypes.c_int]),
    ("MGparm_setupPara", ctypes.c_int,
     [ctypes.POINTER(MGparm), ctypes.c_int, ctypes.POINTER(ctypes.POINTER(MGparm)), ctypes.c_int]),
    ("Vgrid_ctor", ctypes.POINTER(Vgrid),
     [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [_dp]),
    ("Vgrid_dtor", None, [ctypes.POINTER(ctypes.POINTER(Vgrid))]),
    ("Vgrid_value", ctypes.c_int, [ctypes.POINTER(Vgrid), _dp, _dp]),
    ("Vpmgp_ctor", ctypes.POINTER(Vpmgp), [ctypes.POINTER(MGparm)]),
    ("Vpmgp_dtor", None, [ctypes.POINTER(ctypes.POINTER(Vpmgp))]),
    ("Vpmgp_size", None, [ctypes.POINTER(Vpmgp)]),
    ("Vpmgp_makeCoarse", None,
     [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    ("Vpmg_ctor", ctypes.POINTER(Vpmg),
     [ctypes.POINTER(Vpmgp), ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Vpmg),
      ctypes.POINTER(MGparm), ctypes.c_int]),
    ("Vpmg_ctor2", ctypes.c_int,
     [ctypes.POINTER(Vpmg), ctypes.POINTER(Vpmgp), ctypes.c_void_p, ctypes.c_int,
      ctypes.POINTER(Vpmg), ctypes.POINTER(MGparm), ctypes.c_int]),
    ("Vpmg_dtor", None, [ctypes.POINTER(ctypes.POINTER(Vpmg))]),
    ("Vpmg_dtor2", None, [ctypes.POINTER(Vpmg)]),
    ("Vpmg_fillco", ctypes.c_int,
     [ctypes.POINTER(Vpmg), ctypes.c_int, ctypes.c_double, ctypes.c_int] +
     [ctypes.c_int, ctypes.POINTER(Vgrid)] * 6),
    ("Vpmg_solve", ctypes.c_int, [ctypes.POINTER(Vpmg)]),
    ("Vpmg_energy", ctypes.c_double, [ctypes.POINTER(Vpmg), ctypes.c_int]),
    ("Vpmg_qfEnergy", ctypes.c_double, [ctypes.POINTER(Vpmg), ctypes.c_int]),
    ("Vpmg_qmEnergy", ctypes.c_double, [ctypes.POINTER(Vpmg), ctypes.c_int]),
    ("Vpmg_dielEnergy", ctypes.c_double, [ctypes.POINTER(Vpmg), ctypes.c_int]),
    ("Vpmg_qfAtomEnergy", ctypes.c_double, [ctypes.POINTER(Vpmg), ctypes.POINTER(Vatom)]),
    ("Vpmg_force", ctypes.c_int, [ctypes.POINTER(Vpmg), _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("Vpmg_qfForce", ctypes.c_int, [ctypes.POINTER(Vpmg), _dp, ctypes.c_int, ctypes.c_int]),
    ("Vpmg_ibForce", ctypes.c_int, [ctypes.POINTER(Vpmg), _dp, ctypes.c_int, ctypes.c_int]),
    ("Vpmg_dbForce", ctypes.c_int, [ctypes.POINTER(Vpmg), _dp, ctypes.c_int, ctypes.c_int]),
    ("Vpmg_setPart", None, [ctypes.POINTER(Vpmg), _dp, _dp, ctypes.POINTER(ctypes.c_int)]),
    ("Vpmg_unsetPart", None, [ctypes.POINTER(Vpmg)]),
    ("Vpmg_fillArray", ctypes.c_int,
     [ctypes.POINTER(Vpmg), _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]),
    ("Vpmg_memChk", ctypes.c_ulong, [ctypes.POINTER(Vpmg)]),
    ("Vpmg_setPrecision", ctypes.c_int, [ctypes.POINTER(Vpmg), ctypes.c_int, ctypes.c_double]),
    ("Vpmg_setHostMirror", ctypes.c_int, [ctypes.POINTER(Vpmg), ctypes.c_int]),
    ("Vpmg_getStats", ctypes.c_int, [ctypes.POINTER(Vpmg), _dp, ctypes.c_int]),
    ("Vpmg_copyArray", ctypes.c_int, [ctypes.POINTER(Vpmg), ctypes.c_int, _dp, ctypes.c_size_t]),
    ("Vpmg_setKernelTiming", ctypes.c_int, [ctypes.POINTER(Vpmg), ctypes.c_int]),
    ("Vpmg_getKernelTiming", ctypes.c_int, [ctypes.POINTER(Vpmg), _dp, ctypes.c_int]),
    ("Vpmg_energyNoAlloc", ctypes.c_int, [ctypes.POINTER(Vpmg), ctypes.c_int, _dp]),
    ("Vpmg_solveBatch", ctypes.c_int,
     [ctypes.POINTER(ctypes.POINTER(Vpmg)), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    ("apbs_pmg_synchronize", ctypes.c_int, []),
    ("PmgServer_ctor", ctypes.c_void_p, [ctypes.c_int, ctypes.c_size_t]),
    ("PmgServer_submit", ctypes.c_long, [ctypes.c_void_p, ctypes.POINTER(PmgRequest)]),
    ("PmgServer_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PmgResult), ctypes.c_int]),
    ("PmgServer_drain", None, [ctypes.c_void_p]),
    ("PmgServer_stats", None, [ctypes.c_void_p, _dp]),
    ("PmgServer_dtor", None, [ctypes.POINTER(ctypes.c_void_p)]),
    ("PmgServer_footprint", ctypes.c_size_t, [ctypes.POINTER(PmgRequest)]),
    ("PmgServer_setBatching", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("PmgServer_batchStats", None,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
    ("apbs_pmg_now", ctypes.c_double, []),
    ("apbs_pmg_poolStats", None, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                  ctypes.POINTER(ctypes.c_size_t), ctypes.c_int]),
    ("apbs_pmg_version", ctypes.c_char_p, []),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


def lib():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} not found: build it with `python -m apbs_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
