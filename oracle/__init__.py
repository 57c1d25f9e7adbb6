"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker.  The product
path (baum_amd) never imports this package.

Arrays follow BAM's SoA layout: a state is a numpy array of shape
(nvars, o, n, m) in C order, i.e. one contiguous double[m*n*o] per grid
function with ijk = i + m*j + m*n*k.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NV = 25
VAR_NAMES = ["gxx", "gxy", "gxz", "gyy", "gyz", "gzz", "chi",
             "Axx", "Axy", "Axz", "Ayy", "Ayz", "Azz", "Khat",
             "Gx", "Gy", "Gz", "Theta", "alpha",
             "betax", "betay", "betaz", "Bx", "By", "Bz"]

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class Box(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("o", C.c_int),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("skip", C.c_int * 6)]


class Z4cParams(C.Structure):
    _fields_ = [("order_centered", C.c_int), ("order_advection", C.c_int),
                ("advection_lopsided", C.c_int), ("order_dissipation", C.c_int),
                ("dissfactor", C.c_double), ("kappa1", C.c_double), ("kappa2", C.c_double),
                ("zeroThetaRHS", C.c_double), ("subtractA", C.c_int), ("normalizedetg", C.c_int),
                ("oploglapse", C.c_int), ("oploglapse2", C.c_int), ("oplogwithshift", C.c_int),
                ("harmoniclapse", C.c_int), ("lapseharmonicf", C.c_double),
                ("gamma0factor", C.c_double), ("gamma2factor", C.c_double),
                ("withGadv", C.c_double), ("withShiftadv", C.c_double),
                ("withBadv", C.c_double), ("withB", C.c_double),
                ("shiftalphapower", C.c_double), ("shiftgammacoeff", C.c_double),
                ("shiftdriver", C.c_double), ("chiDivFloor", C.c_double),
                ("chipsipower", C.c_double)]


class C2PParams(C.Structure):
    _fields_ = [("eos", C.c_int), ("gamma", C.c_double), ("K", C.c_double),
                ("rhoatm", C.c_double), ("fatm", C.c_double), ("epslatm", C.c_double),
                ("patm", C.c_double), ("vmax", C.c_double), ("errormax", C.c_double),
                ("countmax", C.c_int)]


class HrscParams(C.Structure):
    _fields_ = [("eos", C.c_int), ("gamma", C.c_double), ("K", C.c_double),
                ("vmax", C.c_double), ("wlormax", C.c_double), ("riemann", C.c_int), ("rec", C.c_int)]


def hrsc_params(eos=0, gamma=2.0, K=100.0, vmax=0.999, wlormax=1e5, riemann="LLF", rec="WENOZ"):
    """grhd_vmax / grhd_Wlor_max / hrsc_flux / hrsc_rec of the reference's BNS parameter file"""
    return HrscParams(eos, gamma, K, vmax, wlormax, {"LLF": 0, "HLL": 1}[riemann], {"WENOZ": 0, "WENO5": 1}[rec])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.ob_z4c_rhs.restype = C.c_int
        L.ob_z4c_rhs.argtypes = [C.POINTER(Box), C.POINTER(Z4cParams), C.POINTER(DP),
                                 C.POINTER(DP), C.c_double, C.POINTER(DP)]
        L.ob_z4c_radiative_bc.argtypes = [C.POINTER(Box), C.POINTER(Z4cParams), C.POINTER(DP),
                                          C.POINTER(DP), C.c_double, C.POINTER(DP)]
        L.ob_periodic_fill.argtypes = [C.POINTER(Box), C.c_int * 3, C.c_int, DP]
        L.ob_z4c_algcon.argtypes = [C.POINTER(Box), C.c_int, C.c_int, C.POINTER(DP), C.POINTER(DP)]
        L.ob_adm_to_z4c.argtypes = [C.POINTER(Box), C.c_double, C.POINTER(DP), C.POINTER(DP), C.POINTER(DP), C.c_int]
        L.ob_punctures_bl.argtypes = [C.POINTER(Box), C.c_int, DP, DP, C.POINTER(DP), C.POINTER(DP), DP]
        L.ob_z4c_to_adm.argtypes = [C.POINTER(Box), C.c_double, C.POINTER(DP), C.POINTER(DP), C.POINTER(DP)]
        L.ob_z4c_rk4_step.restype = C.c_int
        L.ob_z4c_rk4_step.argtypes = [C.POINTER(Box), C.POINTER(Z4cParams), C.c_double, C.c_int,
                                      C.c_int, C.POINTER(DP), C.POINTER(DP), C.c_int]
        L.ob_hrsc_rec.restype = C.c_double
        L.ob_hrsc_rec.argtypes = [C.c_int, C.c_int, DP, C.c_int]
        L.ob_grhd_interface_flux.restype = None
        L.ob_grhd_interface_flux.argtypes = [C.POINTER(HrscParams), DP, DP, DP, DP]
        L.ob_grhd_fluxes.restype = None
        L.ob_grhd_fluxes.argtypes = [C.POINTER(Box), C.POINTER(HrscParams), C.c_double, C.POINTER(DP),
                                     C.POINTER(DP), C.POINTER(DP)]
        L.ob_grhd_sources.restype = None
        L.ob_grhd_sources.argtypes = [C.POINTER(Box), C.POINTER(HrscParams), C.c_double, C.c_int, C.c_double,
                                      C.POINTER(DP), C.POINTER(DP), C.POINTER(DP), C.POINTER(DP), C.POINTER(DP)]
        L.ob_z4c_matter_source.restype = None
        L.ob_z4c_matter_source.argtypes = [C.POINTER(Box), C.POINTER(Z4cParams), C.c_double, C.POINTER(DP),
                                           C.POINTER(DP), C.POINTER(DP)]
        L.ob_reduce_0d.restype = None
        L.ob_reduce_0d.argtypes = [C.c_int, C.POINTER(Box), C.POINTER(DP), C.c_int, C.c_int, DP]
        L.ob_line_1d.restype = C.c_int
        L.ob_line_1d.argtypes = [C.c_int, C.POINTER(Box), C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                 DP, IP, C.POINTER(C.c_long)]
        L.ob_interp_point.restype = C.c_double
        L.ob_interp_point.argtypes = [C.c_int, C.POINTER(Box), C.POINTER(DP), C.c_double, C.c_double, C.c_double,
                                      C.c_int, C.POINTER(C.c_int)]
        L.ob_track_puncture.restype = None
        L.ob_track_puncture.argtypes = [C.c_int, C.POINTER(Box), C.POINTER(DP), C.POINTER(DP), C.c_double, C.c_int,
                                        C.c_int, DP, DP]
        L.ob_z4c_ham_l2.restype = C.c_double
        L.ob_z4c_ham_l2.argtypes = [C.POINTER(Box), C.POINTER(DP), C.c_double, DP]
        L.ob_z4c_constraints_l2.restype = None
        L.ob_z4c_constraints_l2.argtypes = [C.POINTER(Box), C.POINTER(DP), C.c_double, DP, C.POINTER(DP), DP]
        L.ob_lagrange_coeffs.argtypes = [C.c_int, C.c_double, DP]
        L.ob_prolong_1d.argtypes = [C.c_int, C.c_int, DP, DP]
        L.ob_prolong_box.argtypes = [C.c_int, IP, IP, IP, DP, DP, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ob_restrict_box.argtypes = [C.c_int, IP, IP, DP, DP, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ob_rp_flags.argtypes = [C.c_int, C.POINTER(Box), C.c_int, C.POINTER(Box), IP, C.c_int, DP, DP]
        L.ob_prolong_region.restype = C.c_int
        L.ob_prolong_region.argtypes = [C.POINTER(Box), C.POINTER(Box), DP, C.c_int, C.c_int, C.c_int,
                                        IP, IP, IP]
        L.ob_restrict_region.restype = C.c_int
        L.ob_restrict_region.argtypes = [C.POINTER(Box), C.POINTER(Box), DP, C.c_int, C.c_int, IP, IP]
        L.ob_prolong_level.restype = C.c_int
        L.ob_prolong_level.argtypes = [C.c_int, C.POINTER(Box), C.c_int, C.POINTER(Box), IP, DP, C.c_int,
                                       C.c_int, C.c_int, C.c_int, DP, DP]
        L.ob_restrict_level.restype = C.c_int
        L.ob_restrict_level.argtypes = [C.c_int, C.POINTER(Box), C.c_int, C.POINTER(Box), IP, DP, C.c_int,
                                        C.c_int, C.c_int, DP, DP]
        L.ob_c2p_hybrid.argtypes = [C.POINTER(C2PParams), C.c_long, DP, DP, DP, IP, IP]
        L.ob_c2p_proot.argtypes = [C.POINTER(C2PParams), C.c_long, DP, DP, DP, IP, IP]
        L.ob_c2p_bracketed.argtypes = [C.POINTER(C2PParams), C.c_long, DP, DP, DP, IP, IP]
        L.ob_c2p_vanal.argtypes = [C.POINTER(C2PParams), C.c_long, DP, DP, DP, IP, DP]
        L.ob_p2c.argtypes = [C.POINTER(C2PParams), C.c_long, DP, DP, DP]
        L.ob_grhd_atm_set_u.restype = C.c_long
        L.ob_grhd_atm_set_u.argtypes = [C.POINTER(C2PParams), C.c_long, DP, DP, DP]
        L.ob_try_hot.restype = C.c_int
        L.ob_try_hot.argtypes = [C.POINTER(C2PParams), C.c_double, C.c_double, C.c_double,
                                 DP, DP, DP, C.c_double, IP]
        _LIB = L
    return _LIB


# --------------------------------------------------------------------------
# parameters: BAM parameter-file semantics (whole-word Getv) -> struct
def _getv(value, word):
    return 1 if word in str(value).split() else 0


Z4C_DEFAULTS = {
    # src/main/main/bam_main.c:54-73, src/projects/z4/bam_z4.c:30-70,
    # src/main/amr/bam_amr.c:50 (advection_lopsided)
    "order_centered": 4, "order_advection": 4, "advection_lopsided": 1,
    "order_dissipation": 4, "dissipation_factor": 0.1,
    "z4_kappa1": 0.0, "z4_kappa2": 0.0, "z4_zeroThetaRHS": "no",
    "z4_subtractA": "yes", "z4_normalizedetg": "yes",
    "z4_lapse": "1+log withshift", "z4_lapseharmonicf": 2.0,
    "z4_shift": "withShiftadv", "z4_shiftalphapower": 0.0, "z4_shiftgammacoeff": 1.0,
    "z4_shiftdriver": 2.0, "z4_chi_div_floor": -1000.0, "z4_chi_psipower": -4.0,
}


def z4c_params(**over):
    d = dict(Z4C_DEFAULTS)
    d.update(over)
    p = Z4cParams()
    p.order_centered = int(d["order_centered"])
    p.order_advection = int(d["order_advection"])
    p.advection_lopsided = int(d["advection_lopsided"])
    p.order_dissipation = int(d["order_dissipation"])
    p.dissfactor = float(d["dissipation_factor"]) / 2.0 ** p.order_dissipation
    p.kappa1 = float(d["z4_kappa1"])
    p.kappa2 = float(d["z4_kappa2"])
    p.zeroThetaRHS = float(_getv(d["z4_zeroThetaRHS"], "no"))
    p.subtractA = _getv(d["z4_subtractA"], "yes")
    p.normalizedetg = _getv(d["z4_normalizedetg"], "yes")
    p.oploglapse = _getv(d["z4_lapse"], "1+log")
    p.oploglapse2 = _getv(d["z4_lapse"], "1+log2")
    p.oplogwithshift = _getv(d["z4_lapse"], "withshift")
    p.harmoniclapse = _getv(d["z4_lapse"], "harmonic")
    p.lapseharmonicf = float(d["z4_lapseharmonicf"])
    p.gamma0factor = float(_getv(d["z4_shift"], "gamma0"))
    p.gamma2factor = float(_getv(d["z4_shift"], "gamma2"))
    p.withGadv = float(_getv(d["z4_shift"], "withGadv"))
    p.withShiftadv = float(_getv(d["z4_shift"], "withShiftadv"))
    p.withBadv = float(_getv(d["z4_shift"], "withBadv"))
    p.withB = float(_getv(d["z4_shift"], "withB"))
    p.shiftalphapower = float(d["z4_shiftalphapower"])
    p.shiftgammacoeff = float(d["z4_shiftgammacoeff"])
    p.shiftdriver = float(d["z4_shiftdriver"])
    p.chiDivFloor = float(d["z4_chi_div_floor"])
    p.chipsipower = float(d["z4_chi_psipower"])
    return p


def make_box(m, n, o, x0, y0, z0, dx, dy=None, dz=None, skip=(1, 1, 1, 1, 1, 1)):
    b = Box()
    b.m, b.n, b.o = m, n, o
    b.x0, b.y0, b.z0 = x0, y0, z0
    b.dx = dx
    b.dy = dx if dy is None else dy
    b.dz = dx if dz is None else dz
    for f in range(6):
        b.skip[f] = skip[f]
    return b


def _ptrs(arr):
    """array of double* to the leading-axis slices of a C-contiguous array"""
    assert arr.dtype == np.float64 and arr.flags.c_contiguous
    nv = arr.shape[0]
    stride = arr[0].size * 8
    base = arr.ctypes.data
    return (DP * nv)(*[C.cast(base + i * stride, DP) for i in range(nv)])


def z4c_rhs(box, params, u, unew=None, upre=None, c=0.0):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u) if unew is None else unew
    pre = _ptrs(upre) if upre is not None else _ptrs(out)
    nbad = lib().ob_z4c_rhs(C.byref(box), C.byref(params), _ptrs(out), pre, c, _ptrs(u))
    return out, nbad


def z4c_radiative_bc(box, params, rhs, u):
    lib().ob_z4c_radiative_bc(C.byref(box), C.byref(params), _ptrs(rhs), _ptrs(rhs), 0.0, _ptrs(u))


def periodic_fill(box, u, g, per=(1, 1, 1)):
    arr = (C.c_int * 3)(*per)
    for v in range(u.shape[0]):
        lib().ob_periodic_fill(C.byref(box), arr, g, u[v].ctypes.data_as(DP))


def z4c_algcon(box, u, subtractA=1, normalizedetg=1):
    lib().ob_z4c_algcon(C.byref(box), subtractA, normalizedetg, _ptrs(u[0:6]), _ptrs(u[7:13]))


def adm_to_z4c(box, gadm, kadm, chipsipower=-4.0, phybound=False):
    """z4_init; phybound: extrapolate G on the outermost layer (z4_setbound_G)"""
    u = np.zeros((NV,) + gadm.shape[1:])
    lib().ob_adm_to_z4c(C.byref(box), chipsipower, _ptrs(np.ascontiguousarray(gadm)),
                        _ptrs(np.ascontiguousarray(kadm)), _ptrs(u), 1 if phybound else 0)
    return u


def punctures_bl(box, masses, positions):
    """Brill-Lindquist ADM data: (g_ij (6, o, n, m), K_ij, psi)"""
    shape = (box.o, box.n, box.m)
    g, k, psi = np.zeros((6,) + shape), np.zeros((6,) + shape), np.zeros(shape)
    m = np.ascontiguousarray(masses, dtype=np.float64)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    lib().ob_punctures_bl(C.byref(box), len(m), m.ctypes.data_as(DP), pos.ctypes.data_as(DP), _ptrs(g), _ptrs(k),
                          psi.ctypes.data_as(DP))
    return g, k, psi


def z4c_rk4_step(box, params, dt, u, bc=1, pgh=2, algcon=True, work=None):
    if work is None:
        work = np.zeros((3 * NV,) + u.shape[1:])
    nbad = lib().ob_z4c_rk4_step(C.byref(box), C.byref(params), dt, bc, pgh, _ptrs(u), _ptrs(work),
                                 1 if algcon else 0)
    return nbad


def z4c_ham_l2(box, u, chipsipower=-4.0, want_field=False):
    ham = np.zeros(u.shape[1:]) if want_field else None
    l2 = lib().ob_z4c_ham_l2(C.byref(box), _ptrs(u), chipsipower,
                             ham.ctypes.data_as(DP) if want_field else None)
    return (l2, ham) if want_field else l2


def z4c_constraints_l2(box, u, chipsipower=-4.0, want_fields=False):
    """L2 norms (ham, mom_x, mom_y, mom_z) over the inner points; with
    want_fields also the fields ham (shape of a variable) and mom (3, ...)"""
    l2 = np.zeros(4)
    ham = np.zeros(u.shape[1:]) if want_fields else None
    mom = np.zeros((3,) + u.shape[1:]) if want_fields else None
    lib().ob_z4c_constraints_l2(C.byref(box), _ptrs(u), chipsipower,
                                ham.ctypes.data_as(DP) if want_fields else None,
                                _ptrs(mom) if want_fields else None, l2.ctypes.data_as(DP))
    return (l2, ham, mom) if want_fields else l2


def lagrange_coeffs(n, x):
    c = np.zeros(n)
    lib().ob_lagrange_coeffs(n, x, c.ctypes.data_as(DP))
    return c


def prolong_1d(order, uc):
    uc = np.ascontiguousarray(uc, dtype=np.float64)
    uf = np.zeros(2 * uc.size)
    lib().ob_prolong_1d(order, uc.size, uc.ctypes.data_as(DP), uf.ctypes.data_as(DP))
    return uf


# ---------------------------------------------------------------- FMR coupling
def _ivec(vals):
    return (C.c_int * len(vals))(*vals)


def _boxes(boxes):
    return (Box * len(boxes))(*boxes)


def level_points(boxes):
    return sum(b.m * b.n * b.o for b in boxes)


def prolong_box(order, cbox, fbox, fbox0, cv, fv):
    """prolong_boxtobox_N on single boxes: cv (o,n,m) coarse, fv (o,n,m) fine
    (updated in place); cbox/fbox/fbox0 are 6-int index boxes"""
    assert cv.flags.c_contiguous and fv.flags.c_contiguous
    lib().ob_prolong_box(order, _ivec(cbox), _ivec(fbox), _ivec(fbox0), cv.ctypes.data_as(DP),
                         fv.ctypes.data_as(DP), cv.shape[2], cv.shape[2] * cv.shape[1], fv.shape[2],
                         fv.shape[2] * fv.shape[1])


def restrict_box(order, fbox, cbox, fv, cv):
    assert cv.flags.c_contiguous and fv.flags.c_contiguous
    lib().ob_restrict_box(order, _ivec(fbox), _ivec(cbox), fv.ctypes.data_as(DP), cv.ctypes.data_as(DP),
                          fv.shape[2], fv.shape[2] * fv.shape[1], cv.shape[2], cv.shape[2] * cv.shape[1])


def rp_flags(fboxes, cboxes, parent, nbuffer):
    """setRPflags: returns (flagprolong over the fine level, flagrestrict over
    the coarse level), each a flat array stored box after box"""
    fp = np.zeros(level_points(fboxes))
    fr = np.zeros(level_points(cboxes))
    lib().ob_rp_flags(len(fboxes), _boxes(fboxes), len(cboxes), _boxes(cboxes), _ivec(parent), nbuffer,
                      fp.ctypes.data_as(DP), fr.ctypes.data_as(DP))
    return fp, fr


def prolong_region(fbox, cbox, fp_box, nbuffer, zone, order):
    """returns (n, ifbox, ifbox0, icbox)"""
    a, b, c = (C.c_int * 6)(), (C.c_int * 6)(), (C.c_int * 6)()
    fp_box = np.ascontiguousarray(fp_box, dtype=np.float64)
    n = lib().ob_prolong_region(C.byref(fbox), C.byref(cbox), fp_box.ctypes.data_as(DP), nbuffer, zone,
                                order, a, b, c)
    return n, list(a), list(b), list(c)


def restrict_region(fbox, cbox, fr_box, nbuffer, order):
    """returns (n, icbox, ifbox)"""
    a, b = (C.c_int * 6)(), (C.c_int * 6)()
    fr_box = np.ascontiguousarray(fr_box, dtype=np.float64)
    n = lib().ob_restrict_region(C.byref(fbox), C.byref(cbox), fr_box.ctypes.data_as(DP), nbuffer, order,
                                 a, b)
    return n, list(a), list(b)


def prolong_level(fboxes, cboxes, parent, fp, nbuffer, zone, order, cv, fv):
    """prolong_varlist_boxtobox_N for nvar variables: cv (nvar, ncoarse),
    fv (nvar, nfine) updated in place; returns the status (0 or -(b+1))"""
    assert cv.flags.c_contiguous and fv.flags.c_contiguous and cv.shape[0] == fv.shape[0]
    return lib().ob_prolong_level(len(fboxes), _boxes(fboxes), len(cboxes), _boxes(cboxes), _ivec(parent),
                                  np.ascontiguousarray(fp).ctypes.data_as(DP), nbuffer, zone, order,
                                  cv.shape[0], cv.ctypes.data_as(DP), fv.ctypes.data_as(DP))


def restrict_level(fboxes, cboxes, parent, fr, nbuffer, order, fv, cv):
    assert cv.flags.c_contiguous and fv.flags.c_contiguous and cv.shape[0] == fv.shape[0]
    return lib().ob_restrict_level(len(fboxes), _boxes(fboxes), len(cboxes), _boxes(cboxes), _ivec(parent),
                                   np.ascontiguousarray(fr).ctypes.data_as(DP), nbuffer, order, cv.shape[0],
                                   fv.ctypes.data_as(DP), cv.ctypes.data_as(DP))


# ---------------------------------------------------------------- con2prim
C2P_OK, C2P_ATM, C2P_COLD, C2P_FAIL, C2P_NR = 2, 1, 0, 3, 4


def c2p_params(eos="ideal", gamma=2.0, K=100.0, rhoatm=1e-20, fatm=100.0, vmax=0.999,
               errormax=1e-10, countmax=100):
    P = C2PParams()
    P.eos = 0 if eos == "ideal" else 1
    P.gamma, P.K = gamma, K
    P.rhoatm, P.fatm = rhoatm, fatm
    # atmosphere from the cold EOS (grhd_atm.c:74-75): polytrope p, eps
    kr = K * rhoatm ** (gamma - 1.0)
    P.patm = rhoatm * kr
    P.epslatm = kr / (gamma - 1.0)
    P.vmax, P.errormax, P.countmax = vmax, errormax, countmax
    return P


def p2c(P, prim5, metric):
    prim5 = np.ascontiguousarray(prim5, dtype=np.float64)
    metric = np.ascontiguousarray(metric, dtype=np.float64)
    n = prim5.shape[1]
    cons = np.zeros((5, n))
    lib().ob_p2c(C.byref(P), n, prim5.ctypes.data_as(DP), metric.ctypes.data_as(DP), cons.ctypes.data_as(DP))
    return cons


def c2p(P, cons, metric, pguess, rhoguess=None, method="hybrid"):
    """returns (prim[8,n], status[n], iters[n], cons_after[5,n])"""
    cons = np.array(cons, dtype=np.float64, order="C", copy=True)
    metric = np.ascontiguousarray(metric, dtype=np.float64)
    n = cons.shape[1]
    prim = np.zeros((8, n))
    prim[5] = pguess
    if rhoguess is not None:
        prim[0] = rhoguess
    st = np.zeros(n, dtype=np.int32)
    it = np.zeros(n, dtype=np.int32)
    fn = {"hybrid": lib().ob_c2p_hybrid, "plain": lib().ob_c2p_proot,
          "bracketed": lib().ob_c2p_bracketed}[method]
    fn(C.byref(P), n, cons.ctypes.data_as(DP), metric.ctypes.data_as(DP), prim.ctypes.data_as(DP),
       st.ctypes.data_as(IP), it.ctypes.data_as(IP))
    return prim, st, it, cons


def c2p_vanal(P, cons, metric, prim0=None):
    """grhd_c2p_vanal: returns (prim[8,n], status[n], coef[4,n] = a3, a2, a1, a0
    of the quartic in |v|); prim0 (optional) is the primitive array before
    the call (v^2 at prim[6] is left as it was)"""
    cons = np.ascontiguousarray(cons, dtype=np.float64)
    metric = np.ascontiguousarray(metric, dtype=np.float64)
    n = cons.shape[1]
    prim = np.zeros((8, n)) if prim0 is None else np.array(prim0, dtype=np.float64, order="C", copy=True)
    st = np.zeros(n, dtype=np.int32)
    coef = np.zeros((4, n))
    lib().ob_c2p_vanal(C.byref(P), n, cons.ctypes.data_as(DP), metric.ctypes.data_as(DP), prim.ctypes.data_as(DP),
                       st.ctypes.data_as(IP), coef.ctypes.data_as(DP))
    return prim, st, coef


def grhd_atm_set_u(P, cons, prim):
    """grhd_atm_set_u on u_p: cons[5, n] (D, tau, S_i) and prim[7] (detg)
    modified in place from prim[0] (rho); returns the count of points left
    non-finite (errorexit in the reference)"""
    assert cons.flags.c_contiguous and prim.flags.c_contiguous and cons.dtype == np.float64
    n = cons.shape[1]
    rho = np.ascontiguousarray(prim[0], dtype=np.float64)
    detg = prim[7]
    assert detg.flags.c_contiguous
    return int(lib().ob_grhd_atm_set_u(C.byref(P), n, cons.ctypes.data_as(DP), rho.ctypes.data_as(DP),
                                       detg.ctypes.data_as(DP)))


# ---------------------------------------------------------------- punctures
def _box_ptrs(fields):
    """double* array of per-box arrays (fields: list over components of lists over boxes)"""
    flat = [np.ascontiguousarray(a, dtype=np.float64) for comp in fields for a in comp]
    arr = (DP * len(flat))(*[a.ctypes.data_as(DP) for a in flat])
    return arr, flat


def interp_point(boxes, var, x, y, z, order=4):
    """interpolate_xyz_scalar: var is a list over boxes; returns (value, found)"""
    ptrs, keep = _box_ptrs([var])
    f = C.c_int()
    v = lib().ob_interp_point(len(boxes), _boxes(boxes), ptrs, x, y, z, order, C.byref(f))
    return v, bool(f.value)


def track_puncture(boxes, beta, beta_p, dt, pold, order=4, pin_z=False):
    """one step of track_moving_puncture_integrate inside track_puncture
    (old position kept when the new one leaves the level; pin_z keeps z);
    beta, beta_p: 3 lists over boxes"""
    pb, kb = _box_ptrs(beta)
    pp, kp = _box_ptrs(beta_p)
    po = np.ascontiguousarray(pold, dtype=np.float64)
    pn = np.zeros(3)
    lib().ob_track_puncture(len(boxes), _boxes(boxes), pb, pp, dt, order, 1 if pin_z else 0, po.ctypes.data_as(DP),
                            pn.ctypes.data_as(DP))
    return pn


# ---------------------------------------------------------------- GRHD fluxes
def hrsc_rec(kind, side, u, i):
    """kind "WENOZ" / "WENO5" / "LAG4"; side 0: value at i+1/2 from the left
    (rec1d_p_*), 1: value at i-1/2 from the right (rec1d_m_*)"""
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().ob_hrsc_rec({"WENOZ": 0, "WENO5": 1, "LAG4": 2}[kind], side, u.ctypes.data_as(DP), i)


def grhd_interface_flux(par, g, wl, wr):
    """grhd_flx1d_fvpr at one interface: g = (alpha, beta^x, gxx..gzz), w = (rho, epsl, vx, vy, vz)"""
    g, wl, wr = (np.ascontiguousarray(a, dtype=np.float64) for a in (g, wl, wr))
    F = np.zeros(5)
    lib().ob_grhd_interface_flux(C.byref(par), g.ctypes.data_as(DP), wl.ctypes.data_as(DP),
                                 wr.ctypes.data_as(DP), F.ctypes.data_as(DP))
    return F


def grhd_fluxes(box, par, metric, prim, rhs, c=0.0):
    """matter_fluxes on one box: metric (10, o, n, m), prim (5, o, n, m);
    rhs (5, o, n, m) is updated in place"""
    metric = np.ascontiguousarray(metric, dtype=np.float64)
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
    lib().ob_grhd_fluxes(C.byref(box), C.byref(par), c, _ptrs(metric), _ptrs(prim), _ptrs(rhs))
    return rhs


def grhd_sources(box, par, metric, Kij, prim, rhs, madm, c=0.0, rhoatm=0.0, order_centered=4):
    """grhd_sources on one box: metric (10, o, n, m), Kij (6, ...), prim (6,
    ...: rho, epsl, v^i, p); rhs (5, ...) updated and madm (11, ...) written
    in place at the inner points"""
    metric, Kij, prim = (np.ascontiguousarray(a, dtype=np.float64) for a in (metric, Kij, prim))
    for a in (rhs, madm):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().ob_grhd_sources(C.byref(box), C.byref(par), rhoatm, order_centered, c, _ptrs(metric), _ptrs(Kij),
                          _ptrs(prim), _ptrs(rhs), _ptrs(madm))
    return rhs, madm


def z4c_matter_source(box, par, u, madm, r, c=0.0):
    """z4_rhs_add_source on one box: u (25, o, n, m) Z4c fields, madm (11, ...)
    matter ADM variables; r (25, ...) updated in place at the inner points"""
    u, madm = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, madm))
    assert r.dtype == np.float64 and r.flags.c_contiguous
    lib().ob_z4c_matter_source(C.byref(box), C.byref(par), c, _ptrs(u), _ptrs(madm), _ptrs(r))
    return r


def z4c_to_adm(box, u, chipsipower=-4.0):
    """z4_to_adm_vl on one box: returns (g_ij (6, ...), K_ij (6, ...))"""
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.zeros((6,) + u.shape[1:])
    k = np.zeros((6,) + u.shape[1:])
    lib().ob_z4c_to_adm(C.byref(box), chipsipower, _ptrs(u), _ptrs(g), _ptrs(k))
    return g, k


# ---------------------------------------------------------------- 0d / 1d output
REDUCE_TYPES = {"max": 0, "min": 1, "norm": 2, "norminf": 3, "integral": 4}


def reduce_0d(boxes, fields, kind):
    """write_level_0d's reduction of each variable: fields is a list over
    variables of lists over boxes"""
    ptrs, keep = _box_ptrs(fields)
    out = np.zeros(len(fields))
    lib().ob_reduce_0d(len(boxes), _boxes(boxes), ptrs, len(fields), REDUCE_TYPES[kind], out.ctypes.data_as(DP))
    return out


def line_1d(boxes, line, origin=(0.0, 0.0, 0.0), maxp=100000):
    """the 1d output line ("x", "y", "z"): (coordinates, box of each point or -1, index in that box)"""
    crd = np.zeros(maxp)
    bo = np.zeros(maxp, dtype=np.int32)
    ix = np.zeros(maxp, dtype=np.int64)
    n = lib().ob_line_1d(len(boxes), _boxes(boxes), {"x": 1, "y": 2, "z": 3}[line], *origin, maxp,
                         crd.ctypes.data_as(DP), bo.ctypes.data_as(IP), ix.ctypes.data_as(C.POINTER(C.c_long)))
    return crd[:n], bo[:n], ix[:n]
