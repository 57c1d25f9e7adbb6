"""baum_amd: Python view of the MI355X (gfx950) BAM hot-path library.

The product is libbamhip.so (HIP kernels + C host drivers) behind the C ABI
declared in include/bamhip.h.  This module binds that ABI with ctypes so the
tests and the benchmark can drive it with numpy arrays; it holds no compute
of its own, and every entry point fails loudly if the library is missing.

Names follow the reference: a Level mirrors BAM's tL (boxes concatenated in
one index space, one device array per variable index), variable lists are
lists of variable indices as in tVarList, and parameters are given with the
reference's parameter-file names (bam_z4.c, bam_main.c).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libbamhip.so")
_LIB = None

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

NV = 25
VAR_NAMES = ["gxx", "gxy", "gxz", "gyy", "gyz", "gzz", "chi",
             "Axx", "Axy", "Axz", "Ayy", "Ayz", "Azz", "Khat",
             "Gx", "Gy", "Gz", "Theta", "alpha",
             "betax", "betay", "betaz", "Bx", "By", "Bz"]

REFBOUND, NOTBOUND, PHYBOUND, GHOBOUND, SYMBOUND, PERBOUND = -1, 0, 1, 2, 3, 6


class BamhipError(RuntimeError):
    pass


class Box(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("o", C.c_int),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("skip", C.c_int * 6), ("bflag", C.c_int * 6)]


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


def _sig(L, name, res, args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = args


def lib():
    """load libbamhip.so (raises if it is not built)"""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise BamhipError("baum_amd/libbamhip.so is not built; run __graft_entry__.build()")
    L = C.CDLL(LIBPATH)
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    _sig(L, "bamhip_init", i, [i])
    _sig(L, "bamhip_finalize", i, [])
    _sig(L, "bamhip_last_error", C.c_char_p, [])
    _sig(L, "bamhip_stream", vp, [])
    _sig(L, "bamhip_synchronize", i, [])
    _sig(L, "bamhip_malloc", i, [C.POINTER(vp), C.c_size_t])
    _sig(L, "bamhip_free", i, [vp])
    _sig(L, "bamhip_memcpy_htod", i, [vp, vp, C.c_size_t])
    _sig(L, "bamhip_memcpy_dtoh", i, [vp, vp, C.c_size_t])
    _sig(L, "bamhip_level_create", i, [C.POINTER(vp), i, i, C.POINTER(Box), i])
    _sig(L, "bamhip_level_free", i, [vp])
    _sig(L, "bamhip_level_npoints", i, [vp, C.POINTER(C.c_long)])
    _sig(L, "bamhip_level_enable", i, [vp, i])
    _sig(L, "bamhip_level_disable", i, [vp, i])
    _sig(L, "bamhip_level_ptr", vp, [vp, i])
    _sig(L, "bamhip_level_put", i, [vp, i, DP])
    _sig(L, "bamhip_level_get", i, [vp, i, DP])
    _sig(L, "bamhip_level_fill", i, [vp, i, d])
    _sig(L, "bamhip_level_set_check", i, [vp, i])
    _sig(L, "bamhip_level_nonfinite", i, [vp, i, IP])
    _sig(L, "bamhip_level_get_check", i, [vp, IP])
    _sig(L, "bamhip_z4c_rhs", i, [vp, C.POINTER(Z4cParams), IP, IP, d, IP])
    _sig(L, "bamhip_z4c_tune_mode", i, [i])
    _sig(L, "bamhip_z4c_tune_clear", i, [])
    _sig(L, "bamhip_z4c_set_overlap", i, [i])
    _sig(L, "bamhip_z4c_set_rk_fusion", i, [i])
    _sig(L, "bamhip_z4c_set_xcd_map", i, [i])
    _sig(L, "bamhip_z4c_variant_count", i, [i])
    _sig(L, "bamhip_z4c_variant_name", C.c_char_p, [i, i])
    _sig(L, "bamhip_z4c_force_variant", i, [i, i])
    _sig(L, "bamhip_z4c_tune_report", i, [C.c_char_p, i])
    _sig(L, "bamhip_z4c_set_boundary", i, [vp, C.POINTER(Z4cParams), IP, IP, d, IP, i])
    _sig(L, "bamhip_vladd", i, [vp, i, IP, d, IP, d, IP])
    _sig(L, "bamhip_rk4_stage_update", i, [vp, i, IP, IP, IP, IP, d, d])
    _sig(L, "bamhip_z4c_rk4_step", i, [vp, C.POINTER(Z4cParams), d, IP, IP, IP, IP, i, i])
    _sig(L, "bamhip_z4c_rk4_stages", i, [vp, C.POINTER(Z4cParams), d, IP, IP, IP, IP, i])
    _sig(L, "bamhip_z4c_rk4_step_bo", i, [vp, C.POINTER(Z4cParams), d, IP, IP, IP, IP, i, IP, i, i, i])
    _sig(L, "bamhip_z4c_algcon", i, [vp, IP, IP, i, i])
    _sig(L, "bamhip_z4c_ham_l2", i, [vp, IP, d, i, DP])
    _sig(L, "bamhip_z4c_constraints_l2", i, [vp, IP, d, i, IP, DP])
    _sig(L, "bamhip_set_rp_flags", i, [vp, vp, IP, i, i, i])
    _sig(L, "bamhip_restrict_varlist", i, [vp, vp, i, IP, IP, i, i])
    _sig(L, "bamhip_prolong_varlist", i, [vp, vp, i, IP, IP, i, i])
    _sig(L, "bamhip_restrict_prolong_varlist", i, [vp, vp, i, IP, IP, i, i])
    _sig(L, "bamhip_bo_interp_begin", i, [vp, i, i, i, IP, IP, IP, IP, IP, IP])
    _sig(L, "bamhip_bo_interp_end", i, [vp, i, i, i, IP, IP])
    _sig(L, "bamhip_puncture_track", i, [vp, IP, IP, d, i, i, i, vp])
    _sig(L, "bamhip_timer_enable", i, [i])
    _sig(L, "bamhip_timer_reset", i, [])
    _sig(L, "bamhip_timer_get", i, [C.c_char_p, DP, C.POINTER(C.c_long)])
    _LIB = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().bamhip_last_error().decode(errors="replace")
        raise BamhipError("%s failed (%d): %s" % (what, rc, msg))
    return rc


_INIT = False


def init(device=0):
    global _INIT
    if not _INIT:
        check(lib().bamhip_init(device), "bamhip_init")
        _INIT = True


def synchronize():
    check(lib().bamhip_synchronize(), "bamhip_synchronize")


# ---------------------------------------------------------------- parameters
def _getv(value, word):
    """BAM's Getv: whole-word membership in a parameter value list
    (src/main/main/parameters.c:883-910)"""
    return 1 if word in str(value).split() else 0


Z4C_DEFAULTS = {
    "order_centered": 4, "order_advection": 4, "advection_lopsided": 1,
    "order_dissipation": 4, "dissipation_factor": 0.1,
    "z4_kappa1": 0.0, "z4_kappa2": 0.0, "z4_zeroThetaRHS": "no",
    "z4_subtractA": "yes", "z4_normalizedetg": "yes",
    "z4_lapse": "1+log withshift", "z4_lapseharmonicf": 2.0,
    "z4_shift": "withShiftadv", "z4_shiftalphapower": 0.0, "z4_shiftgammacoeff": 1.0,
    "z4_shiftdriver": 2.0, "z4_chi_div_floor": -1000.0, "z4_chi_psipower": -4.0,
}


def z4c_params(**over):
    """parameter snapshot from BAM parameter names (defaults: the puncture
    configuration used throughout, 1+log withshift, Gamma driver eta=2, KO 0.1)"""
    dd = dict(Z4C_DEFAULTS)
    dd.update(over)
    p = Z4cParams()
    p.order_centered = int(dd["order_centered"])
    p.order_advection = int(dd["order_advection"])
    p.advection_lopsided = int(dd["advection_lopsided"])
    p.order_dissipation = int(dd["order_dissipation"])
    p.dissfactor = float(dd["dissipation_factor"]) / 2.0 ** p.order_dissipation
    p.kappa1, p.kappa2 = float(dd["z4_kappa1"]), float(dd["z4_kappa2"])
    p.zeroThetaRHS = float(_getv(dd["z4_zeroThetaRHS"], "no"))
    p.subtractA = _getv(dd["z4_subtractA"], "yes")
    p.normalizedetg = _getv(dd["z4_normalizedetg"], "yes")
    p.oploglapse = _getv(dd["z4_lapse"], "1+log")
    p.oploglapse2 = _getv(dd["z4_lapse"], "1+log2")
    p.oplogwithshift = _getv(dd["z4_lapse"], "withshift")
    p.harmoniclapse = _getv(dd["z4_lapse"], "harmonic")
    p.lapseharmonicf = float(dd["z4_lapseharmonicf"])
    p.gamma0factor = float(_getv(dd["z4_shift"], "gamma0"))
    p.gamma2factor = float(_getv(dd["z4_shift"], "gamma2"))
    p.withGadv = float(_getv(dd["z4_shift"], "withGadv"))
    p.withShiftadv = float(_getv(dd["z4_shift"], "withShiftadv"))
    p.withBadv = float(_getv(dd["z4_shift"], "withBadv"))
    p.withB = float(_getv(dd["z4_shift"], "withB"))
    p.shiftalphapower = float(dd["z4_shiftalphapower"])
    p.shiftgammacoeff = float(dd["z4_shiftgammacoeff"])
    p.shiftdriver = float(dd["z4_shiftdriver"])
    p.chiDivFloor = float(dd["z4_chi_div_floor"])
    p.chipsipower = float(dd["z4_chi_psipower"])
    return p


class DissipationRule(C.Structure):
    _fields_ = [("dissipation_factor", C.c_double), ("order_dissipation", C.c_int),
                ("amr_move_lcube", C.c_int), ("physics_z4d", C.c_int),
                ("factor_level0", C.c_double), ("factor_shells", C.c_double), ("factor_move", C.c_double)]


def dissipation_factor(l, dissipation_factor=0.1, order_dissipation=4, amr_move_lcube=0, physics="z4",
                       dissipation_factor_level0=-1.0, dissipation_factor_shells=-1.0,
                       dissipation_factor_move=-1.0, shells=False):
    """get_dissipation_factor(level l) (evolve.c:486-530) from BAM parameter
    names; the value for z4c_params' dissipation factor per level"""
    R = DissipationRule(float(dissipation_factor), int(order_dissipation), int(amr_move_lcube),
                        _getv(physics, "Z4d"), float(dissipation_factor_level0),
                        float(dissipation_factor_shells), float(dissipation_factor_move))
    out = C.c_double()
    L = lib()
    L.bamhip_dissipation_factor.restype = C.c_int
    L.bamhip_dissipation_factor.argtypes = [C.POINTER(DissipationRule), C.c_int, C.c_int, DP]
    check(L.bamhip_dissipation_factor(C.byref(R), int(l), 1 if shells else 0, C.byref(out)),
          "bamhip_dissipation_factor")
    return out.value


def make_box(m, n, o, x0, y0, z0, dx, dy=None, dz=None, skip=(1,) * 6, bflag=(PHYBOUND,) * 6):
    b = Box()
    b.m, b.n, b.o = m, n, o
    b.x0, b.y0, b.z0 = x0, y0, z0
    b.dx = dx
    b.dy = dx if dy is None else dy
    b.dz = dx if dz is None else dz
    for f in range(6):
        b.skip[f] = skip[f]
        b.bflag[f] = bflag[f]
    return b


def _iarr(lst):
    return (C.c_int * len(lst))(*[int(x) for x in lst])


class Level:
    """device mirror of one BAM level (tL) with `nvariables` variable slots"""

    def __init__(self, boxes, nvariables, l=0):
        init()
        self.boxes = list(boxes)
        arr = (Box * len(self.boxes))(*self.boxes)
        h = C.c_void_p()
        check(lib().bamhip_level_create(C.byref(h), l, len(self.boxes), arr, nvariables), "bamhip_level_create")
        self.h = h
        n = C.c_long()
        check(lib().bamhip_level_npoints(self.h, C.byref(n)), "bamhip_level_npoints")
        self.npoints = n.value
        self.nvariables = nvariables
        self.offsets = []
        off = 0
        for b in self.boxes:
            self.offsets.append(off)
            off += b.m * b.n * b.o

    @classmethod
    def adopt(cls, handle, nvariables):
        """wrap a level the library created (a level bamhip_move_box made)"""
        L = lib()
        _sig(L, "bamhip_level_info", C.c_int, [C.c_void_p, IP, IP])
        _sig(L, "bamhip_level_box", C.c_int, [C.c_void_p, C.c_int, C.POINTER(Box), C.POINTER(C.c_long)])
        self = cls.__new__(cls)
        self.h = C.c_void_p(handle)
        lv, nb = C.c_int(), C.c_int()
        check(L.bamhip_level_info(self.h, C.byref(lv), C.byref(nb)), "bamhip_level_info")
        self.boxes, self.offsets = [], []
        for b in range(nb.value):
            bx, off = Box(), C.c_long()
            check(L.bamhip_level_box(self.h, b, C.byref(bx), C.byref(off)), "bamhip_level_box")
            self.boxes.append(bx)
            self.offsets.append(off.value)
        n = C.c_long()
        check(L.bamhip_level_npoints(self.h, C.byref(n)), "bamhip_level_npoints")
        self.npoints = n.value
        self.nvariables = nvariables
        return self

    def free(self):
        if self.h:
            lib().bamhip_level_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def enable(self, vars_):
        for v in vars_:
            check(lib().bamhip_level_enable(self.h, int(v)), "bamhip_level_enable")

    def put(self, var, data):
        a = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        assert a.size == self.npoints, (a.size, self.npoints)
        check(lib().bamhip_level_put(self.h, int(var), a.ctypes.data_as(DP)), "bamhip_level_put")

    def get(self, var):
        a = np.empty(self.npoints)
        check(lib().bamhip_level_get(self.h, int(var), a.ctypes.data_as(DP)), "bamhip_level_get")
        return a

    def fill(self, var, value):
        check(lib().bamhip_level_fill(self.h, int(var), float(value)), "bamhip_level_fill")

    def put_state(self, idx, fields):
        """fields: array (len(idx), npoints) or list of per-box (nv, o, n, m)"""
        for k, v in enumerate(idx):
            self.put(v, fields[k])

    def get_state(self, idx):
        return np.array([self.get(v) for v in idx])

    def box_view(self, flat, b):
        bx = self.boxes[b]
        o = self.offsets[b]
        return flat[..., o:o + bx.m * bx.n * bx.o].reshape(flat.shape[:-1] + (bx.o, bx.n, bx.m))

    # -- Z4c
    def z4c_rhs(self, params, unew, ucur, upre=None, c=0.0):
        check(lib().bamhip_z4c_rhs(self.h, C.byref(params), _iarr(unew), _iarr(upre) if upre is not None else None,
                                   c, _iarr(ucur)), "bamhip_z4c_rhs")

    def z4c_set_boundary(self, params, unew, ucur, upre=None, c=0.0, pgh=0):
        check(lib().bamhip_z4c_set_boundary(self.h, C.byref(params), _iarr(unew),
                                            _iarr(upre) if upre is not None else None, c, _iarr(ucur), pgh),
              "bamhip_z4c_set_boundary")

    def vladd(self, r, ca, a, cb, b):
        check(lib().bamhip_vladd(self.h, len(r), _iarr(r), ca, _iarr(a) if a is not None else None, cb,
                                 _iarr(b) if b is not None else None), "bamhip_vladd")

    def z4c_rk4_step(self, params, dt, uc, up, r, v, pgh=0, algcon=True):
        check(lib().bamhip_z4c_rk4_step(self.h, C.byref(params), dt, _iarr(uc), _iarr(up), _iarr(r), _iarr(v),
                                        pgh, 1 if algcon else 0), "bamhip_z4c_rk4_step")

    def z4c_rk4_stages(self, params, dt, uc, up, r, v, pgh=0):
        check(lib().bamhip_z4c_rk4_stages(self.h, C.byref(params), dt, _iarr(uc), _iarr(up), _iarr(r), _iarr(v), pgh),
              "bamhip_z4c_rk4_stages")

    def z4c_rk4_step_bo(self, params, dt, uc, up, r, v, upp, timeinterpolation, iteration, pgh=0, algcon=True):
        """RK4 step with the Berger-Oliger list operations around it (upp=None: finest level)"""
        check(lib().bamhip_z4c_rk4_step_bo(self.h, C.byref(params), dt, _iarr(uc), _iarr(up), _iarr(r), _iarr(v), pgh,
                                           _iarr(upp) if upp is not None else None, timeinterpolation, iteration,
                                           1 if algcon else 0), "bamhip_z4c_rk4_step_bo")

    def puncture_track(self, beta, beta_p, dt, pos, order=4, pin_z=False):
        """one tracking step of the punctures whose coordinates are in the
        device array pos (shape (np, 3)), updated in place"""
        check(lib().bamhip_puncture_track(self.h, _iarr(beta), _iarr(beta_p), dt, order, 1 if pin_z else 0,
                                          pos.shape[0], pos.ptr), "bamhip_puncture_track")

    def grhd_fluxes(self, par, unew, metric, prim, c=0.0):
        """matter_fluxes: unew (D, tau, S_i) -= flux divergence; metric: alpha,
        beta^i, ADM g_ij (10 indices); prim: rho, epsl, v^i (5 indices)"""
        L = lib()
        _sig(L, "bamhip_grhd_fluxes", C.c_int, [C.c_void_p, C.POINTER(HrscParams), IP, C.c_double, IP, IP])
        check(L.bamhip_grhd_fluxes(self.h, C.byref(par), _iarr(unew), c, _iarr(metric), _iarr(prim)),
              "bamhip_grhd_fluxes")

    def grhd_sources(self, par, unew, metric, Kij, prim, madm, c=0.0, rhoatm=0.0, order_centered=4):
        """grhd_sources: unew (D, tau, S_i) += source terms at the inner
        points; madm (rho, S_i, S_ij, S) written"""
        L = lib()
        _sig(L, "bamhip_grhd_sources", C.c_int, [C.c_void_p, C.POINTER(HrscParams), C.c_double, C.c_int, IP,
                                                  C.c_double, IP, IP, IP, IP])
        check(L.bamhip_grhd_sources(self.h, C.byref(par), rhoatm, order_centered, _iarr(unew), c, _iarr(metric),
                                    _iarr(Kij), _iarr(prim), _iarr(madm)), "bamhip_grhd_sources")

    def z4c_matter_source(self, par, unew, ucur, madm, c=0.0):
        """z4_rhs_add_source: the matter terms of the Z4c right-hand sides"""
        L = lib()
        _sig(L, "bamhip_z4c_matter_source", C.c_int, [C.c_void_p, C.POINTER(Z4cParams), IP, C.c_double, IP, IP])
        check(L.bamhip_z4c_matter_source(self.h, C.byref(par), _iarr(unew), c, _iarr(ucur), _iarr(madm)),
              "bamhip_z4c_matter_source")

    def punctures_bl(self, masses, positions, adm_g, adm_K):
        """Brill-Lindquist ADM data (psi^4 delta_ij, K_ij = 0)"""
        L = lib()
        _sig(L, "bamhip_punctures_bl", C.c_int, [C.c_void_p, C.c_int, DP, DP, IP, IP])
        m = np.ascontiguousarray(masses, dtype=np.float64)
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        check(L.bamhip_punctures_bl(self.h, len(m), m.ctypes.data_as(DP), p.ctypes.data_as(DP), _iarr(adm_g),
                                    _iarr(adm_K)), "bamhip_punctures_bl")

    def tov_set(self, star, adm_g, adm_K, alpha, prim, center=(0.0, 0.0, 0.0), riso=None, rsch=None, phi=None,
                h=None):
        """set_TOV: the star of tov_h(...) (or explicit tables) on the level;
        prim = rho, epsl, vx, vy, vz, p"""
        L = lib()
        _sig(L, "bamhip_tov_set", C.c_int, [C.c_void_p, C.c_int, DP, DP, DP, DP, C.c_double, C.c_double,
                                             C.c_double, DP, IP, IP, C.c_int, IP])
        t = [np.ascontiguousarray(star[k] if a is None else a, dtype=np.float64)
             for k, a in (("riso", riso), ("r", rsch), ("phi", phi), ("h", h))]
        c = np.ascontiguousarray(center, dtype=np.float64)
        check(L.bamhip_tov_set(self.h, t[0].size, *[a.ctypes.data_as(DP) for a in t], star["M"], star["K"],
                               star["Gamma"], c.ctypes.data_as(DP), _iarr(adm_g), _iarr(adm_K), int(alpha),
                               _iarr(prim)), "bamhip_tov_set")

    def adm_to_z4c(self, adm_g, adm_K, u, chipsipower=-4.0, initial_lapse=2, zero_shift=True):
        """z4 initialisation: lapse (0 unchanged, 1 one, 2 precollapsed), shift, z4_init, G boundary"""
        L = lib()
        _sig(L, "bamhip_adm_to_z4c", C.c_int, [C.c_void_p, IP, IP, IP, C.c_double, C.c_int, C.c_int])
        check(L.bamhip_adm_to_z4c(self.h, _iarr(adm_g), _iarr(adm_K), _iarr(u), chipsipower, initial_lapse,
                                  1 if zero_shift else 0), "bamhip_adm_to_z4c")

    def set_check(self, mode):
        """non-finite check mode: 1 synchronise per call, 0 deferred flag"""
        check(lib().bamhip_level_set_check(self.h, mode), "bamhip_level_set_check")

    def get_check(self):
        m = C.c_int()
        check(lib().bamhip_level_get_check(self.h, C.byref(m)), "bamhip_level_get_check")
        return m.value

    def nonfinite(self, reset=True):
        f = C.c_int()
        check(lib().bamhip_level_nonfinite(self.h, 1 if reset else 0, C.byref(f)), "bamhip_level_nonfinite")
        return f.value

    def grhd_c2p(self, c2p, method, cons5, metric6, prim8, status, iters):
        """MATTER.c2p on the level (status / iters: DeviceArray of npoints int32)"""
        L = lib()
        _sig(L, "bamhip_grhd_c2p", C.c_int, [C.c_void_p, C.POINTER(C2PParams), C.c_int, IP, IP, IP, C.c_void_p,
                                             C.c_void_p])
        check(L.bamhip_grhd_c2p(self.h, C.byref(c2p), method, _iarr(cons5), _iarr(metric6), _iarr(prim8),
                                status.ptr, iters.ptr), "bamhip_grhd_c2p")

    def grhd_c2p_learned(self, c2p, mlp, cons5, metric6, prim8, status, iters, fallbacks=None, polish_max=8):
        """MATTER.c2p with the learned guess (the reference hybrid's control flow)"""
        L = lib()
        _sig(L, "bamhip_grhd_c2p_learned", C.c_int, [C.c_void_p, C.POINTER(C2PParams), C.c_void_p, C.c_int, IP, IP,
                                                     IP, C.c_void_p, C.c_void_p, C.c_void_p])
        check(L.bamhip_grhd_c2p_learned(self.h, C.byref(c2p), mlp.h, polish_max, _iarr(cons5), _iarr(metric6),
                                        _iarr(prim8), status.ptr, iters.ptr,
                                        fallbacks.ptr if fallbacks is not None else None), "bamhip_grhd_c2p_learned")

    def grhd_atm_set_prerhs(self, c2p, upre5, prim8):
        """grhd_atm_set_prerhs: the cold static atmosphere on the conserved
        variables upre5 (D, tau, S_i) where prim8's rho is below fatm*rhoatm"""
        L = lib()
        _sig(L, "bamhip_grhd_atm_set_prerhs", C.c_int, [C.c_void_p, C.POINTER(C2PParams), IP, IP])
        check(L.bamhip_grhd_atm_set_prerhs(self.h, C.byref(c2p), _iarr(upre5), _iarr(prim8)),
              "bamhip_grhd_atm_set_prerhs")

    def z4c_to_adm(self, u, adm_g, adm_K, chipsipower=-4.0):
        """z4_to_adm_vl: ADM g_ij, K_ij from the Z4c fields"""
        L = lib()
        _sig(L, "bamhip_z4c_to_adm", C.c_int, [C.c_void_p, IP, C.c_double, IP, IP])
        check(L.bamhip_z4c_to_adm(self.h, _iarr(u), chipsipower, _iarr(adm_g), _iarr(adm_K)), "bamhip_z4c_to_adm")

    @staticmethod
    def _z4c_grhd(z4, c2p, hrsc, adm_g, adm_K, prim, madm, status, iters, method, mlp, polish_max, fallbacks, pgh,
                  atm_prerhs=True):
        keep = [_iarr(x) for x in (adm_g, adm_K, prim, madm)]
        S = Z4cGrhdStruct(C.pointer(z4), C.pointer(c2p), C.pointer(hrsc), method, mlp.h if mlp is not None else None,
                          polish_max, fallbacks.ptr if fallbacks is not None else None, pgh,
                          *[C.cast(k, IP) for k in keep], status.ptr, iters.ptr, 1 if atm_prerhs else 0)
        return S, keep

    def z4c_grhd_rhs(self, z4, c2p, hrsc, unew, ucur, adm_g, adm_K, prim, madm, status, iters, upre=None, c=0.0,
                     method=0, mlp=None, polish_max=8, fallbacks=None, pgh=0, atm_prerhs=True):
        """evolve_rhs of the coupled Z4c + GRHD system; lists of 30 indices
        (25 Z4c fields, then D, tau, S_i); status / iters: DeviceArray int32;
        atm_prerhs: the atmosphere reset of upre after the c2p (when upre is given)"""
        L = lib()
        _sig(L, "bamhip_z4c_grhd_rhs", C.c_int, [C.c_void_p, C.POINTER(Z4cGrhdStruct), IP, IP, C.c_double, IP])
        S, keep = self._z4c_grhd(z4, c2p, hrsc, adm_g, adm_K, prim, madm, status, iters, method, mlp, polish_max,
                                 fallbacks, pgh, atm_prerhs)
        check(L.bamhip_z4c_grhd_rhs(self.h, C.byref(S), _iarr(unew), _iarr(upre) if upre is not None else None, c,
                                    _iarr(ucur)), "bamhip_z4c_grhd_rhs")

    def z4c_grhd_rk4_step(self, z4, c2p, hrsc, dt, uc, up, r, v, adm_g, adm_K, prim, madm, status, iters,
                          method=0, mlp=None, polish_max=8, fallbacks=None, pgh=0, algcon=True, atm_prerhs=True):
        """one RK4 step of the coupled system (lists of 30 indices)"""
        L = lib()
        _sig(L, "bamhip_z4c_grhd_rk4_step", C.c_int,
             [C.c_void_p, C.POINTER(Z4cGrhdStruct), C.c_double, IP, IP, IP, IP, C.c_int])
        S, keep = self._z4c_grhd(z4, c2p, hrsc, adm_g, adm_K, prim, madm, status, iters, method, mlp, polish_max,
                                 fallbacks, pgh, atm_prerhs)
        check(L.bamhip_z4c_grhd_rk4_step(self.h, C.byref(S), dt, _iarr(uc), _iarr(up), _iarr(r), _iarr(v),
                                         1 if algcon else 0), "bamhip_z4c_grhd_rk4_step")

    def reduce(self, vars, kind):
        """write_level_0d's reduction ("max", "min", "norm", "norminf", "integral") of each variable"""
        L = lib()
        _sig(L, "bamhip_level_reduce", C.c_int, [C.c_void_p, C.c_int, IP, C.c_int, DP])
        out = (C.c_double * len(vars))()
        check(L.bamhip_level_reduce(self.h, len(vars), _iarr(vars), OUTPUT_0D.index(kind), out),
              "bamhip_level_reduce")
        return np.array(out[:])


    def output_0d(self, vars, names, kind, time, outdir):
        """append one 0d line per variable in the reference's format"""
        L = lib()
        _sig(L, "bamhip_output_0d", C.c_int, [C.c_void_p, C.c_int, IP, C.POINTER(C.c_char_p), C.c_int, C.c_double,
                                               C.c_char_p])
        nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
        check(L.bamhip_output_0d(self.h, len(vars), _iarr(vars), nm, OUTPUT_0D.index(kind), time,
                                 str(outdir).encode()), "bamhip_output_0d")

    def output_1d(self, vars, names, line, time, outdir, origin=(0.0, 0.0, 0.0)):
        """append one 1d block per variable ("x", "y" or "z" line) in the reference's format"""
        L = lib()
        _sig(L, "bamhip_output_1d", C.c_int, [C.c_void_p, C.c_int, IP, C.POINTER(C.c_char_p), C.c_int, C.c_double,
                                               C.c_double, C.c_double, C.c_double, C.c_char_p])
        nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
        check(L.bamhip_output_1d(self.h, len(vars), _iarr(vars), nm, "dxyz".index(line), *origin, time,
                                 str(outdir).encode()), "bamhip_output_1d")

    def z4c_algcon(self, g, A, subtractA=1, normalizedetg=1):
        check(lib().bamhip_z4c_algcon(self.h, _iarr(g), _iarr(A), subtractA, normalizedetg), "bamhip_z4c_algcon")

    def z4c_ham_l2(self, u, chipsipower=-4.0, ham=-1):
        out = C.c_double()
        check(lib().bamhip_z4c_ham_l2(self.h, _iarr(u), chipsipower, ham, C.byref(out)), "bamhip_z4c_ham_l2")
        return out.value

    def z4c_constraints_l2(self, u, chipsipower=-4.0, ham=-1, mom=None):
        """(|ham|, |mom_x|, |mom_y|, |mom_z|) over the inner points
        (adm.c:14-74); ham / mom: optional variable indices for the fields"""
        out = (C.c_double * 4)()
        check(lib().bamhip_z4c_constraints_l2(self.h, _iarr(u), chipsipower, ham,
                                              _iarr(mom) if mom is not None else None, out),
              "bamhip_z4c_constraints_l2")
        return np.array(out[:])

    # -- Berger-Oliger time interpolation (evolve.c:298-340, 367-380)
    def bo_interp_begin(self, timeinterpolation, iteration, upp, uc, up=None, up2=None, up3=None, up4=None):
        opt = [_iarr(x) if x is not None else None for x in (up, up2, up3, up4)]
        check(lib().bamhip_bo_interp_begin(self.h, timeinterpolation, iteration, len(upp), _iarr(upp), _iarr(uc),
                                           *opt), "bamhip_bo_interp_begin")

    def bo_interp_end(self, timeinterpolation, iteration, upp, uc):
        check(lib().bamhip_bo_interp_end(self.h, timeinterpolation, iteration, len(upp), _iarr(upp), _iarr(uc)),
              "bamhip_bo_interp_end")


# ---------------------------------------------------------------- FMR coupling
def set_rp_flags(lc, lf, parent, nbuffer, fp, fr):
    """setRPflags(lc, lf): links lf to lc (fine box b in coarse box parent[b])"""
    check(lib().bamhip_set_rp_flags(lc.h, lf.h, _iarr(parent), nbuffer, fp, fr), "bamhip_set_rp_flags")


def restrict_varlist(lf, lc, uf, uc, nbuffer=3, order=4):
    assert len(uf) == len(uc)
    check(lib().bamhip_restrict_varlist(lf.h, lc.h, len(uf), _iarr(uf), _iarr(uc), nbuffer, order),
          "bamhip_restrict_varlist")


def prolong_varlist(lc, lf, uc, uf, nbuffer=3, order=4):
    assert len(uf) == len(uc)
    check(lib().bamhip_prolong_varlist(lc.h, lf.h, len(uc), _iarr(uc), _iarr(uf), nbuffer, order),
          "bamhip_prolong_varlist")


def restrict_prolong_varlist(lc, lf, uc, uf, nbuffer=3, order=4):
    assert len(uf) == len(uc)
    check(lib().bamhip_restrict_prolong_varlist(lc.h, lf.h, len(uc), _iarr(uc), _iarr(uf), nbuffer, order),
          "bamhip_restrict_prolong_varlist")


# ---------------------------------------------------------------- hierarchy
class FmrHierarchyStruct(C.Structure):
    _fields_ = [("lmax", C.c_int), ("levels", C.POINTER(C.c_void_p)), ("params", C.POINTER(Z4cParams)),
                ("dt0", C.c_double), ("iteration", C.POINTER(C.c_int)),
                ("u", IP), ("up", IP), ("upp", IP), ("r", IP), ("v", IP),
                ("order_RP", C.c_int), ("timeinterpolation", C.c_int), ("pgh", C.c_int),
                ("npunc", C.c_int), ("punc", C.c_void_p), ("punc_flags", C.c_int),
                ("move_nxyz", C.c_int), ("move_lcube", C.c_int),
                ("nbuffer", C.c_int), ("fp", C.c_int), ("fr", C.c_int),
                ("nx", C.c_int), ("nz", C.c_int), ("n_amr_nxyz", C.c_int), ("amr_nxyz", IP),
                ("addextra", C.c_int)]


class NestedSizesStruct(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("move_nxyz", C.c_int), ("move_lcube", C.c_int),
                ("nlist", C.c_int), ("list", IP), ("nbuffer", C.c_int), ("addextra", C.c_int)]


class NestedSizes:
    """the box-size parameters of find_nested_bboxes: nx, nz of the coarsest
    grid, amr_move_nxyz, amr_move_lcube, the amr_nxyz list, amr_nbuffer and
    amr_buffer_addextra (the reference default: yes)"""

    def __init__(self, nx, nz=None, move_nxyz=0, move_lcube=0, nxyz_list=(), nbuffer=6, addextra=True):
        self.nx, self.nz = int(nx), int(nx if nz is None else nz)
        self.move_nxyz, self.move_lcube = int(move_nxyz), int(move_lcube)
        self.nxyz_list = [int(x) for x in nxyz_list]
        self.nbuffer, self.addextra = int(nbuffer), bool(addextra)
        self._list = (C.c_int * max(len(self.nxyz_list), 1))(*self.nxyz_list)
        self.s = NestedSizesStruct(self.nx, self.nz, self.move_nxyz, self.move_lcube, len(self.nxyz_list),
                                   C.cast(self._list, IP), self.nbuffer, 1 if self.addextra else 0)


def nested_box_size(sizes, lfine):
    """bamhip_nested_box_size (host only): coarse points per direction of the
    boxes of fine level lfine"""
    L = lib()
    _sig(L, "bamhip_nested_box_size", C.c_int, [C.POINTER(NestedSizesStruct), C.c_int, IP])
    n = (C.c_int * 3)()
    check(L.bamhip_nested_box_size(C.byref(sizes.s), int(lfine), n), "bamhip_nested_box_size")
    return list(n)


def find_nested_bboxes(sizes, lfine, points, h):
    """bamhip_find_nested_bboxes (host only): the coarse bounding boxes
    (nboxes, 6) and index boxes of fine level lfine around the points (np, 3)
    on the coarse spacing h"""
    L = lib()
    _sig(L, "bamhip_find_nested_bboxes", C.c_int,
         [C.POINTER(NestedSizesStruct), C.c_int, C.c_int, DP, DP, IP, DP, IP])
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    hh = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (3,)))
    nb = C.c_int()
    bb = (C.c_double * 60)()
    ib = (C.c_int * 60)()
    check(L.bamhip_find_nested_bboxes(C.byref(sizes.s), int(lfine), p.shape[0], p.ctypes.data_as(DP),
                                      hh.ctypes.data_as(DP), C.byref(nb), bb, ib), "bamhip_find_nested_bboxes")
    n = nb.value
    return np.array(bb[:6 * n]).reshape(n, 6), np.array(ib[:6 * n]).reshape(n, 6)


def level_boxes_of_bboxes(bb, ib, h):
    """make_level_bboxes (box.c:465-560): the fine boxes of coarse bounding
    boxes (2 (cells + 1) points per direction, the first point half a fine
    spacing inside) as (m, n, o, x0, y0, z0, dx) tuples"""
    out = []
    hh = np.broadcast_to(np.asarray(h, dtype=np.float64), (3,))
    for b in range(bb.shape[0]):
        m, n_, o = (2 * (int(ib[b, 2 * d + 1]) + 1) for d in range(3))
        out.append((m, n_, o, bb[b, 0] - hh[0] / 4, bb[b, 2] - hh[1] / 4, bb[b, 4] - hh[2] / 4, hh[0] / 2))
    return out


def tov_h(hc, K=100.0, Gamma=2.0, npts=100000):
    """bamhip_tov_h (host only): the TOV star of the polytrope integrated in
    the log enthalpy; returns the tables and M, R, Riso, Mb"""
    L = lib()
    _sig(L, "bamhip_tov_h", C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int] + [DP] * 8)
    t = {k: np.empty(npts) for k in ("r", "m", "h", "rho", "pre", "phi", "riso")}
    out = np.empty(4)
    check(L.bamhip_tov_h(hc, K, Gamma, npts, *[t[k].ctypes.data_as(DP) for k in ("r", "m", "h", "rho", "pre", "phi",
                                                                              "riso")],
                         out.ctypes.data_as(DP)), "bamhip_tov_h")
    t.update(M=out[0], R=out[1], Riso=out[2], Mb=out[3], K=K, Gamma=Gamma)
    return t


def copy_overlap(lold, lnew, vars_):
    """bamhip_level_copy_overlap: the points the levels share, old -> new"""
    L = lib()
    _sig(L, "bamhip_level_copy_overlap", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, IP])
    check(L.bamhip_level_copy_overlap(lold.h, lnew.h, len(vars_), _iarr(vars_)), "bamhip_level_copy_overlap")


class Hierarchy:
    """Berger-Oliger hierarchy of device levels (bamhip_fmr_hierarchy).
    levels[l] are Level objects, linked with set_rp_flags beforehand;
    params: one Z4cParams per level."""

    def __init__(self, levels, params, dt0, u, up, upp, r, v, order_RP=4, timeinterpolation=3, pgh=0,
                 punctures=None, pin_z=False, sizes=None, rp_flags=(3, -1, -1)):
        L = lib()
        hp = C.POINTER(FmrHierarchyStruct)
        _sig(L, "bamhip_advance_levelandsublevels", C.c_int, [hp, C.c_int])
        _sig(L, "bamhip_fmr_set_check", C.c_int, [hp, C.c_int])
        _sig(L, "bamhip_fmr_nonfinite", C.c_int, [hp, C.c_int, IP])
        _sig(L, "bamhip_graph_begin", C.c_int, [])
        _sig(L, "bamhip_graph_end", C.c_int, [C.POINTER(C.c_void_p)])
        _sig(L, "bamhip_graph_launch", C.c_int, [C.c_void_p])
        _sig(L, "bamhip_graph_free", C.c_int, [C.c_void_p])
        self.levels = list(levels)
        n = len(self.levels)
        self._lv = (C.c_void_p * n)(*[lv.h for lv in self.levels])
        self._par = (Z4cParams * n)(*params)
        self.iteration = (C.c_int * n)()
        self._lists = [_iarr(x) for x in (u, up, upp, r, v)]
        s = FmrHierarchyStruct()
        s.lmax = n - 1
        s.levels = C.cast(self._lv, C.POINTER(C.c_void_p))
        s.params = C.cast(self._par, C.POINTER(Z4cParams))
        s.dt0 = dt0
        s.iteration = C.cast(self.iteration, C.POINTER(C.c_int))
        s.u, s.up, s.upp, s.r, s.v = [C.cast(a, IP) for a in self._lists]
        s.order_RP, s.timeinterpolation, s.pgh = order_RP, timeinterpolation, pgh
        # punctures: device array (DeviceArray, shape (np, 3)) tracked on the finest level
        self.punctures = punctures
        if punctures is not None:
            s.npunc, s.punc, s.punc_flags = punctures.shape[0], punctures.ptr, 1 if pin_z else 0
        # sizes: NestedSizes of the moving boxes (None: the boxes stay)
        self.sizes = sizes
        if sizes is not None:
            assert sizes.move_nxyz > 0, "moving boxes need amr_move_nxyz > 0"
            assert sizes.nbuffer == rp_flags[0], "amr_nbuffer of the sizes and of the flags differ"
            s.move_nxyz, s.move_lcube = sizes.move_nxyz, sizes.move_lcube
            s.nx, s.nz, s.n_amr_nxyz, s.amr_nxyz = sizes.nx, sizes.nz, len(sizes.nxyz_list), C.cast(sizes._list, IP)
            s.addextra = 1 if sizes.addextra else 0
        s.nbuffer, s.fp, s.fr = rp_flags
        self.s = s

    def _adopt_moved(self):
        """wrap the levels bamhip_move_box replaced (the old ones are freed)"""
        for l, lv in enumerate(self.levels):
            if self._lv[l] != lv.h.value:
                lv.h = None
                self.levels[l] = Level.adopt(self._lv[l], lv.nvariables)

    def move_box(self, l):
        """move_box(g, l): the levels from l down follow the punctures;
        returns whether a level moved"""
        L = lib()
        _sig(L, "bamhip_move_box", C.c_int, [C.POINTER(FmrHierarchyStruct), C.c_int, IP])
        moved = C.c_int()
        rc = L.bamhip_move_box(C.byref(self.s), l, C.byref(moved))
        self._adopt_moved()
        check(rc, "bamhip_move_box")
        return bool(moved.value)

    def set_check(self, mode):
        check(lib().bamhip_fmr_set_check(C.byref(self.s), mode), "bamhip_fmr_set_check")

    def advance(self, l=0):
        """advance_levelandsublevels(l): one step of level l and its sublevels"""
        rc = lib().bamhip_advance_levelandsublevels(C.byref(self.s), l)
        if self.s.move_nxyz > 0:
            self._adopt_moved()
        check(rc, "bamhip_advance_levelandsublevels")

    def nonfinite(self, reset=True):
        f = C.c_int()
        check(lib().bamhip_fmr_nonfinite(C.byref(self.s), 1 if reset else 0, C.byref(f)), "bamhip_fmr_nonfinite")
        return f.value

    def capture(self, l=0):
        """capture one advance(l) into a graph (iteration counters advance as
        in an eager call); returns the graph handle"""
        check(lib().bamhip_graph_begin(), "bamhip_graph_begin")
        try:
            self.advance(l)
        finally:
            g = C.c_void_p()
            check(lib().bamhip_graph_end(C.byref(g)), "bamhip_graph_end")
        return g

    @staticmethod
    def launch(graph):
        check(lib().bamhip_graph_launch(graph), "bamhip_graph_launch")

    @staticmethod
    def free_graph(graph):
        check(lib().bamhip_graph_free(graph), "bamhip_graph_free")


# ---------------------------------------------------------------- con2prim
C2P_COLD, C2P_ATM, C2P_OK, C2P_FAIL, C2P_NR = 0, 1, 2, 3, 4
C2P_HYBRID, C2P_PLAIN, C2P_BRACKETED, C2P_LEARNED, C2P_VANAL = 0, 1, 2, 3, 4


class C2PParams(C.Structure):
    _fields_ = [("eos", C.c_int), ("gamma", C.c_double), ("K", C.c_double),
                ("rhoatm", C.c_double), ("fatm", C.c_double), ("epslatm", C.c_double),
                ("patm", C.c_double), ("vmax", C.c_double), ("errormax", C.c_double),
                ("countmax", C.c_int)]


class CheckpointStruct(C.Structure):
    _fields_ = [("lmin", C.c_int), ("lmax", C.c_int), ("levels", C.POINTER(C.c_void_p)), ("iteration", IP),
                ("time", DP), ("nparams", C.c_int), ("param_names", C.POINTER(C.c_char_p)),
                ("param_values", C.POINTER(C.c_char_p)), ("nvars", C.c_int), ("vars", IP),
                ("var_names", C.POINTER(C.c_char_p)), ("dt", DP)]


def _checkpoint_struct(levels, iterations, times, vars, names, params, lmin):
    n = len(levels)
    keep = dict(lv=(C.c_void_p * n)(*[lv.h for lv in levels]), it=(C.c_int * n)(*iterations),
                tm=(C.c_double * n)(*times), pn=(C.c_char_p * max(len(params), 1))(*[k.encode() for k in params]),
                pv=(C.c_char_p * max(len(params), 1))(*[str(v).encode() for v in params.values()]),
                vr=(C.c_int * max(len(vars), 1))(*vars), vn=(C.c_char_p * max(len(names), 1))(*[x.encode() for x in names]))
    S = CheckpointStruct(lmin, lmin + n - 1, C.cast(keep["lv"], C.POINTER(C.c_void_p)), C.cast(keep["it"], IP),
                         C.cast(keep["tm"], DP), len(params), keep["pn"], keep["pv"], len(vars),
                         C.cast(keep["vr"], IP), keep["vn"])
    return S, keep


def checkpoint_write(path, levels, iterations, times, vars, names, params=None, lmin=0):
    """checkpoint_write_local for device levels (the reference's file layout)"""
    L = lib()
    _sig(L, "bamhip_checkpoint_write", C.c_int, [C.c_char_p, C.POINTER(CheckpointStruct)])
    S, keep = _checkpoint_struct(levels, iterations, times, vars, names, params or {}, lmin)
    check(L.bamhip_checkpoint_write(str(path).encode(), C.byref(S)), "bamhip_checkpoint_write")


def checkpoint_read(path, levels, vars, names, lmin=0, dt=None):
    """checkpoint_read into device levels; returns (iterations, times) per
    level; with dt (per level) the times are iteration * dt"""
    L = lib()
    _sig(L, "bamhip_checkpoint_read", C.c_int, [C.c_char_p, C.POINTER(CheckpointStruct), IP, DP])
    n = len(levels)
    S, keep = _checkpoint_struct(levels, [0] * n, [0.0] * n, vars, names, {}, lmin)
    if dt is not None:
        keep["dt"] = (C.c_double * n)(*dt)
        S.dt = C.cast(keep["dt"], DP)
    it, tm = (C.c_int * n)(), (C.c_double * n)()
    check(L.bamhip_checkpoint_read(str(path).encode(), C.byref(S), it, tm), "bamhip_checkpoint_read")
    return list(it), list(tm)


class HrscParams(C.Structure):
    _fields_ = [("eos", C.c_int), ("gamma", C.c_double), ("K", C.c_double), ("vmax", C.c_double),
                ("wlormax", C.c_double), ("riemann", C.c_int), ("rec", C.c_int)]


def hrsc_params(eos="ideal", gamma=2.0, K=100.0, vmax=0.999, wlormax=1e5, riemann="LLF", rec="WENOZ"):
    """bamhip_hrsc_params: eos_*, grhd_vmax, grhd_Wlor_max, hrsc_flux, hrsc_rec"""
    return HrscParams(0 if eos == "ideal" else 1, gamma, K, vmax, wlormax, {"LLF": 0, "HLL": 1}[riemann],
                      {"WENOZ": 0, "WENO5": 1}[rec])


C2P_LEARNED_METHOD = 3
OUTPUT_0D = ["max", "min", "norm", "norminf", "integral"]


class Z4cGrhdStruct(C.Structure):
    _fields_ = [("z4", C.POINTER(Z4cParams)), ("c2p", C.POINTER(C2PParams)), ("hrsc", C.POINTER(HrscParams)),
                ("c2p_method", C.c_int), ("mlp", C.c_void_p), ("polish_max", C.c_int), ("fallbacks", C.c_void_p),
                ("pgh", C.c_int), ("adm_g", IP), ("adm_K", IP), ("prim", IP), ("madm", IP),
                ("status", C.c_void_p), ("iters", C.c_void_p), ("atm_prerhs", C.c_int)]


def c2p_params(eos="ideal", gamma=2.0, K=100.0, rhoatm=1e-20, fatm=100.0, vmax=0.999, errormax=1e-10,
               countmax=100):
    """EOS/atmosphere snapshot; the atmosphere state follows the cold EOS
    (grhd_atm.c:74-75)"""
    P = C2PParams()
    P.eos = 0 if eos == "ideal" else 1
    P.gamma, P.K = gamma, K
    P.rhoatm, P.fatm = rhoatm, fatm
    kr = K * rhoatm ** (gamma - 1.0)
    P.patm = rhoatm * kr
    P.epslatm = kr / (gamma - 1.0)
    P.vmax, P.errormax, P.countmax = vmax, errormax, countmax
    return P


class DeviceArray:
    """a device buffer owned by the library allocator"""

    def __init__(self, shape, dtype=np.float64):
        init()
        self.shape = tuple(shape) if np.iterable(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().bamhip_malloc(C.byref(p), self.nbytes), "bamhip_malloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        d.put(a)
        return d

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().bamhip_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes), "bamhip_memcpy_htod")

    def get(self):
        a = np.empty(self.shape, self.dtype)
        check(lib().bamhip_memcpy_dtoh(a.ctypes.data, self.ptr, self.nbytes), "bamhip_memcpy_dtoh")
        return a

    def free(self):
        if self.ptr:
            lib().bamhip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def c2p_batch(P, method, cons, metric, prim, status, iters):
    """device-array c2p: cons (5,n) in/out, metric (6,n), prim (8,n) in/out
    (prim[5] = pressure guess, prim[0] = density guess), status/iters (n,)"""
    n = cons.shape[1]
    L = lib()
    L.bamhip_c2p_batch.restype = C.c_int
    L.bamhip_c2p_batch.argtypes = [C.POINTER(C2PParams), C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
    check(L.bamhip_c2p_batch(C.byref(P), method, n, cons.ptr, metric.ptr, prim.ptr, status.ptr, iters.ptr),
          "bamhip_c2p_batch")


MLP_NPARAM = 4481


class Mlp:
    """the learned pressure guess (bamhip_mlp)"""

    def __init__(self, seed=0):
        init()
        L = lib()
        vp, i, l, f = C.c_void_p, C.c_int, C.c_long, C.c_float
        FP = C.POINTER(C.c_float)
        _sig(L, "bamhip_mlp_create", i, [C.POINTER(vp), C.c_ulonglong])
        _sig(L, "bamhip_mlp_free", i, [vp])
        _sig(L, "bamhip_mlp_get_params", i, [vp, FP, FP])
        _sig(L, "bamhip_mlp_set_params", i, [vp, FP, FP])
        _sig(L, "bamhip_mlp_grad", i, [vp, l, vp, vp, FP, FP])
        _sig(L, "bamhip_mlp_train", i, [vp, l, vp, vp, i, l, f, C.c_ulonglong, FP])
        _sig(L, "bamhip_mlp_predict", i, [vp, l, vp, vp])
        h = vp()
        check(L.bamhip_mlp_create(C.byref(h), seed), "bamhip_mlp_create")
        self.h = h

    def free(self):
        if self.h:
            lib().bamhip_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def get(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        nrm = np.zeros(8, np.float32)
        FP = C.POINTER(C.c_float)
        check(lib().bamhip_mlp_get_params(self.h, p.ctypes.data_as(FP), nrm.ctypes.data_as(FP)),
              "bamhip_mlp_get_params")
        return p, nrm

    def set(self, params=None, norm=None):
        FP = C.POINTER(C.c_float)
        pa = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        na = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        check(lib().bamhip_mlp_set_params(self.h, pa.ctypes.data_as(FP) if pa is not None else None,
                                          na.ctypes.data_as(FP) if na is not None else None),
              "bamhip_mlp_set_params")

    def grad(self, feat, target):
        """feat: DeviceArray (n,3) float32, target: DeviceArray (n,) float32"""
        n = target.shape[0]
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_float()
        check(lib().bamhip_mlp_grad(self.h, n, feat.ptr, target.ptr, g.ctypes.data_as(C.POINTER(C.c_float)),
                                    C.byref(loss)), "bamhip_mlp_grad")
        return g, loss.value

    def train(self, feat, target, epochs, batch=16384, lr=2e-3, seed=1):
        n = target.shape[0]
        el = np.zeros(epochs, np.float32)
        check(lib().bamhip_mlp_train(self.h, n, feat.ptr, target.ptr, epochs, batch, lr, seed,
                                     el.ctypes.data_as(C.POINTER(C.c_float))), "bamhip_mlp_train")
        return el

    def predict(self, feat, y):
        check(lib().bamhip_mlp_predict(self.h, y.shape[0], feat.ptr, y.ptr), "bamhip_mlp_predict")


def mlp_prepare(P, cons, metric, prim, status, feat, target):
    """raw features (n,3) and targets (n,) of solved points (device arrays)"""
    n = cons.shape[1]
    L = lib()
    vp = C.c_void_p
    _sig(L, "bamhip_mlp_prepare", C.c_int, [C.POINTER(C2PParams), C.c_long, vp, vp, vp, vp, vp, vp])
    check(L.bamhip_mlp_prepare(C.byref(P), n, cons.ptr, metric.ptr, prim.ptr, status.ptr, feat.ptr, target.ptr),
          "bamhip_mlp_prepare")


def c2p_learned_batch(P, mlp, cons, metric, prim, status, iters, fallbacks=None, polish_max=8):
    """learned-guess c2p on device arrays; fallbacks: DeviceArray (1,) int32 accumulating the count"""
    n = cons.shape[1]
    L = lib()
    vp = C.c_void_p
    _sig(L, "bamhip_c2p_learned_batch", C.c_int,
         [C.POINTER(C2PParams), vp, C.c_int, C.c_long, vp, vp, vp, vp, vp, vp])
    check(L.bamhip_c2p_learned_batch(C.byref(P), mlp.h, polish_max, n, cons.ptr, metric.ptr, prim.ptr, status.ptr,
                                     iters.ptr, fallbacks.ptr if fallbacks is not None else None),
          "bamhip_c2p_learned_batch")


def p2c_batch(P, prim5, metric, cons):
    """device-array prim -> cons: prim5 (5,n), metric (6,n), cons (5,n) out"""
    n = cons.shape[1]
    L = lib()
    L.bamhip_p2c_batch.restype = C.c_int
    L.bamhip_p2c_batch.argtypes = [C.POINTER(C2PParams), C.c_long, C.c_void_p, C.c_void_p, C.c_void_p]
    check(L.bamhip_p2c_batch(C.byref(P), n, prim5.ptr, metric.ptr, cons.ptr), "bamhip_p2c_batch")


# ------------------------------------------------------------------ timers
def timer_enable(on=True):
    check(lib().bamhip_timer_enable(1 if on else 0), "bamhip_timer_enable")


def timer_reset():
    check(lib().bamhip_timer_reset(), "bamhip_timer_reset")


def timer_get(name):
    ms = C.c_double()
    n = C.c_long()
    check(lib().bamhip_timer_get(name.encode(), C.byref(ms), C.byref(n)), "bamhip_timer_get")
    return ms.value, n.value


# ------------------------------------------------------- RHS kernel tuner
BULK, SHELL = 0, 1


def tune_mode(mode):
    """0: default tile kernels; 1: time every variant on each new box set"""
    rc = lib().bamhip_z4c_tune_mode(mode)
    check(min(rc, 0), "bamhip_z4c_tune_mode")
    return rc


def tune_clear():
    check(lib().bamhip_z4c_tune_clear(), "bamhip_z4c_tune_clear")


def set_rk_fusion(mode=2):
    """RK4 stage update inside the RHS kernels: 0 never, 1 always, 2 on
    levels of at least 2^19 points (default)"""
    return lib().bamhip_z4c_set_rk_fusion(int(mode))


def set_xcd_map(on=True):
    """XCD-aware tile numbering in the RHS tile kernels"""
    return lib().bamhip_z4c_set_xcd_map(1 if on else 0)


def set_overlap(on=True):
    """run the box-edge shell kernels concurrently with the bulk kernel"""
    return lib().bamhip_z4c_set_overlap(1 if on else 0)


def variants(kind):
    n = lib().bamhip_z4c_variant_count(kind)
    check(min(n, 0), "bamhip_z4c_variant_count")
    return [lib().bamhip_z4c_variant_name(kind, v).decode() for v in range(n)]


def force_variant(kind, idx):
    """pin variant idx of kind (BULK, SHELL); -1 hands the choice back to the tuner"""
    check(lib().bamhip_z4c_force_variant(kind, idx), "bamhip_z4c_force_variant")


def tune_report():
    n = lib().bamhip_z4c_tune_report(None, 0)
    buf = C.create_string_buffer(n)
    lib().bamhip_z4c_tune_report(buf, n)
    return buf.value.decode()


def gpu_available():
    """True when a HIP device is visible (does not initialise the library)"""
    try:
        return lib().bamhip_init(0) == 0
    except Exception:
        return False
