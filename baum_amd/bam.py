"""ctypes view of the BAM-shaped adapters (include/bamhip_bam.h).

BamLevel builds a BAM level (tL, src/main/amr/bam_amr.h:42-70) whose tL::v
holds host numpy arrays, as BAM's storage does, and VarList a tVarList
(bam_amr.h:127-131) over it.  hooks() returns the adapters as the
function-pointer table `bamhip_bam` the library exports, typed as the
reference's hooks (evolve.c:22-29, bam_matter.h:49, inject.c:66, rk.c:216),
so a caller invokes them exactly as BAM's evolve_rhs, matter_c2p, main and
evolve_rk do.  No compute here: every call goes into libbamhip.so.
"""
import ctypes as C

import numpy as np

from . import BamhipError, C2PParams, DissipationRule, Z4cParams, check, lib

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class tB(C.Structure):
    pass


class tL(C.Structure):
    pass


class tG(C.Structure):
    pass


tB._fields_ = [("pr", C.POINTER(tB)), ("level", C.POINTER(tL)), ("l", C.c_int), ("i", C.c_int),
               ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dd", C.c_double * 3),
               ("di", C.c_int), ("dj", C.c_int), ("dk", C.c_int), ("id", C.c_int * 3),
               ("bbox", C.c_double * 6), ("ibbox", C.c_int * 6), ("com", C.c_void_p),
               ("npoints", C.c_int), ("noffset", C.c_int), ("x0", C.c_double), ("y0", C.c_double),
               ("z0", C.c_double), ("m", C.c_int), ("n", C.c_int), ("o", C.c_int), ("imax", C.c_int),
               ("jmax", C.c_int), ("kmax", C.c_int), ("bflag", C.c_int * 6)]
tL._fields_ = [("grid", C.POINTER(tG)), ("prlocal", C.POINTER(tL)), ("l", C.c_int), ("shells", C.c_int),
               ("nboxes", C.c_int), ("box", C.POINTER(C.POINTER(tB))), ("npoints", C.c_int),
               ("nnodes", C.c_int), ("boundary", IP), ("nvariables", C.c_int), ("v", C.POINTER(DP)),
               ("iteration", C.c_int), ("time", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
               ("dz", C.c_double), ("dt", C.c_double), ("bbox", C.c_double * 6), ("ibbox", C.c_int * 6),
               ("sbox", C.c_void_p), ("nsboxes", C.c_int), ("com", C.c_void_p)]
tG._fields_ = [("level", C.POINTER(C.POINTER(tL))), ("nlevels", C.c_int), ("lmin", C.c_int), ("lmax", C.c_int),
               ("ltop", C.c_int), ("lbo", C.c_int), ("nvariables", C.c_int), ("half", C.c_int * 3),
               ("symmetric", C.c_int * 3), ("full", C.c_int), ("bitant", C.c_int), ("rotant", C.c_int),
               ("quadrant", C.c_int), ("octant", C.c_int), ("qreflect", C.c_int), ("npunc", C.c_int),
               ("lmaxpunc", IP), ("puncpos", C.POINTER(DP))]


class tVarList(C.Structure):
    _fields_ = [("level", C.POINTER(tL)), ("n", C.c_int), ("index", IP)]


RHS_FN = C.CFUNCTYPE(None, C.POINTER(tVarList), C.POINTER(tVarList), C.c_double, C.POINTER(tVarList))
C2P_FN = C.CFUNCTYPE(None, C.POINTER(tVarList), C.POINTER(tVarList))
RP_FN = C.CFUNCTYPE(None, C.POINTER(tG), C.c_int, C.c_int, C.POINTER(tVarList), C.POINTER(tVarList), C.c_int)
STEP_FN = C.CFUNCTYPE(None, C.POINTER(tL), C.c_int)
ERR_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_int, C.c_char_p)
REG_FN = C.CFUNCTYPE(None, C.c_char_p, RHS_FN)


class Hooks(C.Structure):
    _fields_ = [("rhs_geometry", RHS_FN), ("rhs_boundary", RHS_FN), ("rhs_geometry_adm", RHS_FN),
                ("rhs_geometry_source", RHS_FN), ("c2p", C2P_FN), ("restrict_prolong_varlist", RP_FN),
                ("evolve_rk4", STEP_FN)]


_BOUND = False


def _bind():
    global _BOUND
    L = lib()
    if _BOUND:
        return L
    i, vp = C.c_int, C.c_void_p
    L.bamhip_bam_set_errorexit.restype = None
    L.bamhip_bam_set_errorexit.argtypes = [ERR_FN]
    for name, args in [("bamhip_bam_attach", [C.POINTER(tL), vp]), ("bamhip_bam_detach", [C.POINTER(tL)]),
                       ("bamhip_bam_set_fmr", [i]),
                       ("bamhip_bam_set_adm", [i, i, i, i, i, i]),
                       ("bamhip_bam_set_rk_lists", [C.POINTER(tVarList)] * 4 + [i]),
                       ("bamhip_bam_put", [C.POINTER(tVarList)]), ("bamhip_bam_get", [C.POINTER(tVarList)])]:
        f = getattr(L, name)
        f.restype = i
        f.argtypes = args
    L.bamhip_bam_device.restype = vp
    L.bamhip_bam_device.argtypes = [C.POINTER(tL)]
    L.bamhip_bam_set_z4c.restype = i
    L.bamhip_bam_set_z4c.argtypes = [C.POINTER(Z4cParams), C.POINTER(DissipationRule), i]
    L.bamhip_bam_set_grhd.restype = i
    L.bamhip_bam_set_grhd.argtypes = [C.POINTER(C2PParams), i, i, i]
    L.bamhip_bam_register.restype = None
    L.bamhip_bam_register.argtypes = [REG_FN]
    _BOUND = True
    return L


def set_z4c(params, rule, pgh=0):
    """params: baum_amd.z4c_params(); rule: baum_amd.DissipationRule"""
    check(_bind().bamhip_bam_set_z4c(C.byref(params), C.byref(rule), pgh), "bamhip_bam_set_z4c")


def set_grhd(c2p, method, indx_var_q, indx_var_gxx):
    check(_bind().bamhip_bam_set_grhd(C.byref(c2p), method, indx_var_q, indx_var_gxx), "bamhip_bam_set_grhd")


def set_fmr(order_rp):
    check(_bind().bamhip_bam_set_fmr(order_rp), "bamhip_bam_set_fmr")


def set_rk_lists(uc, up, r, v, algcon=True):
    check(_bind().bamhip_bam_set_rk_lists(uc.ref(), up.ref(), r.ref(), v.ref(), 1 if algcon else 0),
          "bamhip_bam_set_rk_lists")


def hooks():
    """the exported adapter table, typed as the reference's function pointers"""
    return Hooks.in_dll(_bind(), "bamhip_bam")


class Errors:
    """errorexit hook that records instead of stopping the process"""

    def __init__(self):
        self.messages = []
        self._cb = ERR_FN(self._hook)
        _bind().bamhip_bam_set_errorexit(self._cb)

    def _hook(self, file, line, msg):
        self.messages.append(msg.decode())

    def raise_if_any(self):
        if self.messages:
            raise BamhipError("; ".join(self.messages))


class BamLevel:
    """a BAM level with host storage: tL::v[index] -> float64 array of
    npoints (allocated for the indices given), plus the device mirror
    attached to it.  boxes: bamhip Box records of the mirror (same point
    order); dev: a baum_amd.Level with the same boxes and variable slots."""

    def __init__(self, dev, l=0, dt=0.0, shells=0):
        self.dev = dev
        self.nvariables = dev.nvariables
        self.npoints = dev.npoints
        self.arrays = {}
        self._vp = (DP * self.nvariables)()
        self.tl = tL()
        self.tl.l = l
        self.tl.shells = shells
        self.tl.npoints = self.tl.nnodes = self.npoints
        self.tl.nvariables = self.nvariables
        self.tl.v = C.cast(self._vp, C.POINTER(DP))
        self.tl.dt = dt
        self.tl.dx, self.tl.dy, self.tl.dz = dev.boxes[0].dx, dev.boxes[0].dy, dev.boxes[0].dz
        check(_bind().bamhip_bam_attach(C.byref(self.tl), dev.h), "bamhip_bam_attach")

    def alloc(self, indices):
        for v in indices:
            if v not in self.arrays:
                a = np.zeros(self.npoints)
                self.arrays[v] = a
                self._vp[v] = a.ctypes.data_as(DP)

    def set(self, v, data):
        self.alloc([v])
        self.arrays[v][:] = np.asarray(data, dtype=np.float64).reshape(-1)

    def varlist(self, indices):
        return VarList(self, indices)

    def detach(self):
        _bind().bamhip_bam_detach(C.byref(self.tl))


class VarList:
    """tVarList over a BamLevel (its host arrays are allocated on creation)"""

    def __init__(self, level, indices):
        self.level = level
        level.alloc(indices)
        self.indices = list(indices)
        self._idx = (C.c_int * len(self.indices))(*self.indices)
        self.vl = tVarList(C.pointer(level.tl), len(self.indices), C.cast(self._idx, IP))

    def ref(self):
        return C.byref(self.vl)

    def put(self):
        check(_bind().bamhip_bam_put(self.ref()), "bamhip_bam_put")

    def get(self):
        check(_bind().bamhip_bam_get(self.ref()), "bamhip_bam_get")
        return np.stack([self.level.arrays[v] for v in self.indices])


def grid(levels, lmin=0):
    """a tG over BamLevels (index = level number - lmin)"""
    g = tG()
    arr = (C.POINTER(tL) * len(levels))(*[C.pointer(L.tl) for L in levels])
    g.level = C.cast(arr, C.POINTER(C.POINTER(tL)))
    g.nlevels = len(levels)
    g.lmin = lmin
    g.lmax = lmin + len(levels) - 1
    g._keep = (arr, levels)
    return g
