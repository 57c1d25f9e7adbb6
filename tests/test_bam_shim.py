"""The BAM-shaped adapters (include/bamhip_bam.h) on the CPU: the stand-in
structures have the member offsets of BAM's own tB, tL, tG and tVarList
(src/main/amr/bam_amr.h, compiled here when the reference sources are
present), the adapter table and set_rhs_func_register installation work,
and an adapter called on a level without device mirror reports through the
errorexit hook without touching a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

import baum_amd as B
import baum_amd.bam as BB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_HDR = "/root/reference/src/main/amr/bam_amr.h"
STRUCTS = {"tB": BB.tB, "tL": BB.tL, "tG": BB.tG, "tVarList": BB.tVarList}


def _offsets_from_c(tmp_path, include_line, tag):
    """compile a probe that prints sizeof and offsetof of every member the
    ctypes mirror names, against the given header"""
    src = ["#include <stdio.h>", "#include <stddef.h>", include_line, "int main(void) {"]
    for name, cls in STRUCTS.items():
        src.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (name, name))
        for f in cls._fields_:
            src.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (name, f[0], name, f[0]))
    src.append("return 0; }")
    c = tmp_path / ("probe_%s.c" % tag)
    c.write_text("\n".join(src))
    exe = tmp_path / ("probe_%s" % tag)
    subprocess.run(["gcc", "-std=gnu99", "-o", str(exe), str(c)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict((tuple(l.split()[:2]), int(l.split()[2])) for l in out.splitlines())


def _offsets_ctypes():
    d = {}
    for name, cls in STRUCTS.items():
        d[(name, "sizeof")] = C.sizeof(cls)
        for f in cls._fields_:
            d[(name, f[0])] = getattr(cls, f[0]).offset
    return d


def test_mirror_header_matches_ctypes(tmp_path):
    got = _offsets_from_c(tmp_path, '#include "%s/include/bamhip_bam.h"' % ROOT, "mirror")
    assert got == _offsets_ctypes()


@pytest.mark.skipif(not os.path.exists(REF_HDR), reason="reference sources not present")
def test_mirror_layout_matches_reference_header(tmp_path):
    # bam_amr.h declares only structures, macros and an extern; it needs no
    # other BAM header for offsetof
    got = _offsets_from_c(tmp_path, '#include "%s"' % REF_HDR, "ref")
    assert got == _offsets_ctypes()


def test_adapters_are_exported_and_registered():
    hooks = BB.hooks()
    names = []
    addr = {}

    @BB.REG_FN
    def register(name, f):
        names.append(name.decode())
        addr[name.decode()] = C.cast(f, C.c_void_p).value

    BB._bind().bamhip_bam_register(register)
    assert names == ["geometry", "geometry_adm", "geometry_source", "geometry_bound"]
    L = B.lib()
    for hook, sym in [("rhs_geometry", "z4_rhs_hip"), ("rhs_boundary", "set_boundary_hip"),
                      ("rhs_geometry_adm", "z4_to_adm_hip"), ("rhs_geometry_source", "z4_rhs_add_source_hip"),
                      ("c2p", "grhd_c2p_hip"), ("restrict_prolong_varlist", "restrict_prolong_varlist_hip"),
                      ("evolve_rk4", "evolve_rk_rk4_hip")]:
        assert C.cast(getattr(hooks, hook), C.c_void_p).value == C.cast(getattr(L, sym), C.c_void_p).value
    assert addr["geometry"] == C.cast(L.z4_rhs_hip, C.c_void_p).value
    assert addr["geometry_bound"] == C.cast(L.set_boundary_hip, C.c_void_p).value


def test_header_declares_every_adapter():
    text = open(os.path.join(ROOT, "include", "bamhip_bam.h")).read()
    decl = set(re.findall(r"^void ([a-z0-9_]+_hip)\(", text, re.M))
    assert decl == {"z4_rhs_hip", "set_boundary_hip", "z4_to_adm_hip", "z4_rhs_add_source_hip", "grhd_c2p_hip",
                    "restrict_prolong_varlist_hip", "evolve_rk_rk4_hip"}
    L = B.lib()
    assert all(hasattr(L, n) for n in decl)


def test_errorexit_hook_on_unattached_level():
    err = BB.Errors()
    tl = BB.tL()
    tl.l = 3
    idx = (C.c_int * 25)(*range(25))
    vl = BB.tVarList(C.pointer(tl), 25, C.cast(idx, BB.IP))
    BB.hooks().rhs_geometry(C.byref(vl), None, 0.0, C.byref(vl))
    assert len(err.messages) == 1 and "level 3 has no device mirror" in err.messages[0]
    short = BB.tVarList(C.pointer(tl), 24, C.cast(idx, BB.IP))
    BB.hooks().rhs_boundary(C.byref(short), None, 0.0, C.byref(short))
    assert "24 variables" in err.messages[1]
    BB.hooks().evolve_rk4(C.pointer(tl), 0)
    assert len(err.messages) == 3
