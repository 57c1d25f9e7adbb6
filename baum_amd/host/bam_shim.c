/* bam_shim.c -- BAM-shaped adapters (void f(tVarList *, tVarList *, double,
 * tVarList *) and friends) over the bamhip C ABI.
 *
 * Each adapter reads what the reference function reads from its arguments
 * (the level through vl->level, the data through vl->index) and forwards to
 * the device mirror attached to that level.  Parameters the reference reads
 * from its database on every call are snapshots set once (bamhip_bam_set_*).
 * References: src/utility/evolve/evolve.c:22-29,185-215 (hook types and
 * registration), src/projects/z4/z4.c:93-124,245-248, z4_rhs_movpunc_N.c:23,
 * src/projects/matter/matter.c:309-321,400-420 (c2p call sites),
 * src/projects/grhd/grhd_c2p_v2.c:265-300 (c2p argument use),
 * src/main/amr/inject.c:66-110 (R/P), src/utility/evolve/rk.c:216-257. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/bamhip_bam.h"

#define SHIM_MAX_LEVELS 64

typedef struct {
  const tL *level;
  bamhip_level *dev;
  int *status, *iters; /* c2p outputs per point, allocated on first use */
  long npoints;
} shim_level;

static shim_level g_lev[SHIM_MAX_LEVELS];

static struct {
  int have_z4c;
  bamhip_z4c_params z4c;
  bamhip_dissipation_rule rule;
  int pgh;
  int have_grhd;
  bamhip_c2p_params c2p;
  int c2p_method, indx_q, indx_gxx;
  int have_adm;
  int adm_g, adm_K, adm_rho, adm_S, adm_SS, adm_ST;
  int order_rp;
  tVarList *uc, *up, *r, *v;
  int algcon;
} g_cfg; /* zero until the set-up calls; order_rp 0 means 4 (the default of order_RP) */

static void default_errorexit(const char *file, int line, const char *msg) {
  fprintf(stderr, "Error in %s, line %d:\n  %s\n", file, line, msg);
  exit(1);
}
static void (*g_errorexit)(const char *, int, const char *) = default_errorexit;

void bamhip_bam_set_errorexit(void (*hook)(const char *file, int line, const char *msg)) {
  g_errorexit = hook ? hook : default_errorexit;
}

/* report a failure through the errorexit hook, with the library's message */
static void fail_at(const char *file, int line, const char *what, int code) {
  char msg[512];
  if (code)
    snprintf(msg, sizeof msg, "%s: %s", what, bamhip_last_error());
  else
    snprintf(msg, sizeof msg, "%s", what);
  g_errorexit(file, line, msg);
}
#define FAIL(what, code) fail_at(__FILE__, __LINE__, (what), (code))

static shim_level *find(const tL *level) {
  if (!level) return NULL;
  for (int k = 0; k < SHIM_MAX_LEVELS; k++)
    if (g_lev[k].level == level && g_lev[k].dev) return &g_lev[k];
  return NULL;
}

static void release(shim_level *s) {
  if (s->status) bamhip_free(s->status);
  if (s->iters) bamhip_free(s->iters);
  memset(s, 0, sizeof *s);
}

int bamhip_bam_attach(tL *level, bamhip_level *dev) {
  if (!level || !dev) return BAMHIP_EINVAL;
  shim_level *s = find(level);
  if (!s) {
    for (int k = 0; k < SHIM_MAX_LEVELS && !s; k++)
      if (!g_lev[k].dev) s = &g_lev[k];
    if (!s) return BAMHIP_ENOMEM;
  } else {
    release(s);
  }
  long np = 0;
  int e = bamhip_level_npoints(dev, &np);
  if (e) return e;
  s->level = level;
  s->dev = dev;
  s->npoints = np;
  return BAMHIP_OK;
}

int bamhip_bam_detach(tL *level) {
  shim_level *s = find(level);
  if (!s) return BAMHIP_EINVAL;
  release(s);
  return BAMHIP_OK;
}

bamhip_level *bamhip_bam_device(const tL *level) {
  shim_level *s = find(level);
  return s ? s->dev : NULL;
}

int bamhip_bam_set_z4c(const bamhip_z4c_params *par, const bamhip_dissipation_rule *rule, int pgh) {
  if (!par || !rule || pgh < 0) return BAMHIP_EINVAL;
  g_cfg.z4c = *par;
  g_cfg.rule = *rule;
  g_cfg.pgh = pgh;
  g_cfg.have_z4c = 1;
  return BAMHIP_OK;
}

int bamhip_bam_set_grhd(const bamhip_c2p_params *par, int method, int indx_var_q, int indx_var_gxx) {
  if (!par || indx_var_q < 0 || indx_var_gxx < 0) return BAMHIP_EINVAL;
  if (method != BAMHIP_C2P_HYBRID && method != BAMHIP_C2P_PLAIN && method != BAMHIP_C2P_BRACKETED &&
      method != BAMHIP_C2P_VANAL)
    return BAMHIP_EINVAL;
  g_cfg.c2p = *par;
  g_cfg.c2p_method = method;
  g_cfg.indx_q = indx_var_q;
  g_cfg.indx_gxx = indx_var_gxx;
  g_cfg.have_grhd = 1;
  return BAMHIP_OK;
}

int bamhip_bam_set_adm(int adm_gxx, int adm_Kxx, int adm_rho, int adm_Sx, int adm_SSxx, int adm_ST) {
  if (adm_gxx < 0 || adm_Kxx < 0 || adm_rho < 0 || adm_Sx < 0 || adm_SSxx < 0 || adm_ST < 0)
    return BAMHIP_EINVAL;
  g_cfg.adm_g = adm_gxx;
  g_cfg.adm_K = adm_Kxx;
  g_cfg.adm_rho = adm_rho;
  g_cfg.adm_S = adm_Sx;
  g_cfg.adm_SS = adm_SSxx;
  g_cfg.adm_ST = adm_ST;
  g_cfg.have_adm = 1;
  return BAMHIP_OK;
}

int bamhip_bam_set_fmr(int order_rp) {
  if (order_rp != 2 && order_rp != 4 && order_rp != 6 && order_rp != 8) return BAMHIP_EINVAL;
  g_cfg.order_rp = order_rp;
  return BAMHIP_OK;
}

int bamhip_bam_set_rk_lists(tVarList *uc, tVarList *up, tVarList *r, tVarList *v, int algcon) {
  if (!uc || !up || !r || !v) return BAMHIP_EINVAL;
  g_cfg.uc = uc;
  g_cfg.up = up;
  g_cfg.r = r;
  g_cfg.v = v;
  g_cfg.algcon = algcon != 0;
  return BAMHIP_OK;
}

int bamhip_bam_put(tVarList *vl) {
  if (!vl || !vl->level || !vl->level->v) return BAMHIP_EINVAL;
  shim_level *s = find(vl->level);
  if (!s) return BAMHIP_EINVAL;
  for (int i = 0; i < vl->n; i++) {
    const int var = vl->index[i];
    if (var < 0 || var >= vl->level->nvariables || !vl->level->v[var]) return BAMHIP_EINVAL;
    int e = bamhip_level_enable(s->dev, var);
    if (!e) e = bamhip_level_put(s->dev, var, vl->level->v[var]);
    if (e) return e;
  }
  return BAMHIP_OK;
}

int bamhip_bam_get(tVarList *vl) {
  if (!vl || !vl->level || !vl->level->v) return BAMHIP_EINVAL;
  shim_level *s = find(vl->level);
  if (!s) return BAMHIP_EINVAL;
  for (int i = 0; i < vl->n; i++) {
    const int var = vl->index[i];
    if (var < 0 || var >= vl->level->nvariables || !vl->level->v[var]) return BAMHIP_EINVAL;
    int e = bamhip_level_get(s->dev, var, vl->level->v[var]);
    if (e) return e;
  }
  return BAMHIP_OK;
}

/* the Z4c snapshot with get_dissipation_factor(level) */
static int z4c_params_for(const tL *level, bamhip_z4c_params *par) {
  if (!g_cfg.have_z4c) {
    FAIL("bamhip_bam: Z4c parameters not set (bamhip_bam_set_z4c)", 0);
    return -1;
  }
  *par = g_cfg.z4c;
  int e = bamhip_dissipation_factor(&g_cfg.rule, level->l, level->shells, &par->dissfactor);
  if (e) {
    FAIL("get_dissipation_factor", e);
    return -1;
  }
  return 0;
}

/* the device mirror behind a Z4c list (25 variables) */
static shim_level *z4c_level(tVarList *vl, const char *who) {
  char msg[160];
  if (!vl || !vl->level) {
    snprintf(msg, sizeof msg, "%s: variable list without level", who);
    FAIL(msg, 0);
    return NULL;
  }
  if (vl->n != BAMHIP_Z4C_NVARS) {
    snprintf(msg, sizeof msg, "%s: the Z4c list has %d variables, expected %d", who, vl->n, BAMHIP_Z4C_NVARS);
    FAIL(msg, 0);
    return NULL;
  }
  shim_level *s = find(vl->level);
  if (!s) {
    snprintf(msg, sizeof msg, "%s: level %d has no device mirror (bamhip_bam_attach)", who, vl->level->l);
    FAIL(msg, 0);
  }
  return s;
}

void z4_rhs_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur) {
  shim_level *s = z4c_level(ucur, "z4_rhs_hip");
  bamhip_z4c_params par;
  if (!s || z4c_params_for(ucur->level, &par)) return;
  if (!unew || unew->n != BAMHIP_Z4C_NVARS || (c != 0.0 && (!upre || upre->n != BAMHIP_Z4C_NVARS))) {
    FAIL("z4_rhs_hip: unew/upre must be Z4c lists", 0);
    return;
  }
  const int e = bamhip_z4c_rhs(s->dev, &par, unew->index, upre ? upre->index : NULL, c, ucur->index);
  if (e) FAIL("z4_rhs_hip", e);
}

void set_boundary_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur) {
  shim_level *s = z4c_level(ucur, "set_boundary_hip");
  bamhip_z4c_params par;
  if (!s || z4c_params_for(ucur->level, &par)) return;
  if (!unew || unew->n != BAMHIP_Z4C_NVARS || (c != 0.0 && (!upre || upre->n != BAMHIP_Z4C_NVARS))) {
    FAIL("set_boundary_hip: unew/upre must be Z4c lists", 0);
    return;
  }
  const int e = bamhip_z4c_set_boundary(s->dev, &par, unew->index, upre ? upre->index : NULL, c, ucur->index,
                                        g_cfg.pgh);
  if (e) FAIL("set_boundary_hip", e);
}

static void seq(int *idx, int first, int n) {
  for (int k = 0; k < n; k++) idx[k] = first + k;
}

void z4_to_adm_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur) {
  shim_level *s = z4c_level(ucur, "z4_to_adm_hip");
  if (!s) return;
  if (!g_cfg.have_z4c || !g_cfg.have_adm) {
    FAIL("z4_to_adm_hip: parameters not set (bamhip_bam_set_z4c, bamhip_bam_set_adm)", 0);
    return;
  }
  int g[6], K[6];
  seq(g, g_cfg.adm_g, 6);
  seq(K, g_cfg.adm_K, 6);
  const int e = bamhip_z4c_to_adm(s->dev, ucur->index, g_cfg.z4c.chipsipower, g, K);
  if (e) FAIL("z4_to_adm_hip", e);
}

void z4_rhs_add_source_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur) {
  shim_level *s = z4c_level(ucur, "z4_rhs_add_source_hip");
  bamhip_z4c_params par;
  if (!s || z4c_params_for(ucur->level, &par)) return;
  if (!g_cfg.have_adm) {
    FAIL("z4_rhs_add_source_hip: ADM indices not set (bamhip_bam_set_adm)", 0);
    return;
  }
  if (!unew || unew->n != BAMHIP_Z4C_NVARS) {
    FAIL("z4_rhs_add_source_hip: unew must be a Z4c list", 0);
    return;
  }
  int madm[11];
  madm[0] = g_cfg.adm_rho;
  seq(madm + 1, g_cfg.adm_S, 3);
  seq(madm + 4, g_cfg.adm_SS, 6);
  madm[10] = g_cfg.adm_ST;
  const int e = bamhip_z4c_matter_source(s->dev, &par, unew->index, c, ucur->index, madm);
  if (e) FAIL("z4_rhs_add_source_hip", e);
}

void grhd_c2p_hip(tVarList *ucur, tVarList *primVars) {
  if (!ucur || !ucur->level || !primVars) {
    FAIL("grhd_c2p_hip: NULL variable list", 0);
    return;
  }
  if (!g_cfg.have_grhd) {
    FAIL("grhd_c2p_hip: parameters not set (bamhip_bam_set_grhd)", 0);
    return;
  }
  if (g_cfg.indx_q + 5 > ucur->n || primVars->n < 8) {
    FAIL("grhd_c2p_hip: ucur lacks the 5 conserved variables or primVars has fewer than 8", 0);
    return;
  }
  shim_level *s = find(ucur->level);
  if (!s) {
    FAIL("grhd_c2p_hip: level has no device mirror (bamhip_bam_attach)", 0);
    return;
  }
  if (!s->status) {
    int e = bamhip_malloc((void **)&s->status, sizeof(int) * (size_t)(s->npoints ? s->npoints : 1));
    if (!e) e = bamhip_malloc((void **)&s->iters, sizeof(int) * (size_t)(s->npoints ? s->npoints : 1));
    if (e) {
      FAIL("grhd_c2p_hip: status buffers", e);
      return;
    }
  }
  int metric[6];
  seq(metric, g_cfg.indx_gxx, 6);
  /* vldataptr(ucur, INDX_VAR_q + k) = level->v[ucur->index[INDX_VAR_q + k]] */
  const int *cons = ucur->index + g_cfg.indx_q;
  int e = bamhip_grhd_c2p(s->dev, &g_cfg.c2p, g_cfg.c2p_method, cons, metric, primVars->index, s->status, s->iters);
  if (e) {
    FAIL("grhd_c2p_hip", e);
    return;
  }
  if ((g_cfg.c2p_method == BAMHIP_C2P_PLAIN || g_cfg.c2p_method == BAMHIP_C2P_VANAL) && s->npoints > 0) {
    /* grhd_c2p_proot and grhd_c2p_vanal stop at a point they cannot invert
     * (errorexit at grhd_c2p.c:178-222, 287-363) */
    int *st = malloc(sizeof(int) * (size_t)s->npoints);
    if (!st) {
      FAIL("grhd_c2p_hip: host buffer", 0);
      return;
    }
    e = bamhip_memcpy_dtoh(st, s->status, sizeof(int) * (size_t)s->npoints);
    long bad = -1;
    for (long k = 0; !e && k < s->npoints && bad < 0; k++)
      if (st[k] != BAMHIP_C2P_OK) bad = k;
    free(st);
    if (e) {
      FAIL("grhd_c2p_hip: status copy", e);
    } else if (bad >= 0) {
      char msg[128];
      snprintf(msg, sizeof msg, "grhd_c2p_hip: con2prim failed at point %ld of level %d", bad, ucur->level->l);
      FAIL(msg, 0);
    }
  }
}

void restrict_prolong_varlist_hip(tG *g, int lcoarse, int lfine, tVarList *uc, tVarList *uf, int nbuffer) {
  if (!g || !g->level || lcoarse < 0 || lfine < 0 || lcoarse >= g->nlevels || lfine >= g->nlevels || !uc ||
      !uf || uc->n != uf->n) {
    FAIL("restrict_prolong_varlist_hip: bad arguments", 0);
    return;
  }
  shim_level *c = find(g->level[lcoarse]), *f = find(g->level[lfine]);
  if (!c || !f) {
    FAIL("restrict_prolong_varlist_hip: level without device mirror (bamhip_bam_attach)", 0);
    return;
  }
  if (g->level[lcoarse]->shells) {
    FAIL("restrict_prolong_varlist_hip: shell levels are not supported", 0);
    return;
  }
  const int e =
      bamhip_restrict_prolong_varlist(c->dev, f->dev, uc->n, uc->index, uf->index, nbuffer,
                                                  g_cfg.order_rp ? g_cfg.order_rp : 4);
  if (e) FAIL("restrict_prolong_varlist_hip", e);
}

void evolve_rk_rk4_hip(tL *level, int comp) {
  (void)comp;
  if (!g_cfg.uc) {
    FAIL("evolve_rk_rk4_hip: RK lists not set (bamhip_bam_set_rk_lists)", 0);
    return;
  }
  /* evolve_rk points its lists at the level being stepped (rk.c:308-310) */
  g_cfg.uc->level = g_cfg.up->level = g_cfg.r->level = g_cfg.v->level = level;
  shim_level *s = z4c_level(g_cfg.uc, "evolve_rk_rk4_hip");
  bamhip_z4c_params par;
  if (!s || z4c_params_for(level, &par)) return;
  const int e = bamhip_z4c_rk4_step(s->dev, &par, level->dt, g_cfg.uc->index, g_cfg.up->index, g_cfg.r->index,
                                    g_cfg.v->index, g_cfg.pgh, g_cfg.algcon);
  if (e) FAIL("evolve_rk_rk4_hip", e);
}

/* The adapters in hooks of the reference's types: a mismatch of any
 * signature is a compile error here. */
const bamhip_bam_hooks bamhip_bam = {
    z4_rhs_hip, set_boundary_hip, z4_to_adm_hip, z4_rhs_add_source_hip,
    grhd_c2p_hip, restrict_prolong_varlist_hip, evolve_rk_rk4_hip,
};

void bamhip_bam_register(void (*set_rhs_func_register)(char *name, bamhip_bam_rhs_fn f)) {
  if (!set_rhs_func_register) return;
  set_rhs_func_register("geometry", z4_rhs_hip);
  set_rhs_func_register("geometry_adm", z4_to_adm_hip);
  set_rhs_func_register("geometry_source", z4_rhs_add_source_hip);
  set_rhs_func_register("geometry_bound", set_boundary_hip);
}
