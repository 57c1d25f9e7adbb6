/* bamhip_bam.h -- BAM-shaped adapters over the bamhip C ABI.
 *
 * BAM calls its evolution hot path through function pointers of fixed
 * types: the geometry right-hand side and its boundary
 * (void (*)(tVarList *, tVarList *, double, tVarList *),
 * src/utility/evolve/evolve.c:22-29, registered with set_rhs_func_register
 * at src/projects/z4/z4.c:245-248), the con2prim hook MATTER.c2p
 * (called as MATTER.c2p(ucur, primVars), src/projects/matter/bam_matter.h:49,
 * matter.c:317,411), the R/P call restrict_prolong_varlist(tG *, int, int,
 * tVarList *, tVarList *, int) (src/main/amr/inject.c:66) and the RK4 step
 * evolve_rk_rk4(tL *, int) (src/utility/evolve/rk.c:216).  The adapters
 * below have exactly these types and forward to the device path of
 * bamhip.h, so a BAM build installs them where it installs its CPU
 * functions.
 *
 * Types.  A BAM build compiles bam_shim.c against BAM's own bam.h.  Outside
 * BAM (tests, other hosts) the structures below stand in for it: they have
 * the member order and types of tB, tL, tG and tVarList
 * (src/main/amr/bam_amr.h:10-75,127-131), so code written against BAM's
 * headers and code written against these see the same layout.  Members the
 * adapters never read are kept (as opaque pointers where their types are
 * BAM-internal) so that the offsets agree.
 *
 * Data.  tL::v holds host arrays as in BAM; the adapters work on the device
 * mirror attached to the level (bamhip_bam_attach), which holds the same
 * variable indices.  bamhip_bam_put / bamhip_bam_get move a list between the
 * two (the sync points BAM would place after initial data and before
 * output).
 *
 * Errors.  The reference stops with errorexit(file, line, msg).  The
 * adapters report every nonzero bamhip return through the hook set with
 * bamhip_bam_set_errorexit (default: message on stderr and exit(1)); a BAM
 * build passes BAM's errorexit.  If the hook returns, the adapter returns
 * without further work.
 */
#ifndef BAMHIP_BAM_H
#define BAMHIP_BAM_H

#include "bamhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef BAMHIP_BAM_HAVE_TYPES
/* src/main/amr/bam_amr.h:10-31 */
typedef struct tBOX {
  struct tBOX *pr;
  struct tLEVEL *level;
  int l;
  int i;
  double dx, dy, dz;
  double dd[3];
  int di, dj, dk;
  int id[3];
  double bbox[6];
  int ibbox[6];
  void *com;
  int npoints;
  int noffset;
  double x0, y0, z0;
  int m, n, o;
  int imax, jmax, kmax;
  int bflag[6];
} tB;

/* src/main/amr/bam_amr.h:42-70 */
typedef struct tLEVEL {
  struct tGRID *grid;
  struct tLEVEL *prlocal;
  int l;
  int shells;
  int nboxes;
  tB **box;
  int npoints;
  int nnodes;
  int *boundary;
  int nvariables;
  double **v;
  int iteration;
  double time;
  double dx, dy, dz;
  double dt;
  double bbox[6];
  int ibbox[6];
  void *sbox;
  int nsboxes;
  void *com;
} tL;

/* src/main/amr/bam_amr.h:76-94 */
typedef struct tGRID {
  tL **level;
  int nlevels;
  int lmin;
  int lmax;
  int ltop;
  int lbo;
  int nvariables;
  int half[3];
  int symmetric[3];
  int full, bitant, rotant, quadrant, octant, qreflect;
  int npunc;
  int *lmaxpunc;
  double **puncpos;
} tG;

/* src/main/amr/bam_amr.h:127-131 */
typedef struct tVARLIST {
  struct tLEVEL *level;
  int n;
  int *index;
} tVarList;
#endif

/* the hook types of the reference */
typedef void (*bamhip_bam_rhs_fn)(tVarList *, tVarList *, double, tVarList *); /* evolve.c:22-29 */
typedef void (*bamhip_bam_c2p_fn)(tVarList *, tVarList *);                       /* MATTER.c2p */
typedef void (*bamhip_bam_rp_fn)(tG *, int, int, tVarList *, tVarList *, int);   /* inject.c:66 */
typedef void (*bamhip_bam_step_fn)(tL *, int);                                   /* rk.c:216 */

/* ---------------------------------------------------------------- set-up */
/* errorexit(file, line, msg) of the caller (bam_main.h:112) */
void bamhip_bam_set_errorexit(void (*hook)(const char *file, int line, const char *msg));

/* Attach the device mirror dev to the BAM level (one per level; a regrid
 * attaches the new mirror, which replaces the old one).  The mirror's boxes
 * and variable indices are those of the level.  detach also frees the
 * adapters' per-level buffers (c2p status and iteration counts). */
int bamhip_bam_attach(tL *level, bamhip_level *dev);
int bamhip_bam_detach(tL *level);
bamhip_level *bamhip_bam_device(const tL *level);

/* Parameter snapshots, taken where BAM reads them with Getd/Getv (once at
 * startup instead of on every call).  The Z4c snapshot's dissfactor is
 * replaced per level by get_dissipation_factor's rule (evolve.c:486-530);
 * pgh is the periodic ghost width for set_boundary (0: not periodic). */
int bamhip_bam_set_z4c(const bamhip_z4c_params *par, const bamhip_dissipation_rule *rule, int pgh);
/* MATTER.INDX_VAR_q: position of D in the conserved list (then tau, S_i);
 * MATTER.INDX_VAR_gxx: variable index of adm_gxx (then gxy .. gzz);
 * method: BAMHIP_C2P_HYBRID (grhd_c2p_proot_ColdStaticAtm_hybrid),
 * BAMHIP_C2P_PLAIN (grhd_c2p_proot: stops on a failed point),
 * BAMHIP_C2P_BRACKETED or BAMHIP_C2P_VANAL (grhd_c2p_vanal: stops on a
 * failed point) */
int bamhip_bam_set_grhd(const bamhip_c2p_params *par, int method, int indx_var_q, int indx_var_gxx);
/* variable indices of adm_gxx, adm_Kxx (6 components each, z4_to_adm_vl,
 * z4.c:97-98) and of the matter ADM variables adm_rho, adm_Sx (3), adm_SSxx
 * (6), adm_ST (z4_rhs_add_source.c:26-40) */
int bamhip_bam_set_adm(int adm_gxx, int adm_Kxx, int adm_rho, int adm_Sx, int adm_SSxx, int adm_ST);
/* order_RP of the R/P operators (2, 4, 6 or 8) */
int bamhip_bam_set_fmr(int order_rp);
/* the RK registers evolve_rk_rk4 reads (evolve.c:15-17: u_c, u_p, and the
 * stage lists u_k[0], u_k[1] that evolve_rk adds, rk.c:300-303) and whether
 * the algebraic constraints follow the step (z4_algcon_wrapper, registered
 * after evolve at bam_z4.c:41-44) */
int bamhip_bam_set_rk_lists(tVarList *uc, tVarList *up, tVarList *r, tVarList *v, int algcon);

/* enable the list's variables on the device mirror and upload them from
 * level->v (put), or download them into level->v (get) */
int bamhip_bam_put(tVarList *vl);
int bamhip_bam_get(tVarList *vl);

/* -------------------------------------------------------------- adapters */
/* z4_rhs_movpunc_N(unew, upre, c, ucur) (z4_rhs_movpunc_N.c:23) */
void z4_rhs_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur);
/* set_boundary(unew, upre, c, ucur) for the Z4c list (boundary.c:25) */
void set_boundary_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur);
/* z4_to_adm_vl(unew, upre, c, ucur) (z4.c:93-124), "geometry_adm" */
void z4_to_adm_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur);
/* z4_rhs_add_source(unew, upre, c, ucur) (z4_rhs_add_source.c), "geometry_source" */
void z4_rhs_add_source_hip(tVarList *unew, tVarList *upre, double c, tVarList *ucur);
/* MATTER.c2p(ucur, primVars) (grhd_c2p_v2.c:265, grhd_c2p.c:102) */
void grhd_c2p_hip(tVarList *ucur, tVarList *primVars);
/* restrict_prolong_varlist(g, lcoarse, lfine, uc, uf, nbuffer) (inject.c:66) */
void restrict_prolong_varlist_hip(tG *g, int lcoarse, int lfine, tVarList *uc, tVarList *uf, int nbuffer);
/* evolve_rk_rk4(level, comp) (rk.c:216) with the Z4c right-hand side and
 * boundary above, followed by z4_algcon when enabled */
void evolve_rk_rk4_hip(tL *level, int comp);

/* The adapters as the reference's hooks hold them. */
typedef struct {
  bamhip_bam_rhs_fn rhs_geometry;
  bamhip_bam_rhs_fn rhs_boundary;
  bamhip_bam_rhs_fn rhs_geometry_adm;
  bamhip_bam_rhs_fn rhs_geometry_source;
  bamhip_bam_c2p_fn c2p;
  bamhip_bam_rp_fn restrict_prolong_varlist;
  bamhip_bam_step_fn evolve_rk4;
} bamhip_bam_hooks;
extern const bamhip_bam_hooks bamhip_bam;

/* Install the four Z4c adapters through the caller's set_rhs_func_register
 * (evolve.c:185) under the names z4_startup uses for its CPU functions
 * (z4.c:245-248): geometry, geometry_adm, geometry_source, geometry_bound. */
void bamhip_bam_register(void (*set_rhs_func_register)(char *name, bamhip_bam_rhs_fn f));

#ifdef __cplusplus
}
#endif
#endif
