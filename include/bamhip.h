/* bamhip.h -- C ABI of the MI355X (gfx950) Z4c / RK4 / FMR / con2prim path.
 *
 * The library keeps a device mirror of BAM levels.  A level is a set of boxes
 * whose points are concatenated in one linear index space (tL::box[b]->noffset,
 * src/main/amr/bam_amr.h:10-75) and, per BAM variable index, one device array
 * of doubles of length npoints (the device counterpart of tL::v[index],
 * enabled on demand like enablevarcomp, src/main/main/storage.c:79-95).
 * Variable lists are passed as plain int arrays of BAM variable indices: the
 * `index` member of tVarList (bam_amr.h:127-131).
 *
 * All functions return 0 on success and a negative code on error; the text
 * of the last error is available from bamhip_last_error().  Errors that BAM
 * reports with errorexit() (non-finite RHS, c2p failure with no atmosphere,
 * boxes outside their parent) are returned as codes instead.
 *
 * Threading: one host thread per device, like BAM's one MPI rank per process.
 * All device work is issued on the library's stream, which the caller can
 * query to order its own work.  Entry points synchronise only where they
 * return data to the host.
 */
#ifndef BAMHIP_H
#define BAMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BAMHIP_OK = 0,
  BAMHIP_EINVAL = -1,    /* bad argument (size, index, parameter) */
  BAMHIP_EDEVICE = -2,   /* HIP runtime error */
  BAMHIP_ENOMEM = -3,
  BAMHIP_ENONFINITE = -4 /* non-finite values produced (errorexit in BAM) */
};

/* ---------------------------------------------------------------- runtime */
/* select the device and create the library stream (bampi_initialize) */
int bamhip_init(int device);
int bamhip_finalize(void);
const char *bamhip_last_error(void);
/* the hipStream_t all library work runs on */
void *bamhip_stream(void);
int bamhip_synchronize(void);
/* device buffers for callers without their own allocator */
int bamhip_malloc(void **ptr, size_t bytes);
int bamhip_free(void *ptr);
int bamhip_memcpy_htod(void *dst, const void *src, size_t bytes);
int bamhip_memcpy_dtoh(void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------ levels */
/* One box (tB, src/main/amr/bam_amr.h:10-31).  skip[f] = number of flagged
 * point layers at face f (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z): the flags BAM keeps
 * in tL::boundary (PHYBOUND = 1 layer, PERBOUND = 2, GHOBOUND = nghosts,
 * REFBOUND = 1).  bflag[f] is the reference's per-face flag (bam_amr.h:114-122)
 * and selects the boundary treatment in bamhip_z4c_set_boundary. */
typedef struct {
  int m, n, o;            /* points per direction */
  double x0, y0, z0;      /* coordinates of point (0,0,0) */
  double dx, dy, dz;
  int skip[6];
  int bflag[6];
} bamhip_box;

enum { BAMHIP_REFBOUND = -1, BAMHIP_NOTBOUND = 0, BAMHIP_PHYBOUND = 1, BAMHIP_GHOBOUND = 2,
       BAMHIP_SYMBOUND = 3, BAMHIP_PERBOUND = 6 };

typedef struct bamhip_level bamhip_level;

/* create a level of nboxes boxes with room for nvariables variable indices;
 * box b occupies points [noffset_b, noffset_b + m*n*o) with noffset
 * accumulated in box order, as BAM does */
int bamhip_level_create(bamhip_level **lev, int l, int nboxes, const bamhip_box *boxes,
                        int nvariables);
int bamhip_level_free(bamhip_level *lev);
int bamhip_level_npoints(const bamhip_level *lev, long *npoints);
/* enablevarcomp / disablevarcomp on the device mirror */
int bamhip_level_enable(bamhip_level *lev, int var);
int bamhip_level_disable(bamhip_level *lev, int var);
/* device pointer of variable var (NULL if not enabled) */
double *bamhip_level_ptr(bamhip_level *lev, int var);
/* bamhip_{sync_to_device,sync_to_host}: copy npoints doubles */
int bamhip_level_put(bamhip_level *lev, int var, const double *host);
int bamhip_level_get(bamhip_level *lev, int var, double *host);
int bamhip_level_fill(bamhip_level *lev, int var, double value);
/* the level number and box count, and box b with its noffset (host copies) */
int bamhip_level_info(const bamhip_level *lev, int *l, int *nboxes);
int bamhip_level_box(const bamhip_level *lev, int b, bamhip_box *box, long *noffset);
/* Reductions of write_level_0d (src/utility/output/output0d.c:28-118)
 * through bampi_allreduce (src/main/bampi/reduce.c:36-131) over the inner
 * points of the level: type 0 max, 1 min, 2 l2 norm (sum of squares over
 * the count of all nv variables, times sqrt(nv)), 3 max |u|, 4 sum * dx^3.
 * result[nv] on the host; deterministic (fixed blocks, fixed sum order). */
int bamhip_level_reduce(bamhip_level *lev, int nv, const int *vars, int type, double *result);
/* bamhip_level_reduce without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels, block layout and results (bitwise) as
 * bamhip_level_reduce; the block partials live in a fixed device array set
 * up by bamhip_init on each device and kept until the process ends, so a
 * level may need at most 4096 reduction blocks (each box takes
 * min(1024, ceil(inner points / 256))), at most 64 boxes and 64 variables;
 * larger calls return BAMHIP_EINVAL.  Calls that overlap on one device share
 * the array and race. */
int bamhip_level_reduce_noalloc(bamhip_level *lev, int nv, const int *vars, int type, double *result);
/* host out[t] = var[idx[t]] for n level indices (bufferorderedpoints) */
int bamhip_level_gather(bamhip_level *lev, int var, long n, const long *idx, double *out);

/* ------------------------------------------------------------------- Z4c */
/* Snapshot of the parameters z4_rhs_movpunc_N reads on every call
 * (src/projects/z4/z4_rhs_movpunc_N.c:108-142); flags hold the value of the
 * corresponding Getv() test, dissfactor the value of get_dissipation_factor()
 * (src/utility/evolve/evolve.c:489-530) for the level. */
typedef struct {
  int order_centered;     /* 2 or 4 */
  int order_advection;    /* 2 or 4 */
  int advection_lopsided; /* 0, 1 or 2 */
  int order_dissipation;  /* 0 or 4 */
  double dissfactor;
  double kappa1, kappa2;
  double zeroThetaRHS;
  int subtractA, normalizedetg;
  int oploglapse, oploglapse2, oplogwithshift, harmoniclapse;
  double lapseharmonicf;
  double gamma0factor, gamma2factor, withGadv, withShiftadv, withBadv, withB;
  double shiftalphapower, shiftgammacoeff, shiftdriver;
  double chiDivFloor, chipsipower;
} bamhip_z4c_params;

#define BAMHIP_Z4C_NVARS 25

/* The parameters get_dissipation_factor reads (src/utility/evolve/
 * evolve.c:486-530): dissipation_factor, order_dissipation, amr_move_lcube,
 * whether physics contains "Z4d", and dissipation_factor_level0, _shells,
 * _move, where -1 means unset (the defaults, bam_main.c:69-72,
 * bam_amr.c:85-90). */
typedef struct {
  double dissipation_factor;
  int order_dissipation;
  int amr_move_lcube;
  int physics_z4d;
  double factor_level0, factor_shells, factor_move;
} bamhip_dissipation_rule;
/* get_dissipation_factor(level) for level index l (shells: the level holds
 * shell coordinates): the value for bamhip_z4c_params.dissfactor.  Host only. */
int bamhip_dissipation_factor(const bamhip_dissipation_rule *R, int l, int shells, double *factor);

/* Replaces z4_rhs_movpunc_N(unew, upre, c, ucur) (z4_rhs_movpunc_N.c:23)
 * registered through set_rhs_func_register("geometry", ...) (evolve.c:185,
 * z4.c:245).  unew/upre/ucur: the 25 variable indices of each list in the
 * order of z4.c:234-242.  c == 0: unew = RHS at all inner points; c != 0:
 * unew = upre + c*RHS with trace removal / det normalisation.  upre may be
 * NULL when c == 0.  Returns BAMHIP_ENONFINITE if any RHS value is not finite
 * (the reference calls errorexit, z4_rhs_movpunc_N.c:10473-10479). */
int bamhip_z4c_rhs(bamhip_level *lev, const bamhip_z4c_params *par,
                   const int *unew, const int *upre, double c, const int *ucur);

/* Autotuner of the RHS tile kernels (no counterpart in the reference, whose
 * RHS is one OpenMP loop nest, z4_rhs_movpunc_N.c:23).  kind 0: the bulk
 * kernel (points >= 3 from every box edge); kind 1: the box-edge shell.
 * A variant fixes the tile shape (its z depth included) and the LDS staging;
 * all variants give bitwise the same RHS.  In mode 1 (default) the first
 * bamhip_z4c_rhs call on a new set of boxes, outside graph capture and with
 * unew not aliasing upre, times every variant and keeps the fastest for that
 * set; mode 0 uses variant 0.  bamhip_z4c_force_variant(kind, i) pins variant
 * i (-1: back to the tuner).  tune_report writes one line per tuned set with
 * the time of each variant and returns the length needed (with the NUL). */
int bamhip_z4c_tune_mode(int mode);
/* on != 0 (default): the shell kernels run on a second stream concurrently
 * with the bulk kernel (both in a captured graph too); returns the old value */
int bamhip_z4c_set_overlap(int on);
/* on != 0 (default): the tile kernels number their tiles so that each of the
 * 8 XCDs (workgroup w runs on XCD w mod 8) takes one contiguous range of
 * tiles, whose shared halo points then hit in that XCD's L2; returns the old
 * value */
int bamhip_z4c_set_xcd_map(int on);
int bamhip_z4c_tune_clear(void);
int bamhip_z4c_variant_count(int kind);
const char *bamhip_z4c_variant_name(int kind, int idx);
int bamhip_z4c_force_variant(int kind, int idx);
int bamhip_z4c_tune_report(char *buf, int len);

/* Replaces set_boundary(unew, upre, c, ucur) for the Z4c list
 * (src/utility/boundary/boundary.c:25-224): radiative condition with the far
 * limits and speeds of z4_startup (z4.c:200-228) on PHYBOUND faces, then the
 * periodic copy of pgh layers on PERBOUND face pairs. */
int bamhip_z4c_set_boundary(bamhip_level *lev, const bamhip_z4c_params *par,
                            const int *unew, const int *upre, double c, const int *ucur,
                            int pgh);

/* vladd / vladdto / vlcopy over the n variables of the lists
 * (src/main/main/variables.c:575-745): r = ca*a + cb*b on all points */
int bamhip_vladd(bamhip_level *lev, int n, const int *r, double ca, const int *a,
                 double cb, const int *b);

/* fused vladdto + vladd of one RK4 stage (rk.c:231-252) over the n variables
 * of the lists: uc += wc*r, then v = up + wv*r (skipped when wv == 0) */
int bamhip_rk4_stage_update(bamhip_level *lev, int n, const int *uc, const int *up, const int *r,
                            const int *v, double wc, double wv);

/* non-finite RHS detection mode of a level: 1 (default) = bamhip_z4c_rhs
 * synchronises and returns BAMHIP_ENONFINITE, as the reference's errorexit;
 * 0 = no synchronisation, the flag accumulates until bamhip_level_nonfinite
 * reads (and optionally resets) it */
int bamhip_level_set_check(bamhip_level *lev, int mode);
int bamhip_level_get_check(const bamhip_level *lev, int *mode);
int bamhip_level_nonfinite(bamhip_level *lev, int reset, int *flag);

/* Replaces evolve_rk_rk4 (src/utility/evolve/rk.c:220-257) for the Z4c list:
 * one full RK4 step of size dt on the level with the RHS and boundary above,
 * using fused stage kernels.  uc: state (updated), up: saved state u_p,
 * r, v: stage registers.  Algebraic constraints (z4_algcon_wrapper) are
 * applied afterwards when algcon != 0.  pgh: periodic layers (0 = none).
 * The non-finite check runs once at the end of the step; the level's check
 * mode is restored on return. */
int bamhip_z4c_rk4_step(bamhip_level *lev, const bamhip_z4c_params *par, double dt,
                        const int *uc, const int *up, const int *r, const int *v,
                        int pgh, int algcon);

/* z4_algcon (z4_algcon.m) on all points: g: 6 indices of gt, A: 6 of At */
int bamhip_z4c_algcon(bamhip_level *lev, const int *g, const int *A, int subtractA,
                      int normalizedetg);

/* L2 norm sqrt(sum H^2/n) of the Hamiltonian constraint over inner points
 * (adm_constraints_N order 4 on z4_to_adm data, reduce.c:25-30,209-217);
 * ham: optional variable index receiving H (-1: not stored) */
int bamhip_z4c_ham_l2(bamhip_level *lev, const int *u, double chipsipower, int ham,
                      double *l2);

/* Hamiltonian and momentum constraints of computeadmconstraints
 * (src/projects/adm/adm.c:14-74 -> adm_constraints_N.m:14-116, usepsi off,
 * no matter; order 4, order 2 one point from the box edge) with the L2 norms
 * of bampi_allreduce_norm2 (reduce.c:25-30,209-217) over inner points:
 * l2[0] = |ham|, l2[1..3] = |mom_x|, |mom_y|, |mom_z|.  ham: optional variable
 * index receiving H (-1: not stored); mom: optional 3 indices receiving
 * mom_i (NULL: not stored). */
int bamhip_z4c_constraints_l2(bamhip_level *lev, const int *u, double chipsipower, int ham,
                              const int *mom, double *l2);

/* --------------------------------------------------------------- con2prim */
/* EOS and atmosphere snapshot read by the reference c2p (GRHD.ATM_*,
 * HRSC_VMAX, grhd_C2P_NewtonRaphsonTR/NR; src/projects/grhd/grhd_atm.c:55-87,
 * bam_grhd.c:46-62, eos_analytic.c:52-72,423-434) */
typedef struct {
  int eos;          /* 0: ideal gas (cold part = polytrope), 1: polytrope */
  double gamma, K;  /* Gamma (and K of the polytrope) */
  double rhoatm, fatm, epslatm, patm;
  double vmax;      /* grhd_vmax, compared with v^2 as in the reference */
  double errormax;  /* grhd_C2P_NewtonRaphsonTR */
  int countmax;     /* grhd_C2P_NewtonRaphsonNR */
} bamhip_c2p_params;

/* methods: the production hybrid p-root with cold static atmosphere
 * (grhd_c2p_proot_ColdStaticAtm_hybrid, grhd_c2p_v2.c:265), the plain
 * p-root without atmosphere (grhd_c2p_proot, grhd_c2p.c:102), and this
 * library's bracketed p-root (Newton safeguarded by bisection, stopping at
 * the rounding floor of the residual; no reference counterpart), which
 * converges where the reference's relative-change test cannot be met */
enum { BAMHIP_C2P_HYBRID = 0, BAMHIP_C2P_PLAIN = 1, BAMHIP_C2P_BRACKETED = 2 };
/* the analytic Gamma-law recovery grhd_c2p_vanal (grhd_c2p.c:239-379,
 * selected by grhd_C2P = v_anal, grhd.c:160): |v| from the closed-form
 * root of a quartic (resolvent cubic, then a quadratic factor), clamped to
 * [0, 1 - 1e-15], then rho, p, eps in closed form.  Ideal gas only.  The
 * reference's stops (detg <= 0, non-finite input, eps < 0) give status
 * FAIL with the outputs other than detg untouched; v^2 is not written;
 * iters = 0.  Method value 4 (3 is BAMHIP_C2P_LEARNED). */
#define BAMHIP_C2P_VANAL 4
/* status per point: the reference's codes (grhd_c2p_v2.c:14-20) */
enum { BAMHIP_C2P_COLD = 0, BAMHIP_C2P_ATM = 1, BAMHIP_C2P_OK = 2, BAMHIP_C2P_FAIL = 3, BAMHIP_C2P_NR = 4 };

/* Batched c2p on flat device arrays of n points (SoA):
 * cons[5n] = densitised D, tau, S_x, S_y, S_z (rewritten on COLD/ATM, as the
 * reference does), metric[6n] = gamma_ij (xx xy xz yy yz zz),
 * prim[8n] = rho, eps, v^x, v^y, v^z, p, v^2, detg (on entry p and rho hold
 * the previous values used as initial guesses), status[n], iters[n]. */
int bamhip_c2p_batch(const bamhip_c2p_params *par, int method, long n, double *cons,
                     const double *metric, double *prim, int *status, int *iters);

/* primitives -> conserved variables on flat device arrays of n points:
 * grhd_compute_v2_pt + EOS + grhd_compute_q_pt (src/projects/grhd/
 * grhd_p2c.c:11-25,68-86).  prim5[5n] = rho, eps, v^x, v^y, v^z
 * (contravariant), metric[6n], cons[5n] = densitised D, tau, S_i. */
int bamhip_p2c_batch(const bamhip_c2p_params *par, long n, const double *prim5, const double *metric,
                     double *cons);

/* Replaces MATTER.c2p(ucur, primVars) (src/projects/matter/bam_matter.h:49,
 * called at matter.c:317,411) on a level.  BAMHIP_C2P_BRACKETED here takes
 * the hybrid's control flow: the bracketed solve from the previous pressure
 * replaces try_hot where D >= fatm rhoatm and it returns OK above the
 * atmosphere threshold with eps >= 0 and v^2 <= vmax; every other point runs
 * the reference hybrid (cold branch, atmosphere, conserved rewrites).
 * cons5 / metric6 / prim8 are the
 * variable indices of the conserved variables, of adm_gxx.. and of the
 * primVars list (rho, epsl, vx, vy, vz, p, v2, detg; grhd.c:46-53).
 * status / iters: device arrays of npoints ints. */
int bamhip_grhd_c2p(bamhip_level *lev, const bamhip_c2p_params *par, int method, const int *cons5,
                    const int *metric6, const int *prim8, int *status, int *iters);

/* Replaces MATTER.set_atm_prerhs = grhd_atm_set_prerhs(unew, upre, c, ucur,
 * primVars) (src/projects/grhd/grhd_atm.c:268-278, registered for the
 * ColdStatic atmosphere at grhd.c:99-110 and called after the c2p at
 * matter.c:414-419) with grhd_use_atmosphere_mask = no (its default):
 * grhd_atm_set_u(upre, primVars) (grhd_atm.c:217-262) at every point of the
 * level.  Where rho < fatm * rhoatm: detg = max(detg, 1e-5) in prim8[7],
 * D = sqrt(detg) * rhoatm, S_i = 0, tau = sqrt(detg) * (rhoatm * epslatm)
 * (sqrt(detg) = 1 if not finite).  upre5: D, tau, S_x, S_y, S_z; prim8 as
 * bamhip_grhd_c2p (rho read, detg read and written).  A point non-finite
 * after the reset is zeroed (detg = 1) and raises the level's non-finite
 * flag (the reference's errorexit): in check mode 1 the call synchronises
 * and returns BAMHIP_ENONFINITE, in mode 0 the flag is deferred. */
int bamhip_grhd_atm_set_prerhs(bamhip_level *lev, const bamhip_c2p_params *par, const int *upre5,
                               const int *prim8);

/* ----------------------------------------------- learned pressure guess */
/* A 3 -> 64 -> 64 -> 1 ReLU network predicting log(p / ((Gamma-1) Delta)),
 * Delta = tau + D - sqrt(S^2 + D^2) the state's thermal content (zero for a
 * cold flow), from log D, |S|/(tau+D) and log(Delta/(tau+D)), all formed in
 * FP64 from the undensitised state (inputs and output normalised by
 * the 8 floats `norm`: input shift[3], input scale[3], output shift, output
 * scale).  bf16 weights with FP32 masters, FP32 accumulation, trained on
 * the device with Adam.  No reference counterpart: it replaces only the
 * initial pressure guess of the bracketed solver, which polishes it in
 * FP64 and remains the fallback. */
typedef struct bamhip_mlp bamhip_mlp;
#define BAMHIP_MLP_NPARAM 4481  /* W1[64][3] b1[64] W2[64][64] b2[64] W3[64] b3 */

int bamhip_mlp_create(bamhip_mlp **m, unsigned long long seed);
int bamhip_mlp_free(bamhip_mlp *m);
/* host copies of the FP32 master parameters and of norm[8] (either may be NULL) */
int bamhip_mlp_get_params(const bamhip_mlp *m, float *params, float *norm);
/* set parameters (resets the optimiser) and/or norm */
int bamhip_mlp_set_params(bamhip_mlp *m, const float *params, const float *norm);
/* training data from solved points (device arrays of n points, layouts of
 * bamhip_c2p_batch): feat[3n] raw features, target[n] (NaN unless status OK) */
int bamhip_mlp_prepare(const bamhip_c2p_params *par, long n, const double *cons, const double *metric,
                       const double *prim, const int *status, float *feat, float *target);
/* gradient of the mean squared error over all n samples (host grad[NPARAM],
 * mean loss), without changing the model */
int bamhip_mlp_grad(bamhip_mlp *m, long n, const float *feat, const float *target, float *grad, float *loss);
/* Adam (cosine-decayed learning rate) over `epochs` passes in minibatches;
 * epoch_loss[epochs] (may be NULL) receives the last minibatch's loss */
int bamhip_mlp_train(bamhip_mlp *m, long n, const float *feat, const float *target, int epochs, long batch,
                     float lr, unsigned long long seed, float *epoch_loss);
/* network output log(p / p_hi) for raw features (device arrays) */
int bamhip_mlp_predict(const bamhip_mlp *m, long n, const float *feat, float *y);
/* con2prim with the learned guess, polished by the bracketed solver in at
 * most polish_max iterations; failed polishes fall back to the cold-started
 * bracketed solve and are counted in *fallbacks (device int, may be NULL).
 * Arrays as bamhip_c2p_batch; ideal-gas EOS only. */
int bamhip_c2p_learned_batch(const bamhip_c2p_params *par, const bamhip_mlp *m, int polish_max, long n,
                             double *cons, const double *metric, double *prim, int *status, int *iters,
                             int *fallbacks);
/* MATTER.c2p with the learned guess on the variables of a level (arrays as
 * bamhip_grhd_c2p): the reference hybrid grhd_c2p_proot_ColdStaticAtm_hybrid
 * (grhd_c2p_v2.c:265-464) whose hot solve (try_hot, the reference's Newton
 * iteration and stopping rule) starts at the network's pressure and runs at
 * most polish_max iterations.  A point where that does not return OK runs
 * the reference hybrid from the previous pressure prim8[5] (counted in
 * *fallbacks), so the cold branch, the atmosphere and the conserved
 * rewrites are the reference's.  States with D < fatm rhoatm (atmosphere
 * whatever the pressure) or no thermal content skip the network. */
int bamhip_grhd_c2p_learned(bamhip_level *lev, const bamhip_c2p_params *par, const bamhip_mlp *m, int polish_max,
                            const int *cons5, const int *metric6, const int *prim8, int *status, int *iters,
                            int *fallbacks);

/* ----------------------------------------------------------- FMR coupling */
/* Replaces setRPflags(lc, lf) (src/main/amr/flag.c:122-262) and links the
 * fine level lf to its parent level lc (tL::prlocal, tB::pr; box.c:560-652):
 * fine box b lies in box parent[b] of lc with half its spacing.  Writes
 * flagprolong (variable fp of lf: depth 1..nbuffer from the fine box
 * boundary, else 0) and flagrestrict (variable fr of lc: 3, 2, 1 or 0) and
 * records the transfer regions the varlist calls below use, derived from
 * these flags as prolong_varlist_boxtobox_N / restrict_varlist_boxtobox_N
 * derive them (interpolate_boxtobox_N.c:890-935, 1240-1275).  nbuffer is
 * amr_nbuffer, also the prolongation zone.  Call again after any regrid or
 * change of the flags.  Fine boxes must have no symmetry or process-ghost
 * faces. */
int bamhip_set_rp_flags(bamhip_level *lc, bamhip_level *lf, const int *parent, int nbuffer, int fp, int fr);

/* Replaces restrict_varlist(g, lfine, lcoarse, uf, uc, nbuffer)
 * (interpolate_boxtobox_N.c:92-180, Lagrange part): order-N Lagrange
 * restriction (N = order_RP = 2, 4, 6, 8) of the n variables uf of lf into
 * uc of lc over the bounding box of the coarse points with flagrestrict >= 3
 * (nbuffer >= 3) or >= 2 (nbuffer == 1).  Bitwise the reference's sweeps. */
int bamhip_restrict_varlist(bamhip_level *lf, bamhip_level *lc, int n, const int *uf, const int *uc, int nbuffer,
                            int order);

/* Replaces prolong_varlist(g, lcoarse, lfine, uc, uf, nbuffer) (:184-279):
 * order-N Lagrange prolongation of uc of lc into uf of lf at the fine points
 * with 0 < flagprolong <= amr_nbuffer (nbuffer >= 3 or -3), flagprolong == 1
 * (nbuffer == 1) or everywhere (nbuffer == -1). */
int bamhip_prolong_varlist(bamhip_level *lc, bamhip_level *lf, int n, const int *uc, const int *uf, int nbuffer,
                           int order);

/* Replaces restrict_prolong_varlist(g, lcoarse, lfine, uc, uf, nbuffer)
 * (src/main/amr/inject.c:66-110): restriction, then prolongation. */
int bamhip_restrict_prolong_varlist(bamhip_level *lc, bamhip_level *lf, int n, const int *uc, const int *uf,
                                    int nbuffer, int order);

/* Berger-Oliger time interpolation of a coarse level (evolve.c:298-340):
 * before the level's step, upp = the order-o extrapolation weights times
 * uc, up, up2.. (o = min(timeinterpolation, iteration + 2), 2..6), then the
 * history lists shift (up4 <- up3 <- up2 <- up).  Lists that order o does
 * not use may be NULL. */
int bamhip_bo_interp_begin(bamhip_level *lev, int timeinterpolation, int iteration, int n, const int *upp,
                           const int *uc, const int *up, const int *up2, const int *up3, const int *up4);
/* after the step (evolve.c:367-380): upp += c_o * uc */
int bamhip_bo_interp_end(bamhip_level *lev, int timeinterpolation, int iteration, int n, const int *upp,
                         const int *uc);

/* ------------------------------------------------ FMR hierarchy evolution */
/* The levels of a Berger-Oliger hierarchy (tG with lmin = 0): levels[l] for
 * l = 0..lmax, each fine level linked to its parent by bamhip_set_rp_flags.
 * Level l steps with dt0 / 2^l; iteration[l] is tL::iteration (updated by
 * the calls).  u, up, upp, r, v: the 25 Z4c indices of the state, the saved
 * state, the parent's time-interpolated state and the RK4 registers, the
 * same on every level.  params[l]: the RHS parameters of level l (the
 * dissipation factor may differ per level).  pgh: periodic ghosts of level 0
 * (0: radiative boundary).  npunc > 0: punc is a device array of 3*npunc
 * puncture coordinates, tracked on level lmax after each of its steps
 * (track_puncture in analyze_level, main.c:415-423, with lmaxpunc = lmax)
 * with the shift u[19..21] / up[19..21], interpolation order order_RP and
 * punc_flags as bamhip_puncture_track's flags.  move_nxyz > 0 (amr = move,
 * amr_move_nxyz): the fine levels follow the punctures (bamhip_move_box),
 * which replaces entries of levels[]; move_lcube is amr_move_lcube (levels
 * up to it stay centred on the origin); nbuffer, fp, fr are
 * bamhip_set_rp_flags' amr_nbuffer and flag variables for the new levels. */
typedef struct {
  int lmax;
  bamhip_level **levels;
  const bamhip_z4c_params *params;
  double dt0;
  int *iteration;
  const int *u, *up, *upp, *r, *v;
  int order_RP;            /* prolongation / restriction order (4) */
  int timeinterpolation;   /* order_timeinterpolation: 2 or 3 */
  int pgh;
  int npunc;               /* 0: no tracking */
  double *punc;            /* device, 3*npunc */
  int punc_flags;
  int move_nxyz;           /* 0: boxes stay where they are */
  int move_lcube;
  int nbuffer, fp, fr;
  /* box sizes of the moved levels (bamhip_nested_sizes): nx, nz of the
   * coarsest grid, the amr_nxyz list (n_amr_nxyz entries, may be 0) and
   * amr_buffer_addextra */
  int nx, nz;
  int n_amr_nxyz;
  const int *amr_nxyz;
  int addextra;
} bamhip_fmr_hierarchy;

/* Replaces advance_levelandsublevels(g, l) (src/main/main/main.c:397-453)
 * for the Z4c list: one step of level l (time interpolation of u_pp around
 * an RK4 step, evolve.c:298-380, then algcon), two steps of every finer
 * level, then restrict_prolong_evolve with the parent (inject.c:116-132),
 * then on level lmax the puncture tracker, then move_box(l + 1)
 * (main.c:425-436) when move_nxyz > 0.  Without moving boxes it issues
 * device work only: the RHS non-finite check is deferred (levels must be in
 * check mode 0, see bamhip_fmr_set_check), so the call can be captured; a
 * box move synchronises and allocates, so such calls cannot. */
int bamhip_advance_levelandsublevels(const bamhip_fmr_hierarchy *H, int l);
int bamhip_fmr_set_check(const bamhip_fmr_hierarchy *H, int mode);
/* OR of the levels' non-finite flags (synchronises); reset clears them */
int bamhip_fmr_nonfinite(const bamhip_fmr_hierarchy *H, int reset, int *flag);

/* ---------------------------------------------------------- moving boxes */
/* The box-size parameters find_nested_bboxes reads (src/main/amr/
 * move.c:233-267; defaults in bam_amr.c:41-69, bam_main.c). */
typedef struct {
  int nx, nz;      /* "nx", "nz" of the coarsest grid (n_y takes nx, as the reference does) */
  int move_nxyz;   /* amr_move_nxyz: size of the moving cubes (0: nx) */
  int move_lcube;  /* amr_move_lcube: fine levels up to it are cubes around the origin */
  int nlist;       /* entries of amr_nxyz (0: none) */
  const int *list; /* amr_nxyz: per-level sizes, multiples of 4 */
  int nbuffer;     /* amr_nbuffer */
  int addextra;    /* amr_buffer_addextra (the reference default: yes = 1) */
} bamhip_nested_sizes;
/* n[3] coarse points per direction of the boxes of fine level lfine
 * (move.c:233-267): nx (nz) of the coarsest grid, amr_move_nxyz above
 * amr_move_lcube, else the amr_nxyz entry min(nlist - 1, lfine), plus
 * 2 nbuffer with addextra, then halved and rounded up to even.  Host only. */
int bamhip_nested_box_size(const bamhip_nested_sizes *S, int lfine, int *n);
/* find_nested_bboxes (move.c:121-465) for fine level lfine, without
 * symmetries or forced motion: around each of the np (<= 10) points pos[3*j..]
 * (all at the origin when lfine <= move_lcube) snapped to the coarse
 * spacing h[3] by truncation, a box of bamhip_nested_box_size points;
 * overlapping boxes are joined in the reference's pairwise order.  Out:
 * *nboxes, bbox[6*b] (coarse coordinates of the first and last point per
 * direction), ibbox[6*b] (0, cells, ...).  The same function lays out the
 * initial grid (make_level_nested_boxes, box.c:437-450).  Host only. */
int bamhip_find_nested_bboxes(const bamhip_nested_sizes *S, int lfine, int np, const double *pos, const double *h,
                              int *nboxes, double *bbox, int *ibbox);
/* move_overlap_nonparallel (move.c:589-676) with copyall: for every pair
 * (old box, new box) in that order, copy the n variables at the points the
 * two boxes share (index ranges of box_overlap_range, move.c:487-533);
 * variables not enabled on both levels are skipped.  Same spacing and
 * aligned points required. */
int bamhip_level_copy_overlap(bamhip_level *lold, bamhip_level *lnew, int n, const int *vars);
/* move_box(g, l) (move.c:808-988) for the hierarchy: returns at once
 * unless move_nxyz > 0, 1 <= l <= lmax and level l-1 is the coarsest of its
 * time (time[l-1] < time[l-2]); then from l to lmax the nested boxes of
 * each level are found on its parent around the punctures (the origin up
 * to move_lcube) with the hierarchy's box sizes (nx, nz, move_nxyz,
 * amr_nxyz, nbuffer, addextra); a level whose boxes differ is replaced by a
 * new one
 * (make_level_bboxes, box.c:465-560: 2*(cells+1) points per direction,
 * first point half a fine spacing inside the coarse box, REFBOUND faces)
 * with the old level's enabled variables, filled by order_RP prolongation
 * everywhere from the parent (prolong_varlist with nbuffer -1) of u, up and
 * upp, then by the copy of the overlap with the old level (the reference's
 * global and local methods give these values); a level kept in place is
 * linked again to a replaced parent.  The box comparison is with the
 * level's present boxes.  A non-finite flag of a replaced level passes to
 * its successor.  *moved (may be NULL) = 1 if a level moved. */
int bamhip_move_box(const bamhip_fmr_hierarchy *H, int l, int *moved);

/* RK4 stages without algcon and without the non-finite check, and algcon on
 * the 12 fields of a Z4c list (building blocks of the drivers above) */
int bamhip_z4c_rk4_stages(bamhip_level *lev, const bamhip_z4c_params *par, double dt, const int *uc, const int *up,
                          const int *r, const int *v, int pgh);
int bamhip_z4c_algcon_list(bamhip_level *lev, const bamhip_z4c_params *par, const int *uc);

/* One level step of the Berger-Oliger evolution (evolve.c:298-380 around
 * evolve_rk_rk4, then z4_algcon_wrapper, z4.c:49-71): time interpolation
 * begin into upp (NULL: finest level, none), the RK4 stages, interpolation
 * end and (algcon != 0) algcon, with orders 2 and 3 (o = min(
 * timeinterpolation, iteration + 2)).  With separate stage updates the list
 * operations are folded into the first and last update passes
 * (bamhip_rk4_first_update: upp = ba*uc + bb*up, up = uc, uc += wc*r,
 * v = up + wv*r; bamhip_rk4_last_update: uc += wc*r, upp += bc*uc, algcon);
 * the results are bitwise those of the separate calls. */
int bamhip_z4c_rk4_step_bo(bamhip_level *lev, const bamhip_z4c_params *par, double dt, const int *uc, const int *up,
                           const int *r, const int *v, int pgh, const int *upp, int timeinterpolation, int iteration,
                           int algcon);
int bamhip_rk4_first_update(bamhip_level *lev, const int *uc, const int *up, const int *r, const int *v, double wc,
                            double wv, const int *upp, double ba, double bb);
int bamhip_rk4_last_update(bamhip_level *lev, const int *uc, const int *r, double wc, const int *upp, double bc,
                           int algcon, int subtractA, int normalizedetg);

/* One RK4 stage of evolve_rk_rk4 (rk.c:231-252: evolve_rhs, set_boundary,
 * vladdto, vladd) with the update fused into the RHS kernels: at every
 * computed point accout = acc + wc * RHS(src), when vout != NULL
 * vout = base + wv * RHS(src), and when copy != NULL copy = base.  The
 * boundary call does the same on PHYBOUND faces (radiative RHS), copies the
 * periodic layers of accout and vout, and with fill_skip sets vout, copy and
 * (when it is another list than acc) accout to base at the other points no
 * RHS kernel computes.  No output may alias src or another output; acc and
 * base are read at the point only and may alias outputs.  scratch
 * (optional, 25 indices, not an input): a list the tile-kernel tuner may
 * overwrite with a plain RHS the first time a set of boxes is seen.  With
 * fusion selected, bamhip_z4c_rk4_stages rotates uc, up, r and v through
 * these roles so that no pass copies up = uc (its comment lists the roles);
 * results are bitwise those of the unfused stages when the periodic layers
 * of uc start consistent.
 * bamhip_z4c_set_rk_fusion(mode): 0 never, 1 always, 2 (default) on levels of
 * at least 2^19 points; returns the previous mode. */
int bamhip_z4c_rhs_rk4_stage(bamhip_level *lev, const bamhip_z4c_params *par, const int *src,
                             const int *acc, const int *accout, const int *base, const int *vout,
                             const int *copy, double wc, double wv, const int *scratch);
int bamhip_z4c_bc_rk4_stage(bamhip_level *lev, const bamhip_z4c_params *par, const int *src,
                            const int *acc, const int *accout, const int *base, const int *vout,
                            const int *copy, double wc, double wv, int pgh, int fill_skip);
int bamhip_z4c_set_rk_fusion(int mode);
/* value at the points of every box that no RHS kernel computes (the skip
 * layers) for the n <= 25 variables vars; no reference counterpart.  The
 * fused stages end with it on r, whose skip points the separate passes take
 * as 0 (a list that only ever holds right-hand sides there). */
int bamhip_skip_fill(bamhip_level *lev, int n, const int *vars, double value);

/* -------------------------------------------------------- GRHD fluxes */
/* hrsc_flux / hrsc_rec / EOS choices of grhd_flx1d_fvpr (grhd.c:193-209,
 * hrsc.c:10-100, eos.c:37-72); the metric is reconstructed with LAG4
 * (hrsc_rec_metric), the atmosphere mask is off (its default). */
typedef struct {
  int eos;          /* 0 ideal gas, 1 polytrope (EOS.COLD: no tau flux) */
  double gamma, K;  /* eos_Gamma, eos_K */
  double vmax;      /* grhd_vmax */
  double wlormax;   /* grhd_Wlor_max */
  int riemann;      /* hrsc_flux: 0 LLF, 1 HLL */
  int rec;          /* hrsc_rec: 0 WENOZ, 1 WENO5 */
} bamhip_hrsc_params;

/* Replaces matter_fluxes(unew, upre, c, ucur, primVars, otherVars)
 * (src/projects/matter/matter_fluxes.c:20-86) with MATTER.hrsc_flx1d =
 * grhd_flx1d_fvpr (grhd_flx1d.c:25-250): at every point of every box,
 * unew -= (F_{i+1/2} - F_{i-1/2}) / h for x, then y, then z (times c when
 * c != 0), lines extended by flat extrapolation, non-finite interface
 * fluxes zeroed.  unew: 5 indices (D, tau, S_x, S_y, S_z); metric: 10
 * (alpha, beta^x..z, ADM gxx, gxy, gxz, gyy, gyz, gzz); prim: 5 (rho,
 * epsl, v^x..z; grhd.c:46-53). */
int bamhip_grhd_fluxes(bamhip_level *lev, const bamhip_hrsc_params *par, const int *unew, double c,
                       const int *metric, const int *prim);

/* Replaces MATTER.sources = grhd_sources(unew, upre, c, ucur, primVars,
 * otherVars, matterADMVars) (src/projects/grhd/grhd_sources.m:8-190,
 * generated grhd_sources.c; called at matter.c:438-442) with
 * grhd_source_computation = prim and grhd_cooling = 0 (bam_grhd.c:51-52):
 * at the inner points, unew (D, tau, S_i) += the geometric source terms
 * (times c when c != 0) and madm (11 indices: rho, S_i, S_ij, S) is written.
 * metric: 10 (alpha, beta^i, ADM g_ij), Kij: 6 (ADM K_ij), prim: 6 (rho,
 * epsl, v^i, p).  rhoatm: GRHD.ATM_RHOATM (no momentum source below it);
 * par->eos 1 (cold EOS): no tau source; order_centered 2 or 4. */
int bamhip_grhd_sources(bamhip_level *lev, const bamhip_hrsc_params *par, double rhoatm, int order_centered,
                        const int *unew, double c, const int *metric, const int *Kij, const int *prim,
                        const int *madm);

/* Replaces z4_rhs_add_source(unew, upre, c, ucur) (src/projects/z4/
 * z4_rhs_add_source.m:14-77; "geometry_source" at z4.c:247, called after
 * the matter RHS, evolve.c:72-76): at the inner points the matter terms of
 * the A_ij, Khat, G^i, Theta and B^i right-hand sides from the matter ADM
 * variables (madm: 11 indices rho, S^i, S_ij, S, as written by
 * bamhip_grhd_sources), times c when c != 0.  unew/ucur: the 25 Z4c
 * indices. */
int bamhip_z4c_matter_source(bamhip_level *lev, const bamhip_z4c_params *par, const int *unew, double c,
                             const int *ucur, const int *madm);

/* Brill-Lindquist puncture data (SetMWBL, src/projects/punctures/
 * ConformalFactor.c:84-200, with the conformal factor absorbed into the
 * metric, adm_undo_conftrans, src/projects/adm/adm.c:96-111): at every point
 * psi = 1 + sum_i m_i / (2 r_i), adm_g = psi^4 delta_ij, adm_K = 0.
 * pos[3 n]: puncture positions, n <= 16. */
int bamhip_punctures_bl(bamhip_level *lev, int n, const double *mass, const double *pos, const int *adm_g,
                        const int *adm_K);
/* The z4 project's POST_INITIALDATA (src/projects/z4/z4.c:280-330): lapse
 * (initial_lapse 0 unchanged, 1 one, 2 precollapsed = 1/sqrt(adm_gxx)),
 * shift and B^i zeroed when zero_shift, then adm_to_z4 (z4.c:142-171) =
 * z4_init (z4_init.m:9-80): the Z4c fields from adm_g / adm_K at every point,
 * G^i at the inner points, and z4_setbound_G's extrapolation
 * (extrapolate.c:22-47) on boxes whose faces are all PHYBOUND. */
int bamhip_adm_to_z4c(bamhip_level *lev, const int *adm_g, const int *adm_K, const int *ucur, double chipsipower,
                      int initial_lapse, int zero_shift);

/* Replaces z4_to_adm_vl(unew, upre, c, ucur) (src/projects/z4/z4.c:93-124),
 * the "geometry_adm" step of evolve_rhs: at every point psi =
 * chi^(1/chipsipower), adm_g = psi^4 g~_ij, adm_K = psi^4 A~_ij + Khat g_ij / 3.
 * ucur: the 25 Z4c indices; adm_g, adm_K: 6 indices each (written). */
int bamhip_z4c_to_adm(bamhip_level *lev, const int *ucur, double chipsipower, const int *adm_g, const int *adm_K);

/* The coupled Z4c + GRHD right-hand side: evolve_rhs (src/utility/evolve/
 * evolve.c:37-111) for physics z4 + matter + grhd as one call:
 * bamhip_z4c_rhs, bamhip_z4c_to_adm, the c2p of c2p_method (BAMHIP_C2P_*
 * on the hybrid/plain/bracketed solvers, or BAMHIP_C2P_LEARNED with mlp),
 * the atmosphere reset of upre (atm_prerhs, upre != NULL),
 * matter_rhs_prepare, bamhip_grhd_sources, bamhip_grhd_fluxes,
 * bamhip_z4c_matter_source and the Z4c boundary.  The lists unew / upre /
 * ucur hold 30 indices: the 25 Z4c fields, then D, tau, S_x, S_y, S_z. */
#define BAMHIP_C2P_LEARNED 3
typedef struct {
  const bamhip_z4c_params *z4;
  const bamhip_c2p_params *c2p;    /* EOS, atmosphere (its rhoatm is GRHD.ATM_RHOATM) */
  const bamhip_hrsc_params *hrsc;
  int c2p_method;
  const bamhip_mlp *mlp;           /* BAMHIP_C2P_LEARNED only */
  int polish_max;
  int *fallbacks;                  /* device int or NULL */
  int pgh;                         /* periodic ghost width (0: radiative boundary) */
  const int *adm_g, *adm_K;        /* 6 + 6 indices (adm_gxx.., adm_Kxx..), written */
  const int *prim;                 /* 8: primVars (rho, epsl, v^i, p, v2, detg) */
  const int *madm;                 /* 11: matter ADM variables, written */
  int *status, *iters;             /* device arrays of npoints ints for the c2p */
  int atm_prerhs;                  /* grhd_use_atmosphere_prerhs (reference default yes): after the c2p,
                                      bamhip_grhd_atm_set_prerhs on upre (matter.c:414-419) */
} bamhip_z4c_grhd;
int bamhip_z4c_grhd_rhs(bamhip_level *lev, const bamhip_z4c_grhd *S, const int *unew, const int *upre, double c,
                        const int *ucur);
/* One RK4 step of the coupled system: evolve_rk_rk4 (src/utility/evolve/
 * rk.c:220-257) over the 30 fields with bamhip_z4c_grhd_rhs(r, up, 0, src)
 * per stage (u_p is upre, so the atmosphere reset acts on it), then algcon
 * on the Z4c fields (z4.c:49-71); lists of 30 indices as above.  The level's
 * check mode is restored on return. */
int bamhip_z4c_grhd_rk4_step(bamhip_level *lev, const bamhip_z4c_grhd *S, double dt, const int *uc, const int *up,
                             const int *r, const int *v, int algcon);

/* ---------------------------------------------------------- punctures */
/* Replaces track_moving_puncture_integrate (src/projects/Gauge/
 * moving_punctures_integrate.c:29-85) on the finest level of a puncture: three
 * ICN iterations x = x_old - dt (beta(x) + beta_p(x_old)) / 2 with the shift
 * interpolated as interpolate_xyz_scalar does (interpolate.c:419-470,
 * Lagrange of `order` points, 0 where no box holds the cube).  beta, beta_p:
 * 3 variable indices of the shift at the new and the previous time level;
 * pos: device array of 3*np coordinates, updated in place (np <= 64), so the
 * tracker can run inside a captured graph.  As track_puncture
 * (moving_punctures.c:111-196) does around it, a new position outside every
 * box of the level is dropped for the old one, and flags
 * BAMHIP_PUNCTURE_PIN_Z keeps z (moving_puncture_fixz = BH / BHNS). */
#define BAMHIP_PUNCTURE_PIN_Z 1
int bamhip_puncture_track(bamhip_level *lev, const int *beta, const int *beta_p, double dt, int order, int flags,
                          int np, double *pos);

/* --------------------------------------------------------------- output */
/* write_level_0d (src/utility/output/output0d.c:28-118, write_scalar_N
 * :11-25): for each variable the line "time  value  " (%22.15e) appended to
 * <outdir>/<names[j]>_<max|min|norm|norminf|integral>.l<l> (output.c:219);
 * type as bamhip_level_reduce. */
int bamhip_output_0d(bamhip_level *lev, int nv, const int *vars, const char *const *names, int type, double time,
                     const char *outdir);
/* write_level_1d for the x, y, z lines (line 1, 2, 3) through the output
 * origin (outputx0..z0) snapped to the next grid point, without
 * interpolation (1doutinterpolate = no; output1d.c:18-103,300-420,
 * points.c:278-335): "\"Time = t\"", rows "coordinate  value" (%22.15e, -1
 * where no box holds the point) and a blank line, appended to
 * <outdir>/<names[j]>.<x|y|z><l>. */
int bamhip_output_1d(bamhip_level *lev, int nv, const int *vars, const char *const *names, int line, double x0,
                     double y0, double z0, double time, const char *outdir);

/* tov_h (src/projects/BBIDdataReader/tov.c:104-236, BBID_tov_integrate = h)
 * for the polytrope p = K rho^Gamma: npts points from the centre enthalpy
 * hc to the surface, per point the Schwarzschild radius r, the mass m, the
 * log enthalpy h, rho, pre, the potential phi and the isotropic radius
 * riso.  out[4] (may be NULL): M, R (Schwarzschild), R (isotropic) and the
 * baryonic mass (computeMb, tov.c:10-24).  Host only. */
int bamhip_tov_h(double hc, double K, double Gamma, int npts, double *r, double *m, double *h, double *rho,
                 double *pre, double *phi, double *riso, double *out);
/* set_TOV (BBIDdataReader.c:483-690) without boost or perturbation: the
 * star of the host table (n points of bamhip_tov_h's riso, r, phi, h; mass
 * M) centred at center[3] on every point of the level: inside the surface
 * riso[n-1] the table read by interp_lag4 (:280-318) at the isotropic
 * radius, alpha = exp(phi), g_ij = (r/riso)^2 delta_ij and the matter of the
 * enthalpy; outside, the exterior Schwarzschild solution in isotropic
 * coordinates and no matter.  adm_g[6], adm_K[6] (set to 0), alpha, prim[6]
 * = rho, epsl, vx, vy, vz (0), p. */
int bamhip_tov_set(bamhip_level *lev, int n, const double *riso, const double *rsch, const double *phi,
                   const double *h, double M, double K, double Gamma, const double *center, const int *adm_g,
                   const int *adm_K, int alpha, const int *prim);

/* ----------------------------------------------------------- checkpoints */
/* checkpoint_write_local / checkpoint_read (src/utility/checkpoint/
 * checkpoint_write.c:91-205, checkpoint_read.c:96-275) for device levels:
 * the reference's file layout (text header with iteration and time per
 * level, the parameter database as name = value lines, then per level and
 * enabled variable a "$variable = name : level = l , level->nnodes = n" line
 * and n binary doubles).  levels[l - lmin], iteration[] / time[] per level,
 * vars / var_names the variables to write or to restore by name.  Reading
 * fills iteration / time (may be NULL) and the listed, enabled variables; a
 * point count other than the level's is an error, the parameter database
 * and unlisted variables are skipped. */
typedef struct {
  int lmin, lmax;
  bamhip_level *const *levels;
  const int *iteration;
  const double *time;
  int nparams;
  const char *const *param_names, *const *param_values;
  int nvars;
  const int *vars;
  const char *const *var_names;
  const double *dt;   /* per-level time step or NULL: reading returns time = iteration * dt, as the reference
                         does (checkpoint_read.c:117-124); NULL: the header's time (printed with %g) */
} bamhip_checkpoint;
int bamhip_checkpoint_write(const char *path, const bamhip_checkpoint *C);
int bamhip_checkpoint_read(const char *path, const bamhip_checkpoint *C, int *iteration, double *time);

/* --------------------------------------------------------------- graphs */
/* Capture the device work issued on the library stream between begin and
 * end into an executable graph (hipGraph), replayed by launch.  Timers are
 * paused during capture; work that synchronises with the host (check mode
 * 1, put/get) fails while capturing. */
int bamhip_graph_begin(void);
int bamhip_graph_end(void **graph);
int bamhip_graph_launch(void *graph);
int bamhip_graph_free(void *graph);

/* ----------------------------------------------------------------- timers */
/* accumulated device time per kernel family (BAM's timer_start/timer_stop
 * keys, src/main/bampi/timer.c:60-104), measured with HIP events on the
 * library stream; names: "z4_rhs", "z4_rhs_edge", "rk4_update", "set_boundary",
 * "algcon", "prolong", "restrict", "c2p", "mlp" */
int bamhip_timer_enable(int on);
int bamhip_timer_reset(void);
int bamhip_timer_get(const char *name, double *ms_total, long *launches);

#ifdef __cplusplus
}
#endif
#endif
