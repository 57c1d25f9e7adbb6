/* oracle.h -- CPU restatement of the BAM Z4c / RK4 / FMR / con2prim hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path in
 * baum_amd/ is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never links it.
 *
 * Every function restates the reference algorithm from the reference's own
 * sources (paths relative to the reference tree, cited at each function) in
 * plain C written for clarity: full 3x3 index loops, no hand expansion.
 * Layout mirrors BAM: one double[m*n*o] per grid function component, index
 * ijk = i + m*j + m*n*k (bam_amr.h:10-31, bam_amr_loops_box.h:117-139).
 */
#ifndef BAMHIP_ORACLE_H
#define BAMHIP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One BAM box (subset of tB, src/main/amr/bam_amr.h:10-31).  skip[f] is the
 * number of flagged point layers on face f (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z);
 * flagged points (PHYBOUND/GHOBOUND/PERBOUND/REFBOUND) are not touched by
 * the RHS loop, which in BAM is forinnerpoints_ijk_openmp = !boundary[ijk]. */
typedef struct {
  int m, n, o;
  double x0, y0, z0;
  double dx, dy, dz;
  int skip[6];
} ob_box;

/* Snapshot of the parameters z4_rhs_movpunc_N reads on every call
 * (src/projects/z4/z4_rhs_movpunc_N.c:108-142, defaults bam_z4.c:30-70). */
typedef struct {
  int order_centered;      /* 2 or 4 */
  int order_advection;     /* 2 or 4 */
  int advection_lopsided;  /* 0,1,2 */
  int order_dissipation;   /* 0 or 4 */
  double dissfactor;       /* get_dissipation_factor(): sigma / 2^order */
  double kappa1, kappa2;
  double zeroThetaRHS;     /* 1 keeps the Theta RHS (Getv(...,"no") = 1) */
  int subtractA, normalizedetg;
  int oploglapse, oploglapse2, oplogwithshift, harmoniclapse;
  double lapseharmonicf;
  double gamma0factor, gamma2factor, withGadv, withShiftadv, withBadv, withB;
  double shiftalphapower, shiftgammacoeff, shiftdriver;
  double chiDivFloor, chipsipower;
} ob_z4c_params;

/* variable order of the Z4c evolution list (z4.c:234-242) */
enum {
  OV_GXX = 0, OV_GXY, OV_GXZ, OV_GYY, OV_GYZ, OV_GZZ,
  OV_CHI,
  OV_AXX, OV_AXY, OV_AXZ, OV_AYY, OV_AYZ, OV_AZZ,
  OV_KHAT,
  OV_GX, OV_GY, OV_GZ,
  OV_THETA,
  OV_ALPHA,
  OV_BETAX, OV_BETAY, OV_BETAZ,
  OV_BX, OV_BY, OV_BZ,
  OV_NVARS
};

/* Z4c right-hand side on one box; returns the number of inner points whose
 * RHS was not finite (the reference errorexits on the first such point). */
int ob_z4c_rhs(const ob_box *b, const ob_z4c_params *p,
               double *const *unew, const double *const *upre, double c,
               const double *const *ucur);

/* radiative (Sommerfeld) condition on the PHYBOUND layer of one variable */
void ob_radiative_bc(const ob_box *b, double *unew, const double *upre, double c,
                     const double *ucur, double var0, double v);
/* all 25 Z4c variables with the far limits / speeds of z4_startup */
void ob_z4c_radiative_bc(const ob_box *b, const ob_z4c_params *p,
                         double *const *unew, const double *const *upre, double c,
                         const double *const *ucur);
/* BAM-style periodic copy of g layers in the directions with per[d] != 0 */
void ob_periodic_fill(const ob_box *b, const int per[3], int g, double *u);

/* algebraic constraints (z4_algcon.m) on all points of one box */
void ob_z4c_algcon(const ob_box *b, int subtractA, int normalizedetg,
                   double *const *g6, double *const *A6);

/* ADM -> Z4c (z4_init.m) and Z4c -> ADM (z4.c:93-124) */
void ob_adm_to_z4c(const ob_box *b, double chipsipower, const double *const *gadm, const double *const *kadm,
                   double *const *u, int phybound);
/* Brill-Lindquist ADM data (SetMWBL, adm_undo_conftrans) */
void ob_punctures_bl(const ob_box *b, int n, const double *mass, const double *pos, double *const *gadm,
                     double *const *kadm, double *psi_out);
void ob_z4c_to_adm(const ob_box *b, double chipsipower, const double *const *u,
                   double *const *gadm, double *const *kadm);

/* full RK4 step on one box (rk.c:220-257 + z4_algcon_wrapper) with either
 * radiative (bc=1) or periodic (bc=2, g layers) boundaries.  u holds the
 * state (25 fields), work must hold 3*25 fields of the same size. */
int ob_z4c_rk4_step(const ob_box *b, const ob_z4c_params *p, double dt, int bc,
                    int pgh, double *const *u, double *const *work, int algcon);

/* vector ops of variables.c:575-745 on npts points */
void ob_vladd(long npts, double *r, double ca, const double *a, double cb, const double *b);

/* Hamiltonian constraint of the Z4c data (via ADM, adm_constraints_N order 4)
 * at inner points; returns the L2 norm sqrt(sum H^2 / n) over them */
double ob_z4c_ham_l2(const ob_box *b, const double *const *u, double chipsipower, double *ham);
void ob_z4c_constraints_l2(const ob_box *b, const double *const *u, double chipsipower, double *ham,
                           double *const *mom, double l2[4]);

/* ---------------- interpolation (src/utility/interpolate) ----------------- */
void ob_lagrange_coeffs(int n, double x, double *c);
void ob_prolong_1d(int order, int nc, const double *uc, double *uf);

/* ---------------- FMR coupling (amr/flag.c, interpolate_boxtobox_N.c) ------ */
void ob_prolong_box(int N, const int *cbox, const int *fbox, const int *fbox0, const double *cv, double *fv,
                    int cdj, int cdk, int fdj, int fdk);
void ob_restrict_box(int N, const int *fbox, const int *cbox, const double *fv, double *cv, int fdj, int fdk,
                     int cdj, int cdk);
void ob_rp_flags(int nf, const ob_box *fb, int nc, const ob_box *cb, const int *parent, int nbuffer, double *fp,
                 double *fr);
int ob_prolong_region(const ob_box *f, const ob_box *c, const double *fp, int nbuffer, int zone, int order,
                      int *ifbox, int *ifbox0, int *icbox);
int ob_restrict_region(const ob_box *f, const ob_box *c, const double *fr, int nbuffer, int order, int *icbox,
                       int *ifbox);
int ob_prolong_level(int nf, const ob_box *fb, int nc, const ob_box *cb, const int *parent, const double *fp,
                     int nbuffer, int zone, int order, int nvar, const double *cv, double *fv);
int ob_restrict_level(int nf, const ob_box *fb, int nc, const ob_box *cb, const int *parent, const double *fr,
                      int nbuffer, int order, int nvar, const double *fv, double *cv);

/* ---------------- con2prim (src/projects/grhd) ---------------------------- */
typedef struct {
  int eos;               /* 0 ideal gas (cold part polytrope), 1 polytrope */
  double gamma, K;       /* Gamma and K of the (cold) polytrope / ideal gas */
  double rhoatm, fatm, epslatm, patm;
  double vmax;           /* grhd_vmax (HRSC_VMAX) */
  double errormax;       /* grhd_C2P_NewtonRaphsonTR */
  int countmax;          /* grhd_C2P_NewtonRaphsonNR */
} ob_c2p_params;

enum { OB_C2P_COLD = 0, OB_C2P_ATM, OB_C2P_OK, OB_C2P_FAIL, OB_C2P_NR };

double ob_invg(double gxx, double gxy, double gxz, double gyy, double gyz, double gzz,
               double *ixx, double *ixy, double *ixz, double *iyy, double *iyz, double *izz);
int ob_try_hot(const ob_c2p_params *P, double D, double S2, double T,
               double *rho, double *eps, double *p, double init, int *iters);
int ob_try_cold(const ob_c2p_params *P, double D, double S2, double *T,
                double *rho, double *eps, double *p, double init, int *iters);
/* point loop of grhd_c2p_proot_ColdStaticAtm_hybrid.  cons: 5*n (D,T,Sx,Sy,Sz,
 * densitised, rewritten on ATM/COLD), metric: 6*n, prim: 8*n (rho,eps,vx,vy,vz,
 * p,v2,detg; p and rho hold the previous values on entry), status/iters: n */
void ob_c2p_hybrid(const ob_c2p_params *P, long n, double *cons, const double *metric,
                   double *prim, int *status, int *iters);
/* grhd_atm_set_u (grhd_atm.c:217-262), the body of grhd_atm_set_prerhs
 * (:268-278) on u_p: cons 5*n (D,T,Sx,Sy,Sz) reset to the cold static
 * atmosphere where rho < fatm*rhoatm, detg[n] floored at DETGMIN = 1e-5
 * there.  Returns the number of points left non-finite (errorexit in the
 * reference; zeroed here, detg = 1). */
long ob_grhd_atm_set_u(const ob_c2p_params *P, long n, double *cons, const double *rho, double *detg);
/* prim -> cons (grhd_p2c.c:11-86) */
void ob_p2c(const ob_c2p_params *P, long n, const double *prim5, const double *metric,
            double *cons);
/* plain p-root (grhd_c2p.c:102-237): no atmosphere; the reference's
 * errorexit cases become status FAIL (bad input) or NR (no convergence) */
void ob_c2p_proot(const ob_c2p_params *P, long n, const double *cons, const double *metric,
                  double *prim, int *status, int *iters);
/* bracketed Newton on the pressure with a rounding-floor stop (no reference
 * counterpart: the safeguarded solver of this build); prim[5n..] holds the
 * initial guess (<= 0: start at the bracket's lower end) */
int ob_c2p_bracketed_pt(const ob_c2p_params *P, double D, double S2, double T, double guess,
                        double *rho, double *eps, double *p, int *iters);
void ob_c2p_bracketed(const ob_c2p_params *P, long n, const double *cons, const double *metric,
                      double *prim, int *status, int *iters);
/* analytic Gamma-law recovery grhd_c2p_vanal (grhd_c2p.c:239-379); coef
 * (optional, 4*n): the quartic's a3, a2, a1, a0 per point */
void ob_c2p_vanal(const ob_c2p_params *P, long n, const double *cons, const double *metric,
                  double *prim, int *status, double *coef);

/* GRHD flux divergence (grhd_flux_oracle.c): matter_fluxes with
 * grhd_flx1d_fvpr.  eos 0 ideal gas, 1 polytrope; riemann 0 LLF, 1 HLL;
 * rec 0 WENOZ, 1 WENO5 (primitives; the metric is always LAG4). */
typedef struct {
  int eos;
  double gamma, K;
  double vmax, wlormax; /* grhd_vmax, grhd_Wlor_max */
  int riemann, rec;
} ob_hrsc_params;
/* kind 0 WENOZ, 1 WENO5, 2 LAG4; side 0 rec1d_p_*(u, i), 1 rec1d_m_*(u, i) */
double ob_hrsc_rec(int kind, int side, const double *u, int i);
/* g: alpha, beta^x, gxx..gzz at the interface; wl/wr: rho, epsl, v^x,y,z */
void ob_grhd_interface_flux(const ob_hrsc_params *P, const double g[8], const double wl[5], const double wr[5],
                            double F[5]);
/* metric: alpha, beta^x,y,z, gxx, gxy, gxz, gyy, gyz, gzz (10 arrays);
 * prim: rho, epsl, v^x,y,z; rhs: D, tau, S_x,y,z, updated in place */
void ob_grhd_fluxes(const ob_box *B, const ob_hrsc_params *P, double c, const double *const *metric,
                    const double *const *prim, double *const *rhs);

/* GRHD source terms and matter ADM variables (grhd_source_oracle.c),
 * grhd_sources at the inner points.  metric: alpha, beta^x,y,z, g_ij (10);
 * Kij: K_ij (6); prim: rho, epsl, v^x,y,z, p (6); rhs: D, tau, S_i (5,
 * updated); madm: rho, S_i, S_ij, S (11, written).  P->eos 1: cold EOS. */
void ob_grhd_sources(const ob_box *B, const ob_hrsc_params *P, double rhoatm, int order_centered, double c,
                     const double *const *metric, const double *const *Kij, const double *const *prim,
                     double *const *rhs, double *const *madm);

/* matter terms of the Z4c right-hand side (z4c_matter_oracle.c),
 * z4_rhs_add_source at the inner points: u the 25 Z4c fields, madm the 11
 * matter ADM variables (rho, S^i, S_ij, S), r the 25 right-hand sides
 * (A_ij, Khat, G^i, Theta, B^i updated) */
void ob_z4c_matter_source(const ob_box *B, const ob_z4c_params *P, double c, const double *const *u,
                          const double *const *madm, double *const *r);

/* 0d reductions and 1d point lists (output_oracle.c).  var[v * nb + b]: the
 * array of variable v on box b (nv <= 64).  ob_line_1d: the points of the x,
 * y or z line (line 1..3), their coordinate along the line, the box holding
 * each (-1: none) and the index in that box; returns the point count. */
void ob_reduce_0d(int nb, const ob_box *boxes, const double *const *var, int nv, int type, double *result);
int ob_line_1d(int nb, const ob_box *boxes, int line, double x0, double y0, double z0, int maxp, double *coord,
               int *box_of, long *index);

/* moving-puncture tracking (puncture_oracle.c): interpolate_xyz_scalar on a
 * level of nb boxes (var[b] = the variable on box b; found = 0 and value 0
 * when no box holds the interpolation cube), and one step of
 * track_moving_puncture_integrate with beta[3*nb] / beta_p[3*nb] the shift
 * components of the current / previous time level, box by box; flags bit 0
 * keeps z (moving_puncture_fixz BH); a step that leaves every box keeps the
 * old position (track_puncture) */
double ob_interp_point(int nb, const ob_box *boxes, const double *const *var, double x, double y, double z, int order,
                       int *found);
void ob_track_puncture(int nb, const ob_box *boxes, const double *const *beta, const double *const *beta_p, double dt,
                       int order, int flags, const double pold[3], double pnew[3]);

#ifdef __cplusplus
}
#endif
#endif
