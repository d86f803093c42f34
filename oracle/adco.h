/*
 * adco.h -- CPU restatement ("oracle") of the reference's 2DDI per-timestep
 * hot path.  TEST INFRASTRUCTURE ONLY: the product never links or calls this.
 *
 * Every function here follows a reference routine statement by statement, in
 * the reference's evaluation order, so that it is bit-identical to the
 * reference compiled without FMA contraction:
 *
 *   adco_jcg            src/itpackv.F:63-456   (jcg, itjcg, chgcon, eigvss,
 *                                               zbrent, determ, pstop_nrms,
 *                                               parcon, pjac, scal, unscal,
 *                                               peror, sdot)
 *   adco_step           src/timestep.F:231-237,257-431,577-593,1069-1103
 *     friction          src/nodalattr.F:2387-2411, 2700-2705
 *     GWCE LHS/RHS/BC   src/gwce.F:419-743, 1000-1862, 2000-2008
 *     wet/dry           src/wetdry.F:140-1470 (+ computeAlpha :1561-1652)
 *     momentum          src/momentum.F:298-1182
 *     tidal potential   src/timestep.F:412-413, 841, 1507-1562 (NTIP = 1)
 *     peak surge        src/write_output.F:6357-6366
 *
 * A second reduction order (ADCO_ORDER_BLOCKED) replaces the reference's
 * sequential sums (EP, JCG inner products) by the fixed blocked tree the GPU
 * kernels use; everything else is unchanged, so the same state can be checked
 * bit-for-bit against the GPU as well.
 */
#ifndef ADCO_H
#define ADCO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ADCO_ORDER_SEQUENTIAL = 0, ADCO_ORDER_BLOCKED = 1 };

/* Blocked reduction shape shared with the GPU kernels: 2048-element chunks
 * of 512 threads x 4 elements; the chunk values are combined by 256 threads. */
#define ADCO_RED_THREADS 512
#define ADCO_RED_PER_THREAD 4
#define ADCO_RED_CHUNK (ADCO_RED_THREADS * ADCO_RED_PER_THREAD)
#define ADCO_RED_XTHREADS 256

/* ITPACK's COMMON /itcom1-3/ plus the SAVEd locals of itjcg and chgcon. */
typedef struct {
  int in, is, isym, itmax, level;
  int adapt, betadt, caseii, halt, partad;
  double bdelnm, betab, cme, delnnm, delsnm, ff, gamma, omega, qa, qt, rho,
      rrr, sige, sme, specr, spr, srelpr, stptst, udnm, zeta;
  /* SAVE in itjcg */
  int q1;
  double gamold, rhoold;
  /* SAVE in chgcon */
  double cmold;
  int order; /* ADCO_ORDER_* used by the inner products */
} adco_itpack;

void adco_itpack_init(adco_itpack *s, int order);
void adco_dfault(int *iparm, double *rparm);
/* JCG on an ELLPACK system: coef(ndim,maxnz), jcoef(ndim,maxnz) column-major,
 * 1-based column indices (0 = padding), exactly the reference contract. */
int adco_jcg(adco_itpack *s, int n, int ndim, int maxnz, int *jcoef,
             double *coef, double *rhs, double *u, int nw, double *wksp,
             int *iparm, double *rparm);

double adco_dot(int order, int n, const double *x, const double *y);

/* ---------------------------------------------------------------- model */
typedef struct {
  /* sizes */
  int np, ne, mnei, neta, nvel, nvelme, nbfr;
  /* mesh, as produced by the reference's mesh set-up (1-based node ids) */
  const int *nm;          /* [ne][3] */
  const double *x, *y, *dp;
  const double *fdxe;     /* [ne][3] */
  const double *fdye;     /* [ne][3] */
  const double *areas;    /* [ne] (twice the element area) */
  const int *nneigh;      /* [np] */
  const int *neitab;      /* column-major (np, mnei), 1-based, 0 padding */
  const int *csr;         /* [ne][3][2] CSR_POSITION */
  const int *nbd;         /* [neta] */
  const int *nbv;         /* [nvel] */
  const int *lbcodei;     /* [nvel] */
  const double *csii, *siii; /* [nvel] */
  const int *me2gw;       /* [nvelme] */
  const double *emo, *efa; /* column-major (nbfr, neta), efa in radians */
  const double *amig, *ff, *face; /* [nbfr], face in radians */
  const double *tau0var, *evm, *fric, *corif; /* [np] */
  /* parameters */
  double dtdp, statim, reftim, g, h0, velmin, a00, b00, c00, convcr;
  double drampelev, prbckgrnd_mh2o;
  int nramp, nolifa, nolica, nolicat, nolibf, itmax, nws;
  /* initial state (cold start) */
  const int *nodecode0, *noff0;
  const double *eta2_0;
  /* GWCE LHS: 0 consistent (JCG), 1 lumped diagonal (explicit solve) */
  int ilump;
  /* hybrid friction (NOLIBF = 2): HBREAK, FTHETA, FGAMMA of fort.15 */
  double hbreak, ftheta, fgamma;
} adco_setup;

typedef struct adco_model adco_model;

adco_model *adco_create(const adco_setup *s, int order);
void adco_destroy(adco_model *m);
/* One timestep.  wsx2/wsy2/pr2 are this step's met forcing (NULL if nws==0).
 * Returns the JCG error code of the step's solve. */
int adco_step(adco_model *m, int it, const double *wsx2, const double *wsy2,
              const double *pr2);
/* Access to state arrays by name ("eta1","eta2","etas","uu2","vv2","qx2",
 * "qy2","tk","etamax","umax","totalarea","coef","gwce_lv","h2","alphal"). */
double *adco_array(adco_model *m, const char *name);
int *adco_iarray(adco_model *m, const char *name); /* nodecode, noff, mju, etamax_it, umax_it */
int adco_numitr(const adco_model *m);
int adco_ncchange(const adco_model *m);
uint64_t adco_state_hash(const adco_model *m);

/* Tidal potential (NTIP = 1): the constituents (FACET in radians), the nodes'
 * longitude and latitude SLAM/SFEA in radians.  From the next step on, TIP2 is
 * formed per step and enters the GWCE and momentum pressure terms
 * (src/timestep.F:412-413, 841, 1507-1562; src/gwce.F:1181-1185;
 * src/momentum.F:443-450).  Returns 0, or -1 for NTIF outside 1..ADCO_MAXTIF. */
#define ADCO_MAXTIF 32
int adco_set_tidepot(adco_model *m, int ntif, const double *tpk, const double *amigt,
                     const double *etrf, const double *fft, const double *facet,
                     const double *slam, const double *sfea);
/* RampTip of the last step */
double adco_ramptip(const adco_model *m);

/* Specified normal flux boundaries (NFLUXF = 1; LBCODEI 2, 12, 22 among land
 * codes, all <= 29): BNDLEN2O3(NVEL) from the mesh set-up (src/mesh.F:2303).
 * Each later step shifts QN0 <- QN1 <- QN2 (src/timestep.F:605-612), takes QN2
 * from the last adco_set_qn2 (zero without one), adds the flux line integrals
 * to GWCE_LV (src/gwce.F:1737-1834) and uses -QN2/H2 as the normal velocity of
 * codes 0-19 (src/momentum.F:810-848).  Returns 0, or -1 for a code above 29. */
int adco_set_flux(adco_model *m, const double *bndlen2o3);
void adco_set_qn2(adco_model *m, const double *qn2);
/* QN2 of periodic flux forcing at one time (src/timestep.F:861-878, ICS = 1):
 * qnam/qnph (NFFR, NVEL) column-major, FFACE and QNPH in radians. */
void adco_flux_bc(int nvel, const int *lbcodei, int nffr, const double *famig,
                  const double *fff, const double *fface, const double *qnam,
                  const double *qnph, double timeh, double rampextflux, double *out);

/* elevation boundary value for one boundary node at one time (gwce.F:1638-1651) */
void adco_eta_bc(const adco_setup *s, double timeh, double rampelev, double *out);

/* ------------------------------------------------------------- utilities */
int64_t adco_fnv1a(const void *p, int64_t nbytes, int64_t seed);
/* Holland (1980) vortex: wind stress / rho0 (m^2/s^2) and pressure head (m).
 * prm: [0] Pc hPa, [1] Pn hPa, [2] Rmax m, [3] B, [4] x0, [5] y0, [6] vx,
 *      [7] vy, [8] f, [9] ramp days (0 = none), [10] rho_air,
 *      [11] surface wind factor, [12] inflow angle deg */
void adco_holland(int np, const double *x, const double *y, double t,
                  const double *prm, double *wsx, double *wsy, double *pr);

#ifdef __cplusplus
}
#endif
#endif
