/*
 * adcgpu.h -- C ABI of the MI355X (gfx950) implementation of ADCIRC's 2DDI
 * per-timestep hot path: GWCE assembly, ITPACKV Jacobi-CG, wet/dry
 * classification and the explicit momentum update.
 *
 * The reference has no foreign-function layer on this path: its timestep
 * driver calls Fortran module procedures that work on module arrays.  These
 * entry points are what an ISO_C_BINDING interface for those call sites binds
 * (see fortran/adcgpu_iface.F90 and INTEGRATION.md):
 *
 *   adcgpu_jcg                 <- CALL JCG(...)           src/itpackv.F:63-64,
 *                                                         called at src/gwce.F:2003-2004
 *   adcgpu_solve_gwce          <- CALL solveGWCE(...)     src/timestep.F:1069 -> src/gwce.F:166
 *   adcgpu_wetdry              <- CALL computeWettingAndDrying(IT)
 *                                                         src/timestep.F:1088 -> src/wetdry.F:140
 *   adcgpu_momentum            <- CALL solveMomentumEq()  src/timestep.F:1103 -> src/momentum.F:118
 *   adcgpu_friction            <- Apply2DBottomFriction + Apply2DInternalWaveDrag
 *                                                         src/timestep.F:577-588
 *   adcgpu_begin_step          <- time-level rotation     src/timestep.F:385-431
 *   adcgpu_timestep            <- the 2DDI subset of TIMESTEP, src/timestep.F:55-1162
 *
 * Conventions (all mirror the reference):
 *   - node and element ids are 1-based; NM is (NE,3) and NEITAB (NP,MNEI)
 *     column-major exactly as in src/mesh.F; CSR_POSITION is (NE,3,2).
 *   - every call returns 0 on success or a negative ADCGPU_E_* code; JCG
 *     returns ITPACK's ier (0, 13 = itmax reached) through its own argument.
 *   - one host thread per context; all device work runs on the context's
 *     stream; a context owns one copy of the state on one GPU.
 *   - arrays passed to adcgpu_create are copied; the caller keeps ownership.
 */
#ifndef ADCGPU_H
#define ADCGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADCGPU_OK 0
#define ADCGPU_E_ARG (-1)         /* invalid argument or size */
#define ADCGPU_E_UNSUPPORTED (-2) /* configuration outside the implemented subset */
#define ADCGPU_E_HIP (-3)         /* HIP runtime error (message via adcgpu_last_error) */
#define ADCGPU_E_DIAG (-4)        /* GWCE diagonal <= 0 (src/gwce.F:678-707 terminates) */
#define ADCGPU_E_CAPACITY (-5)    /* the mesh has more rows than the resident JCG grid holds */
#define ADCGPU_E_NONFINITE (-6)   /* non-finite JCG residual */

typedef struct adcgpu_ctx adcgpu_ctx;

/* Mesh and boundary tables, as built by the reference's mesh set-up
 * (src/mesh.F: initializeMesh :1458, NEIGHB :3089, initializeBoundaries :1769,
 * compute_csr_positions :5261). */
typedef struct {
  int32_t np, ne, mnei, neta, nvel, nvelme;
  const int32_t *nm;        /* NM(NE,3), column-major, 1-based */
  const double *x, *y, *dp; /* X, Y, DP (after initializeMesh) */
  const double *fdxe, *fdye;/* FDXE(3,NE), FDYE(3,NE), column-major */
  const double *areas;      /* AREAS(NE) = 2 x element area */
  const int32_t *nneigh;    /* NNEIGH(NP) including the node itself */
  const int32_t *neitab;    /* NEITAB(NP,MNEI), column-major, 1-based, 0 padded */
  const int32_t *csr_position; /* CSR_POSITION(NE,3,2), column-major */
  const int32_t *nbd;       /* NBD(NETA) elevation-boundary nodes */
  const int32_t *nbv;       /* NBV(NVEL) flow-boundary nodes */
  const int32_t *lbcodei;   /* LBCODEI(NVEL) */
  const double *csii, *siii;/* CSII(NVEL), SIII(NVEL) boundary normals */
  const int32_t *me2gw;     /* ME2GW(NVELME) */
} adcgpu_mesh;

/* Run parameters (fort.15 subset; src/read_input.F). */
typedef struct {
  double dtdp, g, h0, velmin, a00, b00, c00, convcr, prbckgrnd_mh2o;
  int32_t nolifa, nolica, nolicat, nolibf, itmax, nws;
  int32_t precision;        /* ADCGPU_FP64 or ADCGPU_MIXED */
  const double *tau0var;    /* Tau0Var(NP) */
  const double *evm;        /* EVM(NP) lateral eddy viscosity */
  const double *fric;       /* FRIC(NP) bottom friction coefficient */
  const double *corif;      /* CORIF(NP) */
  int32_t ilump;            /* ILump: 0 consistent GWCE LHS solved by JCG, 1 lumped
                               diagonal LHS solved explicitly (src/gwce.F:141-155) */
  int32_t warm_start;       /* JCG initial guess: 0 the previous increment ETAS, as the
                               reference (src/gwce.F:2002-2004); 1 the increment
                               extrapolated from the last two, 2*ETAS - ETAS_prev (same
                               stopping test, results within the JCG tolerance, not
                               bit-identical to the reference) */
} adcgpu_params;

#define ADCGPU_FP64 0
#define ADCGPU_MIXED 1

/* State fields exchanged with the host (ADCIRC module arrays, src/global.F). */
enum adcgpu_field {
  ADCGPU_ETA1 = 0, ADCGPU_ETA2, ADCGPU_ETAS, ADCGPU_UU1, ADCGPU_VV1, ADCGPU_UU2,
  ADCGPU_VV2, ADCGPU_QX1, ADCGPU_QY1, ADCGPU_QX2, ADCGPU_QY2, ADCGPU_H1, ADCGPU_H2,
  ADCGPU_TK, ADCGPU_ETAMAX, ADCGPU_TOTALAREA, ADCGPU_WSX1, ADCGPU_WSY1, ADCGPU_WSX2,
  ADCGPU_WSY2, ADCGPU_PR1, ADCGPU_PR2, ADCGPU_GWCE_LV,
  ADCGPU_UMAX,      /* peak sqrt(UU2^2+VV2^2) at wet nodes, -99999 before any
                       (src/write_output.F:6367-6376) */
  ADCGPU_NODECODE = 100, ADCGPU_NOFF, ADCGPU_MJU,
  ADCGPU_ETAMAX_IT, /* step IT of each node's last ETAMAX rise, 0 if none: the
                       reference's ETAMAX_Time is IT*DTDP + STATIM*86400
                       (src/write_output.F:6357-6364) */
  ADCGPU_UMAX_IT,   /* step IT of each node's last UMAX rise, 0 if none (UMAX_Time) */
  ADCGPU_COEF = 200, /* COEF(NP,MNEI) column-major */
  ADCGPU_ALPHAL = 300
};

/* Create a context on `device` (HIP ordinal) from the mesh and parameters.
 * Returns ADCGPU_E_UNSUPPORTED for configurations outside the implemented
 * subset (hybrid friction, 3D, sponge, subgrid, weirs). */
int adcgpu_create(const adcgpu_mesh *mesh, const adcgpu_params *prm, int device,
                  adcgpu_ctx **out);
void adcgpu_destroy(adcgpu_ctx *ctx);
const char *adcgpu_last_error(void);

/* Copy a state field between host and device (count elements of its type). */
int adcgpu_set_field(adcgpu_ctx *ctx, int field, const void *host, int64_t count);
int adcgpu_get_field(adcgpu_ctx *ctx, int field, void *host, int64_t count);
/* Cold start: eta2 (= eta1), nodecode and noff from the reference's
 * setColdWetDryStateAndWaterSurfaceElevation, TotalArea/MJU computed here
 * (src/mesh.F:3360 totalAreaCalc), NCChange = 1.  All other state starts
 * cold as well (velocities, fluxes, ETAS, TK, forcing at rest, ALPHAL = 1,
 * the peak-surge and peak-speed maxima, the JCG saved state), so a context
 * may be cold-started again. */
int adcgpu_cold_start(adcgpu_ctx *ctx, const double *eta2, const int32_t *nodecode,
                      const int32_t *noff);

/* ---- step-level entry points (the reference's per-step call sites) ---- */
/* time-level rotation of U, V, Q, WS, PR (src/timestep.F:385-431); it==1 also
 * sets H1 = H2 = DP + IFNLFA*ETA2 (src/timestep.F:231-237) */
int adcgpu_begin_step(adcgpu_ctx *ctx, int it);
int adcgpu_friction(adcgpu_ctx *ctx);
/* This step's met forcing (WSX2, WSY2, PR2; NP each), host pointers. */
int adcgpu_set_forcing(adcgpu_ctx *ctx, const double *wsx2, const double *wsy2,
                       const double *pr2);
/* GWCE_New: LHS when NCChange > 0, RHS, elevation BCs, JCG, ETA2.  eta_bc are
 * this step's elevation-boundary values at NBD (src/gwce.F:1638-1651 summed
 * over constituents, host libm), NETA doubles. */
int adcgpu_solve_gwce(adcgpu_ctx *ctx, int it, const double *eta_bc);
int adcgpu_wetdry(adcgpu_ctx *ctx, int it);
int adcgpu_momentum(adcgpu_ctx *ctx);
/* The whole 2DDI step in the reference's order: begin_step, friction, met
 * forcing (wsx2/wsy2/pr2 host pointers, required when NWS != 0, else NULL),
 * solve_gwce, wetdry (H2 only when NOLIFA != 2), momentum. */
int adcgpu_timestep(adcgpu_ctx *ctx, int it, const double *eta_bc, const double *wsx2,
                    const double *wsy2, const double *pr2);
/* Iterations of the last JCG solve (NUMITR, src/gwce.F:2005) and its ier. */
int adcgpu_last_numitr(adcgpu_ctx *ctx, int *numitr, int *ier);
/* Waits for the context's stream. */
int adcgpu_synchronize(adcgpu_ctx *ctx);

/* ---- resident multi-step run (no host round trip per step) ----
 * Runs `nsteps` steps from it0 with elevation-boundary values taken from
 * eta_bc_table[(it-it0)*NETA + I] (host pointer, copied once) and, when
 * holland != NULL, met forcing generated on the device by a Holland vortex
 * (13 parameters, see adcgpu_holland).  Each step is one captured hipGraph
 * replay when use_graph != 0. */
int adcgpu_run(adcgpu_ctx *ctx, int it0, int nsteps, const double *eta_bc_table,
               const double *holland, int use_graph);
/* Sum of JCG iterations over the last adcgpu_run. */
int64_t adcgpu_run_iterations(adcgpu_ctx *ctx);
/* Steps of the last adcgpu_run that rebuilt the GWCE matrix (NCChange > 0,
 * src/gwce.F:427). */
int64_t adcgpu_run_rebuilds(adcgpu_ctx *ctx);

/* Timing of the JCG launches with HIP events on the context's stream:
 * adcgpu_profile(ctx, n) with n > 0 clears and starts recording (event pairs
 * for max(n, 1024) launches are created at once, outside any timed region),
 * adcgpu_profile(ctx, 0) stops; adcgpu_profile_read returns the summed kernel
 * time (ms) and the number of launches recorded. */
int adcgpu_profile(adcgpu_ctx *ctx, int enable);
int adcgpu_profile_read(adcgpu_ctx *ctx, double *jcg_ms, int64_t *launches);

/* Launch shape of the JCG solver (no counterpart in the reference; for
 * tuning): shape[0] = blocks per launch (one per 2048-row chunk plus one),
 * shape[1] = the most blocks the device keeps resident at once (the size
 * limit), shape[2] = 2048-row chunks of the system. */
int adcgpu_jcg_shape(adcgpu_ctx *ctx, int *shape3);

/* Elevation-boundary values at the NBD nodes for one time (src/gwce.F:
 * 1638-1651): sum over constituents J of EMO(J,I)*FF(J)*RampElev*cos(AMIG(J)*
 * (timeh - NCYC*PER(J)) + FACE(J) - EFA(J,I)); emo/efa are (NBFR,NETA)
 * column-major, face/efa in radians.  Host computation (C library cos). */
int adcgpu_elevation_bc(int neta, int nbfr, const double *amig, const double *ff,
                        const double *face, const double *emo, const double *efa,
                        double timeh, double rampelev, double *out);

/* ---- drop-in JCG (ITPACKV 2D jcg, src/itpackv.F:63-456) ----
 * Same arguments and semantics as the reference's JCG for iparm(9) = -2,
 * iparm(10) = 0: coef/jcoef are (ndim,maxnz) column-major, jcoef 1-based with
 * 0 padding; coef, rhs and u are scaled and unscaled in place exactly as the
 * reference does, iparm(1) returns the iteration count, rparm(2) the
 * eigenvalue estimate.  ITPACK's saved state (the stopping-test flag q1 and
 * the eigenvalue history) persists across calls on the same device, like the
 * reference's SAVE/COMMON data.  Host pointers; returns ier. */
int adcgpu_jcg(int n, int ndim, int maxnz, int32_t *jcoef, double *coef, double *rhs,
               double *u, int32_t *iparm, double *rparm);

/* ---- synthetic forcing ---- */
/* Holland (1980) vortex wind stress / rho0 and pressure head at time t on
 * the context's nodes, written into WSX2, WSY2, PR2 on the device. */
int adcgpu_holland(adcgpu_ctx *ctx, double t, const double *prm13);

#ifdef __cplusplus
}
#endif
#endif
