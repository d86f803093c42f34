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
#define ADCGPU_E_CAPACITY (-5)    /* the mesh exceeds the kernels' 32-bit offsets, or a JCG
                                     block limit below two */
#define ADCGPU_E_NONFINITE (-6)   /* a non-finite velocity (UU2 or VV2) at some node: the
                                     reference's NaN velocity check, src/timestep.F:1413-1471 */
#define ADCGPU_E_ERRORELEV (-7)   /* |ETA2| above ErrorElev at a wet node: the reference
                                     stops the run, src/timestep.F:1242-1251, 1381-1395 */

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
                               extrapolated from the last two, 2*ETAS - ETAS_prev; 2 the
                               better of ETAS and the learned guess (adcgpu_ml_*).  1 and
                               2 keep the stopping test: results within the JCG
                               tolerance, not bit-identical to the reference */
  double errorelev;         /* ErrorElev (fort.15 namelist WarnElevControl,
                               src/read_input.F:794-812); 0 selects the reference's
                               default, 1000 m */
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
  ADCGPU_ETAS_PREV, /* the increment before ETAS (kept when warm_start != 0 or
                       the learned start is set up) */
  ADCGPU_TIP1,      /* tidal potential at time levels s and s+1 (only after */
  ADCGPU_TIP2,      /* adcgpu_set_tide_potential; TIP1, TIP2 in src/global.F) */
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
 * subset (3D, sponge, subgrid, weirs, flow-boundary codes above 29 other than
 * 30 and 52).  Specified-flux codes (2, 12, 22) carry a zero flux until
 * adcgpu_set_flux_boundary. */
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
/* Iterations of the last JCG solve (NUMITR, src/gwce.F:2005) and its ier.
 * Also reports the sticky error conditions of the steps run so far
 * (ADCGPU_E_DIAG, ADCGPU_E_ERRORELEV, ADCGPU_E_NONFINITE, ADCGPU_E_HIP for a
 * JCG wait that timed out); adcgpu_cold_start clears them. */
int adcgpu_last_numitr(adcgpu_ctx *ctx, int *numitr, int *ier);
/* Waits for the context's stream. */
int adcgpu_synchronize(adcgpu_ctx *ctx);

/* Whole-state snapshot between steps, into / from a host buffer of
 * adcgpu_state_bytes() bytes: every array a later step reads (time levels,
 * flags, maxima, the GWCE matrix and its boundary columns, EP, ITPACK's saved
 * state, the device flags) and the step reached.  Loading it into a context
 * of the same mesh (the same or another warm_start / precision) continues the
 * run exactly as the saved context would: the hot start of this library
 * (the reference's fort.67/68 records, src/write_output.F:4429-5582 and
 * src/hstart.F, hold less: its hot-started run rebuilds the matrix and starts
 * ITPACK's saved state afresh). */
int64_t adcgpu_state_bytes(adcgpu_ctx *ctx);
int adcgpu_state_save(adcgpu_ctx *ctx, void *buf, int64_t bytes);
int adcgpu_state_load(adcgpu_ctx *ctx, const void *buf, int64_t bytes);

/* ---- resident multi-step run (no host round trip per step) ----
 * Runs `nsteps` steps from it0 with elevation-boundary values taken from
 * eta_bc_table[(it-it0)*NETA + I] (host pointer, copied once) and, when
 * holland != NULL, met forcing generated on the device by a Holland vortex
 * (13 parameters, see adcgpu_holland).  use_graph must be 0 (the steps are
 * issued as stream launches; ADCGPU_E_UNSUPPORTED otherwise).
 * The reference's run-time checks are made on the device every step and read
 * every 64 steps and at the end: ErrorElev (ADCGPU_E_ERRORELEV) and
 * non-finite velocities (ADCGPU_E_NONFINITE); the run returns the error at
 * the next read, and adcgpu_run_stats gives the first step that raised it. */
int adcgpu_run(adcgpu_ctx *ctx, int it0, int nsteps, const double *eta_bc_table,
               const double *holland, int use_graph);
/* Statistics of the last adcgpu_run: out[0] JCG iterations, out[1] GWCE
 * matrix rebuilds, out[2] solves that reached ITMAX (ier = 13; the reference
 * continues, src/gwce.F:2003-2005), out[3] the first step that raised a
 * run-time error (0 if none), out[4] solves that started from the learned
 * guess (warm_start = 2). */
int adcgpu_run_stats(adcgpu_ctx *ctx, int64_t *out5);
/* The last JCG solve's end: out[0] the stopping-test value STPTST it ended
 * with and out[1] the bound ZETA it was held to (pstop_nrms, src/itpackv.F:809-890:
 * converged when out[0] < out[1]); out[2], out[3] the squared norms of the
 * initial pseudo-residuals of the reference's start and of the learned start
 * (warm_start = 2; the smaller one was used), 0 otherwise.  Synchronises the
 * context's stream.  No counterpart in the reference: a reporting call. */
int adcgpu_last_solve(adcgpu_ctx *ctx, double *out4);
/* Limits the blocks of the context's JCG launches to `blocks` (>= 2; 0: as
 * many as are resident on the device, the default).  A system with more
 * 4,096-row tiles than the blocks allow, by this limit or beyond 1,044,480
 * rows on MI355X, is solved memory-resident (iterates in HBM, products
 * gathered, two grid barriers per iteration) with the same results.  Not
 * with warm_start = 2, nor for contexts that belong to a batch.  No
 * counterpart in the reference: for sharing the device and for tests of the
 * memory-resident solve. */
int adcgpu_set_jcg_blocks(adcgpu_ctx *ctx, int blocks);

/* Tidal potential forcing, NTIP = 1 (replaces applyTidePotentialForcing,
 * src/timestep.F:1507-1562, with CTIP's terms in src/gwce.F:1181-1185 and
 * src/momentum.F:443-450).  ntif constituents as fort.15 gives them (TPK,
 * AMIGT rad/s, ETRF, FFT) with FACET in radians; slam/sfea: every node's
 * longitude and latitude in radians (SLAM, SFEA); the ramp as NRAMP and
 * DRampTip (days), and STATIM/REFTIM (days) for the step's times.  From the
 * next adcgpu_begin_step on, each step moves TIP2 to TIP1 and forms TIP2 on
 * the device (the C library's cos and the device's differ in the last bit of
 * some arguments; a caller that needs the host's values overwrites
 * ADCGPU_TIP2 after adcgpu_begin_step).  ntif = 0 turns the forcing off.
 * Keeps the current TIP1/TIP2 at zero, as a cold start does. */
int adcgpu_set_tide_potential(adcgpu_ctx *ctx, int ntif, const double *tpk, const double *amigt,
                              const double *etrf, const double *fft, const double *facet,
                              const double *slam, const double *sfea, int nramp,
                              double dramptip, double statim, double reftim);

/* Hybrid bottom friction (NOLIBF = 2): fort.15's HBREAK, FTHETA, FGAMMA, the
 * break depth and exponents of FRIC*(1 + (HBREAK/H)**FTHETA)**(FGAMMA/FTHETA)
 * (src/nodalattr.F:2400-2411; the wetting criterion's friction too,
 * src/wetdry.F:415-451).  A context created with NOLIBF = 2 starts from
 * HBREAK = 1, FTHETA = 10, FGAMMA = 1/3.  The device's pow can differ from
 * the C library's in the last bit, so hybrid runs agree with the reference
 * to rounding, not bit for bit. */
int adcgpu_set_hybrid_friction(adcgpu_ctx *ctx, double hbreak, double ftheta, double fgamma);

/* Specified normal flux boundaries (NFLUXF = 1; LBCODEI 2, 12 and 22 among the
 * land codes, and 52, river radiation with no flux-settling time: its
 * reference elevation is the cold-start ETA2, src/cstart.F:1188-1192) and
 * outward-radiating ones (NFLUXRBC = 1; LBCODEI 30).  30 and 52 need this
 * call before the first step.  bndlen2o3: BNDLEN2O3(NVEL), two thirds of each flow-boundary
 * edge to the next entry, zero after each segment's last node, as the
 * reference's mesh set-up forms it (src/mesh.F:2300-2303); NULL turns the
 * flux boundaries off.  From then on each adcgpu_begin_step shifts
 * QN0 <- QN1 <- QN2 and zeroes QN2 (src/timestep.F:605-612), the GWCE takes
 * the flux line integrals of every boundary edge (src/gwce.F:1737-1834,
 * codes up to 29) and momentum imposes -QN2/H2 as the normal velocity of
 * codes 0-19 (src/momentum.F:810-848).  QN0..QN2 start at zero.
 * Radiating entries take QN1 from the last velocities each step
 * (src/timestep.F:929-940) and the radiation QFORCE (src/gwce.F:1750-1754);
 * they need no QN2. */
int adcgpu_set_flux_boundary(adcgpu_ctx *ctx, const double *bndlen2o3);
/* This step's QN2(NVEL) (src/timestep.F:861-903: the caller's periodic or
 * fort.20 flux, e.g. from adcgpu_flux_bc), host pointer; between
 * adcgpu_begin_step and adcgpu_solve_gwce.  Entries whose code takes no
 * specified flux should hold zero, as in the reference. */
int adcgpu_set_flux(adcgpu_ctx *ctx, const double *qn2);
/* QN2 rows for the resident runs (adcgpu_run, adcgpu_batch_run) of a context
 * with flux boundaries: qn2_table[s*NVEL + I] is QN2 of the run's step
 * it0 + s (host pointer, copied at once); a run of more steps than the table
 * holds is refused. */
int adcgpu_set_flux_table(adcgpu_ctx *ctx, int nsteps, const double *qn2_table);
/* Periodic specified normal flux at one time (src/timestep.F:861-878, ICS = 1):
 * QN2(I) = sum over constituents J of QNAM(J,I)*FFF(J)*RampExtFlux*
 * cos(FAMIG(J)*(timeh - NCYC*FPER(J)) + FFACE(J) - QNPH(J,I)) at the entries
 * whose code is 2, 12, 22 or 52, zero elsewhere; qnam/qnph are (NFFR, NVEL)
 * column-major, fface/qnph in radians.  Host computation (C library cos). */
int adcgpu_flux_bc(int nvel, const int32_t *lbcodei, int nffr, const double *famig,
                   const double *fff, const double *fface, const double *qnam,
                   const double *qnph, double timeh, double rampextflux, double *out);

/* ---- batches: several contexts stepped together (ensembles, uneven meshes)
 * No counterpart in the reference (an ensemble there is separate runs of
 * the TIMESTEP loop, src/timestep.F).  The members are contexts on one
 * device, of the same mesh or of different meshes; each keeps its own
 * state, matrices and wet/dry flags.  adcgpu_batch_run advances all of them
 * by nsteps steps: every member's passes are issued in turn on the batch's
 * stream, and the members' JCG solves run together as one cooperative launch
 * (more than one when their blocks would not all be resident at once), each
 * system's blocks synchronising only among themselves and leaving at its own
 * convergence.  Each member's results are bit-identical to adcgpu_run on that
 * member alone.  eta_bc_tables[k] and holland[k] are what adcgpu_run takes for
 * member k (either array may be NULL when no member needs it);
 * adcgpu_run_stats then reports each member's run.  Members must not be used
 * by another thread during a batch run. */
typedef struct adcgpu_batch adcgpu_batch;
int adcgpu_batch_create(adcgpu_ctx *const *members, int n, adcgpu_batch **out);
int adcgpu_batch_run(adcgpu_batch *batch, int it0, int nsteps,
                     const double *const *eta_bc_tables, const double *const *holland);
void adcgpu_batch_destroy(adcgpu_batch *batch);

/* ---- learned JCG start (warm_start = 2) ----
 * No counterpart in the reference: these replace only the initial guess of
 * CALL JCG (ETAS, src/gwce.F:2002-2004); the stopping test
 * (src/itpackv.F:809-890) and everything after it are unchanged.  A per-node
 * network, 48 inputs -> 128 -> 128 -> 1 with GELU, BF16 weights and FP32
 * accumulation, predicts the correction to ETAS from the node's stencil (the
 * node and up to 12 NEITAB neighbours: last two increments, elevation
 * differences), depth, wind stress, pressure change, friction and wet/dry
 * flags.  Each solve starts from whichever of ETAS and ETAS + correction has
 * the smaller initial pseudo-residual.  Training runs on the device, on
 * samples of the model's own solves. */
#define ADCGPU_ML_NPARAM 22913 /* W1[48][128], b1[128], W2[128][128], b2, w3[128], b3 (FP32) */
/* Set (in != NULL; the Adam moments restart) and/or read (out != NULL) the
 * FP32 parameters.  A new context starts from a fixed-seed Glorot draw with a
 * zero output layer: correction 0, the reference's start. */
int adcgpu_ml_params(adcgpu_ctx *ctx, const float *in, float *out);
/* From the next step on, record per_step node snapshots of every solve as
 * training samples (features before the solve, target (ETAS_new - ETAS) / s
 * after it), in a ring of 2^20; 0 stops. */
int adcgpu_ml_collect(adcgpu_ctx *ctx, int per_step);
int64_t adcgpu_ml_samples(adcgpu_ctx *ctx);
/* Adam (lr) over the held samples, minibatches of 4096 in a seeded random
 * order, loss = sum w (y - t)^2 / 4096 with w = s^2 / mean(s^2); the mean
 * loss of the last epoch in *loss. */
int adcgpu_ml_train(adcgpu_ctx *ctx, int epochs, double lr, uint64_t seed, double *loss);
/* Loss and gradient (ADCGPU_ML_NPARAM floats, the parameters' order) of one
 * batch of n <= 4096 samples: x[n][48] in the kernels' input order, targets y,
 * weights w, loss = sum w (y - t)^2 / n.  No update.  Overwrites the first n
 * samples of the ring. */
int adcgpu_ml_grad(adcgpu_ctx *ctx, int n, const float *x, const float *y, const float *w,
                   float *grad, double *loss);
/* The network's inputs for every node of the current state (x[NP][48], BF16
 * values, the kernels' input order) and each node's scale s; overwrites the
 * first NP samples of the ring. */
int adcgpu_ml_features(adcgpu_ctx *ctx, float *x, double *s);
/* The correction s * y the network gives every node of the current state
 * (0 on dry and elevation-boundary nodes). */
int adcgpu_ml_correction(adcgpu_ctx *ctx, double *corr);
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

/* ---- drop-in JCG that allocates nothing ----
 * No counterpart in the reference: for callers that cannot allocate memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, semantics and return values as adcgpu_jcg, for
 * systems of at most ADCGPU_JCG_NOALLOC_MAX_N rows and
 * ADCGPU_JCG_NOALLOC_MAX_NZ columns (larger ones return ADCGPU_E_ARG before
 * touching the device).  Every device buffer lives in one static device array
 * of ADCGPU_JCG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends;
 * the call makes no hipMalloc/hipFree and no host allocation.  The array keeps
 * what the last call left, ITPACK's saved state included (it carries from call
 * to call on a device, separately from adcgpu_jcg's).  Calls that overlap on
 * one GPU race on the array.  Work runs on the null stream. */
#define ADCGPU_JCG_NOALLOC_MAX_N 32768
#define ADCGPU_JCG_NOALLOC_MAX_NZ 16
#define ADCGPU_JCG_NOALLOC_SCRATCH_BYTES 17999616
int adcgpu_jcg_noalloc(int n, int ndim, int maxnz, int32_t *jcoef, double *coef, double *rhs,
                       double *u, int32_t *iparm, double *rparm);

/* Upper bound, in seconds of device wall clock, of every wait inside the
 * persistent JCG kernel (grid barrier, neighbour flags); default 0.5.  A wait
 * that runs out sets an error flag that every block sees, so the launch drains
 * and the call returns ADCGPU_E_HIP.  No counterpart in the reference; for
 * failure-handling tests (a tiny bound forces the error path from the host). */
int adcgpu_set_wait_limit(double seconds);

/* ---- synthetic forcing ---- */
/* Holland (1980) vortex wind stress / rho0 and pressure head at time t on
 * the context's nodes, written into WSX2, WSY2, PR2 on the device. */
int adcgpu_holland(adcgpu_ctx *ctx, double t, const double *prm13);

#ifdef __cplusplus
}
#endif
#endif
