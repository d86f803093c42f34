/*
 * One 2DDI barotropic timestep, restated from the reference for the oracle
 * (test infrastructure; never linked into the product).  Order of work and
 * every floating-point expression follow the reference statement by
 * statement; build with -ffp-contract=off.
 *
 * Supported subset (the one the test configurations select): IM = 0
 * (consistent GWCE, Kolar-Gray lateral stress, non-conservative advection,
 * IBPV momentum lateral stress, corrected area integration), Cartesian mesh
 * (ICS = 1: every spherical factor is exactly 1, TANPHI = 0), periodic
 * elevation boundaries, no-normal-flow land boundaries (IBTYPE 0, 10, 20),
 * specified normal flux boundaries (IBTYPE 2, 12, 22), NOLIFA 0/1/2, NOLIBF
 * 0/1/2 with nodal friction, optional met forcing passed in.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "adco.h"

struct adco_model {
  adco_setup s;
  int order;
  int np, ne, mnei, neta, nvel, nvelme, nbfr;
  /* mutable copies */
  int *neitab;
  double *fric;
  /* nodal state */
  double *eta1, *eta2, *etas, *uu1, *vv1, *uu2, *vv2, *qx1, *qy1, *qx2, *qy2;
  double *h1, *h2, *tk, *tkm, *gwce_lv, *mom_lv_x, *mom_lv_y;
  double *wsx1, *wsy1, *wsx2, *wsy2, *pr1, *pr2, *etamax, *totalarea, *auv;
  double *tk_tmp;
  int *nodecode, *nnodecode, *mju, *nibcnt, *etamax_it, *umax_it;
  double *umax;
  /* element state */
  int *noff, *noffold;
  double *alphal;
  /* GWCE system */
  double *coef, *obccoef, ep;
  adco_itpack itp;
  int iparm[12];
  double rparm[12];
  int nw;
  double *wksp;
  int numitr, ncchange, ifnlfa, ifnlct, ifnlcat, ifwind;
  int iflinbf, ifnlbf, ifhybf;
  double hbreak, ftheta, fgamma;
  double habsmin, hoff;
  double dt, dto2, go2;
  double ramp, rampelev, rampmete;
  /* tidal potential (NTIP = 1, src/timestep.F:1507-1562): constituents, the
   * species factors L_N(0:2, I) (src/adcirc.F:295-308), longitudes, TIP1/2 */
  int ctip, ntif;
  double tpk[ADCO_MAXTIF], amigt[ADCO_MAXTIF], etrf[ADCO_MAXTIF], fft[ADCO_MAXTIF];
  double facet[ADCO_MAXTIF], pert[ADCO_MAXTIF];
  double *ln, *slam, *tip1, *tip2;
  /* specified normal flux boundaries (NFLUXF = 1): BNDLEN2O3 and the flux
   * at three time levels QN0/QN1/QN2 per flow-boundary entry
   * (src/global.F, src/boundaries.F); qn_next is the next step's QN2 */
  int nfluxf, qn_pending;
  double *bndlen, *qn0, *qn1, *qn2, *qn_next;
  double *elevdisc; /* river radiation (52): the cold-start elevation (ElevDisc) */
};

#define NM(ie, k) (m->s.nm[(size_t)(ie)*3 + (k)] - 1)
#define FDXE(ie, k) (m->s.fdxe[(size_t)(ie)*3 + (k)])
#define FDYE(ie, k) (m->s.fdye[(size_t)(ie)*3 + (k)])
#define COEF(i, j) m->coef[(size_t)((j)-1) * m->np + (i)]     /* i 0-based */
#define NEITAB(i, j) m->neitab[(size_t)((j)-1) * m->np + (i)] /* 1-based value */
#define OBC(i, j) m->obccoef[(size_t)(j) * m->neta + (i)]   /* i,j 0-based */
#define CSR(ie, k, l) (m->s.csr[((size_t)(ie)*3 + (k)) * 2 + (l)])

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static int *ialloc(size_t n) { return (int *)calloc(n ? n : 1, sizeof(int)); }

adco_model *adco_create(const adco_setup *s, int order) {
  adco_model *m = (adco_model *)calloc(1, sizeof(adco_model));
  m->s = *s;
  m->order = order;
  int np = m->np = s->np, ne = m->ne = s->ne;
  m->mnei = s->mnei;
  m->neta = s->neta;
  m->nvel = s->nvel;
  m->nvelme = s->nvelme;
  m->nbfr = s->nbfr;
  m->neitab = ialloc((size_t)np * s->mnei);
  memcpy(m->neitab, s->neitab, sizeof(int) * (size_t)np * s->mnei);
  m->fric = dalloc(np);
  memcpy(m->fric, s->fric, sizeof(double) * np);
  double **d[] = {&m->eta1, &m->eta2, &m->etas, &m->uu1, &m->vv1, &m->uu2, &m->vv2,
                  &m->qx1, &m->qy1, &m->qx2, &m->qy2, &m->h1, &m->h2, &m->tk,
                  &m->gwce_lv, &m->mom_lv_x, &m->mom_lv_y, &m->wsx1, &m->wsy1,
                  &m->wsx2, &m->wsy2, &m->pr1, &m->pr2, &m->etamax, &m->totalarea,
                  &m->tk_tmp};
  for (size_t k = 0; k < sizeof(d) / sizeof(d[0]); ++k) *d[k] = dalloc(np);
  m->tkm = dalloc((size_t)3 * np);
  m->auv = dalloc((size_t)4 * np);
  m->nodecode = ialloc(np);
  m->nnodecode = ialloc(np);
  m->mju = ialloc(np);
  m->etamax_it = ialloc(np);
  m->umax_it = ialloc(np);
  m->umax = dalloc(np);
  m->nibcnt = ialloc(np);
  m->noff = ialloc(ne);
  m->noffold = ialloc(ne);
  m->alphal = dalloc(ne);
  m->coef = dalloc((size_t)np * s->mnei);
  m->obccoef = dalloc((size_t)(s->neta > 0 ? s->neta : 1) * s->mnei);
  adco_itpack_init(&m->itp, order);
  /* initGWCE, src/gwce.F:123-157 */
  adco_dfault(m->iparm, m->rparm);
  m->iparm[0] = s->itmax;
  m->iparm[1] = 0;  /* ISLDIA */
  m->iparm[3] = 16; /* log unit */
  m->rparm[0] = s->convcr;
  m->nw = 4 * np + 4 * s->itmax;
  m->wksp = dalloc((size_t)m->nw);
  m->dt = s->dtdp;
  m->dto2 = m->dt / 2.0;
  m->go2 = s->g / 2.0;
  m->ifnlfa = s->nolifa == 0 ? 0 : 1;
  m->ifnlct = s->nolica == 0 ? 0 : 1;
  m->ifnlcat = s->nolicat == 0 ? 0 : 1;
  m->ifwind = 1;
  /* InitNodalAttr friction switches, src/nodalattr.F:1974-1988 */
  m->iflinbf = s->nolibf == 0 ? 1 : 0;
  m->ifnlbf = (s->nolibf == 1 || s->nolibf == 3) ? 1 : 0;
  m->ifhybf = s->nolibf == 2 ? 1 : 0;
  m->hbreak = s->hbreak;
  m->ftheta = s->ftheta;
  m->fgamma = s->fgamma;
  /* initializeWettingAndDrying, src/wetdry.F:114-130 */
  m->habsmin = 0.8 * s->h0;
  m->hoff = 1.2 * s->h0;
  /* cold start state */
  for (int i = 0; i < np; ++i) {
    m->eta2[i] = s->eta2_0[i];
    m->eta1[i] = s->eta2_0[i];
    m->nodecode[i] = s->nodecode0[i];
    m->nnodecode[i] = s->nodecode0[i];
    m->pr1[i] = s->prbckgrnd_mh2o;
    m->pr2[i] = s->prbckgrnd_mh2o;
    m->etamax[i] = -99999.0;
    m->umax[i] = -99999.0;
  }
  for (int ie = 0; ie < ne; ++ie) {
    m->noff[ie] = s->noff0[ie];
    m->noffold[ie] = 1;
    m->alphal[ie] = 1.0; /* src/global.F:877 */
  }
  /* totalAreaCalc, src/mesh.F:3360-3400 */
  for (int ie = 0; ie < ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    int ncele = m->nodecode[n1] * m->nodecode[n2] * m->nodecode[n3] * m->noff[ie];
    m->mju[n1] += ncele;
    m->mju[n2] += ncele;
    m->mju[n3] += ncele;
    double areaele = ncele * s->areas[ie] / 2.0;
    m->totalarea[n1] = m->totalarea[n1] + areaele;
    m->totalarea[n2] = m->totalarea[n2] + areaele;
    m->totalarea[n3] = m->totalarea[n3] + areaele;
  }
  for (int i = 0; i < np; ++i)
    if (m->mju[i] == 0) m->mju[i] = 1;
  m->ncchange = 1;
  return m;
}

void adco_destroy(adco_model *m) {
  if (!m) return;
  void *p[] = {m->neitab, m->fric, m->eta1, m->eta2, m->etas, m->uu1, m->vv1, m->uu2,
               m->vv2, m->qx1, m->qy1, m->qx2, m->qy2, m->h1, m->h2, m->tk, m->gwce_lv,
               m->mom_lv_x, m->mom_lv_y, m->wsx1, m->wsy1, m->wsx2, m->wsy2, m->pr1,
               m->pr2, m->etamax, m->totalarea, m->tk_tmp, m->tkm, m->auv, m->nodecode,
               m->nnodecode, m->mju, m->etamax_it, m->umax_it, m->umax, m->nibcnt, m->noff, m->noffold, m->alphal, m->coef,
               m->obccoef, m->wksp, m->ln, m->slam, m->tip1, m->tip2, m->bndlen, m->qn0,
               m->qn1, m->qn2, m->qn_next, m->elevdisc};
  for (size_t k = 0; k < sizeof(p) / sizeof(p[0]); ++k) free(p[k]);
  free(m);
}

int adco_set_tidepot(adco_model *m, int ntif, const double *tpk, const double *amigt,
                     const double *etrf, const double *fft, const double *facet,
                     const double *slam, const double *sfea) {
  if (ntif < 1 || ntif > ADCO_MAXTIF) return -1;
  const int np = m->np;
  m->ctip = 1;
  m->ntif = ntif;
  for (int j = 0; j < ntif; ++j) {
    m->tpk[j] = tpk[j];
    m->amigt[j] = amigt[j];
    m->etrf[j] = etrf[j];
    m->fft[j] = fft[j];
    m->facet[j] = facet[j];
    /* src/read_input.F:3349-3353 */
    m->pert[j] = amigt[j] == 0.0 ? 0.0 : 2.0 * 3.141592653589793 / amigt[j];
  }
  if (!m->ln) {
    m->ln = dalloc(3 * (size_t)np);
    m->slam = dalloc(np);
    m->tip1 = dalloc(np);
    m->tip2 = dalloc(np);
  }
  for (int i = 0; i < np; ++i) {
    /* src/adcirc.F:303-307 */
    m->ln[i] = 1.5 * cos(sfea[i]) * cos(sfea[i]) - 1.0;
    m->ln[np + i] = sin(2.0 * sfea[i]);
    m->ln[2 * (size_t)np + i] = cos(sfea[i]) * cos(sfea[i]);
    m->slam[i] = slam[i];
    m->tip1[i] = 0.0;
    m->tip2[i] = 0.0;
  }
  return 0;
}

/* applyTidePotentialForcing, the sum over constituents without the full
 * formula and without self-attraction/loading (NTIP = 1; SALTAMP is zero,
 * src/read_input.F:6460, so its term adds a zero to a sum that starts at +0
 * and changes no bit) (src/timestep.F:1527-1557) */
static void tide_potential(adco_model *m, double timeh, double ramptip) {
  const int np = m->np;
  for (int j = 0; j < m->ntif; ++j) {
    const int ncyc = m->pert[j] == 0.0 ? 0 : (int)(timeh / m->pert[j]);
    const double argt = m->amigt[j] * (timeh - ncyc * m->pert[j]) + m->facet[j];
    const double tpmul = ramptip * m->etrf[j] * m->tpk[j] * m->fft[j];
    int na = (int)lround(m->amigt[j] / 7e-5);
    if (na > 2) na = 2;
    const double *l = m->ln + (size_t)na * np;
    for (int i = 0; i < np; ++i) {
      const double argtp = argt + na * m->slam[i];
      m->tip2[i] = m->tip2[i] + tpmul * l[i] * cos(argtp);
    }
  }
}

double adco_ramptip(const adco_model *m) { return m->ramp; }

int adco_set_flux(adco_model *m, const double *bndlen2o3) {
  const int nvel = m->nvel;
  if (nvel < 1) return -1;
  /* the codes whose fluxes are specified or zero (LBCODEI <= 29) and the
   * outward radiation (30) */
  for (int k = 0; k < nvel; ++k)
    if (m->s.lbcodei[k] < 0 ||
        (m->s.lbcodei[k] > 29 && m->s.lbcodei[k] != 30 && m->s.lbcodei[k] != 52))
      return -1;
  if (!m->bndlen) {
    m->bndlen = dalloc(nvel);
    m->qn0 = dalloc(nvel);
    m->qn1 = dalloc(nvel);
    m->qn2 = dalloc(nvel);
    m->qn_next = dalloc(nvel);
    m->elevdisc = dalloc(nvel);
  }
  /* river radiation's reference elevation: with no flux-settling time the
   * first step takes the cold-start elevation kept in EtaDisc
   * (src/cstart.F:1188-1192, src/timestep.F:917-925) */
  for (int k = 0; k < nvel; ++k)
    if (m->s.lbcodei[k] == 52) m->elevdisc[k] = m->eta2[m->s.nbv[k] - 1];
  memcpy(m->bndlen, bndlen2o3, sizeof(double) * nvel);
  m->nfluxf = 1;
  m->qn_pending = 0;
  return 0;
}

void adco_set_qn2(adco_model *m, const double *qn2) {
  if (!m->nfluxf) return;
  memcpy(m->qn_next, qn2, sizeof(double) * m->nvel);
  m->qn_pending = 1;
}

/* Periodic specified normal flux (src/timestep.F:861-878, NFFR > 0, ICS = 1 so
 * rotate_normal_flux returns its argument): QN2 starts at zero (the shift of
 * src/timestep.F:605-612) and sums each constituent in turn at the entries
 * whose LBCODEI is 2, 12, 22 or 52.  qnam/qnph are (NFFR, NVEL) column-major,
 * phases in radians. */
void adco_flux_bc(int nvel, const int *lbcodei, int nffr, const double *famig,
                  const double *fff, const double *fface, const double *qnam,
                  const double *qnph, double timeh, double rampextflux, double *out) {
  for (int i = 0; i < nvel; ++i) out[i] = 0.0;
  for (int j = 0; j < nffr; ++j) {
    /* src/read_input.F:3551-3555 */
    const double fper = famig[j] == 0.0 ? 0.0 : 2.0 * 3.141592653589793 / famig[j];
    const int ncyc = fper == 0.0 ? 0 : (int)(timeh / fper);
    const double argj = famig[j] * (timeh - ncyc * fper) + fface[j];
    const double rff = fff[j] * rampextflux;
    for (int i = 0; i < nvel; ++i) {
      const int c = lbcodei[i];
      if (c == 2 || c == 12 || c == 22 || c == 52)
        out[i] = out[i] + qnam[(size_t)i * nffr + j] * rff * cos(argj - qnph[(size_t)i * nffr + j]);
    }
  }
}

/* The specified-flux line integrals into GWCE_LV (src/gwce.F:1737-1834) for
 * boundary codes up to 29: each boundary edge J-1..J adds
 * NCI*NCJ*BNDLEN2O3(J-1)/4 * (2 QFORCE(own) + QFORCE(other)) to both of its
 * nodes, in edge order, I's node first. */
static double qforce(const adco_model *m, int K, int b) {
  const double *t0 = m->s.tau0var;
  if (m->s.lbcodei[K] == 30) {
    /* radiation: src/gwce.F:1750-1754, 1785-1789 (H1 is the saved H1, the
     * last step's H2; ETAS the last increment) */
    const double celerity = sqrt(m->s.g * m->h1[b]);
    return -celerity * m->etas[b] / m->dt - t0[b] * m->qn1[K];
  }
  double q = (m->qn2[K] - m->qn0[K]) / (m->dt * 2.0) + t0[b] * m->qn1[K];
  if (m->s.lbcodei[K] == 52) {
    /* river radiation after the flux-settling steps (none here, IT > 0):
     * src/gwce.F:1767-1775, 1808-1816 */
    const double celerity = sqrt(m->s.g * m->h1[b]);
    q = q - celerity * (m->etas[b] / m->dt + t0[b] * (m->eta1[b] - m->elevdisc[K]));
  }
  return q;
}

static void gwce_flux(adco_model *m) {
  const int *nbv = m->s.nbv;
  int nbdj = nbv[0] - 1;
  double qforcej = qforce(m, 0, nbdj);
  for (int J = 1; J < m->nvel; ++J) {
    const int nbdi = nbdj;
    nbdj = nbv[J] - 1;
    const double qforcei = qforcej;
    qforcej = qforce(m, J, nbdj);
    const int nci = m->nodecode[nbdi], ncj = m->nodecode[nbdj];
    const double w = (double)(nci * ncj) * m->bndlen[J - 1] / 4.0;
    m->gwce_lv[nbdi] = m->gwce_lv[nbdi] + w * (2.0 * qforcei + qforcej);
    m->gwce_lv[nbdj] = m->gwce_lv[nbdj] + w * (2.0 * qforcej + qforcei);
  }
}

/* flang's NORM2: running maximum with scaled sum of squares */
static double norm2_3(const double *v) {
  double mx = 0.0, sum = 0.0;
  for (int k = 0; k < 3; ++k) {
    double a = fabs(v[k]);
    if (mx == 0.0) {
      mx = a;
    } else if (a > mx) {
      double t = mx / a;
      double tsq = t * t;
      sum = sum * tsq;
      sum = sum + tsq;
      mx = a;
    } else {
      double t = a / mx;
      sum = sum + t * t;
    }
  }
  return mx * sqrt(1.0 + sum);
}

/* Apply2DBottomFriction (src/nodalattr.F:2387-2411) + Apply2DInternalWaveDrag
 * without internal-wave drag (src/nodalattr.F:2700-2705) */
static void friction(adco_model *m) {
  const double *dp = m->s.dp;
  for (int i = 0; i < m->np; ++i) {
    double uv1 = sqrt(m->uu1[i] * m->uu1[i] + m->vv1[i] * m->vv1[i]);
    double h1 = dp[i] + m->ifnlfa * m->eta2[i];
    double hyb = m->ifhybf * pow(1 + pow(m->hbreak / h1, m->ftheta), m->fgamma / m->ftheta);
    double a = m->fric[i] * (m->iflinbf + (uv1 / h1) * (m->ifnlbf + hyb));
    double b = (double)2.0f / m->s.dtdp;
    m->tk[i] = fmin(a, b);
  }
  for (int i = 0; i < m->np; ++i) {
    m->tkm[3 * i + 0] = m->tk[i];
    m->tkm[3 * i + 1] = m->tk[i];
    m->tkm[3 * i + 2] = 0.0;
  }
}

/* GWCE LHS, src/gwce.F:419-743 (ILump = 0) */
/* Lumped LHS (ILump = 1): the diagonal COEFD only, src/gwce.F:596-635,
 * with EP, the boundary rows and the zero-diagonal fix of :709-743; COEFD
 * is kept in COEF's diagonal column. */
static int gwce_lhs_lumped(adco_model *m) {
  const int np = m->np, ne = m->ne;
  const double OnDiag = 4.0, DT = m->dt; /* (1+ILump)*2 */
  memset(m->coef, 0, sizeof(double) * (size_t)np * m->mnei);
  const double *t0 = m->s.tau0var;
  for (int ie = 0; ie < ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    double AreaIE2 = m->s.areas[ie];
    double AreaIE = AreaIE2 * 0.5;
    int ncele = m->nodecode[n1] * m->nodecode[n2] * m->nodecode[n3] * m->noff[ie];
    double NCELE = ncele;
    double Tau0Avg = (t0[n1] + t0[n2] + t0[n3]) / 3.0;
    double Tau0Avg_S = Tau0Avg;
    double MsOn = OnDiag * AreaIE * (1.0 / DT + Tau0Avg_S / 2.0) / DT / 12.0;
    double SPM1 = 1.0, SPM2 = 1.0, SPM3 = 1.0;
    COEF(n1, 1) = COEF(n1, 1) + MsOn * SPM1 * NCELE;
    COEF(n2, 1) = COEF(n2, 1) + MsOn * SPM2 * NCELE;
    COEF(n3, 1) = COEF(n3, 1) + MsOn * SPM3 * NCELE;
  }
  if (m->order == ADCO_ORDER_BLOCKED) {
    m->ep = adco_dot(ADCO_ORDER_BLOCKED, np, m->coef, m->coef);
  } else {
    m->ep = 0.0;
    for (int i = 0; i < np; ++i) m->ep = m->ep + COEF(i, 1) * COEF(i, 1);
  }
  m->ep = sqrt(m->ep / np);
  for (int I = 0; I < m->neta; ++I) COEF(m->s.nbd[I] - 1, 1) = m->ep;
  int die = 0;
  for (int i = 0; i < np; ++i) {
    if (COEF(i, 1) == 0.0) COEF(i, 1) = m->ep;
    if (COEF(i, 1) < 0.0) die = 1;
  }
  return die ? -1 : 0;
}

static int gwce_lhs(adco_model *m) {
  const int np = m->np, ne = m->ne;
  const double OnDiag = 2.0, OffDiag = 1.0, DT = m->dt;
  const double A00 = m->s.a00, G = m->s.g;
  if (m->s.ilump) return gwce_lhs_lumped(m);
  memset(m->coef, 0, sizeof(double) * (size_t)np * m->mnei);
  for (int ie = 0; ie < ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    int ncele = m->nodecode[n1] * m->nodecode[n2] * m->nodecode[n3] * m->noff[ie];
    double NCELE = ncele;
    double FDX1 = FDXE(ie, 0) * 1.0, FDX2 = FDXE(ie, 1) * 1.0, FDX3 = FDXE(ie, 2) * 1.0;
    double FDY1 = FDYE(ie, 0) * 1.0, FDY2 = FDYE(ie, 1) * 1.0, FDY3 = FDYE(ie, 2) * 1.0;
    double SFCXADJ = 1.0, SFCYADJ = 1.0;
    double AreaIE2 = m->s.areas[ie];
    double AreaIE = AreaIE2 * 0.5;
    double AreaIE4 = AreaIE2 * 2.0;
    const double *dp = m->s.dp;
    double DPAvg = (dp[n1] + dp[n2] + dp[n3]) / 3.0 +
                   (double)(m->ifnlfa * 0) * (m->eta2[n1] + m->eta2[n2] + m->eta2[n3]) / 3.0;
    double Cfac = 1.0;
    double GA = m->alphal[ie] * A00 * G * DPAvg * Cfac / AreaIE4;
    const double *t0 = m->s.tau0var;
    double Tau0Avg = (t0[n1] + t0[n2] + t0[n3]) / 3.0;
    double Tau0Avg_S = Tau0Avg;
    double MsOn = OnDiag * AreaIE * (1.0 / DT + Tau0Avg_S / 2.0) / DT / 12.0;
    double MsOff = OffDiag * AreaIE * (1.0 / DT + Tau0Avg_S / 2.0) / DT / 12.0;
    int J12 = CSR(ie, 0, 0), J13 = CSR(ie, 0, 1), J21 = CSR(ie, 1, 0);
    int J23 = CSR(ie, 1, 1), J31 = CSR(ie, 2, 0), J32 = CSR(ie, 2, 1);
    double SPM1 = 1.0, SPM2 = 1.0, SPM3 = 1.0;
    COEF(n1, 1) = COEF(n1, 1) + (MsOn * SPM1 + GA * (SFCXADJ * FDX1 * FDX1 + SFCYADJ * FDY1 * FDY1)) * NCELE;
    COEF(n1, J12) = COEF(n1, J12) + (MsOff * SPM2 + GA * (SFCXADJ * FDX1 * FDX2 + SFCYADJ * FDY1 * FDY2)) * NCELE;
    COEF(n1, J13) = COEF(n1, J13) + (MsOff * SPM3 + GA * (SFCXADJ * FDX1 * FDX3 + SFCYADJ * FDY1 * FDY3)) * NCELE;
    COEF(n2, J21) = COEF(n2, J21) + (MsOff * SPM1 + GA * (SFCXADJ * FDX2 * FDX1 + SFCYADJ * FDY2 * FDY1)) * NCELE;
    COEF(n2, 1) = COEF(n2, 1) + (MsOn * SPM2 + GA * (SFCXADJ * FDX2 * FDX2 + SFCYADJ * FDY2 * FDY2)) * NCELE;
    COEF(n2, J23) = COEF(n2, J23) + (MsOff * SPM3 + GA * (SFCXADJ * FDX2 * FDX3 + SFCYADJ * FDY2 * FDY3)) * NCELE;
    COEF(n3, J31) = COEF(n3, J31) + (MsOff * SPM1 + GA * (SFCXADJ * FDX3 * FDX1 + SFCYADJ * FDY3 * FDY1)) * NCELE;
    COEF(n3, J32) = COEF(n3, J32) + (MsOff * SPM2 + GA * (SFCXADJ * FDX3 * FDX2 + SFCYADJ * FDY3 * FDY2)) * NCELE;
    COEF(n3, 1) = COEF(n3, 1) + (MsOn * SPM3 + GA * (SFCXADJ * FDX3 * FDX3 + SFCYADJ * FDY3 * FDY3)) * NCELE;
  }
  /* EP and boundary rows/columns, src/gwce.F:643-707 */
  if (m->order == ADCO_ORDER_BLOCKED) {
    m->ep = adco_dot(ADCO_ORDER_BLOCKED, np, m->coef, m->coef);
  } else {
    m->ep = 0.0;
    for (int i = 0; i < np; ++i) m->ep = m->ep + COEF(i, 1) * COEF(i, 1);
  }
  m->ep = sqrt(m->ep / np);
  const int *nbd = m->s.nbd, *nneigh = m->s.nneigh;
  for (int I = 0; I < m->neta; ++I) {
    int b = nbd[I] - 1;
    COEF(b, 1) = m->ep;
    for (int J = 2; J <= nneigh[b]; ++J) COEF(b, J) = 0.0;
  }
  memset(m->obccoef, 0, sizeof(double) * (size_t)(m->neta > 0 ? m->neta : 1) * m->mnei);
  for (int I = 0; I < m->neta; ++I) {
    int b = nbd[I] - 1;
    for (int J = 2; J <= nneigh[b]; ++J) {
      int nb = NEITAB(b, J) - 1;
      for (int IJ = 2; IJ <= nneigh[nb]; ++IJ) {
        if (b + 1 == NEITAB(nb, IJ)) {
          OBC(I, J - 2) = COEF(nb, IJ);
          COEF(nb, IJ) = 0.0;
        }
      }
    }
  }
  int die = 0;
  for (int i = 0; i < np; ++i) {
    if (COEF(i, 1) == 0.0) COEF(i, 1) = m->ep;
    if (COEF(i, 1) < 0.0) die = 1;
  }
  return die ? -1 : 0;
}

/* GWCE RHS element loop, src/gwce.F:1000-1615 */
static void gwce_rhs(adco_model *m) {
  const int ne = m->ne;
  /* OnDiag = (1+ILump)*2, OffDiag = 1-ILump (src/gwce.F:419-420) */
  const double OnDiag = m->s.ilump ? 4.0 : 2.0, OffDiag = m->s.ilump ? 0.0 : 1.0;
  const double DT = m->dt, G = m->s.g;
  const double A00 = m->s.a00, B00 = m->s.b00, C00 = m->s.c00;
  const double *dp = m->s.dp, *evm = m->s.evm, *t0 = m->s.tau0var, *corif = m->s.corif;
  const int nws = m->s.nws;
  const double sponge = 1.0; /* sponge_dis_mthd */
  const double IFNLFA = m->ifnlfa, IFNLCT = m->ifnlct, IFNLCAT = m->ifnlcat;
  const double IFNL_HDP = 0.0;
  for (int ie = 0; ie < ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    int ncele = m->nodecode[n1] * m->nodecode[n2] * m->nodecode[n3] * m->noff[ie];
    double E0N1 = m->eta1[n1], E0N2 = m->eta1[n2], E0N3 = m->eta1[n3];
    double E1N1 = m->eta2[n1], E1N2 = m->eta2[n2], E1N3 = m->eta2[n3];
    double E1N1SQ = E1N1 * E1N1, E1N2SQ = E1N2 * E1N2, E1N3SQ = E1N3 * E1N3;
    double ESN1 = m->etas[n1], ESN2 = m->etas[n2], ESN3 = m->etas[n3];
    double U1N1 = m->uu1[n1], U1N2 = m->uu1[n2], U1N3 = m->uu1[n3];
    double V1N1 = m->vv1[n1], V1N2 = m->vv1[n2], V1N3 = m->vv1[n3];
    double QX1N1 = m->qx1[n1], QX1N2 = m->qx1[n2], QX1N3 = m->qx1[n3];
    double QY1N1 = m->qy1[n1], QY1N2 = m->qy1[n2], QY1N3 = m->qy1[n3];
    double H1N1 = m->h2[n1], H1N2 = m->h2[n2], H1N3 = m->h2[n3];
    double EVMN1 = evm[n1], EVMN2 = evm[n2], EVMN3 = evm[n3];
    double T0N1 = t0[n1], T0N2 = t0[n2], T0N3 = t0[n3];
    double E0AN1 = 0, E0AN2 = 0, E0AN3 = 0, ESAN1 = 0, ESAN2 = 0, ESAN3 = 0;
    double ABC1 = 0, ABC2 = 0, ABC3 = 0, U1AN1 = 0, U1AN2 = 0, U1AN3 = 0;
    double V1AN1 = 0, V1AN2 = 0, V1AN3 = 0;
    double Sg_eta1 = 0, Sg_eta2 = 0, Sg_eta3 = 0, Sg_mnx1 = 0, Sg_mnx2 = 0, Sg_mnx3 = 0;
    double Sg_mny1 = 0, Sg_mny2 = 0, Sg_mny3 = 0;
    (void)E0AN1; (void)E0AN2; (void)E0AN3;
    double PR1N1 = 0, PR1N2 = 0, PR1N3 = 0;
    if (nws != 0) {
      PR1N1 = m->pr1[n1];
      PR1N2 = m->pr1[n2];
      PR1N3 = m->pr1[n3];
    }
    double Off1N1 = 0, Off1N2 = 0, Off1N3 = 0, TiPN1 = 0, TiPN2 = 0, TiPN3 = 0;
    if (m->ctip) { /* src/gwce.F:1181-1185 */
      TiPN1 = m->tip1[n1];
      TiPN2 = m->tip1[n2];
      TiPN3 = m->tip1[n3];
    }
    const double *tkm = m->tkm;
    double BSXN1 = tkm[3 * n1 + 0] * QX1N1 + tkm[3 * n1 + 2] * QY1N1;
    double BSYN1 = tkm[3 * n1 + 2] * QX1N1 + tkm[3 * n1 + 1] * QY1N1;
    double BSXN2 = tkm[3 * n2 + 0] * QX1N2 + tkm[3 * n2 + 2] * QY1N2;
    double BSYN2 = tkm[3 * n2 + 2] * QX1N2 + tkm[3 * n2 + 1] * QY1N2;
    double BSXN3 = tkm[3 * n3 + 0] * QX1N3 + tkm[3 * n3 + 2] * QY1N3;
    double BSYN3 = tkm[3 * n3 + 2] * QX1N3 + tkm[3 * n3 + 1] * QY1N3;
    double AreaIE2 = m->s.areas[ie];
    double AreaIE = AreaIE2 / 2.0;
    double AreaIE4 = 2.0 * AreaIE2;
    double TANPHIAvg = 0.0, SFCXADJ = 1.0, SFCYADJ = 1.0;
    double FDX1 = FDXE(ie, 0) * 1.0, FDX2 = FDXE(ie, 1) * 1.0, FDX3 = FDXE(ie, 2) * 1.0;
    double FDY1 = FDYE(ie, 0) * 1.0, FDY2 = FDYE(ie, 1) * 1.0, FDY3 = FDYE(ie, 2) * 1.0;
    /* zero with the lumped LHS (src/gwce.F:1233-1238) */
    double E0XGrad2A = 0.0, E0YGrad2A = 0.0;
    if (!m->s.ilump) {
      E0XGrad2A = E0N1 * FDX1 + E0N2 * FDX2 + E0N3 * FDX3;
      E0YGrad2A = E0N1 * FDY1 + E0N2 * FDY2 + E0N3 * FDY3;
    }
    double E1XGrad2A = E1N1 * FDX1 + E1N2 * FDX2 + E1N3 * FDX3;
    double E1YGrad2A = E1N1 * FDY1 + E1N2 * FDY2 + E1N3 * FDY3;
    double Tau0XGrad2A = T0N1 * FDX1 + T0N2 * FDX2 + T0N3 * FDX3;
    double Tau0YGrad2A = T0N1 * FDY1 + T0N2 * FDY2 + T0N3 * FDY3;
    Tau0XGrad2A = SFCXADJ * Tau0XGrad2A;
    Tau0YGrad2A = SFCYADJ * Tau0YGrad2A;
    /* Kolar-Gray lateral stress */
    double EVMXGrad = (EVMN1 * FDX1 + EVMN2 * FDX2 + EVMN3 * FDX3) / AreaIE2;
    double EVMYGrad = (EVMN1 * FDY1 + EVMN2 * FDY2 + EVMN3 * FDY3) / AreaIE2;
    double EVMAvgODT = ((EVMN1 + EVMN2 + EVMN3) / 3.0) / DT;
    double MX = (EVMXGrad * (QX1N1 * FDX1 + QX1N2 * FDX2 + QX1N3 * FDX3) +
                 EVMYGrad * (QY1N1 * FDX1 + QY1N2 * FDX2 + QY1N3 * FDX3) -
                 EVMAvgODT * (ESN1 * FDX1 + ESN2 * FDX2 + ESN3 * FDX3)) /
                AreaIE2;
    double MY = (EVMXGrad * (QX1N1 * FDY1 + QX1N2 * FDY2 + QX1N3 * FDY3) +
                 EVMYGrad * (QY1N1 * FDY1 + QY1N2 * FDY2 + QY1N3 * FDY3) -
                 EVMAvgODT * (ESN1 * FDY1 + ESN2 * FDY2 + ESN3 * FDY3)) /
                AreaIE2;
    double CorifAvg = (corif[n1] + corif[n2] + corif[n3]) / 3.0;
    double Tau0Avg = (T0N1 + T0N2 + T0N3) / 3.0;
    double Tau0QXAvg = (T0N1 * QX1N1 + T0N2 * QX1N2 + T0N3 * QX1N3) / 3.0;
    double Tau0QYAvg = (T0N1 * QY1N1 + T0N2 * QY1N2 + T0N3 * QY1N3) / 3.0;
    double U1Avg = (U1N1 + U1N2 + U1N3) / 3.0;
    double V1Avg = (V1N1 + V1N2 + V1N3) / 3.0;
    double QX1Avg = (QX1N1 + QX1N2 + QX1N3) / 3.0;
    double QY1Avg = (QY1N1 + QY1N2 + QY1N3) / 3.0;
    double ESAvg = (ESN1 + ESN2 + ESN3) / 3.0;
    int ihdp = m->ifnlfa * 0; /* IFNLFA*IFNL_HDP */
    double DPAvg = (dp[n1] + dp[n2] + dp[n3]) / 3.0 + ihdp * (E1N1 + E1N2 + E1N3) / 3.0;
    double DPAvgS0 = (dp[n1] + dp[n2] + dp[n3]) / 3.0 + ihdp * (E0N1 + E0N2 + E0N3) / 3.0;
    double Cfac = 1.0, CfacS0 = 1.0;
    double GDPAvgOAreaIE4 = m->alphal[ie] * G * DPAvg * Cfac / AreaIE4;
    double GDPAvgOAreaIE4_S0 = m->alphal[ie] * G * DPAvgS0 * CfacS0 / AreaIE4;
    double HAvg = (H1N1 + H1N2 + H1N3) / 3.0;
    double GHAvg = m->alphal[ie] * G * HAvg;
    double GHAvgOAreaIE2 = GHAvg / AreaIE2;
    double BSXAvg = (BSXN1 + BSXN2 + BSXN3) / 3.0;
    double BSYAvg = (BSYN1 + BSYN2 + BSYN3) / 3.0;
    double MXAvg = MX, MYAvg = MY;
    CorifAvg = CorifAvg + IFNLCT * TANPHIAvg * U1Avg;
    double SigmaAvg = (Sg_eta1 + Sg_eta2 + Sg_eta3) / 3.0;
    double Tau0SigmaAvg = (Sg_eta1 * T0N1 + Sg_eta2 * T0N2 + Sg_eta3 * T0N3) / 3.0;
    double SigmaQxAvg = (Sg_mnx1 * QX1N1 + Sg_mnx2 * QX1N2 + Sg_mnx3 * QX1N3) / 3.0;
    double SigmaQyAvg = (Sg_mny1 * QY1N1 + Sg_mny2 * QY1N2 + Sg_mny3 * QY1N3) / 3.0;
    double AvgSigmaHAbsU = (Sg_mnx1 * H1N1 * U1AN1 + Sg_mnx2 * H1N2 * U1AN2 + Sg_mnx3 * H1N3 * U1AN3) / 3.0;
    double AvgSigmaHAbsV = (Sg_mny1 * H1N1 * V1AN1 + Sg_mny2 * H1N2 * V1AN2 + Sg_mny3 * H1N3 * V1AN3) / 3.0;
    double WSXAvg = 0.0, WSYAvg = 0.0;
    if (nws != 0) {
      double f1 = 1.0, f2 = 1.0, f3 = 1.0;
      WSXAvg = (f1 * m->wsx1[n1] + f2 * m->wsx1[n2] + f3 * m->wsx1[n3]) / 3.0;
      WSYAvg = (f1 * m->wsy1[n1] + f2 * m->wsy1[n2] + f3 * m->wsy1[n3]) / 3.0;
    }
    double DispXAvg = 0.0, DispYAvg = 0.0, BCXAvg = 0.0, BCYAvg = 0.0;
    double Tau0Avg_S = Tau0Avg + ((A00 + B00) * 2.0 * Tau0SigmaAvg * DT + SigmaAvg) * sponge;
    double MsFacR = AreaIE * (1.0 / DT - Tau0Avg_S / 2.0) / DT / 12.0;
    double MsFacR_S = (-(A00 + B00 + C00) * AreaIE * Tau0SigmaAvg / 12.0) * sponge;
    double MsFacRS_ABC00 = (Tau0SigmaAvg * AreaIE / 12.0) * sponge;
    double MsFacRS_T = ((SigmaAvg / (2.0 * DT)) * AreaIE / 12.0) * sponge;
    double GOAreaIE4 = m->alphal[ie] * G / AreaIE4;
    double Tau0SpaVar = (QX1Avg * Tau0XGrad2A + QY1Avg * Tau0YGrad2A) / 6.0;
    double A00pB00 = A00 + B00;
    double JXAvg = CorifAvg * QY1Avg -
                   IFNLFA * GOAreaIE4 * (1 - IFNL_HDP) * (E1N1SQ * FDX1 + E1N2SQ * FDX2 + E1N3SQ * FDX3) -
                   GHAvgOAreaIE2 * ((PR1N1 - TiPN1 - Off1N1) * FDX1 + (PR1N2 - TiPN2 - Off1N2) * FDX2 +
                                    (PR1N3 - TiPN3 - Off1N3) * FDX3) +
                   WSXAvg - BSXAvg + MXAvg - DispXAvg - BCXAvg + Tau0QXAvg - SigmaQxAvg * sponge +
                   AvgSigmaHAbsU * sponge;
    double JYAvg = -CorifAvg * QX1Avg -
                   IFNLFA * GOAreaIE4 * (1 - IFNL_HDP) * (E1N1SQ * FDY1 + E1N2SQ * FDY2 + E1N3SQ * FDY3) -
                   GHAvgOAreaIE2 * ((PR1N1 - TiPN1 - Off1N1) * FDY1 + (PR1N2 - TiPN2 - Off1N2) * FDY2 +
                                    (PR1N3 - TiPN3 - Off1N3) * FDY3) +
                   WSYAvg - BSYAvg + MYAvg - DispYAvg - BCYAvg + Tau0QYAvg - SigmaQyAvg * sponge +
                   AvgSigmaHAbsV * sponge;
    /* non-conservative advection */
    JXAvg = JXAvg -
            IFNLCT * (QX1Avg * (U1N1 * FDX1 + U1N2 * FDX2 + U1N3 * FDX3) +
                      QY1Avg * (U1N1 * FDY1 + U1N2 * FDY2 + U1N3 * FDY3)) /
                AreaIE2 +
            IFNLCAT * U1Avg * ESAvg / DT;
    JYAvg = JYAvg -
            IFNLCT * (QX1Avg * (V1N1 * FDX1 + V1N2 * FDX2 + V1N3 * FDX3) +
                      QY1Avg * (V1N1 * FDY1 + V1N2 * FDY2 + V1N3 * FDY3)) /
                AreaIE2 +
            IFNLCAT * V1Avg * ESAvg / DT;
    double SPM1 = 1.0, SPM2 = 1.0, SPM3 = 1.0;
    double A1 = (OnDiag * ESN1 * SPM1 + OffDiag * (ESN2 * SPM2 + ESN3 * SPM3)) * MsFacR -
                GDPAvgOAreaIE4_S0 * (C00 * (SFCXADJ * FDX1 * E0XGrad2A + SFCYADJ * FDY1 * E0YGrad2A)) -
                GDPAvgOAreaIE4 * (A00pB00 * (SFCXADJ * FDX1 * E1XGrad2A + SFCYADJ * FDY1 * E1YGrad2A)) +
                (SFCXADJ * JXAvg * FDX1 + SFCYADJ * JYAvg * FDY1) * 0.5 + Tau0SpaVar +
                (OnDiag * E0N1 * SPM1 + OffDiag * (E0N2 * SPM2 + E0N3 * SPM3)) * MsFacR_S +
                (OnDiag * ESAN1 * SPM1 + OffDiag * (ESAN2 * SPM2 + ESAN3 * SPM3)) * MsFacRS_T +
                (OnDiag * ABC1 * SPM1 + OffDiag * (ABC2 * SPM2 + ABC3 * SPM3)) * MsFacRS_ABC00;
    double A2 = (OnDiag * ESN2 * SPM2 + OffDiag * (ESN1 * SPM1 + ESN3 * SPM3)) * MsFacR -
                GDPAvgOAreaIE4_S0 * (C00 * (SFCXADJ * FDX2 * E0XGrad2A + SFCYADJ * FDY2 * E0YGrad2A)) -
                GDPAvgOAreaIE4 * (A00pB00 * (SFCXADJ * FDX2 * E1XGrad2A + SFCYADJ * FDY2 * E1YGrad2A)) +
                (SFCXADJ * JXAvg * FDX2 + SFCYADJ * JYAvg * FDY2) * 0.5 + Tau0SpaVar +
                (OnDiag * E0N2 * SPM2 + OffDiag * (E0N1 * SPM1 + E0N3 * SPM3)) * MsFacR_S +
                (OnDiag * ESAN2 * SPM2 + OffDiag * (ESAN1 * SPM1 + ESAN3 * SPM3)) * MsFacRS_T +
                (OnDiag * ABC2 * SPM2 + OffDiag * (ABC1 * SPM1 + ABC3 * SPM3)) * MsFacRS_ABC00;
    double A3 = (OnDiag * ESN3 * SPM3 + OffDiag * (ESN1 * SPM1 + ESN2 * SPM2)) * MsFacR -
                GDPAvgOAreaIE4_S0 * (C00 * (SFCXADJ * FDX3 * E0XGrad2A + SFCYADJ * FDY3 * E0YGrad2A)) -
                GDPAvgOAreaIE4 * (A00pB00 * (SFCXADJ * FDX3 * E1XGrad2A + SFCYADJ * FDY3 * E1YGrad2A)) +
                (SFCXADJ * JXAvg * FDX3 + SFCYADJ * JYAvg * FDY3) * 0.5 + Tau0SpaVar +
                (OnDiag * E0N3 * SPM3 + OffDiag * (E0N1 * SPM1 + E0N2 * SPM2)) * MsFacR_S +
                (OnDiag * ESAN3 * SPM3 + OffDiag * (ESAN1 * SPM1 + ESAN2 * SPM2)) * MsFacRS_T +
                (OnDiag * ABC3 * SPM3 + OffDiag * (ABC1 * SPM1 + ABC2 * SPM2)) * MsFacRS_ABC00;
    double NCE = ncele;
    m->gwce_lv[n1] = m->gwce_lv[n1] + A1 * NCE;
    m->gwce_lv[n2] = m->gwce_lv[n2] + A2 * NCE;
    m->gwce_lv[n3] = m->gwce_lv[n3] + A3 * NCE;
  }
}

/* periodic elevation forcing, src/gwce.F:1638-1651 */
static void eta_bc_periodic(adco_model *m, double timeh, double *eta2) {
  for (int J = 0; J < m->nbfr; ++J) {
    double PER = m->s.amig[J] == 0.0 ? 0.0 : 2.0 * 3.141592653589793 / m->s.amig[J];
    int ncyc = 0;
    if (PER != 0.0) ncyc = (int)(timeh / PER);
    double ARGJ = m->s.amig[J] * (timeh - ncyc * PER) + m->s.face[J];
    double RFF = m->s.ff[J] * m->rampelev;
    for (int I = 0; I < m->neta; ++I) {
      double ARG = ARGJ - m->s.efa[(size_t)I * m->nbfr + J];
      int b = m->s.nbd[I] - 1;
      eta2[b] = eta2[b] + m->s.emo[(size_t)I * m->nbfr + J] * RFF * cos(ARG);
    }
  }
}

void adco_eta_bc(const adco_setup *s, double timeh, double rampelev, double *out) {
  adco_model tmp;
  memset(&tmp, 0, sizeof(tmp));
  tmp.s = *s;
  tmp.nbfr = s->nbfr;
  tmp.neta = s->neta;
  tmp.rampelev = rampelev;
  int mx = 0;
  for (int I = 0; I < s->neta; ++I) if (s->nbd[I] > mx) mx = s->nbd[I];
  double *e = dalloc((size_t)mx + 1);
  eta_bc_periodic(&tmp, timeh, e);
  for (int I = 0; I < s->neta; ++I) out[I] = e[s->nbd[I] - 1];
  free(e);
}

/* GWCE_New, src/gwce.F:237-2111 (consistent LHS path) */
static int gwce_new(adco_model *m, double timeh) {
  const int np = m->np;
  if (m->ncchange > 0) {
    m->ncchange = 0;
    if (gwce_lhs(m) != 0) return -1;
  }
  gwce_rhs(m);
  for (int i = 0; i < np; ++i) {
    m->eta1[i] = m->eta2[i];
    m->h1[i] = m->h2[i];
    m->eta2[i] = 0.0;
  }
  eta_bc_periodic(m, timeh, m->eta2);
  if (m->nfluxf) gwce_flux(m);
  const int *nbd = m->s.nbd, *nneigh = m->s.nneigh;
  if (m->s.ilump) {
    /* lumped: boundary rows from COEFD, explicit solve (src/gwce.F:1856-1862,
     * 2009-2022) */
    for (int I = 0; I < m->neta; ++I) {
      int b = nbd[I] - 1;
      m->etas[b] = m->eta2[b] - m->eta1[b];
      m->gwce_lv[b] = m->etas[b] * m->nodecode[b] * COEF(b, 1);
    }
    for (int i = 0; i < np; ++i) {
      double rdiag = COEF(i, 1) == 0.0 ? 0.0 : 1.0 / COEF(i, 1);
      m->etas[i] = m->gwce_lv[i] * rdiag;
    }
    m->numitr = 0;
    for (int i = 0; i < np; ++i)
      m->eta2[i] = m->nodecode[i] * (m->etas[i] + 0.0) + m->eta1[i];
    return 0;
  }
  for (int I = 0; I < m->neta; ++I) {
    int b = nbd[I] - 1;
    m->etas[b] = m->eta2[b] - m->eta1[b];
    m->gwce_lv[b] = m->etas[b] * m->nodecode[b] * m->ep;
    for (int J = 2; J <= nneigh[b]; ++J) {
      int k = NEITAB(b, J) - 1;
      m->gwce_lv[k] = m->gwce_lv[k] - m->etas[b] * OBC(I, J - 2);
    }
  }
  m->iparm[0] = m->s.itmax;
  int ier = adco_jcg(&m->itp, np, np, m->mnei, m->neitab, m->coef, m->gwce_lv, m->etas,
                     m->nw, m->wksp, m->iparm, m->rparm);
  m->numitr = m->iparm[0];
  for (int i = 0; i < np; ++i)
    m->eta2[i] = m->nodecode[i] * (m->etas[i] + 0.0) + m->eta1[i];
  return ier;
}

/* computeWettingAndDrying, src/wetdry.F:140-1470 (2DDI, no subgrid/weirs) */
static void wetdry(adco_model *m) {
  const int np = m->np, ne = m->ne;
  const double *dp = m->s.dp, *X = m->s.x, *Y = m->s.y;
  const double H0 = m->s.h0, HOFF = m->hoff, HABSMIN = m->habsmin, G = m->s.g;
  const double VELMIN = m->s.velmin, DTDP = m->s.dtdp, TKNF = 1.0;
  double *eta2 = m->eta2, *tk = m->tk, *tkm = m->tkm, *fric = m->fric;
  int *nodecode = m->nodecode, *nnodecode = m->nnodecode;
  for (int i = 0; i < np; ++i) m->tk_tmp[i] = 0.0;
  for (int i = 0; i < np; ++i) m->nibcnt[i] = 0;
  for (int i = 0; i < ne; ++i) {
    m->noffold[i] = m->noff[i];
    m->noff[i] = 1;
  }
  /* D1 */
  for (int i = 0; i < np; ++i) {
    if (nodecode[i] == 1) {
      double htot = dp[i] + eta2[i];
      if (htot <= H0) {
        if (htot < HABSMIN) eta2[i] = HABSMIN - dp[i];
        nnodecode[i] = 0;
        nodecode[i] = 0;
        m->ncchange = m->ncchange + 1;
      }
    }
  }
  /* W1 / W2a */
  for (int ie = 0; ie < ne; ++ie) {
    int n[3] = {NM(ie, 0), NM(ie, 1), NM(ie, 2)};
    int nctot = nodecode[n[0]] + nodecode[n[1]] + nodecode[n[2]];
    if (nctot == 2) {
      double htotn[3];
      for (int k = 0; k < 3; ++k) htotn[k] = dp[n[k]] + eta2[n[k]];
      /* the three cases of src/wetdry.F:368, 515, 662: (wet a, wet b, dry c) */
      int a = -1, b = -1, c = -1;
      if (nodecode[n[0]] == 1 && nodecode[n[1]] == 1) {
        a = 0; b = 1; c = 2;
      } else if (nodecode[n[1]] == 1 && nodecode[n[2]] == 1) {
        a = 1; b = 2; c = 0;
      } else if (nodecode[n[2]] == 1 && nodecode[n[0]] == 1) {
        a = 2; b = 0; c = 1;
      }
      if (a >= 0 && htotn[a] >= HOFF && htotn[b] >= HOFF) {
        int na = n[a], nb = n[b], nc = n[c];
        int nm123 = na;
        if (eta2[nb] > eta2[na]) nm123 = nb;
        double dx = Y[nc] - Y[nm123], dy = X[nc] - X[nm123];
        double deldist = sqrt(dx * dx + dy * dy);
        double deleta = eta2[nm123] - eta2[nc];
        if (deleta < 0.0) deleta = 0.0;
        double H1 = eta2[nm123] + dp[nm123];
        double hyb = m->ifhybf * pow(1.0 + pow(m->hbreak / H1, m->ftheta), m->fgamma / m->ftheta);
        double tkwet = fmax(fric[nm123] * (m->iflinbf + (VELMIN / H1) * (m->ifnlbf + hyb)), 0.0001);
        double vel = G * (deleta / deldist) / tkwet;
        if (vel > VELMIN) {
          nnodecode[nc] = 1;
          double hyb2 = m->ifhybf * pow(1.0 + pow(m->hbreak / H1, m->ftheta), m->fgamma / m->ftheta);
          m->tk_tmp[nm123] = fmin(fric[nm123] * (m->iflinbf + (TKNF * vel / H1) * (m->ifnlbf + hyb2)), 2.0 / DTDP);
          tk[nm123] = fmax(m->tk_tmp[nm123], tk[nm123]);
          tkm[3 * nm123 + 0] = tk[nm123];
          tkm[3 * nm123 + 1] = tk[nm123];
          /* partner: the other wet node */
          int partner = (nm123 == nb) ? na : nb;
          m->tk_tmp[partner] = tk[nm123];
          tk[partner] = fmax(m->tk_tmp[partner], tk[partner]);
          tkm[3 * partner + 0] = tk[partner];
          tkm[3 * partner + 1] = tk[partner];
          if (m->tk_tmp[partner] < tk[partner]) {
            tk[nm123] = tk[partner];
            tkm[3 * nm123 + 0] = tk[nm123];
            tkm[3 * nm123 + 1] = tk[nm123];
          }
        }
      }
    }
    /* NIBNODECODE is zero without barriers: NIBCNT stays zero */
  }
  /* DE1 (NIBCNT = 0 everywhere, noffActive) */
  for (int ie = 0; ie < ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    double e1 = eta2[n1], e2 = eta2[n2], e3 = eta2[n3];
    double h1 = dp[n1] + eta2[n1], h2 = dp[n2] + eta2[n2], h3 = dp[n3] + eta2[n3];
    if ((e1 >= e2) && (e2 > e3)) if ((h1 < HOFF) || (h2 < HOFF)) m->noff[ie] = 0;
    if ((e2 >= e3) && (e3 > e1)) if ((h2 < HOFF) || (h3 < HOFF)) m->noff[ie] = 0;
    if ((e3 >= e1) && (e1 > e2)) if ((h3 < HOFF) || (h1 < HOFF)) m->noff[ie] = 0;
    if ((e1 >= e3) && (e3 > e2)) if ((h1 < HOFF) || (h3 < HOFF)) m->noff[ie] = 0;
    if ((e2 >= e1) && (e1 > e3)) if ((h2 < HOFF) || (h1 < HOFF)) m->noff[ie] = 0;
    if ((e3 >= e2) && (e2 > e1)) if ((h3 < HOFF) || (h2 < HOFF)) m->noff[ie] = 0;
  }
  /* D2: MJU, TotalArea, landlocked nodes */
  for (int i = 0; i < np; ++i) {
    m->mju[i] = 0;
    m->totalarea[i] = 0.0;
  }
  for (int ie = 0; ie < ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    int ncele = nnodecode[n1] * nnodecode[n2] * nnodecode[n3] * m->noff[ie];
    double areaele = ncele * m->s.areas[ie] / 2.0;
    m->mju[n1] += ncele;
    m->mju[n2] += ncele;
    m->mju[n3] += ncele;
    m->totalarea[n1] = m->totalarea[n1] + areaele;
    m->totalarea[n2] = m->totalarea[n2] + areaele;
    m->totalarea[n3] = m->totalarea[n3] + areaele;
  }
  for (int i = 0; i < np; ++i) {
    if (nnodecode[i] == 1 && m->mju[i] == 0) nnodecode[i] = 0;
    if (m->mju[i] == 0) m->mju[i] = 1;
  }
  for (int i = 0; i < np; ++i) {
    if (nnodecode[i] != nodecode[i]) {
      nodecode[i] = nnodecode[i];
      m->ncchange = m->ncchange + 1;
    }
  }
  for (int ie = 0; ie < ne; ++ie)
    if (m->noff[ie] != m->noffold[ie]) m->ncchange = m->ncchange + 1;
}

/* computeAlpha, src/wetdry.F:1561-1652 */
static void compute_alpha(adco_model *m) {
  const double *dp = m->s.dp, *eta2 = m->eta2;
  const double slim = 1.0e9;
  for (int ie = 0; ie < m->ne; ++ie) m->alphal[ie] = 1.0;
  for (int ie = 0; ie < m->ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    int ncele = m->nodecode[n1] * m->nodecode[n2] * m->nodecode[n3] * m->noff[ie];
    if (ncele != 1) continue;
    double hn1 = eta2[n1] + dp[n1], hn2 = eta2[n2] + dp[n2], hn3 = eta2[n3] + dp[n3];
    if (!((hn1 <= 5.0e-1) || (hn2 <= 5.0e-1) || (hn3 <= 5.0e-1))) continue;
    double FDX1 = FDXE(ie, 0) * 1.0, FDX2 = FDXE(ie, 1) * 1.0, FDX3 = FDXE(ie, 2) * 1.0;
    double FDY1 = FDYE(ie, 0) * 1.0, FDY2 = FDYE(ie, 1) * 1.0, FDY3 = FDYE(ie, 2) * 1.0;
    double DPDXA = dp[n1] * FDX1 + dp[n2] * FDX2 + dp[n3] * FDX3;
    double DPDYA = dp[n1] * FDY1 + dp[n2] * FDY2 + dp[n3] * FDY3;
    double DETADXA = eta2[n1] * FDX1 + eta2[n2] * FDX2 + eta2[n3] * FDX3;
    double DETADYA = eta2[n1] * FDY1 + eta2[n2] * FDY2 + eta2[n3] * FDY3;
    double area = m->s.areas[ie];
    double DPM = sqrt(DPDXA * DPDXA + DPDYA * DPDYA) / area;
    double a;
    if (DPM > 1e-5) {
      double DDTMP = DETADXA * (-DPDXA) + DETADYA * (-DPDYA);
      if (fabs(DDTMP) > 1e-12)
        a = slim * (sqrt(DPDXA * DPDXA + DPDYA * DPDYA) / ((1.0 / area) * (DDTMP)));
      else
        a = 1.0;
    } else {
      double DETAM = sqrt(DETADXA * DETADXA + DETADYA * DETADYA) / area;
      if (fabs(DETAM) > 1e-12)
        a = slim / DETAM;
      else
        a = 1.0;
    }
    if (a <= 0.0) a = 1.0;
    if (a >= 1.0) a = 1.0;
    m->alphal[ie] = a;
  }
}

/* Mom_Eqs_New_NC, src/momentum.F:166-1208 */
static void momentum(adco_model *m) {
  const int np = m->np, ne = m->ne;
  const double DT = m->dt, GO2 = m->go2, DTO2 = m->dto2;
  const double THIRD = 1.0 / 3.0, IFNLCT = m->ifnlct;
  const double *evm = m->s.evm;
  const int nws = m->s.nws;
  for (int ie = 0; ie < ne; ++ie) {
    int n1 = NM(ie, 0), n2 = NM(ie, 1), n3 = NM(ie, 2);
    int ncele = m->nodecode[n1] * m->nodecode[n2] * m->nodecode[n3] * m->noff[ie];
    double U1N1 = m->uu1[n1], U1N2 = m->uu1[n2], U1N3 = m->uu1[n3];
    double V1N1 = m->vv1[n1], V1N2 = m->vv1[n2], V1N3 = m->vv1[n3];
    double H1N1 = m->h1[n1], H1N2 = m->h1[n2], H1N3 = m->h1[n3];
    double AreaIE2 = m->s.areas[ie];
    double AreaIE = 0.5 * AreaIE2;
    double FDX1 = FDXE(ie, 0) * 1.0, FDX2 = FDXE(ie, 1) * 1.0, FDX3 = FDXE(ie, 2) * 1.0;
    double FDY1 = FDYE(ie, 0) * 1.0, FDY2 = FDYE(ie, 1) * 1.0, FDY3 = FDYE(ie, 2) * 1.0;
    double U1Avg = THIRD * (U1N1 + U1N2 + U1N3);
    double V1Avg = THIRD * (V1N1 + V1N2 + V1N3);
    double DU1DXA = (U1N1 * FDX1 + U1N2 * FDX2 + U1N3 * FDX3) * 0.5;
    double DU1DYA = (U1N1 * FDY1 + U1N2 * FDY2 + U1N3 * FDY3) * 0.5;
    double DV1DXA = (V1N1 * FDX1 + V1N2 * FDX2 + V1N3 * FDX3) * 0.5;
    double DV1DYA = (V1N1 * FDY1 + V1N2 * FDY2 + V1N3 * FDY3) * 0.5;
    double EVMEle = THIRD * (evm[n1] + evm[n2] + evm[n3]);
    double al = m->alphal[ie];
    double B1N1 = al * m->eta1[n1], B2N1 = al * m->eta2[n1];
    double B1N2 = al * m->eta1[n2], B2N2 = al * m->eta2[n2];
    double B1N3 = al * m->eta1[n3], B2N3 = al * m->eta2[n3];
    if (nws != 0) {
      B1N1 = B1N1 + m->pr1[n1];
      B2N1 = B2N1 + m->pr2[n1];
      B1N2 = B1N2 + m->pr1[n2];
      B2N2 = B2N2 + m->pr2[n2];
      B1N3 = B1N3 + m->pr1[n3];
      B2N3 = B2N3 + m->pr2[n3];
    }
    if (m->ctip) { /* src/momentum.F:443-450 */
      B1N1 = B1N1 - m->tip1[n1];
      B2N1 = B2N1 - m->tip2[n1];
      B1N2 = B1N2 - m->tip1[n2];
      B2N2 = B2N2 - m->tip2[n2];
      B1N3 = B1N3 - m->tip1[n3];
      B2N3 = B2N3 - m->tip2[n3];
    }
    double DBTPDX1A = B1N1 * FDX1 + B1N2 * FDX2 + B1N3 * FDX3;
    double DBTPDX2A = B2N1 * FDX1 + B2N2 * FDX2 + B2N3 * FDX3;
    double DBTPDY1A = B1N1 * FDY1 + B1N2 * FDY2 + B1N3 * FDY3;
    double DBTPDY2A = B2N1 * FDY1 + B2N2 * FDY2 + B2N3 * FDY3;
    double Cfac1 = 1.0, Cfac2 = 1.0;
    double DBTPDXA = 0.5 * (Cfac1 * DBTPDX1A + Cfac2 * DBTPDX2A);
    double DBTPDYA = 0.5 * (Cfac1 * DBTPDY1A + Cfac2 * DBTPDY2A);
    double DBC = 0.0;
    double U1AvgDU1DXA = U1Avg * DU1DXA, V1AvgDU1DYA = V1Avg * DU1DYA;
    double U1AvgDV1DXA = U1Avg * DV1DXA, V1AvgDV1DYA = V1Avg * DV1DYA;
    double DU1DX = DU1DXA / AreaIE, DU1DY = DU1DYA / AreaIE;
    double DV1DX = DV1DXA / AreaIE, DV1DY = DV1DYA / AreaIE;
    double EH1 = EVMEle * H1N1, EH2 = EVMEle * H1N2, EH3 = EVMEle * H1N3;
    double LSXX1 = EH1 * DU1DX, LSXX2 = EH2 * DU1DX, LSXX3 = EH3 * DU1DX;
    double LSXY1 = EH1 * DU1DY, LSXY2 = EH2 * DU1DY, LSXY3 = EH3 * DU1DY;
    double LSYX1 = EH1 * DV1DX, LSYX2 = EH2 * DV1DX, LSYX3 = EH3 * DV1DX;
    double LSYY1 = EH1 * DV1DY, LSYY2 = EH2 * DV1DY, LSYY3 = EH3 * DV1DY;
    double TP = 0.0; /* TANPHI */
    double NCDT = ncele * DT;
    double TA1 = NCDT * (-IFNLCT * (U1AvgDU1DXA + V1AvgDU1DYA) - GO2 * DBTPDXA - DBC -
                         (1.5 * (LSXX1 * FDX1 + LSXY1 * FDY1) - AreaIE * TP * (LSXY1 + LSYX1)) / H1N1);
    double TA2 = NCDT * (-IFNLCT * (U1AvgDU1DXA + V1AvgDU1DYA) - GO2 * DBTPDXA - DBC -
                         (1.5 * (LSXX2 * FDX2 + LSXY2 * FDY2) - AreaIE * TP * (LSXY2 + LSYX2)) / H1N2);
    double TA3 = NCDT * (-IFNLCT * (U1AvgDU1DXA + V1AvgDU1DYA) - GO2 * DBTPDXA - DBC -
                         (1.5 * (LSXX3 * FDX3 + LSXY3 * FDY3) - AreaIE * TP * (LSXY3 + LSYX3)) / H1N3);
    double TB1 = NCDT * (-IFNLCT * (U1AvgDV1DXA + V1AvgDV1DYA) - GO2 * DBTPDYA - DBC -
                         (1.5 * (LSYX1 * FDX1 + LSYY1 * FDY1) + AreaIE * TP * (LSXX1 - LSYY1)) / H1N1);
    double TB2 = NCDT * (-IFNLCT * (U1AvgDV1DXA + V1AvgDV1DYA) - GO2 * DBTPDYA - DBC -
                         (1.5 * (LSYX2 * FDX2 + LSYY2 * FDY2) + AreaIE * TP * (LSXX2 - LSYY2)) / H1N2);
    double TB3 = NCDT * (-IFNLCT * (U1AvgDV1DXA + V1AvgDV1DYA) - GO2 * DBTPDYA - DBC -
                         (1.5 * (LSYX3 * FDX3 + LSYY3 * FDY3) + AreaIE * TP * (LSXX3 - LSYY3)) / H1N3);
    m->mom_lv_x[n1] = m->mom_lv_x[n1] + TA1;
    m->mom_lv_x[n2] = m->mom_lv_x[n2] + TA2;
    m->mom_lv_x[n3] = m->mom_lv_x[n3] + TA3;
    m->mom_lv_y[n1] = m->mom_lv_y[n1] + TB1;
    m->mom_lv_y[n2] = m->mom_lv_y[n2] + TB2;
    m->mom_lv_y[n3] = m->mom_lv_y[n3] + TB3;
  }
  double WSX = 0.0, WSY = 0.0;
  double *auv = m->auv;
  for (int i = 0; i < np; ++i) {
    int nci = m->nodecode[i];
    if (m->totalarea[i] != 0.0) {
      m->mom_lv_x[i] = m->mom_lv_x[i] / m->totalarea[i];
      m->mom_lv_y[i] = m->mom_lv_y[i] / m->totalarea[i];
    }
    if (nws != 0) {
      WSX = DTO2 * m->ifwind * (m->wsx1[i] / m->h1[i] + m->wsx2[i] / m->h2[i]);
      WSY = DTO2 * m->ifwind * (m->wsy1[i] / m->h1[i] + m->wsy2[i] / m->h2[i]);
    }
    double VCOEFXX = DTO2 * (m->tkm[3 * i + 0] + 0.0);
    double VCOEFYY = DTO2 * (m->tkm[3 * i + 1] + 0.0);
    double CorifA = m->s.corif[i] + IFNLCT * 0.0 * m->uu1[i];
    double VCOEFXY = DTO2 * (m->tkm[3 * i + 2] - CorifA);
    double VCOEFYX = DTO2 * (m->tkm[3 * i + 2] + CorifA);
    m->mom_lv_x[i] = nci * (m->mom_lv_x[i] + WSX + (1.0 - VCOEFXX) * m->uu1[i] - VCOEFXY * m->vv1[i]);
    m->mom_lv_y[i] = nci * (m->mom_lv_y[i] + WSY + (1.0 - VCOEFYY) * m->vv1[i] - VCOEFYX * m->uu1[i]);
    auv[4 * i + 0] = 1.0 + VCOEFXX * nci;
    auv[4 * i + 1] = 1.0 + VCOEFYY * nci;
    auv[4 * i + 2] = VCOEFXY * nci;
    auv[4 * i + 3] = VCOEFYX * nci;
  }
  /* velocity boundary conditions, src/momentum.F:810-934 */
  for (int J = 0; J < m->nvelme; ++J) {
    int I = m->s.me2gw[J] - 1;
    int b = m->s.nbv[I] - 1;
    int nci = m->nodecode[b];
    int code = m->s.lbcodei[I];
    double CS = m->s.csii[I], SI = m->s.siii[I];
    double *A = auv + 4 * b;
    const double qn2 = m->nfluxf ? m->qn2[I] : 0.0;
    if (code >= 0 && code <= 9) {
      double VelNorm = -qn2 / m->h2[b];
      double eps = 2.220446049250313e-16;
      if (fabs(norm2_3(m->tkm + 3 * b) - sqrt(2.0) * m->tk[b]) < eps) {
        m->mom_lv_x[b] = (SI * m->mom_lv_x[b] - CS * m->mom_lv_y[b] - VelNorm * A[2]) * nci;
        m->mom_lv_y[b] = VelNorm * A[0] * nci;
        A[2] = -CS * A[0];
        A[3] = -A[2];
        A[0] = SI * A[0];
        A[1] = A[0];
      } else {
        m->mom_lv_x[b] = (SI * m->mom_lv_x[b] - CS * m->mom_lv_y[b]) * nci;
        m->mom_lv_y[b] = VelNorm * nci;
        A[0] = A[0] * SI - A[3] * CS;
        A[2] = A[2] * SI - A[1] * CS;
        A[3] = CS;
        A[1] = SI;
      }
    }
    if (code >= 10 && code <= 19) {
      double VelNorm = -qn2 / m->h2[b];
      double VelTan = 0.0;
      m->mom_lv_x[b] = VelTan * nci;
      m->mom_lv_y[b] = VelNorm * nci;
      A[0] = SI;
      A[2] = -CS;
      A[3] = CS;
      A[1] = SI;
    }
  }
  for (int i = 0; i < np; ++i) {
    double *A = auv + 4 * i;
    double DDU = A[0] * A[1] - A[2] * A[3];
    m->uu2[i] = (m->mom_lv_x[i] * A[1] - m->mom_lv_y[i] * A[2]) / DDU;
    m->vv2[i] = (m->mom_lv_y[i] * A[0] - m->mom_lv_x[i] * A[3]) / DDU;
  }
  for (int i = 0; i < np; ++i) {
    m->qx2[i] = m->uu2[i] * m->h2[i];
    m->qy2[i] = m->vv2[i] * m->h2[i];
  }
}

int adco_step(adco_model *m, int it, const double *wsx2, const double *wsy2,
              const double *pr2) {
  const int np = m->np;
  const double *dp = m->s.dp;
  if (it == 1) {
    for (int i = 0; i < np; ++i) {
      m->h2[i] = dp[i] + m->ifnlfa * m->eta2[i];
      m->h1[i] = m->h2[i];
    }
  }
  double timeloc = it * m->s.dtdp + m->s.statim * 86400.0;
  double timeh = it * m->s.dtdp + (m->s.statim - m->s.reftim) * 86400.0;
  (void)timeloc;
  if (m->s.nramp == 0) {
    m->ramp = m->rampelev = m->rampmete = 1.0;
  } else {
    m->ramp = tanh((2.0 * timeloc / 86400.0) / m->s.drampelev);
    m->rampelev = m->ramp;
    m->rampmete = m->ramp;
  }
  for (int i = 0; i < np; ++i) {
    m->qx1[i] = m->qx2[i];
    m->qy1[i] = m->qy2[i];
    m->uu1[i] = m->uu2[i];
    m->vv1[i] = m->vv2[i];
    m->gwce_lv[i] = 0.0;
    m->mom_lv_x[i] = 0.0;
    m->mom_lv_y[i] = 0.0;
    if (m->s.nws != 0) {
      m->wsx1[i] = m->wsx2[i];
      m->wsx2[i] = 0.0;
      m->wsy1[i] = m->wsy2[i];
      m->wsy2[i] = 0.0;
      m->pr1[i] = m->pr2[i];
      m->pr2[i] = m->s.prbckgrnd_mh2o;
    } else {
      m->pr2[i] = m->s.prbckgrnd_mh2o;
    }
    if (m->ctip) {
      /* src/timestep.F:412-413 */
      m->tip1[i] = m->tip2[i];
      m->tip2[i] = 0.0;
    }
  }
  /* src/timestep.F:841; RampTip shares the elevation ramp's length here */
  if (m->ctip) tide_potential(m, timeh, m->ramp);
  friction(m);
  if (m->nfluxf) {
    /* src/timestep.F:605-612, then this step's QN2 (src/timestep.F:861-903) */
    for (int k = 0; k < m->nvel; ++k) {
      m->qn0[k] = m->qn1[k];
      m->qn1[k] = m->qn2[k];
      m->qn2[k] = m->qn_pending ? m->qn_next[k] : 0.0;
    }
    m->qn_pending = 0;
    /* outward radiation (LBCODEI 30): the flux leaving through the boundary
     * from the last velocities, src/timestep.F:929-940 */
    for (int k = 0; k < m->nvel; ++k)
      if (m->s.lbcodei[k] == 30) {
        const int b = m->s.nbv[k] - 1;
        const double un1 = m->uu1[b] * m->s.csii[k] + m->vv1[b] * m->s.siii[k];
        m->qn1[k] = m->h2[b] * un1;
      }
  }
  if (m->s.nws != 0 && wsx2 && wsy2 && pr2) {
    memcpy(m->wsx2, wsx2, sizeof(double) * np);
    memcpy(m->wsy2, wsy2, sizeof(double) * np);
    memcpy(m->pr2, pr2, sizeof(double) * np);
  }
  int ier = gwce_new(m, timeh);
  if (ier < 0) return ier;
  if (m->s.nolifa == 2) {
    wetdry(m);
    compute_alpha(m);
  }
  for (int i = 0; i < np; ++i) m->h2[i] = dp[i] + m->ifnlfa * m->eta2[i];
  momentum(m);
  /* collectMinMaxData (src/write_output.F:6357-6364): the step stands for
   * ETAMAX_Time = IT*DTDP + STATIM*86400 */
  for (int i = 0; i < np; ++i)
    if (m->nodecode[i] == 1 && m->eta2[i] > m->etamax[i]) {
      m->etamax[i] = m->eta2[i];
      m->etamax_it[i] = it;
    }
  /* the peak speed at wet nodes (src/write_output.F:6367-6376) */
  for (int i = 0; i < np; ++i)
    if (m->nodecode[i] == 1) {
      const double ut = sqrt(m->uu2[i] * m->uu2[i] + m->vv2[i] * m->vv2[i]);
      if (ut > m->umax[i]) {
        m->umax[i] = ut;
        m->umax_it[i] = it;
      }
    }
  return ier;
}

double *adco_array(adco_model *m, const char *name) {
  struct { const char *n; double *p; } t[] = {
      {"eta1", m->eta1}, {"eta2", m->eta2}, {"etas", m->etas}, {"uu2", m->uu2},
      {"vv2", m->vv2}, {"qx2", m->qx2}, {"qy2", m->qy2}, {"tk", m->tk},
      {"etamax", m->etamax}, {"umax", m->umax}, {"totalarea", m->totalarea}, {"coef", m->coef},
      {"gwce_lv", m->gwce_lv}, {"h2", m->h2}, {"h1", m->h1}, {"alphal", m->alphal},
      {"uu1", m->uu1}, {"vv1", m->vv1}, {"wsx2", m->wsx2}, {"wsy2", m->wsy2},
      {"pr2", m->pr2}, {"wsx1", m->wsx1}, {"wsy1", m->wsy1}, {"pr1", m->pr1},
      {"tkm", m->tkm}, {"mom_lv_x", m->mom_lv_x}, {"mom_lv_y", m->mom_lv_y},
      {"tip1", m->tip1}, {"tip2", m->tip2}, {"qn0", m->qn0}, {"qn1", m->qn1},
      {"qn2", m->qn2}};
  for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); ++k)
    if (!strcmp(t[k].n, name)) return t[k].p;
  if (!strcmp(name, "ep")) return &m->ep;
  return NULL;
}

int *adco_iarray(adco_model *m, const char *name) {
  if (!strcmp(name, "nodecode")) return m->nodecode;
  if (!strcmp(name, "noff")) return m->noff;
  if (!strcmp(name, "mju")) return m->mju;
  if (!strcmp(name, "etamax_it")) return m->etamax_it;
  if (!strcmp(name, "umax_it")) return m->umax_it;
  if (!strcmp(name, "neitab")) return m->neitab;
  return NULL;
}

int adco_numitr(const adco_model *m) { return m->numitr; }
int adco_ncchange(const adco_model *m) { return m->ncchange; }

uint64_t adco_state_hash(const adco_model *m) {
  int64_t h = (int64_t)0xCBF29CE484222325ULL;
  h = adco_fnv1a(m->eta2, 8LL * m->np, h);
  h = adco_fnv1a(m->uu2, 8LL * m->np, h);
  h = adco_fnv1a(m->vv2, 8LL * m->np, h);
  h = adco_fnv1a(m->etas, 8LL * m->np, h);
  h = adco_fnv1a(m->tk, 8LL * m->np, h);
  h = adco_fnv1a(m->nodecode, 4LL * m->np, h);
  h = adco_fnv1a(m->noff, 4LL * m->ne, h);
  return (uint64_t)h;
}
