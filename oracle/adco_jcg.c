/*
 * ITPACKV 2D Jacobi-conjugate-gradient, restated from src/itpackv.F for the
 * oracle (test infrastructure; never linked into the product).
 *
 * Arrays keep the reference layout: coef(ndim,maxnz) and jcoef(ndim,maxnz)
 * column-major with 1-based column numbers, so COEF(i,j) is coef[(j-1)*ndim +
 * (i-1)].  Every floating-point expression keeps the reference's evaluation
 * order; build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "adco.h"

#define C(i, j) coef[(size_t)((j)-1) * ndim + ((i)-1)]
#define J(i, j) jcoef[(size_t)((j)-1) * ndim + ((i)-1)]

/* Blocked tree reduction; the GPU kernels use exactly this shape:
 * chunks of 2048 elements, thread t (0..511) of a chunk sums elements
 * t + 512*r (r = 0..3) sequentially, the 64 lanes of a wave combine with an
 * xor butterfly (offsets 32..1), the 8 wave sums of a chunk combine as
 * ((w0+w1)+(w2+w3))+((w4+w5)+(w6+w7)); the chunk sums are then reduced by
 * 256 threads (thread t sums chunk sums t, t+256, ...), the butterfly, and
 * (w0+w1)+(w2+w3) over those 4 waves. */
static double wave_butterfly(const double *v) {
  double a[64], b[64];
  memcpy(a, v, sizeof(a));
  for (int off = 32; off >= 1; off >>= 1) {
    for (int i = 0; i < 64; ++i) b[i] = a[i] + a[i ^ off];
    memcpy(a, b, sizeof(a));
  }
  return a[0];
}

static double combine4(const double *thr /* [256] */) {
  double w0 = wave_butterfly(thr), w1 = wave_butterfly(thr + 64);
  double w2 = wave_butterfly(thr + 128), w3 = wave_butterfly(thr + 192);
  return (w0 + w1) + (w2 + w3);
}

static double combine8(const double *thr /* [512] */) {
  return combine4(thr) + combine4(thr + 256);
}

static double dot_blocked(int n, const double *x, const double *y) {
  int nch = (n + ADCO_RED_CHUNK - 1) / ADCO_RED_CHUNK;
  if (nch == 0) return 0.0;
  double thr[ADCO_RED_THREADS];
  double stackbuf[4096];
  double *part = (nch <= 4096) ? stackbuf : (double *)malloc(sizeof(double) * nch);
  for (int c = 0; c < nch; ++c) {
    for (int t = 0; t < ADCO_RED_THREADS; ++t) {
      double acc = 0.0;
      for (int r = 0; r < ADCO_RED_PER_THREAD; ++r) {
        long k = (long)c * ADCO_RED_CHUNK + t + (long)ADCO_RED_THREADS * r;
        if (k < n) acc = acc + x[k] * y[k];
      }
      thr[t] = acc;
    }
    part[c] = combine8(thr);
  }
  for (int t = 0; t < ADCO_RED_XTHREADS; ++t) {
    double acc = 0.0;
    for (int c = t; c < nch; c += ADCO_RED_XTHREADS) acc = acc + part[c];
    thr[t] = acc;
  }
  double r = combine4(thr);
  if (part != stackbuf) free(part);
  return r;
}

/* sdot, src/itpackv.F:2828-2845 */
static double dot_seq(int n, const double *x, const double *y) {
  double d = 0.0;
  for (int i = 0; i < n; ++i) d = d + x[i] * y[i];
  return d;
}

double adco_dot(int order, int n, const double *x, const double *y) {
  return order == ADCO_ORDER_BLOCKED ? dot_blocked(n, x, y) : dot_seq(n, x, y);
}

void adco_itpack_init(adco_itpack *s, int order) {
  memset(s, 0, sizeof(*s));
  s->order = order;
}

/* dfault, src/itpackv.F:1071-1170 */
void adco_dfault(int *iparm, double *rparm) {
  volatile double srelpr = 1.0, temp;
  do {
    srelpr = 0.5 * srelpr;
    temp = srelpr + 1.0;
  } while (temp > 1.0);
  srelpr = 2.0 * srelpr;
  static const int ip[12] = {100, -1, 0, 6, 0, 1, 1, 0, -2, 0, 0, 3};
  for (int i = 0; i < 12; ++i) iparm[i] = ip[i];
  rparm[0] = 5.0e-6;
  rparm[1] = 0.0;
  rparm[2] = 0.0;
  rparm[3] = .75;
  rparm[4] = 1.0;
  rparm[5] = 0.0;
  rparm[6] = .25;
  rparm[7] = 128.0 * srelpr;
  for (int i = 8; i < 12; ++i) rparm[i] = 0.0;
}

static double srelpr_value(void) {
  volatile double srelpr = 1.0, temp;
  do {
    srelpr = 0.5 * srelpr;
    temp = srelpr + 1.0;
  } while (temp > 1.0);
  return 2.0 * srelpr;
}

/* echall(icall=1), src/itpackv.F:1172-1255 */
static void echall(adco_itpack *s, const int *iparm, const double *rparm) {
  s->zeta = rparm[0];
  s->cme = rparm[1];
  s->sme = rparm[2];
  s->ff = rparm[3];
  s->omega = rparm[4];
  s->specr = rparm[5];
  s->betab = rparm[6];
  s->itmax = iparm[0];
  s->level = iparm[1];
  s->isym = iparm[4];
  s->adapt = 0;
  s->partad = 0;
  s->betadt = 0;
  if (iparm[5] == 1 || iparm[5] == 3) s->adapt = 1;
  if (iparm[5] == 1) s->betadt = 1;
  if (iparm[5] == 2) s->partad = 1;
  s->caseii = 0;
  if (iparm[6] == 2) s->caseii = 1;
  if (s->caseii) s->sme = -s->cme;
  if (!s->caseii && s->sme == 0.0) s->sme = -1.0;
  s->spr = s->sme;
  s->in = 0;
  s->is = 0;
  s->halt = 0;
  s->bdelnm = s->delnnm = s->delsnm = s->gamma = s->qa = s->qt = s->rho = 0.0;
  s->rrr = s->sige = s->stptst = s->udnm = 0.0;
}

/* determ, src/itpackv.F:1029-1068.  tri(ldt,*) column-major. */
static double determ(int ldt, int n, const double *tri, double *wk1,
                     double *wk2, double xlmda) {
#define TRI(i, k) tri[(size_t)((k)-1) * ldt + ((i)-1)]
  for (int i = 1; i <= n; ++i) wk1[i - 1] = TRI(i, 1) - xlmda;
  wk2[n - 1] = wk1[n - 1];
  wk2[n - 2] = wk1[n - 2] * wk2[n - 1] + TRI(n, 2);
  if (n != 2) {
    for (int l = n - 2; l >= 1; --l)
      wk2[l - 1] = wk1[l - 1] * wk2[l] + TRI(l + 1, 2) * wk2[l + 1];
  }
  return wk2[0];
#undef TRI
}

/* zbrent, src/itpackv.F:2521-2781 (eps = 0) */
static void zbrent(adco_itpack *s, int n, double *tri, double eps, int nsig,
                   double *pa, double *pb, int *pmaxfn, int *ier) {
  double a = *pa, b = *pb, c = 0, d = 0, e = 0, p, q, r, ss, t, fa, fb, fc = 0,
         tol, rm, rone, temp;
  int maxfn = *pmaxfn, ic;
  int itmax = s->itmax;
  double *wk1 = tri + 2 * (size_t)itmax, *wk2 = tri + 3 * (size_t)itmax;
  *ier = 0;
  t = pow(10.0, (double)(-nsig));
  ic = 2;
  fa = determ(itmax, n, tri, wk1, wk2, a);
  fb = determ(itmax, n, tri, wk1, wk2, b);
  ss = b;
  if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0.0)) {
    *ier = 502;
    maxfn = ic;
    goto done;
  }
L10:
  c = a;
  fc = fa;
  d = b - c;
  e = d;
L20:
  if (!(fabs(fc) >= fabs(fb))) {
    a = b;
    b = c;
    c = a;
    fa = fb;
    fb = fc;
    fc = fa;
  }
  tol = t * fmax(fabs(b), 0.1);
  rm = (c - b) * 0.5;
  if (fabs(fb) <= eps) goto L80;
  if (fabs(c - b) <= tol) goto L80;
  if (ic >= maxfn) goto L90;
  if (fabs(e) < tol) goto L60;
  if (fabs(fa) <= fabs(fb)) goto L60;
  ss = fb / fa;
  if (a != c) goto L40;
  p = (c - b) * ss;
  q = 1.0 - ss;
  goto L50;
L40:
  q = fa / fc;
  r = fb / fc;
  rone = r - 1.0;
  p = ss * ((c - b) * q * (q - r) - (b - a) * rone);
  q = (q - 1.0) * rone * (ss - 1.0);
L50:
  if (p > 0.0) q = -q;
  if (p < 0.0) p = -p;
  ss = e;
  e = d;
  if (p + p >= 3.0 * rm * q) goto L60;
  if (p + p >= fabs(ss * q)) goto L60;
  d = p / q;
  goto L70;
L60:
  e = rm;
  d = e;
L70:
  a = b;
  fa = fb;
  temp = d;
  if (fabs(temp) <= 0.5 * tol) temp = copysign(0.5 * tol, rm);
  b = b + temp;
  ss = b;
  fb = determ(itmax, n, tri, wk1, wk2, ss);
  ic = ic + 1;
  if (fb == 0.0 || fc == 0.0) goto L20;
  if ((fb > 0.0 && fc < 0.0) || (fb < 0.0 && fc > 0.0)) goto L20;
  goto L10;
L80:
  a = c;
  maxfn = ic;
  goto done;
L90:
  *ier = 501;
  a = c;
  maxfn = ic;
done:
  *pa = a;
  *pb = b;
  *pmaxfn = maxfn;
}

/* eigvss, src/itpackv.F:1361-1399 */
static double eigvss(adco_itpack *s, int n, double *tri, double start,
                     double zeta, int itmax, int *ier) {
  int itmp = (int)(-log10(fabs(zeta)));
  int nsig = itmp > 4 ? itmp : 4;
  int maxfn = itmax > 50 ? itmax : 50;
  double a = start, b = 1.0;
  zbrent(s, n, tri, 0.0, nsig, &a, &b, &maxfn, ier);
  return b;
}

/* chgcon(ibmth=1), src/itpackv.F:893-1027 (isym = 0 path) */
static void chgcon(adco_itpack *s, int ldt, double *tri, double gamold,
                   double rhoold) {
#define TRI(i, k) tri[(size_t)((k)-1) * ldt + ((i)-1)]
  double start = s->cme, end1;
  int ip = s->in, ier = 0;
  if (ip >= 2) goto L60;
  if (ip == 1) goto L50;
  end1 = 0.0;
  s->cmold = 0.0;
  goto L110;
L50:
  end1 = 1.0 - 1.0 / s->gamma;
  TRI(1, 1) = end1;
  TRI(1, 2) = 0.0;
  goto L110;
L60:
  if (fabs(start - s->cmold) <= s->zeta * start) {
    s->adapt = 0;
    s->partad = 0;
    return;
  }
  s->cmold = start;
  TRI(ip, 1) = 1.0 - 1.0 / s->gamma;
  TRI(ip, 2) = (1.0 - s->rho) / (s->rho * rhoold * s->gamma * gamold);
  /* isym /= 0 would use eigvns; ADCIRC's SPD systems never negate rows */
  end1 = eigvss(s, ip, tri, start, s->zeta, s->itmax, &ier);
  if (ier != 0) {
    s->adapt = 0;
    s->partad = 0;
    return;
  }
L110:
  s->cme = end1;
#undef TRI
}

/* pstop_nrms(iflag=1), src/itpackv.F:809-890 */
static void pstop_nrms(adco_itpack *s, double unrm, double dnrm, double ccon) {
  double con = ccon, uold, tr, tl;
  s->halt = 0;
  if (s->in >= 1) goto L10;
  s->q1 = 0;
  s->udnm = 1.0;
  s->stptst = 1000.0;
L10:
  if (s->q1) goto L20;
  if (s->in > 5 && (s->in % 5) != 0) goto L20;
  uold = s->udnm;
  s->udnm = unrm;
  if (s->udnm == 0.0) s->udnm = 1.0;
  if (s->in > 5 && fabs(s->udnm - uold) <= s->udnm * s->zeta) s->q1 = 1;
L20:
  tr = sqrt(s->udnm);
  tl = 1.0;
  if (con == 1.0) goto L40;
  tl = sqrt(dnrm);
  tr = tr * (1.0 - con);
L40:
  s->stptst = tl / tr;
  if (tl >= tr * s->zeta) return;
  s->halt = 1;
}

/* parcon(ibmth=1), src/itpackv.F:1774-1851 */
static void parcon(adco_itpack *s, double dtnrm, double *c1, double *c2,
                   double *c3, double *c4, double *gamold, double *rhotmp) {
  int ip = s->in - s->is;
  double rhoold = s->rho;
  *gamold = s->gamma;
  s->gamma = 1.0 / (1.0 - dtnrm / s->delnnm);
  s->rho = 1.0;
  if (ip != 0)
    s->rho = 1.0 / (1.0 - s->gamma * s->delnnm / (*gamold * s->delsnm * rhoold));
  s->delsnm = s->delnnm;
  *rhotmp = rhoold;
  *c1 = s->rho * s->gamma;
  *c2 = s->rho;
  *c3 = 1.0 - s->rho;
  *c4 = s->rho * (1.0 - s->gamma);
}

/* pjac, src/itpackv.F:1932-1967: rhs = rhs - N*u, column by column */
static void pjac(int n, int ndim, int maxnz, const int *jcoef,
                 const double *coef, const double *u, double *rhs) {
  for (int j = 2; j <= maxnz; ++j)
    for (int i = 1; i <= n; ++i)
      rhs[i - 1] = rhs[i - 1] - C(i, j) * u[J(i, j) - 1];
}

/* itjcg, src/itpackv.F:459-604 */
static void itjcg(adco_itpack *s, int n, int ndim, int maxnz, const int *jcoef,
                  const double *coef, double *u, double *u1, double *d,
                  double *d1, double *dtwd, double *tri) {
  double delnnm, dtnrm, unorm, c1, c2, c3, c4;
  if (s->adapt) chgcon(s, s->itmax, tri, s->gamold, s->rhoold);
  for (int i = 0; i < n; ++i) dtwd[i] = 0.0;
  pjac(n, ndim, maxnz, jcoef, coef, d, dtwd);
  if (s->q1 || (s->in > 5 && (s->in % 5) != 0)) {
    delnnm = adco_dot(s->order, n, d, d);
    dtnrm = adco_dot(s->order, n, d, dtwd);
    unorm = 1.0;
  } else {
    delnnm = adco_dot(s->order, n, d, d);
    dtnrm = adco_dot(s->order, n, d, dtwd);
    unorm = adco_dot(s->order, n, u, u);
  }
  s->delnnm = delnnm;
  pstop_nrms(s, unorm, delnnm, s->cme);
  if (s->halt) return;
  parcon(s, dtnrm, &c1, &c2, &c3, &c4, &s->gamold, &s->rhoold);
  for (int i = 0; i < n; ++i) {
    u1[i] = c1 * d[i] + c2 * u[i] + c3 * u1[i];
    d1[i] = c1 * dtwd[i] + c4 * d[i] + c3 * d1[i];
  }
}

/* scal, src/itpackv.F:2281-2421 */
static int scal(adco_itpack *s, int n, int ndim, int maxnz, int *jcoef,
                double *coef, double *rhs, double *u, double *work) {
  int nsgncg = 0;
  if (n <= 0) return 0;
  for (int i = 1; i <= n; ++i) {
    if (J(i, 1) != i) {
      if (maxnz < 2) return 402;
      for (int j = 2; j <= maxnz; ++j) {
        if (J(i, j) == i) {
          double sv = C(i, j);
          C(i, j) = C(i, 1);
          J(i, j) = J(i, 1);
          C(i, 1) = sv;
          J(i, 1) = i;
        }
      }
    }
    if (C(i, 1) < 0.0) {
      for (int j = 1; j <= maxnz; ++j) C(i, j) = -C(i, j);
      rhs[i - 1] = -rhs[i - 1];
      nsgncg = nsgncg + 1;
    } else if (C(i, 1) == 0.0) {
      return 401;
    }
  }
  if (maxnz >= 2) {
    for (int j = 2; j <= maxnz; ++j)
      for (int i = 1; i <= n; ++i)
        if (J(i, j) <= 0) J(i, j) = i;
  }
  for (int i = 0; i < n; ++i) work[i] = sqrt(C(i + 1, 1));
  for (int i = 0; i < n; ++i) u[i] = u[i] * work[i];
  for (int i = 0; i < n; ++i) work[i] = 1.0 / work[i];
  for (int i = 0; i < n; ++i) C(i + 1, 1) = work[i];
  for (int i = 0; i < n; ++i) rhs[i] = rhs[i] * work[i];
  if (maxnz < 2) return 0;
  for (int j = 2; j <= maxnz; ++j)
    for (int i = 1; i <= n; ++i)
      C(i, j) = C(i, j) * work[i - 1] * work[J(i, j) - 1];
  if (nsgncg > 0 && nsgncg < n) s->isym = 1;
  return 0;
}

/* unscal, src/itpackv.F:2424-2479 */
static void unscal(int n, int ndim, int maxnz, int *jcoef, double *coef,
                   double *rhs, double *u, double *work) {
  for (int i = 0; i < n; ++i) work[i] = C(i + 1, 1);
  for (int i = 0; i < n; ++i) u[i] = u[i] * work[i];
  for (int i = 0; i < n; ++i) work[i] = 1.0 / work[i];
  for (int i = 0; i < n; ++i) rhs[i] = rhs[i] * work[i];
  if (maxnz >= 2) {
    for (int j = 2; j <= maxnz; ++j)
      for (int i = 1; i <= n; ++i)
        C(i, j) = C(i, j) * work[i - 1] * work[J(i, j) - 1];
    for (int j = 2; j <= maxnz; ++j)
      for (int i = 1; i <= n; ++i)
        if (J(i, j) == i) J(i, j) = 0;
  }
  for (int i = 0; i < n; ++i) C(i + 1, 1) = work[i] * work[i];
}

/* peror, src/itpackv.F:606-700 (diagnostic digits only; kept so the CPU
 * baseline does the reference's full per-solve work) */
static void peror(adco_itpack *s, int n, int ndim, int maxnz, const int *jcoef,
                  const double *coef, const double *rhs, const double *u,
                  double *work, double *digit1, double *digit2) {
  double bnrm, rnrm, temp;
  *digit1 = 0.0;
  *digit2 = 0.0;
  if (n <= 0) return;
  *digit1 = -log10(fabs(s->srelpr));
  if (s->stptst > 0.0) *digit1 = -log10(fabs(s->stptst));
  for (int i = 0; i < n; ++i) work[i] = rhs[i] / C(i + 1, 1);
  bnrm = adco_dot(s->order, n, work, work);
  if (bnrm == 0.0) goto L30;
  /* pmult: work = u - N u */
  for (int i = 0; i < n; ++i) work[i] = u[i];
  for (int j = 2; j <= maxnz; ++j)
    for (int i = 1; i <= n; ++i)
      work[i - 1] = work[i - 1] - C(i, j) * u[J(i, j) - 1];
  for (int i = 0; i < n; ++i) work[i] = (rhs[i] - work[i]) / C(i + 1, 1);
  rnrm = adco_dot(s->order, n, work, work);
  temp = rnrm / bnrm;
  if (temp == 0.0) goto L30;
  *digit2 = -log10(fabs(temp)) / 2.0;
  return;
L30:
  *digit2 = -log10(fabs(s->srelpr));
}

/* jcg, src/itpackv.F:63-456 for iparm(9) = -2, iparm(10) = 0 */
int adco_jcg(adco_itpack *s, int n, int ndim, int maxnz, int *jcoef,
             double *coef, double *rhs, double *u, int nw, double *wksp,
             int *iparm, double *rparm) {
  int ier = 0, ib1, ib2, ib3, ib4, ib5, itmax1;
  double temp, digit1 = 0, digit2 = 0;
  s->srelpr = srelpr_value();
  s->level = iparm[1];
  if (iparm[0] <= 0) return 0;
  echall(s, iparm, rparm);
  temp = 512.0 * s->srelpr;
  if (!(s->zeta >= temp)) s->zeta = temp;
  if (n <= 0) return 11;
  ier = scal(s, n, ndim, maxnz, jcoef, coef, rhs, u, wksp);
  if (ier != 0) return ier; /* go to 370 */
  ib1 = 0;
  ib2 = ib1 + n;
  ib3 = ib2 + n;
  ib4 = ib3 + n;
  ib5 = ib4 + n;
  iparm[7] = 4 * n + 4 * s->itmax;
  if (nw < iparm[7]) {
    ier = 12;
    goto L340;
  }
  for (int i = 0; i < nw; ++i) wksp[i] = 0.0;
  memcpy(wksp + ib2, rhs, sizeof(double) * n);
  pjac(n, ndim, maxnz, jcoef, coef, u, wksp + ib2);
  for (int i = 0; i < n; ++i) wksp[n + i] = wksp[n + i] - u[i];
  itmax1 = s->itmax + 1;
  {
    int loop;
    for (loop = 1; loop <= itmax1; ++loop) {
      s->in = loop - 1;
      if (s->in % 2 == 1) {
        itjcg(s, n, ndim, maxnz, jcoef, coef, wksp + ib1, u, wksp + ib3,
              wksp + ib2, wksp + ib4, wksp + ib5);
      } else {
        itjcg(s, n, ndim, maxnz, jcoef, coef, u, wksp + ib1, wksp + ib2,
              wksp + ib3, wksp + ib4, wksp + ib5);
      }
      if (s->halt) break;
    }
    if (!s->halt) {
      ier = 13;
      if (iparm[2] == 0) rparm[0] = s->stptst;
    }
  }
  if (s->in % 2 == 1) memcpy(u, wksp, sizeof(double) * n);
L340:
  if (iparm[11] >= 0)
    peror(s, n, ndim, maxnz, jcoef, coef, rhs, u, wksp, &digit1, &digit2);
  unscal(n, ndim, maxnz, jcoef, coef, rhs, u, wksp);
  iparm[7] = iparm[7] - 4 * (s->itmax - s->in);
  if (iparm[2] == 0) {
    iparm[0] = s->in;
    iparm[8] = -2;
    rparm[1] = s->cme;
    rparm[2] = s->sme;
    rparm[10] = digit1;
    rparm[11] = digit2;
  }
  return ier;
}
