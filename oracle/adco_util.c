/*
 * Oracle utilities (test infrastructure): a state hash and the synthetic
 * Holland (1980) vortex forcing used by the storm test cases.
 */
#include <math.h>
#include <stdint.h>
#include "adco.h"

int64_t adco_fnv1a(const void *p, int64_t nbytes, int64_t seed) {
  uint64_t h = (uint64_t)seed;
  const unsigned char *b = (const unsigned char *)p;
  for (int64_t i = 0; i < nbytes; ++i) {
    h ^= b[i];
    h *= 0x100000001B3ULL;
  }
  return (int64_t)h;
}

/* Holland (1980) parametric vortex: pressure P(r) = Pc + dP exp(-(Rmax/r)^B),
 * gradient wind from the cyclostrophic/geostrophic balance, reduced to the
 * surface by a constant factor and turned inward by a constant inflow angle;
 * stress by the Garratt (1977) drag law capped at 0.0035.  The profile is the
 * one src/nws08.F90:616-713 evaluates; the surface reduction and drag here are
 * this build's own fixed choices for synthetic storms. */
void adco_holland(int np, const double *x, const double *y, double t,
                  const double *prm, double *wsx, double *wsy, double *pr) {
  const double pc = prm[0] * 100.0, pn = prm[1] * 100.0, rmax = prm[2], B = prm[3];
  const double xc = prm[4] + prm[6] * t, yc = prm[5] + prm[7] * t, f = prm[8];
  const double rampd = prm[9], rhoa = prm[10], wfac = prm[11];
  const double beta = prm[12] * 3.141592653589793 / 180.0;
  const double rho0 = 1000.0, g = 9.80665;
  const double ramp = rampd > 0.0 ? tanh((2.0 * t / 86400.0) / rampd) : 1.0;
  const double dpr = pn - pc, cb = cos(beta), sb = sin(beta);
  for (int i = 0; i < np; ++i) {
    double dx = x[i] - xc, dy = y[i] - yc;
    double r = sqrt(dx * dx + dy * dy);
    if (r < 1.0) r = 1.0;
    double a = pow(rmax / r, B);
    double e = exp(-a);
    double p = pc + dpr * e;
    double rf2 = 0.5 * r * f;
    double vg = sqrt(B / rhoa * a * dpr * e + rf2 * rf2) - rf2;
    double w = wfac * vg;
    double tx = -dy / r, ty = dx / r; /* counterclockwise tangent */
    double rx = dx / r, ry = dy / r;
    double wx = w * (cb * tx - sb * rx), wy = w * (cb * ty - sb * ry);
    double wm = sqrt(wx * wx + wy * wy);
    double cd = 0.001 * (0.75 + 0.067 * wm);
    if (cd > 0.0035) cd = 0.0035;
    double s = ramp * cd * (rhoa / rho0) * wm;
    wsx[i] = s * wx;
    wsy[i] = s * wy;
    pr[i] = (pn + ramp * (p - pn)) / (rho0 * g);
  }
}
