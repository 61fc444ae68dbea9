// CPU oracle for the PB-SAM mutual-polarization path -- TEST INFRASTRUCTURE.
//
// A restatement, in this project's own flat layout, of the reference algorithms
// on the hot path.  Each function cites the reference file:line it follows
// (paths relative to the reference's src/).  Only tests/, the smoke check and
// the benchmark's CPU baseline load this library; the product never does.
//
// Floating-point evaluation order follows the reference expressions so that
// function-level results agree to a few ulps.
#include "pbsam_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

typedef std::complex<double> cplx;
typedef std::vector<double> vec;

static const double SOLVE_TOL = 1e-8;      // expcenter.cpp:17 (__ACCURATE__)
static const int SOLVE_MAX_CT = 100;       // expcenter.cpp:18
static const double MAX_ERROR = 1e-6;      // xforms.h:14
static const double MINSEPDIST = 5.0;      // molecule.h:22
static const double PREC_LIMIT = 1e-30;    // expansion.h:11
static const double EPS_SIN_THETA = 1e-12; // rotcoeff.cpp:3

// ----------------------------------------------------------------------------
// context and tables
// ----------------------------------------------------------------------------
struct orc_ctx {
  int PM;  // max p supported
  int NL;  // 2*PM: SH/rotation table rows
  double kappa, sdiel, kapsqr;
  vec c1, c2, c3, c6, c4;  // expansion.cpp:46-96
  vec zeta, eta;           // rotcoeff.cpp:16-33
  vec Q;                   // rotcoeff.cpp:950-983
  vec alpha, beta, gam, del; // transcoeff.cpp:16-38
  double C1(int n, int m) const { return c1[n * NL + m]; }
  double C2(int n, int m) const { return c2[n * NL + m]; }
  double C3(int n, int m) const { return c3[n * NL + m]; }
  double ZETA(int n, int m) const { return zeta[n * NL + m]; }
  double ETA(int n, int m) const { return eta[n * (2 * NL) + NL + m]; }
  double ALPHA(int n, int m) const { return alpha[n * PM + m]; }
  double BETA(int n, int m) const { return beta[n * PM + m]; }
  double GAMMA(int n, int m) const { return gam[n * (2 * PM) + PM + m]; }
  double DELTA(int n, int m) const { return del[n * (2 * PM) + PM + m]; }
  double QC(int n, int m, int k) const { return Q[(n * PM + m) * 2 + k]; }
};

extern "C" orc_ctx *orc_create(double kappa, double eps_out, int pmax) {
  orc_ctx *c = new orc_ctx();
  c->PM = pmax;
  c->NL = 2 * pmax;
  c->kappa = kappa;
  c->sdiel = eps_out;
  c->kapsqr = (kappa == 0 ? 0 : 1.0);
  const int NL = c->NL;
  c->c1.assign(NL * NL, 0.0);
  c->c2.assign(NL * NL, 0.0);
  c->c3.assign(NL * NL, 0.0);
  c->c6.assign(NL, 0.0);
  c->c4.assign(NL, 0.0);
  // expansion.cpp:76-95
  vec temp(2 * NL);
  temp[0] = 1.0;
  for (int i = 1; i < 2 * NL; i++) temp[i] = temp[i - 1] * sqrt((double)i);
  for (int n = 0; n < NL; n++)
    for (int m = 0; m <= n; m++) {
      c->c1[n * NL + m] = (2 * n - 1) / (double)(n - m);
      c->c2[n * NL + m] = (n + m - 1) / (double)(n - m);
      c->c3[n * NL + m] = temp[n - m] / temp[n + m];
    }
  c->c6[0] = 1.0;
  c->c6[1] = 1.0;
  for (int n = 2; n < NL; n++) c->c6[n] = c->c6[n - 1] * (2 * n - 1);
  // expansion.cpp:46-47
  for (int n = 0; n < NL; n++) c->c4[n] = 1.0 / ((2 * n - 1) * (2 * n - 3));
  // rotcoeff.cpp:18-30 (table sized so that ZETA(2N-1,2N-1) stays in bounds)
  c->zeta.assign(NL * NL, 0.0);
  c->eta.assign(NL * 2 * NL, 0.0);
  for (int n = 0; n < NL; n++)
    for (int m = 0; m <= n; m++)
      c->zeta[n * NL + m] = sqrt((double)(n + m + 1) * (n - m + 1) / ((2 * n + 1) * (2 * n + 3)));
  for (int n = 0; n < NL; n++)
    for (int m = 0; m <= n; m++) {
      c->eta[n * 2 * NL + NL + m] = sqrt((double)(n - m - 1) * (n - m) / ((2 * n + 1) * (2 * n - 1)));
      if (m != 0)
        c->eta[n * 2 * NL + NL - m] =
            -sqrt((double)(n + m - 1) * (n + m) / ((2 * n + 1) * (2 * n - 1)));
    }
  // rotcoeff.cpp:950-983 (specialSH expansion.cpp:526-546)
  {
    int PM = pmax;
    c->Q.assign((2 * PM - 1) * PM * 2, 0.0);
    int n2 = 2 * PM - 1;
    vec SH(n2);
    if (n2 > 0) SH[0] = 0;
    if (n2 > 1) SH[1] = -1;
    if (n2 > 2) SH[2] = -1.0 * 3;
    for (int l = 3; l < n2; l++) SH[l] = (1.0 * (2 * l - 1) * SH[l - 1] - l * SH[l - 2]) / (l - 1);
    for (int l = 1; l < n2; l++) SH[l] *= (-c->C3(l, 1));
    auto Qr = [&](int n, int m, int k) -> double & { return c->Q[(n * PM + m) * 2 + k]; };
    for (int n = 1; n < n2; n++) {
      Qr(n, 0, 0) = 0.0;
      Qr(n, 0, 1) = SH[n];
    }
    for (int n = 2; n < n2; n++) {
      Qr(n - 1, 1, 0) = (c->ETA(n, -1) * Qr(n, 0, 1) + c->ZETA(n - 1, 0)) / c->ETA(n, 0);
      if (n > 2)
        Qr(n - 1, 1, 1) = (c->ETA(n, 1) / c->ETA(n, 0)) * Qr(n, 0, 1);
      else
        Qr(n - 1, 1, 1) = 0.0;
    }
    for (int m = 1; m < PM - 1; m++)
      for (int n = m + 2; n < 2 * PM - m - 1; n++) {
        Qr(n - 1, m + 1, 0) = (c->ETA(n, m - 1) * Qr(n, m, 0) + c->ZETA(n - 1, m)) / c->ETA(n, m);
        if (n > m + 2)
          Qr(n - 1, m + 1, 1) = (c->ETA(n, m + 1) / c->ETA(n, m)) * Qr(n, m, 1);
        else
          Qr(n - 1, m + 1, 1) = 0.0;
      }
  }
  // transcoeff.cpp:18-33
  {
    int PM = pmax;
    c->alpha.assign(NL * PM, 0.0);
    c->beta.assign(NL * PM, 0.0);
    c->gam.assign(NL * 2 * PM, 0.0);
    c->del.assign(NL * 2 * PM, 0.0);
    double kapsqr = c->kapsqr;
    for (int n = 0; n < NL; n++)
      for (int m = 0; m < PM; m++) {
        double a = sqrt((double)(n + m + 1) * (n - m + 1));
        c->alpha[n * PM + m] = a;
        c->beta[n * PM + m] = kapsqr * a / ((2 * n + 1) * (2 * n + 3));
        double g = sqrt((double)(n - m - 1) * (n - m));
        c->gam[n * 2 * PM + PM + m] = g;
        c->del[n * 2 * PM + PM + m] = kapsqr * g / ((2 * n + 1) * (2 * n - 1));
        if (m != 0) {
          double gn = -sqrt((double)(n + m - 1) * (n + m));
          c->gam[n * 2 * PM + PM - m] = gn;
          c->del[n * 2 * PM + PM - m] = kapsqr * gn / ((2 * n + 1) * (2 * n - 1));
        }
      }
  }
  return c;
}

extern "C" void orc_destroy(orc_ctx *c) { delete c; }

// ----------------------------------------------------------------------------
// geometry helpers (util.h:404-468)
// ----------------------------------------------------------------------------
static inline void cart_to_sph(const double *c, double &rho, double &theta, double &phi) {
  rho = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  if (rho < fabs(c[2])) rho = fabs(c[2]);
  theta = (rho == 0.0) ? 0.0 : acos(c[2] / rho);
  phi = (c[0] == 0.0 && c[1] == 0.0) ? 0.0 : atan2(c[1], c[0]);
}
static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double *a, const double *b, double *o) {
  double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
  o[0] = x; o[1] = y; o[2] = z;
}
// CQuat * CPnt (util.h:463-468); q = (w,x,y,z)
static inline void quat_rot(const double *q, const double *r, double *o) {
  double qr[3], t[3], u[3];
  cross3(q + 1, r, qr);
  cross3(q + 1, qr, t);
  for (int i = 0; i < 3; i++) u[i] = q[0] * qr[i] + t[i];
  double x = r[0] + 2 * u[0], y = r[1] + 2 * u[1], z = r[2] + 2 * u[2];
  o[0] = x; o[1] = y; o[2] = z;
}
static inline void quat_normalize(double *q) {  // util.h:261-265
  double in = 1.0 / sqrt((q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + q[0] * q[0]);
  for (int i = 1; i < 4; i++) q[i] *= in;
  q[0] *= in;
}

// ----------------------------------------------------------------------------
// spherical harmonics and Bessel functions
// ----------------------------------------------------------------------------
// CSHExpan::computecoeff + legendre (non-SSE): expansion.cpp:483-518, 634-675
static void sh_eval(const orc_ctx *c, int P, double theta, double phi, double *M) {
  double negterm = 1.0;
  double cost = cos(theta), sint = sin(theta);
  double sqrtterm = 1.0;
  for (int i = 0; i < P * P; i++) M[i] = 0.0;
  M[0] = c->c6[0];
  negterm = -negterm;
  sqrtterm *= sint;
  if (P > 1) {
    M[1] = cost * M[0];
    for (int n = 2; n < P; n++)
      M[n * n] = cost * c->C1(n, 0) * M[(n - 1) * (n - 1)] - c->C2(n, 0) * M[(n - 2) * (n - 2)];
  }
  for (int m = 1, k = 1; m < P; m++, k += 2) {
    M[m * m + k] = negterm * c->c6[m] * sqrtterm;
    M[m * m + k + 1] = M[m * m + k];
    negterm = -negterm;
    sqrtterm *= sint;
    if (m < P - 1) {
      M[(m + 1) * (m + 1) + k] = cost * (2 * m + 1) * M[m * m + k];
      M[(m + 1) * (m + 1) + k + 1] = M[(m + 1) * (m + 1) + k];
      for (int n = m + 2; n < P; n++) {
        M[n * n + k] = cost * c->C1(n, m) * M[(n - 1) * (n - 1) + k] - c->C2(n, m) * M[(n - 2) * (n - 2) + k];
        M[n * n + k + 1] = M[n * n + k];
      }
    }
  }
  std::vector<double> cosp(P > 0 ? P : 1), sinp(P > 0 ? P : 1);
  double ang = 0;
  for (int m = 0; m < P; m++, ang += phi) {
    cosp[m] = cos(ang);
    sinp[m] = sin(ang);
  }
  double *it = M;
  for (int n = 0; n < P; n++) {
    *(it++) *= c->C3(n, 0);
    double s = -1.0;
    for (int m = 1; m <= n; m++, s = -s) {
      *(it++) *= s * c->C3(n, m) * cosp[m];
      *(it++) *= s * c->C3(n, m) * sinp[m];
    }
  }
}

// CLExpan::BESSEL expansion.cpp:422-436 (K pre-filled with 1.0)
static void bessel_k(const orc_ctx *c, int L, double val, double *K) {
  for (int i = 0; i < L; i++) K[i] = 1.0;
  if (val != 0.0) {
    if (L > 1) {
      K[1] += val;
      double valsqr = val * val;
      for (int n = 2; n < L; n++) K[n] = K[n - 1] + valsqr * K[n - 2] * c->c4[n];
    }
  }
}
// CMExpan::BESSEL expansion.cpp:348-365
static void bessel_i(int L, double val, double *K) {
  for (int i = 0; i < L; i++) K[i] = 1.0;
  if (val != 0.0) {
    double z = 0.5 * val * val;
    for (int n = 0; n < L; n++) {
      double t = z / (2 * n + 3);
      for (int j = 1; j <= 20; j++) {
        K[n] += t;
        t *= (z / ((j + 1) * (2 * (n + j) + 3)));
        if (t < 1e-20) break;
      }
    }
  }
}

extern "C" void orc_sh(orc_ctx *c, int P, double theta, double phi, double *out) {
  sh_eval(c, P, theta, phi, out);
}
extern "C" void orc_bessel_k(orc_ctx *c, int L, double val, double *out) { bessel_k(c, L, val, out); }
extern "C" void orc_bessel_i(orc_ctx *c, int L, double val, double *out) {
  (void)c;
  bessel_i(L, val, out);
}

// single-charge local expansion added into L (CLExpan, expansion.cpp:289-297,305-316)
static void add_local_1(const orc_ctx *c, double ch, const double *pos, int p, bool bKappa,
                        double scale, double *L, double *Y, double *bes) {
  double rho, th, ph;
  cart_to_sph(pos, rho, th, ph);
  sh_eval(c, p, th, ph, Y);
  double val = bKappa ? rho * c->kappa : 0.0;
  bessel_k(c, p + 1, val, bes);
  double mr = 1.0 / rho;
  double mk = ch;
  mk *= mr * (bKappa ? exp(-val) : 1.0);
  mr *= scale;
  const double *it = Y;
  for (int n = 0; n < p; n++) {
    double f = bes[n] * mk;
    for (int m = 0; m < 2 * n + 1; m++) L[n * n + m] += *(it++) * f;
    mk *= mr;
  }
}
// single-charge multipole expansion added into M (CMExpan, expansion.cpp:323-328)
static void add_mul_1(const orc_ctx *c, double ch, const double *pos, int p, bool bKappa,
                      double scale, double *M, double *Y, double *bes) {
  double rho, th, ph;
  cart_to_sph(pos, rho, th, ph);
  sh_eval(c, p, th, ph, Y);
  double val = bKappa ? rho * c->kappa : 0.0;
  bessel_i(p + 1, val, bes);
  double mscale = 1.0 / scale;
  double mr = rho * mscale;
  double mk = ch;
  const double *it = Y;
  for (int n = 0; n < p; n++) {
    double f = bes[n] * mk;
    for (int m = 0; m < 2 * n + 1; m++) M[n * n + m] += *(it++) * f;
    mk *= mr;
  }
}
// CLocalExpan(ch,pos,...) expansion.cpp:227-264 ; CMulExpan expansion.cpp:203-220
static void local_expan(const orc_ctx *c, int n, const double *q, const double *xyz, int p,
                        bool bKappa, double scale, double *out) {
  for (int i = 0; i < p * p; i++) out[i] = 0.0;
  vec Y(p * p), bes(p + 1), tmp(p * p);
  for (int i = 0; i < n; i++) {
    if (q[i] == 0.0) continue;
    for (int k = 0; k < p * p; k++) tmp[k] = 0.0;
    add_local_1(c, q[i], xyz + 3 * i, p, bKappa, scale, tmp.data(), Y.data(), bes.data());
    for (int k = 0; k < p * p; k++) out[k] += tmp[k];
  }
}
static void mul_expan(const orc_ctx *c, int n, const double *q, const double *xyz, int p,
                      bool bKappa, double scale, double *out) {
  for (int i = 0; i < p * p; i++) out[i] = 0.0;
  vec Y(p * p), bes(p + 1), tmp(p * p);
  for (int i = 0; i < n; i++) {
    if (q[i] == 0.0) continue;
    for (int k = 0; k < p * p; k++) tmp[k] = 0.0;
    add_mul_1(c, q[i], xyz + 3 * i, p, bKappa, scale, tmp.data(), Y.data(), bes.data());
    for (int k = 0; k < p * p; k++) out[k] += tmp[k];
  }
}
extern "C" void orc_local_expan(orc_ctx *c, int n, const double *q, const double *xyz, int p,
                                int bkappa, double scale, double *out) {
  local_expan(c, n, q, xyz, p, bkappa != 0, scale, out);
}
extern "C" void orc_mul_expan(orc_ctx *c, int n, const double *q, const double *xyz, int p,
                              int bkappa, double scale, double *out) {
  mul_expan(c, n, q, xyz, p, bkappa != 0, scale, out);
}

// CGradExpan(g,pos,p,bKappa,scale,false) triexpan.cpp:127-153 (local variant); the
// reference skips charges with |g|^2 < 1e-15 (triexpan.cpp:136,146).  The threshold is a
// setting of the oracle so that tests can measure what the skip leaves out of the forces.
static double g_grad_skip = 1e-15;
extern "C" void orc_set_grad_skip(double thr) { g_grad_skip = thr; }
static void grad_local(const orc_ctx *c, int n, const double *g, const double *xyz, int p,
                       bool bKappa, double scale, double *out3) {
  int len = p * p;
  for (int i = 0; i < 3 * len; i++) out3[i] = 0.0;
  vec Y(len), bes(p + 1), L(len);
  for (int i = 0; i < n; i++) {
    const double *gi = g + 3 * i;
    if (dot3(gi, gi) < g_grad_skip || dot3(gi, gi) == 0.0) continue;
    for (int k = 0; k < len; k++) L[k] = 0.0;
    add_local_1(c, 1.0, xyz + 3 * i, p, bKappa, scale, L.data(), Y.data(), bes.data());
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < len; k++) out3[d * len + k] += gi[d] * L[k];
  }
}
extern "C" void orc_grad_local(orc_ctx *c, int npts, const double *g3, const double *xyz, int p,
                               int bkappa, double scale, double *out3) {
  grad_local(c, npts, g3, xyz, p, bkappa != 0, scale, out3);
}

// CXFormBase::computeOrder xforms.cpp:402-414, with N_POLES replaced by the run's p
static int g_dbg_full_order = 0;  // diagnostics: 1 = disable adaptive truncation
extern "C" void orc_debug_full_order(int on) { g_dbg_full_order = on; }
static int compute_order(double rho, double tq, double scale, int p) {
  if (g_dbg_full_order) return p;
  if (tq < 1e-10) return 1;
  double cc = rho / scale - 1;
  if (cc <= 1) return p;
  int q = int(log(tq / (MAX_ERROR * scale * (cc - 1))) / log(cc) - 1 + 0.5);
  if (q < 1) q = 1;
  if (q > p) q = p;
  return q;
}
extern "C" int orc_compute_order(double rho, double tq, double scale, int p) {
  return compute_order(rho, tq, scale, p);
}

// ----------------------------------------------------------------------------
// rotation coefficients (rotcoeff.cpp)
// ----------------------------------------------------------------------------
struct Rot {
  int p = 0, nmax = 0;
  bool sing = false, grad = false;
  double theta, phi, xi, sint, cost, cott;
  cplx exphi, exiphi, exxi;
  std::vector<cplx> R, dR;
  std::vector<int> off;
  cplx &ROT(int n, int s, int m) { return R[off[n] + (s + n) * (n + 1) + m]; }
  const cplx &ROT(int n, int s, int m) const { return R[off[n] + (s + n) * (n + 1) + m]; }
  cplx &dROT(int n, int s, int m) { return dR[off[n] + (s + n) * (n + 1) + m]; }
  const cplx &dROT(int n, int s, int m) const { return dR[off[n] + (s + n) * (n + 1) + m]; }
};
static cplx sh_comp(const double *SH, int n, int m) {  // expansion.h:521-536
  if (m == 0) return cplx(SH[n * n], 0.0);
  int i = n * n + 2 * abs(m) - 1;
  if (m > 0) return cplx(SH[i], SH[i + 1]);
  return cplx(SH[i], -SH[i + 1]);
}
// CRotCoeff::initParams rotcoeff.cpp:209-244
static void rot_params(Rot &r, double theta, double phi, double xi) {
  r.theta = theta; r.phi = phi; r.xi = xi;
  r.sint = sin(r.theta);
  if (r.sint < EPS_SIN_THETA) {
    r.sing = true;
    if (r.theta < M_PI / 2) {
      r.theta = 0.0; r.phi = 0.0; r.sint = 0.0; r.cost = 1.0;
    } else {
      r.theta = M_PI; r.phi = 0.0; r.sint = 0.0; r.cost = -1.0;
    }
    r.exphi = cplx(1.0, 0.0);
  } else {
    r.sing = false;
    r.cost = cos(r.theta);
    r.cott = r.cost / r.sint;
    r.exphi = cplx(cos(phi), sin(phi));
  }
  r.exiphi = std::conj(r.exphi);
  r.exxi = cplx(cos(xi), sin(xi));
}
// CRotCoeff::reset rotcoeff.cpp:155-171 with computeCoeff :780-810 and computeGradCoeff :818-850
static void rot_reset(const orc_ctx *c, Rot &r, double theta, double phi, double xi, int p, bool grad) {
  rot_params(r, theta, phi, xi);
  r.p = p;
  r.grad = grad;
  int nmax = 2 * p - 1;
  r.nmax = nmax;
  r.off.assign(nmax + 1, 0);
  for (int n = 1; n <= nmax; n++) r.off[n] = r.off[n - 1] + (2 * (n - 1) + 1) * n;
  r.R.assign(r.off[nmax], cplx(0, 0));
  if (grad) r.dR.assign(r.off[nmax], cplx(0, 0));
  if (r.sing) return;
  vec SH(nmax * nmax);
  sh_eval(c, nmax, r.theta, r.phi, SH.data());
  r.ROT(0, 0, 0) = sh_comp(SH.data(), 0, 0);
  for (int n = 1; n < nmax; n++) {
    r.ROT(n, 0, 0) = sh_comp(SH.data(), n, 0);
    for (int s = 1; s <= n; s++) {
      r.ROT(n, s, 0) = sh_comp(SH.data(), n, -s);
      r.ROT(n, -s, 0) = sh_comp(SH.data(), n, s);
    }
  }
  double r1 = -0.5 * (1 + r.cost), r2 = 0.5 * (1 - r.cost), r3 = -r.sint;
  for (int m = 0; m < p - 1; m++)
    for (int n = m + 2; n < 2 * p - m - 1; n++)
      for (int s = -n + 1; s < n; s++) {
        cplx d1 = (r1 * c->ETA(n, s - 1)) * r.ROT(n, s - 1, m);
        cplx d2 = (r2 * c->ETA(n, -s - 1)) * r.ROT(n, s + 1, m);
        cplx d3 = (r3 * c->ZETA(n - 1, abs(s))) * r.ROT(n, s, m);
        r.ROT(n - 1, s, m + 1) = r.exxi * (d1 * r.exiphi + d2 * r.exphi + d3) / c->ETA(n, m);
      }
  if (!grad) return;
  r.dROT(0, 0, 0) = cplx(0.0, 0.0);
  for (int n = 1; n < nmax; n++) {
    r.dROT(n, 0, 0) = -sqrt((double)n * (n + 1)) * r.exiphi * sh_comp(SH.data(), n, 1);
    for (int s = 1; s < n; s++) {
      r.dROT(n, -s, 0) = (s * r.cott) * sh_comp(SH.data(), n, s) -
                         sqrt((double)(n - s) * (n + s + 1)) * r.exiphi * sh_comp(SH.data(), n, s + 1);
      r.dROT(n, s, 0) = std::conj(r.dROT(n, -s, 0));
    }
    r.dROT(n, -n, 0) = (n * r.cott) * sh_comp(SH.data(), n, n);
    r.dROT(n, n, 0) = std::conj(r.dROT(n, -n, 0));
  }
  double dr1 = 0.5 * r.sint, dr2 = dr1, dr3 = -r.cost;
  for (int m = 0; m < p - 1; m++)
    for (int n = m + 2; n < 2 * p - m - 1; n++)
      for (int s = -n + 1; s < n; s++) {
        cplx d1 = c->ETA(n, s - 1) * (r1 * r.dROT(n, s - 1, m) + dr1 * r.ROT(n, s - 1, m));
        cplx d2 = c->ETA(n, -s - 1) * (r2 * r.dROT(n, s + 1, m) + dr2 * r.ROT(n, s + 1, m));
        cplx d3 = c->ZETA(n - 1, abs(s)) * (r3 * r.dROT(n, s, m) + dr3 * r.ROT(n, s, m));
        r.dROT(n - 1, s, m + 1) = r.exxi * (d1 * r.exiphi + d2 * r.exphi + d3) / c->ETA(n, m);
      }
}
// CRotCoeff::reset(CQuat) rotcoeff.cpp:181-201
static void rot_reset_quat(const orc_ctx *c, Rot &r, const double *q, int p) {
  double w = q[0], x = q[1], y = q[2], z = q[3];
  double theta = acos(1.0 - 2 * x * x - 2 * y * y);
  double phi, xi;
  if (fabs(theta) > 1e-12) {
    phi = atan2(y * z + w * x, x * z - w * y);
    xi = atan2(y * z - w * x, x * z + w * y);
  } else {
    theta = 0.0;
    phi = 0.0;
    xi = 2 * acos(w);
  }
  rot_reset(c, r, theta, phi, xi, p, false);
}
// CRotCoeff::rotate (RSR transform) rotcoeff.cpp:357-441
static void rot_apply(const Rot &r, const double *Min, double *Mout, int p2, bool bFor) {
  for (int i = 0; i < p2 * p2; i++) Mout[i] = 0.0;
  if (r.sing) {
    if (r.cost == -1.0) {
      for (int n = 0; n < p2; n++) {
        Mout[n * n] = (n % 2 == 0 ? Min[n * n] : -Min[n * n]);
        int s = (n % 2 == 0 ? 1 : -1);
        for (int m = 1; m < 2 * n + 1; m++, s = -s) Mout[n * n + m] = s * Min[n * n + m];
      }
    } else {
      for (int n = 0; n < p2; n++)
        for (int m = 0; m < 2 * n + 1; m++) Mout[n * n + m] = Min[n * n + m];
    }
    return;
  }
  if (bFor) {
    for (int n = 0; n < p2; n++)
      for (int m = 0; m <= n; m++) {
        cplx inc(Min[n * n], 0.0);
        cplx out = r.ROT(n, 0, m) * inc;
        for (int s = 1; s <= n; s++) {
          inc = cplx(Min[n * n + 2 * s - 1], Min[n * n + 2 * s]);
          out += (r.ROT(n, s, m) * inc + r.ROT(n, -s, m) * std::conj(inc));
        }
        if (m == 0)
          Mout[n * n] = out.real();
        else {
          Mout[n * n + 2 * m - 1] = out.real();
          Mout[n * n + 2 * m] = out.imag();
        }
      }
  } else {
    for (int n = 0; n < p2; n++)
      for (int s = 0; s <= n; s++) {
        cplx inc(Min[n * n], 0.0);
        cplx out = 1.0 * r.ROT(n, -s, 0) * inc;
        for (int m = 1; m <= n; m++) {
          inc = cplx(Min[n * n + 2 * m - 1], Min[n * n + 2 * m]);
          out += std::conj(r.ROT(n, s, m)) * inc + r.ROT(n, -s, m) * std::conj(inc);
        }
        if (s == 0)
          Mout[n * n] = out.real() * 1.0;
        else {
          Mout[n * n + 2 * s - 1] = out.real();
          Mout[n * n + 2 * s] = out.imag();
        }
      }
  }
}
// CRotCoeff::rotateWithXi rotcoeff.cpp:250-350
static void rot_apply_xi(const Rot &r, const double *Min, double *Mout, int p2, bool bFor) {
  if (!r.sing) {
    rot_apply(r, Min, Mout, p2, bFor);
    return;
  }
  for (int i = 0; i < p2 * p2; i++) Mout[i] = 0.0;
  double xi = (bFor ? r.xi : -r.xi);
  if (r.cost == -1.0) {
    for (int n = 0; n < p2; n++) {
      Mout[n * n] = (n % 2 == 0 ? Min[n * n] : -Min[n * n]);
      int s = (n % 2 == 0 ? 1 : -1);
      for (int m = 1; m <= n; m++) {
        cplx inc(Min[n * n + 2 * m - 1], Min[n * n + 2 * m]);
        cplx ex(cos(m * xi), sin(m * xi));
        cplx out = ex * inc;
        Mout[n * n + 2 * m - 1] = s * out.real();
        Mout[n * n + 2 * m] = -s * out.imag();
      }
    }
  } else {
    for (int n = 0; n < p2; n++) {
      Mout[n * n] = Min[n * n];
      for (int m = 1; m <= n; m++) {
        cplx inc(Min[n * n + 2 * m - 1], Min[n * n + 2 * m]);
        cplx ex(cos(m * xi), sin(m * xi));
        cplx out = ex * inc;
        Mout[n * n + 2 * m - 1] = out.real();
        Mout[n * n + 2 * m] = out.imag();
      }
    }
  }
}
static inline cplx derv(const cplx &c, int s) { return cplx(-s * c.imag(), s * c.real()); }
// CRotCoeff::dRotateTSing rotcoeff.cpp:617-691
static void rot_dT_sing(const orc_ctx *c, const Rot &r, const double *Min, double *Mout, int p2) {
  Mout[0] = 0.0;
  if (r.cost == 1.0) {
    for (int n = 1; n < p2; n++) {
      Mout[n * n] = 2.0 * c->QC(n, 0, 1) * Min[n * n + 1] * 1.0;
      for (int m = 1; m < n; m++) {
        cplx i0, out;
        if (m == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
        else i0 = cplx(Min[n * n + 2 * (m - 1) - 1], Min[n * n + 2 * (m - 1)]);
        cplx i1(Min[n * n + 2 * (m + 1) - 1], Min[n * n + 2 * (m + 1)]);
        out = c->QC(n, m, 0) * i0 + c->QC(n, m, 1) * i1;
        Mout[n * n + 2 * m - 1] = out.real();
        Mout[n * n + 2 * m] = out.imag();
      }
      cplx i0;
      if (n == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
      else i0 = cplx(Min[n * n + 2 * (n - 1) - 1], Min[n * n + 2 * (n - 1)]);
      cplx out = c->QC(n, n, 0) * i0;
      Mout[n * n + 2 * n - 1] = out.real();
      Mout[n * n + 2 * n] = out.imag();
    }
  } else {
    double s = -1.0;
    for (int n = 1; n < p2; n++, s = -s) {
      Mout[n * n] = 2.0 * s * c->QC(n, 0, 1) * Min[n * n + 1] * 1.0;
      for (int m = 1; m < n; m++) {
        cplx i0, out;
        if (m == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
        else i0 = cplx(Min[n * n + 2 * (m - 1) - 1], Min[n * n + 2 * (m - 1)]);
        cplx i1(Min[n * n + 2 * (m + 1) - 1], Min[n * n + 2 * (m + 1)]);
        out = s * (c->QC(n, m, 0) * std::conj(i0) + c->QC(n, m, 1) * std::conj(i1));
        Mout[n * n + 2 * m - 1] = out.real();
        Mout[n * n + 2 * m] = out.imag();
      }
      cplx i0;
      if (n == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
      else i0 = cplx(Min[n * n + 2 * (n - 1) - 1], Min[n * n + 2 * (n - 1)]);
      cplx out = s * c->QC(n, n, 0) * std::conj(i0);
      Mout[n * n + 2 * n - 1] = out.real();
      Mout[n * n + 2 * n] = out.imag();
    }
  }
}
// CRotCoeff::dRotatePSing rotcoeff.cpp:704-771
static void rot_dP_sing(const orc_ctx *c, const Rot &r, const double *Min, double *Mout, int p2) {
  Mout[0] = 0.0;
  if (r.cost == 1.0) {
    for (int n = 1; n < p2; n++) {
      Mout[n * n] = 2.0 * c->QC(n, 0, 1) * Min[n * n + 2] * 1.0;
      for (int m = 1; m < n; m++) {
        cplx i0, out;
        if (m == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
        else i0 = cplx(Min[n * n + 2 * (m - 1) - 1], Min[n * n + 2 * (m - 1)]);
        cplx i1(Min[n * n + 2 * (m + 1) - 1], Min[n * n + 2 * (m + 1)]);
        out = cplx(0.0, c->QC(n, m, 0)) * i0 - cplx(0.0, c->QC(n, m, 1)) * i1;
        Mout[n * n + 2 * m - 1] = out.real();
        Mout[n * n + 2 * m] = out.imag();
      }
      cplx i0;
      if (n == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
      else i0 = cplx(Min[n * n + 2 * (n - 1) - 1], Min[n * n + 2 * (n - 1)]);
      cplx out = cplx(0.0, c->QC(n, n, 0)) * i0;
      Mout[n * n + 2 * n - 1] = out.real();
      Mout[n * n + 2 * n] = out.imag();
    }
  } else {
    double s = 1.0;
    for (int n = 1; n < p2; n++, s = -s) {
      Mout[n * n] = s * 2 * c->QC(n, 0, 1) * Min[n * n + 2] * 1.0;
      for (int m = 1; m < n; m++) {
        cplx i0, out;
        if (m == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
        else i0 = cplx(Min[n * n + 2 * (m - 1) - 1], Min[n * n + 2 * (m - 1)]);
        cplx i1(Min[n * n + 2 * (m + 1) - 1], Min[n * n + 2 * (m + 1)]);
        out = s * (cplx(0.0, c->QC(n, m, 1)) * std::conj(i1) - cplx(0.0, c->QC(n, m, 0)) * std::conj(i0));
        Mout[n * n + 2 * m - 1] = out.real();
        Mout[n * n + 2 * m] = out.imag();
      }
      cplx i0;
      if (n == 1) i0 = cplx(Min[n * n], 0.0) * 1.0;
      else i0 = cplx(Min[n * n + 2 * (n - 1) - 1], Min[n * n + 2 * (n - 1)]);
      cplx out = cplx(0.0, -s * c->QC(n, n, 0)) * std::conj(i0);
      Mout[n * n + 2 * n - 1] = out.real();
      Mout[n * n + 2 * n] = out.imag();
    }
  }
}
// CRotCoeff::dRotateT rotcoeff.cpp:450-522
static void rot_dT(const orc_ctx *c, const Rot &r, const double *Min, double *Mout, int p2, bool bFor) {
  for (int i = 0; i < p2 * p2; i++) Mout[i] = 0.0;
  if (r.sing) {
    rot_dT_sing(c, r, Min, Mout, p2);
    if (!bFor)
      for (int i = 0; i < p2 * p2; i++) Mout[i] = -Mout[i];
    return;
  }
  if (bFor) {
    for (int n = 0; n < p2; n++)
      for (int m = 0; m <= n; m++) {
        cplx inc(Min[n * n], 0.0);
        cplx out = 1.0 * r.dROT(n, 0, m) * inc;
        for (int s = 1; s <= n; s++) {
          inc = cplx(Min[n * n + 2 * s - 1], Min[n * n + 2 * s]);
          out += (r.dROT(n, s, m) * inc + r.dROT(n, -s, m) * std::conj(inc));
        }
        if (m == 0) Mout[n * n] = out.real() * 1.0;
        else { Mout[n * n + 2 * m - 1] = out.real(); Mout[n * n + 2 * m] = out.imag(); }
      }
  } else {
    for (int n = 0; n < p2; n++)
      for (int s = 0; s <= n; s++) {
        cplx inc(Min[n * n], 0.0);
        cplx out = 1.0 * r.dROT(n, -s, 0) * inc;
        for (int m = 1; m <= n; m++) {
          inc = cplx(Min[n * n + 2 * m - 1], Min[n * n + 2 * m]);
          out += std::conj(r.dROT(n, s, m)) * inc + r.dROT(n, -s, m) * std::conj(inc);
        }
        if (s == 0) Mout[n * n] = out.real() * 1.0;
        else { Mout[n * n + 2 * s - 1] = out.real(); Mout[n * n + 2 * s] = out.imag(); }
      }
  }
}
// CRotCoeff::dRotateP rotcoeff.cpp:531-606
static void rot_dP(const orc_ctx *c, const Rot &r, const double *Min, double *Mout, int p2, bool bFor) {
  for (int i = 0; i < p2 * p2; i++) Mout[i] = 0.0;
  if (r.sing) {
    rot_dP_sing(c, r, Min, Mout, p2);
    if (!bFor && r.cost == 1.0)
      for (int i = 0; i < p2 * p2; i++) Mout[i] = -Mout[i];
    return;
  }
  if (bFor) {
    for (int n = 0; n < p2; n++)
      for (int m = 0; m <= n; m++) {
        cplx inc;
        cplx out(0.0, 0.0);
        for (int s = 1; s <= n; s++) {
          inc = cplx(Min[n * n + 2 * s - 1], Min[n * n + 2 * s]);
          out -= (r.ROT(n, s, m) * derv(inc, s) + r.ROT(n, -s, m) * derv(std::conj(inc), -s));
        }
        if (m == 0) Mout[n * n] = out.real() * 1.0;
        else { Mout[n * n + 2 * m - 1] = out.real(); Mout[n * n + 2 * m] = out.imag(); }
      }
  } else {
    for (int n = 0; n < p2; n++)
      for (int s = 0; s <= n; s++) {
        cplx inc(Min[n * n], 0);
        cplx out = 1.0 * r.ROT(n, -s, 0) * derv(inc, s);
        for (int m = 1; m <= n; m++) {
          inc = cplx(Min[n * n + 2 * m - 1], Min[n * n + 2 * m]);
          out += std::conj(r.ROT(n, s, m)) * derv(inc, s) + r.ROT(n, -s, m) * derv(std::conj(inc), s);
        }
        if (s == 0) Mout[n * n] = out.real() * 1.0;
        else { Mout[n * n + 2 * s - 1] = out.real(); Mout[n * n + 2 * s] = out.imag(); }
      }
  }
}

extern "C" int orc_rot_coeff(orc_ctx *c, double theta, double phi, double xi, int p, int nmax,
                             double *out, double *dout) {
  Rot r;
  rot_reset(c, r, theta, phi, xi, p, dout != nullptr);
  if (nmax > 2 * p - 1) nmax = 2 * p - 1;
  for (int n = 0; n < nmax; n++)
    for (int s = -n; s <= n; s++)
      for (int m = 0; m <= n; m++) {
        int idx = r.off[n] + (s + n) * (n + 1) + m;
        out[2 * idx] = r.R[idx].real();
        out[2 * idx + 1] = r.R[idx].imag();
        if (dout) {
          dout[2 * idx] = r.dR[idx].real();
          dout[2 * idx + 1] = r.dR[idx].imag();
        }
      }
  return r.sing ? 1 : 0;
}
extern "C" void orc_rotate(orc_ctx *c, double theta, double phi, double xi, int p, int p2, int fwd,
                           const double *in, double *out) {
  Rot r;
  rot_reset(c, r, theta, phi, xi, p, false);
  rot_apply(r, in, out, p2, fwd != 0);
}
extern "C" void orc_rotate_quat(orc_ctx *c, const double *q, int p, int fwd, const double *in,
                                double *out) {
  Rot r;
  rot_reset_quat(c, r, q, p);
  rot_apply_xi(r, in, out, p, fwd != 0);
}
extern "C" void orc_rotate_grad(orc_ctx *c, double theta, double phi, double xi, int p, int which,
                                int fwd, const double *in, double *out) {
  Rot r;
  rot_reset(c, r, theta, phi, xi, p, true);
  if (which == 0) rot_dT(c, r, in, out, p, fwd != 0);
  else rot_dP(c, r, in, out, p, fwd != 0);
}

// ----------------------------------------------------------------------------
// translation coefficients (transcoeff.cpp)
// ----------------------------------------------------------------------------
struct Trans {
  int p = 0;
  bool grad = false;
  double scalei, scaleo;
  std::vector<double> T, dT, fact;
  std::vector<int> off;
  double &U(std::vector<double> &A, int l, int n, int m) { return A[off[l] + n * (n + 1) / 2 + m]; }
  double TR(int l, int n, int m) const { return T[off[l] + n * (n + 1) / 2 + m]; }
  double dTR(int l, int n, int m) const { return dT[off[l] + n * (n + 1) / 2 + m]; }
};
// computeCoeff_ transcoeff.cpp:403-449
static void trans_rec(const orc_ctx *c, Trans &t, std::vector<double> &A) {
  int p = t.p;
  double kappasqr = c->kappa * c->kappa;
  double kpio = kappasqr * t.scalei * t.scaleo;
  double kpoo = kappasqr * t.scaleo * t.scaleo;
  double o_i = t.scaleo / t.scalei;
  for (int l = 1; l < 2 * p - 2; l++)
    t.U(A, l, 1, 0) = -(kpio * c->BETA(l - 1, 0) * t.U(A, l - 1, 0, 0) +
                        o_i * c->ALPHA(l, 0) * t.U(A, l + 1, 0, 0)) / c->ALPHA(0, 0);
  for (int n = 1; n < p - 1; n++)
    for (int l = n + 1; l < 2 * p - 2 - n; l++)
      t.U(A, l, n + 1, 0) = -(kpio * c->BETA(l - 1, 0) * t.U(A, l - 1, n, 0) +
                              kpoo * c->BETA(n - 1, 0) * t.U(A, l, n - 1, 0) +
                              o_i * c->ALPHA(l, 0) * t.U(A, l + 1, n, 0)) / c->ALPHA(n, 0);
  for (int m = 1; m < p; m++) {
    for (int l = m; l < 2 * p - 1 - m; l++)
      t.U(A, l, m, m) = -(kpio * c->DELTA(l, -m) * t.U(A, l - 1, m - 1, m - 1) +
                          o_i * c->GAMMA(l + 1, m - 1) * t.U(A, l + 1, m - 1, m - 1)) / c->GAMMA(m, -m);
    int n = m;
    for (int l = n + 1; l < 2 * p - 2 - n; l++)
      t.U(A, l, n + 1, m) = -(kpio * c->BETA(l - 1, m) * t.U(A, l - 1, n, m) +
                              o_i * c->ALPHA(l, m) * t.U(A, l + 1, n, m)) / c->ALPHA(n, m);
    for (int n = m + 1; n < p - 1; n++)
      for (int l = n + 1; l < 2 * p - 2 - n; l++)
        t.U(A, l, n + 1, m) = -(kpio * c->BETA(l - 1, m) * t.U(A, l - 1, n, m) +
                                kpoo * c->BETA(n - 1, m) * t.U(A, l, n - 1, m) +
                                o_i * c->ALPHA(l, m) * t.U(A, l + 1, n, m)) / c->ALPHA(n, m);
  }
}
// reset + initParams + initScale + computeCoeff transcoeff.cpp:151-203, 372-395
static void trans_reset(const orc_ctx *c, Trans &t, double rho, double kappa, int p, double scalei,
                        double scaleo, bool grad) {
  t.p = p;
  t.grad = grad;
  t.scalei = scalei;
  t.scaleo = scaleo;
  t.fact.assign(std::max(p, 1), 0.0);
  t.fact[0] = 1.0;
  for (int n = 1; n < p; n++) t.fact[n] = t.fact[n - 1] * (scaleo / scalei);
  int L = 2 * p;  // l < 2p-1 needed, allocate one spare level
  t.off.assign(L + 1, 0);
  for (int l = 1; l <= L; l++) t.off[l] = t.off[l - 1] + l * (l + 1) / 2;
  t.T.assign(t.off[L], 0.0);
  if (grad) t.dT.assign(t.off[L], 0.0);
  double d = rho, kd = kappa * d, ir = 0, exkid = 0;
  if (d > 0.0) {
    ir = 1.0 / d;
    exkid = exp(-kd) * ir;
  }
  vec K(2 * p);
  bessel_k(c, 2 * p, kd, K.data());
  t.U(t.T, 0, 0, 0) = exkid * K[0];
  exkid *= ir;
  if (grad) t.U(t.dT, 0, 0, 0) = -exkid * K[1];
  exkid *= scalei;
  for (int l = 1; l < 2 * p - 1; l++) {
    t.U(t.T, l, 0, 0) = exkid * K[l];
    exkid *= ir;
    if (grad) t.U(t.dT, l, 0, 0) = exkid * (l * K[l] - (2 * l + 1) * K[l + 1]);
    exkid *= scalei;
  }
  trans_rec(c, t, t.T);
  if (grad) trans_rec(c, t, t.dT);
}
static inline bool EVEN(int n) { return (n % 2) == 0; }
// translate / dTranslate (forward only) transcoeff.cpp:211-364
static void trans_apply(const Trans &t, const double *Min, double *Mout, int p, bool useD) {
  for (int i = 0; i < p * p; i++) Mout[i] = 0.0;
  auto TT = [&](int l, int n, int m) { return useD ? t.dTR(l, n, m) : t.TR(l, n, m); };
  for (int n = 0; n < p; n++) {
    int l;
    for (l = 0; l <= n; l++) {
      double T = (EVEN(n + l) ? TT(n, l, 0) : -TT(n, l, 0));
      Mout[n * n] += t.fact[n - l] * T * Min[l * l];
    }
    for (; l < p; l++) Mout[n * n] += TT(l, n, 0) * Min[l * l];
    for (int m = 1; m <= n; m++) {
      for (l = m; l <= n; l++) {
        double T = (EVEN(n + l) ? TT(n, l, m) : -TT(n, l, m));
        Mout[n * n + 2 * m - 1] += t.fact[n - l] * T * Min[l * l + 2 * m - 1];
        Mout[n * n + 2 * m] += t.fact[n - l] * T * Min[l * l + 2 * m];
      }
      for (; l < p; l++) {
        Mout[n * n + 2 * m - 1] += TT(l, n, m) * Min[l * l + 2 * m - 1];
        Mout[n * n + 2 * m] += TT(l, n, m) * Min[l * l + 2 * m];
      }
    }
  }
}
extern "C" void orc_trans_coeff(orc_ctx *c, double rho, double kappa_seed, int p, double scalei,
                                double scaleo, double *T, double *dT) {
  Trans t;
  trans_reset(c, t, rho, kappa_seed, p, scalei, scaleo, dT != nullptr);
  // out layout: [m][n][l] for m<=n<=l<p (only what the apply uses)
  int k = 0;
  for (int m = 0; m < p; m++)
    for (int n = m; n < p; n++)
      for (int l = n; l < p; l++) {
        T[k] = t.TR(l, n, m);
        if (dT) dT[k] = t.dTR(l, n, m);
        k++;
      }
}
extern "C" void orc_translate(orc_ctx *c, double rho, double kappa_seed, int p, double scalei,
                              double scaleo, int grad, const double *in, double *out) {
  Trans t;
  trans_reset(c, t, rho, kappa_seed, p, scalei, scaleo, grad != 0);
  trans_apply(t, in, out, p, grad != 0);
}

// ----------------------------------------------------------------------------
// re-expansion transforms (xforms.cpp)
// ----------------------------------------------------------------------------
// CXFormA::reset + xformH xforms.cpp:51-77, 123-132.  P = target - source.
static void xform_h(const orc_ctx *c, const double *P, double a_src, double a_tgt, int p, int pfull,
                    const double *in, double *out) {
  double rho, th, ph;
  cart_to_sph(P, rho, th, ph);
  Rot r;
  rot_reset(c, r, th, ph, M_PI, p, false);
  Trans t;
  trans_reset(c, t, rho, c->kappa, p, a_src, a_tgt, false);
  vec t1(p * p), t2(p * p), o(p * p);
  rot_apply(r, in, t1.data(), p, true);
  trans_apply(t, t1.data(), t2.data(), p, false);
  rot_apply(r, t2.data(), o.data(), p, false);
  for (int i = 0; i < pfull * pfull; i++) out[i] = (i < p * p) ? o[i] : 0.0;
}
// CXFormAIntra::reset + xformF + xformH xforms.cpp:214-227, 261-270
static void xform_intra(const orc_ctx *c, const double *P, double a_src, double a_tgt, int pH, int pF,
                        int pfull, const double *H, const double *F, double *outH, double *outF) {
  double rho, th, ph;
  cart_to_sph(P, rho, th, ph);
  int maxp = (pH > pF ? pH : pF);
  Rot r;
  rot_reset(c, r, th, ph, M_PI, maxp, false);
  Trans tH, tF;
  trans_reset(c, tH, rho, c->kappa, pH, a_src, a_tgt, false);
  trans_reset(c, tF, rho, 0.0, pF, a_src, a_tgt, false);
  vec t1(maxp * maxp), t2(maxp * maxp), o(maxp * maxp);
  for (int i = 0; i < pfull * pfull; i++) { outH[i] = 0.0; outF[i] = 0.0; }
  rot_apply(r, F, t1.data(), pF, true);
  trans_apply(tF, t1.data(), t2.data(), pF, false);
  rot_apply(r, t2.data(), o.data(), pF, false);
  for (int i = 0; i < pF * pF; i++) outF[i] = o[i];
  rot_apply(r, H, t1.data(), pH, true);
  trans_apply(tH, t1.data(), t2.data(), pH, false);
  rot_apply(r, t2.data(), o.data(), pH, false);
  for (int i = 0; i < pH * pH; i++) outH[i] = o[i];
}
// CXFormA::xformH(CExpan, CGradExpan) + sphToCart xforms.cpp:86-115, 165-192
static void xform_h_grad(const orc_ctx *c, const double *P, double a_src, double a_tgt, int p, int pfull,
                         const double *Min, double *G3) {
  double rho, th, ph;
  cart_to_sph(P, rho, th, ph);
  Rot r;
  rot_reset(c, r, th, ph, M_PI, p, true);
  Trans t;
  trans_reset(c, t, rho, c->kappa, p, a_src, a_tgt, true);
  double sint = r.sint, cost = r.cost, sinp = r.exphi.imag(), cosp = r.exphi.real();
  double ir = 1.0 / rho;
  double JR[3][3];
  if (r.sing) {
    double a[3][3] = {{0.0, ir * cost, 0.0}, {0.0, 0.0, ir}, {cost, 0.0, 0.0}};
    memcpy(JR, a, sizeof(a));
  } else {
    double irst = ir / sint, irct = ir * cost;
    double a[3][3] = {{sint * cosp, irct * cosp, -irst * sinp},
                      {sint * sinp, irct * sinp, irst * cosp},
                      {cost, -ir * sint, 0.0}};
    memcpy(JR, a, sizeof(a));
  }
  int len = p * p;
  vec tM1(len), tL(len), Gr(len), Gt(len), Gp(len), tmp(len);
  rot_apply(r, Min, tM1.data(), p, true);
  trans_apply(t, tM1.data(), tL.data(), p, true);
  rot_apply(r, tL.data(), Gr.data(), p, false);
  trans_apply(t, tM1.data(), tL.data(), p, false);
  rot_dT(c, r, tL.data(), Gt.data(), p, false);
  rot_dP(c, r, tL.data(), Gp.data(), p, false);
  rot_dT(c, r, Min, tM1.data(), p, true);
  trans_apply(t, tM1.data(), tL.data(), p, false);
  rot_apply(r, tL.data(), tmp.data(), p, false);
  for (int i = 0; i < len; i++) Gt[i] += tmp[i];
  rot_dP(c, r, Min, tM1.data(), p, true);
  trans_apply(t, tM1.data(), tL.data(), p, false);
  rot_apply(r, tL.data(), tmp.data(), p, false);
  for (int i = 0; i < len; i++) Gp[i] += tmp[i];
  int lf = pfull * pfull;
  for (int i = 0; i < 3 * lf; i++) G3[i] = 0.0;
  for (int i = 0; i < len; i++) {
    double ar[3] = {Gr[i], Gt[i], Gp[i]};
    for (int d = 0; d < 3; d++) G3[d * lf + i] = dot3(JR[d], ar);
  }
}
// CXFormNIntra::xformFH xforms.cpp:528-550
static void xform_num(const orc_ctx *c, const double *P, int npts, const double *xyz, const double *qF,
                      const double *qH, double tQF, double tQH, double a_src, double a_tgt, int p,
                      double *outF, double *outH) {
  double rho = sqrt(dot3(P, P));
  vec d(3 * npts);
  for (int i = 0; i < npts; i++)
    for (int k = 0; k < 3; k++) d[3 * i + k] = xyz[3 * i + k] - P[k];
  int pF = compute_order(rho, tQF, a_src, p);
  vec o(p * p);
  local_expan(c, npts, qF, d.data(), pF, false, a_tgt, o.data());
  for (int i = 0; i < p * p; i++) outF[i] = (i < pF * pF) ? o[i] : 0.0;
  int pH = compute_order(rho, tQH, a_src, p);
  local_expan(c, npts, qH, d.data(), pH, true, a_tgt, o.data());
  for (int i = 0; i < p * p; i++) outH[i] = (i < pH * pH) ? o[i] : 0.0;
}
// CXFormNIntra::xformGFH xforms.cpp:558-582
static void xform_num_grad(const orc_ctx *c, const double *P, int npts, const double *xyz,
                           const double *gF, const double *gH, double tGF, double tGH, double a_src,
                           double a_tgt, int p, double *outF3, double *outH3) {
  double rho = sqrt(dot3(P, P));
  vec d(3 * npts);
  for (int i = 0; i < npts; i++)
    for (int k = 0; k < 3; k++) d[3 * i + k] = xyz[3 * i + k] - P[k];
  int len = p * p;
  int pF = compute_order(rho, tGF, a_src, p);
  vec o(3 * p * p);
  grad_local(c, npts, gF, d.data(), pF, false, a_tgt, o.data());
  for (int dd = 0; dd < 3; dd++)
    for (int i = 0; i < len; i++) outF3[dd * len + i] = (i < pF * pF) ? o[dd * pF * pF + i] : 0.0;
  int pH = compute_order(rho, tGH, a_src, p);
  grad_local(c, npts, gH, d.data(), pH, true, a_tgt, o.data());
  for (int dd = 0; dd < 3; dd++)
    for (int i = 0; i < len; i++) outH3[dd * len + i] = (i < pH * pH) ? o[dd * pH * pH + i] : 0.0;
}

extern "C" void orc_xform_h(orc_ctx *c, const double *P, double a_src, double a_tgt, int p,
                            const double *in, double *out) {
  xform_h(c, P, a_src, a_tgt, p, p, in, out);
}
extern "C" void orc_xform_intra(orc_ctx *c, const double *P, double a_src, double a_tgt, int pH, int pF,
                                const double *H, const double *F, double *outH, double *outF) {
  int pf = pH > pF ? pH : pF;
  xform_intra(c, P, a_src, a_tgt, pH, pF, pf, H, F, outH, outF);
}
extern "C" void orc_xform_h_grad(orc_ctx *c, const double *P, double a_src, double a_tgt, int p,
                                 const double *in, double *out3) {
  xform_h_grad(c, P, a_src, a_tgt, p, p, in, out3);
}
extern "C" void orc_xform_num(orc_ctx *c, const double *P, int npts, const double *xyz, const double *qF,
                              const double *qH, double tQF, double tQH, double a_src, double a_tgt, int p,
                              double *outF, double *outH) {
  xform_num(c, P, npts, xyz, qF, qH, tQF, tQH, a_src, a_tgt, p, outF, outH);
}

// ----------------------------------------------------------------------------
// surface points, IMat (molecule.cpp, expcenter.cpp)
// ----------------------------------------------------------------------------
// CMolecule::spherePts molecule.cpp:2273-2294 and point scaling :973-975
extern "C" void orc_sphere_points(int N, double r, double *xyz) {
  vec th(N, 0.0), ph(N, 0.0);
  for (int i = 1; i <= N; i++) {
    double h = -1.0 + 2.0 * (i - 1.0) / (N - 1.0);
    th[i - 1] = acos(h);
    if (i == 1 || i == N)
      ph[i - 1] = 0;
    else
      ph[i - 1] = (ph[i - 2] + 3.6 / sqrt(N * (1.0 - h * h)));
    while (ph[i - 1] > 2 * M_PI) ph[i - 1] -= 2 * M_PI;
    while (ph[i - 1] < -2 * M_PI) ph[i - 1] += 2 * M_PI;
  }
  for (int i = 0; i < N; i++) {
    double x = 1 * sin(th[i]) * cos(ph[i]);
    double y = 1 * sin(th[i]) * sin(ph[i]);
    double z = 1 * cos(th[i]);
    xyz[3 * i] = r * x;
    xyz[3 * i + 1] = r * y;
    xyz[3 * i + 2] = r * z;
  }
}
// point count rule of getXFormSpherePoints molecule.cpp:957-966
extern "C" int orc_npts_rule(double r, int nin) {
  const int minimumPoints = 200;
  const double cutoff = 1.5, maxrad = 10;
  const double gradient = nin / ((maxrad - cutoff) * (maxrad - cutoff));
  int N;
  if (r < cutoff)
    N = minimumPoints;
  else {
    N = int(minimumPoints + gradient * (r - cutoff) * (r - cutoff));
    if (N > nin) N = nin;
  }
  return N;
}

// computeIntegralE/B + build_IMat expcenter.cpp:219-335, 667-902.
// Integral index kY runs over (l,s,n,m) with n<=l and (n<l or m<=s).
static void imat_build(const orc_ctx *c, int p, int npe, const double *spe, int npb, const double *spb,
                       double *IMat) {
  int PSQH = p * (p + 1) / 2;
  int PQUADH = PSQH * (PSQH + 1) / 2;
  vec Yrr(PQUADH, 0.0), Yri(PQUADH, 0.0), Yir(PQUADH, 0.0), Yii(PQUADH, 0.0);
  std::vector<int> id(p), idk(PSQH);
  for (int l = 0; l < p; l++) id[l] = l * (l + 1) / 2;
  idk[0] = 0;
  for (int l = 1; l < p; l++)
    for (int s = 0; s <= l; s++) {
      int k = id[l] + s;
      idk[k] = idk[k - 1] + k;
    }
  int np = npb + npe;
  bool useE = (npe <= npb);
  const double *pts = useE ? spe : spb;
  int npts = useE ? npe : npb;
  vec Y(p * p);
  for (int h = 0; h < npts; h++) {
    double rho, th, ph;
    cart_to_sph(pts + 3 * h, rho, th, ph);
    sh_eval(c, p, th, ph, Y.data());
    int kY = 0;
    for (int l = 0; l < p; l++)
      for (int s = 0; s <= l; s++) {
        double Ylsr, Ylsi;
        if (s == 0) { Ylsr = Y[l * l]; Ylsi = 0.0; }
        else { Ylsr = Y[l * l + 2 * s - 1]; Ylsi = Y[l * l + 2 * s]; }
        for (int n = 0; n <= l; n++)
          for (int m = 0; m <= n; m++) {
            if (n == l && m > s) break;
            double Ynmr, Ynmi;
            if (m == 0) { Ynmr = Y[n * n]; Ynmi = 0.0; }
            else { Ynmr = Y[n * n + 2 * m - 1]; Ynmi = Y[n * n + 2 * m]; }
            double w;
            if (m == 0 && s == 0 && (n + l) % 2 == 0) {
              if (h > 0 && h < np) w = (h % 2 == 0) ? 2.0 / 3.0 : 4.0 / 3.0;
              else w = 1.0 / 3.0;
            } else
              w = 1.0;
            Yrr[kY] += w * Ylsr * Ynmr;
            Yri[kY] += w * Ylsr * Ynmi;
            Yir[kY] += w * Ylsi * Ynmr;
            Yii[kY] += w * Ylsi * Ynmi;
            kY++;
          }
      }
  }
  double dA = 4 * M_PI / (double)(np - 1);
  for (int l = 0, kY = 0; l < p; l++) {
    double delta = 1.0 / ((2 * l + 1) / (4.0 * M_PI));
    for (int s = 0; s <= l; s++)
      for (int n = 0; n <= l; n++)
        for (int m = 0; m <= n; m++) {
          if (n == l && m > s) break;
          if (useE) {
            if (n == l && m == s) {
              if (m == 0) { Yrr[kY] *= dA; Yii[kY] = 0.0; }
              else { Yrr[kY] *= dA; Yii[kY] *= dA; }
              Yri[kY] *= dA;
              Yir[kY] *= dA;
            } else {
              Yrr[kY] *= dA; Yri[kY] *= dA; Yir[kY] *= dA; Yii[kY] *= dA;
            }
          } else {
            if (n == l && m == s) {
              if (m == 0) { Yrr[kY] = delta - dA * Yrr[kY]; Yii[kY] = 0.0; }
              else { Yrr[kY] = 0.5 * delta - dA * Yrr[kY]; Yii[kY] = 0.5 * delta - dA * Yii[kY]; }
              Yri[kY] *= -dA;
              Yir[kY] *= -dA;
            } else {
              Yrr[kY] *= -dA; Yri[kY] *= -dA; Yir[kY] *= -dA; Yii[kY] *= -dA;
            }
          }
          kY++;
        }
  }
  auto YY = [&](vec &A, int n, int m, int l, int s) { return A[idk[id[l] + s] + id[n] + m]; };
  int k = 0;
  const int sNeg = -1, mNeg = -1;
  for (int l = 0; l < p; l++) {
    double ytemp;
    for (int n = 0; n < p; n++) {
      double fact_n = (2 * n + 1) / (4.0 * M_PI);
      for (int m = 0; m <= n; m++) {
        bool bUpper = (n < l || (n == l && m <= 0));
        ytemp = (bUpper ? YY(Yrr, n, m, l, 0) : YY(Yrr, l, 0, n, m));
        IMat[k++] = fact_n * ytemp;
        if (m != 0) {
          ytemp = (bUpper ? YY(Yri, n, m, l, 0) : YY(Yir, l, 0, n, m));
          IMat[k++] = mNeg * -1.0 * fact_n * ytemp;
        }
      }
    }
    for (int s = 1; s <= l; s++) {
      for (int n = 0; n < p; n++) {
        double fact_n = (2 * n + 1) / (4.0 * M_PI);
        for (int m = 0; m <= n; m++) {
          bool bUpper = (n < l || (n == l && m <= s));
          ytemp = (bUpper ? YY(Yrr, n, m, l, s) : YY(Yrr, l, s, n, m));
          IMat[k++] = 2.0 * fact_n * ytemp;
          if (m != 0) {
            ytemp = (bUpper ? YY(Yri, n, m, l, s) : YY(Yir, l, s, n, m));
            IMat[k++] = mNeg * -2.0 * fact_n * ytemp;
          }
        }
      }
      for (int n = 0; n < p; n++) {
        double fact_n = (2 * n + 1) / (4.0 * M_PI);
        for (int m = 0; m <= n; m++) {
          bool bUpper = (n < l || (n == l && m <= s));
          ytemp = (bUpper ? YY(Yir, n, m, l, s) : YY(Yri, l, s, n, m));
          IMat[k++] = sNeg * -2.0 * fact_n * ytemp;
          if (m != 0) {
            ytemp = (bUpper ? YY(Yii, n, m, l, s) : YY(Yii, l, s, n, m));
            IMat[k++] = 2.0 * fact_n * ytemp;
          }
        }
      }
    }
  }
}
extern "C" void orc_imat_build(orc_ctx *c, int p, int npe, const double *spe, int npb, const double *spb,
                               double *imat) {
  imat_build(c, p, npe, spe, npb, spb, imat);
}

#include "pbsam_oracle_solver.inc"
#include "pbsam_oracle_bd.inc"
