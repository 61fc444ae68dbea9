// Device-side data layouts and small math helpers shared by the HIP kernels.
//
// Coefficient blocks follow the reference layout (src/expansion.cpp:57-63): for
// order n the 2n+1 reals [c(n,0), Re c(n,1), Im c(n,1), ...] start at n*n.
#pragma once
#include <hip/hip_runtime.h>

namespace pbsam {

constexpr int PMAX = 32;        // largest pole order supported
constexpr int NLT = 2 * PMAX;   // rows of the SH / rotation tables

// Constant tables (built on the host, see pbsam_host.cpp: build_tables)
struct Tabs {
  const double *c1, *c2, *c3;   // [NLT][NLT]  Legendre recursion and sqrt((n-m)!/(n+m)!)
  const double *c6;             // [NLT]       (2m-1)!!
  const double *c4;             // [NLT]       1/((2n-1)(2n-3))  (k-hat recursion)
  const double *zeta;           // [NLT][NLT]
  const double *eta;            // [NLT][2*NLT] index m+NLT
  const double *alpha, *beta;   // [NLT][PMAX]
  const double *gam, *del;      // [NLT][2*PMAX] index m+PMAX
  const double *Q;              // [2*PMAX-1][PMAX][2]  singular dR
  double kappa;
};

__device__ __forceinline__ double C1(const Tabs &t, int n, int m) { return t.c1[n * NLT + m]; }
__device__ __forceinline__ double C2(const Tabs &t, int n, int m) { return t.c2[n * NLT + m]; }
__device__ __forceinline__ double C3(const Tabs &t, int n, int m) { return t.c3[n * NLT + m]; }
__device__ __forceinline__ double ZETA(const Tabs &t, int n, int m) { return t.zeta[n * NLT + m]; }
__device__ __forceinline__ double ETA(const Tabs &t, int n, int m) { return t.eta[n * 2 * NLT + NLT + m]; }
__device__ __forceinline__ double ALPHA(const Tabs &t, int n, int m) { return t.alpha[n * PMAX + m]; }
__device__ __forceinline__ double BETA(const Tabs &t, int n, int m) { return t.beta[n * PMAX + m]; }
__device__ __forceinline__ double GAMMA(const Tabs &t, int n, int m) { return t.gam[n * 2 * PMAX + PMAX + m]; }
__device__ __forceinline__ double DELTA(const Tabs &t, int n, int m) { return t.del[n * 2 * PMAX + PMAX + m]; }

struct cd {  // complex double with the reference's (std::complex) operation order
  double x, y;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd cscale(double s, cd a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ cd cconj(cd a) { return {a.x, -a.y}; }

// CartToSph (util.h:418-438)
__device__ __forceinline__ void cart_to_sph(double x, double y, double z, double &rho, double &th,
                                            double &ph) {
  rho = sqrt(x * x + y * y + z * z);
  if (rho < fabs(z)) rho = fabs(z);
  th = (rho == 0.0) ? 0.0 : acos(z / rho);
  ph = (x == 0.0 && y == 0.0) ? 0.0 : atan2(y, x);
}

// CXFormBase::computeOrder (xforms.cpp:402-414), clamped to the run's p
__device__ __forceinline__ int compute_order(double rho, double tq, double scale, int p) {
  if (tq < 1e-10) return 1;
  double cc = rho / scale - 1;
  if (cc <= 1) return p;
  int q = int(log(tq / (1e-6 * scale * (cc - 1))) / log(cc) - 1 + 0.5);
  if (q < 1) q = 1;
  if (q > p) q = p;
  return q;
}

// Associated Legendre column m (orders m..P-1) scaled to the reference's real SH:
// CSHExpan::legendre + computecoeff (expansion.cpp:483-518, 634-675).
// Calls emit(n, re, im) for n in [m, P); for m == 0, im is unused.
template <class F>
__device__ __forceinline__ void sh_column(const Tabs &t, int P, int m, double cost, double sint, double phi,
                                          F emit) {
  if (m >= P) return;
  double cm = 1.0, sm = 0.0;
  if (m > 0) {
    double ang = 0.0;
    for (int k = 0; k < m; k++) ang += phi;
    sincos(ang, &sm, &cm);
  }
  double pm2 = 0.0, pm1;
  if (m == 0) {
    pm1 = t.c6[0];
  } else {
    double sq = sint;
    for (int k = 1; k < m; k++) sq *= sint;
    double neg = (m & 1) ? -1.0 : 1.0;
    pm1 = neg * t.c6[m] * sq;
  }
  double s = (m & 1) ? -1.0 : 1.0;
  auto out = [&](int n, double pv) {
    if (m == 0)
      emit(n, pv * C3(t, n, 0), 0.0);
    else
      emit(n, pv * (s * C3(t, n, m) * cm), pv * (s * C3(t, n, m) * sm));
  };
  out(m, pm1);
  if (m + 1 < P) {
    double pv = cost * (2 * m + 1) * pm1;
    pm2 = pm1;
    pm1 = pv;
    out(m + 1, pv);
    for (int n = m + 2; n < P; n++) {
      double v = cost * C1(t, n, m) * pm1 - C2(t, n, m) * pm2;
      pm2 = pm1;
      pm1 = v;
      out(n, v);
    }
  }
}

// CLExpan::BESSEL (expansion.cpp:422-436): k-hat_n, computed sequentially.
__device__ __forceinline__ double bessel_k_next(const Tabs &t, int n, double km1, double km2, double valsqr) {
  return km1 + valsqr * km2 * t.c4[n];
}

// ---- job records (built on the host) ----
struct XJob {       // analytical re-expansion of one source sphere onto one target
  int src, tgt;     // sphere indices in the state arrays
  int slot;         // output slot
  int kind;         // 0: H only (CXFormA), 1: H and F (CXFormAIntra)
  const double *TH; // translation matrices (k_trans layout) for H (screened seeds)
  const double *TF; // translation matrices for F (unscreened seeds), kind 1 only
  double P[3];      // target centre - source centre
  double a_src, a_tgt;
  int tord;         // order the translation matrices were made at (0: this job's H order)
  int pad;
  const double *rot;  // stored rotation coefficients of the pair's direction (k_rotcoeff
                      // layout, xi = pi, rows n < p), or null: generated in the kernel
};
struct NJob {       // numerical P2L of a source sphere's surface charges onto a target
  int src, tgt, slot, pad;
  double P[3];
  double a_src, a_tgt;
};
struct TJob {       // translation matrices of one pair
  double P[3];
  double a_src, a_tgt;
  double kappa_seed;
  long long off;
  int grad;
  int src;          // source sphere whose total sets the order (used with a totals array)
};
struct RJob {       // rotation coefficients stored to HBM (molecule orientations)
  double theta, phi, xi;
  long long off;
  int sing;         // 0 regular, 1 theta=0, -1 theta=pi (set by the kernel)
  int pad;
};

// ---- algorithmic work per unit, counted from the reference's loop bounds ----
// Used only by the optional kernel instrumentation (flops and HBM bytes a launch must do
// by the algorithm, independent of how a kernel schedules them).
// rotation recursion (rotcoeff.cpp:398-441): levels m = 0..q-2, each step makes one complex
// coefficient from three (33 flops)
__host__ __device__ inline double work_rotgen(int q) {
  double s = 0;
  for (int m = 0; m + 1 < q; m++) s += (double)(2 * q - m - 2) * (2 * q - m - 2) - (double)(m + 1) * (m + 1);
  return 33.0 * s;
}
// rotation apply (rotcoeff.cpp:250-441): one complex MAC (8 flops) per (n, s, m>=0)
__host__ __device__ inline double work_rotapply(int q) {
  double s = 0;
  for (int n = 0; n < q; n++) s += (double)(n + 1) * (2 * n + 1);
  return 8.0 * s;
}
// coaxial translation apply (transcoeff.cpp:211-364): per level m a (q-m)^2 real-by-complex
// product (4 flops per term)
__host__ __device__ inline double work_transapply(int q) {
  double s = 0;
  for (int m = 0; m < q; m++) s += (double)(q - m) * (q - m);
  return 4.0 * s;
}
// one analytical re-expansion R^H T R of a pair with block orders pH and (intra) pF
__host__ __device__ inline double work_xform(int pH, int pF) {
  int q = pH > pF ? pH : pF;
  double w = work_rotgen(q) + 2 * work_rotapply(pH) + work_transapply(pH);
  if (pF > 0) w += 2 * work_rotapply(pF) + work_transapply(pF);
  return w;
}
// point-to-expansion term of one point at order q: harmonics (7 flops per (n, m>=0)),
// radial factors (3 per n) and the accumulation (2 per real coefficient)
__host__ __device__ inline double work_p2l_point(int q) {
  return 7.0 * q * (q + 1) / 2 + 3.0 * q + 2.0 * q * q;
}
// surface charges of one point from F and H at order p (expcenter.cpp:1420-1440):
// harmonics once, 3 flops per real coefficient for each block
__host__ __device__ inline double work_surface_point(int p) { return 7.0 * p * (p + 1) / 2 + 6.0 * p * p; }

// offset of translation matrix of level m (dense (p-m)^2) in a pair's block
__host__ __device__ __forceinline__ long long tmat_off(int p, int m) {
  // sum_{k<m} (p-k)^2
  const long long P = p, M = m;
  return M * P * P - P * M * (M - 1) + (M - 1) * M * (2 * M - 1) / 6;
}
__host__ __device__ __forceinline__ long long tmat_size(int p) { return tmat_off(p, p); }
// stored rotation coefficients: row n holds (n+1) x (2n+1) complex, index m*(2n+1)+(s+n)
__host__ __device__ __forceinline__ long long rot_off(int n) {
  long long o = 0;
  for (int k = 0; k < n; k++) o += (long long)(k + 1) * (2 * k + 1);
  return o;
}

}  // namespace pbsam
