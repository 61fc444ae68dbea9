// HIP kernels for the PB-SAM mutual-polarization path (gfx950, FP64).
//
// Kernel map (reference functions restated, paths relative to the reference src/):
//   k_trans        CTransCoeff::reset/computeCoeff/computeCoeff_   transcoeff.cpp:151-203,372-449
//   k_xform        CXFormA/CXFormAIntra::reset+xformH/xformF      xforms.cpp:51-77,123-132,214-270
//                  (CRotCoeff::reset/computeCoeff/rotate rotcoeff.cpp:155-171,357-441,780-810,
//                   CTransCoeff::translate transcoeff.cpp:211-286) -- fused per pair
//   k_p2l          CXFormNIntra::xformFH / CLocalExpan / CMulExpan xforms.cpp:528-550,
//                  expansion.cpp:203-340
//   k_rotcoeff     CRotCoeff::reset(CQuat) rotcoeff.cpp:181-201 (coefficients stored)
//   k_rotapply     CRotCoeff::rotateWithXi rotcoeff.cpp:250-350
//   k_sumslots     LS / LHS / LFS assembly molecule.cpp:1214-1297,1414-1473
//   k_xorder       CXFormBase::computeOrder per pair xforms.cpp:402-414 (order classes)
//   k_solve        CSolExpCenter::solveFH expcenter.cpp:1080-1370
//   k_surface      computeExposedSurfaceChargesFH expcenter.cpp:1420-1440
#include "pbsam_device.h"

namespace pbsam {

// ----------------------------------------------------------------------------
// translation matrices: one wavefront per pair, recursion over (l, n) per level m
// Output per pair: for each m, dense M_m[n-m][l-m] (n,l in [m,q)) holding the factor the
// forward translate multiplies Min(l,m) by:  l<=n: FACT(n-l)*(-1)^(n+l)*T(n,l,m), l>n: T(l,n,m).
// The order q is pmax, or with a totals array the pair's truncation order
// (CXFormBase::computeOrder) -- the values are those of the order-pmax recursion
// truncated, since row l, column n depends only on rows <= l+n.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_trans(Tabs tb, const TJob *jobs, const int *sel, const double *TQ,
                                              int pmax, double *Tout) {
  extern __shared__ double sh[];
  const TJob J = jobs[sel ? sel[blockIdx.x] : blockIdx.x];
  const int tid = threadIdx.x;
  double rho, th, ph;
  cart_to_sph(J.P[0], J.P[1], J.P[2], rho, th, ph);
  const int p = TQ ? compute_order(rho, TQ[J.src], J.a_src, pmax) : pmax;
  const int L = 2 * p - 1;  // recursion rows l in [0, 2p-2]
  double *U = sh;           // [(L+1)][p]   U[l*p+n] for the current level (columns n>=m)
  double *Ks = U + (L + 1) * p;   // [2p]
  double *fact = Ks + 2 * p;      // [p]
  const double scalei = J.a_src, scaleo = J.a_tgt;
  double *out = Tout + J.off;
  for (int pass = 0; pass < (J.grad ? 2 : 1); pass++) {
    // seeds T(l,0,0) (transcoeff.cpp:374-391): sequential product chain on lane 0
    if (tid == 0) {
      double d = rho, kd = J.kappa_seed * d, ir = 0, exkid = 0;
      if (d > 0.0) {
        ir = 1.0 / d;
        exkid = exp(-kd) * ir;
      }
      for (int i = 0; i < 2 * p; i++) Ks[i] = 1.0;
      if (kd != 0.0 && 2 * p > 1) {
        Ks[1] += kd;
        double vs = kd * kd;
        for (int n = 2; n < 2 * p; n++) Ks[n] = bessel_k_next(tb, n, Ks[n - 1], Ks[n - 2], vs);
      }
      if (pass == 0) U[0] = exkid * Ks[0];
      exkid *= ir;
      if (pass == 1) U[0] = -exkid * Ks[1];
      exkid *= scalei;
      for (int l = 1; l < 2 * p - 1; l++) {
        if (pass == 0) U[l * p] = exkid * Ks[l];
        exkid *= ir;
        if (pass == 1) U[l * p] = exkid * (l * Ks[l] - (2 * l + 1) * Ks[l + 1]);
        exkid *= scalei;
      }
      fact[0] = 1.0;
      for (int n = 1; n < p; n++) fact[n] = fact[n - 1] * (scaleo / scalei);
    }
    __syncthreads();
    const double kappasqr = tb.kappa * tb.kappa;
    const double kpio = kappasqr * scalei * scaleo;
    const double kpoo = kappasqr * scaleo * scaleo;
    const double o_i = scaleo / scalei;
    double *o = out + pass * tmat_size(p);
    for (int m = 0; m < p; m++) {
      if (m == 0) {
        for (int l = 1 + tid; l < 2 * p - 2; l += 64)
          U[l * p + 1] = -(kpio * BETA(tb, l - 1, 0) * U[(l - 1) * p + 0] +
                           o_i * ALPHA(tb, l, 0) * U[(l + 1) * p + 0]) / ALPHA(tb, 0, 0);
        __syncthreads();
        for (int n = 1; n < p - 1; n++) {
          for (int l = n + 1 + tid; l < 2 * p - 2 - n; l += 64)
            U[l * p + n + 1] = -(kpio * BETA(tb, l - 1, 0) * U[(l - 1) * p + n] +
                                 kpoo * BETA(tb, n - 1, 0) * U[l * p + n - 1] +
                                 o_i * ALPHA(tb, l, 0) * U[(l + 1) * p + n]) / ALPHA(tb, n, 0);
          __syncthreads();
        }
      } else {
        // diagonal U(l,m,m) from column m-1 of the previous level
        for (int l = m + tid; l < 2 * p - 1 - m; l += 64)
          U[l * p + m] = -(kpio * DELTA(tb, l, -m) * U[(l - 1) * p + m - 1] +
                           o_i * GAMMA(tb, l + 1, m - 1) * U[(l + 1) * p + m - 1]) / GAMMA(tb, m, -m);
        __syncthreads();
        {
          int n = m;
          for (int l = n + 1 + tid; l < 2 * p - 2 - n; l += 64)
            U[l * p + n + 1] = -(kpio * BETA(tb, l - 1, m) * U[(l - 1) * p + n] +
                                 o_i * ALPHA(tb, l, m) * U[(l + 1) * p + n]) / ALPHA(tb, n, m);
          __syncthreads();
        }
        for (int n = m + 1; n < p - 1; n++) {
          for (int l = n + 1 + tid; l < 2 * p - 2 - n; l += 64)
            U[l * p + n + 1] = -(kpio * BETA(tb, l - 1, m) * U[(l - 1) * p + n] +
                                 kpoo * BETA(tb, n - 1, m) * U[l * p + n - 1] +
                                 o_i * ALPHA(tb, l, m) * U[(l + 1) * p + n]) / ALPHA(tb, n, m);
          __syncthreads();
        }
      }
      // emit M_m: rows n, cols l in [m,p)
      const int w = p - m;
      double *Mm = o + tmat_off(p, m);
      for (int e = tid; e < w * w; e += 64) {
        int n = m + e / w, l = m + e % w;
        double v;
        if (l <= n) {
          double T = U[n * p + l];
          T = ((n + l) % 2 == 0) ? T : -T;
          v = fact[n - l] * T;
        } else {
          v = U[l * p + n];
        }
        Mm[e] = v;
      }
      __syncthreads();
    }
  }
}

// ----------------------------------------------------------------------------
// fused rotate -> translate -> rotate-back, one workgroup per pair.
// LDS: one rotation level (2q-1)^2 complex (advanced in place), the source blocks and the
// per-level vectors (the host's xform_lds).  NPL consecutive source blocks (planes: the x,y,z
// gradient blocks of one sphere) share one generation of the rotation coefficients.
// ----------------------------------------------------------------------------
constexpr int RT = 512;  // threads per workgroup of k_rotcoeff
constexpr int XG = 16;   // lanes per dot product in the forward rotate / translate

__device__ __forceinline__ double seg_reduce(double v) {
  for (int o = XG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, XG);
  return v;
}

struct XArgs {
  const XJob *jobs;
  const double *H;     // source H blocks (Hrot for inter pairs)
  const double *F;     // source F blocks (launches of kind-1 jobs)
  const double *TQH;   // source totals driving the truncation order
  const double *TQF;
  double *slotH, *slotF;
  const int *ordH, *ordF;  // explicit orders per batch position (null: adaptive from the totals)
  const int *sel;          // job index of each batch position (null: the position itself)
  const int *cls;          // batch position of each workgroup (null: the workgroup index)
  const int *tqmap;        // index into TQH/TQF of each source block (null: the block index)
  int p;
  int full;            // 1: use order p for both blocks (no adaptive truncation)
  unsigned long long *work;  // instrumentation: [flops, bytes] (null: off)
  int planes;          // consecutive source blocks / output slots per job (0 or 1: one)
  int split;           // 1: one plane per workgroup, plane = blockIdx.y (single-plane kernel)
  const int *cnt;      // with cls: device count of the class (workgroups past it leave)
  int nosplit;         // host hint: keep the planes of a job in one workgroup (jobs with
                       // stored rotations read their coefficients once for all planes)
  int allrot;          // host: every job of the launch carries stored rotations (k_xapply)
};

// Orders of a batch of re-expansion jobs (the same rule as k_xform), per batch position,
// and the split of the positions into classes by q = max(pH, pF): class c holds
// q <= XCLASS[c].  cnt[c] is advanced with atomics, so the order inside a class varies;
// each job writes only its own slot, so results do not depend on it.
constexpr int NXCLASS = 4;
__constant__ int XCLASS[NXCLASS] = {8, 16, 24, PMAX};
__global__ void __launch_bounds__(256) k_xorder(const XJob *jobs, const int *sel, int n, const double *TQH,
                                                 const double *TQF, const int *tqmap, int p, int full, int *ordH,
                                                 int *ordF, int *cnt, int *lists) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  int c = -1;
  if (e < n) {
    const XJob J = jobs[sel ? sel[e] : e];
    double rho, th, ph;
    cart_to_sph(J.P[0], J.P[1], J.P[2], rho, th, ph);
    const int tqi = tqmap ? tqmap[J.src] : J.src;
    const int pH = full ? p : compute_order(rho, TQH[tqi], J.a_src, p);
    const int pF = (J.kind == 1) ? (full ? p : compute_order(rho, TQF[tqi], J.a_src, p)) : 0;
    ordH[e] = pH;
    ordF[e] = pF;
    const int q = pH > pF ? pH : pF;
    c = 0;
    while (c < NXCLASS - 1 && q > XCLASS[c]) c++;
  }
  // one atomic per (wavefront, class): the first lane of the class reserves its lanes' slots
  const int lane = threadIdx.x % 64;
  for (int k = 0; k < NXCLASS; k++) {
    const unsigned long long mask = __ballot(c == k);
    if (!mask) continue;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&cnt[k], __popcll(mask));
    base = __shfl(base, leader, 64);
    if (c == k) {
      const int rank = __popcll(mask & ((1ull << lane) - 1));
      lists[(size_t)k * n + base + rank] = e;
    }
  }
}

// Mout(n, s) of one block for the singular rotations (theta = 0: identity; theta = pi:
// Mout(n,mm) = s Min(n,mm), s = (-1)^n at mm = 0 and mm = 1, alternating over mm) as a
// complex (n, m) value: re sign sg, im sign sgi (rotcoeff.cpp:368-393)
__device__ __forceinline__ void sing_signs(int sing, int n, int m, double &sg, double &sgi) {
  sg = 1.0;
  sgi = 1.0;
  if (sing == -1) {
    const int s0 = (n % 2 == 0) ? 1 : -1;
    sg = s0;
    if (m > 0) sgi = -s0;
  }
}

// A job whose rotation coefficients are stored (XJob::rot, the pair direction's ROT(n,s,m)
// for n < p as k_rotcoeff writes them, with xi = pi): forward-rotate every level at once,
// translate, then back-rotate, three barriers in all instead of the in-LDS recursion's
// four per level.  Each sum runs over the same terms in the same order as the generating
// path (XG-lane segments for the forward rotation and the translation, m in order for the
// back rotation), and the stored coefficients are the recursion's own values, so the
// outputs are the same.  LDS: a/b values [NPL][F,H][q(q+1)/2] complex, then the source
// blocks (fits the host's xform_lds for any q no larger than the launch's).
template <int XT, int NPL>
__device__ __forceinline__ void xapply_job(const XArgs &A, const XJob &J, int pH, int pF, int q, int tq) {
  extern __shared__ double sh[];
  constexpr int QM = (XT <= 128) ? 8 : ((XT == 256) ? 16 : PMAX);
  constexpr int NO = (QM * (QM + 1) / 2 + XT - 1) / XT;
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  const bool intra = (J.kind == 1);
  const int nfb = A.F ? NPL : 0;
  const int nl = q * (q + 1) / 2;  // (n, m) values with m <= n < q, index n(n+1)/2 + m
  cd *aH = reinterpret_cast<cd *>(sh);  // [NPL][nl], then [nfb][nl] for F
  cd *aF = aH + NPL * nl;
  double *hs = reinterpret_cast<double *>(aF + nfb * nl);  // [NPL][q^2]
  double *fs = hs + NPL * q * q;                            // [nfb][q^2]
  for (int i = tid; i < NPL * q * q; i += XT) {
    const int r = i / (q * q), c = i % (q * q);
    hs[i] = (c < pH * pH) ? A.H[(size_t)(J.src + r) * P2 + c] : 0.0;
    if (nfb) fs[i] = (intra && c < pF * pF) ? A.F[(size_t)(J.src + r) * P2 + c] : 0.0;
  }
  const cd *ROT = reinterpret_cast<const cd *>(J.rot);
  __syncthreads();
  const int g = tid / XG, lane = tid % XG;
  // (a) forward rotation of every level: a(n, m) = sum_s ROT(n,s,m) c(n,s)
  for (int t = g; t < nl; t += XT / XG) {
    int n = (int)((sqrtf(8.0f * t + 1.0f) - 1.0f) * 0.5f);
    while (n * (n + 1) / 2 > t) n--;
    while ((n + 1) * (n + 2) / 2 <= t) n++;
    const int m = t - n * (n + 1) / 2;
    const cd *row = ROT + rot_off(n) + m * (2 * n + 1) + n;  // row[s] = ROT(n,s,m)
    cd sH[NPL], sF[NPL];
#pragma unroll
    for (int r = 0; r < NPL; r++) sH[r] = sF[r] = {0.0, 0.0};
    for (int s = lane; s <= n; s += XG) {
      const cd rp = row[s], rm = row[-s];
#pragma unroll
      for (int r = 0; r < NPL; r++) {
        const double *h = hs + r * q * q, *f = fs + r * q * q;
        if (s == 0) {
          if (n < pH) sH[r] = cadd(sH[r], cmul(rp, cd{h[n * n], 0.0}));
          if (intra && n < pF) sF[r] = cadd(sF[r], cmul(rp, cd{f[n * n], 0.0}));
        } else {
          if (n < pH) {
            const cd c = {h[n * n + 2 * s - 1], h[n * n + 2 * s]};
            sH[r] = cadd(sH[r], cadd(cmul(rp, c), cmul(rm, cconj(c))));
          }
          if (intra && n < pF) {
            const cd c = {f[n * n + 2 * s - 1], f[n * n + 2 * s]};
            sF[r] = cadd(sF[r], cadd(cmul(rp, c), cmul(rm, cconj(c))));
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NPL; r++) {
      sH[r].x = seg_reduce(sH[r].x);
      sH[r].y = seg_reduce(sH[r].y);
      if (nfb) {
        sF[r].x = seg_reduce(sF[r].x);
        sF[r].y = seg_reduce(sF[r].y);
      }
      if (lane == 0) {
        if (m == 0) { sH[r].y = 0.0; sF[r].y = 0.0; }  // Mout(n,0) keeps the real part
        aH[r * nl + t] = sH[r];
        if (nfb) aF[r * nl + t] = sF[r];
      }
    }
  }
  __syncthreads();
  // (b) translation, one level per XG-lane group: the level's a values move to the lanes'
  // registers (lane k holds l = m + k, m + k + XG, ...), then b(n, m) = sum_l M_m[n][l] a(l, m)
  // for n in order is written over a(n, m) (only this group touches level m)
  constexpr int NL = (QM + XG - 1) / XG;
  for (int m = g; m < q; m += XT / XG) {
    cd rH[NL][NPL], rF[NL][NPL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int l = m + lane + k * XG;
#pragma unroll
      for (int r = 0; r < NPL; r++) {
        rH[k][r] = (l < q) ? aH[r * nl + l * (l + 1) / 2 + m] : cd{0.0, 0.0};
        rF[k][r] = (nfb && l < q) ? aF[r * nl + l * (l + 1) / 2 + m] : cd{0.0, 0.0};
      }
    }
    const int w = tq - m;
    for (int n = m; n < q; n++) {
      cd sH[NPL], sF[NPL];
#pragma unroll
      for (int r = 0; r < NPL; r++) sH[r] = sF[r] = {0.0, 0.0};
      if (n < pH) {
        const double *Mm = J.TH + tmat_off(tq, m) + (size_t)(n - m) * w;
#pragma unroll
        for (int k = 0; k < NL; k++) {
          const int l = m + lane + k * XG;
          if (l < pH) {
            const double tv = Mm[l - m];
#pragma unroll
            for (int r = 0; r < NPL; r++) sH[r] = cadd(sH[r], cscale(tv, rH[k][r]));
          }
        }
      }
      if (intra && n < pF) {
        const double *Mm = J.TF + tmat_off(tq, m) + (size_t)(n - m) * w;
#pragma unroll
        for (int k = 0; k < NL; k++) {
          const int l = m + lane + k * XG;
          if (l < pF) {
            const double tv = Mm[l - m];
#pragma unroll
            for (int r = 0; r < NPL; r++) sF[r] = cadd(sF[r], cscale(tv, rF[k][r]));
          }
        }
      }
      const int t = n * (n + 1) / 2 + m;
#pragma unroll
      for (int r = 0; r < NPL; r++) {
        sH[r].x = seg_reduce(sH[r].x);
        sH[r].y = seg_reduce(sH[r].y);
        if (nfb) {
          sF[r].x = seg_reduce(sF[r].x);
          sF[r].y = seg_reduce(sF[r].y);
        }
        if (lane == 0) {
          if (m == 0) { sH[r].y = 0.0; sF[r].y = 0.0; }
          aH[r * nl + t] = sH[r];
          if (nfb) aF[r * nl + t] = sF[r];
        }
      }
    }
  }
  __syncthreads();
  // (c) back rotation (rotcoeff.cpp:420-438): each thread's outputs (n, s), m in order
  const int nout = nl;
  for (int k = 0; k < NO; k++) {
    const int ek = tid + k * XT;
    if (ek >= nout) break;
    int n = (int)((sqrtf(8.0f * ek + 1.0f) - 1.0f) * 0.5f);
    while (n * (n + 1) / 2 > ek) n--;
    while ((n + 1) * (n + 2) / 2 <= ek) n++;
    const int s = ek - n * (n + 1) / 2;
    cd accH[NPL], accF[NPL];
#pragma unroll
    for (int r = 0; r < NPL; r++) accH[r] = accF[r] = {0.0, 0.0};
    const cd *rown = ROT + rot_off(n) + n;
    for (int m = 0; m <= n; m++) {
      const cd *row = rown + m * (2 * n + 1);
      const cd rp = cconj(row[s]), rm = row[-s];
      const int t = n * (n + 1) / 2 + m;
#pragma unroll
      for (int r = 0; r < NPL; r++) {
        if (n < pH) {
          const cd b = aH[r * nl + t];
          if (m == 0) accH[r] = cadd(accH[r], cmul(cscale(1.0, rm), cd{b.x, 0.0}));
          else accH[r] = cadd(accH[r], cadd(cmul(rp, b), cmul(rm, cconj(b))));
        }
        if (intra && n < pF) {
          const cd b = aF[r * nl + t];
          if (m == 0) accF[r] = cadd(accF[r], cmul(cscale(1.0, rm), cd{b.x, 0.0}));
          else accF[r] = cadd(accF[r], cadd(cmul(rp, b), cmul(rm, cconj(b))));
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NPL; r++) {
      double *oH = A.slotH + (size_t)(J.slot + r) * P2;
      if (n < pH) {
        if (s == 0) oH[n * n] = accH[r].x;
        else { oH[n * n + 2 * s - 1] = accH[r].x; oH[n * n + 2 * s] = accH[r].y; }
      } else {
        if (s == 0) oH[n * n] = 0.0;
        else { oH[n * n + 2 * s - 1] = 0.0; oH[n * n + 2 * s] = 0.0; }
      }
      if (intra) {
        double *oF = A.slotF + (size_t)(J.slot + r) * P2;
        if (n < pF) {
          if (s == 0) oF[n * n] = accF[r].x;
          else { oF[n * n + 2 * s - 1] = accF[r].x; oF[n * n + 2 * s] = accF[r].y; }
        } else {
          if (s == 0) oF[n * n] = 0.0;
          else { oF[n * n + 2 * s - 1] = 0.0; oF[n * n + 2 * s] = 0.0; }
        }
      }
    }
  }
  // orders q..p-1 of the slots are zero
#pragma unroll
  for (int r = 0; r < NPL; r++) {
    double *oH = A.slotH + (size_t)(J.slot + r) * P2;
    double *oF = intra ? A.slotF + (size_t)(J.slot + r) * P2 : nullptr;
    for (int i = q * q + tid; i < P2; i += XT) {
      oH[i] = 0.0;
      if (oF) oF[i] = 0.0;
    }
  }
}

// XT threads per workgroup: 64 (one wavefront) for q <= 8, 256 for q <= 16, 512 above
// (each thread owns at most NO of the q(q+1)/2 back-rotated outputs)
// one job: batch position e, plane offset pl (split launches)
template <int XT, int NPL>
__device__ __forceinline__ void xform_job(const Tabs &tb, const XArgs &A, int e, int pl) {
  extern __shared__ double sh[];
  constexpr int QM = (XT <= 128) ? 8 : ((XT == 256) ? 16 : PMAX);
  constexpr int NV = ((2 * QM - 2) * (2 * QM - 2) + XT - 1) / XT;  // staged recursion values per thread
  constexpr int NO = (QM * (QM + 1) / 2 + XT - 1) / XT;            // back-rotated outputs per thread
  const int job = A.sel ? A.sel[e] : e;
  XJob J = A.jobs[job];
  const int p = A.p, P2 = p * p;
  const int tid = threadIdx.x;
  const bool intra = (J.kind == 1);
  const int tqsrc = J.src;  // totals index of the job's source sphere (its first plane)
  J.src += pl;
  J.slot += pl;
  double rho, th, ph;
  cart_to_sph(J.P[0], J.P[1], J.P[2], rho, th, ph);
  int pH, pF = 0;
  if (A.ordH) {
    pH = A.ordH[e];
    if (intra) pF = A.ordF[e];
  } else {
    const int tqi = A.tqmap ? A.tqmap[tqsrc] : tqsrc;
    pH = A.full ? p : compute_order(rho, A.TQH[tqi], J.a_src, p);
    if (intra) pF = A.full ? p : compute_order(rho, A.TQF[tqi], J.a_src, p);
  }
  const int q = pH > pF ? pH : pF;
  const int NR = 2 * q - 1;  // rotation rows n < 2q-1
  const int tq = J.tord ? J.tord : pH;  // layout order of the translation matrices
  if (A.work && tid == 0) {
    double w = (J.rot ? 0.0 : work_rotgen(q)) + NPL * (2 * work_rotapply(pH) + work_transapply(pH));
    if (intra) w += NPL * (2 * work_rotapply(pF) + work_transapply(pF));
    atomicAdd(&A.work[0], (unsigned long long)w);
    atomicAdd(&A.work[1], (unsigned long long)(8 * NPL * (pH * pH + pF * pF + (intra ? 2 : 1) * P2)));
  }
  if (J.rot) {  // stored rotation coefficients (uniform across the workgroup)
    xapply_job<XT, NPL>(A, J, pH, pF, q, tq);
    return;
  }
  // LDS carve (the host's xform_lds): R, source blocks, level vectors
  const int nfb = A.F ? NPL : 0;
  cd *R = reinterpret_cast<cd *>(sh);                        // ROT level m, row n at n^2 + n + s
  double *hs = reinterpret_cast<double *>(R + NR * NR);     // [NPL][q^2] source H
  double *fs = hs + NPL * q * q;                             // [nfb][q^2] source F
  cd *aH = reinterpret_cast<cd *>(fs + nfb * q * q);         // [NPL][q] forward-rotated
  cd *aF = aH + NPL * q;                                     // [nfb][q]
  cd *bH = aF + nfb * q;                                     // [NPL][q] translated
  cd *bF = bH + NPL * q;                                     // [nfb][q]
  for (int i = tid; i < NPL * q * q; i += XT) {
    const int r = i / (q * q), c = i % (q * q);
    hs[i] = (c < pH * pH) ? A.H[(size_t)(J.src + r) * P2 + c] : 0.0;
    if (nfb) fs[i] = (intra && c < pF * pF) ? A.F[(size_t)(J.src + r) * P2 + c] : 0.0;
  }
  // rotation parameters (CRotCoeff::initParams rotcoeff.cpp:209-244), xi = pi
  double sint = sin(th), cost;
  int sing = 0;
  cd exphi, exxi;
  if (sint < 1e-12) {
    sing = (th < M_PI / 2) ? 1 : -1;
    sint = 0.0;
    cost = (sing == 1) ? 1.0 : -1.0;
    exphi = {1.0, 0.0};
  } else {
    cost = cos(th);
    double s_, c_;
    sincos(ph, &s_, &c_);
    exphi = {c_, s_};
  }
  const cd exiphi = cconj(exphi);
  {
    double s_, c_;
    sincos(M_PI, &s_, &c_);
    exxi = {c_, s_};
  }
  // back-rotated accumulators: each thread owns up to two (n,s) outputs
  const int nout = q * (q + 1) / 2;
  int on[NO], os[NO];
  cd accH[NO][NPL], accF[NO][NPL];
#pragma unroll
  for (int k = 0; k < NO; k++) {
    int ek = tid + k * XT;
    on[k] = -1;
    os[k] = 0;
#pragma unroll
    for (int r = 0; r < NPL; r++) accH[k][r] = accF[k][r] = {0.0, 0.0};
    if (ek < nout) {
      int n = (int)((sqrtf(8.0f * ek + 1.0f) - 1.0f) * 0.5f);
      while (n * (n + 1) / 2 > ek) n--;
      while ((n + 1) * (n + 2) / 2 <= ek) n++;
      on[k] = n;
      os[k] = ek - n * (n + 1) / 2;
    }
  }
  if (sing != 0) {
    // singular rotation: rotate() reduces to identity / parity maps (rotcoeff.cpp:368-393)
    __syncthreads();
    for (int m = 0; m < q; m++) {
      // forward "rotation" value a(n,m) = sign * c(n,m)
      for (int i = tid; i < NPL * q; i += XT) {
        const int r = i / q, n = i % q;
        if (n < m) continue;
        double sg, sgi;
        sing_signs(sing, n, m, sg, sgi);
        const double *h = hs + r * q * q;
        if (n < pH) aH[r * q + n] = (m == 0) ? cd{sg * h[n * n], 0.0}
                                             : cd{sg * h[n * n + 2 * m - 1], sgi * h[n * n + 2 * m]};
        else aH[r * q + n] = {0.0, 0.0};
        if (nfb) {
          const double *f = fs + r * q * q;
          if (intra && n < pF) aF[r * q + n] = (m == 0) ? cd{sg * f[n * n], 0.0}
                                                        : cd{sg * f[n * n + 2 * m - 1], sgi * f[n * n + 2 * m]};
          else aF[r * q + n] = {0.0, 0.0};
        }
      }
      __syncthreads();
      for (int i = tid; i < NPL * q; i += XT) {
        const int r = i / q, n = i % q;
        if (n < m) continue;
        const int w = tq - m;
        cd sH = {0, 0}, sF = {0, 0};
        if (n < pH) {
          const double *Mm = J.TH + tmat_off(tq, m) + (size_t)(n - m) * w;
          for (int l = m; l < pH; l++) sH = cadd(sH, cscale(Mm[l - m], aH[r * q + l]));
        }
        if (intra && n < pF) {
          const double *Mm = J.TF + tmat_off(tq, m) + (size_t)(n - m) * w;
          for (int l = m; l < pF; l++) sF = cadd(sF, cscale(Mm[l - m], aF[r * q + l]));
        }
        bH[r * q + n] = sH;
        if (nfb) bF[r * q + n] = sF;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NO; k++) {
        const int n = on[k], s = os[k];
        if (n < 0 || s != m) continue;
        double sg, sgi;
        sing_signs(sing, n, m, sg, sgi);
#pragma unroll
        for (int r = 0; r < NPL; r++) {
          const cd b = bH[r * q + n];
          accH[k][r] = (m == 0) ? cd{sg * b.x, 0.0} : cd{sg * b.x, sgi * b.y};
          if (nfb) {
            const cd bf = bF[r * q + n];
            accF[k][r] = (m == 0) ? cd{sg * bf.x, 0.0} : cd{sg * bf.x, sgi * bf.y};
          }
        }
      }
      __syncthreads();
    }
  } else {
    // level 0 seeds: ROT(n,s,0) = conj-ordered SH of order 2q-1 (rotcoeff.cpp:785-795)
    for (int m = tid; m < NR; m += XT) {
      sh_column(tb, NR, m, cost, sint, ph, [&](int n, double re, double im) {
        if (m == 0) {
          R[n * n + n] = {re, 0.0};
        } else {
          R[n * n + n + m] = {re, -im};  // ROT(n, m,0) = comp(n,-m)
          R[n * n + n - m] = {re, im};   // ROT(n,-m,0) = comp(n, m)
        }
      });
    }
    const double r1 = -0.5 * (1 + cost), r2 = 0.5 * (1 - cost), r3 = -sint;
    __syncthreads();
    for (int m = 0; m < q; m++) {
      // (a) forward rotation of this level: a(n) for n in [m, q)
      {
        const int g = tid / XG, lane = tid % XG;
        for (int n = m + g; n < q; n += XT / XG) {
          const cd *row = R + n * n + n;  // row[s] = ROT(n,s,m)
          cd sH[NPL], sF[NPL];
#pragma unroll
          for (int r = 0; r < NPL; r++) sH[r] = sF[r] = {0.0, 0.0};
          for (int s = lane; s <= n; s += XG) {
            const cd rp = row[s], rm = row[-s];
#pragma unroll
            for (int r = 0; r < NPL; r++) {
              const double *h = hs + r * q * q, *f = fs + r * q * q;
              if (s == 0) {
                if (n < pH) sH[r] = cadd(sH[r], cmul(rp, cd{h[n * n], 0.0}));
                if (intra && n < pF) sF[r] = cadd(sF[r], cmul(rp, cd{f[n * n], 0.0}));
              } else {
                if (n < pH) {
                  const cd c = {h[n * n + 2 * s - 1], h[n * n + 2 * s]};
                  sH[r] = cadd(sH[r], cadd(cmul(rp, c), cmul(rm, cconj(c))));
                }
                if (intra && n < pF) {
                  const cd c = {f[n * n + 2 * s - 1], f[n * n + 2 * s]};
                  sF[r] = cadd(sF[r], cadd(cmul(rp, c), cmul(rm, cconj(c))));
                }
              }
            }
          }
#pragma unroll
          for (int r = 0; r < NPL; r++) {
            sH[r].x = seg_reduce(sH[r].x);
            sH[r].y = seg_reduce(sH[r].y);
            if (nfb) {
              sF[r].x = seg_reduce(sF[r].x);
              sF[r].y = seg_reduce(sF[r].y);
            }
            if (lane == 0) {
              if (m == 0) { sH[r].y = 0.0; sF[r].y = 0.0; }  // Mout(n,0) keeps the real part
              aH[r * q + n] = sH[r];
              if (nfb) aF[r * q + n] = sF[r];
            }
          }
        }
      }
      __syncthreads();
      // (b) translate this level: b(n) = sum_l M_m[n][l] a(l)
      {
        const int g = tid / XG, lane = tid % XG;
        const int w = tq - m;
        for (int n = m + g; n < q; n += XT / XG) {
          cd sH[NPL], sF[NPL];
#pragma unroll
          for (int r = 0; r < NPL; r++) sH[r] = sF[r] = {0.0, 0.0};
          if (n < pH) {
            const double *Mm = J.TH + tmat_off(tq, m) + (size_t)(n - m) * w;
            for (int l = m + lane; l < pH; l += XG) {
              const double t = Mm[l - m];
#pragma unroll
              for (int r = 0; r < NPL; r++) sH[r] = cadd(sH[r], cscale(t, aH[r * q + l]));
            }
          }
          if (intra && n < pF) {
            const double *Mm = J.TF + tmat_off(tq, m) + (size_t)(n - m) * w;
            for (int l = m + lane; l < pF; l += XG) {
              const double t = Mm[l - m];
#pragma unroll
              for (int r = 0; r < NPL; r++) sF[r] = cadd(sF[r], cscale(t, aF[r * q + l]));
            }
          }
#pragma unroll
          for (int r = 0; r < NPL; r++) {
            sH[r].x = seg_reduce(sH[r].x);
            sH[r].y = seg_reduce(sH[r].y);
            if (nfb) {
              sF[r].x = seg_reduce(sF[r].x);
              sF[r].y = seg_reduce(sF[r].y);
            }
            if (lane == 0) {
              if (m == 0) { sH[r].y = 0.0; sF[r].y = 0.0; }
              bH[r * q + n] = sH[r];
              if (nfb) bF[r * q + n] = sF[r];
            }
          }
        }
      }
      __syncthreads();
      // (c) back-rotation contribution of level m (rotcoeff.cpp:420-438)
#pragma unroll
      for (int k = 0; k < NO; k++) {
        const int n = on[k], s = os[k];
        if (n < m) continue;
        const cd *row = R + n * n + n;
        const cd rp = cconj(row[s]), rm = row[-s];
#pragma unroll
        for (int r = 0; r < NPL; r++) {
          if (n < pH) {
            const cd b = bH[r * q + n];
            if (m == 0) accH[k][r] = cadd(accH[k][r], cmul(cscale(1.0, rm), cd{b.x, 0.0}));
            else accH[k][r] = cadd(accH[k][r], cadd(cmul(rp, b), cmul(rm, cconj(b))));
          }
          if (intra && n < pF) {
            const cd b = bF[r * q + n];
            if (m == 0) accF[k][r] = cadd(accF[k][r], cmul(cscale(1.0, rm), cd{b.x, 0.0}));
            else accF[k][r] = cadd(accF[k][r], cadd(cmul(rp, b), cmul(rm, cconj(b))));
          }
        }
      }
      // (d) recursion to level m+1 in place (rotcoeff.cpp:801-809), rows n-1 for n in
      // [m+2, 2q-m-2]: the new values are staged in registers, then written over level m
      if (m < q - 1) {
        const int base = (m + 1) * (m + 1);
        const int total = (2 * q - m - 2) * (2 * q - m - 2) - (m + 1) * (m + 1);
        cd vR[NV];
        int vi[NV];
        int cnt = 0;
#pragma unroll
        for (int t0 = 0; t0 < NV; t0++) {
          const int ee = tid + t0 * XT;
          if (ee >= total) break;
          int t = ee + base;
          int n = (int)sqrtf((float)t);  // exact floor for t < 2^22 after the corrections
          while (n * n > t) n--;
          while ((n + 1) * (n + 1) <= t) n++;
          n += 1;  // element t belongs to row n: start (n-1)^2
          const int s = t - (n - 1) * (n - 1) - (n - 1);
          const cd *row = R + n * n + n;
          cd d1 = cscale(r1 * ETA(tb, n, s - 1), row[s - 1]);
          cd d2 = cscale(r2 * ETA(tb, n, -s - 1), row[s + 1]);
          cd d3 = cscale(r3 * ZETA(tb, n - 1, abs(s)), row[s]);
          cd v = cadd(cadd(cmul(d1, exiphi), cmul(d2, exphi)), d3);
          v = cmul(exxi, v);
          const double em = ETA(tb, n, m);
          vR[t0] = {v.x / em, v.y / em};
          vi[t0] = (n - 1) * (n - 1) + (n - 1) + s;
          cnt = t0 + 1;
        }
        __syncthreads();
#pragma unroll
        for (int t0 = 0; t0 < NV; t0++)
          if (t0 < cnt) R[vi[t0]] = vR[t0];
      }
      __syncthreads();
    }
  }
  // write slots (orders beyond the truncation are zero)
#pragma unroll
  for (int r = 0; r < NPL; r++) {
    double *oH = A.slotH + (size_t)(J.slot + r) * P2;
    double *oF = intra ? A.slotF + (size_t)(J.slot + r) * P2 : nullptr;
    for (int i = tid; i < P2; i += XT) {
      oH[i] = 0.0;
      if (oF) oF[i] = 0.0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NO; k++) {
    const int n = on[k], s = os[k];
    if (n < 0) continue;
#pragma unroll
    for (int r = 0; r < NPL; r++) {
      double *oH = A.slotH + (size_t)(J.slot + r) * P2;
      if (n < pH) {
        if (s == 0) oH[n * n] = accH[k][r].x;
        else { oH[n * n + 2 * s - 1] = accH[k][r].x; oH[n * n + 2 * s] = accH[k][r].y; }
      }
      if (intra && n < pF) {
        double *oF = A.slotF + (size_t)(J.slot + r) * P2;
        if (s == 0) oF[n * n] = accF[k][r].x;
        else { oF[n * n + 2 * s - 1] = accF[k][r].x; oF[n * n + 2 * s] = accF[k][r].y; }
      }
    }
  }
}
// The stored-rotation path of k_xform on its own (launches whose jobs all carry stored
// rotations, J.rot): without the level recursion's registers more workgroups fit on a CU.
// Same batch positions, orders, planes and instrumentation as k_xform; the LDS holds only the
// source blocks and level vectors.
template <int XT, int NPL>
__global__ void __launch_bounds__(XT) k_xapply(XArgs A) {
  if (A.cnt && (int)blockIdx.x >= *A.cnt) return;
  const int e = A.cls ? A.cls[blockIdx.x] : blockIdx.x;
  const int pl = (NPL == 1 && A.split) ? blockIdx.y : 0;
  XJob J = A.jobs[A.sel ? A.sel[e] : e];
  const int p = A.p;
  const bool intra = (J.kind == 1);
  const int tqsrc = J.src;
  J.src += pl;
  J.slot += pl;
  int pH, pF = 0;
  if (A.ordH) {
    pH = A.ordH[e];
    if (intra) pF = A.ordF[e];
  } else if (A.full) {
    pH = p;
    if (intra) pF = p;
  } else {
    double rho, th, ph;
    cart_to_sph(J.P[0], J.P[1], J.P[2], rho, th, ph);
    const int tqi = A.tqmap ? A.tqmap[tqsrc] : tqsrc;
    pH = compute_order(rho, A.TQH[tqi], J.a_src, p);
    if (intra) pF = compute_order(rho, A.TQF[tqi], J.a_src, p);
  }
  const int q = pH > pF ? pH : pF;
  if (A.work && threadIdx.x == 0) {
    double w = NPL * (2 * work_rotapply(pH) + work_transapply(pH));
    if (intra) w += NPL * (2 * work_rotapply(pF) + work_transapply(pF));
    atomicAdd(&A.work[0], (unsigned long long)w);
    atomicAdd(&A.work[1], (unsigned long long)(8 * NPL * (pH * pH + pF * pF + (intra ? 2 : 1) * p * p)));
  }
  xapply_job<XT, NPL>(A, J, pH, pF, q, J.tord ? J.tord : pH);
}
template <int XT, int NPL>
__global__ void __launch_bounds__(XT) k_xform(Tabs tb, XArgs A) {
  // one job per workgroup; a class of a binned batch is launched for the whole batch and
  // its workgroups past the class's device-side count leave at once
  if (A.cnt && (int)blockIdx.x >= *A.cnt) return;
  const int e = A.cls ? A.cls[blockIdx.x] : blockIdx.x;
  xform_job<XT, NPL>(tb, A, e, (NPL == 1 && A.split) ? blockIdx.y : 0);
}


// ----------------------------------------------------------------------------
// point-to-expansion sums: CLocalExpan / CMulExpan of point charges
// (expansion.cpp:203-340), used for the numerical re-expansion of surface charges
// (CXFormNIntra::xformFH, xforms.cpp:528-550) and for fixed-charge expansions.
// One workgroup per job, or (nsplit > 1) per (job, contiguous range of spl points); points
// are processed in tiles held in LDS, one thread per coefficient accumulates over the
// points of its range in their input order.  Split launches write per-range partial sums
// that k_p2lpsum adds in range order (a fixed order, so results do not vary run to run).
// ----------------------------------------------------------------------------
constexpr int PT = 256;

struct PJob {
  int pt_off, npts;  // point range (xyz and charges share the index)
  int slot;          // output slot
  int src;           // source sphere for the truncation order (-1: full order)
  double P[3];       // target centre - source centre (points are relative to the source)
  double a_src, a_tgt;
};
struct PArgs {
  const PJob *jobs;
  const double *xyz;
  const double *q1, *q2;     // two charge sets on the same points (q2 may be null)
  const double *TQ1, *TQ2;   // source totals for the orders (null: full order)
  double *out1, *out2;
  int kap1, kap2;            // screened (k-hat with kappa) for each set
  int multipole;             // 0: local (CLExpan), 1: multipole (CMExpan)
  int tile;                  // points per tile
  const int *sel;            // job indices to run (null: job = workgroup index)
  int p;
  unsigned long long *work;  // instrumentation: [flops, bytes] (null: off)
  // gradient charges: per-point squared norms of the 3-vector charge of each set; a point
  // whose norm is below 1e-15 is skipped (triexpan.cpp:127-153).  Null: skip zero charges.
  const double *n1, *n2;
  // point split (nsplit > 1): workgroup (j, s) takes points [s*spl, (s+1)*spl) of job j and
  // writes its partial sums to part[(j*nsplit + s)*2*P2 ...] (set 1, then set 2)
  int nsplit, spl;
  double *part;
};

__global__ void __launch_bounds__(PT) k_p2l(Tabs tb, PArgs A) {
  extern __shared__ double sh[];
  const PJob J = A.jobs[A.sel ? A.sel[blockIdx.x] : blockIdx.x];
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  const bool two = (A.q2 != nullptr);
  double rho0 = sqrt(J.P[0] * J.P[0] + J.P[1] * J.P[1] + J.P[2] * J.P[2]);
  int p1 = p, p2 = two ? p : 0;
  if (J.src >= 0 && A.TQ1) p1 = compute_order(rho0, A.TQ1[J.src], J.a_src, p);
  if (two && J.src >= 0 && A.TQ2) p2 = compute_order(rho0, A.TQ2[J.src], J.a_src, p);
  const int q = p1 > p2 ? p1 : p2;
  const int T = A.tile;
  const bool split = A.nsplit > 1;
  const int pt_lo = split ? min(J.npts, (int)blockIdx.y * A.spl) : 0;
  const int pt_hi = split ? min(J.npts, pt_lo + A.spl) : J.npts;
  if (A.work && tid == 0 && blockIdx.y == 0) {
    const double w = work_p2l_point(p1) + (two ? work_p2l_point(p2) : 0.0);
    atomicAdd(&A.work[0], (unsigned long long)(w * J.npts));
    atomicAdd(&A.work[1], (unsigned long long)(8 * ((size_t)J.npts * (two ? 5 : 4) + (two ? 2 : 1) * P2)));
  }
  double *Y = sh;                    // [T][q*q]
  double *w1 = Y + (size_t)T * q * q; // [T][q]
  double *w2 = w1 + T * q;           // [T][q]
  double *geo = w2 + T * q;          // [T][3]: cost, sint, phi
  double acc1[4] = {0, 0, 0, 0}, acc2[4] = {0, 0, 0, 0};
  const int q2 = q * q;
  for (int t0 = pt_lo; t0 < pt_hi; t0 += T) {
    const int nt = min(T, pt_hi - t0);
    // phase 0: geometry and radial factors of each point
    for (int t = tid; t < nt; t += PT) {
      const double *x = A.xyz + 3 * (size_t)(J.pt_off + t0 + t);
      double dx = x[0] - J.P[0], dy = x[1] - J.P[1], dz = x[2] - J.P[2];
      double rho, th, ph;
      cart_to_sph(dx, dy, dz, rho, th, ph);
      geo[3 * t] = cos(th);
      geo[3 * t + 1] = sin(th);
      geo[3 * t + 2] = ph;
      for (int set = 0; set < (two ? 2 : 1); set++) {
        const double ch = (set == 0 ? A.q1 : A.q2)[J.pt_off + t0 + t];
        const int kap = (set == 0 ? A.kap1 : A.kap2);
        const int qq = (set == 0 ? p1 : p2);
        double *w = (set == 0 ? w1 : w2) + t * q;
        const double *nrm = (set == 0 ? A.n1 : A.n2);
        if (nrm ? (nrm[J.pt_off + t0 + t] < 1e-15) : (ch == 0.0)) {
          for (int n = 0; n < q; n++) w[n] = 0.0;
          continue;
        }
        const double val = kap ? rho * tb.kappa : 0.0;
        if (!A.multipole) {
          // CLExpan::init + BESSEL + applyRadialComponent (expansion.cpp:289-316,422-436)
          double mr = 1.0 / rho;
          double mk = ch;
          mk *= mr * (kap ? exp(-val) : 1.0);
          mr *= J.a_tgt;
          double km2 = 1.0, km1 = 1.0 + val, vs = val * val;
          for (int n = 0; n < qq; n++) {
            double K;
            if (val == 0.0) K = 1.0;
            else if (n == 0) K = 1.0;
            else if (n == 1) K = km1;
            else {
              K = bessel_k_next(tb, n, km1, km2, vs);
              km2 = km1;
              km1 = K;
            }
            w[n] = K * mk;
            mk *= mr;
          }
        } else {
          // CMExpan::init + BESSEL + applyRadialComponent (expansion.cpp:323-365)
          double mscale = 1.0 / J.a_tgt;
          double mr = rho * mscale;
          double mk = ch;
          double z = 0.5 * val * val;
          for (int n = 0; n < qq; n++) {
            double I = 1.0;
            if (val != 0.0) {
              double tt = z / (2 * n + 3);
              for (int j = 1; j <= 20; j++) {
                I += tt;
                tt *= (z / ((j + 1) * (2 * (n + j) + 3)));
                if (tt < 1e-20) break;
              }
            }
            w[n] = I * mk;
            mk *= mr;
          }
        }
        for (int n = qq; n < q; n++) w[n] = 0.0;
      }
    }
    __syncthreads();
    // phase A: spherical harmonics of the tile, one (point, column) per task
    for (int task = tid; task < nt * q; task += PT) {
      const int t = task / q, m = task % q;
      double *Yt = Y + (size_t)t * q2;
      sh_column(tb, q, m, geo[3 * t], geo[3 * t + 1], geo[3 * t + 2], [&](int n, double re, double im) {
        if (m == 0) Yt[n * n] = re;
        else {
          Yt[n * n + 2 * m - 1] = re;
          Yt[n * n + 2 * m] = im;
        }
      });
    }
    __syncthreads();
    // phase B: accumulate in point order
    for (int k = 0; k < 4; k++) {
      const int c = tid + k * PT;
      if (c >= q2) break;
      int n = (int)sqrt((double)c);
      while (n * n > c) n--;
      while ((n + 1) * (n + 1) <= c) n++;
      for (int t = 0; t < nt; t++) {
        double y = Y[(size_t)t * q2 + c];
        if (n < p1) acc1[k] += y * w1[t * q + n];
        if (two && n < p2) acc2[k] += y * w2[t * q + n];
      }
    }
    __syncthreads();
  }
  double *o1 = split ? A.part + ((size_t)blockIdx.x * A.nsplit + blockIdx.y) * 2 * P2 : A.out1 + (size_t)J.slot * P2;
  double *o2 = two ? (split ? o1 + P2 : A.out2 + (size_t)J.slot * P2) : nullptr;
  for (int i = tid; i < P2; i += PT) {
    o1[i] = 0.0;
    if (o2) o2[i] = 0.0;
  }
  __syncthreads();
  for (int k = 0; k < 4; k++) {
    const int c = tid + k * PT;
    if (c >= q2) break;
    int n = (int)sqrt((double)c);
    while (n * n > c) n--;
    while ((n + 1) * (n + 1) <= c) n++;
    if (n < p1) o1[c] = acc1[k];
    if (o2 && n < p2) o2[c] = acc2[k];
  }
}

// Partial sums of a split k_p2l launch added in range order: one thread per coefficient of
// each job's output slot(s).
__global__ void __launch_bounds__(256) k_p2lpsum(PArgs A) {
  const int jb = blockIdx.x;
  const PJob J = A.jobs[A.sel ? A.sel[jb] : jb];
  const int P2 = A.p * A.p, nset = A.q2 ? 2 : 1;
  const int c = blockIdx.y * 256 + threadIdx.x;
  if (c >= nset * P2) return;
  const double *src = A.part + (size_t)jb * A.nsplit * 2 * P2 + c;
  const size_t st = 2 * (size_t)P2;
  double acc = 0.0;
  int k = 0;
  for (; k + 8 <= A.nsplit; k += 8) {  // eight loads in flight, added in range order
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = src[(k + u) * st];
#pragma unroll
    for (int u = 0; u < 8; u++) acc += v[u];
  }
  for (; k < A.nsplit; k++) acc += src[k * st];
  if (c < P2) A.out1[(size_t)J.slot * P2 + c] = acc;
  else A.out2[(size_t)J.slot * P2 + c - P2] = acc;
}

// ----------------------------------------------------------------------------
// rotation coefficients stored to HBM: molecule orientations (CRotCoeff::reset
// rotcoeff.cpp:155-171,780-810).  Stores ROT(n,s,m) for n < p.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(RT) k_rotcoeff(Tabs tb, RJob *jobs, int p, double *out) {
  extern __shared__ double sh[];
  RJob &J = jobs[blockIdx.x];
  const int tid = threadIdx.x, q = p, NR = 2 * q - 1;
  cd *cur = reinterpret_cast<cd *>(sh);
  cd *nxt = cur + NR * NR;
  double th = J.theta, ph = J.phi;
  double sint = sin(th), cost;
  int sing = 0;
  cd exphi, exxi;
  if (sint < 1e-12) {
    sing = (th < M_PI / 2) ? 1 : -1;
    sint = 0.0;
    cost = (sing == 1) ? 1.0 : -1.0;
    exphi = {1.0, 0.0};
  } else {
    cost = cos(th);
    double s_, c_;
    sincos(J.phi, &s_, &c_);
    exphi = {c_, s_};
  }
  const cd exiphi = cconj(exphi);
  {
    double s_, c_;
    sincos(J.xi, &s_, &c_);
    exxi = {c_, s_};
  }
  if (tid == 0) J.sing = sing;
  if (sing != 0) return;
  cd *R = reinterpret_cast<cd *>(out + J.off);
  for (int m = tid; m < NR; m += RT) {
    sh_column(tb, NR, m, cost, sint, ph, [&](int n, double re, double im) {
      if (m == 0) cur[n * n + n] = {re, 0.0};
      else {
        cur[n * n + n + m] = {re, -im};
        cur[n * n + n - m] = {re, im};
      }
    });
  }
  const double r1 = -0.5 * (1 + cost), r2 = 0.5 * (1 - cost), r3 = -sint;
  __syncthreads();
  for (int m = 0; m < q; m++) {
    // store level m rows n in [m, q)
    for (int n = m; n < q; n++)
      for (int s = -n + tid; s <= n; s += RT) R[rot_off(n) + m * (2 * n + 1) + s + n] = cur[n * n + n + s];
    if (m < q - 1) {
      const int base = (m + 1) * (m + 1);
      const int total = (2 * q - m - 2) * (2 * q - m - 2) - (m + 1) * (m + 1);
      for (int e = tid; e < total; e += RT) {
        int t = e + base;
        int n = (int)sqrt((double)t);
        while (n * n > t) n--;
        while ((n + 1) * (n + 1) <= t) n++;
        n += 1;
        int s = t - (n - 1) * (n - 1) - (n - 1);
        const cd *row = cur + n * n + n;
        cd d1 = cscale(r1 * ETA(tb, n, s - 1), row[s - 1]);
        cd d2 = cscale(r2 * ETA(tb, n, -s - 1), row[s + 1]);
        cd d3 = cscale(r3 * ZETA(tb, n - 1, abs(s)), row[s]);
        cd v = cadd(cadd(cmul(d1, exiphi), cmul(d2, exphi)), d3);
        v = cmul(exxi, v);
        double em = ETA(tb, n, m);
        nxt[(n - 1) * (n - 1) + (n - 1) + s] = {v.x / em, v.y / em};
      }
    }
    __syncthreads();
    cd *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
}

// rotateWithXi (rotcoeff.cpp:250-350) of whole blocks with stored coefficients.
struct RAJob {
  int in, out;       // block indices into the in / out arrays
  int rot;           // rotation job index
  int fwd;
};
// Grid (job, output chunk of RAT): the input block is staged in LDS and each thread forms
// one output (n, k) as the reference's sum over s (or m) in order, with the rotation
// coefficients loaded ahead.
constexpr int RAT = 128;
__global__ void __launch_bounds__(RAT) k_rotapply(const RAJob *jobs, const int *sel, const RJob *rj,
                                                  const double *Rst, const double *in, double *out, int p) {
  __shared__ double mi_s[PMAX * PMAX];
  const RAJob J = jobs[sel ? sel[blockIdx.x] : blockIdx.x];
  const RJob R = rj[J.rot];
  const int P2 = p * p;
  const double *Min = in + (size_t)J.in * P2;
  double *Mout = out + (size_t)J.out * P2;
  for (int i = threadIdx.x; i < P2; i += RAT) mi_s[i] = Min[i];
  __syncthreads();
  const cd *C = reinterpret_cast<const cd *>(Rst + R.off);
  const int nout = p * (p + 1) / 2;
  const int e = blockIdx.y * RAT + threadIdx.x;
  if (e >= nout) return;
  int n = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
  while (n * (n + 1) / 2 > e) n--;
  while ((n + 1) * (n + 1 + 1) / 2 <= e) n++;
  const int k = e - n * (n + 1) / 2;  // m (forward) or s (transposed)
  const double *mi = mi_s + n * n;
  if (R.sing != 0) {
    double xi = J.fwd ? R.xi : -R.xi;
    if (k == 0) {
      Mout[n * n] = (R.sing == -1 && (n % 2)) ? -mi[0] : mi[0];
    } else {
      cd inc = {mi[2 * k - 1], mi[2 * k]};
      double sx, cx;
      sincos(k * xi, &sx, &cx);
      cd o = cmul(cd{cx, sx}, inc);
      if (R.sing == -1) {
        int sg = (n % 2 == 0) ? 1 : -1;
        Mout[n * n + 2 * k - 1] = sg * o.x;
        Mout[n * n + 2 * k] = -sg * o.y;
      } else {
        Mout[n * n + 2 * k - 1] = o.x;
        Mout[n * n + 2 * k] = o.y;
      }
    }
    return;
  }
  const cd *row = C + rot_off(n);  // row[m*(2n+1) + s + n]
  const int W = 2 * n + 1;
  cd o;
  if (J.fwd) {
    const int m = k;
    const cd *rm = row + m * W + n;  // rm[s] = ROT(n, s, m)
    o = cmul(cscale(1.0, rm[0]), cd{mi[0], 0.0});
    int s = 1;
    for (; s + 4 <= n + 1; s += 4) {
      cd a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a[u] = rm[s + u];
        b[u] = rm[-(s + u)];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const cd inc = {mi[2 * (s + u) - 1], mi[2 * (s + u)]};
        o = cadd(o, cadd(cmul(a[u], inc), cmul(b[u], cconj(inc))));
      }
    }
    for (; s <= n; s++) {
      const cd inc = {mi[2 * s - 1], mi[2 * s]};
      o = cadd(o, cadd(cmul(rm[s], inc), cmul(rm[-s], cconj(inc))));
    }
  } else {
    const int s = k;
    o = cmul(cscale(1.0, row[0 * W + n - s]), cd{mi[0], 0.0});
    int m = 1;
    for (; m + 4 <= n + 1; m += 4) {
      cd a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a[u] = row[(m + u) * W + n + s];
        b[u] = row[(m + u) * W + n - s];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const cd inc = {mi[2 * (m + u) - 1], mi[2 * (m + u)]};
        o = cadd(o, cadd(cmul(cconj(a[u]), inc), cmul(b[u], cconj(inc))));
      }
    }
    for (; m <= n; m++) {
      const cd inc = {mi[2 * m - 1], mi[2 * m]};
      o = cadd(o, cadd(cmul(cconj(row[m * W + n + s]), inc), cmul(row[m * W + n - s], cconj(inc))));
    }
  }
  if (k == 0) Mout[n * n] = o.x * 1.0;
  else {
    Mout[n * n + 2 * k - 1] = o.x;
    Mout[n * n + 2 * k] = o.y;
  }
}

// ----------------------------------------------------------------------------
// per-sphere local-field assembly + inner F/H iteration (solveFH)
// ----------------------------------------------------------------------------
constexpr int ST = 256;

struct SArgs {
  const int *sph;          // sphere index per workgroup
  const int *lst_off;      // per workgroup: CSR offsets into lst (intra slots)
  const int *lst;          // slot ids
  const double *slotH, *slotF;
  const double *LS;        // [sphere] inter-molecular local expansion, or null
  const double *farH, *farF;  // [sphere] frozen local expansions, or null
  const double *consts;    // [sphere][10][p]
  const double *Dfix, *Vfix;  // [sphere][p^2]
  const double *const *imat;  // [sphere] column-major p^2 x p^2 matrix, null: identity
  double *F, *H;           // [sphere][p^2] in/out
  double *LFS, *LHS;       // [sphere][p^2] out
  double *dev;             // [sphere] out: computeDev2(H_in, H_out)
  int *sweeps;             // [sphere] out (may be null)
  int p;
  unsigned long long *work;  // instrumentation: [flops, bytes] (null: off)
  const int *sel;          // entries of sph/lst_off to run (null: the workgroup index)
};

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x / 64); i++) s += red[i];
  return s;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// a double moved across lanes by a DPP control (lanes without a source read 0.0)
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Sum over the 64 lanes of a wavefront with DPP row shifts and row broadcasts (gfx9
// family), the total broadcast to every lane.  Fixed order: deterministic.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_d<0x111, 0xf>(v);   // row_shr:1
  v += dpp_d<0x112, 0xf>(v);   // row_shr:2
  v += dpp_d<0x114, 0xf>(v);   // row_shr:4
  v += dpp_d<0x118, 0xf>(v);   // row_shr:8   -> lane 15 of each row holds the row sum
  v += dpp_d<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
  v += dpp_d<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3 -> lane 63 holds the total
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// computeDev term of one (n, m) pair given its old (ar, ai) and new (br, bi) values
// (expcenter.cpp:1270-1328); returns false where the pair is skipped
__device__ __forceinline__ bool dev_term(double ar, double ai, double br, double bi, bool mpos, double &s) {
  const double PL = 1e-30;
  if (fabs(ar) < PL && fabs(ai) < PL) {
    if (fabs(br) < PL && fabs(bi) < PL) return false;
    s = 1.0;
  } else if (fabs(br) < PL && fabs(bi) < PL) {
    s = 1.0;
  } else {
    double dr = ar - br, di = ai - bi;
    double top = dr * dr + di * di;
    double bot = ar * ar + ai * ai + br * br + bi * bi;
    s = mpos ? 2 * top / bot : top / bot;
  }
  return true;
}

// Inner F/H iteration of a sphere whose surface-integral matrix is the identity: every
// coefficient pair (F_i, H_i) evolves on its own, so the block lives in registers spread
// over the workgroup's NW wavefronts (KW (n, m) pairs per lane) and only the stopping test
// is reduced across lanes and wavefronts, once per sweep.  Same arithmetic per coefficient
// as the general path (expcenter.cpp:1080-1135, computeDev :1270-1328).
template <int KW>
struct IdState {
  double Fr[KW], Fi[KW], Hr[KW], Hi[KW], fbr[KW], fbi[KW], hbr[KW], hbi[KW], uHr[KW], uHi[KW];
  double kH1[KW], kF1[KW], kH2[KW], kF2[KW], kIH[KW], kInv[KW];
  double c2[KW];  // fast stopping test: the pair alone exceeds the threshold when top > c2 bot
  int nn[KW], i0[KW];
};
// (n, m) pair index ep -> n, and the real index of c(n, m)
__device__ __forceinline__ void pair_nm(int ep, int &n, int &i) {
  n = (int)((sqrt(8.0 * ep + 1.0) - 1.0) * 0.5);
  while (n * (n + 1) / 2 > ep) n--;
  while ((n + 1) * (n + 2) / 2 <= ep) n++;
  const int m = ep - n * (n + 1) / 2;
  i = (m == 0) ? n * n : n * n + 2 * m - 1;
}
template <int KW>
__device__ __forceinline__ void id_consts(IdState<KW> &S, int k, int n, const double *cs, int p) {
  S.kH1[k] = cs[n];
  S.kF1[k] = cs[p + n];
  S.kH2[k] = cs[4 * p + n];
  S.kF2[k] = cs[5 * p + n];
  S.kIH[k] = cs[8 * p + n];
  S.kInv[k] = 1 / (double)(2 * n + 1);
}
// one sweep of this lane's pairs; returns the lane's computeDev partial (old values: H at
// the start of the sweep)
template <int KW>
__device__ __forceinline__ double id_sweep(IdState<KW> &S) {
  double sum = 0.0;
  for (int k = 0; k < KW; k++) {
    if (S.nn[k] < 0) continue;
    const bool mpos = (S.i0[k] != S.nn[k] * S.nn[k]);
    const double oHr = S.Hr[k], oHi = S.Hi[k];
    double xr = S.fbr[k] + S.kH1[k] * S.Hr[k] + S.kF1[k] * S.Fr[k];
    double xi = S.fbi[k] + S.kH1[k] * S.Hi[k] + S.kF1[k] * S.Fi[k];
    S.Fr[k] = xr * S.kInv[k];
    S.Fi[k] = mpos ? xi * S.kInv[k] : 0.0;
    xr = S.hbr[k] + S.kH2[k] * S.Hr[k] + S.kF2[k] * S.Fr[k];
    xi = S.hbi[k] + S.kH2[k] * S.Hi[k] + S.kF2[k] * S.Fi[k];
    S.Hr[k] = xr * S.kIH[k];
    S.Hi[k] = mpos ? xi * S.kIH[k] : 0.0;
    double sv;
    if (dev_term(oHr, mpos ? oHi : 0.0, S.Hr[k], mpos ? S.Hi[k] : 0.0, mpos, sv)) sum += sv;
  }
  return sum;
}
// One sweep as id_sweep without the computeDev sum: true when one of this lane's pairs
// alone makes the sweep's computeDev exceed 1e-8 -- its term s (s >= 0 for every pair)
// exceeds c = 4 p^2 1e-8 with a relative margin of 1e-6, tested as top > c2 bot (c2 = c, or
// c/2 where the term carries the reference's factor 2) without the division.  The pair
// must also have bot > 4e-60, so that one of its old and new values is at least 1e-30: the
// reference skips a pair only when all of them are below 1e-30, and gives the term 1 (above
// c) when only one side is.  The sum of the nonnegative terms is then above the threshold
// whatever the summation order.  The sweep has no branches, so a wavefront issues the next
// sweep's updates while this one's test completes: pairs a lane does not hold (nn < 0) must
// be all zeros, which stay zeros and never pass the test.  m = 0 pairs keep zero imaginary
// parts by themselves (zero right-hand side and start), so no selects are needed here.
template <int KW>
__device__ __forceinline__ bool id_sweep_fast(IdState<KW> &S) {
  bool big = false;
  for (int k = 0; k < KW; k++) {
    const double ar = S.Hr[k], ai = S.Hi[k];
    double xr = S.fbr[k] + S.kH1[k] * S.Hr[k] + S.kF1[k] * S.Fr[k];
    double xi = S.fbi[k] + S.kH1[k] * S.Hi[k] + S.kF1[k] * S.Fi[k];
    S.Fr[k] = xr * S.kInv[k];
    S.Fi[k] = xi * S.kInv[k];
    xr = S.hbr[k] + S.kH2[k] * S.Hr[k] + S.kF2[k] * S.Fr[k];
    xi = S.hbi[k] + S.kH2[k] * S.Hi[k] + S.kF2[k] * S.Fi[k];
    const double br = xr * S.kIH[k], bi = xi * S.kIH[k];
    S.Hr[k] = br;
    S.Hi[k] = bi;
    const double dr = ar - br, di = ai - bi;
    const double top = dr * dr + di * di;
    const double bot = ar * ar + ai * ai + br * br + bi * bi;
    big = big | ((bot > 4e-60) & (top > S.c2[k] * bot));
  }
  return big;
}
// every field of pair k zero (a pair the lane does not hold)
template <int KW>
__device__ __forceinline__ void id_clear(IdState<KW> &S, int k) {
  S.Fr[k] = S.Fi[k] = S.Hr[k] = S.Hi[k] = 0.0;
  S.fbr[k] = S.fbi[k] = S.hbr[k] = S.hbi[k] = S.uHr[k] = S.uHi[k] = 0.0;
  S.kH1[k] = S.kF1[k] = S.kH2[k] = S.kF2[k] = S.kIH[k] = S.kInv[k] = S.c2[k] = 0.0;
  S.nn[k] = -1;
  S.i0[k] = 0;
}

// Sweeps until the summed computeDev of all NW wavefronts' pairs <= 1e-8 (or 100 sweeps),
// the reference's loop (expcenter.cpp:1097-1135).  The stopping test is evaluated in
// chunks of IC sweeps: each chunk runs IC sweeps from a saved state, exchanges the IC
// per-sweep partials with one barrier, and if the loop ends inside the chunk restores the
// state and reruns the sweeps up to the stopping one -- the same arithmetic in the same
// order as sweep-by-sweep testing.  A chunk first runs with id_sweep_fast: when every
// sweep of it whose test can stop the loop is shown to continue by a single pair, the
// exact sums are not needed; otherwise it is rerun from the saved state with them.
// red holds 2 * IC * NW doubles, then 2 * NW doubles the caller uses after the loop,
// then 2 * (NW + 1) ints of flags.  Returns the sweeps.
constexpr int IC = 10;  // divides the 100-sweep cap
template <int KW, int NW>
__device__ int id_iterate(IdState<KW> &S, int p, double *red) {
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  int *flags = reinterpret_cast<int *>(red + 2 * IC * NW + 2 * NW);
  const double cthr = 4.0 * p * p * 1e-8 * (1.0 + 1e-6);
  for (int k = 0; k < KW; k++)
    if (S.nn[k] >= 0) S.c2[k] = (S.i0[k] != S.nn[k] * S.nn[k]) ? 0.5 * cthr : cthr;
  int ct = 0;
  for (int chunk = 0;; chunk++) {
    double sFr[KW], sFi[KW], sHr[KW], sHi[KW];
    for (int k = 0; k < KW; k++) {
      sFr[k] = S.Fr[k]; sFi[k] = S.Fi[k]; sHr[k] = S.Hr[k]; sHi[k] = S.Hi[k];
    }
    {
      // fast pass: bit t of `mine` set when this lane shows sweep t continues
      unsigned mine = 0;
      for (int t = 0; t < IC; t++)
        if (id_sweep_fast(S)) mine |= 1u << t;
      unsigned wb = 0;
      for (int t = 0; t < IC; t++)
        if (__ballot((mine >> t) & 1u)) wb |= 1u << t;
      int *fb = flags + (chunk & 1) * (NW + 1);  // double-buffered with the sums below
      if (lane == 0) fb[w] = (int)wb;
      __syncthreads();
      unsigned all = 0;
      for (int x = 0; x < NW; x++) all |= (unsigned)fb[x];
      // sweeps whose test can stop the loop: done = ct + t + 1 < 100
      unsigned need = 0;
      for (int t = 0; t < IC; t++)
        if (ct + t + 1 < 100) need |= 1u << t;
      if ((all & need) == need) {
        ct += IC;
        if (ct >= 100) return ct;
        continue;
      }
      for (int k = 0; k < KW; k++) {
        S.Fr[k] = sFr[k]; S.Fi[k] = sFi[k]; S.Hr[k] = sHr[k]; S.Hi[k] = sHi[k];
      }
    }
    double part[IC];
    for (int t = 0; t < IC; t++) part[t] = id_sweep(S);
    for (int t = 0; t < IC; t++) part[t] = wave_sum_dpp(part[t]);
    double *rb = red + (chunk & 1) * IC * NW;  // double-buffered: one barrier per chunk
    if (lane == 0)
      for (int t = 0; t < IC; t++) rb[t * NW + w] = part[t];
    __syncthreads();
    int stop = -1;
    for (int t = 0; t < IC; t++) {
      double tot = 0.0;
      for (int x = 0; x < NW; x++) tot += rb[t * NW + x];
      const double dev = tot / (4 * p * p);
      const int done = ct + t + 1;
      if (!(dev > 1e-8 && done < 100)) {
        stop = t;
        break;
      }
    }
    if (stop < 0) {
      ct += IC;
      continue;
    }
    if (stop < IC - 1) {
      for (int k = 0; k < KW; k++) {
        S.Fr[k] = sFr[k]; S.Fi[k] = sFi[k]; S.Hr[k] = sHr[k]; S.Hi[k] = sHi[k];
      }
      for (int t = 0; t <= stop; t++) (void)id_sweep(S);
    }
    ct += stop + 1;
    return ct;
  }
}
// computeDev2 terms of this lane's pairs (expcenter.cpp:1336-1370)
template <int KW>
__device__ __forceinline__ void id_dev2(const IdState<KW> &S, double &td, double &ts) {
  td = 0.0;
  ts = 0.0;
  for (int k = 0; k < KW; k++) {
    if (S.nn[k] < 0) continue;
    const bool mpos = (S.i0[k] != S.nn[k] * S.nn[k]);
    const double w = mpos ? 2.0 : 1.0;
    td += w * ((S.uHr[k] - S.Hr[k]) * (S.uHr[k] - S.Hr[k]));
    ts += w * (S.uHr[k] * S.uHr[k] + S.Hr[k] * S.Hr[k]);
    if (mpos) {
      td += w * ((S.uHi[k] - S.Hi[k]) * (S.uHi[k] - S.Hi[k]));
      ts += w * (S.uHi[k] * S.uHi[k] + S.Hi[k] * S.Hi[k]);
    }
  }
}

// one right-hand side: the workgroup's ST/64 wavefronts share the sphere's pairs
template <int KW>
__device__ void solve_identity(const SArgs &A, int g, int lst0, int lst1) {
  constexpr int NW = ST / 64;
  __shared__ double red[2 * IC * NW + 2 * NW + NW + 1];
  const int p = A.p, P2 = p * p;
  const double *cs = A.consts + (size_t)g * 10 * p;
  const double *cLH1 = cs + 2 * p, *cLF1 = cs + 3 * p, *cLH2 = cs + 6 * p, *cLF2 = cs + 7 * p;
  const int npair = p * (p + 1) / 2;
  IdState<KW> S;
  for (int k = 0; k < KW; k++) {
    const int ep = threadIdx.x + ST * k;
    id_clear(S, k);
    if (ep >= npair) continue;
    int n, i;
    pair_nm(ep, n, i);
    S.nn[k] = n;
    S.i0[k] = i;
    id_consts(S, k, n, cs, p);
    const int cnt = (i == n * n) ? 1 : 2;
    double fb[2] = {0, 0}, hb[2] = {0, 0};
    for (int c = 0; c < cnt; c++) {
      const int ii = i + c;
      double lh = A.LS ? A.LS[(size_t)g * P2 + ii] : 0.0;
      double lf = 0.0;
      if (A.LS) lh = 0.0 + lh;
      for (int e2 = lst0; e2 < lst1; e2++) {
        lf += A.slotF[(size_t)A.lst[e2] * P2 + ii];
        lh += A.slotH[(size_t)A.lst[e2] * P2 + ii];
      }
      if (A.farH) {
        lh += A.farH[(size_t)g * P2 + ii];
        lf += A.farF[(size_t)g * P2 + ii];
      }
      A.LFS[(size_t)g * P2 + ii] = lf;
      A.LHS[(size_t)g * P2 + ii] = lh;
      fb[c] = A.Vfix[(size_t)g * P2 + ii] + cLH1[n] * lh + cLF1[n] * lf;
      hb[c] = A.Dfix[(size_t)g * P2 + ii] + cLH2[n] * lh + cLF2[n] * lf;
    }
    S.fbr[k] = fb[0];
    S.hbr[k] = hb[0];
    S.fbi[k] = (cnt == 2) ? fb[1] : 0.0;
    S.hbi[k] = (cnt == 2) ? hb[1] : 0.0;
    S.Fr[k] = A.F[(size_t)g * P2 + i];
    S.Hr[k] = A.H[(size_t)g * P2 + i];
    S.Fi[k] = (cnt == 2) ? A.F[(size_t)g * P2 + i + 1] : 0.0;
    S.Hi[k] = (cnt == 2) ? A.H[(size_t)g * P2 + i + 1] : 0.0;
    S.uHr[k] = S.Hr[k];
    S.uHi[k] = S.Hi[k];
  }
  const int ct = id_iterate<KW, NW>(S, p, red);
  double td, ts;
  id_dev2(S, td, ts);
  for (int k = 0; k < KW; k++) {
    if (S.nn[k] < 0) continue;
    const int i = S.i0[k];
    A.F[(size_t)g * P2 + i] = S.Fr[k];
    A.H[(size_t)g * P2 + i] = S.Hr[k];
    if (i != S.nn[k] * S.nn[k]) {
      A.F[(size_t)g * P2 + i + 1] = S.Fi[k];
      A.H[(size_t)g * P2 + i + 1] = S.Hi[k];
    }
  }
  td = wave_sum(td);
  ts = wave_sum(ts);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  double *r2 = red + 2 * IC * NW;
  if (lane == 0) {
    r2[w] = td;
    r2[NW + w] = ts;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double TD = 0.0, TS = 0.0;
    for (int x = 0; x < NW; x++) {
      TD += r2[x];
      TS += r2[NW + x];
    }
    A.dev[g] = TD / (4 * TS) / 1.0;
    if (A.sweeps) A.sweeps[g] = ct;
    if (A.work) {
      atomicAdd(&A.work[0], (unsigned long long)(ct * 12.0 * P2));
      atomicAdd(&A.work[1], (unsigned long long)(8.0 * 10 * P2));
    }
  }
}

__global__ void __launch_bounds__(ST) k_solve(SArgs A) {
  extern __shared__ double sh[];
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  const int e0 = A.sel ? A.sel[blockIdx.x] : blockIdx.x;
  const int g = A.sph[e0];
  if (!A.imat[g]) {  // uniform across the workgroup
    const int l0 = A.lst_off[e0], l1 = A.lst_off[e0 + 1];
    // KW pairs per lane: ST * KW >= p(p+1)/2
    if (p <= 22) solve_identity<1>(A, g, l0, l1);
    else if (p <= 31) solve_identity<2>(A, g, l0, l1);
    else solve_identity<3>(A, g, l0, l1);
    return;
  }
  double *F = sh, *H = F + P2, *X = H + P2, *red = X + P2;
  const double *cs = A.consts + (size_t)g * 10 * p;
  const double *cH1 = cs, *cF1 = cs + p, *cLH1 = cs + 2 * p, *cLF1 = cs + 3 * p, *cH2 = cs + 4 * p,
               *cF2 = cs + 5 * p, *cLH2 = cs + 6 * p, *cLF2 = cs + 7 * p, *cInvH = cs + 8 * p;
  const int l0 = A.lst_off[e0], l1 = A.lst_off[e0 + 1];
  const double *IM = A.imat[g];
  constexpr int KPT = 4;  // up to 1024 coefficients per sphere
  double fb[KPT], hb[KPT], oldH[KPT], outerH[KPT];
  int nn[KPT];
  for (int k = 0; k < KPT; k++) {
    const int i = tid + k * ST;
    nn[k] = -1;
    if (i >= P2) continue;
    int n = (int)sqrt((double)i);
    while (n * n > i) n--;
    while ((n + 1) * (n + 1) <= i) n++;
    nn[k] = n;
    double LH = A.LS ? A.LS[(size_t)g * P2 + i] : 0.0;
    double LF = 0.0;
    if (A.LS) LH = 0.0 + LH;
    for (int e = l0; e < l1; e++) {
      LF += A.slotF[(size_t)A.lst[e] * P2 + i];
      LH += A.slotH[(size_t)A.lst[e] * P2 + i];
    }
    if (A.farH) {
      LH += A.farH[(size_t)g * P2 + i];
      LF += A.farF[(size_t)g * P2 + i];
    }
    A.LFS[(size_t)g * P2 + i] = LF;
    A.LHS[(size_t)g * P2 + i] = LH;
    fb[k] = A.Vfix[(size_t)g * P2 + i] + cLH1[n] * LH + cLF1[n] * LF;
    hb[k] = A.Dfix[(size_t)g * P2 + i] + cLH2[n] * LH + cLF2[n] * LF;
    F[i] = A.F[(size_t)g * P2 + i];
    H[i] = A.H[(size_t)g * P2 + i];
    oldH[k] = H[i];
    outerH[k] = H[i];
  }
  __syncthreads();
  int ct = 0;
  double dev = 1e300;
  while (ct == 0 || (dev > 1e-8 && ct < 100)) {
    // F <- diag(1/(2n+1)) IMat (fbase + cH1 H + cF1 F)
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * ST;
      if (nn[k] < 0) continue;
      X[i] = fb[k] + cH1[nn[k]] * H[i] + cF1[nn[k]] * F[i];
    }
    __syncthreads();
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * ST;
      if (nn[k] < 0) continue;
      double y;
      if (IM) {
        y = 0.0;
        for (int c = 0; c < P2; c++) y += IM[i + (size_t)c * P2] * X[c];
      } else
        y = X[i];
      F[i] = y * (1 / (double)(2 * nn[k] + 1));
    }
    __syncthreads();
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * ST;
      if (nn[k] < 0) continue;
      X[i] = hb[k] + cH2[nn[k]] * H[i] + cF2[nn[k]] * F[i];
    }
    __syncthreads();
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * ST;
      if (nn[k] < 0) continue;
      double y;
      if (IM) {
        y = 0.0;
        for (int c = 0; c < P2; c++) y += IM[i + (size_t)c * P2] * X[c];
      } else
        y = X[i];
      H[i] = y * cInvH[nn[k]];
    }
    __syncthreads();
    // computeDev(oldH, H) over (n,m) pairs (expcenter.cpp:1270-1328); oldH staged in X
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * ST;
      if (nn[k] < 0) continue;
      X[i] = oldH[k];
    }
    __syncthreads();
    double sum = 0.0;
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * ST;
      if (nn[k] < 0) continue;
      const int n = nn[k], mm = i - n * n;
      if (mm > 0 && (mm % 2) == 0) continue;
      double ar = X[i], ai = (mm > 0) ? X[i + 1] : 0.0;
      double br = H[i], bi = (mm > 0) ? H[i + 1] : 0.0;
      double s;
      const double PL = 1e-30;
      if (fabs(ar) < PL && fabs(ai) < PL) {
        if (fabs(br) < PL && fabs(bi) < PL) continue;
        s = 1.0;
      } else if (fabs(br) < PL && fabs(bi) < PL) {
        s = 1.0;
      } else {
        double dr = ar - br, di = ai - bi;
        double top = dr * dr + di * di;
        double bot = ar * ar + ai * ai + br * br + bi * bi;
        s = (mm != 0) ? 2 * top / bot : top / bot;
      }
      sum += s;
    }
    sum = block_sum(sum, red);
    dev = sum / (4 * p * p);
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * ST;
      if (nn[k] < 0) continue;
      oldH[k] = H[i];
    }
    ct++;
    __syncthreads();
  }
  // computeDev2(outerOldH, H) (expcenter.cpp:1336-1370)
  double td = 0.0, ts = 0.0;
  for (int k = 0; k < KPT; k++) {
    const int i = tid + k * ST;
    if (nn[k] < 0) continue;
    const int mm = i - nn[k] * nn[k];
    double a = outerH[k], b = H[i];
    double w = (mm == 0) ? 1.0 : 2.0;
    td += w * ((a - b) * (a - b));
    ts += w * (a * a + b * b);
  }
  td = block_sum(td, red);
  ts = block_sum(ts, red);
  for (int k = 0; k < KPT; k++) {
    const int i = tid + k * ST;
    if (nn[k] < 0) continue;
    A.F[(size_t)g * P2 + i] = F[i];
    A.H[(size_t)g * P2 + i] = H[i];
  }
  if (tid == 0) {
    A.dev[g] = td / (4 * ts) / 1.0;
    if (A.sweeps) A.sweeps[g] = ct;
    if (A.work) {
      const double P4 = (double)P2 * P2;
      atomicAdd(&A.work[0], (unsigned long long)(ct * ((IM ? 4 * P4 : 0.0) + 12.0 * P2)));
      atomicAdd(&A.work[1], (unsigned long long)(ct * (IM ? 16 * P4 : 0.0) + 8.0 * 10 * P2));
    }
  }
}

// ----------------------------------------------------------------------------
// exposed surface charges (computeExposedSurfaceChargesFH expcenter.cpp:1420-1440)
// ----------------------------------------------------------------------------
struct QArgs {
  const int *sph;
  const double *xyz;            // exposed points, relative to the sphere centre
  const int *pt_off, *pt_cnt, *nspx;
  const double *consts;         // [sphere][10][p] (cChgH at row 9)
  const double *F, *H;
  double *qF, *qH, *TQF, *TQH;
  int p;
  unsigned long long *work;  // instrumentation: [flops, bytes] (null: off)
  double *part;              // [workgroup][2] partial sums of |qF|, |qH|
  int nchunk;                // workgroups per sphere (ST / lanes points each)
  int lanes;                 // lanes per point (1, 2, 4, 8 or 16; 0 means 1)
};
// One workgroup per (sphere, chunk of ST/lanes points); `lanes` adjacent threads per point
// share its harmonic columns m (lane l takes m = kL + l for even k and kL + L-1-l for odd
// k, which balances the columns' lengths p - m) and add their sums with lane shuffles.
// The totals TQF/TQH are finished by k_surface_totals from the per-workgroup partial sums.
__global__ void __launch_bounds__(ST) k_surface(Tabs tb, QArgs A) {
  extern __shared__ double sh[];
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  const int L = A.lanes > 1 ? A.lanes : 1, lane = tid % L, PW = ST / L;
  const int ch = blockIdx.x % A.nchunk;
  const int g = A.sph[blockIdx.x / A.nchunk];
  double *Fs = sh, *Hs = Fs + P2, *red = Hs + P2;
  const double *cChgH = A.consts + (size_t)g * 10 * p + 9 * p;
  const int off = A.pt_off[g], cnt = A.pt_cnt[g];
  if (ch * PW >= cnt) {  // uniform across the workgroup
    if (tid == 0) {
      A.part[2 * blockIdx.x] = 0.0;
      A.part[2 * blockIdx.x + 1] = 0.0;
    }
    return;
  }
  for (int i = tid; i < P2; i += ST) {
    Fs[i] = A.F[(size_t)g * P2 + i];
    Hs[i] = A.H[(size_t)g * P2 + i];
  }
  __syncthreads();
  const double dA = 4 * M_PI / double(A.nspx[g]);
  double tF = 0.0, tH = 0.0;
  const int h = ch * PW + tid / L;
  if (h < cnt) {  // uniform across each point's lanes
    const double *x = A.xyz + 3 * (size_t)(off + h);
    double rho, th, ph;
    cart_to_sph(x[0], x[1], x[2], rho, th, ph);
    const double cost = cos(th), sint = sin(th);
    double sF = 0.0, sH = 0.0;
    for (int k = 0; k * L < p; k++) {
      const int m = k * L + ((k & 1) ? L - 1 - lane : lane);
      if (m >= p) continue;
      sh_column(tb, p, m, cost, sint, ph, [&](int n, double re, double im) {
        const double c2 = (2 * n + 1) / (4.0 * M_PI), chH = cChgH[n];
        if (m == 0) {
          sF += Fs[n * n] * (re * c2);
          sH += Hs[n * n] * (re * chH);
        } else {
          const int i = n * n + 2 * m - 1;
          sF += 2.0 * (Fs[i] * (re * c2) + Fs[i + 1] * (im * c2));
          sH += 2.0 * (Hs[i] * (re * chH) + Hs[i + 1] * (im * chH));
        }
      });
    }
    for (int o = 1; o < L; o <<= 1) {  // the point's lanes are adjacent and all active
      sF += __shfl_xor(sF, o, 64);
      sH += __shfl_xor(sH, o, 64);
    }
    if (lane == 0) {
      const double qf = sF * dA, qh = sH * dA;
      A.qF[off + h] = qf;
      A.qH[off + h] = qh;
      tF = fabs(qf);
      tH = fabs(qh);
    }
  }
  tF = block_sum(tF, red);
  tH = block_sum(tH, red);
  if (tid == 0) {
    A.part[2 * blockIdx.x] = tF;
    A.part[2 * blockIdx.x + 1] = tH;
    if (A.work && ch == 0) {
      atomicAdd(&A.work[0], (unsigned long long)(work_surface_point(p) * cnt));
      atomicAdd(&A.work[1], (unsigned long long)(40 * (size_t)cnt + 16 * P2));
    }
  }
}
// TQF/TQH of each listed sphere: its chunks' partial sums in chunk order
__global__ void __launch_bounds__(64) k_surface_totals(const int *sph, int n, int nchunk, const double *part,
                                                       double *TQF, double *TQH) {
  const int e = blockIdx.x * 64 + threadIdx.x;
  if (e >= n) return;
  double f = 0.0, h = 0.0;
  const double *pe = part + 2 * (size_t)e * nchunk;
  int c = 0;
  for (; c + 8 <= nchunk; c += 8) {  // eight chunks' loads in flight, added in chunk order
    double vf[8], vh[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      vf[u] = pe[2 * (c + u)];
      vh[u] = pe[2 * (c + u) + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      f += vf[u];
      h += vh[u];
    }
  }
  for (; c < nchunk; c++) {
    f += pe[2 * c];
    h += pe[2 * c + 1];
  }
  TQF[sph[e]] = f;
  TQH[sph[e]] = h;
}

// max over listed spheres of dev; also returns the sum
__global__ void __launch_bounds__(256) k_maxdev(const int *sph, int n, const double *dev, double *out) {
  __shared__ double red[8];
  double mx = 0.0, sm = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    double d = dev[sph[i]];
    mx = fmax(mx, d);
    sm += d;
  }
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, o, 64));
    sm += __shfl_xor(sm, o, 64);
  }
  if (threadIdx.x % 64 == 0) {
    red[threadIdx.x / 64] = mx;
    red[4 + threadIdx.x / 64] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    out[1] = red[4] + red[5] + red[6] + red[7];
  }
}

// inter-molecular local expansion: LS = R^T (sum of slots in list order)
struct LArgs {
  const int *sph;        // target sphere per workgroup
  const int *lst_off, *lst;
  const double *slot;
  const double *base;    // [sphere] starting value, or null (zero)
  double *out;           // [sphere][p^2]
  const int *sel;        // entries of sph/lst_off to run (null: workgroup index)
  int p;
  int bs;                // doubles per slot and per output (0: p^2)
  const int *obs;        // output block of each entry (null: its sph index)
};
// grid (entries, element chunks of 256)
__global__ void __launch_bounds__(256) k_sumslots(LArgs A) {
  const int e0 = A.sel ? A.sel[blockIdx.x] : blockIdx.x;
  const int p = A.p, P2 = A.bs ? A.bs : p * p, g = A.sph[e0];
  const int l0 = A.lst_off[e0], l1 = A.lst_off[e0 + 1];
  const int go = A.obs ? A.obs[e0] : g;
  for (int i = blockIdx.y * 256 + threadIdx.x; i < P2; i += 256 * gridDim.y) {
    double s = A.base ? A.base[(size_t)g * P2 + i] : 0.0;
    int e = l0;
    // eight loads in flight, added in list order
    for (; e + 8 <= l1; e += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = A.slot[(size_t)A.lst[e + u] * P2 + i];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; e < l1; e++) s += A.slot[(size_t)A.lst[e] * P2 + i];
    A.out[(size_t)go * P2 + i] = s;
  }
}

}  // namespace pbsam

#include "pbsam_kernels_grad.hip"
