// Kernels of the adjoint (transposed) mutual-polarization fixed point.
//
// The mutual solve's linear operator is built from per-pair re-expansions X(target <- source),
// molecule rotations R, per-sphere inner solves and the numerical re-expansion of surface
// charges.  With W = diag(1 for m = 0, 2 for the real and imaginary parts of m > 0) -- the
// weights of the reference's block inner product (expansion.h:706-741) -- the transposes of the
// pair operators are again pair operators:
//   X(b <- a)^T = W X(a <- b) W^-1     (screened and unscreened, any orders, any radii)
//   R_fwd^T     = W R_back W^-1
// so the adjoint solve, written in the W-dual variable lt = W^-1 l, runs the forward kernels on
// the same job lists (every job list holds both directions of a pair).  Only the per-level inner
// solve, the numerical re-expansion and the surface charges need the kernels below.
namespace pbsam {

__device__ __forceinline__ int coef_level(int c) {
  int n = (int)sqrt((double)c);
  while (n * n > c) n--;
  while ((n + 1) * (n + 1) <= c) n++;
  return n;
}

// Transposed inner solve of each listed sphere.  With the identity surface matrix the inner
// iteration of solveFH (expcenter.cpp:1080-1135) converges, per level n, to
//   A_n (F, H) = B_n (LF, LH) + fixed terms,
//   A_n = [[1 - c3 cF1, -c3 cH1], [-cInvH cF2, 1 - cInvH cH2]],
//   B_n = [[c3 cLF1, c3 cLH1], [cInvH cLF2, cInvH cLH2]],  c3 = 1/(2n+1),
// so (wLF, wLH) = B_n^T A_n^-T (aF, aH) coefficient by coefficient (W is per coefficient and
// commutes).  Grid (spheres, chunks of 256 coefficients).
__global__ void __launch_bounds__(256) k_adj_local(const int *pol, const double *consts, const double *aF,
                                                   const double *aH, double *wF, double *wH, int p) {
  const int P2 = p * p, c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const int g = pol[blockIdx.x], n = coef_level(c);
  const double *k = consts + (size_t)g * 10 * p;
  const double c3 = 1.0 / double(2 * n + 1);
  const double cH1 = k[n], cF1 = k[p + n], cLH1 = k[2 * p + n], cLF1 = k[3 * p + n];
  const double cH2 = k[4 * p + n], cF2 = k[5 * p + n], cLH2 = k[6 * p + n], cLF2 = k[7 * p + n];
  const double cInvH = k[8 * p + n];
  const double a11 = 1.0 - c3 * cF1, a12 = -c3 * cH1, a21 = -cInvH * cF2, a22 = 1.0 - cInvH * cH2;
  const double det = a11 * a22 - a12 * a21;
  const size_t i = (size_t)g * P2 + c;
  const double lf = aF[i], lh = aH[i];
  const double y1 = (a22 * lf - a21 * lh) / det, y2 = (a11 * lh - a12 * lf) / det;
  wF[i] = c3 * cLF1 * y1 + cInvH * cLF2 * y2;
  wH[i] = c3 * cLH1 * y1 + cInvH * cLH2 * y2;
}

// Zero the output slots of the numerical re-expansion jobs (the adjoint's numerical part does
// not go through them).  Grid (jobs, chunks of 256 coefficients).
__global__ void __launch_bounds__(256) k_adj_zero_slots(const PJob *jobs, double *slotH, double *slotF, int P2) {
  const int c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const size_t o = (size_t)jobs[blockIdx.x].slot * P2 + c;
  slotH[o] = 0.0;
  slotF[o] = 0.0;
}

// Zero the listed output slots (H and F).  Grid (slots, chunks of 256 coefficients).
__global__ void __launch_bounds__(256) k_adj_zero_slotlist(const int *slots, double *slotH, double *slotF, int P2) {
  const int c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const size_t o = (size_t)slots[blockIdx.x] * P2 + c;
  slotH[o] = 0.0;
  slotF[o] = 0.0;
}

// Transposed numerical re-expansion (CXFormNIntra::xformFH, xforms.cpp:528-550).  The forward
// job adds to the target's LF/LH the local expansions of the source's surface charges; its
// transpose gives, at every exposed point h of the source, the weighted contraction
//   qT(h) = sum_c N_c(h) W_c wt_c,
// N_c(h) the local-expansion coefficient of a unit charge at h about the target (unscreened for
// the F set, screened for H), i.e. the potential of the target's adjoint block at h.  One
// workgroup per (job, chunk of ST/L points); L adjacent lanes per point share its harmonic
// columns as k_surface does, and add their sums with lane shuffles.
struct TArgs {
  const PJob *jobs;
  const int *tgt;          // target sphere of each job
  const double *xyz;       // exposed points, relative to their sphere's centre
  const double *wF, *wH;   // [sphere][p^2] W-dual adjoint blocks of the targets
  const long long *qoff;   // first output element of each job
  double *outF, *outH;     // [job points]
  int p, lanes, nchunk;
};
__global__ void __launch_bounds__(ST) k_p2lT(Tabs tb, TArgs A) {
  extern __shared__ double sh[];
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  const int L = A.lanes, lane = tid % L, PW = ST / L;
  const int jb = blockIdx.x / A.nchunk, ch = blockIdx.x % A.nchunk;
  const PJob J = A.jobs[jb];
  if (ch * PW >= J.npts) return;  // uniform across the workgroup
  double *Ws = sh, *Wh = Ws + P2, *rad = Wh + P2;  // rad: [PW][F set p | H set p]
  const int g = A.tgt[jb];
  for (int i = tid; i < P2; i += ST) {
    Ws[i] = A.wF[(size_t)g * P2 + i];
    Wh[i] = A.wH[(size_t)g * P2 + i];
  }
  const int pl = tid / L, h = ch * PW + pl;
  const bool act = h < J.npts;  // uniform across each point's lanes
  double cost = 0.0, sint = 0.0, ph = 0.0;
  if (act) {
    const double *x = A.xyz + 3 * (size_t)(J.pt_off + h);
    double rho, th;
    cart_to_sph(x[0] - J.P[0], x[1] - J.P[1], x[2] - J.P[2], rho, th, ph);
    cost = cos(th);
    sint = sin(th);
    if (lane == 0) {  // radial factors of a unit charge (CLExpan, expansion.cpp:289-316,422-436)
      double *rf = rad + (size_t)pl * 2 * p, *rh = rf + p;
      const double val = rho * tb.kappa, vs = val * val, mr = J.a_tgt / rho;
      double mkF = 1.0 / rho, mkH = exp(-val) / rho;
      double km2 = 1.0, km1 = 1.0 + val;
      for (int n = 0; n < p; n++) {
        double K;
        if (val == 0.0 || n == 0) K = 1.0;
        else if (n == 1) K = km1;
        else {
          K = bessel_k_next(tb, n, km1, km2, vs);
          km2 = km1;
          km1 = K;
        }
        rf[n] = mkF;
        rh[n] = K * mkH;
        mkF *= mr;
        mkH *= mr;
      }
    }
  }
  __syncthreads();
  if (!act) return;
  const double *rf = rad + (size_t)pl * 2 * p, *rh = rf + p;
  double sF = 0.0, sH = 0.0;
  for (int k = 0; k * L < p; k++) {
    const int m = k * L + ((k & 1) ? L - 1 - lane : lane);
    if (m >= p) continue;
    sh_column(tb, p, m, cost, sint, ph, [&](int n, double re, double im) {
      if (m == 0) {
        sF += rf[n] * (re * Ws[n * n]);
        sH += rh[n] * (re * Wh[n * n]);
      } else {
        const int i = n * n + 2 * m - 1;
        sF += 2.0 * rf[n] * (re * Ws[i] + im * Ws[i + 1]);
        sH += 2.0 * rh[n] * (re * Wh[i] + im * Wh[i + 1]);
      }
    });
  }
  for (int o = 1; o < L; o <<= 1) {  // the point's lanes are adjacent and all active
    sF += __shfl_xor(sF, o, 64);
    sH += __shfl_xor(sH, o, 64);
  }
  if (lane == 0) {
    A.outF[A.qoff[jb] + h] = sF;
    A.outH[A.qoff[jb] + h] = sH;
  }
}

// Products of stored point matrices with two charge sets: out1[slot][c] = sum_h M1[h][c] q1[h],
// out2[slot][c] = sum_h M2[h][c] q2[h] over the job's points h (point-major npts x p^2
// matrices, charges at pt_off + h), points added in order.  Used for the numerical
// re-expansion with a molecule type's stored maps (M1 the F map, M2 the H map) and for the
// harmonic sums of a sphere's adjoint surface charges (M1 = M2 = the stored harmonics of its
// exposed points).  Grid (jobs, chunks of 256 coefficients); each thread streams one column.
struct GVJob {
  const double *M1, *M2;
  int npts, pt_off, slot, pad;
};
// Split launches (gridDim.z > 1, for few jobs): workgroup z takes points [z spl, (z+1) spl)
// and writes its sums to part[((job nz + z) 2 + set) P2 + c]; k_mapgemv_sum adds them in z order.
// When M1 == M2 (a sphere's harmonics) each matrix element is loaded once.
__global__ void __launch_bounds__(256) k_mapgemv(const GVJob *jobs, const double *q1, const double *q2, double *out1,
                                                 double *out2, int P2, int spl, double *part) {
  const GVJob J = jobs[blockIdx.x];
  const int c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const int nz = gridDim.z, lo = nz > 1 ? min(J.npts, (int)blockIdx.z * spl) : 0;
  const int hi = nz > 1 ? min(J.npts, lo + spl) : J.npts;
  const double *m1 = J.M1 + c, *m2 = J.M2 + c, *a = q1 + J.pt_off, *b = q2 + J.pt_off;
  double s1 = 0.0, s2 = 0.0;
  int h = lo;
  if (J.M1 == J.M2) {
    for (; h + 8 <= hi; h += 8) {  // eight loads in flight, added in point order
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) x[u] = m1[(size_t)(h + u) * P2];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        s1 += x[u] * a[h + u];
        s2 += x[u] * b[h + u];
      }
    }
  } else {
    for (; h + 4 <= hi; h += 4) {  // eight loads in flight, added in point order
      double x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        x[u] = m1[(size_t)(h + u) * P2];
        y[u] = m2[(size_t)(h + u) * P2];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        s1 += x[u] * a[h + u];
        s2 += y[u] * b[h + u];
      }
    }
  }
  for (; h < hi; h++) {
    s1 += m1[(size_t)h * P2] * a[h];
    s2 += m2[(size_t)h * P2] * b[h];
  }
  if (nz > 1) {
    double *o = part + ((size_t)blockIdx.x * nz + blockIdx.z) * 2 * P2 + c;
    o[0] = s1;
    o[P2] = s2;
  } else {
    out1[(size_t)J.slot * P2 + c] = s1;
    out2[(size_t)J.slot * P2 + c] = s2;
  }
}
// the split sums of k_mapgemv added in range order.  Grid (jobs, chunks of 256 coefficients).
__global__ void __launch_bounds__(256) k_mapgemv_sum(const GVJob *jobs, int nz, const double *part, double *out1,
                                                     double *out2, int P2) {
  const int c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const double *o = part + (size_t)blockIdx.x * nz * 2 * P2 + c;
  double s1 = 0.0, s2 = 0.0;
  for (int z = 0; z < nz; z++) {
    s1 += o[(size_t)z * 2 * P2];
    s2 += o[(size_t)z * 2 * P2 + P2];
  }
  const int slot = jobs[blockIdx.x].slot;
  out1[(size_t)slot * P2 + c] = s1;
  out2[(size_t)slot * P2 + c] = s2;
}

// Adjoint surface charges of each listed sphere: at each of its exposed points, the sum over
// the numerical jobs that read that sphere's charges, in job order.  Grid (spheres, chunks of
// 256 points).
__global__ void __launch_bounds__(256) k_adj_qsum(const int *pol, const int *joff, const int *jlist,
                                                  const long long *qoff, const double *qTF, const double *qTH,
                                                  const int *pt_off, const int *pt_cnt, double *lqF, double *lqH) {
  const int e = blockIdx.x, m = pol[e], h = blockIdx.y * 256 + threadIdx.x;
  if (h >= pt_cnt[m]) return;
  double f = 0.0, q = 0.0;
  for (int k = joff[e]; k < joff[e + 1]; k++) {
    const long long o = qoff[jlist[k]] + h;
    f += qTF[o];
    q += qTH[o];
  }
  lqF[pt_off[m] + h] = f;
  lqH[pt_off[m] + h] = q;
}

// The adjoint operator applied to x, as y = x - A~ x, for each listed sphere (one workgroup
// each), x and y compact [entry][F, H][p^2]:
//   (A~ x)_H = (inter part) + (analytic intra slots) + dA cChgH_n sT_H
//   (A~ x)_F =                (analytic intra slots) + dA (2n+1)/(4 pi) sT_F
// (sT: the harmonic sums of the adjoint surface charges, computeExposedSurfaceChargesFH
// expcenter.cpp:1420-1440 transposed).
struct CArgs {
  const int *pol;
  const int *lst_off, *lst;   // analytic and numerical intra slots of each listed sphere
  const double *slotH, *slotF;
  const double *inter;        // [sphere][p^2] inter-molecular part
  const double *sTF, *sTH;    // [list entry][p^2]
  const double *consts;
  const int *nspx;
  const double *x;            // [entry][2][p^2]
  double *y;                  // [entry][2][p^2]
  int p;
};
__global__ void __launch_bounds__(256) k_adj_apply(CArgs A) {
  const int e = blockIdx.x, g = A.pol[e], p = A.p, P2 = p * p;
  const double dA = 4 * M_PI / double(A.nspx[g]);
  const double *cChgH = A.consts + (size_t)g * 10 * p + 9 * p;
  for (int c = threadIdx.x; c < P2; c += 256) {
    const int n = coef_level(c);
    double h = A.inter[(size_t)g * P2 + c], f = 0.0;
    for (int k = A.lst_off[e]; k < A.lst_off[e + 1]; k++) {
      const size_t o = (size_t)A.lst[k] * P2 + c;
      h += A.slotH[o];
      f += A.slotF[o];
    }
    h += dA * cChgH[n] * A.sTH[(size_t)e * P2 + c];
    f += dA * ((2 * n + 1) / (4.0 * M_PI)) * A.sTF[(size_t)e * P2 + c];
    const size_t i = (size_t)e * 2 * P2 + c;
    A.y[i] = A.x[i] - f;
    A.y[i + P2] = A.x[i + P2] - h;
  }
}

// The exact-mode fixed-point map of each listed sphere (one workgroup each): its local
// expansions LH = LS + (intra slots) [+ far], LF = (intra slots) [+ far] and the inner solve's
// limit per level (solveFH, expcenter.cpp:1080-1135, with the identity surface matrix)
//   (F, H) = A_n^-1 [B_n (LF, LH) + (c3 Vfix, cInvH Dfix)]
// (A_n, B_n as in k_adj_local; the fixed terms only for the right-hand side).  With x given,
// y = x - (F, H) (the operator I - A of the linear part), else y = (F, H); x, y compact.
struct EArgs {
  const int *pol;
  const int *lst_off, *lst;
  const double *slotH, *slotF;
  const double *inter;          // [sphere][p^2] LS
  const double *farH, *farF;    // [sphere][p^2] frozen far-pair expansions, or null
  const double *Vfix, *Dfix;    // [sphere][p^2] fixed-charge terms, or null
  const double *consts;
  const double *x;              // [entry][2][p^2] or null
  double *y;                    // [entry][2][p^2]
  int p;
};
__global__ void __launch_bounds__(256) k_exact_apply(EArgs A) {
  const int e = blockIdx.x, g = A.pol[e], p = A.p, P2 = p * p;
  const double *k = A.consts + (size_t)g * 10 * p;
  for (int c = threadIdx.x; c < P2; c += 256) {
    const int n = coef_level(c);
    const size_t gi = (size_t)g * P2 + c;
    double lh = A.inter[gi], lf = 0.0;
    for (int s = A.lst_off[e]; s < A.lst_off[e + 1]; s++) {
      const size_t o = (size_t)A.lst[s] * P2 + c;
      lh += A.slotH[o];
      lf += A.slotF[o];
    }
    if (A.farH) {
      lh += A.farH[gi];
      lf += A.farF[gi];
    }
    const double c3 = 1.0 / double(2 * n + 1);
    const double cH1 = k[n], cF1 = k[p + n], cLH1 = k[2 * p + n], cLF1 = k[3 * p + n];
    const double cH2 = k[4 * p + n], cF2 = k[5 * p + n], cLH2 = k[6 * p + n], cLF2 = k[7 * p + n];
    const double cInvH = k[8 * p + n];
    const double a11 = 1.0 - c3 * cF1, a12 = -c3 * cH1, a21 = -cInvH * cF2, a22 = 1.0 - cInvH * cH2;
    const double det = a11 * a22 - a12 * a21;
    const double r1 = c3 * ((A.Vfix ? A.Vfix[gi] : 0.0) + cLH1 * lh + cLF1 * lf);
    const double r2 = cInvH * ((A.Dfix ? A.Dfix[gi] : 0.0) + cLH2 * lh + cLF2 * lf);
    const double F = (a22 * r1 - a12 * r2) / det, H = (a11 * r2 - a21 * r1) / det;
    const size_t i = (size_t)e * 2 * P2 + c;
    A.y[i] = A.x ? A.x[i] - F : F;
    A.y[i + P2] = A.x ? A.x[i + P2] - H : H;
  }
}

// compact [entry][F, H][p^2] -> the sphere-indexed blocks the re-expansion jobs read
__global__ void __launch_bounds__(256) k_adj_scatter(const int *pol, const double *x, double *aF, double *aH,
                                                     int P2) {
  const int e = blockIdx.x, c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const size_t g = (size_t)pol[e] * P2 + c, i = (size_t)e * 2 * P2 + c;
  aF[g] = x[i];
  aH[g] = x[i + P2];
}
// sphere-indexed F, H blocks of the listed spheres -> compact [entry][F, H][p^2]
__global__ void __launch_bounds__(256) k_adj_gather(const int *pol, const double *F, const double *H, double *x,
                                                    int P2) {
  const int e = blockIdx.x, c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const size_t g = (size_t)pol[e] * P2 + c, i = (size_t)e * 2 * P2 + c;
  x[i] = F[g];
  x[i + P2] = H[g];
}
// the right-hand side in compact form: F part zero; H part of entry e the whole field LS at its
// sphere (mode 0), the part of it from one molecule (mode 1, in part), or zero (mode 2: a group
// without a solve in this pass)
__global__ void __launch_bounds__(256) k_adj_rhs(const int *pol, const int *mode, const double *LS,
                                                 const double *part, double *b, int P2) {
  const int e = blockIdx.x, c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const size_t i = (size_t)e * 2 * P2 + c, g = (size_t)pol[e] * P2 + c;
  const int m = mode[e];
  b[i] = 0.0;
  b[i + P2] = m == 0 ? LS[g] : (m == 1 ? part[g] : 0.0);
}
// H of the spheres flagged in sel, zero elsewhere (sphere-indexed blocks)
__global__ void __launch_bounds__(256) k_adj_srcmask(const int *sel, const double *H, double *out, int P2) {
  const int g = blockIdx.x, c = blockIdx.y * 256 + threadIdx.x;
  if (c >= P2) return;
  const size_t i = (size_t)g * P2 + c;
  out[i] = sel[g] ? H[i] : 0.0;
}

// Krylov vector kernels of the restarted GMRES (columns of V are consecutive vectors of N).
// part[k][b] = sum over block b's range of V_k . w (KB blocks, ranges in order); the host adds
// the blocks in order, so results do not vary run to run.  A non-finite value is passed on.
constexpr int KB = 256;
__global__ void __launch_bounds__(256) k_kdots(const double *V, int ncol, long long N, const double *w,
                                               double *part) {
  __shared__ double red[4];
  const int b = blockIdx.x, k = blockIdx.y;
  const long long per = (N + KB - 1) / KB, lo = b * per, hi = min(N, lo + per);
  const double *v = V + (size_t)k * N;
  double s = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += 256) s += v[i] * w[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[(size_t)k * KB + b] = s;
}
// w = alpha w - sum_k h[k] V_k (h on the device, ncol <= 64)
__global__ void __launch_bounds__(256) k_kaxpy(const double *V, int ncol, long long N, const double *h, double alpha,
                                               double *w) {
  __shared__ double hs[64];
  if (threadIdx.x < ncol) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  double s = alpha * w[i];
  for (int k = 0; k < ncol; k++) s -= hs[k] * V[(size_t)k * N + i];
  w[i] = s;
}
// out = w * s
__global__ void __launch_bounds__(256) k_kscale(long long N, const double *w, double s, double *out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < N) out[i] = w[i] * s;
}

// out[t][a] = <zeta_g + own_t Hrot_g, src_t[a]>_W for target t (sphere g = sph[t]), a = 0..2:
// the adjoint's contraction with the three source planes of one rigid coordinate set.
__global__ void __launch_bounds__(256) k_adj_contract(const int *sph, const int *own, const double *zeta,
                                                      const double *Hrot, const double *src, double *out, int p) {
  __shared__ double red[4];
  const int t = blockIdx.x, g = sph[t], P2 = p * p;
  const bool mine = own[t] != 0;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int c = threadIdx.x; c < P2; c += 256) {
    const int n = coef_level(c);
    double v = zeta[(size_t)g * P2 + c];
    if (mine) v += Hrot[(size_t)g * P2 + c];
    if (c != n * n) v *= 2.0;
    const double *s = src + (size_t)t * 3 * P2 + c;
    s0 += v * s[0];
    s1 += v * s[P2];
    s2 += v * s[2 * P2];
  }
  s0 = block_sum(s0, red);
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (threadIdx.x == 0) {
    out[3 * t] = s0;
    out[3 * t + 1] = s1;
    out[3 * t + 2] = s2;
  }
}

}  // namespace pbsam
