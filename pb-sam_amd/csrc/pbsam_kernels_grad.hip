// HIP kernels of the gradient (force) phase of the PB-SAM mutual-polarization path
// (gfx950, FP64); included by pbsam_kernels.hip.
//
// Kernel map (reference functions restated, paths relative to the reference src/):
//   k_xgrad          CXFormA::xformH(CExpan, CGradExpan) + sphToCart  xforms.cpp:86-115,165-192
//                    (CRotCoeff::computeGradCoeff rotcoeff.cpp:818-850, dRotateT/P(+Sing)
//                     :450-771, CTransCoeff dT :372-395) -- fused per pair
//   k_solve3         CSolExpCenter::solveFH with 3 right-hand sides   expcenter.cpp:1080-1370
//   k_surfgemm       computeExposedSurfaceGradientFH (stored harmonics) expcenter.cpp:1378-1412
//   k_p2lgemv        CXFormNIntra::xformGFH (stored maps)             xforms.cpp:514-582
//   k_force          computeForceOn_0 + the per-sphere force terms    expcenter.cpp:124-149,
//                                                                     moldynamics.cpp:525-550
namespace pbsam {

__device__ __forceinline__ double QC(const Tabs &t, int n, int m, int k) { return t.Q[(n * PMAX + m) * 2 + k]; }
__device__ __forceinline__ cd derv(cd c, int s) { return {-s * c.y, s * c.x}; }  // i*s*c
__device__ __forceinline__ cd blk_c(const double *B, int n, int m) {             // (n, m>=0) entry
  return (m == 0) ? cd{B[n * n], 0.0} : cd{B[n * n + 2 * m - 1], B[n * n + 2 * m]};
}
__device__ __forceinline__ void blk_put(double *B, int n, int m, cd v) {
  if (m == 0) B[n * n] = v.x * 1.0;
  else {
    B[n * n + 2 * m - 1] = v.x;
    B[n * n + 2 * m] = v.y;
  }
}

// ---- singular orientations (sin(theta) < 1e-12): closed forms on whole blocks in LDS ----
// rotate / rotateWithXi with xi = pi at theta = 0 (identity) or pi (parity), rotcoeff.cpp:368-393
__device__ void sing_rot(const double *in, double *out, int q, int sing, int tid, int nt) {
  for (int n = tid; n < q; n += nt) {
    if (sing == 1) {
      for (int mm = 0; mm < 2 * n + 1; mm++) out[n * n + mm] = in[n * n + mm];
    } else {
      out[n * n] = (n % 2 == 0) ? in[n * n] : -in[n * n];
      int s = (n % 2 == 0) ? 1 : -1;
      for (int mm = 1; mm < 2 * n + 1; mm++, s = -s) out[n * n + mm] = s * in[n * n + mm];
    }
  }
}
// dRotateTSing rotcoeff.cpp:610-700 (sign flipped for the transposed direction)
__device__ void sing_dT(const Tabs &tb, const double *in, double *out, int q, int sing, bool fwd, int tid, int nt) {
  const double f = fwd ? 1.0 : -1.0;
  for (int n = tid; n < q; n += nt) {
    if (n == 0) {
      out[0] = 0.0;
      continue;
    }
    if (sing == 1) {
      out[n * n] = f * (2.0 * QC(tb, n, 0, 1) * in[n * n + 1] * 1.0);
      for (int m = 1; m < n; m++) {
        cd i0 = (m == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, m - 1);
        cd i1 = blk_c(in, n, m + 1);
        cd o = cadd(cscale(QC(tb, n, m, 0), i0), cscale(QC(tb, n, m, 1), i1));
        out[n * n + 2 * m - 1] = f * o.x;
        out[n * n + 2 * m] = f * o.y;
      }
      cd i0 = (n == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, n - 1);
      cd o = cscale(QC(tb, n, n, 0), i0);
      out[n * n + 2 * n - 1] = f * o.x;
      out[n * n + 2 * n] = f * o.y;
    } else {
      const double s = (n % 2 == 1) ? -1.0 : 1.0;
      out[n * n] = f * (2.0 * s * QC(tb, n, 0, 1) * in[n * n + 1] * 1.0);
      for (int m = 1; m < n; m++) {
        cd i0 = (m == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, m - 1);
        cd i1 = blk_c(in, n, m + 1);
        cd o = cscale(s, cadd(cscale(QC(tb, n, m, 0), cconj(i0)), cscale(QC(tb, n, m, 1), cconj(i1))));
        out[n * n + 2 * m - 1] = f * o.x;
        out[n * n + 2 * m] = f * o.y;
      }
      cd i0 = (n == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, n - 1);
      cd o = cscale(s * QC(tb, n, n, 0), cconj(i0));
      out[n * n + 2 * n - 1] = f * o.x;
      out[n * n + 2 * n] = f * o.y;
    }
  }
}
// dRotatePSing rotcoeff.cpp:704-771 (sign flipped for the transposed direction at theta = 0)
__device__ void sing_dP(const Tabs &tb, const double *in, double *out, int q, int sing, bool fwd, int tid, int nt) {
  const double f = (!fwd && sing == 1) ? -1.0 : 1.0;
  for (int n = tid; n < q; n += nt) {
    if (n == 0) {
      out[0] = 0.0;
      continue;
    }
    if (sing == 1) {
      out[n * n] = f * (2.0 * QC(tb, n, 0, 1) * in[n * n + 2] * 1.0);
      for (int m = 1; m < n; m++) {
        cd i0 = (m == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, m - 1);
        cd i1 = blk_c(in, n, m + 1);
        cd a = cmul(cd{0.0, QC(tb, n, m, 0)}, i0), b = cmul(cd{0.0, QC(tb, n, m, 1)}, i1);
        out[n * n + 2 * m - 1] = f * (a.x - b.x);
        out[n * n + 2 * m] = f * (a.y - b.y);
      }
      cd i0 = (n == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, n - 1);
      cd o = cmul(cd{0.0, QC(tb, n, n, 0)}, i0);
      out[n * n + 2 * n - 1] = f * o.x;
      out[n * n + 2 * n] = f * o.y;
    } else {
      const double s = (n % 2 == 1) ? 1.0 : -1.0;
      out[n * n] = f * (s * 2 * QC(tb, n, 0, 1) * in[n * n + 2] * 1.0);
      for (int m = 1; m < n; m++) {
        cd i0 = (m == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, m - 1);
        cd i1 = blk_c(in, n, m + 1);
        cd a = cmul(cd{0.0, QC(tb, n, m, 1)}, cconj(i1)), b = cmul(cd{0.0, QC(tb, n, m, 0)}, cconj(i0));
        cd o = cscale(s, cd{a.x - b.x, a.y - b.y});
        out[n * n + 2 * m - 1] = f * o.x;
        out[n * n + 2 * m] = f * o.y;
      }
      cd i0 = (n == 1) ? cscale(1.0, cd{in[n * n], 0.0}) : blk_c(in, n, n - 1);
      cd o = cmul(cd{0.0, -s * QC(tb, n, n, 0)}, cconj(i0));
      out[n * n + 2 * n - 1] = f * o.x;
      out[n * n + 2 * n] = f * o.y;
    }
  }
}
// coaxial translation of a whole block with the level matrices of k_trans (order tq layout)
__device__ void blk_translate(const double *T, int tq, const double *in, double *out, int q, int tid, int nt) {
  for (int e = tid; e < q * q; e += nt) {
    int n = (int)sqrt((double)e);
    while (n * n > e) n--;
    while ((n + 1) * (n + 1) <= e) n++;
    const int mm = e - n * n;
    const int m = (mm + 1) / 2;   // mm = 0 -> m = 0; mm = 2m-1 (re), 2m (im)
    const int w = tq - m;
    const double *Mm = T + tmat_off(tq, m) + (size_t)(n - m) * w;
    double s = 0.0;
    for (int l = m; l < q; l++) s += Mm[l - m] * in[l * l + mm];
    out[e] = s;
  }
}

// ----------------------------------------------------------------------------
// gradient re-expansion: G3[d] = d/dP_d of R^H T R applied to the source block (P = target
// centre - source centre), one workgroup per pair.  The rotation coefficients ROT and
// their theta-derivative dROT advance level by level (m) in place in LDS; each level's
// forward rotations, translations (T and dT from k_trans, grad layout) and back-rotation
// contributions are taken before the level is overwritten.
// LDS (doubles): q^2 source block, 14 q level vectors, then ROT and dROT, (2q-1)^2 complex
// each -- or, at theta = 0 / pi, six q^2 scratch blocks.
// ----------------------------------------------------------------------------
constexpr int GT = 512;
struct GArgs {
  const XJob *jobs;     // kind 0; TH -> T then dT (k_trans grad layout, order tord or pH)
  const double *H;      // source blocks (indexed by J.src)
  const double *TQ;     // source totals for the order (indexed by J.src)
  const int *ord;       // explicit order per job (null: from TQ)
  double *out;          // [slot][3][p^2]
  int p;
  unsigned long long *work;
};

// Job tables of a batch of gradient re-expansions, one (source, target) sphere pair each:
// pair k gets translation matrices at T + 2*TS*k and output slot k.  P is the target's lab
// centre minus the source's, as the host forms it.
__global__ void __launch_bounds__(256) k_dta_jobs(int n, const int *pairs, const double *lab, const double *rad,
                                                  double kappa, double *T, long long TS, TJob *tj, XJob *xj,
                                                  int *lst) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int src = pairs[2 * k], tgt = pairs[2 * k + 1];
  XJob J{};
  TJob M{};
  for (int x = 0; x < 3; x++) {
    const double d = lab[3 * tgt + x] - lab[3 * src + x];
    J.P[x] = d;
    M.P[x] = d;
  }
  J.src = src;
  J.tgt = tgt;
  J.slot = k;
  J.kind = 0;
  J.TH = T + (size_t)k * 2 * TS;
  J.a_src = M.a_src = rad[src];
  J.a_tgt = M.a_tgt = rad[tgt];
  M.kappa_seed = kappa;
  M.off = (long long)k * 2 * TS;
  M.grad = 1;
  M.src = src;
  tj[k] = M;
  xj[k] = J;
  lst[k] = k;
}

// NT threads per pair; QMAX bounds the order (q (q + 1) / 2 <= 2 NT)
template <int NT, int QMAX>
__global__ void __launch_bounds__(NT) k_xgrad(Tabs tb, GArgs A) {
  extern __shared__ double sh[];
  const XJob J = A.jobs[blockIdx.x];
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  double rho, th, ph;
  cart_to_sph(J.P[0], J.P[1], J.P[2], rho, th, ph);
  const int q = A.ord ? A.ord[blockIdx.x] : (A.TQ ? compute_order(rho, A.TQ[J.src], J.a_src, p) : p);
  const int tq = J.tord ? J.tord : q;
  const double *T = J.TH, *dT = J.TH + tmat_size(tq);
  double *G = A.out + (size_t)J.slot * 3 * P2;
  if (A.work && tid == 0) {
    atomicAdd(&A.work[0], (unsigned long long)(3 * work_rotgen(q) + 5 * work_rotapply(q) + 4 * work_transapply(q)));
    atomicAdd(&A.work[1], (unsigned long long)(8 * (q * q + 3 * P2)));
  }
  const int NR = 2 * q - 1;
  double *hs = sh;                             // source block, q^2
  cd *a0 = reinterpret_cast<cd *>(hs + q * q); // R M      at this level
  cd *a1 = a0 + q;                             // dR M
  cd *a2 = a1 + q;                             // dP M
  cd *b0 = a2 + q;                             // dT (R M)
  cd *b1 = b0 + q;                             // T (R M)
  cd *b2 = b1 + q;                             // T (dR M)
  cd *b3 = b2 + q;                             // T (dP M)
  cd *R = b3 + q;                              // ROT level m, row n at n^2 + n + s
  cd *dR = R + NR * NR;                        // dROT level m
  for (int i = tid; i < q * q; i += NT) hs[i] = A.H[(size_t)J.src * P2 + i];
  for (int i = tid; i < 3 * P2; i += NT) G[(i / P2) * P2 + i % P2] = 0.0;
  // CRotCoeff::initParams (rotcoeff.cpp:209-244) with xi = pi
  double sint = sin(th), cost, cott = 0.0;
  int sing = 0;
  cd exphi;
  if (sint < 1e-12) {
    sing = (th < M_PI / 2) ? 1 : -1;
    sint = 0.0;
    cost = (sing == 1) ? 1.0 : -1.0;
    exphi = {1.0, 0.0};
  } else {
    cost = cos(th);
    cott = cost / sint;
    double s_, c_;
    sincos(ph, &s_, &c_);
    exphi = {c_, s_};
  }
  const cd exiphi = cconj(exphi);
  cd exxi;
  {
    double s_, c_;
    sincos(M_PI, &s_, &c_);
    exxi = {c_, s_};
  }
  // Jacobian of (rho, theta, phi) -> (x, y, z) (xforms.cpp:165-192)
  const double ir = 1.0 / rho, sinp = exphi.y, cosp = exphi.x;
  double JR[3][3];
  if (sing) {
    JR[0][0] = 0.0; JR[0][1] = ir * cost; JR[0][2] = 0.0;
    JR[1][0] = 0.0; JR[1][1] = 0.0;       JR[1][2] = ir;
    JR[2][0] = cost; JR[2][1] = 0.0;      JR[2][2] = 0.0;
  } else {
    const double irst = ir / sint, irct = ir * cost;
    JR[0][0] = sint * cosp; JR[0][1] = irct * cosp; JR[0][2] = -irst * sinp;
    JR[1][0] = sint * sinp; JR[1][1] = irct * sinp; JR[1][2] = irst * cosp;
    JR[2][0] = cost;        JR[2][1] = -ir * sint;  JR[2][2] = 0.0;
  }
  __syncthreads();

  if (sing != 0) {
    // closed forms on whole blocks (xforms.cpp:86-115 with rotcoeff.cpp:368-393, 610-771);
    // six scratch blocks in place of the rotation buffers
    double *t1 = reinterpret_cast<double *>(R), *t2 = t1 + q * q, *Gr = t2 + q * q, *Gt = Gr + q * q,
           *Gp = Gt + q * q, *t3 = Gp + q * q;
    sing_rot(hs, t1, q, sing, tid, NT);                      // tM1 = R M
    __syncthreads();
    blk_translate(dT, tq, t1, t2, q, tid, NT);               // dT tM1
    __syncthreads();
    sing_rot(t2, Gr, q, sing, tid, NT);                      // Gr = R^H dT R M
    __syncthreads();
    blk_translate(T, tq, t1, t2, q, tid, NT);                // tL = T R M
    __syncthreads();
    sing_dT(tb, t2, Gt, q, sing, false, tid, NT);            // Gt = dR^H tL
    sing_dP(tb, t2, Gp, q, sing, false, tid, NT);            // Gp = dP^H tL
    sing_dT(tb, hs, t1, q, sing, true, tid, NT);             // dR M
    __syncthreads();
    blk_translate(T, tq, t1, t2, q, tid, NT);
    __syncthreads();
    sing_rot(t2, t3, q, sing, tid, NT);
    __syncthreads();
    for (int i = tid; i < q * q; i += NT) Gt[i] += t3[i];
    sing_dP(tb, hs, t1, q, sing, true, tid, NT);             // dP M
    __syncthreads();
    blk_translate(T, tq, t1, t2, q, tid, NT);
    __syncthreads();
    sing_rot(t2, t3, q, sing, tid, NT);
    __syncthreads();
    for (int i = tid; i < q * q; i += NT) {
      Gp[i] += t3[i];
      const double ar[3] = {Gr[i], Gt[i], Gp[i]};
      for (int d = 0; d < 3; d++) G[d * P2 + i] = JR[d][0] * ar[0] + JR[d][1] * ar[1] + JR[d][2] * ar[2];
    }
    return;
  }

  // level-0 seeds: ROT(n,s,0) = conj-ordered SH of order 2q-1 (rotcoeff.cpp:785-795)
  for (int m = tid; m < NR; m += NT) {
    sh_column(tb, NR, m, cost, sint, ph, [&](int n, double re, double im) {
      if (m == 0) {
        R[n * n + n] = {re, 0.0};
      } else {
        R[n * n + n + m] = {re, -im};
        R[n * n + n - m] = {re, im};
      }
    });
  }
  __syncthreads();
  // dROT level 0 (rotcoeff.cpp:818-836); sh_comp(n, s) = ROT(n, -s, 0) for s >= 0
  for (int n = tid; n < NR; n += NT) {
    if (n == 0) {
      dR[0] = {0.0, 0.0};
      continue;
    }
    const cd *row = R + n * n + n;
    const double c0 = -sqrt((double)n * (n + 1));
    dR[n * n + n] = cmul(cscale(c0, exiphi), row[-1]);
    for (int s = 1; s < n; s++) {
      cd v = cscale(s * cott, row[-s]);
      cd w = cmul(cscale(sqrt((double)(n - s) * (n + s + 1)), exiphi), row[-(s + 1)]);
      cd dm = {v.x - w.x, v.y - w.y};
      dR[n * n + n - s] = dm;
      dR[n * n + n + s] = cconj(dm);
    }
    cd dm = cscale(n * cott, row[-n]);
    dR[n * n + n - n] = dm;
    dR[n * n + n + n] = cconj(dm);
  }
  // back-rotated accumulators: each thread owns up to KO (n, s) outputs (one where the
  // workgroup has a thread per output of order QMAX, which keeps the registers down)
  constexpr int KO = (QMAX * (QMAX + 1) / 2 + NT - 1) / NT;
  const int nout = q * (q + 1) / 2;
  int on[KO], os[KO];
  cd aGr[KO], aGt1[KO], aGt2[KO], aGp1[KO], aGp2[KO];
  for (int k = 0; k < KO; k++) {
    int e = tid + k * NT;
    on[k] = -1;
    os[k] = 0;
    aGr[k] = aGt1[k] = aGt2[k] = aGp1[k] = aGp2[k] = {0.0, 0.0};
    if (e < nout) {
      int n = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
      while (n * (n + 1) / 2 > e) n--;
      while ((n + 1) * (n + 2) / 2 <= e) n++;
      on[k] = n;
      os[k] = e - n * (n + 1) / 2;
    }
  }
  const double r1 = -0.5 * (1 + cost), r2 = 0.5 * (1 - cost), r3 = -sint;
  const double dr1 = 0.5 * sint, dr2 = dr1, dr3 = -cost;
  constexpr int NV = ((2 * QMAX - 2) * (2 * QMAX - 2) + NT - 1) / NT;  // staged values per thread
  __syncthreads();
  for (int m = 0; m < q; m++) {
    // (a) forward rotations of this level: R M, dR M, dP M for rows n in [m, q)
    {
      const int g = tid / XG, lane = tid % XG;
      for (int n = m + g; n < q; n += NT / XG) {
        const cd *row = R + n * n + n, *drow = dR + n * n + n;
        cd s0 = {0, 0}, s1 = {0, 0}, s2 = {0, 0};
        for (int s = lane; s <= n; s += XG) {
          if (s == 0) {
            cd c = {hs[n * n], 0.0};
            s0 = cadd(s0, cmul(row[0], c));
            s1 = cadd(s1, cmul(drow[0], c));
          } else {
            cd c = {hs[n * n + 2 * s - 1], hs[n * n + 2 * s]};
            s0 = cadd(s0, cadd(cmul(row[s], c), cmul(row[-s], cconj(c))));
            s1 = cadd(s1, cadd(cmul(drow[s], c), cmul(drow[-s], cconj(c))));
            cd t = cadd(cmul(row[s], derv(c, s)), cmul(row[-s], derv(cconj(c), -s)));
            s2 = {s2.x - t.x, s2.y - t.y};
          }
        }
        s0.x = seg_reduce(s0.x); s0.y = seg_reduce(s0.y);
        s1.x = seg_reduce(s1.x); s1.y = seg_reduce(s1.y);
        s2.x = seg_reduce(s2.x); s2.y = seg_reduce(s2.y);
        if (lane == 0) {
          if (m == 0) { s0.y = 0.0; s1.y = 0.0; s2.y = 0.0; }  // Mout(n,0) keeps the real part
          a0[n] = s0;
          a1[n] = s1;
          a2[n] = s2;
        }
      }
    }
    __syncthreads();
    // (b) translations of this level: b0 = dT a0, b1 = T a0, b2 = T a1, b3 = T a2
    {
      const int g = tid / XG, lane = tid % XG;
      const int w = tq - m;
      for (int n = m + g; n < q; n += NT / XG) {
        const double *Tm = T + tmat_off(tq, m) + (size_t)(n - m) * w;
        const double *dTm = dT + tmat_off(tq, m) + (size_t)(n - m) * w;
        cd t0 = {0, 0}, t1 = {0, 0}, t2 = {0, 0}, t3 = {0, 0};
        for (int l = m + lane; l < q; l += XG) {
          t0 = cadd(t0, cscale(dTm[l - m], a0[l]));
          t1 = cadd(t1, cscale(Tm[l - m], a0[l]));
          t2 = cadd(t2, cscale(Tm[l - m], a1[l]));
          t3 = cadd(t3, cscale(Tm[l - m], a2[l]));
        }
        t0.x = seg_reduce(t0.x); t0.y = seg_reduce(t0.y);
        t1.x = seg_reduce(t1.x); t1.y = seg_reduce(t1.y);
        t2.x = seg_reduce(t2.x); t2.y = seg_reduce(t2.y);
        t3.x = seg_reduce(t3.x); t3.y = seg_reduce(t3.y);
        if (lane == 0) {
          b0[n] = t0;
          b1[n] = t1;
          b2[n] = t2;
          b3[n] = t3;
        }
      }
    }
    __syncthreads();
    // (c) back-rotation contributions of level m to the owned outputs (n, s), n >= m
    for (int k = 0; k < KO; k++) {
      const int n = on[k], s = os[k];
      if (n < m) continue;
      const cd *row = R + n * n + n, *drow = dR + n * n + n;
      if (m == 0) {
        const cd c0 = {b0[n].x, 0.0}, c1 = {b1[n].x, 0.0}, c2 = {b2[n].x, 0.0}, c3 = {b3[n].x, 0.0};
        aGr[k] = cmul(cscale(1.0, row[-s]), c0);
        aGt1[k] = cmul(cscale(1.0, drow[-s]), c1);
        aGp1[k] = cmul(cscale(1.0, row[-s]), derv(c1, s));
        aGt2[k] = cmul(cscale(1.0, row[-s]), c2);
        aGp2[k] = cmul(cscale(1.0, row[-s]), c3);
      } else {
        const cd Rp = row[s], Rm = row[-s], dRp = drow[s], dRm = drow[-s];
        aGr[k] = cadd(aGr[k], cadd(cmul(cconj(Rp), b0[n]), cmul(Rm, cconj(b0[n]))));
        aGt1[k] = cadd(aGt1[k], cadd(cmul(cconj(dRp), b1[n]), cmul(dRm, cconj(b1[n]))));
        aGp1[k] = cadd(aGp1[k], cadd(cmul(cconj(Rp), derv(b1[n], s)), cmul(Rm, derv(cconj(b1[n]), s))));
        aGt2[k] = cadd(aGt2[k], cadd(cmul(cconj(Rp), b2[n]), cmul(Rm, cconj(b2[n]))));
        aGp2[k] = cadd(aGp2[k], cadd(cmul(cconj(Rp), b3[n]), cmul(Rm, cconj(b3[n]))));
      }
    }
    // (d) advance ROT and dROT to level m+1 in place (rotcoeff.cpp:801-809, 838-848):
    // stage the new rows in registers, then overwrite
    if (m < q - 1) {
      const int base = (m + 1) * (m + 1);
      const int total = (2 * q - m - 2) * (2 * q - m - 2) - (m + 1) * (m + 1);
      cd vR[NV], vD[NV];
      int vi[NV];
      int cnt = 0;
#pragma unroll
      for (int t0 = 0; t0 < NV; t0++) {  // static indices keep the staged values in registers
        const int e = tid + t0 * NT;
        if (e >= total) break;
        cnt = t0 + 1;
        int t = e + base;
        int n = (int)sqrt((double)t);
        while (n * n > t) n--;
        while ((n + 1) * (n + 1) <= t) n++;
        n += 1;
        const int s = t - (n - 1) * (n - 1) - (n - 1);
        const cd *row = R + n * n + n, *drow = dR + n * n + n;
        const double e1 = ETA(tb, n, s - 1), e2 = ETA(tb, n, -s - 1), z3 = ZETA(tb, n - 1, abs(s));
        const double em = ETA(tb, n, m);
        cd d1 = cscale(r1 * e1, row[s - 1]);
        cd d2 = cscale(r2 * e2, row[s + 1]);
        cd d3 = cscale(r3 * z3, row[s]);
        cd v = cmul(exxi, cadd(cadd(cmul(d1, exiphi), cmul(d2, exphi)), d3));
        vR[t0] = {v.x / em, v.y / em};
        cd g1 = cscale(e1, cadd(cscale(r1, drow[s - 1]), cscale(dr1, row[s - 1])));
        cd g2 = cscale(e2, cadd(cscale(r2, drow[s + 1]), cscale(dr2, row[s + 1])));
        cd g3 = cscale(z3, cadd(cscale(r3, drow[s]), cscale(dr3, row[s])));
        cd u = cmul(exxi, cadd(cadd(cmul(g1, exiphi), cmul(g2, exphi)), g3));
        vD[t0] = {u.x / em, u.y / em};
        vi[t0] = (n - 1) * (n - 1) + (n - 1) + s;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NV; k++)
        if (k < cnt) {
          R[vi[k]] = vR[k];
          dR[vi[k]] = vD[k];
        }
    }
    __syncthreads();
  }
  // outputs: Gt = dR^H T R M + R^H T dR M, Gp likewise; G3 = JR (Gr, Gt, Gp)
  for (int k = 0; k < KO; k++) {
    const int n = on[k], s = os[k];
    if (n < 0) continue;
    cd gr = aGr[k], gt = cadd(aGt1[k], aGt2[k]), gp = cadd(aGp1[k], aGp2[k]);
    if (s == 0) {
      const int i = n * n;
      const double ar[3] = {gr.x * 1.0, gt.x * 1.0, gp.x * 1.0};
      for (int d = 0; d < 3; d++) G[d * P2 + i] = JR[d][0] * ar[0] + JR[d][1] * ar[1] + JR[d][2] * ar[2];
    } else {
      const int i = n * n + 2 * s - 1;
      const double re[3] = {gr.x, gt.x, gp.x}, im[3] = {gr.y, gt.y, gp.y};
      for (int d = 0; d < 3; d++) {
        G[d * P2 + i] = JR[d][0] * re[0] + JR[d][1] * re[1] + JR[d][2] * re[2];
        G[d * P2 + i + 1] = JR[d][0] * im[0] + JR[d][1] * im[1] + JR[d][2] * im[2];
      }
    }
  }
}

// ----------------------------------------------------------------------------
// inner F/H iteration with three right-hand sides (the gradient of each sphere's
// coefficients along x, y, z): solveFH with compute_FHbase(3) (expcenter.cpp:1080-1135,
// 1142-1210), computeDev over the three columns and computeDev2(...)/3.
// Right-hand sides come in as LF, LH [sphere entry][3][p^2]; F, H in/out.
// Identity matrix: one wavefront per column, a shared stopping test per sweep.
// General matrix: the three columns of every sweep from LDS.
// ----------------------------------------------------------------------------
struct S3Args {
  const int *sph;          // sphere (for constants and the matrix) per workgroup
  const int *ent;          // entry (for F/H/dev) per workgroup
  const double *LF, *LH;   // [workgroup][3][p^2] right-hand sides
  const double *consts;    // [sphere][10][p]
  const double *const *imat;
  double *F, *H;           // [entry][3][p^2]
  double *dev;             // [entry]
  int p;
  unsigned long long *work;
  const int *sel;          // entries of sph/ent to run (null: the workgroup index)
  const int *lfi;          // right-hand-side block of each entry (null: the workgroup index)
  double *E;               // [block][p^2][6]: the scaled surface right-hand sides of the
                           // result (k_surfgemm's operand), or null
  // identity-matrix entries as S3_PARTS workgroups (grid y; see solve3_split)
  double *spec;                // [row][F, H][3][p^2] tentative states
  unsigned long long *mask;    // [row][S3_PARTS][4] certified-sweep masks and outer-change sums
  unsigned int *cnt;           // [row] arrivals, zero between launches
};

// three right-hand sides: WPC wavefronts per column, column d = wave / WPC
constexpr int WPC = 4;
constexpr int S3T = 3 * WPC * 64;
constexpr int S3_HALF = 2;                // workgroups per column in solve3_split
constexpr int S3_PARTS = 3 * S3_HALF;     // workgroups per entry in solve3_split
// fin (or null): the loop's final F and H ([F, H][3][p^2]) after its 100 sweeps, computed
// by solve3_split; the epilogue then runs on them in place of the loop
template <int KW>
__device__ void solve3_identity(const S3Args &A, int g, int e, int lb, double *red, const double *fin = nullptr) {
  constexpr int NW = 3 * WPC;
  const int p = A.p, P2 = p * p, lane = threadIdx.x % 64, w = threadIdx.x / 64, d = w / WPC, sub = w % WPC;
  const double *cs = A.consts + (size_t)g * 10 * p;
  const double *cLH1 = cs + 2 * p, *cLF1 = cs + 3 * p, *cLH2 = cs + 6 * p, *cLF2 = cs + 7 * p;
  const size_t o = ((size_t)e * 3 + d) * P2, ol = ((size_t)lb * 3 + d) * P2;
  const int npair = p * (p + 1) / 2;
  IdState<KW> S;
  for (int k = 0; k < KW; k++) {
    const int ep = sub * 64 + lane + WPC * 64 * k;
    id_clear(S, k);
    if (ep >= npair) continue;
    int n, i;
    pair_nm(ep, n, i);
    S.nn[k] = n;
    S.i0[k] = i;
    id_consts(S, k, n, cs, p);
    const bool two = (i != n * n);
    S.fbr[k] = cLH1[n] * A.LH[ol + i] + cLF1[n] * A.LF[ol + i];
    S.hbr[k] = cLH2[n] * A.LH[ol + i] + cLF2[n] * A.LF[ol + i];
    S.fbi[k] = two ? cLH1[n] * A.LH[ol + i + 1] + cLF1[n] * A.LF[ol + i + 1] : 0.0;
    S.hbi[k] = two ? cLH2[n] * A.LH[ol + i + 1] + cLF2[n] * A.LF[ol + i + 1] : 0.0;
    S.Fr[k] = A.F[o + i];
    S.Hr[k] = A.H[o + i];
    S.Fi[k] = two ? A.F[o + i + 1] : 0.0;
    S.Hi[k] = two ? A.H[o + i + 1] : 0.0;
    S.uHr[k] = S.Hr[k];
    S.uHi[k] = S.Hi[k];
    if (fin) {
      const double *fF = fin + (size_t)d * P2, *fH = fin + (size_t)(3 + d) * P2;
      S.Fr[k] = fF[i];
      S.Hr[k] = fH[i];
      S.Fi[k] = two ? fF[i + 1] : 0.0;
      S.Hi[k] = two ? fH[i + 1] : 0.0;
    }
  }
  const int ct = fin ? 100 : id_iterate<KW, NW>(S, p, red);
  double td, ts;
  id_dev2(S, td, ts);
  for (int k = 0; k < KW; k++) {
    if (S.nn[k] < 0) continue;
    const int i = S.i0[k], n = S.nn[k];
    A.F[o + i] = S.Fr[k];
    A.H[o + i] = S.Hr[k];
    if (i != n * n) {
      A.F[o + i + 1] = S.Fi[k];
      A.H[o + i + 1] = S.Hi[k];
    }
    if (A.E) {  // w (2n+1)/4pi F and w cChgH[n] H, w = 2 for m > 0
      const double two = (i == n * n) ? 1.0 : 2.0, CF = (2 * n + 1) / (4.0 * M_PI), CH = cs[9 * p + n];
      double *Eo = A.E + ((size_t)lb * P2 + i) * 6;
      Eo[d] = two * (CF * S.Fr[k]);
      Eo[3 + d] = two * (CH * S.Hr[k]);
      if (i != n * n) {
        Eo[6 + d] = two * (CF * S.Fi[k]);
        Eo[6 + 3 + d] = two * (CH * S.Hi[k]);
      }
    }
  }
  td = wave_sum(td);
  ts = wave_sum(ts);
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
    A.dev[e] = TD / (4 * TS) / double(3);
    if (A.work) {
      atomicAdd(&A.work[0], (unsigned long long)(ct * 3 * 12.0 * P2));
      atomicAdd(&A.work[1], (unsigned long long)(8.0 * 3 * 6 * P2));
    }
  }
}

// Part workgroup blockIdx.y of an identity-matrix entry: half h = y % S3_HALF of column
// d = y / S3_HALF's (n, m) pairs, one per thread, through all 100 sweeps of the joint loop
// without its stopping test; bit s of the part's mask is set when one of its pairs alone
// shows that sweep s continues (the test of id_sweep_fast).  Each part writes its pairs'
// results (F, H and the scaled surface right-hand sides E, which are the joint loop's when
// it runs all 100 sweeps), keeps their inputs in the row's backup block, and leaves its
// share of the outer-change sums.  The last of the S3_PARTS workgroups to arrive (a
// device-scope counter; nothing waits) finishes the entry: when every sweep 0..98 is
// certified by some column the joint loop runs all 100 sweeps, so only dev is left (the
// parts' sums in part order); otherwise it restores the inputs and runs the joint loop
// itself.  Either way F, H and E are the joint loop's, bit for bit, from S3_PARTS CUs.
template <int KW>
__device__ void solve3_split(const S3Args &A, int g, int e, int lb, double *red) {
  const int p = A.p, P2 = p * p, tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  const int d = blockIdx.y / S3_HALF, h = blockIdx.y % S3_HALF;
  const int npair = p * (p + 1) / 2, per = (npair + S3_HALF - 1) / S3_HALF;
  const int ep = h * per + tid;
  const bool mine = tid < per && ep < npair;
  const double *cs = A.consts + (size_t)g * 10 * p;
  const double *cLH1 = cs + 2 * p, *cLF1 = cs + 3 * p, *cLH2 = cs + 6 * p, *cLF2 = cs + 7 * p;
  const size_t o = ((size_t)e * 3 + d) * P2, ol = ((size_t)lb * 3 + d) * P2;
  double *spec = A.spec + (size_t)blockIdx.x * 6 * P2;  // inputs: [F, H][3][p^2]
  IdState<1> S;
  id_clear(S, 0);
  int n = 0, i = 0;
  double iFr = 0.0, iFi = 0.0;
  if (mine) {
    pair_nm(ep, n, i);
    S.nn[0] = n;
    S.i0[0] = i;
    id_consts(S, 0, n, cs, p);
    const bool two = (i != n * n);
    S.fbr[0] = cLH1[n] * A.LH[ol + i] + cLF1[n] * A.LF[ol + i];
    S.hbr[0] = cLH2[n] * A.LH[ol + i] + cLF2[n] * A.LF[ol + i];
    S.fbi[0] = two ? cLH1[n] * A.LH[ol + i + 1] + cLF1[n] * A.LF[ol + i + 1] : 0.0;
    S.hbi[0] = two ? cLH2[n] * A.LH[ol + i + 1] + cLF2[n] * A.LF[ol + i + 1] : 0.0;
    S.Fr[0] = iFr = A.F[o + i];
    S.Hr[0] = A.H[o + i];
    S.Fi[0] = iFi = two ? A.F[o + i + 1] : 0.0;
    S.Hi[0] = two ? A.H[o + i + 1] : 0.0;
    S.uHr[0] = S.Hr[0];
    S.uHi[0] = S.Hi[0];
    const double cthr = 4.0 * p * p * 1e-8 * (1.0 + 1e-6);  // as id_iterate
    S.c2[0] = two ? 0.5 * cthr : cthr;
  }
  unsigned long long m0 = 0, m1 = 0;  // sweeps 0..63, 64..99
  if (w * 64 < per) {                 // wavefronts with pairs (uniform per wavefront)
    // each lane records its own certificates; one OR across the wavefront afterwards (no
    // cross-lane step inside the sweep chain)
    for (int t = 0; t < 64; t++) m0 |= (unsigned long long)id_sweep_fast(S) << t;
    for (int t = 64; t < 100; t++) m1 |= (unsigned long long)id_sweep_fast(S) << (t - 64);
    for (int off = 1; off < 64; off <<= 1) {
      m0 |= (unsigned long long)__shfl_xor((long long)m0, off, 64);
      m1 |= (unsigned long long)__shfl_xor((long long)m1, off, 64);
    }
  }
  double td = 0.0, ts = 0.0;
  if (mine) {
    const bool two = (i != n * n);
    spec[(size_t)d * P2 + i] = iFr;  // the inputs, for a rerun of the joint loop
    spec[(size_t)(3 + d) * P2 + i] = S.uHr[0];
    if (two) {
      spec[(size_t)d * P2 + i + 1] = iFi;
      spec[(size_t)(3 + d) * P2 + i + 1] = S.uHi[0];
    }
    if (!two) {  // as the joint loop's epilogue: m = 0 pairs keep zero imaginary parts
      S.Fi[0] = 0.0;
      S.Hi[0] = 0.0;
    }
    A.F[o + i] = S.Fr[0];
    A.H[o + i] = S.Hr[0];
    if (two) {
      A.F[o + i + 1] = S.Fi[0];
      A.H[o + i + 1] = S.Hi[0];
    }
    if (A.E) {  // w (2n+1)/4pi F and w cChgH[n] H, w = 2 for m > 0
      const double tw = two ? 2.0 : 1.0, CF = (2 * n + 1) / (4.0 * M_PI), CH = cs[9 * p + n];
      double *Eo = A.E + ((size_t)lb * P2 + i) * 6;
      Eo[d] = tw * (CF * S.Fr[0]);
      Eo[3 + d] = tw * (CH * S.Hr[0]);
      if (two) {
        Eo[6 + d] = tw * (CF * S.Fi[0]);
        Eo[6 + 3 + d] = tw * (CH * S.Hi[0]);
      }
    }
    id_dev2(S, td, ts);
  }
  td = wave_sum(td);
  ts = wave_sum(ts);
  unsigned long long *wm = reinterpret_cast<unsigned long long *>(red);
  double *wd = red + 2 * (S3T / 64);
  int *flag = reinterpret_cast<int *>(wd + 2 * (S3T / 64));
  if (lane == 0) {
    wm[2 * w] = m0;
    wm[2 * w + 1] = m1;
    wd[2 * w] = td;
    wd[2 * w + 1] = ts;
  }
  __syncthreads();
  unsigned long long *mk = A.mask + (size_t)blockIdx.x * 4 * S3_PARTS;  // [part][m0, m1, td, ts]
  if (tid == 0) {
    unsigned long long a0 = 0, a1 = 0;
    double sd = 0.0, ss = 0.0;
    for (int x = 0; x < S3T / 64; x++) {
      a0 |= wm[2 * x];
      a1 |= wm[2 * x + 1];
      sd += wd[2 * x];
      ss += wd[2 * x + 1];
    }
    mk[4 * blockIdx.y] = a0;
    mk[4 * blockIdx.y + 1] = a1;
    mk[4 * blockIdx.y + 2] = (unsigned long long)__double_as_longlong(sd);
    mk[4 * blockIdx.y + 3] = (unsigned long long)__double_as_longlong(ss);
  }
  __threadfence();  // this workgroup's states, backups and sums, device-wide, before it arrives
  __syncthreads();
  if (tid == 0) flag[0] = (atomicAdd(&A.cnt[blockIdx.x], 1u) == (unsigned)(S3_PARTS - 1));
  __syncthreads();
  if (!flag[0]) return;  // uniform
  __threadfence();       // the other parts' states, backups and sums
  if (tid == 0) {
    (void)atomicExch(&A.cnt[blockIdx.x], 0u);  // for the next launch
    unsigned long long a0 = 0, a1 = 0;
    double TD = 0.0, TS = 0.0;
    for (int x = 0; x < S3_PARTS; x++) {
      a0 |= mk[4 * x];
      a1 |= mk[4 * x + 1];
      TD += __longlong_as_double((long long)mk[4 * x + 2]);
      TS += __longlong_as_double((long long)mk[4 * x + 3]);
    }
    const unsigned long long need1 = (1ull << 35) - 1;  // sweeps 64..98 (sweep 99 ends the loop)
    flag[1] = (a0 == ~0ull) && ((a1 & need1) == need1);
    if (flag[1]) {
      A.dev[e] = TD / (4 * TS) / double(3);
      if (A.work) {
        atomicAdd(&A.work[0], (unsigned long long)(100 * 3 * 12.0 * P2));
        atomicAdd(&A.work[1], (unsigned long long)(8.0 * 3 * 6 * P2));
      }
    }
  }
  __syncthreads();
  if (flag[1]) return;  // uniform
  // the loop stops before sweep 100: restore the inputs and run it jointly
  for (int x = tid; x < 3 * P2; x += S3T) {
    A.F[(size_t)e * 3 * P2 + x] = spec[x];
    A.H[(size_t)e * 3 * P2 + x] = spec[3 * P2 + x];
  }
  __syncthreads();  // red is the joint loop's from here
  solve3_identity<KW>(A, g, e, lb, red, nullptr);
}

__global__ void __launch_bounds__(S3T) k_solve3(S3Args A) {
  extern __shared__ double sh[];
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  const int e0 = A.sel ? A.sel[blockIdx.x] : blockIdx.x;
  const int g = A.sph[e0], e = A.ent[e0], lb = A.lfi ? A.lfi[e0] : blockIdx.x;
  const double *IM = A.imat[g];
  if (!IM) {  // uniform across the workgroup
    __shared__ double red[2 * IC * 3 * WPC + 2 * 3 * WPC + 3 * WPC + 1];
    if (gridDim.y == S3_PARTS) {
      if (p <= 22) solve3_split<1>(A, g, e, lb, red);
      else if (p <= 31) solve3_split<2>(A, g, e, lb, red);
      else solve3_split<3>(A, g, e, lb, red);
    } else {
      if (p <= 22) solve3_identity<1>(A, g, e, lb, red);
      else if (p <= 31) solve3_identity<2>(A, g, e, lb, red);
      else solve3_identity<3>(A, g, e, lb, red);
    }
    return;
  }
  if (blockIdx.y != 0) return;  // general matrix: the first workgroup of the row only
  // general matrix: F, H, X of the three columns in LDS
  double *F = sh, *H = F + 3 * P2, *X = H + 3 * P2, *oldH = X + 3 * P2, *red = oldH + 3 * P2;
  const double *cs = A.consts + (size_t)g * 10 * p;
  const double *cH1 = cs, *cF1 = cs + p, *cLH1 = cs + 2 * p, *cLF1 = cs + 3 * p, *cH2 = cs + 4 * p,
               *cF2 = cs + 5 * p, *cLH2 = cs + 6 * p, *cLF2 = cs + 7 * p, *cInvH = cs + 8 * p;
  const size_t o = (size_t)e * 3 * P2, ol = (size_t)lb * 3 * P2;
  constexpr int KPT = (3 * PMAX * PMAX + S3T - 1) / S3T;
  double fbv[KPT], hbv[KPT], outerH[KPT];
  int nn[KPT];
  for (int k = 0; k < KPT; k++) {
    const int i = tid + k * S3T;
    nn[k] = -1;
    if (i >= 3 * P2) continue;
    const int c = i % P2;
    int n = (int)sqrt((double)c);
    while (n * n > c) n--;
    while ((n + 1) * (n + 1) <= c) n++;
    nn[k] = n;
    fbv[k] = cLH1[n] * A.LH[ol + i] + cLF1[n] * A.LF[ol + i];
    hbv[k] = cLH2[n] * A.LH[ol + i] + cLF2[n] * A.LF[ol + i];
    F[i] = A.F[o + i];
    H[i] = A.H[o + i];
    oldH[i] = H[i];
    outerH[k] = H[i];
  }
  __syncthreads();
  int ct = 0;
  double dev = 1e300;
  while (ct == 0 || (dev > 1e-8 && ct < 100)) {
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * S3T;
      if (nn[k] < 0) continue;
      X[i] = fbv[k] + cH1[nn[k]] * H[i] + cF1[nn[k]] * F[i];
    }
    __syncthreads();
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * S3T;
      if (nn[k] < 0) continue;
      const int d = i / P2, r = i % P2;
      double y = 0.0;
      for (int c = 0; c < P2; c++) y += IM[r + (size_t)c * P2] * X[d * P2 + c];
      F[i] = y * (1 / (double)(2 * nn[k] + 1));
    }
    __syncthreads();
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * S3T;
      if (nn[k] < 0) continue;
      X[i] = hbv[k] + cH2[nn[k]] * H[i] + cF2[nn[k]] * F[i];
    }
    __syncthreads();
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * S3T;
      if (nn[k] < 0) continue;
      const int d = i / P2, r = i % P2;
      double y = 0.0;
      for (int c = 0; c < P2; c++) y += IM[r + (size_t)c * P2] * X[d * P2 + c];
      H[i] = y * cInvH[nn[k]];
    }
    __syncthreads();
    double sum = 0.0;
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * S3T;
      if (nn[k] < 0) continue;
      const int n = nn[k], mm = i % P2 - n * n;
      if (mm > 0 && (mm % 2) == 0) continue;
      double sv;
      if (dev_term(oldH[i], (mm > 0) ? oldH[i + 1] : 0.0, H[i], (mm > 0) ? H[i + 1] : 0.0, mm != 0, sv))
        sum += sv;
    }
    sum = block_sum(sum, red);
    dev = sum / (4 * p * p);
    for (int k = 0; k < KPT; k++) {
      const int i = tid + k * S3T;
      if (nn[k] < 0) continue;
      oldH[i] = H[i];
    }
    ct++;
    __syncthreads();
  }
  double td = 0.0, ts = 0.0;
  for (int k = 0; k < KPT; k++) {
    const int i = tid + k * S3T;
    if (nn[k] < 0) continue;
    const int mm = i % P2 - nn[k] * nn[k];
    const double a = outerH[k], b = H[i], w = (mm == 0) ? 1.0 : 2.0;
    td += w * ((a - b) * (a - b));
    ts += w * (a * a + b * b);
    A.F[o + i] = F[i];
    A.H[o + i] = H[i];
    if (A.E) {
      const int d = i / P2, c = i % P2, n = nn[k];
      const double CF = (2 * n + 1) / (4.0 * M_PI), CH = cs[9 * p + n];
      double *Eo = A.E + ((size_t)lb * P2 + c) * 6;
      Eo[d] = w * (CF * F[i]);
      Eo[3 + d] = w * (CH * H[i]);
    }
  }
  td = block_sum(td, red);
  ts = block_sum(ts, red);
  if (tid == 0) {
    A.dev[e] = td / (4 * ts) / double(3);
    if (A.work) {
      const double P4 = (double)P2 * P2;
      atomicAdd(&A.work[0], (unsigned long long)(ct * 3 * (4 * P4 + 12.0 * P2)));
      atomicAdd(&A.work[1], (unsigned long long)(ct * 16 * P4 + 8.0 * 3 * 6 * P2));
    }
  }
}

// Totals of the exposed-surface gradient charges of each listed entry (the max over
// directions of sum |charge|, expcenter.cpp:1378-1412) from k_surfgemv's per-workgroup
// partials: one wavefront per entry; lane l sums chunks l, l+64, ..., then the lanes are added.
__global__ void __launch_bounds__(64) k_surfgrad_totals(const int *sel, const int *sph, const int *pbuf, int nchunk,
                                                        const double *part, double *maxgF, double *maxgH,
                                                        const double *dev, double *devg) {
  const int e = sel[blockIdx.x];
  if (devg && threadIdx.x == 0) devg[blockIdx.x] = dev[e];  // the entry's change, gathered
  const double *pp = part + (size_t)pbuf[e] * nchunk * 6;
  double t[6] = {0, 0, 0, 0, 0, 0};
  for (int c = threadIdx.x; c < nchunk; c += 64)
    for (int b = 0; b < 6; b++) t[b] += pp[6 * c + b];
  for (int b = 0; b < 6; b++) t[b] = wave_sum(t[b]);
  if (threadIdx.x != 0) return;
  double mf = (t[0] > t[1] ? t[0] : t[1]);
  if (t[2] > mf) mf = t[2];
  double mh = (t[3] > t[4] ? t[3] : t[4]);
  if (t[5] > mh) mh = t[5];
  maxgF[sph[e]] = mf;
  maxgH[sph[e]] = mh;
}

// The intra-molecular sums of a gradient entry in one pass: LGF = sum of its F slots and
// LGH = LGH + (sum of its H slots), both in list order; grid (entry, chunk of 256)
__global__ void __launch_bounds__(256) k_sumfh(const int *sel, const int *obs, const int *lst_off, const int *lst,
                                               const double *slotF, const double *slotH, double *LGF, double *LGH,
                                               int bs) {
  const int e0 = sel[blockIdx.x], go = obs[e0];
  const int l0 = lst_off[e0], l1 = lst_off[e0 + 1];
  for (int i = blockIdx.y * 256 + threadIdx.x; i < bs; i += 256 * gridDim.y) {
    double sF = 0.0, sH = 0.0;
    int e = l0;
    for (; e + 16 <= l1; e += 16) {  // sixteen loads per array in flight, added in list order
      double vF[16], vH[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const size_t o = (size_t)lst[e + u] * bs + i;
        vF[u] = slotF[o];
        vH[u] = slotH[o];
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        sF += vF[u];
        sH += vH[u];
      }
    }
    for (; e + 4 <= l1; e += 4) {
      double vF[4], vH[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const size_t o = (size_t)lst[e + u] * bs + i;
        vF[u] = slotF[o];
        vH[u] = slotH[o];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        sF += vF[u];
        sH += vH[u];
      }
    }
    for (; e < l1; e++) {
      const size_t o = (size_t)lst[e] * bs + i;
      sF += slotF[o];
      sH += slotH[o];
    }
    LGF[(size_t)go * bs + i] = sF;
    LGH[(size_t)go * bs + i] = LGH[(size_t)go * bs + i] + sH;
  }
}

// blocks of bs doubles: dst[didx[e]] = src[sidx[e]] for the listed entries e; grid (entry, chunk)
__global__ void __launch_bounds__(256) k_copyblk(const int *sel, const int *sidx, const int *didx, int bs,
                                                 const double *src, double *dst) {
  const int e = sel[blockIdx.x];
  const double *s = src + (size_t)sidx[e] * bs;
  double *d = dst + (size_t)didx[e] * bs;
  for (int i = blockIdx.y * 256 + threadIdx.x; i < bs; i += 256 * gridDim.y) d[i] = s[i];
}


// ----------------------------------------------------------------------------
// per-sphere force (computeMol_Force_and_Torque moldynamics.cpp:525-550 with
// computeForceOn_0 expcenter.cpp:124-149): one thread per (sphere, direction), the
// inner products in the reference's order (expansion.h:706-741).
//   polarized:      f_d = -<H, gLHN_d> - <gH_d, LS>
//   interact-only:  f_d = -<Hself, gLHN_d>
// ----------------------------------------------------------------------------
__device__ double inprod_blk(const double *M, const double *L, int p) {
  double sum = 0.0;
  for (int n = 0; n < p; n++) {
    double t = 0.0;
    for (int mm = 1; mm < 2 * n + 1; mm++) t += M[n * n + mm] * L[n * n + mm];
    t *= 2.0;
    t += M[n * n] * L[n * n];
    sum += 1.0 * t;
  }
  return sum;
}
__global__ void __launch_bounds__(64) k_force(const int *sph, const int *kind, int n, const double *H,
                                             const double *Hself, const double *LS, const double *gLHN,
                                             const double *gH, double *f, int p) {
  const int e = blockIdx.x * 64 + threadIdx.x;
  if (e >= 3 * n) return;
  const int k = e / 3, d = e % 3, g = sph[k], P2 = p * p;
  double v = 0.0;
  if (kind[k] == 1) {
    v -= inprod_blk(H + (size_t)g * P2, gLHN + ((size_t)g * 3 + d) * P2, p);
    v -= inprod_blk(gH + ((size_t)g * 3 + d) * P2, LS + (size_t)g * P2, p);
  } else {
    v -= inprod_blk(Hself + (size_t)g * P2, gLHN + ((size_t)g * 3 + d) * P2, p);
  }
  f[3 * k + d] = v;
}

// ----------------------------------------------------------------------------
// Numerical re-expansion as a stored linear map.  For an intra-molecular pair whose
// re-expansion is numerical (CXFormNIntra, xforms.cpp:514-582) the local expansion of a
// unit charge at each exposed point of the source, about the target centre, depends on the
// molecule's geometry only.  k_p2lmat stores it once per molecule type, point-major:
//   M[h][c] = Y_c(x_h - P) * k^_n(kr) e^(-kr)/r (a_tgt/r)^n  (expansion.cpp:289-316)
// for F (kappa = 0) and H (kappa) -- the terms CLocalExpan / CGradExpan add per point.
// Applying it is then a streaming product over the points (k_p2lgemv), bounded by HBM.
// ----------------------------------------------------------------------------
struct MBuild {
  long long moff;  // offset of the F matrix; H follows after npts * p^2
  int npts, pt_off;
  double P[3];     // target centre - source centre
  double a_tgt;
};
// one wavefront per point; lane m evaluates harmonic column m
__global__ void __launch_bounds__(256) k_p2lmat(Tabs tb, const MBuild *jobs, const int *sel, const double *xyz,
                                                double *mats, int p) {
  __shared__ double Y[4][PMAX * PMAX];
  __shared__ double wF[4][PMAX], wH[4][PMAX];
  const MBuild J = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
  if (J.npts <= 0) return;  // a source with no exposed point has an empty map
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64, P2 = p * p;
  const int h0 = blockIdx.x * 4 + w;
  const bool valid = h0 < J.npts;
  const int h = valid ? h0 : J.npts - 1;  // idle wavefronts recompute the last point, unstored
  const double *x = xyz + 3 * (size_t)(J.pt_off + h);
  const double dx = x[0] - J.P[0], dy = x[1] - J.P[1], dz = x[2] - J.P[2];
  double rho, th, ph;
  cart_to_sph(dx, dy, dz, rho, th, ph);
  const double cost = cos(th), sint = sin(th);
  if (lane < p) {
    const int m = lane;
    sh_column(tb, p, m, cost, sint, ph, [&](int n, double re, double im) {
      if (m == 0) Y[w][n * n] = re;
      else {
        Y[w][n * n + 2 * m - 1] = re;
        Y[w][n * n + 2 * m] = im;
      }
    });
  }
  if (lane == 63) {
    // radial factors of a unit charge, F unscreened and H screened (CLExpan::init + BESSEL)
    const double val = rho * tb.kappa;
    double mrF = 1.0 / rho, mkF = 1.0;
    mkF *= mrF * 1.0;
    mrF *= J.a_tgt;
    double mrH = 1.0 / rho, mkH = 1.0;
    mkH *= mrH * exp(-val);
    mrH *= J.a_tgt;
    double km2 = 1.0, km1 = 1.0 + val;
    const double vs = val * val;
    for (int n = 0; n < p; n++) {
      double K;
      if (val == 0.0 || n == 0) K = 1.0;
      else if (n == 1) K = km1;
      else {
        K = bessel_k_next(tb, n, km1, km2, vs);
        km2 = km1;
        km1 = K;
      }
      wF[w][n] = 1.0 * mkF;
      wH[w][n] = K * mkH;
      mkF *= mrF;
      mkH *= mrH;
    }
  }
  __syncthreads();
  if (!valid) return;
  double *MF = mats + J.moff + (size_t)h * P2;
  double *MH = mats + J.moff + (size_t)J.npts * P2 + (size_t)h * P2;
  for (int c = lane; c < P2; c += 64) {
    int n = (int)sqrt((double)c);
    while (n * n > c) n--;
    while ((n + 1) * (n + 1) <= c) n++;
    MF[c] = Y[w][c] * wF[w][n];
    MH[c] = Y[w][c] * wH[w][n];
  }
}

// out[r][c] = sum over points h of M[h][c] q_r[h] for the F and H matrices of each job, c
// below the job's order squared (orders from the source totals, CXFormBase::computeOrder);
// points whose charge 3-vector has squared norm < 1e-15 are left out (triexpan.cpp:127-153).
// Grid (job, point split); partial sums go to part and k_p2lsum adds them in split order.
struct MJob {
  const double *M;    // F matrix (npts x p^2, point-major), followed by the H matrix
  int npts, pt_off;   // points of the source (global index into the charge planes)
  int slot;           // output block of r = 0 (blocks slot .. slot+R-1)
  int src;            // source sphere (totals)
  double rho, a_src;
};
struct MArgs {
  const MJob *jobs;
  const int *sel;            // job of each grid row (null: the row index)
  const double *qF, *qH;     // charges, R planes of `plane` doubles
  long long plane;
  const double *nF, *nH;     // squared norms for the skip test (null: none)
  const double *TQF, *TQH;
  int nsplit, spl, p;        // splits per job, points per split (a multiple of MPTS)
  double *part;              // [row][split][2][R][p^2]
  unsigned long long *work;
};
constexpr int MT = 256;
constexpr int MPTS = 32;  // points staged per tile
constexpr int MCW = 2;    // columns per thread, MT apart: a workgroup covers MCW * MT columns
// Grid (job, point split, column chunk).  Each thread owns columns c0 + tid + k MT and
// issues the loads of eight points of both its columns and both maps (32 loads) before it
// adds them, so a launch keeps enough bytes in flight to stream the maps from HBM; the sums
// are formed in the same order as one column at a time.
template <int R>
__global__ void __launch_bounds__(MT) k_p2lgemv(MArgs A) {
  __shared__ double qs[MPTS][2 * R];
  const MJob J = A.jobs[A.sel ? A.sel[blockIdx.x] : blockIdx.x];
  const int p = A.p, P2 = p * p, tid = threadIdx.x;
  // null totals: full order (no adaptive truncation)
  const int pF = A.TQF ? compute_order(J.rho, A.TQF[J.src], J.a_src, p) : p;
  const int pH = A.TQH ? compute_order(J.rho, A.TQH[J.src], J.a_src, p) : p;
  const int nF = pF * pF, nH = pH * pH;
  const int c0 = blockIdx.z * (MCW * MT);
  const int hb = blockIdx.y * A.spl, he = min(J.npts, hb + A.spl);
  double *out = A.part + (((size_t)blockIdx.x * A.nsplit + blockIdx.y) * 2) * R * P2;
  const double *MF = J.M, *MH = MF + (size_t)J.npts * P2;
  bool inF[MCW], inH[MCW];
#pragma unroll
  for (int k = 0; k < MCW; k++) {
    const int c = c0 + tid + k * MT;
    inF[k] = c < nF;
    inH[k] = c < nH;
  }
  double aF[MCW][R], aH[MCW][R];
#pragma unroll
  for (int k = 0; k < MCW; k++)
#pragma unroll
    for (int r = 0; r < R; r++) aF[k][r] = aH[k][r] = 0.0;
  const bool any = c0 < nF || c0 < nH;  // uniform across the workgroup
  for (int h0 = hb; any && h0 < he; h0 += MPTS) {
    const int h1 = min(he, h0 + MPTS);
    __syncthreads();
    // stage this tile's charges (zero where the point is skipped)
    for (int e = tid; e < MPTS * 2 * R; e += MT) {
      const int t = e / (2 * R), v = e % (2 * R), h = h0 + t;
      double q = 0.0;
      if (h < h1) {
        const size_t pt = (size_t)J.pt_off + h;
        if (v < R) q = (A.nF && A.nF[pt] < 1e-15) ? 0.0 : A.qF[v * A.plane + pt];
        else q = (A.nH && A.nH[pt] < 1e-15) ? 0.0 : A.qH[(v - R) * A.plane + pt];
      }
      qs[t][v] = q;
    }
    __syncthreads();
    double tF[MCW][R], tH[MCW][R];
#pragma unroll
    for (int k = 0; k < MCW; k++)
#pragma unroll
      for (int r = 0; r < R; r++) tF[k][r] = tH[k][r] = 0.0;
    int h = h0;
    // groups of 8 points (32 loads in flight), then of 4; each 4 points added as one term
    for (; h + 8 <= h1; h += 8) {
      double mF[MCW][8], mH[MCW][8];
#pragma unroll
      for (int k = 0; k < MCW; k++) {
        const size_t o = (size_t)h * P2 + c0 + tid + k * MT;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          mF[k][u] = inF[k] ? MF[o + (size_t)u * P2] : 0.0;
          mH[k][u] = inH[k] ? MH[o + (size_t)u * P2] : 0.0;
        }
      }
#pragma unroll
      for (int g4 = 0; g4 < 8; g4 += 4) {
        const int t = h - h0 + g4;
#pragma unroll
        for (int k = 0; k < MCW; k++)
#pragma unroll
          for (int r = 0; r < R; r++) {
            tF[k][r] += qs[t][r] * mF[k][g4] + qs[t + 1][r] * mF[k][g4 + 1] + qs[t + 2][r] * mF[k][g4 + 2] +
                        qs[t + 3][r] * mF[k][g4 + 3];
            tH[k][r] += qs[t][R + r] * mH[k][g4] + qs[t + 1][R + r] * mH[k][g4 + 1] +
                        qs[t + 2][R + r] * mH[k][g4 + 2] + qs[t + 3][R + r] * mH[k][g4 + 3];
          }
      }
    }
    for (; h + 4 <= h1; h += 4) {
      double mF[MCW][4], mH[MCW][4];
#pragma unroll
      for (int k = 0; k < MCW; k++) {
        const size_t o = (size_t)h * P2 + c0 + tid + k * MT;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          mF[k][u] = inF[k] ? MF[o + (size_t)u * P2] : 0.0;
          mH[k][u] = inH[k] ? MH[o + (size_t)u * P2] : 0.0;
        }
      }
      const int t = h - h0;
#pragma unroll
      for (int k = 0; k < MCW; k++)
#pragma unroll
        for (int r = 0; r < R; r++) {
          tF[k][r] += qs[t][r] * mF[k][0] + qs[t + 1][r] * mF[k][1] + qs[t + 2][r] * mF[k][2] +
                      qs[t + 3][r] * mF[k][3];
          tH[k][r] += qs[t][R + r] * mH[k][0] + qs[t + 1][R + r] * mH[k][1] + qs[t + 2][R + r] * mH[k][2] +
                      qs[t + 3][R + r] * mH[k][3];
        }
    }
    for (; h < h1; h++) {
#pragma unroll
      for (int k = 0; k < MCW; k++) {
        const size_t o = (size_t)h * P2 + c0 + tid + k * MT;
        const double mf = inF[k] ? MF[o] : 0.0, mh = inH[k] ? MH[o] : 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          tF[k][r] += qs[h - h0][r] * mf;
          tH[k][r] += qs[h - h0][R + r] * mh;
        }
      }
    }
    // tiles are added in point order (a split of one tile keeps the tile's sum)
#pragma unroll
    for (int k = 0; k < MCW; k++)
#pragma unroll
      for (int r = 0; r < R; r++) {
        aF[k][r] = (h0 == hb) ? tF[k][r] : aF[k][r] + tF[k][r];
        aH[k][r] = (h0 == hb) ? tH[k][r] : aH[k][r] + tH[k][r];
      }
  }
#pragma unroll
  for (int k = 0; k < MCW; k++) {
    const int c = c0 + tid + k * MT;
    if (c < P2)
#pragma unroll
      for (int r = 0; r < R; r++) {
        out[(size_t)r * P2 + c] = inF[k] ? aF[k][r] : 0.0;
        out[(size_t)(R + r) * P2 + c] = inH[k] ? aH[k][r] : 0.0;
      }
  }
  if (A.work && tid == 0 && blockIdx.z == 0) {
    const int np = he - hb > 0 ? he - hb : 0;
    atomicAdd(&A.work[0], (unsigned long long)(2.0 * R * (nF + nH) * np));
    atomicAdd(&A.work[1], (unsigned long long)(8.0 * (nF + nH) * np));
  }
}
// The stored maps of a bundle of jobs that share one map (the same target and source
// sphere of one molecule type, in different systems of a batch) as an FP64 matrix-core
// product: out[c][3j + r] = sum_h M[h][c] q_j,r[h] for the F map (blockIdx.z even) or the H
// map (odd).  Workgroup tile: 64 coefficients (one wavefront per 16) x 48 columns (16 jobs
// x 3 planes); the charges of 64 points at a time are staged in LDS with the skip test
// applied, the map is read straight from cache.  Every workgroup writes its rows of the
// jobs' slots, zero at and above each job's own order squared.
typedef double v4d __attribute__((ext_vector_type(4)));
struct PGArgs {
  const MJob *jobs;
  const int *bent, *boff;    // job indices bundle by bundle, [bundle + 1] offsets
  const double *qF, *qH;     // charges, 3 planes of `plane` doubles
  long long plane;
  const double *nF, *nH;     // squared norms for the skip test
  const double *TQF, *TQH;
  double *slotF, *slotH;
  int p;
  unsigned long long *work;
};
constexpr int PGJ = 16;   // jobs per workgroup (48 columns)
constexpr int PGK = 64;   // points per staged chunk
// grid (64-coefficient tile, job tile of the bundle, 2 * bundle + F/H); 256 threads
__global__ void __launch_bounds__(256) k_p2lgemm(PGArgs A) {
  __shared__ double Bs[PGK * 3 * PGJ];
  __shared__ int job_s[PGJ], nq_s[PGJ], pto_s[PGJ];
  __shared__ int maxnq_s;
  const int p = A.p, P2 = p * p, tid = threadIdx.x, w = tid / 64, lane = tid % 64;
  const int hf = blockIdx.z & 1, bund = blockIdx.z >> 1;
  const int b0 = A.boff[bund], nb = A.boff[bund + 1] - b0;
  const int j0 = blockIdx.y * PGJ;
  if (j0 >= nb) return;  // uniform
  const int nj = min(PGJ, nb - j0);
  if (tid < PGJ) {
    int q = 0, jb = -1, po = 0;
    if (tid < nj) {
      jb = A.bent[b0 + j0 + tid];
      const MJob J = A.jobs[jb];
      const double *TQ = hf ? A.TQH : A.TQF;  // null: full order
      const int o = TQ ? compute_order(J.rho, TQ[J.src], J.a_src, p) : p;
      q = o * o;
      po = J.pt_off;
    }
    job_s[tid] = jb;
    nq_s[tid] = q;
    pto_s[tid] = po;
  }
  __syncthreads();
  if (tid == 0) {
    int m = 0;
    for (int t = 0; t < nj; t++) m = max(m, nq_s[t]);
    maxnq_s = m;
  }
  __syncthreads();
  const MJob J0 = A.jobs[job_s[0]];
  const int npts = J0.npts, c0 = blockIdx.x * 64;
  double *slot = hf ? A.slotH : A.slotF;
  const double *q = hf ? A.qH : A.qF;
  const double *nrm = hf ? A.nH : A.nF;
  v4d acc[3];
#pragma unroll
  for (int t = 0; t < 3; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  const int ca = c0 + w * 16 + lane % 16, kq = lane / 16;
  if (c0 < maxnq_s) {  // uniform: rows at or above every job's order stay zero
    const bool cok = ca < P2;
    const double *ma = J0.M + (size_t)hf * npts * P2 + (cok ? ca : 0);
    for (int h0 = 0; h0 < npts; h0 += PGK) {
      __syncthreads();
      // stage the chunk's charges: column col = 3 j + r, 64 points each (coalesced)
      for (int i = tid; i < PGK * 3 * PGJ; i += 256) {
        const int kk = i % PGK, col = i / PGK, jj = col / 3, r = col % 3, h = h0 + kk;
        double v = 0.0;
        if (jj < nj && h < npts) {
          const size_t pt = (size_t)pto_s[jj] + h;
          v = (nrm && nrm[pt] < 1e-15) ? 0.0 : q[(size_t)r * A.plane + pt];  // null: no skip
        }
        Bs[kk * 3 * PGJ + col] = v;
      }
      __syncthreads();
      const int kn = min(PGK, npts - h0);
      if (kn == PGK) {
        double av[PGK / 4];
#pragma unroll
        for (int u = 0; u < PGK / 4; u++) av[u] = cok ? ma[(size_t)(h0 + 4 * u + kq) * P2] : 0.0;
#pragma unroll
        for (int u = 0; u < PGK / 4; u++)
#pragma unroll
          for (int t = 0; t < 3; t++)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], Bs[(4 * u + kq) * 3 * PGJ + t * 16 + lane % 16],
                                                         acc[t], 0, 0, 0);
      } else {
        for (int kk = 0; kk < kn; kk += 4) {  // points past the end are zero in Bs
          const bool hok = kk + kq < kn;
          const double av = (cok && hok) ? ma[(size_t)(h0 + kk + kq) * P2] : 0.0;
#pragma unroll
          for (int t = 0; t < 3; t++)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Bs[(kk + kq) * 3 * PGJ + t * 16 + lane % 16], acc[t],
                                                         0, 0, 0);
        }
      }
    }
  }
  // lane l, register r holds row l/16 + 4 r of the wavefront's 16, column l % 16 of tile t
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const int col = t * 16 + lane % 16, jj = col / 3, rr = col % 3;
    if (jj >= nj) continue;
    const MJob J = A.jobs[job_s[jj]];
    double *o = slot + (size_t)(J.slot + rr) * P2;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int c = c0 + w * 16 + lane / 16 + 4 * r;
      if (c < P2) o[c] = (c < nq_s[jj]) ? acc[t][r] : 0.0;
    }
  }
  if (A.work && tid == 0 && c0 < maxnq_s) {
    double nsum = 0.0;
    for (int t = 0; t < nj; t++) nsum += min(64, max(0, nq_s[t] - c0));
    atomicAdd(&A.work[0], (unsigned long long)(2.0 * 3 * nsum * npts));
    atomicAdd(&A.work[1], (unsigned long long)(8.0 * min(64, maxnq_s - c0) * npts));
  }
}

// slotF/slotH[J.slot + r][c] = sum over splits of the partials, in split order
// grid (job row, output chunks of 256)
__global__ void __launch_bounds__(256) k_p2lsum(const MJob *jobs, const int *sel, int nsplit, int R, int p,
                                               const double *part, double *slotF, double *slotH) {
  const MJob J = jobs[sel ? sel[blockIdx.x] : blockIdx.x];
  const int P2 = p * p;
  for (int e = blockIdx.y * 256 + threadIdx.x; e < 2 * R * P2; e += 256 * gridDim.y) {
    const double *pp = part + ((size_t)blockIdx.x * nsplit * 2) * R * P2 + e;
    const size_t stride = (size_t)2 * R * P2;
    double s = 0.0;
    int sp = 0;
    for (; sp + 16 <= nsplit; sp += 16) {  // sixteen loads in flight, added in split order
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = pp[(sp + u) * stride];
#pragma unroll
      for (int u = 0; u < 16; u++) s += v[u];
    }
    for (; sp + 4 <= nsplit; sp += 4) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = pp[(sp + u) * stride];
#pragma unroll
      for (int u = 0; u < 4; u++) s += v[u];
    }
    for (; sp < nsplit; sp++) s += pp[sp * stride];
    const int half = e / (R * P2), rc = e % (R * P2);
    (half == 0 ? slotF : slotH)[(size_t)J.slot * P2 + rc] = s;
  }
}

// ----------------------------------------------------------------------------
// Exposed-surface gradient charges as a stored linear map.  The harmonics of each exposed
// point about its own sphere centre are fixed by the molecule type; k_surfmat stores them
// point-major (Y[h][c], one row per point, rows in the molecule type's point order), and
// k_surfgemm forms the six charges of each point (F and H gradient along x, y, z) as
//   q[h] = dA * sum_c Y[h][c] w_c E[c],  w_c = (2 if m > 0 else 1) * C_n,
// C_n = (2n+1)/4pi for F and 1/(i^_n(ka)) (2n+1)/4pi for H (expcenter.cpp:1378-1412).
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_surfmat(Tabs tb, const double *xyz, long long npts, double *Y, int p) {
  __shared__ double Ys[4][PMAX * PMAX];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64, P2 = p * p;
  const long long h0 = (long long)blockIdx.x * 4 + w;
  const bool valid = h0 < npts;
  const long long h = valid ? h0 : npts - 1;
  const double *x = xyz + 3 * h;
  double rho, th, ph;
  cart_to_sph(x[0], x[1], x[2], rho, th, ph);
  const double cost = cos(th), sint = sin(th);
  if (lane < p) {
    const int m = lane;
    sh_column(tb, p, m, cost, sint, ph, [&](int n, double re, double im) {
      if (m == 0) Ys[w][n * n] = re;
      else {
        Ys[w][n * n + 2 * m - 1] = re;
        Ys[w][n * n + 2 * m] = im;
      }
    });
  }
  __syncthreads();
  if (!valid) return;
  for (int c = lane; c < P2; c += 64) Y[h * P2 + c] = Ys[w][c];
}

// ----------------------------------------------------------------------------
// Exposed-surface gradient charges of a bundle of entries whose spheres share one molecule
// type's stored harmonics Y (npts x p^2, point-major) as an FP64 matrix-core product
//   C = Y E,  E[c][6t + b] = w_c,b * (gF_t | gH_t)[b][c],
// w = (2 if m > 0 else 1) * C_n, C_n = (2n+1)/4pi (F) or cChgH[n] (H): the six charges of
// every point and entry (expcenter.cpp:1378-1412).  k_solve3 writes each entry's scaled E
// ([c][6], per item) as it finishes; k_surfgemm: one wavefront per 16 points, NT 16-column tiles (entries
// of 6 columns), both operands read straight from cache, K = p^2 in steps of 4.  The
// epilogue stores the charges ([3][all points]), the squared norms of the F and H
// 3-vectors, and per (entry, 64-point tile) the partial sums of |charge| in point order
// (k_surfgrad_totals adds the tiles in order).
// ----------------------------------------------------------------------------
constexpr int SGM_ROWS = 64;  // points per workgroup (4 wavefronts x 16)

struct SGArgs {
  const int *bent;           // entries of the launch, bundle by bundle
  const int *boff;           // [bundle + 1] offsets into bent
  const int *sph;            // [entry] global sphere
  const int *pbuf;           // [entry] item: block of E and of `part`
  const double *const *Y;    // [global sphere] harmonic rows of its exposed points
  const int *pt_off, *pt_cnt, *nspx;  // [global sphere] first point, exposed and spiral counts
  const double *E;           // [item][p^2][6] scaled right-hand sides (written by k_solve3)
  double *gSF, *gSH;         // [3][all points]
  double *nF, *nH;
  long long npts_all;
  double *part;              // [item][chunk][6], chunk = 64-point tile
  int nchunk;                // 64-point tiles per entry (grid x)
  int p;
  unsigned long long *work;
};
// grid (64-point tile, entry tile of the bundle, bundle); 256 threads; NT * 16 / 6 entries
// per workgroup
template <int NT>
__global__ void __launch_bounds__(256) k_surfgemm(SGArgs A) {
  constexpr int NC = NT * 16, NENT = NC / 6;
  __shared__ double Cs[SGM_ROWS * NC];
  __shared__ int ent_s[NENT];
  __shared__ double dA_s[NENT];
  const int p = A.p, P2 = p * p, tid = threadIdx.x, w = tid / 64, lane = tid % 64;
  const int b0 = A.boff[blockIdx.z], nb = A.boff[blockIdx.z + 1] - b0;
  const int t0 = blockIdx.y * NENT;
  if (t0 >= nb) return;  // uniform
  const int nent = min(NENT, nb - t0);
  const int g0 = A.sph[A.bent[b0 + t0]];
  const int npts = A.pt_cnt[g0];
  const int h0 = blockIdx.x * SGM_ROWS;
  if (tid < NENT) {
    const int e = (tid < nent) ? A.bent[b0 + t0 + tid] : -1;
    ent_s[tid] = e;
    dA_s[tid] = (e >= 0) ? 4 * M_PI / double(A.nspx[A.sph[e]]) : 0.0;
  }
  __syncthreads();
  if (h0 >= npts) {  // a tile past this sphere's points: zero partials
    if (tid < nent * 6) A.part[((size_t)A.pbuf[ent_s[tid / 6]] * A.nchunk + blockIdx.x) * 6 + tid % 6] = 0.0;
    return;
  }
  const int nrow = min(SGM_ROWS, npts - h0);
  // A operand: lane l supplies Y[row l%16][k + l/16]; B operand: E[k + l/16][column l%16]
  const int ra = w * 16 + lane % 16, kq = lane / 16;
  const bool rok = ra < nrow;
  const double *ya = A.Y[g0] + (size_t)(h0 + (rok ? ra : 0)) * P2 + kq;
  const double *eb[NT];
  bool cok[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int col = t * 16 + lane % 16, es = col / 6;
    cok[t] = es < nent;
    eb[t] = A.E + ((size_t)A.pbuf[ent_s[cok[t] ? es : 0]] * P2 + kq) * 6 + col % 6;
  }
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  constexpr int UNR = (NT == 1) ? 16 : 8;  // k-steps per iteration: their loads are issued first
  const int K4 = P2 & ~3;
  int k = 0;
  for (; k + 4 * UNR <= K4; k += 4 * UNR) {
    double av[UNR], bv[UNR][NT];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      av[u] = rok ? ya[k + 4 * u] : 0.0;
#pragma unroll
      for (int t = 0; t < NT; t++) bv[u][t] = cok[t] ? eb[t][(size_t)(k + 4 * u) * 6] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u][t], acc[t], 0, 0, 0);
  }
  for (; k < P2; k += 4) {  // remaining steps; k + l/16 past p^2 reads as zero
    const bool kok = k + kq < P2;
    const double av = (rok && kok) ? ya[k] : 0.0;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const double bv = (cok[t] && kok) ? eb[t][(size_t)k * 6] : 0.0;
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
    }
  }
  // C tile to LDS: lane l, register r holds row l/16 + 4 r, column l % 16 of its tile
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) Cs[(w * 16 + lane / 16 + 4 * r) * NC + t * 16 + lane % 16] = acc[t][r];
  __syncthreads();
  // epilogue: wavefront w handles entries w, w + 4, ...; lane = point
  for (int t = w; t < nent; t += 4) {  // uniform per wavefront
    const int e = ent_s[t], g = A.sph[e];
    const double dA = dA_s[t];
    double q[6] = {0, 0, 0, 0, 0, 0};
    if (lane < nrow) {
      const size_t pt = (size_t)A.pt_off[g] + h0 + lane;
      double nf = 0.0, nh = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double gf = Cs[lane * NC + 6 * t + d] * dA, gh = Cs[lane * NC + 6 * t + 3 + d] * dA;
        A.gSF[(size_t)d * A.npts_all + pt] = gf;
        A.gSH[(size_t)d * A.npts_all + pt] = gh;
        q[d] = fabs(gf);
        q[3 + d] = fabs(gh);
        nf += gf * gf;
        nh += gh * gh;
      }
      A.nF[pt] = nf;
      A.nH[pt] = nh;
    }
#pragma unroll
    for (int b = 0; b < 6; b++) q[b] = wave_sum(q[b]);
    if (lane == 0) {
      double *pp = A.part + ((size_t)A.pbuf[e] * A.nchunk + blockIdx.x) * 6;
#pragma unroll
      for (int b = 0; b < 6; b++) pp[b] = q[b];
    }
  }
  if (A.work && tid == 0) {
    atomicAdd(&A.work[0], (unsigned long long)(12.0 * P2 * nrow * nent));
    atomicAdd(&A.work[1], (unsigned long long)(8.0 * P2 * nrow + 64.0 * nrow * nent));
  }
}

// The same product for bundles of one or two entries (a single system): one 16-point tile
// per workgroup, its four wavefronts splitting K = p^2 in four ranges whose partial tiles
// are added in order, so each wavefront's dependent chain is a quarter as long.
constexpr int SG1_ROWS = 16;
// grid (16-point tile, entry pair of the bundle, bundle); 256 threads
template <int NW>  // wavefronts per workgroup, each a K range of the product
__global__ void __launch_bounds__(64 * NW) k_surfgemm_k4(SGArgs A) {
  __shared__ double Cp[NW][16 * 16];
  __shared__ int ent_s[2];
  __shared__ double dA_s[2];
  const int p = A.p, P2 = p * p, tid = threadIdx.x, w = tid / 64, lane = tid % 64;
  const int b0 = A.boff[blockIdx.z], nb = A.boff[blockIdx.z + 1] - b0;
  const int t0 = blockIdx.y * 2;
  if (t0 >= nb) return;  // uniform
  const int nent = min(2, nb - t0);
  const int g0 = A.sph[A.bent[b0 + t0]];
  const int npts = A.pt_cnt[g0];
  const int h0 = blockIdx.x * SG1_ROWS;
  if (tid < 2) {
    const int e = (tid < nent) ? A.bent[b0 + t0 + tid] : -1;
    ent_s[tid] = e;
    dA_s[tid] = (e >= 0) ? 4 * M_PI / double(A.nspx[A.sph[e]]) : 0.0;
  }
  __syncthreads();
  if (h0 >= npts) {  // a tile past this sphere's points: zero partials
    if (tid < nent * 6) A.part[((size_t)A.pbuf[ent_s[tid / 6]] * A.nchunk + blockIdx.x) * 6 + tid % 6] = 0.0;
    return;
  }
  const int nrow = min(SG1_ROWS, npts - h0);
  const int ra = lane % 16, kq = lane / 16;
  const bool rok = ra < nrow;
  const double *ya = A.Y[g0] + (size_t)(h0 + (rok ? ra : 0)) * P2 + kq;
  const int col = lane % 16, es = col / 6;
  const bool cok = es < nent;
  const double *eb = A.E + ((size_t)A.pbuf[ent_s[cok ? es : 0]] * P2 + kq) * 6 + col % 6;
  // this wavefront's K range, in steps of 4
  const int ksteps = (P2 + 3) / 4, kper = (ksteps + NW - 1) / NW;
  const int k0 = 4 * w * kper, k1 = min(P2, 4 * (w + 1) * kper);
  v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
  // UNR steps of 4 per round trip, all loads issued before the products; steps past the
  // wavefront's range and k + l/16 past p^2 read as zero (their products add exact zeros)
  constexpr int UNR = 16;
  for (int k = k0; k < k1; k += 4 * UNR) {
    double av[UNR], bv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int kk = k + 4 * u;
      const bool kok = kk < k1 && kk + kq < P2;
      av[u] = (rok && kok) ? ya[kk] : 0.0;
      bv[u] = (cok && kok) ? eb[(size_t)kk * 6] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
  }
  // lane l, register r holds row l/16 + 4 r, column l % 16
#pragma unroll
  for (int r = 0; r < 4; r++) Cp[w][(lane / 16 + 4 * r) * 16 + lane % 16] = acc[r];
  __syncthreads();
  // epilogue: wavefront t < nent handles entry t, lane = point
  if (w < nent) {
    const int t = w, e = ent_s[t], g = A.sph[e];
    const double dA = dA_s[t];
    double q[6] = {0, 0, 0, 0, 0, 0};
    if (lane < nrow) {
      const size_t pt = (size_t)A.pt_off[g] + h0 + lane;
      double nf = 0.0, nh = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int cf = lane * 16 + 6 * t + d, ch = cf + 3;
        double sf = Cp[0][cf], shv = Cp[0][ch];
#pragma unroll
        for (int x = 1; x < NW; x++) {  // the K ranges' partial products in order
          sf += Cp[x][cf];
          shv += Cp[x][ch];
        }
        const double gf = sf * dA, gh = shv * dA;
        A.gSF[(size_t)d * A.npts_all + pt] = gf;
        A.gSH[(size_t)d * A.npts_all + pt] = gh;
        q[d] = fabs(gf);
        q[3 + d] = fabs(gh);
        nf += gf * gf;
        nh += gh * gh;
      }
      A.nF[pt] = nf;
      A.nH[pt] = nh;
    }
#pragma unroll
    for (int b = 0; b < 6; b++) q[b] = wave_sum(q[b]);
    if (lane == 0) {
      double *pp = A.part + ((size_t)A.pbuf[e] * A.nchunk + blockIdx.x) * 6;
#pragma unroll
      for (int b = 0; b < 6; b++) pp[b] = q[b];
    }
  }
  if (A.work && tid == 0) {
    atomicAdd(&A.work[0], (unsigned long long)(12.0 * P2 * nrow * nent));
    atomicAdd(&A.work[1], (unsigned long long)(8.0 * P2 * nrow + 64.0 * nrow * nent));
  }
}



// blocks listed in blk (B doubles each) multiplied by -1: one workgroup per (block, chunk)
__global__ void __launch_bounds__(256) k_negblk(const int *blk, int B, double *out) {
  const int i = blockIdx.y * 256 + threadIdx.x;
  if (i < B) {
    double *o = out + (size_t)blk[blockIdx.x] * B + i;
    *o = -*o;
  }
}

// ---- rotational derivative of the gradient sweeps (no counterpart in the reference) ----
// Generator of a lab-frame rotation about axis a acting on a block (d/de of the rotation by
// the quaternion (cos e/2, sin e/2 e_a) * q of the rotated block): on the complex
// coefficients c(n,m) (c(n,-m) = conj c(n,m)) it is i conj(L_a), L_+- with the factors
// s_m/s_{m+-1} sqrt((n -+ m)(n +- m + 1)) (s_m = (-1)^m for m >= 0, 1 below; the harmonics
// carry no Condon-Shortley sign) and L_z = m.
__device__ __forceinline__ cd coef_c(const double *B, int n, int m) {
  if (m < -n || m > n) return {0.0, 0.0};
  const int am = m < 0 ? -m : m;
  const cd v = am == 0 ? cd{B[n * n], 0.0} : cd{B[n * n + 2 * am - 1], B[n * n + 2 * am]};
  return m < 0 ? cconj(v) : v;
}
__device__ __forceinline__ double sgm(int m) { return (m >= 0 && (m & 1)) ? -1.0 : 1.0; }
// out[b][a][p^2] = G_a in[blk ? blk[b] : b]; grid (blocks, chunks of 256 coefficients)
__global__ void __launch_bounds__(256) k_rotgen(int p, const int *blk, const double *in, double *out) {
  const int P2 = p * p, i = blockIdx.y * 256 + threadIdx.x;
  if (i >= P2) return;
  const int b = blockIdx.x;
  const double *B = in + (size_t)(blk ? blk[b] : b) * P2;
  int n = (int)sqrt((double)i);
  while (n * n > i) n--;
  while ((n + 1) * (n + 1) <= i) n++;
  const int r = i - n * n, m = (r + 1) / 2;  // r = 0: m = 0 (real); odd: Re c(n,m); even: Im c(n,m)
  const bool im = r > 0 && (r % 2 == 0);
  const cd cl = coef_c(B, n, m - 1), cr = coef_c(B, n, m + 1), c0 = coef_c(B, n, m);
  const double fp = (m - 1 >= -n) ? sgm(m - 1) / sgm(m) * sqrt((double)((n - m + 1) * (n + m))) : 0.0;
  const double fm = (m + 1 <= n) ? sgm(m + 1) / sgm(m) * sqrt((double)((n + m + 1) * (n - m))) : 0.0;
  const cd sx = {fp * cl.x + fm * cr.x, fp * cl.y + fm * cr.y};  // L_+ from m-1, L_- from m+1
  const cd wx = {-0.5 * sx.y, 0.5 * sx.x};                         // i/2 (...)
  const cd wy = {-0.5 * fp * cl.x + 0.5 * fm * cr.x, -0.5 * fp * cl.y + 0.5 * fm * cr.y};
  const cd wz = {-m * c0.y, m * c0.x};                             // i m c
  double *o = out + (size_t)b * 3 * P2 + i;
  o[0] = im ? wx.y : wx.x;
  o[P2] = im ? wy.y : wy.x;
  o[2 * P2] = im ? wz.y : wz.x;
}
// per pair k: the three gradient planes G of its slot become u_k x G (the derivative along
// the displacement omega_a x u_k), plus, where gen[k] >= 0, the three planes of slot2[k]
__global__ void __launch_bounds__(256) k_rot_contract(int p, const double *u, const int *gen, const double *slot2,
                                                      double *slot) {
  const int P2 = p * p, i = blockIdx.y * 256 + threadIdx.x, k = blockIdx.x;
  if (i >= P2) return;
  double *G = slot + (size_t)k * 3 * P2 + i;
  const double g0 = G[0], g1 = G[P2], g2 = G[2 * P2];
  const double ux = u[3 * k], uy = u[3 * k + 1], uz = u[3 * k + 2];
  double r0 = uy * g2 - uz * g1, r1 = uz * g0 - ux * g2, r2 = ux * g1 - uy * g0;
  if (gen[k] >= 0) {
    const double *S2 = slot2 + (size_t)k * 3 * P2 + i;
    r0 += S2[0];
    r1 += S2[P2];
    r2 += S2[2 * P2];
  }
  G[0] = r0;
  G[P2] = r1;
  G[2 * P2] = r2;
}
// re-expansion jobs of the generator planes: pair k's job with source block 3 gen[k] and
// output planes at 3 k
// Rotation jobs of pair directions (CXFormA's rotation, xi = pi): the angles from the pair
// vectors with the same cart_to_sph as k_xform, so that k_rotcoeff stores the coefficients
// k_xform would generate for the pair.
__global__ void __launch_bounds__(256) k_pairrot_jobs(int n, const double *P, long long RS, RJob *jobs) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  double rho, th, ph;
  cart_to_sph(P[3 * j], P[3 * j + 1], P[3 * j + 2], rho, th, ph);
  RJob R{};
  R.theta = th;
  R.phi = ph;
  R.xi = M_PI;
  R.off = (long long)j * RS;
  jobs[j] = R;
}

__global__ void __launch_bounds__(256) k_rot_genjobs(int n, const XJob *xj, const int *gen, XJob *xj2) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  XJob J = xj[k];
  J.src = 3 * (gen[k] >= 0 ? gen[k] : 0);
  J.slot = 3 * k;
  xj2[k] = J;
}
// out[blk[e]][a] -= G_a LSrot[sph[e]] (the turn of a sphere's own back-rotation)
__global__ void __launch_bounds__(256) k_rotgen_sub(int p, const int *blk, const int *sph, const double *gen3,
                                                    double *out) {
  const int P2 = p * p, i = blockIdx.y * 256 + threadIdx.x, e = blockIdx.x;
  if (i >= P2) return;
  const double *g = gen3 + (size_t)sph[e] * 3 * P2 + i;
  double *o = out + (size_t)blk[e] * 3 * P2 + i;
  o[0] -= g[0];
  o[P2] -= g[P2];
  o[2 * P2] -= g[2 * P2];
}

}  // namespace pbsam
#include "pbsam_kernels_imat.hip"
#include "pbsam_kernels_pot.hip"
#include "pbsam_kernels_adjoint.hip"
