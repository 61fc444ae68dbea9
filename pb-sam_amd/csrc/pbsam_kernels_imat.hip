// HIP kernels of the surface-integral matrix build (gfx950, FP64); included by
// pbsam_kernels_grad.hip.
//
// CSolExpCenter::build_IMat with computeIntegralE / computeIntegralB
// (expcenter.cpp:219-335, 667-902): over the smaller of a sphere's exposed or buried
// point sets, the products of its real harmonic columns
//   G[a][b] = sum_h Y_h[a] Y_h[b]
// (one Gram product on the matrix cores), with the reference's Simpson-like point weights
// on the (m = 0, s = 0, n + l even) products (k_imat_gramw), scaled by dA = 4pi/(N-1) --
// and, from the buried points, subtracted from the exact whole-sphere integral -- then
// placed in the reference's column-major layout with its (2n+1)/4pi and factor-2 weights
// (k_imat_assemble).
namespace pbsam {

// G = Y^T Y for the point-major harmonics Y (npts x p^2): one wavefront per 16x16 tile,
// K = points in steps of 4; grid (tile column, tile row)
__global__ void __launch_bounds__(64) k_imat_gram(const double *Y, int npts, int p, double *G) {
  const int P2 = p * p, lane = threadIdx.x;
  const int r0 = blockIdx.y * 16, c0 = blockIdx.x * 16;
  const int ra = r0 + lane % 16, cb = c0 + lane % 16, kq = lane / 16;
  const bool rok = ra < P2, cok = cb < P2;
  v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
  constexpr int UNR = 8;
  const int K4 = npts & ~3;
  int h = 0;
  for (; h + 4 * UNR <= K4; h += 4 * UNR) {
    double av[UNR], bv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const size_t row = (size_t)(h + 4 * u + kq) * P2;
      av[u] = rok ? Y[row + ra] : 0.0;
      bv[u] = cok ? Y[row + cb] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
  }
  for (; h < npts; h += 4) {  // remaining points; rows past the end read as zero
    const bool hok = h + kq < npts;
    const size_t row = (size_t)(h + kq) * P2;
    const double av = (rok && hok) ? Y[row + ra] : 0.0;
    const double bv = (cok && hok) ? Y[row + cb] : 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  // lane l, register r holds row l/16 + 4 r, column l % 16
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int gr = r0 + lane / 16 + 4 * r, gc = c0 + lane % 16;
    if (gr < P2 && gc < P2) G[(size_t)gr * P2 + gc] = acc[r];
  }
}

// the weighted products of the m = 0 columns with n + l even, in point order:
// Gw[l][n] = sum_h w_h Y_h[l^2] Y_h[n^2], w_0 = 1/3, then 4/3 (odd h) and 2/3 (even h)
__global__ void __launch_bounds__(64) k_imat_gramw(const double *Y, int npts, int p, double *Gw) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= p * p) return;
  const int l = i / p, n = i % p, P2 = p * p;
  if ((n + l) % 2) {
    Gw[i] = 0.0;
    return;
  }
  double s = 0.0;
  for (int h = 0; h < npts; h++) {
    const double w = (h > 0) ? ((h % 2 == 0) ? 2.0 / 3.0 : 4.0 / 3.0) : 1.0 / 3.0;
    s += w * Y[(size_t)h * P2 + l * l] * Y[(size_t)h * P2 + n * n];
  }
  Gw[i] = s;
}

// (n, k) of a real column index c = n^2 + k, k in [0, 2n]: m = (k + 1) / 2, imaginary when
// k is even and nonzero
__device__ __forceinline__ void imat_col(int c, int &n, int &m, int &im) {
  n = (int)sqrtf((float)c);  // exact floor for c < 2^22 after the corrections
  while (n * n > c) n--;
  while ((n + 1) * (n + 1) <= c) n++;
  const int k = c - n * n;
  m = (k + 1) / 2;
  im = (k > 0 && k % 2 == 0) ? 1 : 0;
}

// IMat (column-major, p^2 x p^2): column (l, s, y), row (n, m, x) holds
//   (2n+1)/4pi * (s > 0 ? 2 : 1) * G'[(l,s,y)][(n,m,x)],
// G' = dA G (exposed points) or [(n,m,x) == (l,s,y)] (4pi/(2l+1)) (m > 0 ? 1/2 : 1) - dA G
// (buried points), the (m = 0, s = 0, n + l even) products taken weighted (Gw)
__global__ void __launch_bounds__(256) k_imat_assemble(const double *G, const double *Gw, int p, double dA,
                                                        int useE, double *IMat) {
  const int P2 = p * p;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)P2 * P2) return;
  const int col = (int)(idx / P2), row = (int)(idx % P2);
  int l, s, y, n, m, x;
  imat_col(col, l, s, y);
  imat_col(row, n, m, x);
  double raw = (m == 0 && s == 0 && x == 0 && y == 0 && (n + l) % 2 == 0) ? Gw[l * p + n]
                                                                         : G[(size_t)col * P2 + row];
  double v;
  if (useE) {
    v = raw * dA;
  } else if (n == l && m == s && x == y) {
    const double delta = 1.0 / ((2 * l + 1) / (4.0 * M_PI));
    v = ((m == 0) ? delta : 0.5 * delta) - dA * raw;
  } else {
    v = raw * -dA;
  }
  const double fact_n = (2 * n + 1) / (4.0 * M_PI);
  IMat[idx] = ((s > 0) ? 2.0 : 1.0) * fact_n * v;
}

}  // namespace pbsam
