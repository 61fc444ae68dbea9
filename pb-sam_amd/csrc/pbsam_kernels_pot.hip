// Potential at points in space (included by pbsam_kernels_grad.hip).
//
// CMolecule::computePotInSpace (moldynamics.cpp:469-507), the per-sphere branch: for each
// point and sphere, the local expansion of a unit charge at the point (CLExpan with screening,
// order computeOrder(rho, TQH, a), scale a: expansion.cpp:289-316,422-436) contracted with the
// sphere's H (inprod, expansion.h:706-741).  One thread per point walks every sphere; the
// spheres' data are read by all threads of a wave at once (broadcast through the caches).

namespace pbsam {

struct PotArgs {
  const double *pts;  // 3 per point (lab frame)
  const double *mol;  // per molecule: rcen (3), orientation (w, x, y, z)
  const int *sph_mol; // molecule of each sphere
  const double *cen;  // sphere centres in the molecule frame, 3 per sphere
  const double *rad;  // sphere radii
  const double *TQ;   // sphere H totals (order selection)
  const double *H;    // sphere H blocks (molecule frame), p^2 each
  double *out;        // potential per point
  int npts, nsph, p;
  int ref_frame;      // 1: the reference's unrotated displacement P - rcen - cen
};

constexpr int POT_NT = 128;

__global__ void __launch_bounds__(POT_NT) k_potential(Tabs tb, PotArgs A) {
  __shared__ double fr[PMAX][POT_NT];  // radial factors of this thread's current sphere
  const int tid = threadIdx.x, t = blockIdx.x * POT_NT + tid;
  if (t >= A.npts) return;
  const int p = A.p, P2 = p * p;
  const double X = A.pts[3 * t], Y = A.pts[3 * t + 1], Z = A.pts[3 * t + 2];
  double pot = 0.0;
  int cur = -1;
  double d[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < A.nsph; k++) {
    const int mi = A.sph_mol[k];
    if (mi != cur) {
      cur = mi;
      const double *M = A.mol + 7 * mi;
      d[0] = X - M[0];
      d[1] = Y - M[1];
      d[2] = Z - M[2];
      if (!A.ref_frame) {  // conj(q) d: the point in the molecule's frame
        const double w = M[3], vx = -M[4], vy = -M[5], vz = -M[6];
        const double c0 = vy * d[2] - vz * d[1], c1 = vz * d[0] - vx * d[2], c2 = vx * d[1] - vy * d[0];
        const double e0 = vy * c2 - vz * c1, e1 = vz * c0 - vx * c2, e2 = vx * c1 - vy * c0;
        d[0] = d[0] + 2 * (w * c0 + e0);
        d[1] = d[1] + 2 * (w * c1 + e1);
        d[2] = d[2] + 2 * (w * c2 + e2);
      }
    }
    const double sx = d[0] - A.cen[3 * k], sy = d[1] - A.cen[3 * k + 1], sz = d[2] - A.cen[3 * k + 2];
    double rho, th, ph;
    cart_to_sph(sx, sy, sz, rho, th, ph);
    const double a = A.rad[k];
    const int q = compute_order(rho, A.TQ[k], a, p);
    // radial part: k-hat_n(kappa rho) e^(-kappa rho) / rho (a / rho)^n
    {
      const double val = rho * tb.kappa, vs = val * val;
      double mr = 1.0 / rho, mk = 1.0;
      mk *= mr * exp(-val);
      mr *= a;
      double km2 = 1.0, km1 = 1.0 + val;
      for (int n = 0; n < q; n++) {
        double K;
        if (val == 0.0 || n == 0) K = 1.0;
        else if (n == 1) K = km1;
        else {
          K = bessel_k_next(tb, n, km1, km2, vs);
          km2 = km1;
          km1 = K;
        }
        fr[n][tid] = K * mk;
        mk *= mr;
      }
    }
    const double *h = A.H + (size_t)k * P2;
    const double cost = cos(th), sint = sin(th);
    double acc = 0.0;
    for (int m = 0; m < q; m++) {
      sh_column(tb, q, m, cost, sint, ph, [&](int n, double re, double im) {
        const double f = fr[n][tid];
        if (m == 0) acc += f * h[n * n] * re;
        else acc += 2.0 * f * (h[n * n + 2 * m - 1] * re + h[n * n + 2 * m] * im);
      });
    }
    pot += acc;
  }
  A.out[t] = pot;
}

// inprod (expansion.h:706-741) of block pairs: out[i] = <A[ia[i]], B[ib[i]]> over all n < p
// (scale product 1), one wave per pair
__global__ void __launch_bounds__(64) k_blk_inprod(int n, const int *ia, const int *ib, const double *A,
                                                   const double *B, int p, double *out) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const int P2 = p * p;
  const double *a = A + (size_t)ia[i] * P2, *b = B + (size_t)ib[i] * P2;
  double s = 0.0;
  for (int x = threadIdx.x; x < P2; x += 64) {
    int nn = (int)sqrt((double)x);
    while (nn * nn > x) nn--;
    while ((nn + 1) * (nn + 1) <= x) nn++;
    s += ((x == nn * nn) ? 1.0 : 2.0) * a[x] * b[x];  // m > 0 entries stand for +m and -m
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
  if (threadIdx.x == 0) out[i] = s;
}

}  // namespace pbsam
