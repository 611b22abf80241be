// Fused generalized Hermitian eigensolve for small matrices (n <= 32), one
// wave64 workgroup per matrix, every intermediate in LDS (FP64):
//   Cholesky S = L L^H, in-place inverse W = L^-1, C = W H W^H (in place,
//   row-wise then column-wise), unblocked Householder tridiagonalisation,
//   bisection on Sturm counts (lane per eigenvalue).
// Same mathematics as the blocked path (eighb scheme="chol",
// slakonet/utils.py:585-816); one launch per batch instead of O(n) launches.
#include "slk_common.h"

#define SM_N 32
#define SM_T 64

__global__ __launch_bounds__(SM_T) void eigh_small_kernel(const double2* __restrict__ H,
                                                          const double2* __restrict__ S, int n, int ld,
                                                          double* Wout, int wld, int* info) {
  const int b = blockIdx.x;
  const double2* h = H + (size_t)b * ld * ld;
  const double2* s = S + (size_t)b * ld * ld;
  __shared__ double2 A[SM_N][SM_N + 1];
  __shared__ double2 B[SM_N][SM_N + 1];
  __shared__ double2 tv[SM_N];
  __shared__ double d[SM_N], e[SM_N];
  __shared__ double s_x;
  __shared__ double2 s_tau, s_scale;
  __shared__ int s_bad, s_skip;
  const int t = threadIdx.x;
  for (int q = t; q < n * n; q += SM_T) {
    const int r = q / n, c = q % n;
    A[r][c] = h[(size_t)r * ld + c];
    B[r][c] = c <= r ? s[(size_t)r * ld + c] : make_double2(0.0, 0.0);
  }
  if (t == 0) s_bad = 0;
  __syncthreads();
  // ---- Cholesky (lower) ----
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      double dk = B[k][k].x;
      if (!(dk > 0.0)) { if (!s_bad) s_bad = k + 1; dk = 1.0; }
      const double l = sqrt(dk);
      B[k][k] = make_double2(l, 0.0);
      s_x = 1.0 / l;
    }
    __syncthreads();
    if (t > k && t < n) B[t][k] = c_scale(B[t][k], s_x);
    __syncthreads();
    for (int q = t; q < n * n; q += SM_T) {
      const int r = q / n, c = q % n;
      if (c > k && r >= c) B[r][c] = c_sub(B[r][c], c_mulc(B[r][k], B[c][k]));
    }
    __syncthreads();
  }
  // ---- in-place inverse of L (trailing block first) ----
  for (int j = n - 1; j >= 0; --j) {
    if (t > j && t < n) tv[t] = B[t][j];
    if (t == 0) B[j][j] = make_double2(1.0 / B[j][j].x, 0.0);
    __syncthreads();
    const double2 neg = make_double2(-B[j][j].x, 0.0);
    if (t > j && t < n) {
      double2 acc = make_double2(0.0, 0.0);
      for (int k = j + 1; k <= t; ++k) acc = c_add(acc, c_mul(B[t][k], tv[k]));
      B[t][j] = c_mul(acc, neg);
    }
    __syncthreads();
  }
  // ---- T = W H, rows bottom-up in place ----
  for (int i = n - 1; i >= 0; --i) {
    double2 acc = make_double2(0.0, 0.0);
    if (t < n)
      for (int k = 0; k <= i; ++k) acc = c_add(acc, c_mul(B[i][k], A[k][t]));
    __syncthreads();
    if (t < n) A[i][t] = acc;
    __syncthreads();
  }
  // ---- C = T W^H, columns right-to-left in place ----
  for (int j = n - 1; j >= 0; --j) {
    double2 acc = make_double2(0.0, 0.0);
    if (t < n)
      for (int k = 0; k <= j; ++k) acc = c_add(acc, c_mulc(A[t][k], B[j][k]));
    __syncthreads();
    if (t < n) A[t][j] = acc;
    __syncthreads();
  }
  // ---- Householder tridiagonalisation (full storage) ----
  for (int j = 0; j < n - 1; ++j) {
    const int m = n - j - 1;
    if (t == 0) {
      double xn2 = 0.0;
      for (int r = j + 2; r < n; ++r) xn2 += A[r][j].x * A[r][j].x + A[r][j].y * A[r][j].y;
      const double2 al = A[j + 1][j];
      d[j] = A[j][j].x;
      if (xn2 == 0.0 && al.y == 0.0) {
        s_skip = 1;
        e[j] = al.x;
      } else {
        s_skip = 0;
        const double nrm = sqrt(al.x * al.x + al.y * al.y + xn2);
        const double beta = al.x >= 0.0 ? -nrm : nrm;
        e[j] = beta;
        s_tau = make_double2((beta - al.x) / beta, -al.y / beta);
        const double2 den = make_double2(al.x - beta, al.y);
        const double dd = den.x * den.x + den.y * den.y;
        s_scale = make_double2(den.x / dd, -den.y / dd);
      }
    }
    __syncthreads();
    if (s_skip) continue;
    // v in tv (rows j+1..n-1), y = A22 v in B's free column space: use B[.][0] as y? keep a
    // separate vector: reuse column 0..1 of B is unsafe; y is kept in registers per lane.
    if (t < m) tv[t] = t == 0 ? make_double2(1.0, 0.0) : c_mul(A[j + 1 + t][j], s_scale);
    __syncthreads();
    double2 y = make_double2(0.0, 0.0);
    if (t < m)
      for (int c = 0; c < m; ++c) y = c_add(y, c_mulc(tv[c], A[j + 1 + c][j + 1 + t]));   // conj(A[c][r]) v[c]
    // conj(A[c][r]) * v[c] == conj(v[c]^* ... ) : computed as conj(A) v via c_mulc(v, A)
    double p = 0.0;
    if (t < m) p = tv[t].x * y.x + tv[t].y * y.y;
    p = wave_sum(p);
    const double2 tau = s_tau;
    const double t2 = 0.5 * (tau.x * tau.x + tau.y * tau.y) * p;
    double2 w = make_double2(0.0, 0.0);
    if (t < m) {
      const double2 pp = c_mul(tau, y);
      w = make_double2(pp.x - t2 * tv[t].x, pp.y - t2 * tv[t].y);
    }
    __shared__ double2 ws[SM_N];
    if (t < m) ws[t] = w;
    __syncthreads();
    for (int q = t; q < m * m; q += SM_T) {
      const int r = q / m, c = q % m;
      const double2 u1 = c_mulc(tv[r], ws[c]), u2 = c_mulc(ws[r], tv[c]);
      A[j + 1 + r][j + 1 + c] = c_sub(A[j + 1 + r][j + 1 + c], c_add(u1, u2));
    }
    __syncthreads();
  }
  if (t == 0) {
    d[n - 1] = A[n - 1][n - 1].x;
    e[n - 1] = 0.0;
    info[b] = s_bad;
  }
  __syncthreads();
  // ---- bisection, lane per eigenvalue ----
  if (t < n) {
    double lo = 1e300, hi = -1e300, emax = 0.0;
    for (int i = 0; i < n; ++i) {
      const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
      lo = fmin(lo, d[i] - r);
      hi = fmax(hi, d[i] + r);
      emax = fmax(emax, e[i] * e[i]);
    }
    const double span = fmax(fabs(lo), fabs(hi));
    lo -= 2.0 * 2.2e-16 * span * n + 1e-300;
    hi += 2.0 * 2.2e-16 * span * n + 1e-300;
    const double piv = fmax(2.2e-308, 2.2e-308 * emax) * 1e2;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      int cnt = 0;
      double q = d[0] - mid;
      if (fabs(q) < piv) q = -piv;
      if (q < 0.0) ++cnt;
      for (int i = 1; i < n; ++i) {
        q = d[i] - mid - e[i - 1] * e[i - 1] / q;
        if (fabs(q) < piv) q = -piv;
        if (q < 0.0) ++cnt;
      }
      if (cnt > t) hi = mid; else lo = mid;
    }
    Wout[(size_t)b * wld + t] = 0.5 * (lo + hi);
  }
}

extern "C" int slk_eigh_small_f64(const void* H, const void* S, int n, int ld, int batch, double* w, int wld,
                                  int* info, hipStream_t stream) {
  if (n > SM_N) return 3;
  if (batch <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(eigh_small_kernel, dim3(batch), dim3(SM_T), 0, stream, reinterpret_cast<const double2*>(H),
                     reinterpret_cast<const double2*>(S), n, ld, w, wld, info);
  return (int)hipGetLastError();
}
