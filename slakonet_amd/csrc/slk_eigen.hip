// Batched generalized Hermitian eigensolver H z = e S z for MI355X (gfx950).
//
// Pipeline per matrix (same mathematics as eighb(scheme="chol"),
// slakonet/utils.py:585-816):
//   potrf   S = L L^H                          (FP64, one workgroup / matrix)
//   trinv   W = L^-1                           (FP64, column-parallel)
//   reduce  C = W H W^H                        (batched complex GEMMs)
//   hetrd   C = Q T Q^H, T real tridiagonal    (Householder, lower)
//   values  bisection on Sturm counts of T     (FP64, thread per eigenvalue)
//   vectors implicit QL on T with rotations accumulated into Z, then
//           Y = Q Z (reflectors) and z = W^H Y (GEMM)
// Matrices are row-major complex (interleaved re/im) with leading dimension
// ld and batch stride ld*ld.
#include "slk_common.h"
#include "slakonet_amd.h"

#define EIG_THREADS 256

// ---------------------------------------------------------------------------
// Cholesky S = L L^H (lower), right-looking, one workgroup per matrix.
// info[b] = j+1 if the j-th pivot is not positive (S not positive definite).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(EIG_THREADS) void potrf_kernel(const double2* __restrict__ S, double2* L,
                                                            int n, int ld, int* info) {
  const int b = blockIdx.x;
  const size_t mo = (size_t)b * ld * ld;
  const double2* s = S + mo;
  double2* a = L + mo;
  const int tid = threadIdx.x;
  __shared__ double s_inv;
  __shared__ int s_bad;
  for (size_t e = tid; e < (size_t)n * ld; e += blockDim.x) {
    const int r = (int)(e / ld), c = (int)(e % ld);
    a[e] = (c <= r && c < n) ? s[e] : make_double2(0.0, 0.0);
  }
  if (tid == 0) s_bad = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      const double d = a[(size_t)j * ld + j].x;
      if (!(d > 0.0) && s_bad == 0) s_bad = j + 1;
      const double l = sqrt(d > 0.0 ? d : 1.0);
      a[(size_t)j * ld + j] = make_double2(l, 0.0);
      s_inv = 1.0 / l;
    }
    __syncthreads();
    const double inv = s_inv;
    for (int r = j + 1 + tid; r < n; r += blockDim.x) {
      double2 v = a[(size_t)r * ld + j];
      a[(size_t)r * ld + j] = make_double2(v.x * inv, v.y * inv);
    }
    __syncthreads();
    const int m = n - j - 1;
    for (int e = tid; e < m * m; e += blockDim.x) {
      const int r = j + 1 + e / m, c = j + 1 + e % m;
      if (c <= r) {
        const double2 x = a[(size_t)r * ld + j], y = a[(size_t)c * ld + j];
        double2 v = a[(size_t)r * ld + c];
        // v -= x * conj(y)
        v.x -= x.x * y.x + x.y * y.y;
        v.y -= x.y * y.x - x.x * y.y;
        a[(size_t)r * ld + c] = v;
      }
    }
    __syncthreads();
  }
  if (tid == 0) info[b] = s_bad;
}

// ---------------------------------------------------------------------------
// W = L^-1 (lower triangular); each thread owns whole columns.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(EIG_THREADS) void trinv_kernel(const double2* __restrict__ L, double2* W,
                                                            int n, int ld) {
  const int b = blockIdx.y;
  const size_t mo = (size_t)b * ld * ld;
  const double2* l = L + mo;
  double2* w = W + mo;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ld) return;
  for (int r = 0; r < c && r < ld; ++r) w[(size_t)r * ld + c] = make_double2(0.0, 0.0);
  if (c >= n) {
    for (int r = c; r < ld; ++r) w[(size_t)r * ld + c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
    return;
  }
  for (int r = c; r < n; ++r) {
    double2 acc = make_double2(r == c ? 1.0 : 0.0, 0.0);
    for (int k = c; k < r; ++k) {
      const double2 lv = l[(size_t)r * ld + k], wv = w[(size_t)k * ld + c];
      acc.x -= lv.x * wv.x - lv.y * wv.y;
      acc.y -= lv.x * wv.y + lv.y * wv.x;
    }
    const double d = 1.0 / l[(size_t)r * ld + r].x;
    w[(size_t)r * ld + c] = make_double2(acc.x * d, acc.y * d);
  }
  for (int r = n; r < ld; ++r) w[(size_t)r * ld + c] = make_double2(0.0, 0.0);
}

// ---------------------------------------------------------------------------
// Batched complex FP64 GEMM: C = opA(A) opB(B), 64x64 tiles, 4x4 per thread.
// op 0 = N, 1 = conjugate transpose. klim bits restrict the K range for
// triangular operands: 1: opA lower (k <= m), 2: opA upper (k >= m),
// 4: opB upper (k <= n), 8: opB lower (k >= n).
// ---------------------------------------------------------------------------
#define GT 64
#define GK 16
__global__ __launch_bounds__(256) void zgemm_kernel(const double2* __restrict__ A, const double2* __restrict__ B,
                                                    double2* C, int M, int N, int K, int ld, int opA, int opB,
                                                    int klim) {
  __shared__ double2 As[GK][GT + 1];
  __shared__ double2 Bs[GK][GT + 1];
  const int b = blockIdx.z;
  const size_t mo = (size_t)b * ld * ld;
  const double2* a = A + mo;
  const double2* bb = B + mo;
  double2* c = C + mo;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int tid = threadIdx.x;
  const int tr = tid / 16, tc = tid % 16;   // 16 x 16 threads, 4x4 each
  double2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
  int kbeg = 0, kend = K;
  if (klim & 1) kend = min(kend, m0 + GT);
  if (klim & 2) kbeg = max(kbeg, m0);
  if (klim & 4) kend = min(kend, n0 + GT);
  if (klim & 8) kbeg = max(kbeg, n0);
  kbeg = (kbeg / GK) * GK;
  for (int k0 = kbeg; k0 < kend; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      int mm, kk;
      if (opA == 0) { mm = e / GK; kk = e % GK; } else { kk = e / GT; mm = e % GT; }
      const int gm = m0 + mm, gk = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (gm < M && gk < K) {
        if (opA == 0) v = a[(size_t)gm * ld + gk];
        else { v = a[(size_t)gk * ld + gm]; v.y = -v.y; }
      }
      As[kk][mm] = v;
    }
    for (int e = tid; e < GT * GK; e += 256) {
      int nn, kk;
      if (opB == 0) { kk = e / GT; nn = e % GT; } else { nn = e / GK; kk = e % GK; }
      const int gn = n0 + nn, gk = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (gn < N && gk < K) {
        if (opB == 0) v = bb[(size_t)gk * ld + gn];
        else { v = bb[(size_t)gn * ld + gk]; v.y = -v.y; }
      }
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GK; ++kk) {
      double2 av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[kk][tr + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tc + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j].x += av[i].x * bv[j].x - av[i].y * bv[j].y;
          acc[i][j].y += av[i].x * bv[j].y + av[i].y * bv[j].x;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gm = m0 + tr + 16 * i, gn = n0 + tc + 16 * j;
      if (gm < M && gn < N) c[(size_t)gm * ld + gn] = acc[i][j];
    }
}

// ---------------------------------------------------------------------------
// Householder tridiagonalisation of Hermitian C (full storage, updated in
// both triangles). Step j: reflector H = I - tau v v^H with H^H x = beta e1
// (beta real), v[0] = 1 stored in column j below the diagonal; trailing block
// A <- H^H A H = A - v w^H - w v^H with y = A v, p = tau y,
// w = p - (|tau|^2 / 2)(v^H y) v.  d = diag(T), e = subdiag(T).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(EIG_THREADS) void hetrd_kernel(double2* C, int n, int ld, double* D, double* Ed,
                                                            double2* TAU, double2* VW) {
  const int b = blockIdx.x;
  double2* a = C + (size_t)b * ld * ld;
  double* d = D + (size_t)b * ld;
  double* e = Ed + (size_t)b * ld;
  double2* tau_out = TAU + (size_t)b * ld;
  double2* v = VW + (size_t)b * 2 * ld;   // [n]
  double2* w = v + ld;                    // [n]
  __shared__ double red[EIG_THREADS / 64];
  __shared__ double2 s_tau;
  __shared__ double s_beta;
  __shared__ double2 s_scale;
  __shared__ int s_skip;
  const int tid = threadIdx.x;
  for (int j = 0; j < n - 1; ++j) {
    const int m = n - j - 1;          // trailing size
    const size_t col = (size_t)j;
    // norm of x[1:]
    double part = 0.0;
    for (int r = tid + 1; r < m; r += blockDim.x) {
      const double2 x = a[(size_t)(j + 1 + r) * ld + col];
      part += x.x * x.x + x.y * x.y;
    }
    const double xn2 = block_sum(part, red);
    if (tid == 0) {
      const double2 alpha = a[(size_t)(j + 1) * ld + col];
      if (xn2 == 0.0 && alpha.y == 0.0) {
        s_skip = 1;
        s_tau = make_double2(0.0, 0.0);
        s_beta = alpha.x;
      } else {
        s_skip = 0;
        const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        const double beta = alpha.x >= 0.0 ? -nrm : nrm;
        s_beta = beta;
        s_tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
        // scale = 1 / (alpha - beta)
        const double2 den = make_double2(alpha.x - beta, alpha.y);
        const double dd = den.x * den.x + den.y * den.y;
        s_scale = make_double2(den.x / dd, -den.y / dd);
      }
      d[j] = a[(size_t)j * ld + j].x;
      e[j] = s_beta;
      tau_out[j] = s_tau;
    }
    __syncthreads();
    if (s_skip) {
      if (tid == 0) a[(size_t)(j + 1) * ld + col] = make_double2(1.0, 0.0);
      __syncthreads();
      continue;
    }
    const double2 scale = s_scale;
    for (int r = tid; r < m; r += blockDim.x) {
      double2 x = r == 0 ? make_double2(1.0, 0.0) : c_mul(a[(size_t)(j + 1 + r) * ld + col], scale);
      v[r] = x;
      a[(size_t)(j + 1 + r) * ld + col] = x;   // store reflector
    }
    __syncthreads();
    // y = A22 v, A22 Hermitian: y[r] = sum_c conj(A[c][r]) v[c] (coalesced in r)
    for (int r = tid; r < m; r += blockDim.x) {
      double2 acc = make_double2(0.0, 0.0);
      for (int c = 0; c < m; ++c) {
        const double2 av = a[(size_t)(j + 1 + c) * ld + (j + 1 + r)];
        const double2 vv = v[c];
        // conj(av) * vv
        acc.x += av.x * vv.x + av.y * vv.y;
        acc.y += av.x * vv.y - av.y * vv.x;
      }
      w[r] = acc;
    }
    __syncthreads();
    // vhy = v^H y (real)
    double ph = 0.0;
    for (int r = tid; r < m; r += blockDim.x) {
      const double2 vv = v[r], yy = w[r];
      ph += vv.x * yy.x + vv.y * yy.y;
    }
    const double vhy = block_sum(ph, red);
    const double2 tau = s_tau;
    const double t2 = 0.5 * (tau.x * tau.x + tau.y * tau.y) * vhy;
    __syncthreads();
    for (int r = tid; r < m; r += blockDim.x) {
      const double2 p = c_mul(tau, w[r]);
      const double2 vv = v[r];
      w[r] = make_double2(p.x - t2 * vv.x, p.y - t2 * vv.y);
    }
    __syncthreads();
    // A22 -= v w^H + w v^H
    for (int e2 = tid; e2 < m * m; e2 += blockDim.x) {
      const int r = e2 / m, c = e2 % m;
      const double2 vr = v[r], wr = w[r], vc = v[c], wc = w[c];
      double2 x = a[(size_t)(j + 1 + r) * ld + (j + 1 + c)];
      const double2 t1 = c_mulc(vr, wc), t3 = c_mulc(wr, vc);
      x.x -= t1.x + t3.x;
      x.y -= t1.y + t3.y;
      a[(size_t)(j + 1 + r) * ld + (j + 1 + c)] = x;
    }
    __syncthreads();
  }
  if (tid == 0) {
    d[n - 1] = a[(size_t)(n - 1) * ld + (n - 1)].x;
    e[n - 1] = 0.0;
    tau_out[n - 1] = make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Eigenvalues of the real symmetric tridiagonal (d, e) by bisection on the
// Sturm count; thread per eigenvalue index, ascending order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sturm_count(const double* d, const double* e, int n, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

__global__ __launch_bounds__(EIG_THREADS) void bisect_kernel(const double* D, const double* Ed, int n, int ld,
                                                             double* Wout, int wld) {
  const int b = blockIdx.y;
  const double* d = D + (size_t)b * ld;
  const double* e = Ed + (size_t)b * ld;
  extern __shared__ double sm[];
  double* sd = sm;
  double* se = sm + n;
  __shared__ double s_lo, s_hi, s_piv;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sd[i] = d[i];
    se[i] = i < n - 1 ? e[i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double lo = 1e300, hi = -1e300, emax = 0.0;
    for (int i = 0; i < n; ++i) {
      const double r = (i > 0 ? fabs(se[i - 1]) : 0.0) + (i < n - 1 ? fabs(se[i]) : 0.0);
      lo = fmin(lo, sd[i] - r);
      hi = fmax(hi, sd[i] + r);
      emax = fmax(emax, se[i] * se[i]);
    }
    const double span = fmax(fabs(lo), fabs(hi));
    s_lo = lo - 2.0 * 2.2e-16 * span * n - 1e-300;
    s_hi = hi + 2.0 * 2.2e-16 * span * n + 1e-300;
    s_piv = fmax(2.2e-308, 2.2e-308 * emax) * 1e2;
  }
  __syncthreads();
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  double lo = s_lo, hi = s_hi;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(sd, se, n, mid, s_piv) > idx) hi = mid; else lo = mid;
  }
  Wout[(size_t)b * wld + idx] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------------------
// Implicit QL with Wilkinson shift on (d, e); rotations accumulated into the
// real matrix Z (column-major: Zc[col*ld + row]), then sorted ascending.
// Lane 0 computes each sweep's rotation sequence; all threads apply it to
// their rows of Z with a carried column.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(EIG_THREADS) void tql2_kernel(const double* D, const double* Ed, int n, int ld,
                                                           double* Zc, double* Wout, int wld, int* info) {
  const int b = blockIdx.x;
  double* z = Zc + (size_t)b * ld * ld;
  extern __shared__ double sm[];
  double* d = sm;
  double* e = sm + n;
  double* cs = sm + 2 * n;
  double* sn = sm + 3 * n;
  __shared__ int s_l, s_m, s_last, s_done, s_fail;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += blockDim.x) {
    d[i] = D[(size_t)b * ld + i];
    e[i] = i < n - 1 ? Ed[(size_t)b * ld + i] : 0.0;
  }
  for (size_t q = tid; q < (size_t)n * ld; q += blockDim.x) {
    const int c = (int)(q / ld), r = (int)(q % ld);
    z[q] = (r == c) ? 1.0 : 0.0;
  }
  if (tid == 0) s_fail = 0;
  __syncthreads();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      if (tid == 0) {
        int m;
        for (m = l; m < n - 1; ++m) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= 2.2e-16 * dd) break;
        }
        s_m = m;
        s_done = (m == l) || (iter >= 60);
        if (iter >= 60 && m != l) s_fail = l + 1;
        if (!s_done) {
          double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
          double r = hypot(g, 1.0);
          g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
          double s = 1.0, c = 1.0, p = 0.0;
          int i;
          int last = l;
          for (i = m - 1; i >= l; --i) {
            double f = s * e[i];
            const double bb = c * e[i];
            r = hypot(f, g);
            e[i + 1] = r;
            if (r == 0.0) {
              d[i + 1] -= p;
              e[m] = 0.0;
              last = i + 1;
              break;
            }
            s = f / r;
            c = g / r;
            g = d[i + 1] - p;
            r = (d[i] - g) * s + 2.0 * c * bb;
            p = s * r;
            d[i + 1] = g + p;
            g = c * r - bb;
            cs[i] = c;
            sn[i] = s;
          }
          if (i < l) {   // completed sweep
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
            last = l;
          }
          s_last = last;
        }
      }
      __syncthreads();
      if (s_done) break;
      ++iter;
      const int m = s_m, last = s_last;
      // rotations i = m-1 .. last act on columns (i, i+1)
      for (int row = tid; row < n; row += blockDim.x) {
        if (last > m - 1) break;
        double carry = z[(size_t)m * ld + row];
        for (int i = m - 1; i >= last; --i) {
          const double zi = z[(size_t)i * ld + row];
          const double c = cs[i], s = sn[i];
          z[(size_t)(i + 1) * ld + row] = s * zi + c * carry;
          carry = c * zi - s * carry;
        }
        z[(size_t)last * ld + row] = carry;
      }
      __syncthreads();
    }
    __syncthreads();
  }
  // selection sort ascending, swapping columns of Z
  for (int i = 0; i < n - 1; ++i) {
    __shared__ int s_k;
    if (tid == 0) {
      int k = i;
      double p = d[i];
      for (int j = i + 1; j < n; ++j)
        if (d[j] < p) { k = j; p = d[j]; }
      s_k = k;
      if (k != i) { d[k] = d[i]; d[i] = p; }
    }
    __syncthreads();
    const int k = s_k;
    if (k != i) {
      for (int row = tid; row < n; row += blockDim.x) {
        const double t = z[(size_t)i * ld + row];
        z[(size_t)i * ld + row] = z[(size_t)k * ld + row];
        z[(size_t)k * ld + row] = t;
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) Wout[(size_t)b * wld + i] = d[i];
  if (tid == 0) info[b] = s_fail;
}

// ---------------------------------------------------------------------------
// Y = Q Z with Q = H_0 H_1 ... H_{n-2}: apply H_j = I - tau_j v_j v_j^H for
// j = n-2 .. 0. Z is real column-major, Y complex row-major. One workgroup
// per (matrix, 64-column block of Z).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(EIG_THREADS) void apply_q_kernel(const double2* __restrict__ Cref,
                                                              const double2* __restrict__ TAU,
                                                              const double* __restrict__ Zc, double2* Y, int n,
                                                              int ld) {
  const int b = blockIdx.y;
  const int c0 = blockIdx.x * 64;
  const double2* a = Cref + (size_t)b * ld * ld;
  const double2* tau = TAU + (size_t)b * ld;
  const double* z = Zc + (size_t)b * ld * ld;
  double2* y = Y + (size_t)b * ld * ld;
  const int tid = threadIdx.x;
  const int nc = min(64, n - c0);
  // load the column block into Y (complex)
  for (int q = tid; q < n * 64; q += blockDim.x) {
    const int r = q / 64, c = q % 64;
    if (c < nc) y[(size_t)r * ld + c0 + c] = make_double2(z[(size_t)(c0 + c) * ld + r], 0.0);
  }
  __syncthreads();
  __shared__ double2 s_dot[4][64];
  const int lane = tid & 63, wv = tid >> 6;
  for (int j = n - 2; j >= 0; --j) {
    const double2 t = tau[j];
    if (t.x == 0.0 && t.y == 0.0) continue;
    const int m = n - j - 1;
    // dot[c] = v^H Y[j+1:, c]
    double2 acc = make_double2(0.0, 0.0);
    if (lane < nc) {
      for (int r = wv; r < m; r += 4) {
        const double2 vv = r == 0 ? make_double2(1.0, 0.0) : a[(size_t)(j + 1 + r) * ld + j];
        const double2 yy = y[(size_t)(j + 1 + r) * ld + c0 + lane];
        acc.x += vv.x * yy.x + vv.y * yy.y;
        acc.y += vv.x * yy.y - vv.y * yy.x;
      }
    }
    s_dot[wv][lane] = acc;
    __syncthreads();
    if (wv == 0) {
      double2 s = s_dot[0][lane];
      for (int q = 1; q < 4; ++q) s = c_add(s, s_dot[q][lane]);
      s_dot[0][lane] = c_mul(t, s);
    }
    __syncthreads();
    for (int q = tid; q < m * 64; q += blockDim.x) {
      const int r = q / 64, c = q % 64;
      if (c < nc) {
        const double2 vv = r == 0 ? make_double2(1.0, 0.0) : a[(size_t)(j + 1 + r) * ld + j];
        const double2 f = c_mul(vv, s_dot[0][c]);
        double2 yv = y[(size_t)(j + 1 + r) * ld + c0 + c];
        yv.x -= f.x;
        yv.y -= f.y;
        y[(size_t)(j + 1 + r) * ld + c0 + c] = yv;
      }
    }
    __syncthreads();
  }
}

extern "C" {

size_t slk_eigh_workspace_bytes(int n, int ld, int batch, int vectors) {
  (void)n;
  const size_t mat = (size_t)ld * ld * sizeof(double2);
  const size_t vec = (size_t)ld * (2 * sizeof(double) + 3 * sizeof(double2));
  size_t per = 3 * mat + vec;
  if (vectors) per += mat;   // real Z (uses half) + slack
  return per * batch + 256;
}

// H, S: [batch][ld][ld] complex128. w: [batch][wld] ascending (Hartree).
// V (optional): [batch][ld][ld] complex128, eigenvectors in columns.
// info: [batch] int; >0 -> S not positive definite at that pivot (Cholesky),
// <0 -> QL did not converge.
// Stage boundaries for optional timing: potrf | trinv | reduce (2 GEMMs) |
// hetrd | tridiagonal solve | back-transform. ev (nullable) holds 7 events.
static int eigh_f64_impl(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                         void* ws, size_t ws_bytes, int* info, hipStream_t stream, hipEvent_t* ev) {
  if (batch <= 0 || n <= 0) return 0;
  if (ws_bytes < slk_eigh_workspace_bytes(n, ld, batch, V != nullptr)) return 2;
  // LDS-resident tridiagonal data: values need 2n doubles, vectors 4n doubles (<= 64 KiB)
  if (n > 4096 || (V != nullptr && n > 2048)) return 3;
  const size_t matel = (size_t)ld * ld;
  double2* B0 = reinterpret_cast<double2*>(ws);
  double2* B1 = B0 + matel * batch;
  double2* B2 = B1 + matel * batch;
  char* tail = reinterpret_cast<char*>(B2 + matel * batch);
  double* D = reinterpret_cast<double*>(tail);
  double* E = D + (size_t)ld * batch;
  double2* TAU = reinterpret_cast<double2*>(E + (size_t)ld * batch);
  double2* VW = TAU + (size_t)ld * batch;
  double* Z = reinterpret_cast<double*>(VW + (size_t)2 * ld * batch);
  const double2* h = reinterpret_cast<const double2*>(H);
  const double2* s = reinterpret_cast<const double2*>(S);
  if (ev) hipEventRecord(ev[0], stream);
  hipLaunchKernelGGL(potrf_kernel, dim3(batch), dim3(EIG_THREADS), 0, stream, s, B0, n, ld, info);
  if (ev) hipEventRecord(ev[1], stream);
  hipLaunchKernelGGL(trinv_kernel, dim3((ld + 255) / 256, batch), dim3(256), 0, stream, B0, B1, n, ld);
  if (ev) hipEventRecord(ev[2], stream);
  dim3 g((n + GT - 1) / GT, (n + GT - 1) / GT, batch);
  // B2 = W H  (W lower)
  hipLaunchKernelGGL(zgemm_kernel, g, dim3(256), 0, stream, B1, h, B2, n, n, n, ld, 0, 0, 1);
  // B0 = B2 W^H (W^H upper)
  hipLaunchKernelGGL(zgemm_kernel, g, dim3(256), 0, stream, B2, B1, B0, n, n, n, ld, 0, 1, 4);
  if (ev) hipEventRecord(ev[3], stream);
  hipLaunchKernelGGL(hetrd_kernel, dim3(batch), dim3(EIG_THREADS), 0, stream, B0, n, ld, D, E, TAU, VW);
  if (ev) hipEventRecord(ev[4], stream);
  if (V == nullptr) {
    hipLaunchKernelGGL(bisect_kernel, dim3((n + 255) / 256, batch), dim3(256), 2 * n * sizeof(double), stream, D,
                       E, n, ld, w, wld);
    if (ev) hipEventRecord(ev[5], stream);
  } else {
    hipLaunchKernelGGL(tql2_kernel, dim3(batch), dim3(EIG_THREADS), 4 * n * sizeof(double), stream, D, E, n, ld, Z,
                       w, wld, info + batch);
    if (ev) hipEventRecord(ev[5], stream);
    hipLaunchKernelGGL(apply_q_kernel, dim3((n + 63) / 64, batch), dim3(EIG_THREADS), 0, stream, B0, TAU, Z, B2,
                       n, ld);
    // V = W^H Y
    hipLaunchKernelGGL(zgemm_kernel, g, dim3(256), 0, stream, B1, B2, reinterpret_cast<double2*>(V), n, n, n, ld,
                       1, 0, 2);
  }
  if (ev) hipEventRecord(ev[6], stream);
  return (int)hipGetLastError();
}

int slk_eigh_f64(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                 void* ws, size_t ws_bytes, int* info, hipStream_t stream) {
  return eigh_f64_impl(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, nullptr);
}

// Same as slk_eigh_f64, and writes the elapsed milliseconds of the six stages
// (potrf, trinv, reduce, hetrd, tridiagonal solve, back-transform) measured
// with HIP events on `stream`; synchronizes on the last event.
int slk_eigh_f64_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                       void* ws, size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms) {
  hipEvent_t ev[7];
  for (int i = 0; i < 7; ++i) hipEventCreate(&ev[i]);
  int rc = eigh_f64_impl(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, ev);
  if (rc == 0) {
    hipEventSynchronize(ev[6]);
    for (int i = 0; i < 6; ++i) hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]);
  }
  for (int i = 0; i < 7; ++i) hipEventDestroy(ev[i]);
  return rc;
}

}  // extern "C"
