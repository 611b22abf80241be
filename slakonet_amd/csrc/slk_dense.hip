// Blocked dense kernels of the batched eigensolver (gfx950): complex GEMM,
// blocked Cholesky with diagonal-block inverses, blocked triangular inverse,
// and the blocked Householder tridiagonalisation (LAPACK zhetrd/zlatrd
// structure: per column one reflector step + one multi-workgroup trailing
// matrix-vector product; per 32-column panel one rank-2k update on the GEMM).
#include "slk_common.h"
#include "slk_dense.h"

#define GT 64
#define GK 16
#define NB SLK_NB

// ---------------------------------------------------------------------------
// complex FP64 GEMM, 64x64 tile, 4x4 outputs per thread
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 load_op(const double2* p, int ld, int r, int c, int op) {
  // element (r, c) of op(X)
  double2 v = (op & OP_T) ? p[(size_t)c * ld + r] : p[(size_t)r * ld + c];
  if (op & OP_CONJ) v.y = -v.y;
  return v;
}

__global__ __launch_bounds__(256) void zgemm_kernel(int M, int N, int K, double alpha, const double2* __restrict__ A,
                                                    int lda, long long sA, int opA, const double2* __restrict__ B,
                                                    int ldb, long long sB, int opB, double beta, double2* C, int ldc,
                                                    long long sC, int klim, int tri) {
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  if (tri && n0 > m0 + GT - 1) return;
  __shared__ double2 As[GK][GT + 1];
  __shared__ double2 Bs[GK][GT + 1];
  const int b = blockIdx.z;
  const double2* a = A + (size_t)b * sA;
  const double2* bb = B + (size_t)b * sB;
  double2* c = C + (size_t)b * sC;
  const int tid = threadIdx.x;
  const int tr = tid / 16, tc = tid % 16;
  double2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
  int kbeg = 0, kend = K;
  if (klim & KLIM_A_LOWER) kend = min(kend, m0 + GT);
  if (klim & KLIM_A_UPPER) kbeg = max(kbeg, m0);
  if (klim & KLIM_B_UPPER) kend = min(kend, n0 + GT);
  if (klim & KLIM_B_LOWER) kbeg = max(kbeg, n0);
  kbeg = (kbeg / GK) * GK;
  for (int k0 = kbeg; k0 < kend; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      int mm, kk;
      if (opA & OP_T) { kk = e / GT; mm = e % GT; } else { mm = e / GK; kk = e % GK; }
      const int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < K) ? load_op(a, lda, gm, gk, opA) : make_double2(0.0, 0.0);
    }
    for (int e = tid; e < GT * GK; e += 256) {
      int nn, kk;
      if (opB & OP_T) { nn = e / GK; kk = e % GK; } else { kk = e / GT; nn = e % GT; }
      const int gn = n0 + nn, gk = k0 + kk;
      Bs[kk][nn] = (gn < N && gk < K) ? load_op(bb, ldb, gk, gn, opB) : make_double2(0.0, 0.0);
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
      if (gm < M && gn < N) {
        double2 r = make_double2(alpha * acc[i][j].x, alpha * acc[i][j].y);
        if (beta != 0.0) {
          const double2 o = c[(size_t)gm * ldc + gn];
          r.x += beta * o.x;
          r.y += beta * o.y;
        }
        c[(size_t)gm * ldc + gn] = r;
      }
    }
}

void launch_zgemm(hipStream_t st, int batch, int M, int N, int K, double alpha, const double2* A, int lda,
                  long long sA, int opA, const double2* B, int ldb, long long sB, int opB, double beta, double2* C,
                  int ldc, long long sC, int klim, int tri) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  dim3 g((N + GT - 1) / GT, (M + GT - 1) / GT, batch);
  hipLaunchKernelGGL(zgemm_kernel, g, dim3(256), 0, st, M, N, K, alpha, A, lda, sA, opA, B, ldb, sB, opB, beta, C,
                     ldc, sC, klim, tri);
}

__global__ void zero_kernel(uint4* p, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}

void launch_zero(hipStream_t st, void* p, size_t bytes) {
  const size_t n16 = bytes / 16;
  if (n16 == 0) return;
  int blocks = (int)((n16 + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, st, reinterpret_cast<uint4*>(p), n16);
}

// ---------------------------------------------------------------------------
// Cholesky of the NB x NB diagonal block at j0 in LDS, plus its inverse.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void potrf_diag_kernel(double2* L, double2* W, int n, int ld, int j0, int* info) {
  const int b = blockIdx.x;
  double2* a = L + (size_t)b * ld * ld;
  double2* w = W + (size_t)b * ld * ld;
  const int nbk = min(NB, n - j0);
  __shared__ double2 As[NB][NB + 1];
  __shared__ double2 Li[NB][NB + 1];
  __shared__ double s_inv;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx % NB;
    As[r][c] = (r < nbk && c <= r) ? a[(size_t)(j0 + r) * ld + j0 + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int k = 0; k < nbk; ++k) {
    if (tid == 0) {
      double d = As[k][k].x;
      if (!(d > 0.0)) {
        if (info[b] == 0) info[b] = j0 + k + 1;
        d = 1.0;
      }
      const double l = sqrt(d);
      As[k][k] = make_double2(l, 0.0);
      s_inv = 1.0 / l;
    }
    __syncthreads();
    if (tid > k && tid < nbk) {
      const double s = s_inv;
      As[tid][k] = make_double2(As[tid][k].x * s, As[tid][k].y * s);
    }
    __syncthreads();
    for (int idx = tid; idx < nbk * nbk; idx += blockDim.x) {
      const int r = idx / nbk, c = idx % nbk;
      if (c > k && r >= c) {
        const double2 x = As[r][k], y = As[c][k];
        As[r][c].x -= x.x * y.x + x.y * y.y;
        As[r][c].y -= x.y * y.x - x.x * y.y;
      }
    }
    __syncthreads();
  }
  if (tid < nbk) {
    const int c = tid;
    for (int r = 0; r < c; ++r) Li[r][c] = make_double2(0.0, 0.0);
    for (int r = c; r < nbk; ++r) {
      double2 acc = make_double2(r == c ? 1.0 : 0.0, 0.0);
      for (int k = c; k < r; ++k) {
        const double2 lv = As[r][k], wv = Li[k][c];
        acc.x -= lv.x * wv.x - lv.y * wv.y;
        acc.y -= lv.x * wv.y + lv.y * wv.x;
      }
      const double dinv = 1.0 / As[r][r].x;
      Li[r][c] = make_double2(acc.x * dinv, acc.y * dinv);
    }
  }
  __syncthreads();
  for (int idx = tid; idx < nbk * nbk; idx += blockDim.x) {
    const int r = idx / nbk, c = idx % nbk;
    if (c <= r) a[(size_t)(j0 + r) * ld + j0 + c] = As[r][c];
    w[(size_t)(j0 + r) * ld + j0 + c] = Li[r][c];
  }
}

// L[r, j0:j0+NB] = A[r, j0:j0+NB] * Linv^H for the rows below the block.
__global__ __launch_bounds__(256) void trsm_panel_kernel(double2* L, const double2* W, int n, int ld, int j0) {
  const int b = blockIdx.y;
  double2* a = L + (size_t)b * ld * ld;
  const double2* w = W + (size_t)b * ld * ld;
  const int r0 = j0 + NB + blockIdx.x * 32;
  const int nr = min(32, n - r0);
  __shared__ double2 Li[NB][NB + 1];
  __shared__ double2 Ar[32][NB + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int r = idx / NB, c = idx % NB;
    Li[r][c] = w[(size_t)(j0 + r) * ld + j0 + c];
    Ar[r][c] = r < nr ? a[(size_t)(r0 + r) * ld + j0 + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + 256 * q;
    const int r = idx / NB, c = idx % NB;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k <= c; ++k) {
      const double2 x = Ar[r][k], y = Li[c][k];
      acc.x += x.x * y.x + x.y * y.y;
      acc.y += x.y * y.x - x.x * y.y;
    }
    if (r < nr) a[(size_t)(r0 + r) * ld + j0 + c] = acc;
  }
}

void launch_potrf_blocked(hipStream_t st, const double2* S, double2* L, double2* W, int n, int ld, int batch,
                          int* info) {
  // L = S (lower part used), W = 0 (diagonal blocks filled below)
  (void)hipMemcpyAsync(L, S, (size_t)batch * ld * ld * sizeof(double2), hipMemcpyDeviceToDevice, st);
  launch_zero(st, W, (size_t)batch * ld * ld * sizeof(double2));
  for (int j0 = 0; j0 < n; j0 += NB) {
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(batch), dim3(256), 0, st, L, W, n, ld, j0, info);
    const int j1 = j0 + NB;
    if (j1 >= n) break;
    hipLaunchKernelGGL(trsm_panel_kernel, dim3((n - j1 + 31) / 32, batch), dim3(256), 0, st, L, W, n, ld, j0);
    // trailing lower update A22 -= L21 L21^H
    launch_zgemm(st, batch, n - j1, n - j1, NB, -1.0, L + (size_t)j1 * ld + j0, ld, (long long)ld * ld, OP_N,
                 L + (size_t)j1 * ld + j0, ld, (long long)ld * ld, OP_C, 1.0, L + (size_t)j1 * ld + j1, ld,
                 (long long)ld * ld, 0, 1);
  }
}

void launch_trinv_blocked(hipStream_t st, const double2* L, double2* W, double2* T, int n, int ld, int batch) {
  const long long sM = (long long)ld * ld, sT = (long long)NB * ld;
  for (int i0 = NB; i0 < n; i0 += NB) {
    const int mi = min(NB, n - i0);
    // T = L[i0:i0+mi, 0:i0] * W[0:i0, 0:i0]  (W lower)
    launch_zgemm(st, batch, mi, i0, i0, 1.0, L + (size_t)i0 * ld, ld, sM, OP_N, W, ld, sM, OP_N, 0.0, T, ld, sT,
                 KLIM_B_LOWER, 0);
    // W[i0:i0+mi, 0:i0] = -Linv_ii * T
    launch_zgemm(st, batch, mi, i0, mi, -1.0, W + (size_t)i0 * ld + i0, ld, sM, OP_N, T, ld, sT, OP_N, 0.0,
                 W + (size_t)i0 * ld, ld, sM, KLIM_A_LOWER, 0);
  }
}

// ---------------------------------------------------------------------------
// Blocked Householder tridiagonalisation. Panel buffers are column-major:
// Vc[k*ld + r], Wc[k*ld + r] for panel column k. Y: trailing mat-vec result.
// C12: the correction coefficients W^H v (first NB) and V^H v (next NB).
// ---------------------------------------------------------------------------
#define LATRD_THREADS 1024
__global__ __launch_bounds__(LATRD_THREADS) void latrd_col_kernel(double2* A, int n, int ld, int j, int jj,
                                                                  int finalize, int do_col, double2* Vc, double2* Wc,
                                                                  const double2* Y, double2* C12, double2* TAU,
                                                                  double* D, double* E) {
  const int b = blockIdx.x;
  double2* a = A + (size_t)b * ld * ld;
  double2* vc = Vc + (size_t)b * NB * ld;
  double2* wc = Wc + (size_t)b * NB * ld;
  const double2* y = Y + (size_t)b * ld;
  double2* c12 = C12 + (size_t)b * 2 * NB;
  double2* tau = TAU + (size_t)b * ld;
  __shared__ double red[LATRD_THREADS / 64];
  __shared__ double2 s_vj[NB], s_wj[NB];
  __shared__ double2 s_tau, s_scale;
  __shared__ double s_beta;
  __shared__ int s_skip;
  const int tid = threadIdx.x;
  if (finalize) {
    // w_{jj-1} = tau y - |tau|^2/2 (v^H y) v over rows >= j (support of v_{j-1})
    const int kp = jj - 1;
    const double2 tp = tau[j - 1];
    double part = 0.0;
    for (int r = j + tid; r < n; r += blockDim.x) {
      const double2 v = vc[(size_t)kp * ld + r], yy = y[r];
      part += v.x * yy.x + v.y * yy.y;
    }
    const double vhy = block_sum(part, red);
    const double t2 = 0.5 * (tp.x * tp.x + tp.y * tp.y) * vhy;
    for (int r = tid; r < n; r += blockDim.x) {
      double2 o = make_double2(0.0, 0.0);
      if (r >= j) {
        const double2 p = c_mul(tp, y[r]);
        const double2 v = vc[(size_t)kp * ld + r];
        o = make_double2(p.x - t2 * v.x, p.y - t2 * v.y);
      }
      wc[(size_t)kp * ld + r] = o;
    }
    __syncthreads();
  }
  if (!do_col) return;
  // column update with the panel so far (row j of the panel cached in LDS)
  if (tid < jj) {
    s_vj[tid] = vc[(size_t)tid * ld + j];
    s_wj[tid] = wc[(size_t)tid * ld + j];
  }
  __syncthreads();
  for (int r = j + tid; r < n; r += blockDim.x) {
    double2 acc = a[(size_t)r * ld + j];
    for (int k = 0; k < jj; ++k) {
      const double2 t1 = c_mulc(vc[(size_t)k * ld + r], s_wj[k]);
      const double2 t3 = c_mulc(wc[(size_t)k * ld + r], s_vj[k]);
      acc.x -= t1.x + t3.x;
      acc.y -= t1.y + t3.y;
    }
    a[(size_t)r * ld + j] = acc;
  }
  __syncthreads();
  double part = 0.0;
  for (int r = j + 2 + tid; r < n; r += blockDim.x) {
    const double2 x = a[(size_t)r * ld + j];
    part += x.x * x.x + x.y * x.y;
  }
  const double xn2 = block_sum(part, red);
  if (tid == 0) {
    const double2 alpha = a[(size_t)(j + 1) * ld + j];
    if (xn2 == 0.0 && alpha.y == 0.0) {
      s_skip = 1;
      s_tau = make_double2(0.0, 0.0);
      s_beta = alpha.x;
      s_scale = make_double2(0.0, 0.0);
    } else {
      s_skip = 0;
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      const double beta = alpha.x >= 0.0 ? -nrm : nrm;
      s_beta = beta;
      s_tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      const double2 den = make_double2(alpha.x - beta, alpha.y);
      const double dd = den.x * den.x + den.y * den.y;
      s_scale = make_double2(den.x / dd, -den.y / dd);
    }
    D[(size_t)b * ld + j] = a[(size_t)j * ld + j].x;
    E[(size_t)b * ld + j] = s_beta;
    tau[j] = s_tau;
  }
  __syncthreads();
  const double2 scale = s_scale;
  for (int r = tid; r < n; r += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (r == j + 1) v = make_double2(1.0, 0.0);
    else if (r > j + 1) v = s_skip ? make_double2(0.0, 0.0) : c_mul(a[(size_t)r * ld + j], scale);
    vc[(size_t)jj * ld + r] = v;
    if (r > j) a[(size_t)r * ld + j] = v;
  }
  __syncthreads();
  // correction coefficients, one wave per panel column
  const int lane = tid & 63, wv = tid >> 6;
  for (int k = wv; k < jj; k += LATRD_THREADS / 64) {
    double2 s1 = make_double2(0.0, 0.0), s2 = make_double2(0.0, 0.0);
    for (int r = j + 1 + lane; r < n; r += 64) {
      const double2 v = vc[(size_t)jj * ld + r];
      const double2 x1 = c_cmul(wc[(size_t)k * ld + r], v);
      const double2 x2 = c_cmul(vc[(size_t)k * ld + r], v);
      s1.x += x1.x; s1.y += x1.y;
      s2.x += x2.x; s2.y += x2.y;
    }
    s1.x = wave_sum(s1.x); s1.y = wave_sum(s1.y);
    s2.x = wave_sum(s2.x); s2.y = wave_sum(s2.y);
    if (lane == 0) { c12[k] = s1; c12[NB + k] = s2; }
  }
}

// y[r] = sum_c A[r][c] v[c] - sum_k (V_k[r] c1_k + W_k[r] c2_k), rows r > j.
#define HEMV_ROWS 8
__global__ __launch_bounds__(256) void hemv_kernel(const double2* A, int n, int ld, int j, int jj, const double2* Vc,
                                                   const double2* Wc, double2* Y, const double2* C12) {
  const int b = blockIdx.y;
  const double2* a = A + (size_t)b * ld * ld;
  const double2* v = Vc + (size_t)b * NB * ld + (size_t)jj * ld;
  const double2* vc = Vc + (size_t)b * NB * ld;
  const double2* wc = Wc + (size_t)b * NB * ld;
  const double2* c12 = C12 + (size_t)b * 2 * NB;
  double2* y = Y + (size_t)b * ld;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // 2 rows per wave, 8 rows per workgroup: enough workgroups to cover the chip
  const int rbase = j + 1 + blockIdx.x * HEMV_ROWS + wv * 2;
  double2 acc[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) acc[q] = make_double2(0.0, 0.0);
  for (int c = j + 1 + lane; c < n; c += 64) {
    const double2 vv = v[c];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int r = rbase + q;
      if (r < n) {
        const double2 x = a[(size_t)r * ld + c];
        acc[q].x += x.x * vv.x - x.y * vv.y;
        acc[q].y += x.x * vv.y + x.y * vv.x;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    acc[q].x = wave_sum(acc[q].x);
    acc[q].y = wave_sum(acc[q].y);
  }
  if (lane < 2) {
    const int r = rbase + lane;
    if (r < n) {
      double2 s = lane == 0 ? acc[0] : acc[1];
      for (int k = 0; k < jj; ++k) {
        const double2 t1 = c_mul(vc[(size_t)k * ld + r], c12[k]);
        const double2 t2 = c_mul(wc[(size_t)k * ld + r], c12[NB + k]);
        s.x -= t1.x + t2.x;
        s.y -= t1.y + t2.y;
      }
      y[r] = s;
    }
  }
}

__global__ void hetrd_last_kernel(const double2* A, int n, int ld, double* D, double* E, double2* TAU) {
  const int b = blockIdx.x;
  if (threadIdx.x == 0) {
    D[(size_t)b * ld + n - 1] = A[(size_t)b * ld * ld + (size_t)(n - 1) * ld + (n - 1)].x;
    E[(size_t)b * ld + n - 1] = 0.0;
    TAU[(size_t)b * ld + n - 1] = make_double2(0.0, 0.0);
  }
}

void launch_hetrd_blocked(hipStream_t st, double2* C, int n, int ld, int batch, double* D, double* E, double2* TAU,
                          double2* Vc, double2* Wc, double2* Y, double2* C12) {
  const long long sP = (long long)NB * ld, sM = (long long)ld * ld;
  for (int j0 = 0; j0 < n - 1; j0 += NB) {
    const int jn = min(j0 + NB, n - 1);
    for (int j = j0; j < jn; ++j) {
      const int jj = j - j0;
      hipLaunchKernelGGL(latrd_col_kernel, dim3(batch), dim3(LATRD_THREADS), 0, st, C, n, ld, j, jj,
                         jj > 0 ? 1 : 0, 1, Vc, Wc, Y, C12, TAU, D, E);
      const int m = n - j - 1;
      hipLaunchKernelGGL(hemv_kernel, dim3((m + HEMV_ROWS - 1) / HEMV_ROWS, batch), dim3(256), 0, st, C, n, ld, j,
                         jj, Vc, Wc, Y, C12);
    }
    // finalize w of the panel's last column
    hipLaunchKernelGGL(latrd_col_kernel, dim3(batch), dim3(LATRD_THREADS), 0, st, C, n, ld, jn, jn - j0, 1, 0, Vc,
                       Wc, Y, C12, TAU, D, E);
    if (jn < n) {
      const int m = n - jn, kcols = jn - j0;
      // A22 -= V W^H + W V^H on rows/cols >= jn
      launch_zgemm(st, batch, m, m, kcols, -1.0, Vc + jn, ld, sP, OP_T, Wc + jn, ld, sP, OP_CONJ, 1.0,
                   C + (size_t)jn * ld + jn, ld, sM, 0, 0);
      launch_zgemm(st, batch, m, m, kcols, -1.0, Wc + jn, ld, sP, OP_T, Vc + jn, ld, sP, OP_CONJ, 1.0,
                   C + (size_t)jn * ld + jn, ld, sM, 0, 0);
    }
  }
  hipLaunchKernelGGL(hetrd_last_kernel, dim3(batch), dim3(64), 0, st, C, n, ld, D, E, TAU);
}
