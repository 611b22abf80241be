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

