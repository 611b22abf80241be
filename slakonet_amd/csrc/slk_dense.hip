// Dense kernels of the batched eigensolver (gfx950): the batched complex MFMA
// GEMM, and the Cholesky factorisation with its inverse (recursive halving
// over 32x32 leaves factored and inverted in LDS). The tridiagonalisations
// are in slk_hetrd.hip (one-stage) and slk_band.hip (two-stage).
#include "slk_common.h"
#include "slk_dense.h"
#include <hip/hip_ext.h>

#include "slk_sampler.h"

#define GT 64
#define NB SLK_NB

// ---------------------------------------------------------------------------
// Complex GEMM on MFMA: C = alpha opA(A) opB(B) + beta C.
// TM x TM output tile per 256-thread workgroup, 4 waves each owning a
// (TM/2)^2 quadrant as 16x16 MFMA blocks: TM = 64 (2x2 blocks per wave, K
// staged through LDS 16 at a time) for grids that fill the device, TM = 32
// (one block per wave, K 32 at a time) for launches with fewer than 256
// 64x64 tiles, where four times the workgroups shorten the launch. Real and
// imaginary parts sit in separate LDS planes. Complex product = 4 real MFMAs
// per block:
// Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br.
// Operand layout of the 16x16x4 MFMAs (f64 and f32): lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; it receives D[row(l/16, v)][l%16], v < 4.
// ---------------------------------------------------------------------------
template <typename T> struct CT;
// Output row of accumulator element v in lane group g = lane/16 (measured on
// gfx950): f64 16x16x4 interleaves rows (g + 4 v), f32 16x16x4 keeps blocks
// of four (4 g + v); the column is lane % 16 for both.
template <> struct CT<double> {
  using c2 = double2;
  using v4 = __attribute__((ext_vector_type(4))) double;
  static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return g + 4 * v; }
};
template <> struct CT<float> {
  using c2 = float2;
  using v4 = __attribute__((ext_vector_type(4))) float;
  static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
};

// The next K tile is fetched into registers while the current one is
// multiplied from LDS (measured faster than fetch-stage-multiply in turn at
// every size from 64^3 x 8 to 2048^3, FP64 and FP32).
// RE: the operands are real (zero imaginary parts in complex storage): one
// real MFMA per block instead of four, no imaginary LDS planes, and the
// result's imaginary parts are written as zero, or with creal only its real
// parts, packed (element (gm, gn) at real index gm * ldc + gn of the slot).
template <typename T, int TM, bool RE = false>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(int M, int N, int K, T alpha,
                                                        const typename CT<T>::c2* __restrict__ A, int lda,
                                                        long long sA, int opA,
                                                        const typename CT<T>::c2* __restrict__ B, int ldb,
                                                        long long sB, int opB, T beta, typename CT<T>::c2* C,
                                                        int ldc, long long sC, int klim, int tri, int creal,
                                                        unsigned long long* clk) {
  using C2 = typename CT<T>::c2;
  using V4 = typename CT<T>::v4;
  constexpr int KD = 1024 / TM;   // K depth of one LDS stage
  constexpr int BQ = TM / 32;     // 16x16 blocks per wave, per dimension
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TM;
  sample_start(clk);
  if (tri && n0 > m0 + TM - 1) {
    sample_end(clk);
    return;
  }
  __shared__ T Ar[KD][TM + 4], Ai[RE ? 1 : KD][TM + 4], Br[KD][TM + 4], Bi[RE ? 1 : KD][TM + 4];
  const int b = blockIdx.z;
  const C2* a = A + (size_t)b * sA;
  const C2* bb = B + (size_t)b * sB;
  C2* c = C + (size_t)b * sC;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * (TM / 2), wc = (wv & 1) * (TM / 2);
  const int li = lane & 15, lk = lane >> 4;
  V4 accR[BQ][BQ], accI[BQ][BQ];
#pragma unroll
  for (int i = 0; i < BQ; ++i)
#pragma unroll
    for (int j = 0; j < BQ; ++j) {
      accR[i][j] = V4{0, 0, 0, 0};
      accI[i][j] = V4{0, 0, 0, 0};
    }
  int kbeg = 0, kend = K;
  if (klim & KLIM_A_LOWER) kend = min(kend, m0 + TM);
  if (klim & KLIM_A_UPPER) kbeg = max(kbeg, m0);
  if (klim & KLIM_B_UPPER) kend = min(kend, n0 + TM);
  if (klim & KLIM_B_LOWER) kbeg = max(kbeg, n0);
  kbeg = (kbeg / KD) * KD;
  // element e of a K tile: (mm, kk) of op(A), (kk, nn) of op(B), coalesced
  // along the stored rows; loads are unconditional at clamped indices (the
  // values outside op(A), op(B) are zeroed and conjugated at the LDS store,
  // so that no arithmetic waits on a fetch in flight)
  constexpr int NL = TM * KD / 256;
  const bool ta = opA & OP_T, tb = opB & OP_T;
  C2 pa[NL], pb[NL];
  auto a_idx = [&](int e, int& mm, int& kk) {
    if (ta) { kk = e / TM; mm = e % TM; } else { mm = e / KD; kk = e % KD; }
  };
  auto b_idx = [&](int e, int& nn, int& kk) {
    if (tb) { nn = e / KD; kk = e % KD; } else { kk = e / TM; nn = e % TM; }
  };
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      int mm, kk;
      a_idx(tid + 256 * i, mm, kk);
      const int gm = min(m0 + mm, M - 1), gk = min(k0 + kk, K - 1);
      pa[i] = a[ta ? (size_t)gk * lda + gm : (size_t)gm * lda + gk];
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      int nn, kk;
      b_idx(tid + 256 * i, nn, kk);
      const int gn = min(n0 + nn, N - 1), gk = min(k0 + kk, K - 1);
      pb[i] = bb[tb ? (size_t)gn * ldb + gk : (size_t)gk * ldb + gn];
    }
  };
  auto stage = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      int mm, kk;
      a_idx(tid + 256 * i, mm, kk);
      const bool ok = m0 + mm < M && k0 + kk < K;
      Ar[kk][mm] = ok ? pa[i].x : T(0);
      if constexpr (!RE) {
        const T im = ok ? pa[i].y : T(0);
        Ai[kk][mm] = (opA & OP_CONJ) ? -im : im;
      }
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      int nn, kk;
      b_idx(tid + 256 * i, nn, kk);
      const bool ok = n0 + nn < N && k0 + kk < K;
      Br[kk][nn] = ok ? pb[i].x : T(0);
      if constexpr (!RE) {
        const T im = ok ? pb[i].y : T(0);
        Bi[kk][nn] = (opB & OP_CONJ) ? -im : im;
      }
    }
  };
  if (kbeg < kend) fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += KD) {
    stage(k0);
    __syncthreads();
    if (k0 + KD < kend) fetch(k0 + KD);
#pragma unroll
    for (int k4 = 0; k4 < KD; k4 += 4) {
      T ar[BQ], ai[BQ], nai[BQ], br[BQ], bi[BQ];
#pragma unroll
      for (int i = 0; i < BQ; ++i) {
        ar[i] = Ar[k4 + lk][wr + 16 * i + li];
        if constexpr (!RE) {
          ai[i] = Ai[k4 + lk][wr + 16 * i + li];
          nai[i] = -ai[i];
        }
      }
#pragma unroll
      for (int j = 0; j < BQ; ++j) {
        br[j] = Br[k4 + lk][wc + 16 * j + li];
        if constexpr (!RE) bi[j] = Bi[k4 + lk][wc + 16 * j + li];
      }
#pragma unroll
      for (int i = 0; i < BQ; ++i)
#pragma unroll
        for (int j = 0; j < BQ; ++j) {
          accR[i][j] = CT<T>::mfma(ar[i], br[j], accR[i][j]);
          if constexpr (!RE) {
            accR[i][j] = CT<T>::mfma(nai[i], bi[j], accR[i][j]);
            accI[i][j] = CT<T>::mfma(ar[i], bi[j], accI[i][j]);
            accI[i][j] = CT<T>::mfma(ai[i], br[j], accI[i][j]);
          }
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < BQ; ++i)
#pragma unroll
    for (int j = 0; j < BQ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wr + 16 * i + CT<T>::row(lk, v), gn = n0 + wc + 16 * j + li;
        if (RE && creal && gm < M && gn < N) {
          T* cr = reinterpret_cast<T*>(c) + (size_t)gm * ldc + gn;
          T r = alpha * accR[i][j][v];
          if (beta != T(0)) r += beta * *cr;
          *cr = r;
        } else if (gm < M && gn < N) {
          C2 r;
          r.x = alpha * accR[i][j][v];
          r.y = RE ? T(0) : alpha * accI[i][j][v];
          if (beta != T(0)) {
            const C2 o = c[(size_t)gm * ldc + gn];
            r.x += beta * o.x;
            if constexpr (!RE) r.y += beta * o.y;
          }
          c[(size_t)gm * ldc + gn] = r;
        }
      }
  if (clk != nullptr) {
    __syncthreads();
    sample_end(clk);
  }
}

// Flops of the executed TM x TM output tiles (edge tiles clipped to M x N,
// K ranges as limited by klim): 8 real flops per complex multiply-add.
static double gemm_tile_flops(int batch, int M, int N, int K, int klim, int tri, int tm) {
  if (!klim && !tri) return 8.0 * batch * (double)M * N * K;
  double f = 0.0;
  for (int m0 = 0; m0 < M; m0 += tm)
    for (int n0 = 0; n0 < N; n0 += tm) {
      if (tri && n0 > m0 + tm - 1) continue;
      int kbeg = 0, kend = K;
      if (klim & KLIM_A_LOWER) kend = min(kend, m0 + tm);
      if (klim & KLIM_A_UPPER) kbeg = max(kbeg, m0);
      if (klim & KLIM_B_UPPER) kend = min(kend, n0 + tm);
      if (klim & KLIM_B_LOWER) kbeg = max(kbeg, n0);
      if (kend > kbeg) f += 8.0 * min(tm, M - m0) * min(tm, N - n0) * (kend - kbeg);
    }
  return f * batch;
}

// 64x64 tiles once they fill the device's 256 CUs, 32x32 below.
static int gemm_tile(int batch, int M, int N, int tri) {
  const long long tm = (M + GT - 1) / GT, tn = (N + GT - 1) / GT;
  long long tiles = tm * tn;
  if (tri) tiles = 0;
  if (tri)
    for (long long i = 0; i < tm; ++i) tiles += min(tn, i + 1);
  return tiles * batch >= 256 ? GT : 32;
}

void launch_zgemm(hipStream_t st, int batch, int M, int N, int K, double alpha, const double2* A, int lda,
                  long long sA, int opA, const double2* B, int ldb, long long sB, int opB, double beta, double2* C,
                  int ldc, long long sC, int klim, int tri, bool re, bool c_real) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  const int tm = gemm_tile(batch, M, N, tri);
  dim3 g((N + tm - 1) / tm, (M + tm - 1) / tm, batch);
  const int creal = re && c_real;
  SampleSlot sl;
  // a real product is a quarter of the complex flops
  const bool sampled = sampler_enabled() &&
                       sampler_take(SLK_SAMPLE_GEMM_F64,
                                    gemm_tile_flops(batch, M, N, K, klim, tri, tm) * (re ? 0.25 : 1.0), &sl);
  auto kern = re ? (tm == GT ? gemm_mfma_kernel<double, GT, true> : gemm_mfma_kernel<double, 32, true>)
                 : (tm == GT ? gemm_mfma_kernel<double, GT> : gemm_mfma_kernel<double, 32>);
  if (sampled)
    hipExtLaunchKernelGGL(kern, g, dim3(256), 0, st, sl.e0, sl.e1, 0, M, N, K, alpha, A, lda, sA, opA, B, ldb, sB,
                          opB, beta, C, ldc, sC, klim, tri, creal, sl.clk);
  else
    hipLaunchKernelGGL(kern, g, dim3(256), 0, st, M, N, K, alpha, A, lda, sA, opA, B, ldb, sB, opB, beta, C, ldc, sC,
                       klim, tri, creal, (unsigned long long*)nullptr);
}

void launch_cgemm(hipStream_t st, int batch, int M, int N, int K, float alpha, const float2* A, int lda,
                  long long sA, int opA, const float2* B, int ldb, long long sB, int opB, float beta, float2* C,
                  int ldc, long long sC, int klim, int tri, bool re, bool c_real) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  const int tm = gemm_tile(batch, M, N, tri);
  dim3 g((N + tm - 1) / tm, (M + tm - 1) / tm, batch);
  const int creal = re && c_real;
  auto kern = re ? (tm == GT ? gemm_mfma_kernel<float, GT, true> : gemm_mfma_kernel<float, 32, true>)
                 : (tm == GT ? gemm_mfma_kernel<float, GT> : gemm_mfma_kernel<float, 32>);
  hipLaunchKernelGGL(kern, g, dim3(256), 0, st, M, N, K, alpha, A, lda, sA, opA, B, ldb, sB, opB, beta, C, ldc, sC,
                     klim, tri, creal, (unsigned long long*)nullptr);
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
// Leaf of the Cholesky recursion: the m x m diagonal block at j0 (m <= LB, its
// lower part current) is factored, L_jj, and inverted, W_jj = L_jj^-1, by one
// workgroup of 1024 threads per matrix; thread t holds elements
// (t / LB + (1024 / LB) q, t % LB), q < LB^2 / 1024, of A and W in registers.
// Step k reads column k of A and row k of W, published in LDS at the end of
// step k-1, and each thread applies its part of
//   L[r][k] = A[r][k] / l_kk                       (r > k)
//   A[r][c] -= L[r][k] conj(L[c][k])               (k < c <= r)
//   W[k][:] /= l_kk,  W[r][:] -= L[r][k] W[k][:]   (r > k)
// the second line the Cholesky trailing update, the last the same row
// operations applied to the identity (forward substitution). The two LDS
// vectors alternate by step parity, so one barrier per step suffices.
// ---------------------------------------------------------------------------
#define LEAF_THREADS 1024
template <int LB>
__global__ __launch_bounds__(LEAF_THREADS) void potrf_diag_kernel(double2* L, double2* W, int m, int ld, int j0,
                                                                  int* info) {
  constexpr int QN = LB * LB / LEAF_THREADS, RS = LEAF_THREADS / LB;
  const int b = blockIdx.x;
  double2* a = L + (size_t)b * ld * ld;
  double2* w = W + (size_t)b * ld * ld;
  __shared__ double2 s_col[2][LB];
  __shared__ double2 s_row[2][LB];
  const int r0 = threadIdx.x / LB, c = threadIdx.x % LB;
  double2 x[QN], y[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    const int r = r0 + RS * q;
    const bool in = r < m && c < m;
    x[q] = (in && c <= r) ? a[(size_t)(j0 + r) * ld + j0 + c] : make_double2(0.0, 0.0);
    y[q] = make_double2(in && r == c ? 1.0 : 0.0, 0.0);
    if (c == 0) s_col[0][r] = x[q];
    if (r == 0) s_row[0][c] = y[q];
  }
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    const int p = k & 1;
    double d = s_col[p][k].x;
    if (!(d > 0.0)) {
      if (threadIdx.x == 0 && info[b] == 0) info[b] = j0 + k + 1;
      d = 1.0;
    }
    // one reciprocal square root per step instead of a square root and an
    // IEEE division on the step's serial path
    const double il = rsqrt(d), l = d * il;
    const double2 cc = s_col[p][c], wk = s_row[p][c];
    const double2 lc = make_double2(cc.x * il, cc.y * il);
    const double2 wkn = make_double2(wk.x * il, wk.y * il);
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int r = r0 + RS * q;
      const double2 cr = s_col[p][r];
      const double2 lr = make_double2(cr.x * il, cr.y * il);
      if (c == k) {
        if (r == k) x[q] = make_double2(l, 0.0);
        else if (r > k) x[q] = lr;
      } else if (c > k && r >= c) {
        x[q].x -= lr.x * lc.x + lr.y * lc.y;
        x[q].y -= lr.y * lc.x - lr.x * lc.y;
      }
      if (r == k) {
        y[q] = wkn;
      } else if (r > k) {
        y[q].x -= lr.x * wkn.x - lr.y * wkn.y;
        y[q].y -= lr.x * wkn.y + lr.y * wkn.x;
      }
      if (c == k + 1) s_col[p ^ 1][r] = x[q];
      if (r == k + 1) s_row[p ^ 1][c] = y[q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    const int r = r0 + RS * q;
    if (r < m && c < m) {
      const size_t off = (size_t)(j0 + r) * ld + j0 + c;
      if (c <= r) a[off] = x[q];
      w[off] = y[q];
    }
  }
}

// Cholesky S = L L^H together with W = L^-1, by recursive halving of the
// block [o, o+m): factor the leading n1 columns, L21 = A21 W11^H, the
// trailing update A22 -= L21 L21^H (lower tiles, K = n1), factor the
// trailing block, then W21 = -W22 (L21 W11). The NB x NB leaves are factored
// and inverted in LDS by potrf_diag_kernel. The GEMMs have K = n1 instead of
// the NB of a blocked right-looking loop. L is scratch here (only W is read
// afterwards): a level keeps its L21 compact (n2 x n1) at the start of its
// scratch T while the trailing recursion works after it, and L21 W11 goes
// into L's L21 block, which nothing reads any more. A level holds n2 n1 of T
// and the trailing recursion at most its own share: under 2/3 of ld * ld / 2
// per matrix.
static void chol_inv_rec(hipStream_t st, double2* L, double2* W, double2* T, int o, int m, int n, int ld, int batch,
                         long long sM, long long sT, int* info, bool re) {
  // 32-row leaves: a 64-row leaf (64 dependent steps in one workgroup) measured
  // 7% slower over the whole recursion than two 32-row leaves with the four
  // GEMMs between them
  if (m <= NB) {
    hipLaunchKernelGGL(potrf_diag_kernel<NB>, dim3(batch), dim3(LEAF_THREADS), 0, st, L, W, m, ld, o, info);
    return;
  }
  const int n1 = ((m + 1) / 2 + NB - 1) / NB * NB, n2 = m - n1;
  double2* A21 = L + (size_t)(o + n1) * ld + o;
  chol_inv_rec(st, L, W, T, o, n1, n, ld, batch, sM, sT, info, re);
  // L21 = A21 W11^H (W11^H upper: K ends at the column block), compact in T
  launch_zgemm(st, batch, n2, n1, n1, 1.0, A21, ld, sM, OP_N, W + (size_t)o * ld + o, ld, sM, OP_C, 0.0, T, n1, sT,
               KLIM_B_UPPER, 0, re);
  // A22 -= L21 L21^H over the lower tiles
  launch_zgemm(st, batch, n2, n2, n1, -1.0, T, n1, sT, OP_N, T, n1, sT, OP_C, 1.0, L + (size_t)(o + n1) * ld + o + n1,
               ld, sM, 0, 1, re);
  chol_inv_rec(st, L, W, T + (size_t)n2 * n1, o + n1, n2, n, ld, batch, sM, sT, info, re);
  // W21 = -W22 (L21 W11), the product in L's A21 block
  launch_zgemm(st, batch, n2, n1, n1, 1.0, T, n1, sT, OP_N, W + (size_t)o * ld + o, ld, sM, OP_N, 0.0, A21, ld, sM,
               KLIM_B_LOWER, 0, re);
  launch_zgemm(st, batch, n2, n1, n2, -1.0, W + (size_t)(o + n1) * ld + o + n1, ld, sM, OP_N, A21, ld, sM, OP_N, 0.0,
               W + (size_t)(o + n1) * ld + o, ld, sM, KLIM_A_LOWER, 0, re);
}

void launch_chol_inv(hipStream_t st, const double2* S, double2* L, double2* W, double2* T, int n, int ld, int batch,
                     int* info, bool re) {
  // L = S (lower part used), W = 0 (diagonal blocks filled by the leaves)
  (void)hipMemcpyAsync(L, S, (size_t)batch * ld * ld * sizeof(double2), hipMemcpyDeviceToDevice, st);
  launch_zero(st, W, (size_t)batch * ld * ld * sizeof(double2));
  chol_inv_rec(st, L, W, T, 0, n, n, ld, batch, (long long)ld * ld, (long long)ld * ld / 2, info, re);
}



extern "C" {

// Batched complex GEMM entry points (C-ABI), used by the eigensolver and
// exposed for verification: C = alpha opA(A) opB(B) + beta C.
int slk_zgemm(int batch, int M, int N, int K, double alpha, const void* A, int lda, long long sA, int opA,
              const void* B, int ldb, long long sB, int opB, double beta, void* C, int ldc, long long sC,
              hipStream_t stream) {
  launch_zgemm(stream, batch, M, N, K, alpha, reinterpret_cast<const double2*>(A), lda, sA, opA,
               reinterpret_cast<const double2*>(B), ldb, sB, opB, beta, reinterpret_cast<double2*>(C), ldc, sC, 0,
               0);
  return (int)hipGetLastError();
}

int slk_cgemm(int batch, int M, int N, int K, float alpha, const void* A, int lda, long long sA, int opA,
              const void* B, int ldb, long long sB, int opB, float beta, void* C, int ldc, long long sC,
              hipStream_t stream) {
  launch_cgemm(stream, batch, M, N, K, alpha, reinterpret_cast<const float2*>(A), lda, sA, opA,
               reinterpret_cast<const float2*>(B), ldb, sB, opB, beta, reinterpret_cast<float2*>(C), ldc, sC, 0, 0);
  return (int)hipGetLastError();
}

}  // extern "C"

__global__ void to_c64_kernel(const double2* __restrict__ src, float2* __restrict__ dst, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}

void launch_to_c64(hipStream_t st, const double2* src, float2* dst, size_t count) {
  if (count == 0) return;
  int blocks = (int)((count + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(to_c64_kernel, dim3(blocks), dim3(256), 0, st, src, dst, count);
}
