// Two-stage reduction of a batch of Hermitian matrices to real tridiagonal
// form for the eigenvalue path (FP64 or FP32 arithmetic; d, e always FP64).
//
// Stage 1, dense -> band of width BW = 16, per 16-column panel j0
// (rows r0 = j0 + BW .. n-1, m = n - r0; LAPACK zhetrd_he2hb structure),
// on the lower triangle of A only:
//   panel_qr   one workgroup per matrix: Householder QR of the m x BW panel,
//              wave w owning column w, lanes owning rows (in registers up to
//              a size tier); one barrier per column, the reflector published
//              in LDS double-buffered by column parity; the block reflector
//              factor T (Q = I - V T V^H) built by wave 0 one column behind.
//              R goes into the band, the reflectors below it.
//   band_av    (first panel) Y = A22 V on MFMA, 64 rows x 128-column chunks
//              (split K), A22 read as Hermitian from its lower triangle.
//   band_x     X = Y T, per 64-row block partial Z = V^H X.
//   band_w     W = X - V (T^H Z) / 2.
//   update     A22 -= V W^H + W V^H with look-ahead: the next panel's 16
//              columns first (strip_update), then the next panel's QR, then
//              the rest of A22 in one pass that also forms the next panel's
//              partials of A22 V (fused_update_av), on the 32 x 32 blocks
//              touching the lower triangle.
// Stage 2, band -> tridiagonal: band_chase, one workgroup (16 waves) per
// matrix on a compact copy of the band (column j holds A[j..j+2BW-1][j]),
// in LDS when it fits. Sweep s annihilates column s below the subdiagonal
// with one reflector and chases the bulge down in steps of BW rows; each step
// is one wave: reflector from the bulge column, left application to the bulge
// block, two-sided application to the diagonal block, right application to
// the block below. Sweep s may run step c once sweep s-1 has finished step
// c+2 (their regions are then disjoint); waves publish their progress in LDS
// and wait on their predecessor's counter, with a bounded wait.
#include "slk_common.h"
#include "slk_dense.h"
#include <hip/hip_ext.h>

#include <atomic>
#include <mutex>
#include <unordered_map>

#include "slk_sampler.h"

#define BW SLK_BW
#define AV_KT 64                      // k per LDS tile of band_av
#define AV_KCH 128                    // k per workgroup of band_av (split K)
#ifndef CH_THREADS
#define CH_THREADS 1024
#endif
#define CH_WAVES (CH_THREADS / 64)
#define BND_LD (2 * BW)               // compact band: entries per column

template <typename T>
struct BandCtx {
  using C2 = typename C2T<T>::t;
  C2* A;
  int n, ld;
  C2 *Vc, *Wc;   // [batch][2 BW][ld]: [V | W] and [W | V] (column k at Vc[k*ld + r]); X goes to Wc first
  C2* Yp;        // [batch][nch][ld][BW]: split-K partials of A22 V
  C2* Zp;        // [batch][nrb][BW][BW]: per row block partials of V^H X
  C2* TT;        // [batch][npan][BW][BW]: block reflector factor of each panel, row-major
  int npan;      // panels per matrix (stride of TT)
  C2* TAU;       // [batch][ld]: stage-1 tau of column j
  C2* VT;        // nullable [batch][ld][ld]: stage-1 reflector of column j as row j
  C2* BND;       // [batch][ld][BND_LD]: compact band for the chase
  C2* X16;       // [batch][BW][BW]: X of the first BW rows of the current A22, row-major
  double *D, *E;
  int nch, nrb;  // chunk / row-block counts used for the strides of Yp, Zp
  int* info;     // nullable [batch]: set to 2 when the chase gave up a bounded wait
  int spin_max;  // polls per wait before giving up; < 0: give up at the first wait
};

// Element e = r * ld + c of a matrix of A (a: the matrix's origin). A complex
// matrix stores it at a[e]; a real batch (RE) stores the real value alone at
// real index e of the same slot, so the stage-1 passes move half the bytes.
template <typename T, bool RE>
__device__ __forceinline__ typename C2T<T>::t a_ld(const typename C2T<T>::t* a, size_t e, bool ok) {
  if constexpr (RE) {
    T x = 0;
    if (ok) x = reinterpret_cast<const T*>(a)[e];
    return mk<T>(x, T(0));
  } else {
    return load_or_zero(a + e, ok);
  }
}
template <typename T, bool RE>
__device__ __forceinline__ void a_st(typename C2T<T>::t* a, size_t e, typename C2T<T>::t v) {
  if constexpr (RE) reinterpret_cast<T*>(a)[e] = v.x;
  else a[e] = v;
}

template <typename T> struct MF;
template <> struct MF<float> {
  using v4 = __attribute__((ext_vector_type(4))) float;
  static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output row of accumulator element v in lane group g (column = lane % 16)
  static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
};
template <> struct MF<double> {
  using v4 = __attribute__((ext_vector_type(4))) double;
  static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return g + 4 * v; }
};

template <typename C2>
__device__ __forceinline__ C2 sum16c(C2 v) {
  v.x = sum16(v.x);
  v.y = sum16(v.y);
  return v;
}
// Pair sums across rows (lane ^ 16) and halves (lane ^ 32) of a wave with the
// gfx950 permlane swaps (VALU, no LDS round trip): given v twice, the swap
// returns the lane's own value and its partner's, one in each result.
__device__ __forceinline__ unsigned u32_of(float v) { return __builtin_bit_cast(unsigned, v); }
__device__ __forceinline__ float f32_of(unsigned u) { return __builtin_bit_cast(float, u); }
template <int ROWS>
__device__ __forceinline__ float add_partner(float v) {
  const unsigned u = u32_of(v);
  if constexpr (ROWS == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return f32_of(p[0]) + f32_of(p[1]);
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return f32_of(p[0]) + f32_of(p[1]);
  }
}
template <int ROWS>
__device__ __forceinline__ double add_partner(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
  unsigned a0, a1, b0, b1;
  if constexpr (ROWS == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto q = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    a0 = p[0], a1 = p[1], b0 = q[0], b1 = q[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto q = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    a0 = p[0], a1 = p[1], b0 = q[0], b1 = q[1];
  }
  const double x = __builtin_bit_cast(double, ((unsigned long long)b0 << 32) | a0);
  const double y = __builtin_bit_cast(double, ((unsigned long long)b1 << 32) | a1);
  return x + y;
}
// Sums over the 4 lane groups of two values at once (3 swaps for 2 values):
// a row-pair swap of a against b leaves each row pair holding the pair sums
// of a in one row and of b in the other, the half swap completes both totals,
// and a last row-pair swap hands every lane both. Each total is
// (g0 + g1) + (g2 + g3).
__device__ __forceinline__ void sum4g_pair(float& a, float& b) {
  const auto p = __builtin_amdgcn_permlane16_swap(u32_of(a), u32_of(b), false, false);
  const float c = f32_of(p[0]) + f32_of(p[1]);
  const auto q = __builtin_amdgcn_permlane32_swap(u32_of(c), u32_of(c), false, false);
  const float d = f32_of(q[0]) + f32_of(q[1]);
  const auto r = __builtin_amdgcn_permlane16_swap(u32_of(d), u32_of(d), false, false);
  a = f32_of(r[0]);
  b = f32_of(r[1]);
}
__device__ __forceinline__ void sum4g_pair(double& a, double& b) {
  a = add_partner<32>(add_partner<16>(a));
  b = add_partner<32>(add_partner<16>(b));
}
template <typename C2>
__device__ __forceinline__ C2 sum4g_c(C2 v) {
  sum4g_pair(v.x, v.y);
  return v;
}

// Complex multiply-accumulate as explicit FMA chains (4 FMAs each instead of
// the products' 2 multiplies + 2 FMAs + 2 adds).
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }
// acc + a b
template <typename C2>
__device__ __forceinline__ C2 cfma(C2 a, C2 b, C2 acc) {
  acc.x = fmat(-a.y, b.y, fmat(a.x, b.x, acc.x));
  acc.y = fmat(a.y, b.x, fmat(a.x, b.y, acc.y));
  return acc;
}
// acc + conj(a) b
template <typename C2>
__device__ __forceinline__ C2 ccfma(C2 a, C2 b, C2 acc) {
  acc.x = fmat(a.y, b.y, fmat(a.x, b.x, acc.x));
  acc.y = fmat(-a.y, b.x, fmat(a.x, b.y, acc.y));
  return acc;
}
// e - a b
template <typename C2>
__device__ __forceinline__ C2 cfms(C2 a, C2 b, C2 e) {
  e.x = fmat(a.y, b.y, fmat(-a.x, b.x, e.x));
  e.y = fmat(-a.y, b.x, fmat(-a.x, b.y, e.y));
  return e;
}
// e - a conj(b)
template <typename C2>
__device__ __forceinline__ C2 cfmsc(C2 a, C2 b, C2 e) {
  e.x = fmat(-a.y, b.y, fmat(-a.x, b.x, e.x));
  e.y = fmat(a.x, b.y, fmat(-a.y, b.x, e.y));
  return e;
}

// The same products for a real step (RE: every operand has a zero imaginary
// part, which is not read; the result's stays as it is, zero).
template <bool RE, typename C2>
__device__ __forceinline__ C2 cfma_r(C2 a, C2 b, C2 acc) {
  if constexpr (RE) {
    acc.x = fmat(a.x, b.x, acc.x);
    return acc;
  } else {
    return cfma(a, b, acc);
  }
}
template <bool RE, typename C2>
__device__ __forceinline__ C2 ccfma_r(C2 a, C2 b, C2 acc) {
  if constexpr (RE) {
    acc.x = fmat(a.x, b.x, acc.x);
    return acc;
  } else {
    return ccfma(a, b, acc);
  }
}
template <bool RE, typename C2>
__device__ __forceinline__ C2 cfms_r(C2 a, C2 b, C2 e) {
  if constexpr (RE) {
    e.x = fmat(-a.x, b.x, e.x);
    return e;
  } else {
    return cfms(a, b, e);
  }
}
template <bool RE, typename C2>
__device__ __forceinline__ C2 cfmsc_r(C2 a, C2 b, C2 e) {
  if constexpr (RE) {
    e.x = fmat(-a.x, b.x, e.x);
    return e;
  } else {
    return cfmsc(a, b, e);
  }
}

// value of lane 0 (v_readlane: a scalar read, no LDS round trip)
__device__ __forceinline__ float lane0(float v) { return f32_of((unsigned)__builtin_amdgcn_readlane((int)u32_of(v), 0)); }
__device__ __forceinline__ double lane0(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, 0);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 0);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Householder reflector of x = (alpha, rest) with |rest|^2 = xn2 (LAPACK
// zlarfg convention H^H x = beta e0, H = I - tau v v^H, v = (1, rest * scale)).
template <typename T>
struct House {
  typename C2T<T>::t tau, scale;
  T beta;
  bool skip;
};
template <typename T>
__device__ __forceinline__ House<T> house(typename C2T<T>::t alpha, T xn2) {
  House<T> h;
  if (xn2 == T(0) && alpha.y == T(0)) {
    h.skip = true;
    h.beta = alpha.x;
    h.tau = mk<T>(0, 0);
    h.scale = mk<T>(0, 0);
    return h;
  }
  h.skip = false;
  const T nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
  const T bt = alpha.x >= T(0) ? -nrm : nrm;
  h.beta = bt;
  h.tau = mk<T>((bt - alpha.x) / bt, -alpha.y / bt);
  const T dx = alpha.x - bt, dy = alpha.y, dd = dx * dx + dy * dy;
  h.scale = mk<T>(dx / dd, -dy / dd);
  return h;
}

// ---------------------------------------------------------------------------
// Stage 1, panel QR: 16 waves, wave w owns column w of the m x BW panel,
// lane l owns rows l + 64 q (registers for q < RPT, LDS for the QT rows after
// those; RPT = 0 re-reads the panel from global memory). Column k: wave k reduces its norm, forms the
// reflector and publishes v_k (LDS, or column k of Vc for RPT = 0); after one
// barrier, waves j > k apply it to their columns (inner product and update
// within the wave), wave k+1 also forms the next norm, and waves l < k form
// g_l = v_l^H v_k for the block reflector factor T.
// Wave sums: within rows of 16 lanes, then across rows and halves with the
// permlane swaps (VALU, no LDS round trip).
template <typename T>
__device__ __forceinline__ T wave_sum_r(T v) {
  return add_partner<32>(add_partner<16>(sum16(v)));
}
template <typename C2>
__device__ __forceinline__ C2 wave_sum_c(C2 v) {
  v.x = wave_sum_r(v.x);
  v.y = wave_sum_r(v.y);
  return v;
}

// Calls f(q) for the rows q < nq of a lane: fully unrolled in chunks of 8 with
// a scheduling barrier between chunks for register-resident rows (bounds the
// LDS loads the compiler hoists), a plain loop otherwise.
template <int RPT, typename F>
__device__ __forceinline__ void for_rows(int nq, F f) {
  if constexpr (RPT > 0) {
#pragma unroll
    for (int q0 = 0; q0 < RPT; q0 += 8) {
#pragma unroll
      for (int qq = 0; qq < 8; ++qq)
        if (q0 + qq < RPT) f(q0 + qq);
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    for (int q = 0; q < nq; ++q) f(q);
  }
}

template <typename T, int RPT, int QT = 0, bool RE = false>
__global__ __launch_bounds__(1024) void panel_qr_kernel(BandCtx<T> c, int j0) {
  using C2 = typename C2T<T>::t;
  static_assert(QT == 0 || RPT > 0, "LDS rows extend register rows");
  constexpr int NR = RPT > 0 ? RPT : 1;
  constexpr int FR = RPT > 0 ? RPT + QT : 0;   // row slots per lane (0: global)
  const int b = blockIdx.x;
  const int n = c.n, ld = c.ld;
  const int r0 = j0 + BW, m = n - r0;
  const int kb = m < BW ? m : BW;
  C2* a = c.A + (size_t)b * ld * ld;
  C2* vcb = c.Vc + (size_t)b * 2 * BW * ld;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // published reflector, double-buffered by column parity: wave k+1 writes
  // v_{k+1} while other waves may still read v_k (one barrier per column)
  __shared__ C2 s_vb[2][RPT > 0 ? 64 * FR : 1];
  __shared__ C2 s_tau[BW];
  __shared__ C2 s_G[BW][BW];   // s_G[k][l] = v_l^H v_k, l < k
  __shared__ C2 s_T[BW][BW];
  // rows lane + 64 (RPT + t) of column w at s_tl[t][lane][w]
  __shared__ C2 s_tl[QT > 0 ? QT : 1][QT > 0 ? 64 : 1][BW + 1];
  const int nq = RPT > 0 ? FR : (m + 63) / 64;
  const size_t col = (size_t)(r0 + lane) * ld + j0 + w;   // this lane's rows of column w
  C2 p[NR];
  if constexpr (RPT > 0) {
#pragma unroll
    for (int q = 0; q < RPT; ++q) p[q] = a_ld<T, RE>(a, col + (size_t)64 * q * ld, lane + 64 * q < m);
#pragma unroll
    for (int t = 0; t < QT; ++t)
      s_tl[t][lane][w] = a_ld<T, RE>(a, col + (size_t)64 * (RPT + t) * ld, lane + 64 * (RPT + t) < m);
  }
  // q is a compile-time constant in the unrolled row loops (RPT > 0)
  auto get = [&](int q) -> C2 {
    if constexpr (RPT > 0) {
      if (QT > 0 && q >= RPT) return s_tl[q - RPT][lane][w];
      return p[q];
    } else {
      return a_ld<T, RE>(a, col + (size_t)64 * q * ld, lane + 64 * q < m);
    }
  };
  auto put = [&](int q, C2 v) {
    if constexpr (RPT > 0) {
      if (QT > 0 && q >= RPT) s_tl[q - RPT][lane][w] = v;
      else p[q] = v;
    } else if (lane + 64 * q < m) {
      a_st<T, RE>(a, col + (size_t)64 * q * ld, v);
    }
  };
  // T (Q = I - V T V^H): T[k][k] = tau_k, T[i][k] = -tau_k sum_{l=i}^{k-1} T[i][l] g_l^(k);
  // lane i of wave 0 owns row i
  if (tid < BW * BW) s_T[tid / BW][tid % BW] = mk<T>(0, 0);
  auto t_column = [&](int k) {
    if (k < 0 || lane > k) return;
    if (lane == k) {
      s_T[k][k] = s_tau[k];
      return;
    }
    C2 acc = mk<T>(0, 0);
#pragma unroll
    for (int l = 0; l < BW; ++l)
      if (l >= lane && l < k) acc = c_add(acc, c_mul(s_T[lane][l], s_G[k][l]));
    const C2 t = c_mul(s_tau[k], acc);
    s_T[lane][k] = mk<T>(-t.x, -t.y);
  };
  T nrm = 0;
  if (w == 0)
    for_rows<FR>(nq, [&](int q) {
        const int r = lane + 64 * q;
        const C2 e = get(q);
        if (r > 0 && r < m) nrm += RE ? e.x * e.x : e.x * e.x + e.y * e.y;
      });
  for (int k = 0; k < kb; ++k) {
    const C2* vk = vcb + (size_t)k * ld + r0;   // global copy of v_k (RPT = 0)
    C2* s_v = s_vb[k & 1];
    if (w == k) {
      const T xn2 = wave_sum_r(nrm);
      C2 a0 = get(0);   // row k lives in lane k, q = 0 (k < BW <= 64)
      const C2 alpha = mk<T>(__shfl(a0.x, k, 64), RE ? T(0) : __shfl(a0.y, k, 64));
      const House<T> h = house<T>(alpha, xn2);
      for_rows<FR>(nq, [&](int q) {
          const int r = lane + 64 * q;
          if (r >= m) return;
          C2 e = get(q);
          C2 vv = mk<T>(0, 0);
          if (r == k) {
            vv = mk<T>(1, 0);
            put(q, mk<T>(h.beta, 0));
          } else if (r > k) {
            vv = h.skip ? mk<T>(0, 0) : (RE ? mk<T>(e.x * h.scale.x, T(0)) : c_mul(e, h.scale));
            put(q, vv);
          }
          if constexpr (RPT > 0) s_v[r] = vv;
          else const_cast<C2*>(vk)[r] = vv;
        });
      if (lane == 0) s_tau[k] = h.tau;
    }
    __syncthreads();
    if (w > k) {
      const C2 tau = s_tau[k];
      C2 u = mk<T>(0, 0);
      for_rows<FR>(nq, [&](int q) {
          const int r = lane + 64 * q;
          if (r >= k && r < m) {
            const C2 vr = RPT > 0 ? s_v[r] : vk[r];
            if constexpr (RE) {
              u.x = fmat(vr.x, get(q).x, u.x);
            } else {
              const C2 t = c_cmul(vr, get(q));
              u.x += t.x;
              u.y += t.y;
            }
          }
        });
      if constexpr (RE) u.x = wave_sum_r(u.x);
      else u = wave_sum_c(u);
      const C2 cu = RE ? mk<T>(tau.x * u.x, T(0)) : c_mul(c_conj(tau), u);
      nrm = 0;
      for_rows<FR>(nq, [&](int q) {
          const int r = lane + 64 * q;
          if (r >= k && r < m) {
            const C2 vr = RPT > 0 ? s_v[r] : vk[r];
            C2 e = get(q);
            if constexpr (RE) {
              e.x = fmat(-vr.x, cu.x, e.x);
            } else {
              const C2 t = c_mul(vr, cu);
              e.x -= t.x;
              e.y -= t.y;
            }
            put(q, e);
            if (r > k + 1) nrm += RE ? e.x * e.x : e.x * e.x + e.y * e.y;
          }
        });
    } else if (w < k) {
      // g_w = v_w^H v_k over rows >= k (column w holds v_w below row w)
      C2 g = mk<T>(0, 0);
      for_rows<FR>(nq, [&](int q) {
          const int r = lane + 64 * q;
          if (r >= k && r < m) {
            const C2 vr = RPT > 0 ? s_v[r] : vk[r];
            if constexpr (RE) {
              g.x = fmat(get(q).x, vr.x, g.x);
            } else {
              const C2 t = c_cmul(get(q), vr);
              g.x += t.x;
              g.y += t.y;
            }
          }
        });
      if constexpr (RE) g.x = wave_sum_r(g.x);
      else g = wave_sum_c(g);
      if (lane == 0) s_G[k][w] = g;
      // wave 0 meanwhile builds column k-1 of T (its g values are complete)
      if (w == 0) t_column(k - 1);
    }
  }
  __syncthreads();
  if (w == 0) t_column(kb - 1);
  __syncthreads();
  // outputs: panel column w (R in the band, v_w below), V in Vc[w] and Wc[BW + w], VT row j0 + w
  C2* wcv = c.Wc + (size_t)b * 2 * BW * ld + (size_t)BW * ld;
  for_rows<FR>(nq, [&](int q) {
      const int r = lane + 64 * q;
      if (r >= m) return;
      const C2 e = get(q);
      if constexpr (RPT > 0) a_st<T, RE>(a, col + (size_t)64 * q * ld, e);
      C2 v = mk<T>(0, 0);
      if (w < kb) v = r > w ? e : (r == w ? mk<T>(1, 0) : mk<T>(0, 0));
      vcb[(size_t)w * ld + r0 + r] = v;
      wcv[(size_t)w * ld + r0 + r] = v;
      if (c.VT != nullptr && w < kb && r >= w) c.VT[(size_t)b * ld * ld + (size_t)(j0 + w) * ld + r0 + r] = v;
    });
  if (tid < BW) c.TAU[(size_t)b * ld + j0 + tid] = load_or_zero(s_tau + tid, tid < kb);
  if (tid < BW * BW) c.TT[((size_t)b * c.npan + j0 / BW) * BW * BW + tid] = s_T[tid / BW][tid % BW];
}

// ---------------------------------------------------------------------------
// Stage 1, Y = A22 V: workgroup (row block, k chunk, matrix), 4 waves of 16
// rows x 16 columns, complex product as 4 real 16x16x4 MFMAs per k-step.
// The next 64-column tile is fetched into registers while the current one is
// multiplied from LDS.
template <typename T, bool RE = false>
__global__ __launch_bounds__(256, 3) void band_av_kernel(BandCtx<T> c, int r0) {
  using C2 = typename C2T<T>::t;
  using V4 = typename MF<T>::v4;
  const int n = c.n, ld = c.ld;
  const int b = blockIdx.z;
  const int row0 = r0 + blockIdx.x * 64;
  const int kbeg = r0 + blockIdx.y * AV_KCH;
  const int kend = min(kbeg + AV_KCH, n);
  const C2* a = c.A + (size_t)b * ld * ld;
  const C2* vc = c.Vc + (size_t)b * 2 * BW * ld;
  __shared__ T Ar[64][AV_KT + 1], Ai[64][AV_KT + 1];
  __shared__ T Br[AV_KT][BW + 1], Bi[AV_KT][BW + 1];
  __shared__ T s_dump[2][64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  constexpr int NA = 64 * AV_KT / 256, NBT = AV_KT * BW / 256;
  C2 ra[NA], rb[NBT];
  // Only the lower triangle of A22 is current (the rank-2k updates skip the
  // tiles above the diagonal); element (r, k) with k > r is conj(A[k][r]).
  // Row blocks and k tiles share the origin r0, so a 64 x 64 tile is wholly
  // below the diagonal (read as stored), wholly above (fetched transposed,
  // coalesced along r, stored transposed into LDS: mode 1) or a diagonal
  // tile (mode 2: its lower half is fetched and also stored mirrored). The
  // conjugation happens at the LDS store: no arithmetic on a fetched value
  // before the multiply of the current tile, so the fetch stays in flight.
  static_assert(AV_KT == 64, "the tile modes assume square 64 x 64 tiles");
  int mode = 0;
  auto fetch = [&](int k0) {
    mode = k0 >= row0 + 64 ? 1 : (k0 == row0 ? 2 : 0);
    // memory element (R + hi, Cc + lo), hi = e / 64, lo = e % 64: one
    // address stream for all modes (R, Cc swapped for a transposed tile),
    // clamped into the matrix so that every load is unconditional; elements
    // outside the tile's valid part are zeroed at the LDS store
    const int R = mode == 1 ? k0 : row0, Cc = mode == 1 ? row0 : k0;
    const int limR = mode == 1 ? kend : n, limC = mode == 1 ? n : kend;
    const int h0 = tid / 64, lo = tid % 64;
    const size_t p = min(Cc + lo, limC - 1);
#pragma unroll
    for (int i = 0; i < NA; ++i) ra[i] = a_ld<T, RE>(a, p + (size_t)min(R + h0 + 4 * i, limR - 1) * ld, true);
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
      const int e = tid + 256 * i;
      rb[i] = vc[(size_t)(e / AV_KT) * ld + min(k0 + e % AV_KT, kend - 1)];
    }
  };
  V4 accR = V4{0, 0, 0, 0}, accI = V4{0, 0, 0, 0};
  fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += AV_KT) {
    {
      // branch-free stores (hi is uniform over a wave, lo is the lane):
      // element (hi, lo) goes to (hi, lo), or (lo, hi) conjugated for a
      // transposed tile; the upper half of a diagonal tile is not stored
      // (its place takes the mirror of the lower half) and goes to a
      // per-lane dump slot instead
      const bool tr = mode == 1;
      const int R = tr ? k0 : row0, Cc = tr ? row0 : k0;
      const int limR = tr ? kend : n, limC = tr ? n : kend;
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int hi = wv + 4 * i, lo = lane;
        const bool ok = R + hi < limR && Cc + lo < limC;
        const T x = ok ? ra[i].x : T(0), y = ok ? ra[i].y : T(0);
        const bool keep = mode != 2 || lo <= hi;
        const int o = tr ? lo * (AV_KT + 1) + hi : hi * (AV_KT + 1) + lo;
        T* pr = keep ? &Ar[0][0] + o : &s_dump[0][lane];
        T* pi = keep ? &Ai[0][0] + o : &s_dump[1][lane];
        *pr = x;
        *pi = tr ? -y : y;
      }
      if (mode == 2) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
          const int hi = wv + 4 * i, lo = lane;
          const bool ok = R + hi < limR && Cc + lo < limC;
          const bool mir = lo < hi;
          const int o = lo * (AV_KT + 1) + hi;
          T* pr = mir ? &Ar[0][0] + o : &s_dump[0][lane];
          T* pi = mir ? &Ai[0][0] + o : &s_dump[1][lane];
          *pr = ok ? ra[i].x : T(0);
          *pi = ok ? -ra[i].y : T(0);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
      const int e = tid + 256 * i;
      const bool ok = k0 + e % AV_KT < kend;
      Br[e % AV_KT][e / AV_KT] = ok ? rb[i].x : T(0);
      Bi[e % AV_KT][e / AV_KT] = ok ? rb[i].y : T(0);
    }
    __syncthreads();
    if (k0 + AV_KT < kend) fetch(k0 + AV_KT);
#pragma unroll
    for (int k4 = 0; k4 < AV_KT; k4 += 4) {
      const T ar = Ar[16 * wv + li][k4 + lk], ai = Ai[16 * wv + li][k4 + lk];
      const T br = Br[k4 + lk][li], bi = Bi[k4 + lk][li];
      accR = MF<T>::mfma(ar, br, accR);
      accR = MF<T>::mfma(-ai, bi, accR);
      accI = MF<T>::mfma(ar, bi, accI);
      accI = MF<T>::mfma(ai, br, accI);
    }
    __syncthreads();
  }
  C2* yp = c.Yp + ((size_t)b * c.nch + blockIdx.y) * ld * BW;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int gr = row0 + 16 * wv + MF<T>::row(lk, v);
    if (gr < n) a_st<T, RE>(yp, (size_t)gr * BW + li, mk<T>(accR[v], accI[v]));
  }
}

// Stage 1, X = Y T for a block of 64 rows, and the block's partial V^H X.
// RE: the Y partials are packed real values (element e of a (matrix, chunk)
// slot at real index e).
template <typename T, bool RE = false>
__global__ __launch_bounds__(256) void band_x_kernel(BandCtx<T> c, int r0, int nchk) {
  using C2 = typename C2T<T>::t;
  const int n = c.n, ld = c.ld;
  const int b = blockIdx.y;
  const int row0 = r0 + blockIdx.x * 64;
  __shared__ C2 s_T[BW][BW];
  __shared__ C2 s_y[64][BW + 1], s_v[64][BW + 1], s_x[64][BW + 1];
  const int tid = threadIdx.x;
  s_T[tid / BW][tid % BW] = c.TT[((size_t)b * c.npan + (r0 - BW) / BW) * BW * BW + tid];
  const int rl = tid >> 2, c0 = (tid & 3) * 4;
  const int gr = row0 + rl;
  const C2* vc = c.Vc + (size_t)b * 2 * BW * ld;
  C2* wc = c.Wc + (size_t)b * 2 * BW * ld;
  // the 4 columns of this thread are adjacent: one 4-element load per chunk,
  // 8 chunks in flight
  {
    C2 y[4], v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      y[q] = mk<T>(0, 0);
      v[q] = mk<T>(0, 0);
    }
    if (gr < n) {
      // V rows first: in flight together with the partials
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = vc[(size_t)(c0 + q) * ld + gr];
      const size_t cstride = (size_t)ld * BW;
      if constexpr (RE) {
        // 4 adjacent reals: one 16-byte load per chunk
        const T* yp = reinterpret_cast<const T*>(c.Yp + (size_t)b * c.nch * cstride) + (size_t)gr * BW + c0;
#pragma unroll 8
        for (int ch = 0; ch < nchk; ++ch) {
          const T* p = yp + 2 * ch * cstride;
#pragma unroll
          for (int q = 0; q < 4; ++q) y[q].x += p[q];
        }
      } else {
        const C2* yp = c.Yp + ((size_t)b * c.nch * ld + gr) * BW + c0;
#pragma unroll 8
        for (int ch = 0; ch < nchk; ++ch) {
          const C2* p = yp + ch * cstride;
#pragma unroll
          for (int q = 0; q < 4; ++q) y[q] = c_add(y[q], p[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      s_y[rl][c0 + q] = y[q];
      s_v[rl][c0 + q] = v[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int cc = c0 + q;
    C2 x = mk<T>(0, 0);
#pragma unroll
    for (int i = 0; i < BW; ++i) {
      const C2 t = c_mul(s_y[rl][i], s_T[i][cc]);
      x.x += t.x;
      x.y += t.y;
    }
    s_x[rl][cc] = x;
    if (gr < n) wc[(size_t)cc * ld + gr] = x;
    // the strip rows of X, read by every band_w workgroup while workgroup 0
    // replaces them in Wc by W
    if (blockIdx.x == 0 && rl < BW) c.X16[(size_t)b * BW * BW + rl * BW + cc] = x;
  }
  __syncthreads();
  const int i = tid / BW, kk = tid % BW;
  C2 z = mk<T>(0, 0);
  for (int r = 0; r < 64; ++r) {
    const C2 t = c_cmul(s_v[r][i], s_x[r][kk]);
    z.x += t.x;
    z.y += t.y;
  }
  c.Zp[((size_t)b * c.nrb + blockIdx.x) * BW * BW + tid] = z;
}

// Stage 1, W = X - V M / 2 with M = T^H (V^H X) (Hermitian), 64 rows per
// block, and the strip of the next panel: the first min(m, BW) columns of
// A22 (origin r0, order m = n - r0) take, on this block's rows, the rank-2k
// update -= V W_s^H + W V_s^H with V_s, W_s the strip's own rows of V and W.
// Every workgroup forms W_s = X_s - V_s M / 2 itself from the strip rows X_s
// that band_x saved, so the next panel's QR can follow this launch directly.
template <typename T, bool RE = false>
__global__ __launch_bounds__(256) void band_w_kernel(BandCtx<T> c, int r0, int nrbk) {
  using C2 = typename C2T<T>::t;
  const int n = c.n, ld = c.ld;
  const int b = blockIdx.y;
  const int row0 = r0 + blockIdx.x * 64;
  const int m = n - r0, ncol = m < BW ? m : BW;
  __shared__ C2 s_T[BW][BW], s_Z[BW][BW], s_M[BW][BW];
  __shared__ C2 s_v[64][BW + 1], s_w[64][BW + 1];
  __shared__ C2 s_vs[BW][BW + 1], s_ws[BW][BW + 1];
  const int tid = threadIdx.x;
  const int i = tid / BW, kk = tid % BW;
  const int rl = tid >> 2, c0 = (tid & 3) * 4;
  const int gr = row0 + rl;
  const C2* vc = c.Vc + (size_t)b * 2 * BW * ld;
  C2* wc = c.Wc + (size_t)b * 2 * BW * ld;
  // the loads that do not depend on the Zp sum go first, so that they are in
  // flight together with it: T, this block's V and X rows, the strip row i of
  // V (column kk) and of X
  const C2 tt = c.TT[((size_t)b * c.npan + (r0 - BW) / BW) * BW * BW + tid];
  C2 vr[4], xr[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    vr[q] = gr < n ? vc[(size_t)(c0 + q) * ld + gr] : mk<T>(0, 0);
    xr[q] = gr < n ? wc[(size_t)(c0 + q) * ld + gr] : mk<T>(0, 0);
  }
  const C2 vs = i < m ? vc[(size_t)kk * ld + r0 + i] : mk<T>(0, 0);
  const C2 xs = i < m ? c.X16[(size_t)b * BW * BW + tid] : mk<T>(0, 0);
  C2 z = mk<T>(0, 0);
  // unrolled: eight independent loads in flight per step of the (unchanged) sum order
#pragma unroll 8
  for (int q = 0; q < nrbk; ++q) z = c_add(z, c.Zp[((size_t)b * c.nrb + q) * BW * BW + tid]);
  s_Z[i][kk] = z;
  s_T[i][kk] = tt;
#pragma unroll
  for (int q = 0; q < 4; ++q) s_v[rl][c0 + q] = vr[q];
  s_vs[i][kk] = vs;
  __syncthreads();
  C2 mm = mk<T>(0, 0);
#pragma unroll
  for (int l = 0; l < BW; ++l) {
    const C2 t = c_cmul(s_T[l][i], s_Z[l][kk]);
    mm.x += t.x;
    mm.y += t.y;
  }
  s_M[i][kk] = mm;
  __syncthreads();
  {
    // strip row i of W
    C2 x = xs;
#pragma unroll
    for (int l = 0; l < BW; ++l) {
      const C2 t = c_mul(s_vs[i][l], s_M[l][kk]);
      x.x -= T(0.5) * t.x;
      x.y -= T(0.5) * t.y;
    }
    s_ws[i][kk] = x;
  }
  C2* vw = c.Vc + (size_t)b * 2 * BW * ld + (size_t)BW * ld;   // W also goes after V
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int cc = c0 + q;
    C2 x = xr[q];
#pragma unroll
    for (int l = 0; l < BW; ++l) {
      const C2 t = c_mul(s_v[rl][l], s_M[l][cc]);
      x.x -= T(0.5) * t.x;
      x.y -= T(0.5) * t.y;
    }
    s_w[rl][cc] = x;
    if (gr < n) {
      wc[(size_t)cc * ld + gr] = x;
      vw[(size_t)cc * ld + gr] = x;
    }
  }
  __syncthreads();
  if (gr >= n) return;
  // the strip: A[gr][r0 + j] -= sum_k V[gr][k] conj(W_s[j][k]) + W[gr][k] conj(V_s[j][k])
  C2* a = c.A + (size_t)b * ld * ld;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = c0 + q;
    if (j >= ncol) continue;
    const size_t e = (size_t)gr * ld + r0 + j;
    C2 acc = a_ld<T, RE>(a, e, true);
#pragma unroll
    for (int k = 0; k < BW; ++k) {
      if constexpr (RE) {
        acc.x -= s_v[rl][k].x * s_ws[j][k].x + s_w[rl][k].x * s_vs[j][k].x;
      } else {
        const C2 t1 = c_cmul(s_ws[j][k], s_v[rl][k]);
        const C2 t2 = c_cmul(s_vs[j][k], s_w[rl][k]);
        acc.x -= t1.x + t2.x;
        acc.y -= t1.y + t2.y;
      }
    }
    a_st<T, RE>(a, e, acc);
  }
}

// ---------------------------------------------------------------------------
// Stage 2, bulge chasing on the compact band (element (r, col) of the lower
// band at bnd[col * BND_LD + r - col], r - col < BND_LD).
//
// One step (one wave; lane = (li, lg), li = lane % 16, lg = lane / 16), with
// reflector rows q1..q2 (L = q2 - q1 + 1 <= BW) from column c1:
//   (a) reflector of A[q1..q2][c1]; beta stays, the rest becomes 0;
//   (b) left H^H on the bulge block rows q1..q2, columns c1+1 .. q1-1
//       (lane: column c1+1+li, rows lg+4t);
//   (c) two-sided H^H D H on the diagonal block rows/columns q1..q2
//       (lane: row li, columns lg+4t);
//   (d) right H on rows q2+1 .. q2+BW, columns q1..q2 (the next bulge)
//       (lane: row q2+1+li, columns lg+4t).
// v and w reach the lanes that need them by index lg+4t through a per-wave
// LDS slot sv[2 BW]. FULL: L = BW, BW-1 bulge columns and BW rows below
// (compile-time bounds).
template <typename T, bool FULL>
__device__ __forceinline__ T chase_step(typename C2T<T>::t* bnd, typename C2T<T>::t* v2row,
                                        typename C2T<T>::t* sv, int n, int c1, int q1, int q2, int lane) {
  using C2 = typename C2T<T>::t;
  const int li = lane & 15, lg = lane >> 4;
  const int L = FULL ? BW : q2 - q1 + 1;
  const int Le = FULL ? BW : min(BW, n - 1 - q2);
  const int nbc = FULL ? BW - 1 : q1 - c1 - 1;
  const C2 zero = mk<T>(0, 0);
  // every load of the step first
  const C2 x = load_or_zero(bnd + c1 * BND_LD + q1 - c1 + li, FULL || li < L);
  C2 eb[4], ed[4], ee[4];
  const int cb = c1 + 1 + li;   // this lane's bulge column
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = lg + 4 * t;   // row (b) / column (c), (d) index
    eb[t] = load_or_zero(bnd + cb * BND_LD + q1 + r - cb, (FULL || r < L) && li < nbc);
    const bool lo = r <= li;
    const int od = lo ? (q1 + r) * BND_LD + li - r : (q1 + li) * BND_LD + r - li;
    C2 d = load_or_zero(bnd + od, FULL || (li < L && r < L));
    if (!lo) d.y = -d.y;
    ed[t] = d;
    ee[t] = load_or_zero(bnd + (q1 + r) * BND_LD + q2 + 1 + li - q1 - r, (FULL || r < L) && (FULL || li < Le));
  }
  // (a)
  const T xn2 = sum16((li >= 1 && (FULL || li < L)) ? x.x * x.x + x.y * x.y : T(0));
  const C2 alpha = mk<T>(lane0(x.x), lane0(x.y));
  const House<T> h = house<T>(alpha, xn2);
  const C2 tau = h.tau;
  const C2 v = li == 0 ? mk<T>(1, 0) : (((FULL || li < L) && !h.skip) ? c_mul(x, h.scale) : zero);
  if (lg == 0) {
    if (FULL || li < L) {
      bnd[c1 * BND_LD + q1 - c1 + li] = li == 0 ? mk<T>(h.beta, 0) : zero;
      if (v2row != nullptr) v2row[q1 + li] = li == 0 ? tau : v;
    }
    sv[li] = v;
  }
  __builtin_amdgcn_wave_barrier();
  C2 vr[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) vr[t] = sv[lg + 4 * t];
  // (b) d_cb = sum_r conj(v_r) B[r][cb]; B[r][cb] -= conj(tau) v_r d_cb
  C2 db = zero;
#pragma unroll
  for (int t = 0; t < 4; ++t) db = ccfma(vr[t], eb[t], db);
  // (c) y_li = sum_c D[li][c] v_c;  (d) z_li = sum_c E[li][c] v_c
  C2 y = zero, z = zero;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    y = cfma(ed[t], vr[t], y);
    z = cfma(ee[t], vr[t], z);
  }
  db = sum4g_c(db);
  y = sum4g_c(y);
  z = sum4g_c(z);
  const C2 fb = c_mul(c_conj(tau), db);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = lg + 4 * t;
    if ((FULL || r < L) && li < nbc) bnd[cb * BND_LD + q1 + r - cb] = cfms(vr[t], fb, eb[t]);
  }
  const T vhy = sum16((FULL || li < L) ? v.x * y.x + v.y * y.y : T(0));
  const T t2 = T(0.5) * (tau.x * tau.x + tau.y * tau.y) * vhy;
  const C2 ty = c_mul(tau, y);
  const C2 w = (FULL || li < L) ? mk<T>(ty.x - t2 * v.x, ty.y - t2 * v.y) : zero;
  if (lg == 0) sv[BW + li] = w;
  __builtin_amdgcn_wave_barrier();
  const C2 tz = c_mul(tau, z);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = lg + 4 * t;
    const C2 wr = sv[BW + r];
    if ((FULL || li < L) && r <= li) {
      C2 e = cfmsc(v, wr, cfmsc(w, vr[t], ed[t]));
      if (r == li) e.y = T(0);
      bnd[(q1 + r) * BND_LD + li - r] = e;
    }
    if ((FULL || r < L) && (FULL || li < Le))
      bnd[(q1 + r) * BND_LD + q2 + 1 + li - q1 - r] = cfmsc(tz, vr[t], ee[t]);
  }
  return h.beta;
}

// Per-lane constant parts of the element offsets of a full step: every offset
// is a per-step uniform base plus one of these (the band is column-major with
// BND_LD entries per column, so element (r, col) sits at 31 col + r + col).
// Per-lane parts of the element offsets of a full step. Every offset is a
// per-step uniform base plus a lane part; the band is column-major with BND_LD
// entries per column, so element (r, col) sits at 31 col + r + col. For row
// r = lg + 4t of the lane's column li, the bulge and below-block offsets are
// ob0 + 4t and oe0 + 124t (constant steps, folded into the instructions'
// immediate offsets); the diagonal block reads the lower element, which is
// one of the two. Offsets are applied in bytes as unsigned 32-bit values, so
// that a global band is addressed as a uniform base plus one 32-bit VGPR.
struct ChaseLane {
  unsigned ob0;   // bulge block, column c1+1+li, row lg: base 31 (c1+1) + q1
  unsigned oe0;   // block below, row q2+1+li, column q1+lg: base 32 q1 + BW
  int li, lg;
};
__device__ __forceinline__ bool cl_up(const ChaseLane& o, int t) { return o.lg + 4 * t > o.li; }
__device__ __forceinline__ unsigned cl_ob(const ChaseLane& o, int t) { return o.ob0 + 4u * t; }
__device__ __forceinline__ unsigned cl_oe(const ChaseLane& o, int t) { return o.oe0 + 124u * t; }
// diagonal block (lower element of (li, lg+4t)): base 32 q1
__device__ __forceinline__ unsigned cl_od(const ChaseLane& o, int t) { return cl_up(o, t) ? cl_ob(o, t) : cl_oe(o, t); }
template <typename C2>
__device__ __forceinline__ C2& at_off(C2* base, unsigned el) {
  return *reinterpret_cast<C2*>(reinterpret_cast<char*>(base) + el * (unsigned)sizeof(C2));
}
__device__ __forceinline__ ChaseLane chase_lane(int lane) {
  ChaseLane o;
  o.li = lane & 15;
  o.lg = lane >> 4;
  o.ob0 = 31u * o.li + o.lg;
  o.oe0 = 31u * o.lg + o.li;
  return o;
}

// Reciprocal for the FP32 reflector (one-ulp hardware reciprocal; the FP64
// path keeps the IEEE division).
__device__ __forceinline__ float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_t(double x) { return 1.0 / x; }

// FP32 global band through buffer instructions: a uniform descriptor for the
// matrix's band, one per-step uniform byte offset per block (soffset) and the
// lane's constant byte offset (voffset; the t strides become immediates). The
// descriptor's range is the band (n BND_LD elements).
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 band_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return make_float2(f32_of(u[0]), f32_of(u[1]));
}
__device__ __forceinline__ void band_st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float2 v) {
  const u32x2 u = {u32_of(v.x), u32_of(v.y)};
  __builtin_amdgcn_raw_buffer_store_b64(u, r, voff, soff, 0);
}

// A full step (L = BW reflector rows, BW-1 bulge columns, BW rows below),
// with the addressing reduced to uniform bases plus ChaseLane offsets. BUF
// (FP32 global band): every band access is a buffer instruction on rb.
template <typename T, bool BUF, bool RE>
__device__ __forceinline__ T chase_step_full(typename C2T<T>::t* bnd, typename C2T<T>::t* v2row,
                                             typename C2T<T>::t* sv, int c1, int q1, int lane, const ChaseLane& o,
                                             typename C2T<T>::t (*stg)[BW + 1], bool from_stage, bool to_stage,
                                             __amdgpu_buffer_rsrc_t rb) {
  using C2 = typename C2T<T>::t;
  const int li = lane & 15, lg = lane >> 4;
  const C2 zero = mk<T>(0, 0);
  C2* bx = bnd + 31 * c1 + q1;            // column c1, rows q1 + li
  C2* bb = bnd + 31 * (c1 + 1) + q1;      // bulge block
  C2* bd = bnd + 32 * q1;                 // diagonal block
  C2* be = bnd + 32 * q1 + BW;            // block below
  // byte offsets of the same four bases (BUF)
  const unsigned sx = 8u * (31 * c1 + q1), sb = 8u * (31 * (c1 + 1) + q1), sd = 8u * 32 * q1, se = sd + 8u * BW;
  // from_stage: the previous step of this sweep left its updated block below
  // (this step's column c1 and bulge block) in the wave's stage, not in the band
  C2 x;
  if (from_stage)
    x = stg[li][0];
  else if constexpr (BUF)
    x = band_ld(rb, 8u * li, sx);
  else
    x = bx[li];
  C2 eb[4], ed[4], ee[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (from_stage) {
      eb[t] = li < BW - 1 ? stg[lg + 4 * t][1 + li] : zero;
    } else if constexpr (BUF) {
      // lane li = 15 reads column q1 (in the band) and drops it
      const C2 e = band_ld(rb, 8u * cl_ob(o, t), sb);
      eb[t] = li < BW - 1 ? e : zero;
    } else {
      eb[t] = load_or_zero(&at_off(bb, cl_ob(o, t)), li < BW - 1);
    }
    C2 d, e;
    if constexpr (BUF) {
      d = band_ld(rb, 8u * cl_od(o, t), sd);
      e = band_ld(rb, 8u * cl_oe(o, t), se);
    } else {
      d = at_off(bd, cl_od(o, t));
      e = at_off(be, cl_oe(o, t));
    }
    if (cl_up(o, t)) d.y = -d.y;
    ed[t] = d;
    ee[t] = e;
  }
  // (a) reflector
  const T xn2 = sum16(li >= 1 ? (RE ? x.x * x.x : x.x * x.x + x.y * x.y) : T(0));
  const C2 alpha = mk<T>(lane0(x.x), RE ? T(0) : lane0(x.y));
  C2 tau, scale;
  T beta;
  bool skip;
  if (xn2 == T(0) && alpha.y == T(0)) {
    skip = true;
    beta = alpha.x;
    tau = zero;
    scale = zero;
  } else if constexpr (RE) {
    skip = false;
    const T nrm = sqrt(alpha.x * alpha.x + xn2);
    beta = alpha.x >= T(0) ? -nrm : nrm;
    const T ib = rcp_t(beta);
    tau = mk<T>((beta - alpha.x) * ib, T(0));
    const T dx = alpha.x - beta, idd = rcp_t(dx * dx);
    scale = mk<T>(dx * idd, T(0));
  } else {
    skip = false;
    const T nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= T(0) ? -nrm : nrm;
    const T ib = rcp_t(beta);
    tau = mk<T>((beta - alpha.x) * ib, -alpha.y * ib);
    const T dx = alpha.x - beta, dy = alpha.y, idd = rcp_t(dx * dx + dy * dy);
    scale = mk<T>(dx * idd, -dy * idd);
  }
  const C2 v = li == 0 ? mk<T>(1, 0) : (skip ? zero : (RE ? mk<T>(x.x * scale.x, T(0)) : c_mul(x, scale)));
  if (lg == 0) {
    const C2 xn = li == 0 ? mk<T>(beta, 0) : zero;
    if constexpr (BUF)
      band_st(rb, 8u * li, sx, xn);
    else
      bx[li] = xn;
    if (v2row != nullptr) v2row[q1 + li] = li == 0 ? tau : v;
    sv[li] = v;
  }
  __builtin_amdgcn_wave_barrier();
  C2 vr[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) vr[t] = sv[lg + 4 * t];
  C2 db = zero, y = zero, z = zero;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    db = ccfma_r<RE>(vr[t], eb[t], db);
    y = cfma_r<RE>(ed[t], vr[t], y);
    z = cfma_r<RE>(ee[t], vr[t], z);
  }
  if constexpr (RE) {
    sum4g_pair(db.x, y.x);
    z.x = add_partner<32>(add_partner<16>(z.x));
  } else {
    db = sum4g_c(db);
    y = sum4g_c(y);
    z = sum4g_c(z);
  }
  const C2 fb = RE ? mk<T>(tau.x * db.x, T(0)) : c_mul(c_conj(tau), db);
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (li < BW - 1) {
      const C2 nb = cfms_r<RE>(vr[t], fb, eb[t]);
      if constexpr (BUF)
        band_st(rb, 8u * cl_ob(o, t), sb, nb);
      else
        at_off(bb, cl_ob(o, t)) = nb;
    }
  const T vhy = sum16(RE ? v.x * y.x : v.x * y.x + v.y * y.y);
  const T t2 = T(0.5) * (RE ? tau.x * tau.x : tau.x * tau.x + tau.y * tau.y) * vhy;
  const C2 ty = RE ? mk<T>(tau.x * y.x, T(0)) : c_mul(tau, y);
  const C2 w = RE ? mk<T>(ty.x - t2 * v.x, T(0)) : mk<T>(ty.x - t2 * v.x, ty.y - t2 * v.y);
  if (lg == 0) sv[BW + li] = w;
  __builtin_amdgcn_wave_barrier();
  const C2 tz = RE ? mk<T>(tau.x * z.x, T(0)) : c_mul(tau, z);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = lg + 4 * t;
    const C2 wr = sv[BW + r];
    if (!cl_up(o, t)) {
      C2 e = cfmsc_r<RE>(v, wr, cfmsc_r<RE>(w, vr[t], ed[t]));
      if (r == li) e.y = T(0);
      if constexpr (BUF)
        band_st(rb, 8u * cl_od(o, t), sd, e);
      else
        at_off(bd, cl_od(o, t)) = e;
    }
    // to_stage: the next step of this sweep (full, same wave) takes the
    // block below from the stage and writes it back to the band itself
    const C2 ev = cfmsc_r<RE>(tz, vr[t], ee[t]);
    if (to_stage) {
      stg[li][r] = ev;
    } else if constexpr (BUF) {
      band_st(rb, 8u * cl_oe(o, t), se, ev);
    } else {
      at_off(be, cl_oe(o, t)) = ev;
    }
  }
  __builtin_amdgcn_wave_barrier();
  return beta;
}

// Steps of sweep s: step 0, then one per BW rows while at least two rows remain.
__device__ __forceinline__ int chase_nsteps(int n, int s) {
  const int r = n - 3 - s;
  return 1 + (r > 0 ? r / BW : 0);
}

// Schedule (dataflow): wave w runs sweeps w, w + CH_WAVES, ... in order.
// Step c of sweep s may run once sweep s-1 has completed step c+2 (the
// regions of step c of sweep s and step c+3 of sweep s-1 are disjoint), so
// each wave publishes its progress (sweep << 12 | completed steps, 4095 when
// the sweep is done) in LDS with release semantics and waits on its
// predecessor's counter with acquire semantics. A wait is bounded: past
// c.spin_max polls (set by the host, slk_set_chase_spin_limit) the wave stops
// waiting, the matrix is flagged in c.info (2) and its tridiagonal is set to
// NaN; the kernel still runs to its end.
// LDSB: the compact band lives in LDS (dynamic shared memory); otherwise in
// global memory.
template <typename T, bool LDSB, bool RE = false>
__global__ __launch_bounds__(CH_THREADS) void band_chase_kernel(BandCtx<T> c, unsigned long long* clk) {
  using C2 = typename C2T<T>::t;
  sample_start(clk);
  extern __shared__ double ch_smem[];
  __shared__ int s_prog[CH_WAVES];
  __shared__ int s_fail;
  __shared__ C2 s_vw[CH_WAVES][2 * BW];   // per-wave v and w of the current step
  // global band: per-wave hand-over of a full step's block below to the next
  // full step of the same sweep (LDS band: not needed)
  __shared__ C2 s_stage[LDSB ? 1 : CH_WAVES][BW][BW + 1];
  const int b = blockIdx.x;
  const int n = c.n, ld = c.ld;
  C2* a = c.A + (size_t)b * ld * ld;
  C2* bnd = LDSB ? reinterpret_cast<C2*>(ch_smem) : c.BND + (size_t)b * ld * BND_LD;
  double* D = c.D + (size_t)b * ld;
  double* E = c.E + (size_t)b * ld;
  // the wave index is uniform: readfirstlane lets the compiler keep the
  // sweep and step indices (and every address built from them) in SGPRs
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (tid < CH_WAVES) s_prog[tid] = -1;
  if (tid == 0) s_fail = 0;
  for (int e = tid; e < n * BND_LD; e += CH_THREADS) {
    const int col = e / BND_LD, d = e % BND_LD, r = col + d;
    bnd[e] = a_ld<T, RE>(a, (size_t)r * ld + col, d <= BW && r < n);
  }
  __syncthreads();
  const int ns = n - 1;   // sweeps 0 .. n-2
  const int pw = (w + CH_WAVES - 1) % CH_WAVES;   // wave of the predecessor sweep
  const ChaseLane lo = chase_lane(lane);
  // FP32 global band: full steps address it through a buffer descriptor
  constexpr bool BUF = !LDSB && sizeof(T) == 4;
  const __amdgpu_buffer_rsrc_t rb =
      __builtin_amdgcn_make_buffer_rsrc(bnd, 0, BUF ? n * BND_LD * (int)sizeof(C2) : 0, 0x00020000);
  bool ok = true;
  for (int s = w; s < ns; s += CH_WAVES) {
    const int nst = chase_nsteps(n, s);
    int c1 = 0, q1 = 0, q2 = 0;
    bool staged = false;   // the previous step left the block below in s_stage
    for (int cs = 0; cs < nst; ++cs) {
      if (s > 0 && ok) {
        const int need = ((s - 1) << 12) + min(cs + 3, 4095);
        int spins = 0;
        if (c.spin_max < 0) ok = false;
        // sleep (8 x 64 cycles at most) until a wave of the workgroup
        // publishes a step and wakes the sleepers (s_wakeup below)
        while (ok && __hip_atomic_load(&s_prog[pw], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need) {
          __builtin_amdgcn_s_sleep(8);
          if (++spins > c.spin_max) {
            ok = false;
            break;
          }
        }
      }
      if (cs == 0) {
        c1 = s;
        q1 = s + 1;
        q2 = min(s + BW, n - 1);
      } else {
        c1 = q1;
        q1 = q2 + 1;
        q2 = min(q2 + BW, n - 1);
      }
      if (cs == 0 && lane == 0) D[s] = (double)bnd[(size_t)s * BND_LD].x;
      const bool full = cs > 0 && q2 - q1 + 1 == BW && q1 - c1 == BW && q2 + BW <= n - 1;
      // full steps: offsets precomputed per lane (32-bit, in LDS or global).
      // The next step is full too when q1 + 47 <= n - 1 (its q2 + BW): then a
      // global-band step hands its block below over through the wave's stage.
      T beta;
      if (full) {
        const bool to_stage = !LDSB && cs + 1 < nst && q1 + 3 * BW - 1 <= n - 1;
        beta = chase_step_full<T, BUF, RE>(bnd, a + (size_t)s * ld, s_vw[w], c1, q1, lane, lo, s_stage[LDSB ? 0 : w],
                                       staged, to_stage, rb);
        staged = to_stage;
      } else {
        beta = chase_step<T, false>(bnd, a + (size_t)s * ld, s_vw[w], n, c1, q1, q2, lane);
        staged = false;
      }
      if (cs == 0 && lane == 0) E[s] = (double)beta;
      __hip_atomic_store(&s_prog[w], (s << 12) + (cs + 1 == nst ? 4095 : cs + 1), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
      asm volatile("s_wakeup");
    }
  }
  if (!ok && lane == 0) s_fail = 1;
  __syncthreads();
  if (tid == 0) {
    D[n - 1] = (double)bnd[(size_t)(n - 1) * BND_LD].x;
    E[n - 1] = 0.0;
    if (s_fail && c.info) c.info[b] = 2;
  }
  if (s_fail)   // the tridiagonal is unordered garbage: NaN it so no caller takes it as a result
    for (int i = tid; i < n; i += CH_THREADS) D[i] = __builtin_nan("");
  sample_end(clk);
}

// Algorithmic flops of the chase of one n x n band (cached per n): per step
// with L reflector rows, nbc bulge columns and Le rows below, 2 complex
// multiply-adds per bulge element (inner product, update), L^2 for D v and
// L(L+1) for the rank-2 update of D's lower triangle, 2 per element of the
// block below, and ~10 L for the reflector; 8 real flops per multiply-add.
static double chase_flops(int n) {
  static std::mutex mu;
  static std::unordered_map<int, double> cache;
  std::lock_guard<std::mutex> lock(mu);
  const auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  double f = 0.0;
  for (int s = 0; s < n - 1; ++s) {
    const int r = n - 3 - s, nst = 1 + (r > 0 ? r / BW : 0);
    int c1 = s, q1 = s + 1, q2 = min(s + BW, n - 1);
    for (int cs = 0; cs < nst; ++cs) {
      if (cs > 0) {
        c1 = q1;
        q1 = q2 + 1;
        q2 = min(q2 + BW, n - 1);
      }
      const double L = q2 - q1 + 1, nbc = q1 - c1 - 1, Le = min(BW, n - 1 - q2);
      f += 16.0 * nbc * L + 8.0 * L * L + 8.0 * L * (L + 1) + 16.0 * Le * L + 10.0 * L;
    }
  }
  cache[n] = f;
  return f;
}

// partials of A22 V: one slot of ld x BW per 64-row chunk of K
size_t band_partials_elems(int ld) { return (size_t)((ld + 63) / 64) * ld * BW; }
// compact band, Zp, T
int band_panels(int ld) { return (ld + BW - 1) / BW; }
// compact band, Zp, the panels' T factors, the second [V|W], [W|V] pair
size_t band_scratch_elems(int ld) {
  return (size_t)ld * BND_LD + (size_t)((ld + 63) / 64) * BW * BW + (size_t)band_panels(ld) * BW * BW +
         (size_t)2 * (2 * BW) * ld + (size_t)BW * BW;
}
size_t band_tt_offset(int ld, int batch) {
  return (size_t)batch * ld * BND_LD + (size_t)batch * ((ld + 63) / 64) * BW * BW;
}

// ---------------------------------------------------------------------------
// Look-ahead form of stage 1. Panel p's update of A22(p) is split: the next
// panel's 16 columns first (in band_w_kernel, on the rows of each of its
// blocks), so that its QR can run, then the rest, A22(p+1), in one pass that
// also forms the next panel's product Y = A22(p+1) V_{p+1}
// (fused_update_av_kernel). A22(p+1) is read and written once per panel
// instead of read twice and written once.
// ---------------------------------------------------------------------------

// One 64 x 64 tile (I, J), J <= I, of A22(p+1) (origin r0n, order mn):
//   1. A -= [V | W]_p [W | V]_p^H on its 32 x 32 blocks touching the lower
//      triangle, each wave loading its block straight into the MFMA
//      accumulators and the operands straight into registers in MFMA operand
//      layout; the result is stored back and kept in LDS;
//   2. a diagonal tile gets its upper half from the conjugate of the lower
//      one (elements above the diagonal are not kept current in A);
//   3. the partials of Y = A22(p+1) V_{p+1}: rows of block I from A_IJ V_J
//      into chunk J, and (I != J) rows of block J from A_IJ^H V_I into chunk
//      I. Every (row block, chunk) slot is written by exactly one tile.
// Vp, Wp: panel p's [V | W], [W | V] at row r0n; Vn: V_{p+1} at row r0n.
// RE: real data (zero imaginary parts): one real MFMA per block product,
// no imaginary LDS planes (half the LDS, twice the workgroups per CU), and
// zero imaginary parts written.
template <typename T, bool RE = false>
__global__ __launch_bounds__(256) void fused_update_av_kernel(const typename C2T<T>::t* __restrict__ Vp,
                                                              const typename C2T<T>::t* __restrict__ Wp,
                                                              const typename C2T<T>::t* __restrict__ Vn,
                                                              typename C2T<T>::t* __restrict__ A,
                                                              typename C2T<T>::t* __restrict__ Yp, int n, int ld,
                                                              int r0n, int nch, long long sP, long long sM,
                                                              int ntri, int batch, unsigned long long* clk) {
  using C2 = typename C2T<T>::t;
  using V4 = typename MF<T>::v4;
  constexpr int KK = 2 * BW / 4;
  sample_start(clk);
  // One workgroup per lower tile (I, J), J <= I, of every matrix. Workgroups
  // are dealt round-robin to the 8 XCDs, so the launch order is remapped to
  // give each XCD a contiguous run of tiles (row-major in the triangle): tiles
  // of one row block share their V / W rows in that XCD's L2.
  const int G = ntri * batch;
  const int per = (G + 7) / 8;
  const int lin = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (lin >= G) return;
  const int b = lin / ntri, q = lin % ntri;
  int I = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= q) ++I;
  while (I * (I + 1) / 2 > q) --I;
  const int J = q - I * (I + 1) / 2;
  const int mn = n - r0n;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const C2* vp = Vp + (size_t)b * sP;
  const C2* wp = Wp + (size_t)b * sP;
  const C2* vn = Vn + (size_t)b * sP;
  C2* a = A + (size_t)b * sM;
  const size_t o = (size_t)r0n * ld + r0n;   // origin of A22(p+1) in the matrix
  __shared__ T Cr[64][65], Ci[RE ? 1 : 64][65];
  __shared__ T VJr[64][BW + 1], VJi[RE ? 1 : 64][BW + 1], VIr[64][BW + 1], VIi[RE ? 1 : 64][BW + 1];
  // V_{p+1} rows of blocks J and I (zero beyond the matrix). Real data: stored
  // to LDS after the rank-2k, so that these loads are in flight together with
  // the rank-2k's instead of one round trip before them; complex data (twice
  // the registers) is stored at once
  constexpr int NV = 64 * BW / 256;
  C2 xj[NV], xi[NV];
#pragma unroll
  for (int it = 0; it < NV; ++it) {
    const int e = tid + 256 * it, col = e / 64, k = e % 64;
    const int rj = 64 * J + k, ri = 64 * I + k;
    xj[it] = load_or_zero(vn + (size_t)col * ld + rj, rj < mn);
    xi[it] = load_or_zero(vn + (size_t)col * ld + ri, ri < mn);
  }
  auto store_v = [&]() {
#pragma unroll
    for (int it = 0; it < NV; ++it) {
      const int e = tid + 256 * it, col = e / 64, k = e % 64;
      VJr[k][col] = xj[it].x;
      VIr[k][col] = xi[it].x;
      if constexpr (!RE) {
        VJi[k][col] = xj[it].y;
        VIi[k][col] = xi[it].y;
      }
    }
  };
  if constexpr (!RE) store_v();
  // 1. rank-2k on this wave's 32 x 32 block
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  const int m0 = 64 * I + wr, n0 = 64 * J + wc;
  if (n0 <= m0 + 31) {
    C2 av[2][KK], bw[2][KK];
#pragma unroll
    for (int q = 0; q < KK; ++q) {
      const size_t krow = (size_t)(4 * q + lk) * ld;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = m0 + 16 * i + li;
        av[i][q] = load_or_zero(vp + krow + r, r < mn);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = n0 + 16 * j + li;
        bw[j][q] = load_or_zero(wp + krow + col, col < mn);
      }
    }
    V4 accR[2][2], accI[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = m0 + 16 * i + MF<T>::row(lk, v), col = n0 + 16 * j + li;
          const C2 x = a_ld<T, RE>(a, o + (size_t)r * ld + col, r < mn && col < mn);
          accR[i][j][v] = x.x;
          accI[i][j][v] = x.y;
        }
#pragma unroll
    for (int q = 0; q < KK; ++q)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = MF<T>::mfma(-av[i][q].x, bw[j][q].x, accR[i][j]);
          if constexpr (!RE) {
            accR[i][j] = MF<T>::mfma(-av[i][q].y, bw[j][q].y, accR[i][j]);
            accI[i][j] = MF<T>::mfma(-av[i][q].y, bw[j][q].x, accI[i][j]);
            accI[i][j] = MF<T>::mfma(av[i][q].x, bw[j][q].y, accI[i][j]);
          }
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rl = wr + 16 * i + MF<T>::row(lk, v), cl = wc + 16 * j + li;
          const int r = 64 * I + rl, col = 64 * J + cl;
          if (r < mn && col < mn) a_st<T, RE>(a, o + (size_t)r * ld + col, mk<T>(accR[i][j][v], RE ? T(0) : accI[i][j][v]));
          Cr[rl][cl] = accR[i][j][v];
          if constexpr (!RE) Ci[rl][cl] = accI[i][j][v];
        }
  }
  if constexpr (RE) store_v();
  __syncthreads();
  // 2. Hermitian completion of a diagonal tile
  if (I == J) {
    for (int e = tid; e < 64 * 64; e += 256) {
      const int r = e / 64, col = e % 64;
      if (col > r) {
        Cr[r][col] = Cr[col][r];
        if constexpr (!RE) Ci[r][col] = -Ci[col][r];
      }
    }
    __syncthreads();
  }
  // 3. partials of Y: this wave's 16 rows of each product
  const int rw = 16 * wv;
  {
    V4 yR = V4{0, 0, 0, 0}, yI = V4{0, 0, 0, 0};
#pragma unroll
    for (int k4 = 0; k4 < 64; k4 += 4) {
      const T ar = Cr[rw + li][k4 + lk], br = VJr[k4 + lk][li];
      yR = MF<T>::mfma(ar, br, yR);
      if constexpr (!RE) {
        const T ai = Ci[rw + li][k4 + lk], bi = VJi[k4 + lk][li];
        yR = MF<T>::mfma(-ai, bi, yR);
        yI = MF<T>::mfma(ar, bi, yI);
        yI = MF<T>::mfma(ai, br, yI);
      }
    }
    C2* yp = Yp + ((size_t)b * nch + J) * ld * BW;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 64 * I + rw + MF<T>::row(lk, v);
      if (r < mn) a_st<T, RE>(yp, (size_t)(r0n + r) * BW + li, mk<T>(yR[v], RE ? T(0) : yI[v]));
    }
  }
  if (I != J) {
    V4 yR = V4{0, 0, 0, 0}, yI = V4{0, 0, 0, 0};
#pragma unroll
    for (int k4 = 0; k4 < 64; k4 += 4) {
      // (A^H)[i][k] = conj(A[k][i])
      const T ar = Cr[k4 + lk][rw + li], br = VIr[k4 + lk][li];
      yR = MF<T>::mfma(ar, br, yR);
      if constexpr (!RE) {
        const T ai = -Ci[k4 + lk][rw + li], bi = VIi[k4 + lk][li];
        yR = MF<T>::mfma(-ai, bi, yR);
        yI = MF<T>::mfma(ar, bi, yI);
        yI = MF<T>::mfma(ai, br, yI);
      }
    }
    C2* yp = Yp + ((size_t)b * nch + I) * ld * BW;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 64 * J + rw + MF<T>::row(lk, v);
      if (r < mn) a_st<T, RE>(yp, (size_t)(r0n + r) * BW + li, mk<T>(yR[v], RE ? T(0) : yI[v]));
    }
  }
  if (clk != nullptr) {
    __syncthreads();
    sample_end(clk);
  }
}

#define CH_LDS_MAX (150 * 1024)
// Dynamic LDS beyond the default limit has to be allowed per kernel once.
template <typename T, bool RE>
static bool chase_lds_ok() {
  static const bool ok = hipFuncSetAttribute(reinterpret_cast<const void*>(&band_chase_kernel<T, true, RE>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, CH_LDS_MAX) == hipSuccess;
  return ok;
}

template <typename T, bool RE>
static void launch_panel_qr(hipStream_t st, const BandCtx<T>& c, int batch, int j0, int m) {
  const dim3 g(batch), blk(1024);
  // rows per lane in registers: a complex value is 2 (FP32) or 4 (FP64) VGPRs
  const int nq = (m + 63) / 64;
  constexpr bool F = sizeof(T) == 4;
  if (nq <= 2) hipLaunchKernelGGL((panel_qr_kernel<T, 2, 0, RE>), g, blk, 0, st, c, j0);
  else if (nq <= 4) hipLaunchKernelGGL((panel_qr_kernel<T, 4, 0, RE>), g, blk, 0, st, c, j0);
  else if (nq <= 8) hipLaunchKernelGGL((panel_qr_kernel<T, 8, 0, RE>), g, blk, 0, st, c, j0);
  else if (nq <= 16) hipLaunchKernelGGL((panel_qr_kernel<T, 16, 0, RE>), g, blk, 0, st, c, j0);
  else if (nq <= 24 && F) hipLaunchKernelGGL((panel_qr_kernel<T, F ? 24 : 16, 0, RE>), g, blk, 0, st, c, j0);
  else if (nq <= 32 && F) hipLaunchKernelGGL((panel_qr_kernel<T, F ? 32 : 16, 0, RE>), g, blk, 0, st, c, j0);
  // beyond the register tiers: 4 more rows per lane in LDS (FP32: m <= 2304,
  // every serving matrix; FP64: m <= 1280)
  else if (nq <= 36 && F) hipLaunchKernelGGL((panel_qr_kernel<T, F ? 32 : 16, 4, RE>), g, blk, 0, st, c, j0);
  else if (nq <= 20 && !F) hipLaunchKernelGGL((panel_qr_kernel<T, 16, 4, RE>), g, blk, 0, st, c, j0);
  else hipLaunchKernelGGL((panel_qr_kernel<T, 0, 0, RE>), g, blk, 0, st, c, j0);
}

// Polls per chase wait before the wave gives up (process-wide).
static std::atomic<int> g_chase_spin_max{1 << 22};

template <typename T, bool RE>
static void band_impl(hipStream_t st, typename C2T<T>::t* A, int n, int ld, int batch, double* D, double* E,
                      typename C2T<T>::t* TAU, typename C2T<T>::t* Vc, typename C2T<T>::t* Wc,
                      typename C2T<T>::t* Yp, typename C2T<T>::t* scratch, typename C2T<T>::t* VT, int* info) {
  using C2 = typename C2T<T>::t;
  BandCtx<T> c;
  c.info = info;
  c.spin_max = g_chase_spin_max.load();
  c.A = A;
  c.n = n;
  c.ld = ld;
  c.Vc = Vc;
  c.Wc = Wc;
  c.Yp = Yp;
  c.TAU = TAU;
  c.VT = VT;
  c.D = D;
  c.E = E;
  c.nch = (ld + 63) / 64;
  c.nrb = (ld + 63) / 64;
  c.npan = band_panels(ld);
  // scratch: band, then Zp, then the panels' T, then the second [V|W], [W|V]
  // pair, then X16 (band_scratch_elems per matrix)
  c.BND = scratch;
  c.Zp = scratch + (size_t)batch * ld * BND_LD;
  c.TT = c.Zp + (size_t)batch * c.nrb * BW * BW;
  C2* Vc2 = c.TT + (size_t)batch * c.npan * BW * BW;
  C2* Wc2 = Vc2 + (size_t)batch * 2 * BW * ld;
  c.X16 = Wc2 + (size_t)batch * 2 * BW * ld;
  const long long sP = (long long)2 * BW * ld, sM = (long long)ld * ld;
  // panel p works in the buffer pair of its parity: the next panel's QR
  // writes its V while panel p's [V | W] is still needed
  BandCtx<T> cx[2] = {c, c};
  cx[1].Vc = Vc2;
  cx[1].Wc = Wc2;
  // m = 1 gives a one-element (phase) reflector: it leaves the band as it is
  // and defines tau for every column whose reflector starts inside the matrix
  if (n - BW >= 1) {
    const int m = n - BW;
    launch_panel_qr<T, RE>(st, cx[0], batch, 0, m);
    const int nrbk = (m + 63) / 64, nchk = (m + AV_KCH - 1) / AV_KCH;
    hipLaunchKernelGGL((band_av_kernel<T, RE>), dim3(nrbk, nchk, batch), dim3(256), 0, st, cx[0], BW);
    hipLaunchKernelGGL((band_x_kernel<T, RE>), dim3(nrbk, batch), dim3(256), 0, st, cx[0], BW, nchk);
    hipLaunchKernelGGL((band_w_kernel<T, RE>), dim3(nrbk, batch), dim3(256), 0, st, cx[0], BW, nrbk);
    for (int p = 0;; ++p) {
      // the strip of A22(p) (origin BW (p + 1)) is current: the band_w
      // launch that formed panel p's W applied it
      const BandCtx<T>& cp = cx[p & 1];
      const BandCtx<T>& cn = cx[(p + 1) & 1];
      const int r0 = BW * (p + 1);                    // A22(p)
      const int r0n = r0 + BW, mn = n - r0n;          // A22(p + 1)
      if (mn < 1) break;
      launch_panel_qr<T, RE>(st, cn, batch, r0, mn);
      const int t = (mn + 63) / 64;
      // FP32 launches are sampled: 8 (m^2/2) 32 flops for the rank-2k on the
      // lower triangle and 8 m^2 16 for A22 V, per complex matrix; a real
      // matrix takes 2 flops per multiply-add instead of 8
      SampleSlot fs;
      const bool fsam = sizeof(T) == 4 && sampler_enabled() &&
                        sampler_take(RE ? SLK_SAMPLE_FUSED_AV_REAL : SLK_SAMPLE_FUSED_AV,
                                     (double)batch * (RE ? 64.0 : 256.0) * mn * mn, &fs);
      const int ntri = t * (t + 1) / 2;
      const int nwg = (ntri * batch + 7) / 8 * 8;
      if (fsam)
        hipExtLaunchKernelGGL((fused_update_av_kernel<T, RE>), dim3(nwg), dim3(256), 0, st, fs.e0, fs.e1, 0,
                              cp.Vc + r0n, cp.Wc + r0n, cn.Vc + r0n, A, c.Yp, n, ld, r0n, c.nch, sP, sM, ntri,
                              batch, fs.clk);
      else
        hipLaunchKernelGGL((fused_update_av_kernel<T, RE>), dim3(nwg), dim3(256), 0, st, cp.Vc + r0n,
                           cp.Wc + r0n, cn.Vc + r0n, A, c.Yp, n, ld, r0n, c.nch, sP, sM, ntri, batch,
                           (unsigned long long*)nullptr);
      hipLaunchKernelGGL((band_x_kernel<T, RE>), dim3(t, batch), dim3(256), 0, st, cn, r0n, t);
      hipLaunchKernelGGL((band_w_kernel<T, RE>), dim3(t, batch), dim3(256), 0, st, cn, r0n, t);
    }
  }
  // the band (and the schedule) in LDS when it fits, else in global memory
  const size_t lds = (size_t)n * BND_LD * sizeof(C2);
  SampleSlot sl;
  // algorithmic amount per workgroup (one matrix): each workgroup runs on its
  // own compute unit, so flops / window is the per-CU rate
  const bool sampled = sampler_enabled() && sampler_take(RE ? SLK_SAMPLE_CHASE_REAL : SLK_SAMPLE_CHASE,
                                                          RE ? 0.25 * chase_flops(n) : chase_flops(n), &sl);
  unsigned long long* clk = sampled ? sl.clk : nullptr;
  if (lds <= CH_LDS_MAX && chase_lds_ok<T, RE>()) {
    if (sampled)
      hipExtLaunchKernelGGL((band_chase_kernel<T, true, RE>), dim3(batch), dim3(CH_THREADS), lds, st, sl.e0, sl.e1, 0, c,
                            clk);
    else
      hipLaunchKernelGGL((band_chase_kernel<T, true, RE>), dim3(batch), dim3(CH_THREADS), lds, st, c, clk);
  } else {
    if (sampled)
      hipExtLaunchKernelGGL((band_chase_kernel<T, false, RE>), dim3(batch), dim3(CH_THREADS), 0, st, sl.e0, sl.e1, 0, c,
                            clk);
    else
      hipLaunchKernelGGL((band_chase_kernel<T, false, RE>), dim3(batch), dim3(CH_THREADS), 0, st, c, clk);
  }
  (void)sizeof(C2);
}

void launch_band_tridiag(hipStream_t st, double2* C, int n, int ld, int batch, double* D, double* E, double2* TAU,
                         double2* Vc, double2* Wc, double2* Yp, double2* scratch, double2* VT, int* info, bool re) {
  if (re)
    band_impl<double, true>(st, C, n, ld, batch, D, E, TAU, Vc, Wc, Yp, scratch, VT, info);
  else
    band_impl<double, false>(st, C, n, ld, batch, D, E, TAU, Vc, Wc, Yp, scratch, VT, info);
}

void launch_band_tridiag_f32(hipStream_t st, float2* C, int n, int ld, int batch, double* D, double* E, float2* TAU,
                             float2* Vc, float2* Wc, float2* Yp, float2* scratch, float2* VT, int* info, bool re) {
  if (re)
    band_impl<float, true>(st, C, n, ld, batch, D, E, TAU, Vc, Wc, Yp, scratch, VT, info);
  else
    band_impl<float, false>(st, C, n, ld, batch, D, E, TAU, Vc, Wc, Yp, scratch, VT, info);
}

extern "C" int slk_set_chase_spin_limit(int polls) {
  return g_chase_spin_max.exchange(polls);
}
