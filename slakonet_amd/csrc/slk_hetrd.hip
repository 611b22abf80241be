// Blocked Householder tridiagonalisation C = Q T Q^H of a batch of Hermitian
// matrices (full storage, row-major, lower reflectors), LAPACK zhetrd/zlatrd
// structure with 32-column panels, in FP64 or FP32 (the tridiagonal T is
// always returned in FP64):
//
//   per panel:  latrd_start (first column reflector)
//               per column j: hetrd_step = trailing Hermitian mat-vec A22 v_j
//                 over 128x128 tiles of the lower triangle, each read once
//                 and used for both A v and A^H v, writing per-tile partial
//                 products (diagonal tiles also partial W^H v_j, V^H v_j);
//                 hetrd_sum (many workgroups) forms
//                 y = A22 v_j - V (W^H v_j) - W (V^H v_j) and column j+1 with
//                 the panel's older pairs applied; hetrd_tail (one workgroup
//                 per matrix, O(n) work) forms w_j, applies the newest pair
//                 to column j+1 and generates its reflector;
//               rank-2k update A22 -= V W^H + W V^H on the MFMA GEMM, lower
//               triangle tiles only (the upper triangle is never read).
// A kernel boundary orders the tail after the tiles; a device-scope fence
// per tile flushes L2 on this multi-XCD part and costs far more.
// Panel buffers are column-major: Vc[k*ld + r], Wc[k*ld + r].
#include "slk_common.h"
#include "slk_dense.h"
#include <hip/hip_ext.h>

#include "slk_sampler.h"

#define NB SLK_NB
#define HT_THREADS 512
#define TAIL_THREADS 1024
#define HT_TS 128      // square tile of the trailing mat-vec: 2 columns per lane
#define HT_RPW (HT_TS / (HT_THREADS / 64))   // tile rows per wave

template <typename T>
struct HetrdCtx {
  using C2 = typename C2T<T>::t;
  C2* A;
  int n, ld;
  C2 *Vc, *Wc, *Y, *Yp, *C12, *TAU;
  C2* VT;      // optional: reflector j also stored as row j (VT[j*ld + r], r > j)
  C2* COL;     // next column with the older panel pairs applied ([batch][ld])
  C2* C12p;    // per row block partials of W^H v, V^H v ([batch][ncmax][2 NB])
  double *D, *E;
  int ncmax;   // partial-sum chunks per matrix (>= ceil(ld / HT_COLS))
};

// w_{jj-1} = tau y - |tau|^2/2 (v^H y) v over rows >= j (support of v_{j-1}).
template <typename T>
__device__ void latrd_finalize(const HetrdCtx<T>& c, int b, int j, int jj, T* red) {
  using C2 = typename C2T<T>::t;
  const int n = c.n, ld = c.ld;
  const C2* vc = c.Vc + (size_t)b * NB * ld;
  C2* wc = c.Wc + (size_t)b * NB * ld;
  const C2* y = c.Y + (size_t)b * ld;
  const int kp = jj - 1;
  const C2 tp = c.TAU[(size_t)b * ld + j - 1];
  T part = 0;
  for (int r = j + threadIdx.x; r < n; r += blockDim.x) {
    const C2 v = vc[(size_t)kp * ld + r], yy = y[r];
    part += v.x * yy.x + v.y * yy.y;
  }
  const T vhy = block_sum(part, red);
  const T t2 = T(0.5) * (tp.x * tp.x + tp.y * tp.y) * vhy;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    C2 o = mk<T>(0, 0);
    if (r >= j) {
      const C2 p = c_mul(tp, y[r]);
      const C2 v = vc[(size_t)kp * ld + r];
      o = mk<T>(p.x - t2 * v.x, p.y - t2 * v.y);
    }
    wc[(size_t)kp * ld + r] = o;
  }
  __syncthreads();
}

// Column j (panel position jj): apply the panel's pending rank-2 updates,
// generate the reflector, store v (panel buffer and strict lower column j),
// d[j], e[j], tau[j], and the correction coefficients W^H v, V^H v.
template <typename T>
__device__ void latrd_column(const HetrdCtx<T>& c, int b, int j, int jj, T* red) {
  using C2 = typename C2T<T>::t;
  const int n = c.n, ld = c.ld;
  C2* a = c.A + (size_t)b * ld * ld;
  C2* vc = c.Vc + (size_t)b * NB * ld;
  const C2* wc = c.Wc + (size_t)b * NB * ld;
  C2* c12 = c.C12 + (size_t)b * 2 * NB;
  __shared__ C2 s_vj[NB], s_wj[NB];
  __shared__ C2 s_tau, s_scale;
  __shared__ T s_beta;
  __shared__ int s_skip;
  const int tid = threadIdx.x;
  if (tid < jj) {
    s_vj[tid] = vc[(size_t)tid * ld + j];
    s_wj[tid] = wc[(size_t)tid * ld + j];
  }
  __syncthreads();
  T part = 0;
  for (int r = j + tid; r < n; r += blockDim.x) {
    C2 acc = a[(size_t)r * ld + j];
    for (int k = 0; k < jj; ++k) {
      const C2 t1 = c_mulc(vc[(size_t)k * ld + r], s_wj[k]);
      const C2 t3 = c_mulc(wc[(size_t)k * ld + r], s_vj[k]);
      acc.x -= t1.x + t3.x;
      acc.y -= t1.y + t3.y;
    }
    a[(size_t)r * ld + j] = acc;
    if (r >= j + 2) part += acc.x * acc.x + acc.y * acc.y;
  }
  const T xn2 = block_sum(part, red);
  if (tid == 0) {
    const C2 alpha = a[(size_t)(j + 1) * ld + j];
    if (xn2 == T(0) && alpha.y == T(0)) {
      s_skip = 1;
      s_tau = mk<T>(0, 0);
      s_beta = alpha.x;
      s_scale = mk<T>(0, 0);
    } else {
      s_skip = 0;
      const T nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      const T beta = alpha.x >= T(0) ? -nrm : nrm;
      s_beta = beta;
      s_tau = mk<T>((beta - alpha.x) / beta, -alpha.y / beta);
      const C2 den = mk<T>(alpha.x - beta, alpha.y);
      const T dd = den.x * den.x + den.y * den.y;
      s_scale = mk<T>(den.x / dd, -den.y / dd);
    }
    c.D[(size_t)b * ld + j] = (double)a[(size_t)j * ld + j].x;
    c.E[(size_t)b * ld + j] = (double)s_beta;
    c.TAU[(size_t)b * ld + j] = s_tau;
  }
  __syncthreads();
  const C2 scale = s_scale;
  for (int r = tid; r < n; r += blockDim.x) {
    C2 v = mk<T>(0, 0);
    if (r == j + 1) v = mk<T>(1, 0);
    else if (r > j + 1) v = s_skip ? mk<T>(0, 0) : c_mul(a[(size_t)r * ld + j], scale);
    vc[(size_t)jj * ld + r] = v;
    if (r > j) a[(size_t)r * ld + j] = v;
    if (c.VT != nullptr && r > j) c.VT[(size_t)b * ld * ld + (size_t)j * ld + r] = v;
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int k = wv; k < jj; k += nw) {
    C2 s1 = mk<T>(0, 0), s2 = mk<T>(0, 0);
    for (int r = j + 1 + lane; r < n; r += 64) {
      const C2 v = vc[(size_t)jj * ld + r];
      const C2 x1 = c_cmul(wc[(size_t)k * ld + r], v);
      const C2 x2 = c_cmul(vc[(size_t)k * ld + r], v);
      s1.x += x1.x; s1.y += x1.y;
      s2.x += x2.x; s2.y += x2.y;
    }
    s1.x = wave_sum(s1.x); s1.y = wave_sum(s1.y);
    s2.x = wave_sum(s2.x); s2.y = wave_sum(s2.y);
    if (lane == 0) { c12[k] = s1; c12[NB + k] = s2; }
  }
  __syncthreads();
}

// Column j (panel position jj >= 1) from COL (older pairs already applied):
// subtract the newest pair (v_{jj-1}, w_{jj-1}), generate the reflector and
// store it (panel buffer, strict lower column j, VT row). W^H v and V^H v of
// the new reflector are formed by the next mat-vec's diagonal tiles.
template <typename T>
__device__ void latrd_next(const HetrdCtx<T>& c, int b, int j, int jj, T* red) {
  using C2 = typename C2T<T>::t;
  const int n = c.n, ld = c.ld;
  C2* a = c.A + (size_t)b * ld * ld;
  C2* vc = c.Vc + (size_t)b * NB * ld;
  const C2* wc = c.Wc + (size_t)b * NB * ld;
  const C2* col = c.COL + (size_t)b * ld;
  __shared__ C2 s_tau, s_scale;
  __shared__ T s_beta;
  __shared__ int s_skip;
  const int tid = threadIdx.x;
  const int kp = jj - 1;
  const C2 vj = vc[(size_t)kp * ld + j], wj = wc[(size_t)kp * ld + j];
  T part = 0;
  for (int r = j + tid; r < n; r += blockDim.x) {
    C2 acc = col[r];
    const C2 t1 = c_mulc(vc[(size_t)kp * ld + r], wj);
    const C2 t3 = c_mulc(wc[(size_t)kp * ld + r], vj);
    acc.x -= t1.x + t3.x;
    acc.y -= t1.y + t3.y;
    a[(size_t)r * ld + j] = acc;
    if (r >= j + 2) part += acc.x * acc.x + acc.y * acc.y;
  }
  const T xn2 = block_sum(part, red);
  if (tid == 0) {
    const C2 alpha = a[(size_t)(j + 1) * ld + j];
    if (xn2 == T(0) && alpha.y == T(0)) {
      s_skip = 1;
      s_tau = mk<T>(0, 0);
      s_beta = alpha.x;
      s_scale = mk<T>(0, 0);
    } else {
      s_skip = 0;
      const T nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      const T beta = alpha.x >= T(0) ? -nrm : nrm;
      s_beta = beta;
      s_tau = mk<T>((beta - alpha.x) / beta, -alpha.y / beta);
      const C2 den = mk<T>(alpha.x - beta, alpha.y);
      const T dd = den.x * den.x + den.y * den.y;
      s_scale = mk<T>(den.x / dd, -den.y / dd);
    }
    c.D[(size_t)b * ld + j] = (double)a[(size_t)j * ld + j].x;
    c.E[(size_t)b * ld + j] = (double)s_beta;
    c.TAU[(size_t)b * ld + j] = s_tau;
  }
  __syncthreads();
  const C2 scale = s_scale;
  for (int r = tid; r < n; r += blockDim.x) {
    C2 v = mk<T>(0, 0);
    if (r == j + 1) v = mk<T>(1, 0);
    else if (r > j + 1) v = s_skip ? mk<T>(0, 0) : c_mul(a[(size_t)r * ld + j], scale);
    vc[(size_t)jj * ld + r] = v;
    if (r > j) a[(size_t)r * ld + j] = v;
    if (c.VT != nullptr && r > j) c.VT[(size_t)b * ld * ld + (size_t)j * ld + r] = v;
  }
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(TAIL_THREADS) void latrd_start_kernel(HetrdCtx<T> c, int j) {
  __shared__ T red[TAIL_THREADS / 64];
  latrd_column(c, blockIdx.x, j, 0, red);
}

// Trailing Hermitian mat-vec for column j from the lower triangle only.
// The trailing block (rows/columns >= j+1) is cut into HT_TS x HT_TS tiles;
// workgroup (t, b) takes lower tile t = (bi, bj), bi >= bj, of matrix b and
// reads it once: its rows get A v (row partial), its columns get A^H v from
// the same elements (column partial, strictly lower elements only). Partial
// (ch, r) of row r in block bi comes from exactly one tile: ch = bj of tile
// (bi, bj) for bj <= bi, or ch = bi' of tile (bi', bi) for bi' > bi. The
// diagonal tile of a row block also subtracts the panel correction
// V (W^H v) + W (V^H v) for its rows.
template <typename T>
__global__ __launch_bounds__(HT_THREADS) void hetrd_step_kernel(HetrdCtx<T> c, int j, int jj,
                                                                unsigned long long* clk) {
  using C2 = typename C2T<T>::t;
  sample_start(clk);
  __shared__ C2 s_row[HT_TS];
  __shared__ C2 s_col[HT_THREADS / 64][HT_TS];
  const int b = blockIdx.y;
  const int n = c.n, ld = c.ld;
  const C2* a = c.A + (size_t)b * ld * ld;
  const C2* vc = c.Vc + (size_t)b * NB * ld;
  const C2* wc = c.Wc + (size_t)b * NB * ld;
  const C2* v = vc + (size_t)jj * ld;
  C2* yp = c.Yp + (size_t)b * c.ncmax * ld;
  int bi = 0;
  const int t = blockIdx.x;
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  const int bj = t - bi * (bi + 1) / 2;
  const bool diag = bi == bj;
  const int base = j + 1;
  const int row0 = base + bi * HT_TS, col0 = base + bj * HT_TS;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int cA = col0 + 2 * lane, cB = cA + 1;
  const C2 zero = mk<T>(0, 0);
  const C2 vA = cA < n ? v[cA] : zero, vB = cB < n ? v[cB] : zero;
  C2 colA = zero, colB = zero;
#pragma unroll 2
  for (int rr = 0; rr < HT_RPW; ++rr) {
    const int rl = wv * HT_RPW + rr;
    const int r = row0 + rl;
    C2 xA = zero, xB = zero, vr = zero;
    if (r < n) {
      vr = v[r];
      if (cA < n && (!diag || cA <= r)) xA = a[(size_t)r * ld + cA];
      if (cB < n && (!diag || cB <= r)) xB = a[(size_t)r * ld + cB];
    }
    // row partial: A[r][c] v[c]
    T px = xA.x * vA.x - xA.y * vA.y + xB.x * vB.x - xB.y * vB.y;
    T py = xA.x * vA.y + xA.y * vA.x + xB.x * vB.y + xB.y * vB.x;
    px = wave_sum(px);
    py = wave_sum(py);
    if (lane == 0) s_row[rl] = mk<T>(px, py);
    // column partial: conj(A[r][c]) v[r], strictly lower elements
    if (!diag || cA < r) {
      colA.x += xA.x * vr.x + xA.y * vr.y;
      colA.y += xA.x * vr.y - xA.y * vr.x;
    }
    if (!diag || cB < r) {
      colB.x += xB.x * vr.x + xB.y * vr.y;
      colB.y += xB.x * vr.y - xB.y * vr.x;
    }
  }
  s_col[wv][2 * lane] = colA;
  s_col[wv][2 * lane + 1] = colB;
  __syncthreads();
  const int tid = threadIdx.x;
  if (tid < HT_TS) {
    C2 cs = zero;
#pragma unroll
    for (int w = 0; w < HT_THREADS / 64; ++w) cs = c_add(cs, s_col[w][tid]);
    if (diag) {
      const int r = row0 + tid;
      if (r < n) yp[(size_t)bi * ld + r] = c_add(s_row[tid], cs);
    } else {
      const int r = row0 + tid;
      if (r < n) yp[(size_t)bj * ld + r] = s_row[tid];
      const int col = col0 + tid;
      if (col < n) yp[(size_t)bi * ld + col] = cs;
    }
  }
  if (diag && jj > 0) {
    // partial W^H v and V^H v over this row block (reduced by hetrd_sum)
    C2* cp = c.C12p + ((size_t)b * c.ncmax + bi) * 2 * NB;
    for (int k = wv; k < jj; k += HT_THREADS / 64) {
      C2 s1 = zero, s2 = zero;
      for (int rl = lane; rl < HT_TS; rl += 64) {
        const int r = row0 + rl;
        if (r < n) {
          const C2 vv = v[r];
          const C2 x1 = c_cmul(wc[(size_t)k * ld + r], vv);
          const C2 x2 = c_cmul(vc[(size_t)k * ld + r], vv);
          s1.x += x1.x; s1.y += x1.y;
          s2.x += x2.x; s2.y += x2.y;
        }
      }
      s1.x = wave_sum(s1.x); s1.y = wave_sum(s1.y);
      s2.x = wave_sum(s2.x); s2.y = wave_sum(s2.y);
      if (lane == 0) { cp[k] = s1; cp[NB + k] = s2; }
    }
  }
  if (clk != nullptr) {
    __syncthreads();
    sample_end(clk);
  }
}

// Column sum (many workgroups, HS_ROWS rows each): y = sum of the partial
// products - V (W^H v) - W (V^H v), and, inside a panel, the next column with
// the panel's older pairs applied, COL = A[:, j+1] - V_<jj W_<jj[j+1]^H -
// W_<jj V_<jj[j+1]^H (the tail adds the newest pair).
#define HS_ROWS 256
template <typename T>
__global__ __launch_bounds__(HS_ROWS) void hetrd_sum_kernel(HetrdCtx<T> c, int j, int jj, int jn, int nch) {
  using C2 = typename C2T<T>::t;
  const int b = blockIdx.y;
  const int n = c.n, ld = c.ld;
  const C2* yp = c.Yp + (size_t)b * c.ncmax * ld;
  const C2* vc = c.Vc + (size_t)b * NB * ld;
  const C2* wc = c.Wc + (size_t)b * NB * ld;
  const C2* a = c.A + (size_t)b * ld * ld;
  C2* y = c.Y + (size_t)b * ld;
  C2* col = c.COL + (size_t)b * ld;
  __shared__ C2 s_c12[2 * NB];
  __shared__ C2 s_vj[NB], s_wj[NB];
  const int tid = threadIdx.x;
  if (tid < 2 * jj) {
    const int k = tid < jj ? tid : NB + tid - jj;
    const C2* cp = c.C12p + (size_t)b * c.ncmax * 2 * NB;
    C2 acc = mk<T>(0, 0);
    // unrolled: eight independent loads in flight per step of the (unchanged) sum order
#pragma unroll 8
    for (int ch = 0; ch < nch; ++ch) acc = c_add(acc, cp[(size_t)ch * 2 * NB + k]);
    s_c12[k] = acc;
  }
  const bool next = j + 1 < jn;
  if (next && tid < jj) {
    s_vj[tid] = vc[(size_t)tid * ld + j + 1];
    s_wj[tid] = wc[(size_t)tid * ld + j + 1];
  }
  __syncthreads();
  const int r = j + 1 + blockIdx.x * HS_ROWS + tid;
  if (r >= n) return;
  C2 s = mk<T>(0, 0);
#pragma unroll 8
  for (int ch = 0; ch < nch; ++ch) s = c_add(s, yp[(size_t)ch * ld + r]);
  C2 cl = next ? a[(size_t)r * ld + j + 1] : mk<T>(0, 0);
  for (int k = 0; k < jj; ++k) {
    const C2 vk = vc[(size_t)k * ld + r], wk = wc[(size_t)k * ld + r];
    const C2 t1 = c_mul(vk, s_c12[k]);
    const C2 t2 = c_mul(wk, s_c12[NB + k]);
    s.x -= t1.x + t2.x;
    s.y -= t1.y + t2.y;
    if (next) {
      const C2 u1 = c_mulc(vk, s_wj[k]);
      const C2 u2 = c_mulc(wk, s_vj[k]);
      cl.x -= u1.x + u2.x;
      cl.y -= u1.y + u2.y;
    }
  }
  y[r] = s;
  if (next) col[r] = cl;
}

// Column tail (one workgroup per matrix): y = sum of the nch partial
// products, w_j from y, then the reflector of column j+1 if it is in the panel.
template <typename T>
__global__ __launch_bounds__(TAIL_THREADS) void hetrd_tail_kernel(HetrdCtx<T> c, int j, int jj, int jn) {
  __shared__ T red[TAIL_THREADS / 64];
  const int b = blockIdx.x;
  latrd_finalize(c, b, j + 1, jj + 1, red);
  if (j + 1 < jn) latrd_next(c, b, j + 1, jj + 1, red);
}

// Finisher: unblocked Householder on the trailing block (rows/columns >= j0,
// at most HF_MAX of them) of each matrix, one workgroup per matrix, every
// column in one launch. The block's lower triangle is current (all panel
// updates applied); per column: reflector from column j, y = A22 v from the
// lower triangle (waves own rows; the A^H part goes to per-wave column
// accumulators), w = tau y - |tau|^2/2 (v^H y) v, A22 -= v w^H + w v^H on
// the lower triangle. Outputs (d, e, tau, reflectors in column j and VT)
// match the blocked path's.
#define HF_MAX 256
#define HF_THREADS 512
#define HF_WAVES (HF_THREADS / 64)
template <typename T>
__global__ __launch_bounds__(HF_THREADS) void hetrd_finish_kernel(HetrdCtx<T> c, int j0) {
  using C2 = typename C2T<T>::t;
  __shared__ T red[HF_WAVES];
  __shared__ C2 s_v[HF_MAX], s_y[HF_MAX], s_w[HF_MAX];
  __shared__ C2 s_col[HF_WAVES][HF_MAX];
  __shared__ C2 s_tau, s_scale;
  __shared__ T s_beta;
  __shared__ int s_skip;
  const int b = blockIdx.x;
  const int n = c.n, ld = c.ld;
  C2* a = c.A + (size_t)b * ld * ld;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const C2 zero = mk<T>(0, 0);
  for (int j = j0; j < n - 1; ++j) {
    const int base = j + 1;           // first row / column of A22
    const int m = n - base;
    // (1) reflector from column j (rows >= j+1)
    T part = 0;
    for (int r = j + 2 + tid; r < n; r += HF_THREADS) {
      const C2 x = a[(size_t)r * ld + j];
      part += x.x * x.x + x.y * x.y;
    }
    const T xn2 = block_sum(part, red);
    if (tid == 0) {
      const C2 alpha = a[(size_t)base * ld + j];
      if (xn2 == T(0) && alpha.y == T(0)) {
        s_skip = 1;
        s_tau = zero;
        s_beta = alpha.x;
        s_scale = zero;
      } else {
        s_skip = 0;
        const T nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        const T beta = alpha.x >= T(0) ? -nrm : nrm;
        s_beta = beta;
        s_tau = mk<T>((beta - alpha.x) / beta, -alpha.y / beta);
        const C2 den = mk<T>(alpha.x - beta, alpha.y);
        const T dd = den.x * den.x + den.y * den.y;
        s_scale = mk<T>(den.x / dd, -den.y / dd);
      }
      c.D[(size_t)b * ld + j] = (double)a[(size_t)j * ld + j].x;
      c.E[(size_t)b * ld + j] = (double)s_beta;
      c.TAU[(size_t)b * ld + j] = s_tau;
    }
    __syncthreads();
    const C2 scale = s_scale, tau = s_tau;
    const int skip = s_skip;
    for (int i = tid; i < m; i += HF_THREADS) {
      const int r = base + i;
      C2 vv = i == 0 ? mk<T>(1, 0) : (skip ? zero : c_mul(a[(size_t)r * ld + j], scale));
      s_v[i] = vv;
      a[(size_t)r * ld + j] = vv;
      if (c.VT != nullptr) c.VT[(size_t)b * ld * ld + (size_t)j * ld + r] = vv;
    }
    for (int q = tid; q < HF_WAVES * HF_MAX; q += HF_THREADS) (&s_col[0][0])[q] = zero;
    __syncthreads();
    if (tau.x == T(0) && tau.y == T(0)) continue;   // H_j = I: nothing to apply
    // (2) y = A22 v from the lower triangle
    for (int i = wv; i < m; i += HF_WAVES) {
      const int r = base + i;
      const C2 vr = s_v[i];
      T px = 0, py = 0;
      for (int q = lane; q <= i; q += 64) {
        const C2 x = a[(size_t)r * ld + base + q];
        const C2 vq = s_v[q];
        px += x.x * vq.x - x.y * vq.y;
        py += x.x * vq.y + x.y * vq.x;
        if (q < i) {   // conj(A[r][q]) v[r] -> y[q]
          C2 cc = s_col[wv][q];
          cc.x += x.x * vr.x + x.y * vr.y;
          cc.y += x.x * vr.y - x.y * vr.x;
          s_col[wv][q] = cc;
        }
      }
      px = wave_sum(px);
      py = wave_sum(py);
      if (lane == 0) s_y[i] = mk<T>(px, py);
    }
    __syncthreads();
    T vhy_part = 0;
    for (int i = tid; i < m; i += HF_THREADS) {
      C2 yy = s_y[i];
#pragma unroll
      for (int w2 = 0; w2 < HF_WAVES; ++w2) yy = c_add(yy, s_col[w2][i]);
      s_y[i] = yy;
      const C2 vv = s_v[i];
      vhy_part += vv.x * yy.x + vv.y * yy.y;
    }
    // (3) w = tau y - |tau|^2/2 (v^H y) v
    const T vhy = block_sum(vhy_part, red);
    const T t2 = T(0.5) * (tau.x * tau.x + tau.y * tau.y) * vhy;
    for (int i = tid; i < m; i += HF_THREADS) {
      const C2 p = c_mul(tau, s_y[i]);
      const C2 vv = s_v[i];
      s_w[i] = mk<T>(p.x - t2 * vv.x, p.y - t2 * vv.y);
    }
    __syncthreads();
    // (4) A22 -= v w^H + w v^H (lower triangle)
    for (int i = wv; i < m; i += HF_WAVES) {
      const int r = base + i;
      const C2 vr = s_v[i], wr = s_w[i];
      for (int q = lane; q <= i; q += 64) {
        C2 x = a[(size_t)r * ld + base + q];
        const C2 t1 = c_mulc(vr, s_w[q]);
        const C2 t3 = c_mulc(wr, s_v[q]);
        x.x -= t1.x + t3.x;
        x.y -= t1.y + t3.y;
        a[(size_t)r * ld + base + q] = x;
      }
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void hetrd_last_kernel(const typename C2T<T>::t* A, int n, int ld, double* D, double* E,
                                  typename C2T<T>::t* TAU) {
  const int b = blockIdx.x;
  if (threadIdx.x == 0) {
    D[(size_t)b * ld + n - 1] = (double)A[(size_t)b * ld * ld + (size_t)(n - 1) * ld + (n - 1)].x;
    E[(size_t)b * ld + n - 1] = 0.0;
    TAU[(size_t)b * ld + n - 1] = mk<T>(0, 0);
  }
}

size_t hetrd_partials_elems(int ld) { return (size_t)((ld + HT_TS - 1) / HT_TS) * ld; }
size_t hetrd_c12_elems(int ld) { return (size_t)((ld + HT_TS - 1) / HT_TS) * 2 * NB; }

template <typename T>
static void rank2k(hipStream_t st, int batch, int m, int k, typename C2T<T>::t* Vc, typename C2T<T>::t* Wc,
                   typename C2T<T>::t* C, int ld, long long sP, long long sM);
template <>
void rank2k<double>(hipStream_t st, int batch, int m, int k, double2* Vc, double2* Wc, double2* C, int ld,
                    long long sP, long long sM) {
  launch_zgemm(st, batch, m, m, k, -1.0, Vc, ld, sP, OP_T, Wc, ld, sP, OP_CONJ, 1.0, C, ld, sM, 0, 1);
  launch_zgemm(st, batch, m, m, k, -1.0, Wc, ld, sP, OP_T, Vc, ld, sP, OP_CONJ, 1.0, C, ld, sM, 0, 1);
}
template <>
void rank2k<float>(hipStream_t st, int batch, int m, int k, float2* Vc, float2* Wc, float2* C, int ld,
                   long long sP, long long sM) {
  launch_cgemm(st, batch, m, m, k, -1.0f, Vc, ld, sP, OP_T, Wc, ld, sP, OP_CONJ, 1.0f, C, ld, sM, 0, 1);
  launch_cgemm(st, batch, m, m, k, -1.0f, Wc, ld, sP, OP_T, Vc, ld, sP, OP_CONJ, 1.0f, C, ld, sM, 0, 1);
}

template <typename T>
static void hetrd_impl(hipStream_t st, typename C2T<T>::t* C, int n, int ld, int batch, double* D, double* E,
                       typename C2T<T>::t* TAU, typename C2T<T>::t* Vc, typename C2T<T>::t* Wc,
                       typename C2T<T>::t* Y, typename C2T<T>::t* Yp, typename C2T<T>::t* C12,
                       typename C2T<T>::t* VT) {
  HetrdCtx<T> c;
  c.A = C; c.n = n; c.ld = ld; c.Vc = Vc; c.Wc = Wc; c.Y = Y; c.Yp = Yp; c.C12 = C12; c.TAU = TAU; c.VT = VT;
  c.COL = Y + (size_t)batch * ld;   // second half of the Y buffer
  c.C12p = C12;
  c.D = D; c.E = E; c.ncmax = (ld + HT_TS - 1) / HT_TS;
  const long long sP = (long long)NB * ld, sM = (long long)ld * ld;
  // blocked panels until the trailing block fits the one-workgroup finisher
  // (a panel boundary jf with n - jf <= HF_MAX), then the finisher
  int jf = n - HF_MAX;
  jf = jf <= 0 ? 0 : ((jf + NB - 1) / NB) * NB;
  for (int j0 = 0; j0 < jf && j0 < n - 1; j0 += NB) {
    const int jn = min(j0 + NB, n - 1);
    hipLaunchKernelGGL(latrd_start_kernel<T>, dim3(batch), dim3(TAIL_THREADS), 0, st, c, j0);
    for (int j = j0; j < jn; ++j) {
      const int m = n - j - 1;
      const int nbk = (m + HT_TS - 1) / HT_TS;
      dim3 g(nbk * (nbk + 1) / 2, batch);
      // algorithmic bytes: one triangle of the Hermitian trailing block per matrix
      SampleSlot sl;
      if (sampler_take(SLK_SAMPLE_HETRD_STEP, (double)batch * m * (m + 1) / 2 * sizeof(typename C2T<T>::t), &sl))
        hipExtLaunchKernelGGL(hetrd_step_kernel<T>, g, dim3(HT_THREADS), 0, st, sl.e0, sl.e1, 0, c, j, j - j0,
                              sl.clk);
      else
        hipLaunchKernelGGL(hetrd_step_kernel<T>, g, dim3(HT_THREADS), 0, st, c, j, j - j0,
                           (unsigned long long*)nullptr);
      hipLaunchKernelGGL(hetrd_sum_kernel<T>, dim3((m + HS_ROWS - 1) / HS_ROWS, batch), dim3(HS_ROWS), 0, st, c, j,
                         j - j0, jn, nbk);
      hipLaunchKernelGGL(hetrd_tail_kernel<T>, dim3(batch), dim3(TAIL_THREADS), 0, st, c, j, j - j0, jn);
    }
    if (jn < n) rank2k<T>(st, batch, n - jn, jn - j0, Vc + jn, Wc + jn, C + (size_t)jn * ld + jn, ld, sP, sM);
  }
  if (jf < n - 1) hipLaunchKernelGGL(hetrd_finish_kernel<T>, dim3(batch), dim3(HF_THREADS), 0, st, c, jf);
  hipLaunchKernelGGL(hetrd_last_kernel<T>, dim3(batch), dim3(64), 0, st, C, n, ld, D, E, TAU);
}

void launch_hetrd_blocked(hipStream_t st, double2* C, int n, int ld, int batch, double* D, double* E, double2* TAU,
                          double2* Vc, double2* Wc, double2* Y, double2* Yp, double2* C12) {
  hetrd_impl<double>(st, C, n, ld, batch, D, E, TAU, Vc, Wc, Y, Yp, C12, nullptr);
}

void launch_hetrd_blocked_f32(hipStream_t st, float2* C, int n, int ld, int batch, double* D, double* E,
                              float2* TAU, float2* Vc, float2* Wc, float2* Y, float2* Yp, float2* C12,
                              float2* VT) {
  hetrd_impl<float>(st, C, n, ld, batch, D, E, TAU, Vc, Wc, Y, Yp, C12, VT);
}
