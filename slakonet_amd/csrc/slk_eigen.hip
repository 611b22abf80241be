// Batched generalized Hermitian eigensolver H z = e S z for MI355X (gfx950).
//
// Pipeline per matrix (same mathematics as eighb(scheme="chol"),
// slakonet/utils.py:585-816):
//   potrf   S = L L^H                          (FP64, one workgroup / matrix)
//   trinv   W = L^-1                           (FP64, column-parallel)
//   reduce  C = W H W^H                        (batched complex GEMMs)
//   hetrd   C = Q T Q^H, T real tridiagonal    (Householder, lower)
//   values  bisection on Sturm counts of T     (FP64, thread per eigenvalue)
//   vectors inverse iteration on T (thread per eigenvalue, Gram-Schmidt
//           inside clusters) into Z, then Y = Q Z (reflectors) and
//           z = W^H Y (GEMM)
// Matrices are row-major complex (interleaved re/im) with leading dimension
// ld and batch stride ld*ld.
#include <algorithm>

#include "slk_common.h"
#include "slakonet_amd.h"
#include "slk_dense.h"
#include "slk_edges.h"

#define EIG_THREADS 256
#define BISECT_LDS_MAX 4096
#define EIGH_MAX_N 16384

// ---------------------------------------------------------------------------
// Eigenvalues of the real symmetric tridiagonal (d, e) by bisection /
// multisection on Sturm counts, ascending order.
// ---------------------------------------------------------------------------
// 1/q from the hardware approximation (relative error 2^-24.4, measured on
// gfx950) refined by NEWTON Newton steps: 2 give full double precision
// (2^-53), 1 gives 2^-48.7, a relative perturbation of e^2 of that size in
// the Sturm recurrence (|q| >= pivmin keeps it finite).
template <int NEWTON>
__device__ __forceinline__ double rcp_refined(double q) {
  double r = __builtin_amdgcn_rcp(q);
#pragma unroll
  for (int i = 0; i < NEWTON; ++i) r = fma(fma(-q, r, 1.0), r, r);
  return r;
}

// Number of eigenvalues below x; e2[i] = e[i]^2.
template <int NEWTON>
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] * rcp_refined<NEWTON>(q);
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Gershgorin interval [lo, hi] of the tridiagonal (widened by a relative
// margin) and the pivot floor, reduced over the workgroup; with LDS = true d
// and e^2 are staged in sd / se on the way. Every kernel that calls it with
// the same blockDim gets the same bits, so grid points derived from [lo, hi]
// agree between kernels.
struct TriInterval {
  double lo, hi, piv;
  bool bad;   // some d or e is not finite (e.g. a chase that gave up): no eigenvalues
};
template <bool LDS>
__device__ TriInterval tri_interval(const double* d, const double* e, int n, double* sd, double* se) {
  __shared__ double s_red[4][EIG_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double lo = 1e300, hi = -1e300, emax = 0.0, bad = 0.0;
  for (int i = tid; i < n; i += blockDim.x) {
    const double di = d[i];
    const double ei = i < n - 1 ? e[i] : 0.0;
    const double ep = i > 0 ? e[i - 1] : 0.0;
    if (LDS) {
      sd[i] = di;
      se[i] = ei * ei;
    }
    const double r = fabs(ep) + fabs(ei);
    lo = fmin(lo, di - r);
    hi = fmax(hi, di + r);
    emax = fmax(emax, ei * ei);
    if (!isfinite(di) || !isfinite(ei)) bad = 1.0;
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  emax = wave_max(emax);
  bad = wave_max(bad);
  if (lane == 0) { s_red[0][wv] = lo; s_red[1][wv] = hi; s_red[2][wv] = emax; s_red[3][wv] = bad; }
  __syncthreads();
  lo = 1e300; hi = -1e300; emax = 0.0; bad = 0.0;
  for (int q = 0; q < EIG_THREADS / 64; ++q) {
    lo = fmin(lo, s_red[0][q]);
    hi = fmax(hi, s_red[1][q]);
    emax = fmax(emax, s_red[2][q]);
    bad = fmax(bad, s_red[3][q]);
  }
  const double span = fmax(fabs(lo), fabs(hi));
  TriInterval t;
  t.lo = lo - 2.0 * 2.2e-16 * span * n - 1e-300;
  t.hi = hi + 2.0 * 2.2e-16 * span * n + 1e-300;
  t.piv = fmax(2.2e-308, 2.2e-308 * emax) * 1e2;
  t.bad = bad > 0.0;
  return t;
}

// Point j of the equidistant grid lo = x_0 < ... < x_M = hi (explicit fma so
// that every kernel rounds it the same way).
__device__ __forceinline__ double grid_point(const TriInterval& t, int j, int M) {
  return fma(t.hi - t.lo, (double)j / (double)M, t.lo);
}

// Sturm counts at the M + 1 grid points: cnt[b][j] = #eigenvalues below x_j.
// One thread per point; they give every eigenvalue a starting cell of width
// (hi - lo) / M for the bisection, in place of log2(M) halvings per eigenvalue.
template <bool LDS, int NEWTON>
__global__ __launch_bounds__(EIG_THREADS) void sturm_grid_kernel(const double* D, const double* Ed, int n, int ld,
                                                                 int M, int* cnt, const double* E2g) {
  const int b = blockIdx.y;
  extern __shared__ double sm[];
  const double* d = D + (size_t)b * ld;
  double* sd = LDS ? sm : const_cast<double*>(d);
  double* se = LDS ? sm + n : const_cast<double*>(E2g + (size_t)b * ld);
  const TriInterval t = tri_interval<LDS>(d, Ed + (size_t)b * ld, n, sd, se);
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > M) return;
  int c;
  if (j == 0) c = 0;
  else if (j == M) c = n;
  else c = sturm_count<NEWTON>(sd, se, n, grid_point(t, j, M), t.piv);
  cnt[(size_t)b * (M + 1) + j] = c;
}

// Groups of G lanes (G a power of two <= 64) share one eigenvalue; each step
// evaluates G Sturm counts at equidistant interior points and keeps the
// sub-interval where the count crosses the eigenvalue index ((G+1)x shrink per
// step). G = 1 is plain bisection with 64 eigenvalues per wave; larger G
// trades total work for parallelism when few eigenvalues are in flight.
// Eigenvalue idx starts from the grid cell [x_j, x_j+1] whose counts bracket
// it, cnt[j] <= idx < cnt[j+1] (binary search over the grid counts), and
// stops once the interval is below rtol relative (plus the pivot floor).
// LDS = true: d and e^2 staged in LDS (2n doubles, n <= 4096); false: read
// from global memory (d in place, e^2 from E2g written by square_e_kernel).
template <bool LDS, int NEWTON>
__global__ __launch_bounds__(EIG_THREADS) void bisect_kernel(const double* D, const double* Ed, int n, int ld,
                                                             double* Wout, int wld, int G, const double* E2g,
                                                             const int* cnt, int M, double rtol) {
  const int b = blockIdx.y;
  const double* d = D + (size_t)b * ld;
  extern __shared__ double sm[];
  double* sd = LDS ? sm : const_cast<double*>(d);
  double* se = LDS ? sm + n : const_cast<double*>(E2g + (size_t)b * ld);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const TriInterval t = tri_interval<LDS>(d, Ed + (size_t)b * ld, n, sd, se);
  const int per_wave = 64 / G;
  const int grp = lane / G, sub = lane % G;
  const int idx = (blockIdx.x * (EIG_THREADS / 64) + wv) * per_wave + grp;
  if ((blockIdx.x * (EIG_THREADS / 64) + wv) * per_wave >= n) return;   // whole wave idle
  if (t.bad) {   // uniform over the workgroup
    if (sub == 0 && idx < n) Wout[(size_t)b * wld + idx] = __builtin_nan("");
    return;
  }
  double lo = t.lo, hi = t.hi;
  const double piv = t.piv;
  if (idx < n) {
    const int* cb = cnt + (size_t)b * (M + 1);
    if (cb[0] <= idx && idx < cb[M]) {
      int jl = 0, jh = M;
      while (jh - jl > 1) {
        const int jm = (jl + jh) >> 1;
        if (cb[jm] <= idx) jl = jm; else jh = jm;
      }
      lo = grid_point(t, jl, M);
      hi = grid_point(t, jh, M);
    }
  }
  const unsigned long long gmask = G == 64 ? ~0ULL : ((1ULL << G) - 1ULL);
  const double inv = 1.0 / (double)(G + 1);
  bool done = idx >= n;
  for (int it = 0; it < 256; ++it) {
    const double width = hi - lo;
    if (!done && width <= rtol * fmax(fabs(lo), fabs(hi)) + piv) done = true;
    if (__ballot(!done) == 0ULL) break;
    const double x = lo + width * (double)(sub + 1) * inv;
    const bool above = !done && sturm_count<NEWTON>(sd, se, n, x, piv) > idx;
    const unsigned long long mask = (__ballot(above) >> (grp * G)) & gmask;
    if (!done) {
      const int first = mask ? (__ffsll((long long)mask) - 1) : G;
      const double nlo = first == 0 ? lo : lo + width * (double)first * inv;
      const double nhi = first == G ? hi : lo + width * (double)(first + 1) * inv;
      if (nlo == lo && nhi == hi) done = true;
      lo = nlo;
      hi = nhi;
    }
  }
  if (sub == 0 && idx < n) Wout[(size_t)b * wld + idx] = 0.5 * (lo + hi);
}

// e2[b][i] = e[b][i]^2 (i < n-1), 0 at n-1: global Sturm-count input for n > 4096.
__global__ void square_e_kernel(const double* Ed, int n, int ld, double* E2) {
  const int b = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double ei = i < n - 1 ? Ed[(size_t)b * ld + i] : 0.0;
    E2[(size_t)b * ld + i] = ei * ei;
  }
}

// Lanes per eigenvalue so that about 2^14 lanes are busy. G lanes shrink the
// interval (G+1)x per step for G Sturm counts, against 2x for one: G = 32
// spends ~6x the counts of plain bisection for ~4.6x fewer steps. Batches
// share the device with other streams, so the work matters more than the
// latency of this stage (< 1 ms of a large matrix's solve either way).
static int bisect_group(int n, int batch) {
  const double total = (double)n * batch;
  int g = 1;
  while (g < 64 && total * g * 2 <= 16384.0) g *= 2;
  return g;
}

// ---------------------------------------------------------------------------
// Eigenvectors of the FP64 tridiagonal T (d, e) by inverse iteration from the
// bisection eigenvalues (the LAPACK dstein method: LU of T - sigma I with
// partial pivoting, dgttrf / dgttrs, a fixed number of solves from a
// deterministic start vector). One thread per eigenvalue, all eigenvalues of
// all matrices in parallel. Eigenvalues closer than STEIN_CLUSTER * |T| to
// their neighbour form a cluster; its first thread computes all of its
// vectors in order, orthogonalising each against the cluster's earlier
// vectors (modified Gram-Schmidt) after every solve. Outside clusters the
// vectors are orthogonal to eps |T| / gap < 3e-10. A split tridiagonal
// (e_i = 0 after a degenerate spectrum) needs nothing special: degenerate
// eigenvalues are a cluster and the random starts plus Gram-Schmidt give an
// orthonormal basis of their space.
// Z is row-major here, Z[row * ld + col] (column = eigenvector), so that the
// threads of a wave write consecutive addresses. Per-thread scratch (dl, dd,
// du, du2, pivots, y: 6 n doubles) is interleaved with stride nthr (all
// eigenvalue threads of the matrix), so that step i of every thread of a wave
// touches consecutive addresses too.
// ---------------------------------------------------------------------------
#define STEIN_THREADS 256
#define STEIN_CLUSTER 1e-6
#define STEIN_SCR 6

__global__ __launch_bounds__(STEIN_THREADS) void stein_kernel(const double* D, const double* Ed, const double* W,
                                                              int wld, int n, int ld, double* Zr, double* scr,
                                                              int nthr) {
  const int b = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const double* d = D + (size_t)b * ld;
  const double* e = Ed + (size_t)b * ld;
  const double* w = W + (size_t)b * wld;
  double* z = Zr + (size_t)b * ld * ld;
  // |T| = max row sum, reduced over the workgroup
  __shared__ double s_red[STEIN_THREADS / 64];
  double tn = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double ei = i + 1 < n ? e[i] : 0.0;
    const double ep = i > 0 ? e[i - 1] : 0.0;
    tn = fmax(tn, fabs(d[i]) + fabs(ei) + fabs(ep));
  }
  tn = wave_max(tn);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = tn;
  __syncthreads();
  tn = 0.0;
  for (int q = 0; q < STEIN_THREADS / 64; ++q) tn = fmax(tn, s_red[q]);
  if (j >= n) return;
  const double thr = STEIN_CLUSTER * fmax(tn, 1e-300);
  if (j > 0 && w[j] - w[j - 1] <= thr) return;          // not the first of its cluster
  int last = j;
  while (last + 1 < n && w[last + 1] - w[last] <= thr) ++last;
  const double tiny = 2.2e-16 * fmax(tn, 1e-300);
  const size_t st = (size_t)nthr;
  double* base = scr + (size_t)b * STEIN_SCR * nthr * n + j;
  double *dl = base, *dd = base + st * n, *du = base + 2 * st * n, *du2 = base + 3 * st * n;
  double *pv = base + 4 * st * n, *y = base + 5 * st * n;
  const int its = last > j ? 4 : 3;
  // Rows split off as 1 x 1 blocks (e = 0 on both sides, e.g. identity
  // padding) are exact eigenvectors: a cluster member whose eigenvalue equals
  // such a row's diagonal (to the bisection's accuracy) takes that unit
  // vector, so a large degenerate cluster costs O(n) per member instead of
  // Gram-Schmidt. Rows are scanned once, in order; a member without a match
  // goes through inverse iteration.
  const double tol_eig = 16.0 * 2.2e-16 * fmax(tn, 1e-300);
  int ri = 0;
  double sig_prev = -1e300;
  for (int m = j; m <= last; ++m) {
    if (last > j) {
      int irow = -1;
      for (; ri < n; ++ri) {
        const bool iso = (ri == 0 || e[ri - 1] == 0.0) && (ri + 1 == n || e[ri] == 0.0);
        if (iso && fabs(d[ri] - w[m]) <= tol_eig) {
          irow = ri++;
          break;
        }
      }
      if (irow >= 0) {
        for (int r = 0; r < n; ++r) z[(size_t)r * ld + m] = r == irow ? 1.0 : 0.0;
        continue;
      }
    }
    // separate the shifts of a cluster by a few ulps (dstein's perturbation)
    double sig = w[m];
    const double pert = 10.0 * 2.2e-16 * fmax(fabs(sig), tn);
    if (sig - sig_prev < pert) sig = sig_prev + pert;
    sig_prev = sig;
    // LU with partial pivoting of T - sig I (dgttrf); the diagonal carried in a register
    double dcur = d[0] - sig;
    double ucur = n > 1 ? e[0] : 0.0;   // current super-diagonal of row i
    for (int i = 0; i + 1 < n; ++i) {
      const double li = e[i];                       // sub-diagonal below row i
      double dnext = d[i + 1] - sig;
      const double unext = i + 2 < n ? e[i + 1] : 0.0;
      if (fabs(dcur) >= fabs(li)) {
        if (dcur == 0.0) dcur = tiny;
        const double f = li / dcur;
        dl[i * st] = f;
        dd[i * st] = dcur;
        du[i * st] = ucur;
        du2[i * st] = 0.0;
        pv[i * st] = 0.0;
        dnext -= f * ucur;
        ucur = unext;
      } else {
        const double f = dcur / li;
        dd[i * st] = li;
        dl[i * st] = f;
        du[i * st] = dnext;
        du2[i * st] = unext;
        pv[i * st] = 1.0;
        dnext = ucur - f * dnext;
        ucur = -f * unext;
      }
      dcur = dnext;
    }
    if (dcur == 0.0) dcur = tiny;
    dd[(size_t)(n - 1) * st] = dcur;
    // deterministic start vector, different for each member of a cluster
    const double ph = 0.1 + 1.3 * (m - j);
    for (int i = 0; i < n; ++i) y[i * st] = 1.0 + 0.37 * sin(0.7 * i + ph);
    for (int it = 0; it < its; ++it) {
      // forward: row interchanges and multipliers
      double ycur = y[0];
      for (int i = 0; i + 1 < n; ++i) {
        double ynext = y[(i + 1) * st];
        const double f = dl[i * st];
        if (pv[i * st] == 0.0) {
          ynext -= f * ycur;
          y[i * st] = ycur;
        } else {
          y[i * st] = ynext;
          ynext = ycur - f * ynext;
        }
        ycur = ynext;
      }
      y[(size_t)(n - 1) * st] = ycur;
      // back substitution with U (dd, du, du2); scaled so that nothing overflows
      double y1 = 0.0, y2 = 0.0, amax = 0.0;
      for (int i = n - 1; i >= 0; --i) {
        const double ui = i + 1 < n ? du[i * st] : 0.0;
        const double u2 = i + 2 < n ? du2[i * st] : 0.0;
        const double yi = (y[i * st] - ui * y1 - u2 * y2) / dd[i * st];
        y[i * st] = yi;
        amax = fmax(amax, fabs(yi));
        y2 = y1;
        y1 = yi;
      }
      const double sc = amax > 0.0 ? 1.0 / amax : 1.0;
      // orthogonalise against the cluster's earlier vectors (columns j .. m-1)
      for (int k = j; k < m; ++k) {
        double dot = 0.0;
        for (int i = 0; i < n; ++i) dot += z[(size_t)i * ld + k] * y[i * st];
        for (int i = 0; i < n; ++i) y[i * st] -= dot * z[(size_t)i * ld + k];
      }
      double nr = 0.0;
      for (int i = 0; i < n; ++i) {
        const double v = y[i * st] * sc;
        y[i * st] = v;
        nr += v * v;
      }
      const double inv = nr > 0.0 ? 1.0 / sqrt(nr) : 0.0;
      for (int i = 0; i < n; ++i) y[i * st] *= inv;
    }
    for (int i = 0; i < n; ++i) z[(size_t)i * ld + m] = y[i * st];
  }
}

// ---------------------------------------------------------------------------
// Y = Q Z with Q = H_0 H_1 ... H_{n-2}, H_j = I - tau_j v_j v_j^H (v_j[j+1] = 1,
// entries below in column j of C): blocks of QB reflectors as compact WY,
// Q_b = I - V_b T_b V_b^H, applied last block first with three MFMA GEMMs
// each (W = V^H Y, W = T W, Y -= V W). Z is real row-major (columns are
// eigenvectors); Y complex row-major.
// ---------------------------------------------------------------------------
#define QB 32

// Y = complex(Z) on the n x n block.
__global__ void z_to_y_kernel(const double* __restrict__ Zr, double2* __restrict__ Y, int n, int ld) {
  const int b = blockIdx.y;
  const size_t off = (size_t)b * ld * ld;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < (size_t)n * n; q += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(q / n), c = (int)(q % n);
    Y[off + (size_t)r * ld + c] = make_double2(Zr[off + (size_t)r * ld + c], 0.0);
  }
}

// V_b [m x QB] (row stride QB) of reflectors j0 .. j0+nb-1: row r is matrix row
// j0+1+r; reflector c has its unit at local row c, zeros above, C below.
__global__ void qb_pack_kernel(const double2* __restrict__ Cm, int ld, long long sM, int j0, int nb, int m,
                               double2* __restrict__ Vb, long long sV) {
  const int b = blockIdx.y;
  const double2* a = Cm + (size_t)b * sM;
  double2* v = Vb + (size_t)b * sV;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < m * QB; q += gridDim.x * blockDim.x) {
    const int r = q / QB, c = q % QB;
    double2 x = make_double2(0.0, 0.0);
    if (c < nb) {
      if (r == c) x = make_double2(1.0, 0.0);
      else if (r > c) x = a[(size_t)(j0 + 1 + r) * ld + j0 + c];
    }
    v[q] = x;
  }
}

// T_b (QB x QB, upper) from G = V^H V and tau (LAPACK zlarft, forward,
// columnwise): T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] G[0:j, j].
// One workgroup of QB threads per matrix; thread i owns row i.
__global__ void qb_tfactor_kernel(const double2* __restrict__ G, const double2* __restrict__ TAU, int ld, int j0,
                                  int nb, double2* __restrict__ T) {
  const int b = blockIdx.x;
  const double2* g = G + (size_t)b * QB * QB;
  double2* t = T + (size_t)b * QB * QB;
  const double2* tau = TAU + (size_t)b * ld + j0;
  __shared__ double2 s_t[QB][QB + 1];
  const int i = threadIdx.x;
  for (int c = 0; c < QB; ++c) s_t[i][c] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const double2 tj = tau[j];
    if (i < j) {
      double2 acc = make_double2(0.0, 0.0);
      for (int k = i; k < j; ++k) acc = c_add(acc, c_mul(s_t[i][k], g[k * QB + j]));
      const double2 v = c_mul(tj, acc);
      s_t[i][j] = make_double2(-v.x, -v.y);
    } else if (i == j) {
      s_t[i][j] = tj;
    }
    __syncthreads();
  }
  for (int c = 0; c < QB; ++c) t[i * QB + c] = s_t[i][c];
}

// Scratch per matrix: V_b (ld x QB), G and T (QB x QB each), W and W2 (QB x ld each).
static size_t apply_q_scratch_elems(int ld) { return (size_t)ld * QB + 2 * QB * QB + 2 * (size_t)QB * ld; }

static void launch_apply_q(hipStream_t st, const double2* Cm, const double2* TAU, const double* Zr, double2* Y,
                           int n, int ld, int batch, double2* scr, bool re) {
  const long long sM = (long long)ld * ld;
  const long long sV = (long long)ld * QB, sG = (long long)QB * QB, sW = (long long)QB * ld;
  double2* Vb = scr;                          // [batch][ld][QB]
  double2* G = Vb + (size_t)batch * sV;       // [batch][QB][QB]
  double2* T = G + (size_t)batch * sG;        // [batch][QB][QB]
  double2* W = T + (size_t)batch * sG;        // [batch][QB][ld]
  double2* W2 = W + (size_t)batch * sW;       // [batch][QB][ld]
  hipLaunchKernelGGL(z_to_y_kernel, dim3(256, batch), dim3(256), 0, st, Zr, Y, n, ld);
  const int nref = n - 1;   // reflectors 0 .. n-2
  if (nref <= 0) return;
  for (int j0 = ((nref - 1) / QB) * QB; j0 >= 0; j0 -= QB) {
    const int nb = min(QB, nref - j0);
    const int m = n - j0 - 1;
    double2* yr = Y + (size_t)(j0 + 1) * ld;
    hipLaunchKernelGGL(qb_pack_kernel, dim3((m * QB + 255) / 256, batch), dim3(256), 0, st, Cm, ld, sM, j0, nb, m, Vb,
                       sV);
    launch_zgemm(st, batch, QB, QB, m, 1.0, Vb, QB, sV, OP_C, Vb, QB, sV, OP_N, 0.0, G, QB, sG, 0, 0, re);
    hipLaunchKernelGGL(qb_tfactor_kernel, dim3(batch), dim3(QB), 0, st, G, TAU, ld, j0, nb, T);
    launch_zgemm(st, batch, QB, n, m, 1.0, Vb, QB, sV, OP_C, yr, ld, sM, OP_N, 0.0, W, ld, sW, 0, 0, re);
    launch_zgemm(st, batch, QB, n, QB, 1.0, T, QB, sG, OP_N, W, ld, sW, OP_N, 0.0, W2, ld, sW, 0, 0, re);
    launch_zgemm(st, batch, m, n, QB, -1.0, Vb, QB, sV, OP_N, W2, ld, sW, OP_N, 1.0, yr, ld, sM, 0, 0, re);
  }
}

// ---------------------------------------------------------------------------
// Band-edge refinement of the precision form. For each structure, the VBM
// and CBM states picked by the edge kernel get an FP64 Rayleigh quotient
// rho = z^H H z / z^H S z with the FP64 H and S of the build, where z is the
// state's eigenvector carried back from the FP32 reduction:
//   y: inverse iteration on the FP64 tridiagonal T (LU with partial
//      pivoting, LAPACK dgttrf/dgttrs, three solves);
//   x = Q y: the FP32 Householder reflectors, applied in FP64 arithmetic;
//   z = W^H x with W = L^-1 in FP64.
// z carries the FP32 reduction error (~1e-7 relative); the quotient's error
// is second order in it. Four launches per batch: refine_tri (one
// workgroup per (structure, edge): T solve and reflectors), refine_wh
// (z = W^H x over 64-column blocks), refine_rq (partial z^H H z, z^H S z over
// 64-row blocks), refine_commit (sums, writes w and the edge rows).
#define RF_THREADS 256
#define RF_BLK 64           // columns (refine_wh) / rows (refine_rq) per workgroup
#define RF_SCR 10           // scratch doubles per ld per (structure, edge)

struct RefineCtx {
  const double2* H;
  const double2* S;
  const double2* Wm;        // W = L^-1, FP64
  const float2* R;          // FP32 reflectors as rows (row j = reflector j, entries r >= j + off)
  const float2* TAU;
  const float2* V2;         // nullable: band-chase reflectors (row s = sweep s, tau in place of each leading 1)
  int off;                  // first row of reflector j is j + off (1: one-stage, SLK_BW: two-stage)
  int n;                    // dimension of the reduced matrices (chase block partition)
  const double* D;
  const double* E;
  int ld;
  const double* w;
  const SlkSpectrum* spec;
  const double* edges;
  double* scratch;          // [n_spec * 2][RF_SCR * ld]
  int* meta;                // [n_spec * 2][2]: band (-1: skip), matrix index
  double* part;             // [n_spec * 2][nblk][2]: partial num, den
  int nblk;
  int lds_x;                // 1: refine_tri was launched with ld complex doubles of LDS for x
  int split;                // 1: refine_tri stops after the tridiagonal solve (refine_q1_kernel follows)
};

__device__ __forceinline__ double* rf_base(const RefineCtx& c, int v) {
  return c.scratch + (size_t)v * RF_SCR * c.ld;
}

// One wave per (structure, edge). Lane 0 runs the O(m) tridiagonal solves
// with loads issued RF_CH at a time; the wave then applies the reflectors
// (rows of VT, contiguous) with wave reductions only. x lives in LDS when
// the padded dimension ld <= RF_LDS_X, else in the scratch slot.
#define RF_CH 8
#define RF_LDS_X 2560

// NQ > 0: reflector elements per lane in registers (ld <= 64 NQ, x in LDS);
// NQ = 0: generic path (x in LDS or scratch, reflectors re-read).
template <int NQ>
__global__ __launch_bounds__(64) void refine_tri_kernel(RefineCtx c) {
  const int s = blockIdx.x, edge = blockIdx.y, v = s * 2 + edge;
  const SlkSpectrum sp = c.spec[s];
  const double* row = c.edges + (size_t)s * SLK_EDGE_FIELDS;
  const int lane = threadIdx.x;
  const int ld = c.ld;
  extern __shared__ double2 s_x[];
  if (row[E_NVAL] <= 0.0 || row[E_NCON] <= 0.0) {
    if (lane == 0) c.meta[2 * v] = -1;
    return;
  }
  const double mu = row[E_MU];
  const int k = (int)row[edge == 0 ? E_VBMK : E_CBMK];
  const int m = sp.nb;
  const double* wk = c.w + sp.w0 + (size_t)k * sp.wld;
  int cnt = 0;
  for (int b = lane; b < m; b += 64) cnt += (wk[b] * H2E_OUT < mu * H2E_OUT) ? 1 : 0;
  cnt = wave_sum(cnt);
  const int band = edge == 0 ? cnt - 1 : cnt;
  const size_t mat = (size_t)sp.kw0 + k;
  if (lane == 0) {
    c.meta[2 * v] = (band < 0 || band >= m) ? -1 : band;
    c.meta[2 * v + 1] = (int)mat;
  }
  if (band < 0 || band >= m) return;
  const double lam = wk[band];
  const double* d = c.D + mat * ld;
  const double* e = c.E + mat * ld;
  double* base = rf_base(c, v);
  double *dl = base, *dd = base + ld, *du = base + 2 * (size_t)ld, *du2 = base + 3 * (size_t)ld;
  double* piv = base + 4 * (size_t)ld;
  double* y = base + 5 * (size_t)ld;
  double2* x = c.lds_x ? s_x : reinterpret_cast<double2*>(base + 6 * (size_t)ld);

  // copies with the whole wave: dd = d - lam, dl = du = e, du2 = 0, tn = max row sum
  double tn = 0.0;
  for (int i = lane; i < m; i += 64) {
    const double di = d[i];
    const double ei = i + 1 < m ? e[i] : 0.0;
    const double ep = i > 0 ? e[i - 1] : 0.0;
    tn = fmax(tn, fabs(di) + fabs(ei) + fabs(ep));
    dd[i] = di - lam;
    dl[i] = ei;
    du[i] = ei;
    du2[i] = 0.0;
    y[i] = 1.0 + 0.37 * sin(0.7 * i + 0.1);
  }
  tn = wave_max(tn);
  __threadfence_block();
  if (lane == 0) {
    const double tiny = 2.2e-16 * fmax(tn, 1e-300);
    // LU with partial pivoting (dgttrf); the diagonal element carried in a register
    double dcur = dd[0];
    for (int i0 = 0; i0 + 1 < m; i0 += RF_CH) {
      double rdl[RF_CH], rdu[RF_CH], rdu1[RF_CH], rdn[RF_CH];
#pragma unroll
      for (int q = 0; q < RF_CH; ++q) {
        const int i = i0 + q;
        rdl[q] = i + 1 < m ? dl[i] : 0.0;
        rdu[q] = i + 1 < m ? du[i] : 0.0;
        rdu1[q] = i + 2 < m ? du[i + 1] : 0.0;
        rdn[q] = i + 1 < m ? dd[i + 1] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < RF_CH; ++q) {
        const int i = i0 + q;
        if (i + 1 >= m) break;
        double dnext = rdn[q];
        const double dui1 = q + 1 < RF_CH ? rdu[q + 1] : rdu1[q];
        if (fabs(dcur) >= fabs(rdl[q])) {
          if (dcur == 0.0) dcur = tiny;
          const double f = rdl[q] / dcur;
          dl[i] = f;
          dd[i] = dcur;
          dnext -= f * rdu[q];
          piv[i] = 0.0;
        } else {
          const double f = dcur / rdl[q];
          dd[i] = rdl[q];
          dl[i] = f;
          du[i] = dnext;
          dnext = rdu[q] - f * dnext;
          if (i + 2 < m) {
            du2[i] = dui1;
            du[i + 1] = -f * dui1;
            if (q + 1 < RF_CH) rdu[q + 1] = -f * dui1; else rdu1[q] = -f * dui1;
          }
          piv[i] = 1.0;
        }
        dcur = dnext;
      }
    }
    if (dcur == 0.0) dcur = tiny;
    dd[m - 1] = dcur;
  }
  __threadfence_block();
  for (int it = 0; it < 2; ++it) {
    if (lane == 0) {
      // forward: apply the row interchanges and multipliers
      double ycur = y[0];
      for (int i0 = 0; i0 + 1 < m; i0 += RF_CH) {
        double rl[RF_CH], rp[RF_CH], ry[RF_CH];
#pragma unroll
        for (int q = 0; q < RF_CH; ++q) {
          const int i = i0 + q;
          rl[q] = i + 1 < m ? dl[i] : 0.0;
          rp[q] = i + 1 < m ? piv[i] : 0.0;
          ry[q] = i + 1 < m ? y[i + 1] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < RF_CH; ++q) {
          const int i = i0 + q;
          if (i + 1 >= m) break;
          double ynext = ry[q];
          if (rp[q] == 0.0) {
            ynext -= rl[q] * ycur;
            y[i] = ycur;
          } else {
            y[i] = ynext;
            ynext = ycur - rl[q] * ynext;
          }
          ycur = ynext;
        }
      }
      y[m - 1] = ycur;
      // back substitution with U (diagonal dd, super du, du2)
      double y1 = 0.0, y2 = 0.0;   // y[i+1], y[i+2]
      for (int i0 = m - 1; i0 >= 0; i0 -= RF_CH) {
        double rd[RF_CH], ru[RF_CH], ru2[RF_CH], ry[RF_CH];
#pragma unroll
        for (int q = 0; q < RF_CH; ++q) {
          const int i = i0 - q;
          rd[q] = i >= 0 ? dd[i] : 1.0;
          ru[q] = (i >= 0 && i + 1 < m) ? du[i] : 0.0;
          ru2[q] = (i >= 0 && i + 2 < m) ? du2[i] : 0.0;
          ry[q] = i >= 0 ? y[i] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < RF_CH; ++q) {
          const int i = i0 - q;
          if (i < 0) break;
          const double yi = (ry[q] - ru[q] * y1 - ru2[q] * y2) / rd[q];
          y[i] = yi;
          y2 = y1;
          y1 = yi;
        }
      }
    }
    __threadfence_block();
    double nr = 0.0;
    for (int i = lane; i < m; i += 64) nr += y[i] * y[i];
    nr = 1.0 / sqrt(wave_sum(nr));
    for (int i = lane; i < m; i += 64) y[i] *= nr;
    __threadfence_block();
  }
  if (c.V2 != nullptr || c.split) {
    // two-stage reduction, or one-stage with a large ld: refine_band_kernel or
    // refine_q1_kernel carries y back (x slot of scratch)
    double2* xg = reinterpret_cast<double2*>(base + 6 * (size_t)ld);
    for (int r = lane; r < m; r += 64) xg[r] = make_double2(y[r], 0.0);
    return;
  }
  for (int r = lane; r < m; r += 64) x[r] = make_double2(y[r], 0.0);
  __threadfence_block();
  // x = Q1 x, reflectors last to first: x -= v tau (v^H x), v = row j of VT
  const float2* vt = c.R + mat * ld * ld;
  const float2* tau = c.TAU + mat * ld;
  const int off = c.off;
  if (NQ > 0) {
    // each lane keeps its <= NQ elements of the current reflector in registers
    // and prefetches the next reflector while this one is applied (x in LDS)
    float2 vn[NQ > 0 ? NQ : 1];
    auto load = [&](int j, float2* dst) {
      const float2* vr = vt + (size_t)j * ld;
#pragma unroll
      for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) {
        const int r = j + off + lane + 64 * q;
        dst[q] = r < m ? vr[r] : make_float2(0.0f, 0.0f);
      }
    };
    if (m - 1 - off >= 0) load(m - 1 - off, vn);
    for (int j = m - 1 - off; j >= 0; --j) {
      float2 vc[NQ > 0 ? NQ : 1];
#pragma unroll
      for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) vc[q] = vn[q];
      if (j > 0) load(j - 1, vn);
      const float2 tf = tau[j];
      if (tf.x == 0.0f && tf.y == 0.0f) continue;
      double px = 0.0, py = 0.0;
#pragma unroll
      for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) {
        const int r = j + off + lane + 64 * q;
        if (r < m) {
          const double2 xv = x[r];
          px += (double)vc[q].x * xv.x + (double)vc[q].y * xv.y;
          py += (double)vc[q].x * xv.y - (double)vc[q].y * xv.x;
        }
      }
      px = wave_sum(px);
      py = wave_sum(py);
      const double sx = (double)tf.x * px - (double)tf.y * py, sy = (double)tf.x * py + (double)tf.y * px;
#pragma unroll
      for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) {
        const int r = j + off + lane + 64 * q;
        if (r < m) {
          double2 xv = x[r];
          xv.x -= (double)vc[q].x * sx - (double)vc[q].y * sy;
          xv.y -= (double)vc[q].x * sy + (double)vc[q].y * sx;
          x[r] = xv;
        }
      }
    }
  } else {
    for (int j = m - 1 - off; j >= 0; --j) {
      const float2 tf = tau[j];
      if (tf.x == 0.0f && tf.y == 0.0f) continue;
      const float2* vr = vt + (size_t)j * ld;
      double px = 0.0, py = 0.0;
      for (int r = j + off + lane; r < m; r += 64) {
        const float2 vf = vr[r];
        const double2 xv = x[r];
        px += (double)vf.x * xv.x + (double)vf.y * xv.y;
        py += (double)vf.x * xv.y - (double)vf.y * xv.x;
      }
      px = wave_sum(px);
      py = wave_sum(py);
      const double sx = (double)tf.x * px - (double)tf.y * py, sy = (double)tf.x * py + (double)tf.y * px;
      for (int r = j + off + lane; r < m; r += 64) {
        const float2 vf = vr[r];
        double2 xv = x[r];
        xv.x -= (double)vf.x * sx - (double)vf.y * sy;
        xv.y -= (double)vf.x * sy + (double)vf.y * sx;
        x[r] = xv;
      }
      __threadfence_block();   // x[r] written by one lane is read by another for the next reflector
    }
  }
  if (c.lds_x) {
    // hand x to the next kernel through the scratch slot
    double2* xg = reinterpret_cast<double2*>(base + 6 * (size_t)ld);
    for (int r = lane; r < m; r += 64) xg[r] = x[r];
  }
}

// Back-transform after the two-stage reduction, one 1024-thread workgroup per
// (structure, edge), x in LDS (1024 threads: 64 blocks of a sweep per round
// instead of 16; 37% less time than 256 threads on the serving stream):
//   x = Q2 x: chase reflectors, sweeps last to first; the blocks of one sweep
//     are disjoint, one per 16-lane group; row s of the reflectors is staged
//     in LDS while row s-1 is loaded into registers;
//   x = Q1 x: stage-1 reflectors one panel at a time in compact-WY form,
//     x -= V_p (T_p (V_p^H x)), panels last to first (T_p from the reduction).
#define RB_THREADS 1024
#define RB_LDS_MAX 2688     // x (FP64) and one reflector row (FP32) in 64 KB of LDS
template <int NP>
__global__ __launch_bounds__(RB_THREADS) void refine_band_kernel(RefineCtx c, const float2* TT, int npan) {
  const int s = blockIdx.x, edge = blockIdx.y, v = s * 2 + edge;
  if (c.meta[2 * v] < 0) return;
  const size_t mat = (size_t)c.meta[2 * v + 1];
  const int m = c.spec[s].nb, ld = c.ld, nn = c.n;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  extern __shared__ double2 rb_smem[];
  double2* x = rb_smem;
  float2* s_row = reinterpret_cast<float2*>(rb_smem + ld);
  __shared__ double2 s_red[RB_THREADS / 64][SLK_BW];
  __shared__ double2 s_t[SLK_BW];
  double2* xg = reinterpret_cast<double2*>(rf_base(c, v) + 6 * (size_t)ld);
  for (int r = tid; r < m; r += RB_THREADS) x[r] = xg[r];
  const float2* v2 = c.V2 + mat * ld * ld;
  float2 pre[NP];
  auto fetch = [&](int s2) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int r = s2 + 1 + tid + RB_THREADS * q;
      pre[q] = r < m ? v2[(size_t)s2 * ld + r] : make_float2(0.0f, 0.0f);
    }
  };
  if (m >= 2) fetch(m - 2);
  const int li = tid & 15, g = tid >> 4;
  for (int s2 = m - 2; s2 >= 0; --s2) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int r = s2 + 1 + tid + RB_THREADS * q;
      if (r < m) s_row[r] = pre[q];
    }
    if (s2 > 0) fetch(s2 - 1);
    __syncthreads();
    const int rr = nn - 3 - s2;
    const int nst = 1 + (rr > 0 ? rr / SLK_BW : 0);
    for (int t0 = 0; t0 < nst; t0 += RB_THREADS / 16) {
      const int t = t0 + g;
      const int q1 = s2 + 1 + SLK_BW * t;
      const bool act = t < nst && q1 < m;
      const int r = q1 + li;
      const bool in = act && r < m && r < nn;
      const float2 tf = act ? s_row[q1] : make_float2(0.0f, 0.0f);
      float2 vf = make_float2(0.0f, 0.0f);
      if (in) vf = li == 0 ? make_float2(1.0f, 0.0f) : s_row[r];
      const double2 xv = in ? x[r] : make_double2(0.0, 0.0);
      const double px = sum16((double)vf.x * xv.x + (double)vf.y * xv.y);
      const double py = sum16((double)vf.x * xv.y - (double)vf.y * xv.x);
      const double sx = (double)tf.x * px - (double)tf.y * py, sy = (double)tf.x * py + (double)tf.y * px;
      if (in) {
        double2 o = xv;
        o.x -= (double)vf.x * sx - (double)vf.y * sy;
        o.y -= (double)vf.x * sy + (double)vf.y * sx;
        x[r] = o;
      }
    }
    __syncthreads();
  }
  // x = Q1 x, panel by panel (reflector j0+k: row j0+k of VT, entries r >= j0+k+BW)
  const float2* vt = c.R + mat * ld * ld;
  const float2* tt = TT + mat * (size_t)npan * SLK_BW * SLK_BW;
  for (int p = (m - 1 - SLK_BW) / SLK_BW; p >= 0 && m - 1 >= SLK_BW; --p) {
    const int j0 = p * SLK_BW, r0 = j0 + SLK_BW;
    double2 acc[SLK_BW];
#pragma unroll
    for (int k = 0; k < SLK_BW; ++k) acc[k] = make_double2(0.0, 0.0);
    for (int r = r0 + tid; r < m; r += RB_THREADS) {
      const double2 xv = x[r];
#pragma unroll
      for (int k = 0; k < SLK_BW; ++k)
        if (r >= r0 + k) {
          const float2 vf = vt[(size_t)(j0 + k) * ld + r];
          acc[k].x += (double)vf.x * xv.x + (double)vf.y * xv.y;
          acc[k].y += (double)vf.x * xv.y - (double)vf.y * xv.x;
        }
    }
#pragma unroll
    for (int k = 0; k < SLK_BW; ++k) {
      double ax = sum16(acc[k].x), ay = sum16(acc[k].y);
      ax += __shfl_xor(ax, 16, 64);
      ay += __shfl_xor(ay, 16, 64);
      ax += __shfl_xor(ax, 32, 64);
      ay += __shfl_xor(ay, 32, 64);
      if (lane == 0) s_red[wv][k] = make_double2(ax, ay);
    }
    __syncthreads();
    if (tid < SLK_BW) {
      double2 t = make_double2(0.0, 0.0);
#pragma unroll
      for (int l = 0; l < SLK_BW; ++l) {
        double2 yl = make_double2(0.0, 0.0);
#pragma unroll
        for (int w = 0; w < RB_THREADS / 64; ++w) yl = c_add(yl, s_red[w][l]);
        const float2 tf = tt[(size_t)p * SLK_BW * SLK_BW + tid * SLK_BW + l];
        t.x += (double)tf.x * yl.x - (double)tf.y * yl.y;
        t.y += (double)tf.x * yl.y + (double)tf.y * yl.x;
      }
      s_t[tid] = t;
    }
    __syncthreads();
    for (int r = r0 + tid; r < m; r += RB_THREADS) {
      double2 xv = x[r];
#pragma unroll
      for (int k = 0; k < SLK_BW; ++k)
        if (r >= r0 + k) {
          const float2 vf = vt[(size_t)(j0 + k) * ld + r];
          const double2 tk = s_t[k];
          xv.x -= (double)vf.x * tk.x - (double)vf.y * tk.y;
          xv.y -= (double)vf.x * tk.y + (double)vf.y * tk.x;
        }
      x[r] = xv;
    }
    __syncthreads();
  }
  for (int r = tid; r < m; r += RB_THREADS) xg[r] = x[r];
}

// Back-transform after the one-stage reduction for large ld, one 1024-thread
// workgroup per (structure, edge): x = Q y with the FP32 reflectors (rows of
// VT, last to first), x -= v tau (v^H x). Thread t keeps rows t + 1024 q
// (q < NQ) of x in registers; with PF it fetches its elements of the next
// reflector (and the next tau) while the current one is applied. One barrier
// per reflector: the wave partials of v^H x alternate between two LDS rows.
#define RQ_THREADS 1024
#define RQ_MIN_LD (64 * 40)   // below: refine_tri's one-wave paths
template <int NQ, bool PF>
__global__ __launch_bounds__(RQ_THREADS) void refine_q1_kernel(RefineCtx c) {
  const int s = blockIdx.x, edge = blockIdx.y, v = s * 2 + edge;
  if (c.meta[2 * v] < 0) return;
  const size_t mat = (size_t)c.meta[2 * v + 1];
  const int m = c.spec[s].nb, ld = c.ld, off = c.off;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ double2 s_red[2][RQ_THREADS / 64];
  double2* xg = reinterpret_cast<double2*>(rf_base(c, v) + 6 * (size_t)ld);
  double2 x[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = tid + RQ_THREADS * q;
    x[q] = r < m ? xg[r] : make_double2(0.0, 0.0);
  }
  const float2* vt = c.R + mat * ld * ld;
  const float2* tau = c.TAU + mat * ld;
  // elements of reflector j outside rows [j + off, m) are zero
  auto load = [&](int j, float2* dst) {
    const float2* vr = vt + (size_t)j * ld;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int r = tid + RQ_THREADS * q;
      dst[q] = (r >= j + off && r < m) ? vr[r] : make_float2(0.0f, 0.0f);
    }
  };
  float2 vn[NQ];
  float2 tn = make_float2(0.0f, 0.0f);
  const int j1 = m - 1 - off;
  if (PF && j1 >= 0) {
    load(j1, vn);
    tn = tau[j1];
  }
  int par = 0;
  for (int j = j1; j >= 0; --j) {
    float2 vc[NQ];
    float2 tf;
    if (PF) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) vc[q] = vn[q];
      tf = tn;
      if (j > 0) {
        load(j - 1, vn);
        tn = tau[j - 1];
      }
    } else {
      load(j, vc);
      tf = tau[j];
    }
    if (tf.x == 0.0f && tf.y == 0.0f) continue;   // uniform: H_j = I
    double px = 0.0, py = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      px += (double)vc[q].x * x[q].x + (double)vc[q].y * x[q].y;
      py += (double)vc[q].x * x[q].y - (double)vc[q].y * x[q].x;
    }
    px = wave_sum(px);
    py = wave_sum(py);
    if (lane == 0) s_red[par][wv] = make_double2(px, py);
    __syncthreads();
    double dx = 0.0, dy = 0.0;
#pragma unroll
    for (int w = 0; w < RQ_THREADS / 64; ++w) {
      dx += s_red[par][w].x;
      dy += s_red[par][w].y;
    }
    par ^= 1;
    const double sx = (double)tf.x * dx - (double)tf.y * dy, sy = (double)tf.x * dy + (double)tf.y * dx;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      x[q].x -= (double)vc[q].x * sx - (double)vc[q].y * sy;
      x[q].y -= (double)vc[q].x * sy + (double)vc[q].y * sx;
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = tid + RQ_THREADS * q;
    if (r < m) xg[r] = x[q];
  }
}

// z[c] = sum_{r >= c} conj(W[r][c]) x[r] for the RF_BLK columns of this block
// (coalesced row segments; waves split the rows, lanes own the columns).
__global__ __launch_bounds__(RF_THREADS) void refine_wh_kernel(RefineCtx c) {
  const int v = blockIdx.y;
  const int band = c.meta[2 * v];
  if (band < 0) return;
  const int s = v / 2;
  const int m = c.spec[s].nb;
  const int c0 = blockIdx.x * RF_BLK;
  if (c0 >= m) return;
  const int ld = c.ld;
  const size_t mat = (size_t)c.meta[2 * v + 1];
  const double2* wm = c.Wm + mat * ld * ld;
  double* base = rf_base(c, v);
  const double2* x = reinterpret_cast<const double2*>(base + 6 * (size_t)ld);
  double2* z = reinterpret_cast<double2*>(base + 8 * (size_t)ld);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = c0 + lane;
  double ax = 0.0, ay = 0.0;
  if (col < m) {
    for (int r = c0 + wv; r < m; r += RF_THREADS / 64) {
      if (r < col) continue;
      const double2 wr = wm[(size_t)r * ld + col], xr = x[r];
      ax += wr.x * xr.x + wr.y * xr.y;
      ay += wr.x * xr.y - wr.y * xr.x;
    }
  }
  __shared__ double2 acc[RF_THREADS / 64][RF_BLK];
  acc[wv][lane] = make_double2(ax, ay);
  __syncthreads();
  if (wv == 0 && col < m) {
    double2 t = acc[0][lane];
    for (int q = 1; q < RF_THREADS / 64; ++q) t = c_add(t, acc[q][lane]);
    z[col] = t;
  }
}

// Partial Re(z^H H z), Re(z^H S z) over the RF_BLK rows of this block.
__global__ __launch_bounds__(RF_THREADS) void refine_rq_kernel(RefineCtx c) {
  const int v = blockIdx.y;
  const int band = c.meta[2 * v];
  if (band < 0) return;
  const int s = v / 2;
  const int m = c.spec[s].nb;
  const int r0 = blockIdx.x * RF_BLK;
  double* part = c.part + ((size_t)v * c.nblk + blockIdx.x) * 2;
  const int ld = c.ld;
  const size_t mat = (size_t)c.meta[2 * v + 1];
  const double2* hm = c.H + mat * ld * ld;
  const double2* sm = c.S + mat * ld * ld;
  const double2* z = reinterpret_cast<const double2*>(rf_base(c, v) + 8 * (size_t)ld);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ double red[RF_THREADS / 64];
  double num = 0.0, den = 0.0;
  for (int r = r0 + wv; r < min(r0 + RF_BLK, m); r += RF_THREADS / 64) {
    double hx = 0.0, hy = 0.0, sx = 0.0, sy = 0.0;
    for (int q = lane; q < m; q += 64) {
      const double2 zc = z[q], hv = hm[(size_t)r * ld + q], sv = sm[(size_t)r * ld + q];
      hx += hv.x * zc.x - hv.y * zc.y;
      hy += hv.x * zc.y + hv.y * zc.x;
      sx += sv.x * zc.x - sv.y * zc.y;
      sy += sv.x * zc.y + sv.y * zc.x;
    }
    hx = wave_sum(hx); hy = wave_sum(hy); sx = wave_sum(sx); sy = wave_sum(sy);
    const double2 zr = z[r];
    num += zr.x * hx + zr.y * hy;
    den += zr.x * sx + zr.y * sy;
  }
  if (lane != 0) { num = 0.0; den = 0.0; }
  num = block_sum(num, red);
  den = block_sum(den, red);
  if (tid == 0) {
    part[0] = r0 < m ? num : 0.0;
    part[1] = r0 < m ? den : 0.0;
  }
}

// rho per (structure, edge) from the partials; writes w and the edge row.
__global__ void refine_commit_kernel(RefineCtx c, double* w, double* edges, int n_spec) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_spec) return;
  const SlkSpectrum sp = c.spec[s];
  double* row = edges + (size_t)s * SLK_EDGE_FIELDS;
  const double mu = row[E_MU];
  for (int edge = 0; edge < 2; ++edge) {
    const int v = s * 2 + edge;
    const int band = c.meta[2 * v];
    if (band < 0) continue;
    const int nb = (sp.nb + RF_BLK - 1) / RF_BLK;
    double num = 0.0, den = 0.0;
    for (int b = 0; b < nb; ++b) {
      num += c.part[((size_t)v * c.nblk + b) * 2];
      den += c.part[((size_t)v * c.nblk + b) * 2 + 1];
    }
    const double rho = num / den;
    const int k = (int)row[edge == 0 ? E_VBMK : E_CBMK];
    w[sp.w0 + (size_t)k * sp.wld + band] = rho;
    if (edge == 0) {
      row[E_VBM] = rho * H2E_OUT;
      row[E_VB2] = rho;
      row[E_SVBM] = (rho - mu) * H2E_OUT;
    } else {
      row[E_CBM] = rho * H2E_OUT;
      row[E_CB2] = rho;
      row[E_SCBM] = (rho - mu) * H2E_OUT;
    }
  }
}

extern "C" {

// Workspace (per matrix). FP64 path: L/C, W, W*H (and Y for vectors), the
// bisection grid counts, panel V and W, Y, the mat-vec partial sums, C12, TAU,
// D, E, and the real Z for the vectors path. Mixed path: L and W (FP64), the
// grid counts, then FP32 copies W32, H32 (reused for C), W32*H32 and FP32
// panels.
struct EighWs {
  double2 *B0, *B1, *B2, *Vc, *Wc, *Y, *Yp, *C12, *TAU, *Band;
  float2 *F0, *F1, *F2, *fVc, *fWc, *fY, *fYp, *fC12, *fTAU, *fBand;
  double *D, *E, *Z;
  double* SV;   // vectors: inverse-iteration scratch (STEIN_SCR ld^2 doubles per matrix), then the back-transform's
  int* SG;   // Sturm counts on the bisection start grid, 2 ld + 1 per matrix
};

// Carves the workspace (base may be null to only size it); returns bytes.
static size_t carve(char* base, int ld, int batch, int vectors, int mixed, EighWs* w) {
  const size_t matel = (size_t)ld * ld;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return q;
  };
  *w = EighWs{};
  w->B0 = reinterpret_cast<double2*>(take(matel * batch * sizeof(double2)));
  w->B1 = reinterpret_cast<double2*>(take(matel * batch * sizeof(double2)));
  w->SG = reinterpret_cast<int*>(take((size_t)(2 * ld + 1) * batch * sizeof(int)));
  w->D = reinterpret_cast<double*>(take((size_t)ld * batch * sizeof(double)));
  w->E = reinterpret_cast<double*>(take((size_t)ld * batch * sizeof(double)));
  if (!mixed) {
    w->B2 = reinterpret_cast<double2*>(take(matel * batch * sizeof(double2)));
    w->Vc = reinterpret_cast<double2*>(take((size_t)SLK_NB * ld * batch * sizeof(double2)));
    w->Wc = reinterpret_cast<double2*>(take((size_t)SLK_NB * ld * batch * sizeof(double2)));
    w->Y = reinterpret_cast<double2*>(take((size_t)2 * ld * batch * sizeof(double2)));
    w->Yp = reinterpret_cast<double2*>(take(hetrd_partials_elems(ld) * batch * sizeof(double2)));
    w->C12 = reinterpret_cast<double2*>(take(hetrd_c12_elems(ld) * batch * sizeof(double2)));
    w->TAU = reinterpret_cast<double2*>(take((size_t)ld * batch * sizeof(double2)));
    w->Band = reinterpret_cast<double2*>(take(band_scratch_elems(ld) * batch * sizeof(double2)));
    w->Z = vectors ? reinterpret_cast<double*>(take(matel * batch * sizeof(double))) : nullptr;
    // inverse-iteration scratch, reused by the back-transform's blocks
    const size_t sv = std::max((size_t)STEIN_SCR * matel, 2 * apply_q_scratch_elems(ld));
    w->SV = vectors ? reinterpret_cast<double*>(take(sv * batch * sizeof(double))) : nullptr;
  } else {
    w->F0 = reinterpret_cast<float2*>(take(matel * batch * sizeof(float2)));
    w->F1 = reinterpret_cast<float2*>(take(matel * batch * sizeof(float2)));
    w->F2 = reinterpret_cast<float2*>(take(matel * batch * sizeof(float2)));
    w->fVc = reinterpret_cast<float2*>(take((size_t)SLK_NB * ld * batch * sizeof(float2)));
    w->fWc = reinterpret_cast<float2*>(take((size_t)SLK_NB * ld * batch * sizeof(float2)));
    w->fY = reinterpret_cast<float2*>(take((size_t)2 * ld * batch * sizeof(float2)));
    w->fYp = reinterpret_cast<float2*>(take(hetrd_partials_elems(ld) * batch * sizeof(float2)));
    w->fC12 = reinterpret_cast<float2*>(take(hetrd_c12_elems(ld) * batch * sizeof(float2)));
    w->fTAU = reinterpret_cast<float2*>(take((size_t)ld * batch * sizeof(float2)));
    w->fBand = reinterpret_cast<float2*>(take(band_scratch_elems(ld) * batch * sizeof(float2)));
  }
  return off;
}

size_t slk_eigh_workspace_bytes(int n, int ld, int batch, int vectors) {
  (void)n;
  EighWs w;
  return carve(nullptr, ld, batch, vectors, 0, &w);
}

size_t slk_eigh_mixed_workspace_bytes(int n, int ld, int batch) {
  (void)n;
  EighWs w;
  return carve(nullptr, ld, batch, 0, 1, &w);
}

// Tridiagonalisation of the eigenvalue path: 0 automatic (two-stage band
// reduction for values up to BAND_AUTO_MAX, blocked one-stage above and for
// vectors), 1 one-stage, 2 two-stage. The two-stage chase runs on one
// workgroup per matrix; beyond BAND_AUTO_MAX rows its single-CU work outweighs
// the one-stage path's launch chain.
#define BAND_AUTO_MAX 2304
static int g_tridiag_mode = 0;

static bool use_band(int n, int vectors) {
  if (vectors || n <= 32 || g_tridiag_mode == 1) return false;
  return g_tridiag_mode == 2 || n <= BAND_AUTO_MAX;
}

// info[0 .. n) = 0: the solver's failure flags start cleared (a fill kernel,
// so a captured graph replays it too).
__global__ void zero_info_kernel(int* info, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) info[i] = 0;
}

// Stage boundaries for optional timing: potrf | trinv | reduce (2 GEMMs) |
// hetrd | tridiagonal solve | back-transform. ev (nullable) holds 7 events.
// mixed: FP64 Cholesky and inverse, FP32 reduction and tridiagonalisation,
// FP64 bisection (values only).
static int eigh_impl(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                     void* ws, size_t ws_bytes, int* info, hipStream_t stream, hipEvent_t* ev, int mixed,
                     int flags = 0) {
  if (batch <= 0 || n <= 0) return 0;
  // SLK_EIGH_REAL: every H and S of the batch is real symmetric (zero
  // imaginary parts, e.g. k at a time-reversal-invariant point): the kernels
  // that know it skip the imaginary arithmetic and write exact zeros
  const bool re = (flags & SLK_EIGH_REAL) != 0;
  if (mixed && V != nullptr) return 3;
  const size_t need = mixed ? slk_eigh_mixed_workspace_bytes(n, ld, batch)
                            : slk_eigh_workspace_bytes(n, ld, batch, V != nullptr);
  if (ws_bytes < need) return 2;
  // any n up to EIGH_MAX_N (LDS-staged Sturm counts up to 4096, global
  // beyond; vectors by inverse iteration from global scratch)
  if (n > EIGH_MAX_N) return 3;
  hipLaunchKernelGGL(zero_info_kernel, dim3((2 * batch + 255) / 256), dim3(256), 0, stream, info, 2 * batch);
  if (V == nullptr && n <= 32) {
    // whole solve in LDS, one launch (stages collapse into the first interval)
    if (ev) (void)hipEventRecord(ev[0], stream);
    const int rc = slk_eigh_small_f64(H, S, n, ld, batch, w, wld, info, stream);
    if (ev)
      for (int i = 1; i < 7; ++i) (void)hipEventRecord(ev[i], stream);
    return rc;
  }
  EighWs s;
  carve(reinterpret_cast<char*>(ws), ld, batch, V != nullptr, mixed, &s);
  const long long sM = (long long)ld * ld;
  const double2* h = reinterpret_cast<const double2*>(H);
  const double2* sp = reinterpret_cast<const double2*>(S);
  if (ev) (void)hipEventRecord(ev[0], stream);
  // Cholesky and L^-1 in one recursion (the "potrf" stage; "trinv" is empty).
  // The temporary (ld * ld / 2 complex128 per matrix) borrows the product
  // buffer of the reduction, which is free until then.
  launch_chol_inv(stream, sp, s.B0, s.B1, mixed ? reinterpret_cast<double2*>(s.F2) : s.B2, n, ld, batch, info, re);
  if (ev) (void)hipEventRecord(ev[1], stream);
  if (ev) (void)hipEventRecord(ev[2], stream);
  // Both tridiagonalisations (band and one-stage) read only the lower
  // triangle of C = W H W^H. Its lower tiles need W H only on and below the
  // diagonal tiles (C_ij, j <= i, sums (W H)_ik conj(W_jk) over k <= j), so
  // both products skip the tiles above the diagonal (tri = 1).
  const bool band = use_band(n, V != nullptr);
  const int tri = 1;
  if (!mixed) {
    // B2 = W H (W lower), B0 = B2 W^H (W^H upper)
    launch_zgemm(stream, batch, n, n, n, 1.0, s.B1, ld, sM, OP_N, h, ld, sM, OP_N, 0.0, s.B2, ld, sM, KLIM_A_LOWER,
                 tri, re);
    // a real batch bound for the band reduction gets C as packed real values
    launch_zgemm(stream, batch, n, n, n, 1.0, s.B2, ld, sM, OP_N, s.B1, ld, sM, OP_C, 0.0, s.B0, ld, sM,
                 KLIM_B_UPPER, tri, re, band);
    if (ev) (void)hipEventRecord(ev[3], stream);
    if (band)
      // B2 (W H) is free after the reduction: it holds the split-K partials
      launch_band_tridiag(stream, s.B0, n, ld, batch, s.D, s.E, s.TAU, s.Vc, s.Wc, s.B2, s.Band, nullptr,
                          info + batch, re);
    else
      launch_hetrd_blocked(stream, s.B0, n, ld, batch, s.D, s.E, s.TAU, s.Vc, s.Wc, s.Y, s.Yp, s.C12);
  } else {
    launch_to_c64(stream, s.B1, s.F0, (size_t)batch * ld * ld);
    launch_to_c64(stream, h, s.F1, (size_t)batch * ld * ld);
    launch_cgemm(stream, batch, n, n, n, 1.0f, s.F0, ld, sM, OP_N, s.F1, ld, sM, OP_N, 0.0f, s.F2, ld, sM,
                 KLIM_A_LOWER, tri, re);
    launch_cgemm(stream, batch, n, n, n, 1.0f, s.F2, ld, sM, OP_N, s.F0, ld, sM, OP_C, 0.0f, s.F1, ld, sM,
                 KLIM_B_UPPER, tri, re, band);
    if (ev) (void)hipEventRecord(ev[3], stream);
    // F0 (W in FP32) is free after the reduction: it receives the reflector
    // rows; F2 (W32 H32) holds the split-K partials of the band reduction
    if (band)
      launch_band_tridiag_f32(stream, s.F1, n, ld, batch, s.D, s.E, s.fTAU, s.fVc, s.fWc, s.F2, s.fBand, s.F0,
                              info + batch, re);
    else
      launch_hetrd_blocked_f32(stream, s.F1, n, ld, batch, s.D, s.E, s.fTAU, s.fVc, s.fWc, s.fY, s.fYp, s.fC12,
                               s.F0);
  }
  if (ev) (void)hipEventRecord(ev[4], stream);
  {
    const int G = bisect_group(n, batch);
    const int eig_per_block = (EIG_THREADS / 64) * (64 / G);
    const dim3 grid((n + eig_per_block - 1) / eig_per_block, batch);
    const int M = 2 * n;
    const dim3 ggrid((M + 1 + EIG_THREADS - 1) / EIG_THREADS, batch);
    // the mixed path's tridiagonal carries FP32 rounding (~1e-7 of the
    // spectral radius): its eigenvalues are resolved to 2^-36 relative
    // instead of the FP64 path's 4 ulp
    const double rtol = mixed ? 1.4551915228366852e-11 : 4.0 * 2.2e-16;
    // one Newton step on the pivot reciprocals for the mixed path (its
    // tolerance is far above the 2^-48.7 they leave), two for FP64
    if (n <= BISECT_LDS_MAX) {
      const size_t lds = 2 * n * sizeof(double);
      if (mixed) {
        hipLaunchKernelGGL((sturm_grid_kernel<true, 1>), ggrid, dim3(EIG_THREADS), lds, stream, s.D, s.E, n, ld, M,
                           s.SG, (const double*)nullptr);
        hipLaunchKernelGGL((bisect_kernel<true, 1>), grid, dim3(EIG_THREADS), lds, stream, s.D, s.E, n, ld, w, wld, G,
                           (const double*)nullptr, (const int*)s.SG, M, rtol);
      } else {
        hipLaunchKernelGGL((sturm_grid_kernel<true, 2>), ggrid, dim3(EIG_THREADS), lds, stream, s.D, s.E, n, ld, M,
                           s.SG, (const double*)nullptr);
        hipLaunchKernelGGL((bisect_kernel<true, 2>), grid, dim3(EIG_THREADS), lds, stream, s.D, s.E, n, ld, w, wld, G,
                           (const double*)nullptr, (const int*)s.SG, M, rtol);
      }
    } else {
      // e^2 into the (now free) mat-vec vector buffer, ld doubles per matrix
      double* e2 = mixed ? reinterpret_cast<double*>(s.fY) : reinterpret_cast<double*>(s.Y);
      hipLaunchKernelGGL(square_e_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, stream, s.E, n, ld, e2);
      if (mixed) {
        hipLaunchKernelGGL((sturm_grid_kernel<false, 1>), ggrid, dim3(EIG_THREADS), 0, stream, s.D, s.E, n, ld, M,
                           s.SG, (const double*)e2);
        hipLaunchKernelGGL((bisect_kernel<false, 1>), grid, dim3(EIG_THREADS), 0, stream, s.D, s.E, n, ld, w, wld, G,
                           (const double*)e2, (const int*)s.SG, M, rtol);
      } else {
        hipLaunchKernelGGL((sturm_grid_kernel<false, 2>), ggrid, dim3(EIG_THREADS), 0, stream, s.D, s.E, n, ld, M,
                           s.SG, (const double*)e2);
        hipLaunchKernelGGL((bisect_kernel<false, 2>), grid, dim3(EIG_THREADS), 0, stream, s.D, s.E, n, ld, w, wld, G,
                           (const double*)e2, (const int*)s.SG, M, rtol);
      }
    }
  }
  if (V != nullptr) {
    // eigenvectors of T by inverse iteration from the eigenvalues just found
    hipLaunchKernelGGL(stein_kernel, dim3((n + STEIN_THREADS - 1) / STEIN_THREADS, batch), dim3(STEIN_THREADS), 0,
                       stream, s.D, s.E, w, wld, n, ld, s.Z, s.SV, n);
  }
  if (ev) (void)hipEventRecord(ev[5], stream);
  if (V != nullptr) {
    // the inverse-iteration scratch (STEIN_SCR ld^2 doubles per matrix) is
    // free again: it holds the back-transform's blocks
    launch_apply_q(stream, s.B0, s.TAU, s.Z, s.B2, n, ld, batch, reinterpret_cast<double2*>(s.SV), re);
    // V = W^H Y (W^H upper)
    launch_zgemm(stream, batch, n, n, n, 1.0, s.B1, ld, sM, OP_C, s.B2, ld, sM, OP_N, 0.0,
                 reinterpret_cast<double2*>(V), ld, sM, KLIM_A_UPPER, 0, re);
  }
  if (ev) (void)hipEventRecord(ev[6], stream);
  return (int)hipGetLastError();
}

static int eigh_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V, void* ws,
                      size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms, int mixed, int flags = 0) {
  hipEvent_t ev[7];
  for (int i = 0; i < 7; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) return 4;
  int rc = eigh_impl(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, ev, mixed, flags);
  if (rc == 0) {
    rc = (int)hipEventSynchronize(ev[6]);
    for (int i = 0; i < 6 && rc == 0; ++i) rc = (int)hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]);
  }
  for (int i = 0; i < 7; ++i) (void)hipEventDestroy(ev[i]);
  return rc;
}

int slk_set_tridiag_mode(int mode) {
  if (mode < 0 || mode > 2) return 1;
  g_tridiag_mode = mode;
  return 0;
}

int slk_eigh_f64(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                 void* ws, size_t ws_bytes, int* info, hipStream_t stream) {
  return eigh_impl(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, nullptr, 0);
}

int slk_eigh_f64_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                       void* ws, size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms) {
  return eigh_timed(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, stage_ms, 0);
}

int slk_eigh_mixed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* ws,
                   size_t ws_bytes, int* info, hipStream_t stream) {
  return eigh_impl(H, S, n, ld, batch, w, wld, nullptr, ws, ws_bytes, info, stream, nullptr, 1);
}

size_t slk_refine_scratch_bytes(int ld, int n_spec) {
  const size_t nblk = (size_t)(ld + RF_BLK - 1) / RF_BLK;
  return ((size_t)n_spec * 2 * RF_SCR * ld + (size_t)n_spec * 2 * nblk * 2 + (size_t)n_spec * 2) * sizeof(double);
}

int slk_refine_edges_mixed(const void* H, const void* S, int n, int ld, int batch, const void* ws, size_t ws_bytes,
                           double* w, int wld, const SlkSpectrum* spec, int n_spec, double* edges, double* scratch,
                           size_t scratch_bytes, hipStream_t stream) {
  (void)wld;
  if (n_spec <= 0 || n <= 32) return 0;      // n <= 32 was solved in FP64 by the LDS kernel
  if (n > EIGH_MAX_N) return 3;
  if (ws_bytes < slk_eigh_mixed_workspace_bytes(n, ld, batch)) return 2;
  if (scratch_bytes < slk_refine_scratch_bytes(ld, n_spec)) return 2;
  EighWs s;
  carve(reinterpret_cast<char*>(const_cast<void*>(ws)), ld, batch, 0, 1, &s);
  RefineCtx c;
  c.H = reinterpret_cast<const double2*>(H);
  c.S = reinterpret_cast<const double2*>(S);
  c.Wm = s.B1;
  c.R = s.F0;
  c.TAU = s.fTAU;
  const bool band = use_band(n, 0);
  c.V2 = band ? s.F1 : nullptr;
  c.off = band ? SLK_BW : 1;
  c.n = n;
  c.D = s.D;
  c.E = s.E;
  c.ld = ld;
  c.w = w;
  c.spec = spec;
  c.edges = edges;
  c.scratch = scratch;
  c.nblk = (ld + RF_BLK - 1) / RF_BLK;
  c.part = scratch + (size_t)n_spec * 2 * RF_SCR * ld;
  c.meta = reinterpret_cast<int*>(c.part + (size_t)n_spec * 2 * c.nblk * 2);
  // x in LDS when every structure fits (m <= ld <= RF_LDS_X), else in the scratch slot
  c.split = (!band && ld > RQ_MIN_LD) ? 1 : 0;
  c.lds_x = (ld <= RF_LDS_X && !band && !c.split) ? 1 : 0;
  const size_t lds = c.lds_x ? ld * sizeof(double2) : 0;
  const dim3 g(n_spec, 2);
  if (band) {
    // tridiagonal solve only, then the two-stage back-transform
    if (ld > RB_LDS_MAX) return 3;
    const size_t lds2 = ld * (sizeof(double2) + sizeof(float2));
    const float2* TT = s.fBand + band_tt_offset(ld, batch);
    const int npan = band_panels(ld);
    hipLaunchKernelGGL(refine_tri_kernel<0>, g, dim3(64), 0, stream, c);
    // reflector-row elements per thread: ld <= RB_LDS_MAX < 3 RB_THREADS
    static_assert(RB_LDS_MAX <= 3 * RB_THREADS, "refine_band_kernel tiers");
    if (ld <= RB_THREADS)
      hipLaunchKernelGGL(refine_band_kernel<1>, g, dim3(RB_THREADS), lds2, stream, c, TT, npan);
    else if (ld <= RB_THREADS * 2)
      hipLaunchKernelGGL(refine_band_kernel<2>, g, dim3(RB_THREADS), lds2, stream, c, TT, npan);
    else
      hipLaunchKernelGGL(refine_band_kernel<3>, g, dim3(RB_THREADS), lds2, stream, c, TT, npan);
  } else if (c.split) {
    // y on one wave, then x = Q y on 1024 threads (x in registers: m <= n <= NQ * 1024)
    static_assert(EIGH_MAX_N <= 16 * RQ_THREADS, "refine_q1_kernel tiers");
    hipLaunchKernelGGL(refine_tri_kernel<0>, g, dim3(64), 0, stream, c);
    if (n <= 4 * RQ_THREADS)
      hipLaunchKernelGGL((refine_q1_kernel<4, true>), g, dim3(RQ_THREADS), 0, stream, c);
    else if (n <= 8 * RQ_THREADS)
      hipLaunchKernelGGL((refine_q1_kernel<8, true>), g, dim3(RQ_THREADS), 0, stream, c);
    else if (n <= 9 * RQ_THREADS)
      hipLaunchKernelGGL((refine_q1_kernel<9, false>), g, dim3(RQ_THREADS), 0, stream, c);
    else if (n <= 12 * RQ_THREADS)
      hipLaunchKernelGGL((refine_q1_kernel<12, false>), g, dim3(RQ_THREADS), 0, stream, c);
    else
      hipLaunchKernelGGL((refine_q1_kernel<16, false>), g, dim3(RQ_THREADS), 0, stream, c);
  } else if (ld <= 64 * 4)
    hipLaunchKernelGGL(refine_tri_kernel<4>, g, dim3(64), lds, stream, c);
  else if (ld <= 64 * 8)
    hipLaunchKernelGGL(refine_tri_kernel<8>, g, dim3(64), lds, stream, c);
  else if (ld <= 64 * 16)
    hipLaunchKernelGGL(refine_tri_kernel<16>, g, dim3(64), lds, stream, c);
  else if (ld <= 64 * 40)
    hipLaunchKernelGGL(refine_tri_kernel<40>, g, dim3(64), lds, stream, c);
  else
    hipLaunchKernelGGL(refine_tri_kernel<0>, g, dim3(64), lds, stream, c);
  hipLaunchKernelGGL(refine_wh_kernel, dim3(c.nblk, n_spec * 2), dim3(RF_THREADS), 0, stream, c);
  hipLaunchKernelGGL(refine_rq_kernel, dim3(c.nblk, n_spec * 2), dim3(RF_THREADS), 0, stream, c);
  hipLaunchKernelGGL(refine_commit_kernel, dim3((n_spec + 63) / 64), dim3(64), 0, stream, c, w, edges, n_spec);
  return (int)hipGetLastError();
}

int slk_eigh_mixed_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* ws,
                         size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms) {
  return eigh_timed(H, S, n, ld, batch, w, wld, nullptr, ws, ws_bytes, info, stream, stage_ms, 1);
}

int slk_eigh_ex(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V, void* ws,
                size_t ws_bytes, int* info, int flags, float* stage_ms, hipStream_t stream) {
  if (flags & ~(SLK_EIGH_MIXED | SLK_EIGH_REAL)) return 1;
  const int mixed = (flags & SLK_EIGH_MIXED) ? 1 : 0;
  if (stage_ms != nullptr)
    return eigh_timed(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, stage_ms, mixed, flags);
  return eigh_impl(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, nullptr, mixed, flags);
}

}  // extern "C"
