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
#include "slk_dense.h"

#define EIG_THREADS 256

// ---------------------------------------------------------------------------
// Eigenvalues of the real symmetric tridiagonal (d, e) by bisection /
// multisection on Sturm counts, ascending order.
// ---------------------------------------------------------------------------
// 1/q from the hardware approximation refined by two Newton steps (full
// double precision for the pivots; |q| >= pivmin keeps it finite).
__device__ __forceinline__ double rcp_refined(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  r = fma(fma(-q, r, 1.0), r, r);
  return r;
}

// Number of eigenvalues below x; e2[i] = e[i]^2.
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] * rcp_refined(q);
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Groups of G lanes (G a power of two <= 64) share one eigenvalue; each step
// evaluates G Sturm counts at equidistant interior points and keeps the
// sub-interval where the count crosses the eigenvalue index ((G+1)x shrink per
// step). G = 1 is plain bisection with 64 eigenvalues per wave; larger G
// trades total work for parallelism when few eigenvalues are in flight.
__global__ __launch_bounds__(EIG_THREADS) void bisect_kernel(const double* D, const double* Ed, int n, int ld,
                                                             double* Wout, int wld, int G) {
  const int b = blockIdx.y;
  const double* d = D + (size_t)b * ld;
  const double* e = Ed + (size_t)b * ld;
  extern __shared__ double sm[];
  double* sd = sm;
  double* se = sm + n;
  __shared__ double s_red[3][EIG_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double lo = 1e300, hi = -1e300, emax = 0.0;
  for (int i = tid; i < n; i += blockDim.x) {
    const double di = d[i];
    const double ei = i < n - 1 ? e[i] : 0.0;
    const double ep = i > 0 ? e[i - 1] : 0.0;
    sd[i] = di;
    se[i] = ei * ei;
    const double r = fabs(ep) + fabs(ei);
    lo = fmin(lo, di - r);
    hi = fmax(hi, di + r);
    emax = fmax(emax, ei * ei);
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  emax = wave_max(emax);
  if (lane == 0) { s_red[0][wv] = lo; s_red[1][wv] = hi; s_red[2][wv] = emax; }
  __syncthreads();
  lo = 1e300; hi = -1e300; emax = 0.0;
  for (int q = 0; q < EIG_THREADS / 64; ++q) {
    lo = fmin(lo, s_red[0][q]);
    hi = fmax(hi, s_red[1][q]);
    emax = fmax(emax, s_red[2][q]);
  }
  const double span = fmax(fabs(lo), fabs(hi));
  lo = lo - 2.0 * 2.2e-16 * span * n - 1e-300;
  hi = hi + 2.0 * 2.2e-16 * span * n + 1e-300;
  const double piv = fmax(2.2e-308, 2.2e-308 * emax) * 1e2;
  const int per_wave = 64 / G;
  const int grp = lane / G, sub = lane % G;
  const int idx = (blockIdx.x * (EIG_THREADS / 64) + wv) * per_wave + grp;
  if ((blockIdx.x * (EIG_THREADS / 64) + wv) * per_wave >= n) return;   // whole wave idle
  const unsigned long long gmask = G == 64 ? ~0ULL : ((1ULL << G) - 1ULL);
  const double inv = 1.0 / (double)(G + 1);
  bool done = idx >= n;
  for (int it = 0; it < 256; ++it) {
    const double width = hi - lo;
    if (!done && width <= 4.0 * 2.2e-16 * fmax(fabs(lo), fabs(hi)) + piv) done = true;
    if (__ballot(!done) == 0ULL) break;
    const double x = lo + width * (double)(sub + 1) * inv;
    const bool above = !done && sturm_count(sd, se, n, x, piv) > idx;
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

// lanes per eigenvalue so that about 2^17 lanes are busy
static int bisect_group(int n, int batch) {
  const double total = (double)n * batch;
  int g = 1;
  while (g < 64 && total * g * 2 <= 131072.0) g *= 2;
  return g;
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

// Workspace (per matrix). FP64 path: L/C, W, W*H (and Y for vectors), the
// trinv temporary T, panel V and W, Y, the mat-vec partial sums, C12, TAU, D,
// E, and the real Z for the vectors path. Mixed path: L and W (FP64), T, then
// FP32 copies W32, H32 (reused for C), W32*H32 and FP32 panels.
struct EighWs {
  double2 *B0, *B1, *B2, *T, *Vc, *Wc, *Y, *Yp, *C12, *TAU;
  float2 *F0, *F1, *F2, *fVc, *fWc, *fY, *fYp, *fC12, *fTAU;
  double *D, *E, *Z;
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
  w->T = reinterpret_cast<double2*>(take((size_t)SLK_NB * ld * batch * sizeof(double2)));
  w->D = reinterpret_cast<double*>(take((size_t)ld * batch * sizeof(double)));
  w->E = reinterpret_cast<double*>(take((size_t)ld * batch * sizeof(double)));
  if (!mixed) {
    w->B2 = reinterpret_cast<double2*>(take(matel * batch * sizeof(double2)));
    w->Vc = reinterpret_cast<double2*>(take((size_t)SLK_NB * ld * batch * sizeof(double2)));
    w->Wc = reinterpret_cast<double2*>(take((size_t)SLK_NB * ld * batch * sizeof(double2)));
    w->Y = reinterpret_cast<double2*>(take((size_t)ld * batch * sizeof(double2)));
    w->Yp = reinterpret_cast<double2*>(take(hetrd_partials_elems(ld) * batch * sizeof(double2)));
    w->C12 = reinterpret_cast<double2*>(take((size_t)2 * SLK_NB * batch * sizeof(double2)));
    w->TAU = reinterpret_cast<double2*>(take((size_t)ld * batch * sizeof(double2)));
    w->Z = vectors ? reinterpret_cast<double*>(take(matel * batch * sizeof(double))) : nullptr;
  } else {
    w->F0 = reinterpret_cast<float2*>(take(matel * batch * sizeof(float2)));
    w->F1 = reinterpret_cast<float2*>(take(matel * batch * sizeof(float2)));
    w->F2 = reinterpret_cast<float2*>(take(matel * batch * sizeof(float2)));
    w->fVc = reinterpret_cast<float2*>(take((size_t)SLK_NB * ld * batch * sizeof(float2)));
    w->fWc = reinterpret_cast<float2*>(take((size_t)SLK_NB * ld * batch * sizeof(float2)));
    w->fY = reinterpret_cast<float2*>(take((size_t)ld * batch * sizeof(float2)));
    w->fYp = reinterpret_cast<float2*>(take(hetrd_partials_elems(ld) * batch * sizeof(float2)));
    w->fC12 = reinterpret_cast<float2*>(take((size_t)2 * SLK_NB * batch * sizeof(float2)));
    w->fTAU = reinterpret_cast<float2*>(take((size_t)ld * batch * sizeof(float2)));
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

// Stage boundaries for optional timing: potrf | trinv | reduce (2 GEMMs) |
// hetrd | tridiagonal solve | back-transform. ev (nullable) holds 7 events.
// mixed: FP64 Cholesky and inverse, FP32 reduction and tridiagonalisation,
// FP64 bisection (values only).
static int eigh_impl(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                     void* ws, size_t ws_bytes, int* info, hipStream_t stream, hipEvent_t* ev, int mixed) {
  if (batch <= 0 || n <= 0) return 0;
  if (mixed && V != nullptr) return 3;
  const size_t need = mixed ? slk_eigh_mixed_workspace_bytes(n, ld, batch)
                            : slk_eigh_workspace_bytes(n, ld, batch, V != nullptr);
  if (ws_bytes < need) return 2;
  // LDS-resident tridiagonal data: values need 2n doubles, vectors 4n doubles (<= 64 KiB)
  if (n > 4096 || (V != nullptr && n > 2048)) return 3;
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
  launch_potrf_blocked(stream, sp, s.B0, s.B1, n, ld, batch, info);
  if (ev) (void)hipEventRecord(ev[1], stream);
  launch_trinv_blocked(stream, s.B0, s.B1, s.T, n, ld, batch);
  if (ev) (void)hipEventRecord(ev[2], stream);
  if (!mixed) {
    // B2 = W H (W lower), B0 = B2 W^H (W^H upper)
    launch_zgemm(stream, batch, n, n, n, 1.0, s.B1, ld, sM, OP_N, h, ld, sM, OP_N, 0.0, s.B2, ld, sM, KLIM_A_LOWER,
                 0);
    launch_zgemm(stream, batch, n, n, n, 1.0, s.B2, ld, sM, OP_N, s.B1, ld, sM, OP_C, 0.0, s.B0, ld, sM,
                 KLIM_B_UPPER, 0);
    if (ev) (void)hipEventRecord(ev[3], stream);
    launch_hetrd_blocked(stream, s.B0, n, ld, batch, s.D, s.E, s.TAU, s.Vc, s.Wc, s.Y, s.Yp, s.C12);
  } else {
    launch_to_c64(stream, s.B1, s.F0, (size_t)batch * ld * ld);
    launch_to_c64(stream, h, s.F1, (size_t)batch * ld * ld);
    launch_cgemm(stream, batch, n, n, n, 1.0f, s.F0, ld, sM, OP_N, s.F1, ld, sM, OP_N, 0.0f, s.F2, ld, sM,
                 KLIM_A_LOWER, 0);
    launch_cgemm(stream, batch, n, n, n, 1.0f, s.F2, ld, sM, OP_N, s.F0, ld, sM, OP_C, 0.0f, s.F1, ld, sM,
                 KLIM_B_UPPER, 0);
    if (ev) (void)hipEventRecord(ev[3], stream);
    launch_hetrd_blocked_f32(stream, s.F1, n, ld, batch, s.D, s.E, s.fTAU, s.fVc, s.fWc, s.fY, s.fYp, s.fC12);
  }
  if (ev) (void)hipEventRecord(ev[4], stream);
  if (V == nullptr) {
    const int G = bisect_group(n, batch);
    const int eig_per_block = (EIG_THREADS / 64) * (64 / G);
    hipLaunchKernelGGL(bisect_kernel, dim3((n + eig_per_block - 1) / eig_per_block, batch), dim3(EIG_THREADS),
                       2 * n * sizeof(double), stream, s.D, s.E, n, ld, w, wld, G);
    if (ev) (void)hipEventRecord(ev[5], stream);
  } else {
    hipLaunchKernelGGL(tql2_kernel, dim3(batch), dim3(EIG_THREADS), 4 * n * sizeof(double), stream, s.D, s.E, n, ld,
                       s.Z, w, wld, info + batch);
    if (ev) (void)hipEventRecord(ev[5], stream);
    hipLaunchKernelGGL(apply_q_kernel, dim3((n + 63) / 64, batch), dim3(EIG_THREADS), 0, stream, s.B0, s.TAU, s.Z,
                       s.B2, n, ld);
    // V = W^H Y (W^H upper)
    launch_zgemm(stream, batch, n, n, n, 1.0, s.B1, ld, sM, OP_C, s.B2, ld, sM, OP_N, 0.0,
                 reinterpret_cast<double2*>(V), ld, sM, KLIM_A_UPPER, 0);
  }
  if (ev) (void)hipEventRecord(ev[6], stream);
  return (int)hipGetLastError();
}

static int eigh_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V, void* ws,
                      size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms, int mixed) {
  hipEvent_t ev[7];
  for (int i = 0; i < 7; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) return 4;
  int rc = eigh_impl(H, S, n, ld, batch, w, wld, V, ws, ws_bytes, info, stream, ev, mixed);
  if (rc == 0) {
    rc = (int)hipEventSynchronize(ev[6]);
    for (int i = 0; i < 6 && rc == 0; ++i) rc = (int)hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]);
  }
  for (int i = 0; i < 7; ++i) (void)hipEventDestroy(ev[i]);
  return rc;
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

int slk_eigh_mixed_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* ws,
                         size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms) {
  return eigh_timed(H, S, n, ld, batch, w, wld, nullptr, ws, ws_bytes, info, stream, stage_ms, 1);
}

}  // extern "C"
