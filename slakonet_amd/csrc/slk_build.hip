// Fused H(k)/S(k) build for the SlaKoNet hot path on MI355X (gfx950).
//
// One workgroup owns a tile: the orbital rows of one atom i, the orbital
// columns of a run of atoms [j0, j1) (<= 64 columns), and a chunk of <= KC
// k-points of one structure. Inside the workgroup:
//   1. screen every (j, T) image of the tile's column atoms for
//      0 < |R_j + T - R_i| < cutoff (deterministic compaction in LDS);
//   2. interpolate the 20 H/S integrals of both ordered species tables with
//      8-point Lagrange weights on the uniform grid (stencil/clamp/tail rules of
//      slakonet/interpolation.py:107-238);
//   3. form the 9x9 Slater-Koster blocks with the closed-form two-centre table
//      (s, p=(y,z,x), d=(xy,yz,z2,xz,x2-y2); slakonet/slaterkoster.py:1013-1372),
//      choosing per element the direct block or the conjugate-transposed block
//      of the reversed pair exactly as hs_matrix assigns them
//      (slakonet/slaterkoster.py:207-405);
//   4. Bloch-sum with exp(2 pi i k.T) into register accumulators;
//   5. add the on-site diagonal and write the tile rows with coalesced stores.
// Every output element is written exactly once; zero blocks included.
#include "slk_common.h"
#include "slakonet_amd.h"
#include <hip/hip_ext.h>

#include "slk_sampler.h"

#define BUILD_THREADS 256
#define KC 4            // k-points per workgroup
#define MAXPI 24        // pair-images per LDS batch
#define SEG 2048        // candidates screened per segment
#define MAXCOL 64
#define ROWS_MAXSPEC 128     // row panels: species of the table set (slot map in LDS)
#define ROWS_SEG 1024        // row panels: candidates screened per segment (<= 32 per thread)
#define ROWS_MIN_LD 448      // row panels for padded dimensions from here up

__constant__ int c_slot_l[9] = {0, 1, 1, 1, 2, 2, 2, 2, 2};
__constant__ int c_slot_m[9] = {0, 0, 1, 2, 0, 1, 2, 3, 4};
// first canonical integral of the (l1<=l2) pair
__device__ __forceinline__ int lpair_base(int l1, int l2) {
  // (0,0)0 (0,1)1 (0,2)2 (1,1)3 (1,2)5 (2,2)7
  return l1 == 0 ? l2 : (l1 == 1 ? (l2 == 1 ? 3 : 5) : 7);
}

#define SQ3 1.7320508075688772

// Forward-mode dual number with N partial derivatives: exact derivatives of
// the polynomial Slater-Koster table and of the interpolation weights.
template <int N>
struct Dual {
  double v;
  double g[N];
  __device__ Dual() : v(0.0) {
#pragma unroll
    for (int i = 0; i < N; ++i) g[i] = 0.0;
  }
  __device__ Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) g[i] = 0.0;
  }
};
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = -a.g[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double s, const Dual<N>& a) {
  Dual<N> r(s * a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = s * a.g[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double s) { return s * a; }
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double s) { Dual<N> r = a; r.v += s; return r; }
template <int N>
__device__ __forceinline__ Dual<N> operator+(double s, const Dual<N>& a) { return a + s; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double s) { Dual<N> r = a; r.v -= s; return r; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(double s, const Dual<N>& a) { return (-a) + s; }
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double s) { return a * (1.0 / s); }
__device__ __forceinline__ double dval(double x) { return x; }
template <int N>
__device__ __forceinline__ double dval(const Dual<N>& x) { return x.v; }

// Two-centre Slater-Koster element for l1 <= l2: orbital a of shell l1 on the
// row atom, orbital b of shell l2 on the column atom, direction cosines
// (x, y, z) from row to column atom, V = {sigma, pi, delta}.
template <typename R>
__device__ R sk_elem_t(int l1, int l2, int a, int b, R x, R y, R z, const R* V) {
  if (l1 == 0 && l2 == 0) return V[0];
  if (l1 == 0 && l2 == 1) {
    const R c[3] = {y, z, x};
    return c[b] * V[0];
  }
  if (l1 == 0 && l2 == 2) {
    R g;
    switch (b) {
      case 0: g = SQ3 * x * y; break;
      case 1: g = SQ3 * y * z; break;
      case 2: g = z * z - 0.5 * (x * x + y * y); break;
      case 3: g = SQ3 * x * z; break;
      default: g = 0.5 * SQ3 * (x * x - y * y); break;
    }
    return g * V[0];
  }
  if (l1 == 1 && l2 == 1) {
    const R c[3] = {y, z, x};
    return c[a] * c[b] * (V[0] - V[1]) + (a == b ? V[1] : R(0.0));
  }
  const R l = x, m = y, n = z;
  const R l2s = l * l, m2s = m * m, n2s = n * n;
  const R s = V[0], p = V[1];
  if (l1 == 1 && l2 == 2) {
    // a: 0=y 1=z 2=x ; b: 0=xy 1=yz 2=z2 3=xz 4=x2-y2
    const int pa = (a == 2) ? 0 : (a == 0 ? 1 : 2);  // 0=x 1=y 2=z
    switch (pa * 5 + b) {
      case 0: return SQ3 * l2s * m * s + m * (1.0 - 2.0 * l2s) * p;            // x,xy
      case 1: return SQ3 * l * m * n * s - 2.0 * l * m * n * p;                  // x,yz
      case 2: return l * (n2s - 0.5 * (l2s + m2s)) * s - SQ3 * l * n2s * p;      // x,z2
      case 3: return SQ3 * l2s * n * s + n * (1.0 - 2.0 * l2s) * p;            // x,xz
      case 4: return 0.5 * SQ3 * l * (l2s - m2s) * s + l * (1.0 - l2s + m2s) * p;  // x,x2-y2
      case 5: return SQ3 * m2s * l * s + l * (1.0 - 2.0 * m2s) * p;            // y,xy
      case 6: return SQ3 * m2s * n * s + n * (1.0 - 2.0 * m2s) * p;            // y,yz
      case 7: return m * (n2s - 0.5 * (l2s + m2s)) * s - SQ3 * m * n2s * p;      // y,z2
      case 8: return SQ3 * l * m * n * s - 2.0 * l * m * n * p;                  // y,xz
      case 9: return 0.5 * SQ3 * m * (l2s - m2s) * s - m * (1.0 + l2s - m2s) * p;  // y,x2-y2
      case 10: return SQ3 * l * m * n * s - 2.0 * l * m * n * p;                 // z,xy
      case 11: return SQ3 * n2s * m * s + m * (1.0 - 2.0 * n2s) * p;           // z,yz
      case 12: return n * (n2s - 0.5 * (l2s + m2s)) * s + SQ3 * n * (l2s + m2s) * p;  // z,z2
      case 13: return SQ3 * n2s * l * s + l * (1.0 - 2.0 * n2s) * p;           // z,xz
      default: return 0.5 * SQ3 * n * (l2s - m2s) * s - n * (l2s - m2s) * p;   // z,x2-y2
    }
  }
  // d-d, symmetric in (a, b)
  const R d = V[2];
  int ia = a < b ? a : b, ib = a < b ? b : a;
  const R lm = l * m, mn = m * n, nl = n * l;
  const R q = l2s - m2s;                 // l^2 - m^2
  const R r = n2s - 0.5 * (l2s + m2s);   // n^2 - (l^2+m^2)/2
  switch (ia * 5 + ib) {
    case 0: return 3.0 * l2s * m2s * s + (l2s + m2s - 4.0 * l2s * m2s) * p + (n2s + l2s * m2s) * d;
    case 1: return 3.0 * l * m2s * n * s + nl * (1.0 - 4.0 * m2s) * p + nl * (m2s - 1.0) * d;
    case 2: return SQ3 * lm * r * s - 2.0 * SQ3 * lm * n2s * p + 0.5 * SQ3 * lm * (1.0 + n2s) * d;
    case 3: return 3.0 * l2s * m * n * s + mn * (1.0 - 4.0 * l2s) * p + mn * (l2s - 1.0) * d;
    case 4: return 1.5 * lm * q * s - 2.0 * lm * q * p + 0.5 * lm * q * d;
    case 6: return 3.0 * m2s * n2s * s + (m2s + n2s - 4.0 * m2s * n2s) * p + (l2s + m2s * n2s) * d;
    case 7: return SQ3 * mn * r * s + SQ3 * mn * (l2s + m2s - n2s) * p - 0.5 * SQ3 * mn * (l2s + m2s) * d;
    case 8: return 3.0 * m * n2s * l * s + lm * (1.0 - 4.0 * n2s) * p + lm * (n2s - 1.0) * d;
    case 9: return 1.5 * mn * q * s - mn * (1.0 + 2.0 * q) * p + mn * (1.0 + 0.5 * q) * d;
    case 12: return r * r * s + 3.0 * n2s * (l2s + m2s) * p + 0.75 * (l2s + m2s) * (l2s + m2s) * d;
    case 13: return SQ3 * nl * r * s + SQ3 * nl * (l2s + m2s - n2s) * p - 0.5 * SQ3 * nl * (l2s + m2s) * d;
    case 14: return 0.5 * SQ3 * q * r * s - SQ3 * n2s * q * p + 0.25 * SQ3 * (1.0 + n2s) * q * d;
    case 18: return 3.0 * n2s * l2s * s + (n2s + l2s - 4.0 * n2s * l2s) * p + (m2s + n2s * l2s) * d;
    case 19: return 1.5 * nl * q * s + nl * (1.0 - 2.0 * q) * p - nl * (1.0 - 0.5 * q) * d;
    default: return 0.75 * q * q * s + (l2s + m2s - q * q) * p + (n2s + 0.25 * q * q) * d;  // 24
  }
}

__device__ __forceinline__ double sk_elem(int l1, int l2, int a, int b, double x, double y, double z,
                                          const double* V) {
  return sk_elem_t<double>(l1, l2, a, b, x, y, z, V);
}

// Interpolation weights: value = sum_m w[m] * table[base + m]; base = -1 -> 0.
// Follows slakonet/interpolation.py:107-238 (8-point stencil, clamp, tail).
template <typename R>
__device__ void interp_weights_t(R r, int ngrid, double step, double tail, int* base, R* w) {
  const int NI = 8;
  long ind = (long)floor(dval(r) / step);
  if (ind < ngrid) {
    long il = ind + 5;
    if (il > ngrid) il = ngrid;
    if (il < NI + 1) il = NI + 1;
    *base = (int)(il - NI - 1);
    const R t = r / step - (double)(il - NI);   // position in stencil units
    R left[NI], right[NI];
    left[0] = R(1.0);
    for (int m = 1; m < NI; ++m) left[m] = left[m - 1] * (t - (double)(m - 1));
    right[NI - 1] = R(1.0);
    for (int m = NI - 2; m >= 0; --m) right[m] = right[m + 1] * (t - (double)(m + 1));
    // denominators prod_{j != m} (m - j) = (-1)^(7-m) m! (7-m)!
    const double fact[8] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0};
    for (int m = 0; m < NI; ++m) {
      double den = fact[m] * fact[NI - 1 - m] * (((NI - 1 - m) & 1) ? -1.0 : 1.0);
      w[m] = left[m] * right[m] / den;
    }
    return;
  }
  const long max_ind = ngrid - 1 + (long)(tail / step);
  if (ind > max_ind) { *base = -1; return; }
  // fifth-order tail to zero (DFTB+ poly5ToZero) from y, y', y'' at (ngrid-1)*step
  *base = ngrid - NI - 1;
  const double delta = 1e-5;
  const double rmax = (ngrid - 1) * step + tail;
  const double dx = -tail;
  const R xr = (r - rmax) / dx;
  const R x3 = xr * xr * xr;
  const R A = ((6.0 * xr - 15.0) * xr + 10.0) * x3;
  const R B = ((-3.0 * xr + 7.0) * xr - 4.0) * x3;
  const R C = ((0.5 * xr - 1.0) * xr + 0.5) * x3;
  double w0[NI], w2[NI];
  for (int pass = 0; pass < 2; ++pass) {
    const double t = (NI - 1) + (pass == 0 ? -delta : delta) / step;
    double* ww = pass == 0 ? w0 : w2;
    double left[NI], right[NI];
    left[0] = 1.0;
    for (int m = 1; m < NI; ++m) left[m] = left[m - 1] * (t - (m - 1));
    right[NI - 1] = 1.0;
    for (int m = NI - 2; m >= 0; --m) right[m] = right[m + 1] * (t - (m + 1));
    const double fact[8] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0};
    for (int m = 0; m < NI; ++m) {
      double den = fact[m] * fact[NI - 1 - m] * (((NI - 1 - m) & 1) ? -1.0 : 1.0);
      ww[m] = left[m] * right[m] / den;
    }
  }
  for (int m = 0; m < NI; ++m) {
    const double e7 = (m == NI - 1) ? 1.0 : 0.0;
    const double yp = (w2[m] - w0[m]) / (2.0 * delta);
    const double ypp = (w2[m] + w0[m] - 2.0 * e7) / (delta * delta);
    w[m] = A * e7 + B * (dx * yp) + C * (dx * dx * ypp);
  }
}

__device__ __forceinline__ void interp_weights(double r, int ngrid, double step, double tail, int* base, double* w) {
  interp_weights_t<double>(r, ngrid, step, tail, base, w);
}

template <typename OUT>
__device__ __forceinline__ void store_c(OUT* p, double2 v);
template <>
__device__ __forceinline__ void store_c<double2>(double2* p, double2 v) { *p = v; }
template <>
__device__ __forceinline__ void store_c<float2>(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

// One tile (work item wk) of the build, executed by the whole workgroup.
template <typename OUT>
__device__ __forceinline__ void build_tile(const SlkBuildArgs& A, const SlkBuildWork wk) {
  __shared__ int s_list[SEG];            // compacted candidates (packed jrel<<20 | t)
  __shared__ int s_total;
  __shared__ int pi_j[MAXPI];
  __shared__ int pi_t[MAXPI];
  __shared__ double pi_u[MAXPI][3];
  __shared__ int pi_base[MAXPI][2];
  __shared__ double pi_w[MAXPI][2][8];
  __shared__ double pi_int[MAXPI][2][20];
  __shared__ double E[MAXPI][2][81];
  __shared__ double2 ph[MAXPI][KC];
  __shared__ int col_atom[MAXCOL];
  __shared__ int col_slot[MAXCOL];
  __shared__ int col_orb[MAXCOL];
  __shared__ int pi_mask[MAXPI];         // slots present on the column atom
  __shared__ int s_jspec[MAXCOL], s_jnorb[MAXCOL], s_jorb[MAXCOL];
  __shared__ int s_wsum[BUILD_THREADS / 64];
  __shared__ int s_rng[MAXCOL][2];       // per column atom: its pair-images in the batch

  const SlkStructDesc sd = A.structs[wk.s];
  const int tid = threadIdx.x;
  const int ai = sd.atom0 + wk.i;
  const int si = A.atom_spec[ai];
  const int norb_i = A.tab.spec_norb[si];
  const double rix = A.atom_pos[3 * ai], riy = A.atom_pos[3 * ai + 1], riz = A.atom_pos[3 * ai + 2];
  const int nja = wk.j1 - wk.j0;
  // slots (s, p, d components) present on atom i
  int mask_i = 0;
  for (int o = 0; o < norb_i; ++o) mask_i |= 1 << A.tab.spec_slot[si * 9 + o];

  // column map: tile column -> (atom, canonical slot, orbital index). The
  // column atoms' species and offsets are loaded in parallel, then each
  // column finds its atom by a scan of the orbital counts (columns are the
  // consecutive orbitals of atoms j0 .. j1-1, cut at MAXCOL).
  if (tid < nja) {
    const int aj = sd.atom0 + wk.j0 + tid;
    const int sj = A.atom_spec[aj];
    s_jspec[tid] = sj;
    s_jnorb[tid] = A.tab.spec_norb[sj];
    s_jorb[tid] = A.atom_orb[aj];
  }
  __syncthreads();
  if (tid < MAXCOL) {
    int c0 = 0, t = 0;
    for (; t < nja; ++t) {
      if (tid < c0 + s_jnorb[t]) break;
      c0 += s_jnorb[t];
    }
    if (t < nja) {
      const int o = tid - c0;
      col_atom[tid] = wk.j0 + t;
      col_slot[tid] = A.tab.spec_slot[s_jspec[t] * 9 + o];
      col_orb[tid] = s_jorb[t] + o;
    }
  }

  const int nel = norb_i * wk.ncol;
  double2 accH[3][KC], accS[3][KC];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int k = 0; k < KC; ++k) { accH[q][k] = make_double2(0.0, 0.0); accS[q][k] = make_double2(0.0, 0.0); }

  const long long ncand = (long long)nja * sd.nt;
  for (long long seg0 = 0; seg0 < ncand; seg0 += SEG) {
    const int segn = (int)((ncand - seg0) < SEG ? (ncand - seg0) : SEG);
    // each thread screens a contiguous run of candidates -> deterministic order
    const int per = (segn + BUILD_THREADS - 1) / BUILD_THREADS;
    const int beg = tid * per;
    const int end = min(beg + per, segn);
    int mine = 0;
    for (int c = beg; c < end; ++c) {
      const long long g = seg0 + c;
      const int jr = (int)(g / sd.nt), t = (int)(g % sd.nt);
      const int aj = sd.atom0 + wk.j0 + jr;
      const int tt = sd.t0 + t;
      const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
      const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
      const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
      const double r = sqrt(vx * vx + vy * vy + vz * vz);
      if (r < A.cutoff && r > 0.0) ++mine;
    }
    // exclusive prefix over the workgroup: wave scan, then wave offsets
    const int lane = tid & 63, wv = tid >> 6;
    int inc = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(inc, d, 64);
      if (lane >= d) inc += v;
    }
    if (lane == 63) s_wsum[wv] = inc;
    __syncthreads();
    int woff = 0, total_all = 0;
#pragma unroll
    for (int q = 0; q < BUILD_THREADS / 64; ++q) {
      if (q < wv) woff += s_wsum[q];
      total_all += s_wsum[q];
    }
    if (tid == 0) s_total = total_all;
    int pos = woff + inc - mine;
    for (int c = beg; c < end; ++c) {
      const long long g = seg0 + c;
      const int jr = (int)(g / sd.nt), t = (int)(g % sd.nt);
      const int aj = sd.atom0 + wk.j0 + jr;
      const int tt = sd.t0 + t;
      const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
      const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
      const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
      const double r = sqrt(vx * vx + vy * vy + vz * vz);
      if (r < A.cutoff && r > 0.0) s_list[pos++] = (jr << 20) | t;
    }
    __syncthreads();
    const int total = s_total;

    for (int b0 = 0; b0 < total; b0 += MAXPI) {
      const int np = min(MAXPI, total - b0);
      // (a) geometry + interpolation weights per pair-image and table direction
      if (tid < 2 * np) {
        const int p = tid >> 1, dir = tid & 1;
        const int packed = s_list[b0 + p];
        const int jr = packed >> 20, t = packed & 0xFFFFF;
        const int aj = sd.atom0 + wk.j0 + jr;
        const int tt = sd.t0 + t;
        const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
        const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
        const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
        const double r = sqrt(vx * vx + vy * vy + vz * vz);
        const int sj = A.atom_spec[aj];
        const int pid = dir == 0 ? A.tab.pair_index[si * A.tab.nspec + sj]
                                 : A.tab.pair_index[sj * A.tab.nspec + si];
        if (dir == 0) {
          pi_j[p] = wk.j0 + jr;
          pi_t[p] = t;
          pi_u[p][0] = vx / r; pi_u[p][1] = vy / r; pi_u[p][2] = vz / r;
          int mj = 0;
          for (int o = 0; o < A.tab.spec_norb[sj]; ++o) mj |= 1 << A.tab.spec_slot[sj * 9 + o];
          pi_mask[p] = mj;
        }
        int base = -1;
        double w[8];
        if (pid >= 0) {
          interp_weights(r, A.tab.pair_ngrid[pid], A.tab.pair_step[pid], A.tab.tail, &base, w);
          if (base >= 0) base += A.tab.pair_row0[pid];
        }
        pi_base[p][dir] = base;
        for (int m = 0; m < 8; ++m) pi_w[p][dir][m] = (base >= 0) ? w[m] : 0.0;
      }
      // pair-images are in candidate order, i.e. sorted by column atom: the
      // range of each column atom within the batch
      if (tid < nja) {
        int lo = np, hi = np;
        for (int p = 0; p < np; ++p)
          if (s_list[b0 + p] >> 20 == tid) {
            if (lo == np) lo = p;
            hi = p + 1;
          }
        s_rng[tid][0] = lo;
        s_rng[tid][1] = lo == np ? np : hi;
      }
      __syncthreads();
      // (b) integrals
      for (int task = tid; task < np * 40; task += BUILD_THREADS) {
        const int p = task / 40, rem = task % 40, dir = rem / 20, ig = rem % 20;
        const int base = pi_base[p][dir];
        double v = 0.0;
        if (base >= 0) {
          const double* tb = A.tab.tables + (size_t)base * 20 + ig;
#pragma unroll
          for (int m = 0; m < 8; ++m) v += pi_w[p][dir][m] * tb[(size_t)m * 20];
        }
        pi_int[p][dir][ig] = v;
      }
      // (c) phases exp(2 pi i k.T)
      for (int task = tid; task < np * KC; task += BUILD_THREADS) {
        const int p = task / KC, kk = task % KC;
        double2 v = make_double2(0.0, 0.0);
        if (kk < wk.nkc) {
          const int kg = sd.k0 + wk.kc0 + kk;
          const int tt = sd.t0 + pi_t[p];
          const double arg = 2.0 * (A.kfrac[3 * kg] * A.trans_int[3 * tt] +
                                    A.kfrac[3 * kg + 1] * A.trans_int[3 * tt + 1] +
                                    A.kfrac[3 * kg + 2] * A.trans_int[3 * tt + 2]);
          double sn, cs;
          sincospi(arg, &sn, &cs);
          v = make_double2(cs, sn);
        }
        ph[p][kk] = v;
      }
      __syncthreads();
      // (d) effective SK blocks (canonical slots) per pair-image
      for (int task = tid; task < np * 81; task += BUILD_THREADS) {
        const int p = task / 81, ab = task % 81, a = ab / 9, b = ab % 9;
        // slot pairs absent from the (i, j) species pair are never read
        if (!((mask_i >> a) & (pi_mask[p] >> b) & 1)) continue;
        const int la = c_slot_l[a], lb = c_slot_l[b];
        const int j = pi_j[p];
        const bool direct = (la < lb) || (la == lb && wk.i <= j);
        double eh, es;
        const double ux = pi_u[p][0], uy = pi_u[p][1], uz = pi_u[p][2];
        if (direct) {
          const int ib = lpair_base(la, lb);
          eh = sk_elem(la, lb, c_slot_m[a], c_slot_m[b], ux, uy, uz, &pi_int[p][0][ib]);
          es = sk_elem(la, lb, c_slot_m[a], c_slot_m[b], ux, uy, uz, &pi_int[p][0][10 + ib]);
        } else {
          // conjugate of the reversed pair's block: table (Z_j, Z_i), direction -u
          const int ib = lpair_base(lb, la);
          eh = sk_elem(lb, la, c_slot_m[b], c_slot_m[a], -ux, -uy, -uz, &pi_int[p][1][ib]);
          es = sk_elem(lb, la, c_slot_m[b], c_slot_m[a], -ux, -uy, -uz, &pi_int[p][1][10 + ib]);
        }
        E[p][0][ab] = eh;
        E[p][1][ab] = es;
      }
      __syncthreads();
      // (e) accumulate into the tile
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int e = tid + q * BUILD_THREADS;
        if (e < nel) {
          const int ml = e / wk.ncol, c = e % wk.ncol;
          const int a = A.tab.spec_slot[si * 9 + ml];
          const int j = col_atom[c], b = col_slot[c];
          const int p1 = s_rng[j - wk.j0][1];
          for (int p = s_rng[j - wk.j0][0]; p < p1; ++p) {
            const double eh = E[p][0][a * 9 + b], es = E[p][1][a * 9 + b];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
              const double2 f = ph[p][kk];
              accH[q][kk].x += eh * f.x; accH[q][kk].y += eh * f.y;
              accS[q][kk].x += es * f.x; accS[q][kk].y += es * f.y;
            }
          }
        }
      }
      __syncthreads();
    }
  }

  // write the tile (+ on-site on the diagonal)
  OUT* H = reinterpret_cast<OUT*>(A.H);
  OUT* S = reinterpret_cast<OUT*>(A.S);
  const int oi = A.atom_orb[ai];
  const size_t mat_el = (size_t)sd.ld * sd.ld;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int e = tid + q * BUILD_THREADS;
    if (e < nel) {
      const int ml = e / wk.ncol, c = e % wk.ncol;
      const int row = oi + ml, col = col_orb[c];
      double onh = 0.0, ons = 0.0;
      if (row == col) {
        const int a = A.tab.spec_slot[si * 9 + ml];
        onh = A.tab.spec_onsite[si * 3 + c_slot_l[a]];
        ons = 1.0;
      }
#pragma unroll
      for (int kk = 0; kk < KC; ++kk) {
        if (kk < wk.nkc) {
          const size_t off = (size_t)(sd.mat0 + wk.kc0 + kk) * mat_el + (size_t)row * sd.ld + col;
          store_c<OUT>(H + off, make_double2(accH[q][kk].x + onh, accH[q][kk].y));
          store_c<OUT>(S + off, make_double2(accS[q][kk].x + ons, accS[q][kk].y));
        }
      }
    }
  }
}

// Column tiles: one workgroup per work item.
template <typename OUT>
__global__ __launch_bounds__(BUILD_THREADS) void slk_build_kernel(SlkBuildArgs A, unsigned long long* clk) {
  sample_start(clk);
  build_tile<OUT>(A, A.work[blockIdx.x]);
  if (clk != nullptr) {
    __syncthreads();
    sample_end(clk);
  }
}

// Whole-matrix fill of the batch: rows [FILL_ROWS * blockIdx.x, +FILL_ROWS)
// of matrix blockIdx.y. Zeros off the diagonal, on-site energies (S = 1) on
// the diagonal of rows < n_orb, and identity padding (H = pad_h, S = 1) on
// the diagonal of rows >= n_orb. pad_only: only the padded rows and columns.
#define FILL_ROWS 8
template <typename OUT>
__global__ __launch_bounds__(256) void slk_fill_kernel(SlkBuildArgs A, int pad_only, unsigned long long* clk) {
  sample_start(clk);
  const int m = blockIdx.y, ld = A.ld;
  // structure of matrix m: the last s with mat0 <= m
  int lo = 0, hi = A.n_struct - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A.structs[mid].mat0 <= m) lo = mid; else hi = mid - 1;
  }
  const SlkStructDesc sd = A.structs[lo];
  const int n = sd.n_orb;
  OUT* H = reinterpret_cast<OUT*>(A.H) + (size_t)m * ld * ld;
  OUT* S = reinterpret_cast<OUT*>(A.S) + (size_t)m * ld * ld;
  const int r1 = min(ld, (int)(blockIdx.x + 1) * FILL_ROWS);
  for (int row = blockIdx.x * FILL_ROWS; row < r1; ++row) {
    const bool prow = row >= n;
    const int c0 = (pad_only && !prow) ? n : 0;
    for (int col = c0 + threadIdx.x; col < ld; col += 256) {
      double h = 0.0, sv = 0.0;
      if (col == row) {
        sv = 1.0;
        if (prow) {
          h = A.pad_h;
        } else {
          // atom of orbital row: the last atom whose first orbital is <= row
          int a0 = 0, a1 = sd.n_atoms - 1;
          while (a0 < a1) {
            const int mid = (a0 + a1 + 1) >> 1;
            if (A.atom_orb[sd.atom0 + mid] <= row) a0 = mid; else a1 = mid - 1;
          }
          const int sp = A.atom_spec[sd.atom0 + a0];
          const int ml = row - A.atom_orb[sd.atom0 + a0];
          h = A.tab.spec_onsite[sp * 3 + c_slot_l[A.tab.spec_slot[sp * 9 + ml]]];
        }
      }
      store_c<OUT>(H + (size_t)row * ld + col, make_double2(h, 0.0));
      store_c<OUT>(S + (size_t)row * ld + col, make_double2(sv, 0.0));
    }
  }
  if (clk != nullptr) {
    __syncthreads();
    sample_end(clk);
  }
}

// Row panels: one workgroup per (structure, atom i, chunk of <= KC k-points)
// writes the blocks (i, j) of every atom j that has a pair-image within the
// cutoff, over the whole-matrix fill (zeros and the on-site diagonal) of the
// same launch sequence. In a structure whose atoms are not in spatial order,
// almost every 64-column tile holds some neighbour of atom i, so the tile
// form runs one short, latency-bound workgroup per tile; here one workgroup
// screens all (j, T) candidates of atom i in the tile form's order (j-major,
// T-minor), computes weights, integrals, phases and SK blocks per batch of
// pair-images exactly as the tile form does, and walks the pair-images in
// order: each thread keeps the (slot pair, k) entries q = tid, tid + 256 of
// the current atom j's block in registers and writes them when j changes
// (on-site added when j == i). Every element gets the same terms in the same
// order as in the tile form, so the two agree bit for bit.
template <typename OUT>
__global__ __launch_bounds__(BUILD_THREADS, 3) void slk_build_rows_kernel(SlkBuildArgs A, unsigned long long* clk) {
  sample_start(clk);
  constexpr int NQ = (81 * KC + BUILD_THREADS - 1) / BUILD_THREADS;
  __shared__ int s_list[ROWS_SEG + MAXPI];   // pending pair-images (candidate index), in order
  __shared__ int pi_j[MAXPI];
  __shared__ int pi_t[MAXPI];
  __shared__ double pi_u[MAXPI][3];
  __shared__ int pi_base[MAXPI][2];
  __shared__ double pi_w[MAXPI][2][8];
  __shared__ double pi_int[MAXPI][2][20];
  __shared__ double E[MAXPI][2][81];
  __shared__ double2 ph[MAXPI][KC];
  __shared__ int pi_mask[MAXPI];
  __shared__ int pi_sj[MAXPI], pi_oj[MAXPI];        // species and first orbital of the pair-image's atom j
  __shared__ int s_wsum[BUILD_THREADS / 64];
  __shared__ signed char s_inv[ROWS_MAXSPEC * 9];   // species, slot -> local orbital (-1: absent)
  __shared__ int s_smask[ROWS_MAXSPEC];              // slots present per species

  const SlkBuildWork wk = A.work[A.n_work + blockIdx.x];
  const SlkStructDesc sd = A.structs[wk.s];
  const int tid = threadIdx.x;
  const int nt = sd.nt;
  const int ai = sd.atom0 + wk.i;
  const int si = A.atom_spec[ai];
  const int norb_i = A.tab.spec_norb[si];
  const int oi = A.atom_orb[ai];
  const double rix = A.atom_pos[3 * ai], riy = A.atom_pos[3 * ai + 1], riz = A.atom_pos[3 * ai + 2];
  int mask_i = 0;
  for (int o = 0; o < norb_i; ++o) mask_i |= 1 << A.tab.spec_slot[si * 9 + o];
  for (int e = tid; e < A.tab.nspec * 9; e += BUILD_THREADS) s_inv[e] = -1;
  __syncthreads();
  for (int e = tid; e < A.tab.nspec * 9; e += BUILD_THREADS) {
    const int sp = e / 9, o = e % 9;
    if (o < A.tab.spec_norb[sp]) s_inv[sp * 9 + A.tab.spec_slot[sp * 9 + o]] = (signed char)o;
  }
  for (int sp = tid; sp < A.tab.nspec; sp += BUILD_THREADS) {
    int m = 0;
    for (int o = 0; o < A.tab.spec_norb[sp]; ++o) m |= 1 << A.tab.spec_slot[sp * 9 + o];
    s_smask[sp] = m;
  }
  OUT* H = reinterpret_cast<OUT*>(A.H);
  OUT* S = reinterpret_cast<OUT*>(A.S);
  const size_t mat_el = (size_t)sd.ld * sd.ld;
  const size_t kbase = (size_t)(sd.mat0 + wk.kc0) * mat_el;

  double2 accH[NQ], accS[NQ];
#pragma unroll
  for (int u = 0; u < NQ; ++u) { accH[u] = make_double2(0.0, 0.0); accS[u] = make_double2(0.0, 0.0); }
  int cur_j = -1;
  // writes the block of the current atom j (species sj, first orbital oj;
  // rows of atom i) and clears the accumulators
  auto flush = [&](int sj, int oj) {
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int q = tid + u * BUILD_THREADS;
      if (q < 81 * KC) {
        const int kk = q / 81, ab = q % 81, a = ab / 9, b = ab % 9;
        const int ml = s_inv[si * 9 + a], mc = s_inv[sj * 9 + b];
        if (kk < wk.nkc && ml >= 0 && mc >= 0) {
          const int row = oi + ml, col = oj + mc;
          double onh = 0.0, ons = 0.0;
          if (row == col) {
            onh = A.tab.spec_onsite[si * 3 + c_slot_l[a]];
            ons = 1.0;
          }
          const size_t off = kbase + (size_t)kk * mat_el + (size_t)row * sd.ld + col;
          store_c<OUT>(H + off, make_double2(accH[u].x + onh, accH[u].y));
          store_c<OUT>(S + off, make_double2(accS[u].x + ons, accS[u].y));
        }
      }
      accH[u] = make_double2(0.0, 0.0);
      accS[u] = make_double2(0.0, 0.0);
    }
  };

  const int ncand = sd.n_atoms * nt;   // < 2^30 (the planner makes no row panels otherwise)
  int cur_sj = 0, cur_oj = 0;
  // pair-images found by the screening wait in s_list and are processed in
  // full batches of MAXPI (the last batch may be short); pend entries carry
  // over from one screening segment to the next
  int pend = 0;
  for (int seg0 = 0; seg0 < ncand; seg0 += ROWS_SEG) {
    const int segn = (ncand - seg0) < ROWS_SEG ? (ncand - seg0) : ROWS_SEG;
    // each thread screens a contiguous run of <= ROWS_SEG / 256 candidates
    // (deterministic order) and keeps its in-range flags as bits
    const int per = (segn + BUILD_THREADS - 1) / BUILD_THREADS;
    const int beg = tid * per;
    const int end = min(beg + per, segn);
    unsigned bits = 0u;
    int mine = 0;
    for (int c = beg; c < end; ++c) {
      const int g = seg0 + c;
      const int jr = g / nt, t = g % nt;
      const int aj = sd.atom0 + jr;
      const int tt = sd.t0 + t;
      const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
      const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
      const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
      const double r = sqrt(vx * vx + vy * vy + vz * vz);
      if (r < A.cutoff && r > 0.0) {
        bits |= 1u << (c - beg);
        ++mine;
      }
    }
    const int lane = tid & 63, wv = tid >> 6;
    int inc = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(inc, d, 64);
      if (lane >= d) inc += v;
    }
    if (lane == 63) s_wsum[wv] = inc;
    __syncthreads();
    int woff = 0, total_all = 0;
#pragma unroll
    for (int q = 0; q < BUILD_THREADS / 64; ++q) {
      if (q < wv) woff += s_wsum[q];
      total_all += s_wsum[q];
    }
    int pos = pend + woff + inc - mine;
    for (int c = beg; c < end; ++c)
      if ((bits >> (c - beg)) & 1u) s_list[pos++] = seg0 + c;
    __syncthreads();
    const int total = pend + total_all;
    const bool last = seg0 + ROWS_SEG >= ncand;
    int b0 = 0;
    while (total - b0 >= MAXPI || (last && b0 < total)) {
      const int np = min(MAXPI, total - b0);
      // (a) geometry + interpolation weights per pair-image and table direction
      if (tid < 2 * np) {
        const int p = tid >> 1, dir = tid & 1;
        const int g = s_list[b0 + p];
        const int jr = g / nt, t = g % nt;
        const int aj = sd.atom0 + jr;
        const int tt = sd.t0 + t;
        const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
        const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
        const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
        const double r = sqrt(vx * vx + vy * vy + vz * vz);
        const int sj = A.atom_spec[aj];
        const int pid = dir == 0 ? A.tab.pair_index[si * A.tab.nspec + sj]
                                 : A.tab.pair_index[sj * A.tab.nspec + si];
        if (dir == 0) {
          pi_j[p] = jr;
          pi_t[p] = t;
          pi_sj[p] = sj;
          pi_oj[p] = A.atom_orb[aj];
          pi_u[p][0] = vx / r; pi_u[p][1] = vy / r; pi_u[p][2] = vz / r;
          pi_mask[p] = s_smask[sj];
        }
        int base = -1;
        double w[8];
        if (pid >= 0) {
          interp_weights(r, A.tab.pair_ngrid[pid], A.tab.pair_step[pid], A.tab.tail, &base, w);
          if (base >= 0) base += A.tab.pair_row0[pid];
        }
        pi_base[p][dir] = base;
        for (int m = 0; m < 8; ++m) pi_w[p][dir][m] = (base >= 0) ? w[m] : 0.0;
      }
      __syncthreads();
      // (b) integrals
      for (int task = tid; task < np * 40; task += BUILD_THREADS) {
        const int p = task / 40, rem = task % 40, dir = rem / 20, ig = rem % 20;
        const int base = pi_base[p][dir];
        double v = 0.0;
        if (base >= 0) {
          const double* tb = A.tab.tables + (size_t)base * 20 + ig;
#pragma unroll
          for (int m = 0; m < 8; ++m) v += pi_w[p][dir][m] * tb[(size_t)m * 20];
        }
        pi_int[p][dir][ig] = v;
      }
      // (c) phases exp(2 pi i k.T)
      for (int task = tid; task < np * KC; task += BUILD_THREADS) {
        const int p = task / KC, kk = task % KC;
        double2 v = make_double2(0.0, 0.0);
        if (kk < wk.nkc) {
          const int kg = sd.k0 + wk.kc0 + kk;
          const int tt = sd.t0 + pi_t[p];
          const double arg = 2.0 * (A.kfrac[3 * kg] * A.trans_int[3 * tt] +
                                    A.kfrac[3 * kg + 1] * A.trans_int[3 * tt + 1] +
                                    A.kfrac[3 * kg + 2] * A.trans_int[3 * tt + 2]);
          double sn, cs;
          sincospi(arg, &sn, &cs);
          v = make_double2(cs, sn);
        }
        ph[p][kk] = v;
      }
      __syncthreads();
      // (d) effective SK blocks (canonical slots) per pair-image
      for (int task = tid; task < np * 81; task += BUILD_THREADS) {
        const int p = task / 81, ab = task % 81, a = ab / 9, b = ab % 9;
        if (!((mask_i >> a) & (pi_mask[p] >> b) & 1)) continue;
        const int la = c_slot_l[a], lb = c_slot_l[b];
        const int j = pi_j[p];
        const bool direct = (la < lb) || (la == lb && wk.i <= j);
        double eh, es;
        const double ux = pi_u[p][0], uy = pi_u[p][1], uz = pi_u[p][2];
        if (direct) {
          const int ib = lpair_base(la, lb);
          eh = sk_elem(la, lb, c_slot_m[a], c_slot_m[b], ux, uy, uz, &pi_int[p][0][ib]);
          es = sk_elem(la, lb, c_slot_m[a], c_slot_m[b], ux, uy, uz, &pi_int[p][0][10 + ib]);
        } else {
          const int ib = lpair_base(lb, la);
          eh = sk_elem(lb, la, c_slot_m[b], c_slot_m[a], -ux, -uy, -uz, &pi_int[p][1][ib]);
          es = sk_elem(lb, la, c_slot_m[b], c_slot_m[a], -ux, -uy, -uz, &pi_int[p][1][10 + ib]);
        }
        E[p][0][ab] = eh;
        E[p][1][ab] = es;
      }
      __syncthreads();
      // (e) accumulate; a new atom j writes the previous one's block first
      for (int p = 0; p < np; ++p) {
        const int j = pi_j[p];
        if (j != cur_j) {
          if (cur_j >= 0) flush(cur_sj, cur_oj);
          cur_j = j;
          cur_sj = pi_sj[p];
          cur_oj = pi_oj[p];
        }
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
          const int q = tid + u * BUILD_THREADS;
          if (q < 81 * KC) {
            const int kk = q / 81, ab = q % 81;
            const double eh = E[p][0][ab], es = E[p][1][ab];
            const double2 f = ph[p][kk];
            accH[u].x += eh * f.x; accH[u].y += eh * f.y;
            accS[u].x += es * f.x; accS[u].y += es * f.y;
          }
        }
      }
      __syncthreads();
      b0 += np;
    }
    // carry the unprocessed tail (< MAXPI entries) to the front
    const int rem = total - b0;
    const int keep = tid < rem ? s_list[b0 + tid] : 0;
    __syncthreads();
    if (tid < rem) s_list[tid] = keep;
    __syncthreads();
    pend = rem;
  }
  if (cur_j >= 0) flush(cur_sj, cur_oj);
  if (clk != nullptr) {
    __syncthreads();
    sample_end(clk);
  }
}

// Padding rows/columns [n_orb, ld) of every k-matrix: zero off-diagonal,
// H diagonal = pad_h, S diagonal = 1. Keeps padded eigenvalues at pad_h.
template <typename OUT>
__global__ void slk_pad_kernel(const SlkStructDesc* structs, int s, void* Hv, void* Sv, double pad_h) {
  const SlkStructDesc sd = structs[s];
  const int n = sd.n_orb, ld = sd.ld;
  OUT* H = reinterpret_cast<OUT*>(Hv);
  OUT* S = reinterpret_cast<OUT*>(Sv);
  const size_t mat = (size_t)(sd.mat0 + blockIdx.y) * ld * ld;
  const long long padcount = (long long)ld * ld - (long long)n * n;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < padcount;
       idx += (long long)gridDim.x * blockDim.x) {
    int r, c;
    // rows < n: columns [n, ld); rows >= n: all columns
    const long long top = (long long)n * (ld - n);
    if (idx < top) { r = (int)(idx / (ld - n)); c = n + (int)(idx % (ld - n)); }
    else { long long k = idx - top; r = n + (int)(k / ld); c = (int)(k % ld); }
    const size_t off = mat + (size_t)r * ld + c;
    double hv = 0.0, sv = 0.0;
    if (r == c) { hv = pad_h; sv = 1.0; }
    store_c<OUT>(H + off, make_double2(hv, 0.0));
    store_c<OUT>(S + off, make_double2(sv, 0.0));
  }
}

extern "C" {

static int g_build_mode = 0;   // whole-matrix form: 0 row panels, 1 every column tile

int slk_set_build_mode(int mode) {
  if (mode < 0 || mode > 1) return 1;
  g_build_mode = mode;
  return 0;
}

extern "C++" {
template <typename OUT>
static void build_launch(const SlkBuildArgs& a, hipStream_t stream, bool sampled, const SampleSlot& sl) {
  unsigned long long* clk = sampled ? sl.clk : nullptr;
  if (a.n_mat <= 0) {
    if (sampled)
      hipExtLaunchKernelGGL(slk_build_kernel<OUT>, dim3(a.n_work), dim3(BUILD_THREADS), 0, stream, sl.e0, sl.e1, 0,
                            a, clk);
    else
      hipLaunchKernelGGL(slk_build_kernel<OUT>, dim3(a.n_work), dim3(BUILD_THREADS), 0, stream, a, clk);
    return;
  }
  // whole-matrix form: the fill, then the blocks by row panels (or by every
  // column tile); a sampled launch's window and events span both
  // row panels from ROWS_MIN_LD up; below it a structure has few atoms, so
  // few row panels, and the tiles measured faster
  const bool rows = g_build_mode == 0 && a.n_row_work > 0 && a.tab.nspec <= ROWS_MAXSPEC && a.ld >= ROWS_MIN_LD;
  const dim3 fgrid((a.ld + FILL_ROWS - 1) / FILL_ROWS, a.n_mat);
  if (sampled)
    hipExtLaunchKernelGGL(slk_fill_kernel<OUT>, fgrid, dim3(256), 0, stream, sl.e0, nullptr, 0, a, rows ? 0 : 1, clk);
  else
    hipLaunchKernelGGL(slk_fill_kernel<OUT>, fgrid, dim3(256), 0, stream, a, rows ? 0 : 1, clk);
  if (rows) {
    if (sampled)
      hipExtLaunchKernelGGL(slk_build_rows_kernel<OUT>, dim3(a.n_row_work), dim3(BUILD_THREADS), 0, stream, nullptr,
                            sl.e1, 0, a, clk);
    else
      hipLaunchKernelGGL(slk_build_rows_kernel<OUT>, dim3(a.n_row_work), dim3(BUILD_THREADS), 0, stream, a, clk);
  } else {
    if (sampled)
      hipExtLaunchKernelGGL(slk_build_kernel<OUT>, dim3(a.n_work), dim3(BUILD_THREADS), 0, stream, nullptr, sl.e1, 0,
                            a, clk);
    else
      hipLaunchKernelGGL(slk_build_kernel<OUT>, dim3(a.n_work), dim3(BUILD_THREADS), 0, stream, a, clk);
  }
}
}  // extern "C++"

int slk_build_hs_launch(const SlkBuildArgs* args, hipStream_t stream) {
  if (args->n_work <= 0) return 0;
  if (args->n_mat > 0 && (args->n_struct <= 0 || args->ld <= 0)) return (int)hipErrorInvalidValue;
  SampleSlot sl;
  const bool sampled = sampler_take(SLK_SAMPLE_BUILD, -1.0, &sl);
  if (args->out_f32)
    build_launch<float2>(*args, stream, sampled, sl);
  else
    build_launch<double2>(*args, stream, sampled, sl);
  return (int)hipGetLastError();
}

int slk_pad_launch(const SlkStructDesc* structs_dev, int s, int nk, int n_orb, int ld, void* H, void* S,
                   double pad_h, int out_f32, hipStream_t stream) {
  if (ld <= n_orb || nk <= 0) return 0;
  const long long padcount = (long long)ld * ld - (long long)n_orb * n_orb;
  int blocks = (int)((padcount + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  if (out_f32)
    hipLaunchKernelGGL(slk_pad_kernel<float2>, dim3(blocks, nk), dim3(256), 0, stream, structs_dev, s, H, S, pad_h);
  else
    hipLaunchKernelGGL(slk_pad_kernel<double2>, dim3(blocks, nk), dim3(256), 0, stream, structs_dev, s, H, S, pad_h);
  return (int)hipGetLastError();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint of the build with respect to the table values.
//
// Given per-k weighted density matrices Q_H(k) = C diag(g) C^H and
// Q_S(k) = -C diag(g e) C^H (g = dL/de, C the S-orthonormal eigenvectors),
// dL = sum_k Re sum_{mu,nu} Q(k)[nu][mu] dM(k)[mu][nu] for M = H, S. With
// M(k)[mu][nu] = sum_T E(i,j,T)[a][b] exp(2 pi i k.T) this is
// dL = sum_{(i,j,T),a,b} G[a][b] dE[a][b], G = Re sum_k Q(k)[nu][mu] e^{2 pi i k.T}.
// E is linear in the SK integrals (sk_elem with unit integral vectors gives
// the coefficients) and the integrals are linear in the table rows
// (interpolation weights), so G is scattered back onto the table with
// atomic adds. Work decomposition mirrors the build: one workgroup per
// (atom row i, column tile, k range [kc0, kc0 + nkc)); k-points of the range
// are summed inside, ranges and tiles meet in the FP64 atomics (so the last
// bits of the gradient depend on the order the atomics land in).
// ---------------------------------------------------------------------------
#define ADJ_KCH 8

// MODE 0: gradient with respect to the table rows (out = dtab [rows][20]).
// MODE 1: gradient with respect to the atom positions (out = [atoms][3]):
// each pair-image block E(d) depends on d = R_j + T - R_i through the
// direction cosines and the interpolated integrals; dual numbers give
// dE/dd exactly, and -dE/dd / +dE/dd go to atoms i / j.
template <int MODE>
__global__ __launch_bounds__(BUILD_THREADS) void slk_adjoint_kernel(SlkBuildArgs A, const double2* __restrict__ QH,
                                                                    const double2* __restrict__ QS,
                                                                    double* __restrict__ dtab) {
  __shared__ int s_list[SEG];
  __shared__ int s_cnt[BUILD_THREADS];
  __shared__ int s_total;
  __shared__ int pi_j[MAXPI];
  __shared__ int pi_t[MAXPI];
  __shared__ double pi_u[MAXPI][3];
  __shared__ int pi_base[MAXPI][2];
  __shared__ double pi_w[MAXPI][2][8];
  __shared__ double G[MAXPI][2][81];
  __shared__ double dV[MAXPI][2][20];
  __shared__ double2 ph[MAXPI][ADJ_KCH];
  __shared__ double pi_r[MAXPI];
  __shared__ double pi_dw[MAXPI][2][8];
  __shared__ double gd[MAXPI][3];
  __shared__ int col_atom[MAXCOL], col_slot[MAXCOL], col_orb[MAXCOL];

  const SlkBuildWork wk = A.work[blockIdx.x];
  const SlkStructDesc sd = A.structs[wk.s];
  const int tid = threadIdx.x;
  const int ai = sd.atom0 + wk.i;
  const int si = A.atom_spec[ai];
  const int norb_i = A.tab.spec_norb[si];
  const int oi = A.atom_orb[ai];
  const double rix = A.atom_pos[3 * ai], riy = A.atom_pos[3 * ai + 1], riz = A.atom_pos[3 * ai + 2];
  const int nja = wk.j1 - wk.j0;
  // slots (s, p, d components) present on atom i
  int mask_i = 0;
  for (int o = 0; o < norb_i; ++o) mask_i |= 1 << A.tab.spec_slot[si * 9 + o];
  const size_t mat_el = (size_t)sd.ld * sd.ld;
  if (tid == 0) {
    int c = 0;
    for (int j = wk.j0; j < wk.j1; ++j) {
      const int aj = sd.atom0 + j;
      const int sj = A.atom_spec[aj];
      const int nb = A.tab.spec_norb[sj];
      for (int o = 0; o < nb && c < MAXCOL; ++o, ++c) {
        col_atom[c] = j;
        col_slot[c] = A.tab.spec_slot[sj * 9 + o];
        col_orb[c] = A.atom_orb[aj] + o;
      }
    }
  }
  __syncthreads();
  const long long ncand = (long long)nja * sd.nt;
  for (long long seg0 = 0; seg0 < ncand; seg0 += SEG) {
    const int segn = (int)((ncand - seg0) < SEG ? (ncand - seg0) : SEG);
    const int per = (segn + BUILD_THREADS - 1) / BUILD_THREADS;
    const int beg = tid * per;
    const int end = min(beg + per, segn);
    int mine = 0;
    for (int c = beg; c < end; ++c) {
      const long long g = seg0 + c;
      const int jr = (int)(g / sd.nt), t = (int)(g % sd.nt);
      const int aj = sd.atom0 + wk.j0 + jr;
      const int tt = sd.t0 + t;
      const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
      const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
      const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
      const double r = sqrt(vx * vx + vy * vy + vz * vz);
      if (r < A.cutoff && r > 0.0) ++mine;
    }
    s_cnt[tid] = mine;
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int q = 0; q < BUILD_THREADS; ++q) { int v = s_cnt[q]; s_cnt[q] = acc; acc += v; }
      s_total = acc;
    }
    __syncthreads();
    int pos = s_cnt[tid];
    for (int c = beg; c < end; ++c) {
      const long long g = seg0 + c;
      const int jr = (int)(g / sd.nt), t = (int)(g % sd.nt);
      const int aj = sd.atom0 + wk.j0 + jr;
      const int tt = sd.t0 + t;
      const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
      const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
      const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
      const double r = sqrt(vx * vx + vy * vy + vz * vz);
      if (r < A.cutoff && r > 0.0) s_list[pos++] = (jr << 20) | t;
    }
    __syncthreads();
    const int total = s_total;
    for (int b0 = 0; b0 < total; b0 += MAXPI) {
      const int np = min(MAXPI, total - b0);
      if (tid < 2 * np) {
        const int p = tid >> 1, dir = tid & 1;
        const int packed = s_list[b0 + p];
        const int jr = packed >> 20, t = packed & 0xFFFFF;
        const int aj = sd.atom0 + wk.j0 + jr;
        const int tt = sd.t0 + t;
        const double vx = (A.trans_cart[3 * tt] + A.atom_pos[3 * aj]) - rix;
        const double vy = (A.trans_cart[3 * tt + 1] + A.atom_pos[3 * aj + 1]) - riy;
        const double vz = (A.trans_cart[3 * tt + 2] + A.atom_pos[3 * aj + 2]) - riz;
        const double r = sqrt(vx * vx + vy * vy + vz * vz);
        const int sj = A.atom_spec[aj];
        const int pid = dir == 0 ? A.tab.pair_index[si * A.tab.nspec + sj]
                                 : A.tab.pair_index[sj * A.tab.nspec + si];
        if (dir == 0) {
          pi_j[p] = wk.j0 + jr;
          pi_t[p] = t;
          pi_u[p][0] = vx / r; pi_u[p][1] = vy / r; pi_u[p][2] = vz / r;
        }
        int base = -1;
        Dual<1> w[8];
        if (pid >= 0) {
          Dual<1> rd(r);
          rd.g[0] = 1.0;
          interp_weights_t<Dual<1>>(rd, A.tab.pair_ngrid[pid], A.tab.pair_step[pid], A.tab.tail, &base, w);
          if (base >= 0) base += A.tab.pair_row0[pid];
        }
        pi_base[p][dir] = base;
        if (dir == 0) pi_r[p] = r;
        for (int m = 0; m < 8; ++m) {
          pi_w[p][dir][m] = (base >= 0) ? w[m].v : 0.0;
          pi_dw[p][dir][m] = (base >= 0) ? w[m].g[0] : 0.0;
        }
      }
      for (int q = tid; q < np * 2 * 81; q += BUILD_THREADS) (&G[0][0][0])[q] = 0.0;
      for (int q = tid; q < np * 3; q += BUILD_THREADS) (&gd[0][0])[q] = 0.0;
      for (int q = tid; q < np * 2 * 20; q += BUILD_THREADS) (&dV[0][0][0])[q] = 0.0;
      __syncthreads();
      // G[p][H/S][a*9+b] = Re sum_k Q(k)[nu][mu] exp(2 pi i k.T), over k chunks
      for (int k0 = 0; k0 < wk.nkc; k0 += ADJ_KCH) {
        const int nkc = min(ADJ_KCH, wk.nkc - k0);
        for (int task = tid; task < np * ADJ_KCH; task += BUILD_THREADS) {
          const int p = task / ADJ_KCH, kk = task % ADJ_KCH;
          double2 v = make_double2(0.0, 0.0);
          if (kk < nkc) {
            const int kg = sd.k0 + wk.kc0 + k0 + kk;
            const int tt = sd.t0 + pi_t[p];
            const double arg = 2.0 * (A.kfrac[3 * kg] * A.trans_int[3 * tt] +
                                      A.kfrac[3 * kg + 1] * A.trans_int[3 * tt + 1] +
                                      A.kfrac[3 * kg + 2] * A.trans_int[3 * tt + 2]);
            double sn, cs;
            sincospi(arg, &sn, &cs);
            v = make_double2(cs, sn);
          }
          ph[p][kk] = v;
        }
        __syncthreads();
        for (int task = tid; task < np * 81; task += BUILD_THREADS) {
          const int p = task / 81, ab = task % 81, a = ab / 9, b = ab % 9;
          // local orbitals of slots a (atom i) and b (atom j), if present
          int ml = -1, cl = -1;
          for (int o = 0; o < norb_i; ++o)
            if (A.tab.spec_slot[si * 9 + o] == a) ml = o;
          const int j = pi_j[p];
          for (int c = 0; c < wk.ncol; ++c)
            if (col_atom[c] == j && col_slot[c] == b) cl = c;
          if (ml < 0 || cl < 0) continue;
          const size_t mu = (size_t)(oi + ml), nu = (size_t)col_orb[cl];
          double gh = 0.0, gs = 0.0;
          for (int kk = 0; kk < nkc; ++kk) {
            const size_t off = (size_t)(sd.mat0 + wk.kc0 + k0 + kk) * mat_el + nu * sd.ld + mu;
            const double2 f = ph[p][kk];
            const double2 qh = QH[off], qs = QS[off];
            gh += qh.x * f.x - qh.y * f.y;
            gs += qs.x * f.x - qs.y * f.y;
          }
          G[p][0][ab] += gh;
          G[p][1][ab] += gs;
        }
        __syncthreads();
      }
      if (MODE == 0) {
      // contract with the SK coefficients (unit-integral evaluations)
      for (int task = tid; task < np * 81; task += BUILD_THREADS) {
        const int p = task / 81, ab = task % 81, a = ab / 9, b = ab % 9;
        const double gh = G[p][0][ab], gs = G[p][1][ab];
        if (gh == 0.0 && gs == 0.0) continue;
        const int la = c_slot_l[a], lb = c_slot_l[b];
        const int j = pi_j[p];
        const bool direct = (la < lb) || (la == lb && wk.i <= j);
        const double ux = pi_u[p][0], uy = pi_u[p][1], uz = pi_u[p][2];
        int l1, l2, m1, m2, dir;
        double x, y, z;
        if (direct) { l1 = la; l2 = lb; m1 = c_slot_m[a]; m2 = c_slot_m[b]; dir = 0; x = ux; y = uy; z = uz; }
        else { l1 = lb; l2 = la; m1 = c_slot_m[b]; m2 = c_slot_m[a]; dir = 1; x = -ux; y = -uy; z = -uz; }
        const int ib = lpair_base(l1, l2);
        for (int q = 0; q <= l1; ++q) {
          double unit[3] = {0.0, 0.0, 0.0};
          unit[q] = 1.0;
          const double cf = sk_elem(l1, l2, m1, m2, x, y, z, unit);
          if (cf != 0.0) {
            atomicAdd(&dV[p][dir][ib + q], gh * cf);
            atomicAdd(&dV[p][dir][10 + ib + q], gs * cf);
          }
        }
      }
      __syncthreads();
      // scatter onto the table rows through the interpolation weights
      for (int task = tid; task < np * 2 * 20; task += BUILD_THREADS) {
        const int p = task / 40, rem = task % 40, dir = rem / 20, ig = rem % 20;
        const int base = pi_base[p][dir];
        const double g = dV[p][dir][ig];
        if (base < 0 || g == 0.0) continue;
        for (int m = 0; m < 8; ++m) atomicAdd(&dtab[(size_t)(base + m) * 20 + ig], pi_w[p][dir][m] * g);
      }
      } else {
        // dE/dd for every (pair-image, a, b) with dual numbers over d
        for (int task = tid; task < np * 81; task += BUILD_THREADS) {
          const int p = task / 81, ab = task % 81, a = ab / 9, b = ab % 9;
          const double gh = G[p][0][ab], gs = G[p][1][ab];
          if (gh == 0.0 && gs == 0.0) continue;
          const int la = c_slot_l[a], lb = c_slot_l[b];
          const int j = pi_j[p];
          const bool direct = (la < lb) || (la == lb && wk.i <= j);
          const int dir = direct ? 0 : 1;
          const int base = pi_base[p][dir];
          if (base < 0) continue;
          const double r = pi_r[p];
          const double u[3] = {pi_u[p][0], pi_u[p][1], pi_u[p][2]};
          Dual<3> ud[3];
#pragma unroll
          for (int c0 = 0; c0 < 3; ++c0) {
            ud[c0] = Dual<3>(direct ? u[c0] : -u[c0]);
#pragma unroll
            for (int e2 = 0; e2 < 3; ++e2) {
              const double du = ((c0 == e2 ? 1.0 : 0.0) - u[c0] * u[e2]) / r;
              ud[c0].g[e2] = direct ? du : -du;
            }
          }
          int l1, l2, m1, m2;
          if (direct) { l1 = la; l2 = lb; m1 = c_slot_m[a]; m2 = c_slot_m[b]; }
          else { l1 = lb; l2 = la; m1 = c_slot_m[b]; m2 = c_slot_m[a]; }
          const int ib = lpair_base(l1, l2);
          Dual<3> vh[3], vs[3];
          for (int q = 0; q <= l1; ++q) {
            double h0 = 0.0, h1 = 0.0, s0 = 0.0, s1 = 0.0;
            for (int m = 0; m < 8; ++m) {
              const double* tb = A.tab.tables + (size_t)(base + m) * 20;
              h0 += pi_w[p][dir][m] * tb[ib + q];
              h1 += pi_dw[p][dir][m] * tb[ib + q];
              s0 += pi_w[p][dir][m] * tb[10 + ib + q];
              s1 += pi_dw[p][dir][m] * tb[10 + ib + q];
            }
            vh[q] = Dual<3>(h0);
            vs[q] = Dual<3>(s0);
#pragma unroll
            for (int e2 = 0; e2 < 3; ++e2) {
              vh[q].g[e2] = h1 * u[e2];     // dr/dd_e = u_e
              vs[q].g[e2] = s1 * u[e2];
            }
          }
          const Dual<3> eh = sk_elem_t<Dual<3>>(l1, l2, m1, m2, ud[0], ud[1], ud[2], vh);
          const Dual<3> es = sk_elem_t<Dual<3>>(l1, l2, m1, m2, ud[0], ud[1], ud[2], vs);
#pragma unroll
          for (int e2 = 0; e2 < 3; ++e2) atomicAdd(&gd[p][e2], gh * eh.g[e2] + gs * es.g[e2]);
        }
        __syncthreads();
        for (int task = tid; task < np * 3; task += BUILD_THREADS) {
          const int p = task / 3, e2 = task % 3;
          const double g = gd[p][e2];
          if (g == 0.0) continue;
          const int gi = sd.atom0 + wk.i, gj = sd.atom0 + pi_j[p];
          atomicAdd(&dtab[(size_t)gi * 3 + e2], -g);
          atomicAdd(&dtab[(size_t)gj * 3 + e2], g);
        }
      }
      __syncthreads();
    }
  }
}

// Y[b][r][c] = C[b][r][c] * g[b][c] (complex128 C, real g with row stride gld).
__global__ void scale_columns_kernel(const double2* __restrict__ Cm, const double* __restrict__ g, int n, int ld,
                                     int gld, double2* __restrict__ Y, long long total) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const long long b = q / ((long long)ld * ld);
    const int rem = (int)(q % ((long long)ld * ld));
    const int c = rem % ld;
    const double s = c < n ? g[b * gld + c] : 0.0;
    const double2 v = Cm[q];
    Y[q] = make_double2(v.x * s, v.y * s);
  }
}

extern "C" {

int slk_adjoint_tables_launch(const SlkBuildArgs* args, const void* QH, const void* QS, double* dtab,
                              hipStream_t stream) {
  if (args->n_work <= 0) return 0;
  hipLaunchKernelGGL(slk_adjoint_kernel<0>, dim3(args->n_work), dim3(BUILD_THREADS), 0, stream, *args,
                     reinterpret_cast<const double2*>(QH), reinterpret_cast<const double2*>(QS), dtab);
  return (int)hipGetLastError();
}

int slk_forces_launch(const SlkBuildArgs* args, const void* QH, const void* QS, double* dR, hipStream_t stream) {
  if (args->n_work <= 0) return 0;
  hipLaunchKernelGGL(slk_adjoint_kernel<1>, dim3(args->n_work), dim3(BUILD_THREADS), 0, stream, *args,
                     reinterpret_cast<const double2*>(QH), reinterpret_cast<const double2*>(QS), dR);
  return (int)hipGetLastError();
}

int slk_scale_columns(const void* Cm, const double* g, int n, int ld, int batch, int gld, void* Y,
                      hipStream_t stream) {
  const long long total = (long long)batch * ld * ld;
  if (total <= 0) return 0;
  long long blocks = (total + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(scale_columns_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                     reinterpret_cast<const double2*>(Cm), g, n, ld, gld, reinterpret_cast<double2*>(Y), total);
  return (int)hipGetLastError();
}

}  // extern "C"
