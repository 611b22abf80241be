// Fermi level, band edges and Gaussian DOS as workgroup reductions (gfx950).
//
// Fermi level: root of N(mu) = 2 sum_k w_k sum_b f((e_kb - mu)/kT) = Ne,
// f = 1/(exp(x)+1), kT in Hartree (slakonet/fermi.py:139-215), either by
// bisection to full FP64 resolution (mode 0) or by the reference's half-step
// Newton on (N - Ne)^2 with its stopping rule (mode 1).
// Band edges: calculate_band_gap (slakonet/main.py:637-707), band structure
// statistics (slakonet/main.py:1539-1642) and the get_fermi gap block
// (slakonet/optim.py:1011-1059).
// DOS: sum_k w_k sum_b exp(-((E - e')/sigma)^2/2)/(sigma sqrt(2 pi)) with
// e' = (e - mu) * 27.211 eV (slakonet/main.py:709-797).
#include "slk_common.h"
#include "slakonet_amd.h"

#include "slk_edges.h"
#include "slk_sampler.h"
#include <hip/hip_ext.h>

#define POST_THREADS 256

__device__ __forceinline__ double fd(double e, double mu, double kT) { return 1.0 / (exp((e - mu) / kT) + 1.0); }

__global__ __launch_bounds__(POST_THREADS) void fermi_kernel(const double* __restrict__ W,
                                                             const double* __restrict__ KW,
                                                             const SlkSpectrum* spec, double kT, int mode,
                                                             double tol, int max_iter, double* edges) {
  const SlkSpectrum sp = spec[blockIdx.x];
  const int tid = threadIdx.x;
  const int tot = sp.nk * sp.nb;
  __shared__ double red[POST_THREADS / 64];
  // min, max, sum of eigenvalues
  double lo = 1e300, hi = -1e300, sum = 0.0;
  for (int q = tid; q < tot; q += blockDim.x) {
    const int k = q / sp.nb, b = q % sp.nb;
    const double e = W[sp.w0 + (long long)k * sp.wld + b];
    lo = fmin(lo, e);
    hi = fmax(hi, e);
    sum += e;
  }
  __shared__ double s_lo[POST_THREADS / 64], s_hi[POST_THREADS / 64];
  lo = wave_min(lo);
  hi = wave_max(hi);
  if ((tid & 63) == 0) { s_lo[tid >> 6] = lo; s_hi[tid >> 6] = hi; }
  __syncthreads();
  lo = 1e300; hi = -1e300;
  for (int i = 0; i < POST_THREADS / 64; ++i) { lo = fmin(lo, s_lo[i]); hi = fmax(hi, s_hi[i]); }
  const double mean = block_sum(sum, red) / (double)tot;
  double mu;
  int it_used = 0;
  if (mode == 0) {
    double a = lo - 50.0 * kT, b = hi + 50.0 * kT;
    for (int it = 0; it < 400; ++it) {
      const double mid = 0.5 * (a + b);
      if (mid <= a || mid >= b) break;
      double part = 0.0;
      for (int q = tid; q < tot; q += blockDim.x) {
        const int k = q / sp.nb, bb = q % sp.nb;
        part += KW[sp.kw0 + k] * fd(W[sp.w0 + (long long)k * sp.wld + bb], mid, kT);
      }
      const double n = 2.0 * block_sum(part, red);
      if (n > sp.nelec) b = mid; else a = mid;
      it_used = it + 1;
    }
    mu = 0.5 * (a + b);
  } else {
    mu = mean;
    for (int it = 0; it < max_iter; ++it) {
      double pn = 0.0, pd = 0.0;
      for (int q = tid; q < tot; q += blockDim.x) {
        const int k = q / sp.nb, bb = q % sp.nb;
        const double x = (W[sp.w0 + (long long)k * sp.wld + bb] - mu) / kT;
        const double ex = exp(x);
        const double f = 1.0 / (ex + 1.0);
        const double wk = KW[sp.kw0 + k];
        pn += f * wk;
        pd += (f * f * ex / kT) * wk;
      }
      const double n = 2.0 * block_sum(pn, red);
      const double dn = 2.0 * block_sum(pd, red);
      const double loss = (n - sp.nelec) * (n - sp.nelec);
      const double grad = 2.0 * (n - sp.nelec) * dn;
      it_used = it + 1;
      if (fabs(grad) < tol) break;
      mu = mu - loss / (grad + 1e-12);
    }
  }
  if (tid == 0) {
    edges[(size_t)blockIdx.x * SLK_EDGE_FIELDS + E_MU] = mu;
    edges[(size_t)blockIdx.x * SLK_EDGE_FIELDS + E_NITER] = it_used;
  }
}

// Band edges given mu (reads edges[E_MU]).
__global__ __launch_bounds__(POST_THREADS) void edges_kernel(const double* __restrict__ W, const SlkSpectrum* spec,
                                                             double metal_ev, double* edges) {
  const SlkSpectrum sp = spec[blockIdx.x];
  double* out = edges + (size_t)blockIdx.x * SLK_EDGE_FIELDS;
  const double mu = out[E_MU];
  const double fev = mu * H2E_OUT;
  const int tid = threadIdx.x;
  const int tot = sp.nk * sp.nb;
  // per-thread partials
  double vbm = -1e300, cbm = 1e300, vb2 = -1e300, cb2 = 1e300;
  double svbm = -1e300, svmin = 1e300, scbm = 1e300, scmax = -1e300, mindg = 1e300;
  long long vq = (1LL << 62), cq = (1LL << 62);
  int metal = 0, nval = 0, ncon = 0;
  for (int q = tid; q < tot; q += blockDim.x) {
    const int k = q / sp.nb, b = q % sp.nb;
    const double e = W[sp.w0 + (long long)k * sp.wld + b];
    const double ev = e * H2E_OUT;
    if (fabs(ev - fev) < metal_ev) metal = 1;
    if (ev < fev) { if (ev > vbm) vbm = ev; } else { if (ev < cbm) cbm = ev; }
    if (e <= mu) vb2 = fmax(vb2, e); else cb2 = fmin(cb2, e);
    const double s = (e - mu) * H2E_OUT;
    if (s < 0.0) { svbm = fmax(svbm, s); svmin = fmin(svmin, s); ++nval; }
    else { scbm = fmin(scbm, s); scmax = fmax(scmax, s); ++ncon; }
  }
  // per-k direct gaps: thread per k
  for (int k = tid; k < sp.nk; k += blockDim.x) {
    double kv = -1e300, kc = 1e300;
    int hv = 0, hc = 0;
    for (int b = 0; b < sp.nb; ++b) {
      const double s = (W[sp.w0 + (long long)k * sp.wld + b] - mu) * H2E_OUT;
      if (s < 0.0) { kv = fmax(kv, s); hv = 1; }
      if (s > 0.0) { kc = fmin(kc, s); hc = 1; }
    }
    if (hv && hc) mindg = fmin(mindg, kc - kv);
  }
  __shared__ double sd[10][POST_THREADS / 64];
  __shared__ int si[3][POST_THREADS / 64];
  double vals[10] = {vbm, -cbm, vb2, -cb2, svbm, -svmin, -scbm, scmax, -mindg, 0.0};
  const int lane = tid & 63, wv = tid >> 6;
  for (int f = 0; f < 9; ++f) {
    const double m = wave_max(vals[f]);
    if (lane == 0) sd[f][wv] = m;
  }
  const int mt = wave_max(metal), nv = wave_sum(nval), nc = wave_sum(ncon);
  if (lane == 0) { si[0][wv] = mt; si[1][wv] = nv; si[2][wv] = nc; }
  __syncthreads();
  __shared__ double fin[9];
  __shared__ int fini[3];
  if (tid < 9) {
    double m = -1e300;
    for (int i = 0; i < POST_THREADS / 64; ++i) m = fmax(m, sd[tid][i]);
    fin[tid] = m;
  }
  if (tid == 0) {
    int a = 0, b2 = 0, c = 0;
    for (int i = 0; i < POST_THREADS / 64; ++i) { a = max(a, si[0][i]); b2 += si[1][i]; c += si[2][i]; }
    fini[0] = a; fini[1] = b2; fini[2] = c;
  }
  __syncthreads();
  const double VBM = fin[0], CBM = -fin[1];
  // first occurrence (flattened k-major order) of VBM / CBM
  for (int q = tid; q < tot; q += blockDim.x) {
    const int k = q / sp.nb, b = q % sp.nb;
    const double ev = W[sp.w0 + (long long)k * sp.wld + b] * H2E_OUT;
    if (ev == VBM && q < vq) vq = q;
    if (ev == CBM && q < cq) cq = q;
  }
  __shared__ long long sq[2][POST_THREADS / 64];
  long long a1 = wave_min(vq), a2 = wave_min(cq);
  if (lane == 0) { sq[0][wv] = a1; sq[1][wv] = a2; }
  __syncthreads();
  if (tid == 0) {
    long long m1 = (1LL << 62), m2 = (1LL << 62);
    for (int i = 0; i < POST_THREADS / 64; ++i) { m1 = min(m1, sq[0][i]); m2 = min(m2, sq[1][i]); }
    out[E_METAL] = fini[0];
    out[E_VBM] = VBM;
    out[E_CBM] = CBM;
    out[E_VBMK] = (m1 < (1LL << 62)) ? (double)(m1 / sp.nb) : 0.0;
    out[E_CBMK] = (m2 < (1LL << 62)) ? (double)(m2 / sp.nb) : 0.0;
    out[E_SVBM] = fin[4];
    out[E_VBW] = fin[4] + fin[5];      // max - min over valence
    out[E_SCBM] = -fin[6];
    out[E_CBW] = fin[7] + fin[6];      // max - min over conduction
    out[E_MINDG] = -fin[8];
    out[E_VB2] = fin[2];
    out[E_CB2] = -fin[3];
    out[E_NVAL] = fini[1];
    out[E_NCON] = fini[2];
  }
}

#define DOS_CHUNK 1024
// Gaussians of levels more than DOS_CUT sigma from every grid point of a
// workgroup are skipped: each is below exp(-50) ~ 2e-22 of its peak.
#define DOS_CUT 10.0
// One thread per grid point, one workgroup per 256 consecutive points. The
// levels of the structure are streamed through in order; those within
// DOS_CUT sigma of the workgroup's energy window are compacted into LDS
// (order kept by a ballot prefix, so each point sums its terms in level
// order) and summed by every thread of the workgroup.
__global__ __launch_bounds__(POST_THREADS) void dos_kernel(const double* __restrict__ W, const double* __restrict__ KW,
                                                           const SlkDosJob* jobs, const double* __restrict__ grid,
                                                           int ne, double sigma, const double* edges, double* dos,
                                                           unsigned long long* clk) {
  sample_start(clk);
  const SlkDosJob jb = jobs[blockIdx.y];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int e = blockIdx.x * blockDim.x + tid;
  const double E = e < ne ? grid[e] : 0.0;
  __shared__ double s_lo[POST_THREADS / 64], s_hi[POST_THREADS / 64];
  __shared__ int s_cnt[POST_THREADS / 64];
  __shared__ double se[DOS_CHUNK];
  __shared__ double sw[DOS_CHUNK];
  {
    const double lo = wave_min(e < ne ? E : 1e300), hi = wave_max(e < ne ? E : -1e300);
    if (lane == 0) { s_lo[wv] = lo; s_hi[wv] = hi; }
  }
  __syncthreads();
  double blo = 1e300, bhi = -1e300;
  for (int q = 0; q < POST_THREADS / 64; ++q) {
    blo = fmin(blo, s_lo[q]);
    bhi = fmax(bhi, s_hi[q]);
  }
  const double wlo = blo - DOS_CUT * sigma, whi = bhi + DOS_CUT * sigma;
  const double mu = jb.edge >= 0 ? edges[(size_t)jb.edge * SLK_EDGE_FIELDS + E_MU] : 0.0;
  const double inv = 1.0 / sigma;
  const double norm = 1.0 / (sigma * sqrt(2.0 * 3.141592653589793));
  const int tot = jb.nk * jb.nb;
  double acc = 0.0;
  int fill = 0;   // kept levels staged in se / sw (the same on every thread)
  for (int c0 = 0; c0 < tot; c0 += POST_THREADS) {
    const int g = c0 + tid;
    double val = 0.0, wk = 0.0;
    bool keep = false;
    if (g < tot) {
      const int k = g / jb.nb, b = g % jb.nb;
      val = (W[jb.w0 + (long long)k * jb.wld + b] - mu) * H2E_OUT;
      keep = val >= wlo && val <= whi;
      if (keep) wk = KW[jb.kw0 + k];
    }
    const unsigned long long mask = __ballot(keep);
    const int pre = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) s_cnt[wv] = __popcll(mask);
    __syncthreads();
    int off = fill, total = 0;
    for (int q = 0; q < POST_THREADS / 64; ++q) {
      if (q < wv) off += s_cnt[q];
      total += s_cnt[q];
    }
    if (keep) {
      se[off + pre] = val;
      sw[off + pre] = wk;
    }
    fill += total;
    __syncthreads();
    if (fill > DOS_CHUNK - POST_THREADS || c0 + POST_THREADS >= tot) {
      for (int q = 0; q < fill; ++q) {
        const double x = (E - se[q]) * inv;
        acc += sw[q] * (exp(-0.5 * x * x) * norm);
      }
      fill = 0;
      __syncthreads();
    }
  }
  if (e < ne) dos[jb.out0 + e] = acc;
  if (clk != nullptr) {
    __syncthreads();
    sample_end(clk);
  }
}

// Adjoint of dos_kernel: for every eigenvalue e_kb of a job,
//   gW[k][b] = dL/de_kb = H2E * w_k * sum_E gdos(E) N exp(-t^2/2) t / sigma,
//   t = (E - (e_kb - mu) * H2E) / sigma,
// i.e. the derivative at fixed mu (the caller chains mu's own derivative).
// One thread per eigenvalue, the grid and the incoming gradient staged in LDS.
__global__ __launch_bounds__(POST_THREADS) void dos_adjoint_kernel(
    const double* __restrict__ W, const double* __restrict__ KW, const SlkDosJob* jobs,
    const double* __restrict__ grid, int ne, double sigma, const double* edges, const double* __restrict__ gdos,
    double* __restrict__ gW) {
  const SlkDosJob jb = jobs[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int tot = jb.nk * jb.nb;
  const double mu = jb.edge >= 0 ? edges[(size_t)jb.edge * SLK_EDGE_FIELDS + E_MU] : 0.0;
  const double inv = 1.0 / sigma;
  const double norm = 1.0 / (sigma * sqrt(2.0 * 3.141592653589793));
  const int k = q < tot ? q / jb.nb : 0, b = q < tot ? q % jb.nb : 0;
  const double x = q < tot ? (W[jb.w0 + (long long)k * jb.wld + b] - mu) * H2E_OUT : 0.0;
  __shared__ double sE[DOS_CHUNK];
  __shared__ double sg[DOS_CHUNK];
  double acc = 0.0;
  for (int c0 = 0; c0 < ne; c0 += DOS_CHUNK) {
    const int cn = min(DOS_CHUNK, ne - c0);
    __syncthreads();
    for (int i = threadIdx.x; i < cn; i += blockDim.x) {
      sE[i] = grid[c0 + i];
      sg[i] = gdos[jb.out0 + c0 + i];
    }
    __syncthreads();
    for (int i = 0; i < cn; ++i) {
      const double t = (sE[i] - x) * inv;
      acc += sg[i] * (exp(-0.5 * t * t) * t);
    }
  }
  if (q < tot) gW[jb.w0 + (long long)k * jb.wld + b] = acc * norm * inv * H2E_OUT * KW[jb.kw0 + k];
}

// Chain rule through the Fermi level of the bisection (N(mu) = Ne with
// N = 2 sum_k w_k sum_b f((e_kb - mu) / kT)): by the implicit function
// theorem dmu/de_kb = w_k f'_kb / sum w f', f' = f (1 - f). An incoming
// gradient g = dL/de at fixed mu plus dL/dmu = -sum g (every consumer here
// depends on e - mu) gives g_kb -= (sum g) dmu/de_kb. One workgroup per
// structure.
__global__ __launch_bounds__(POST_THREADS) void fermi_chain_kernel(const double* __restrict__ W,
                                                                   const double* __restrict__ KW,
                                                                   const SlkSpectrum* spec, double kT,
                                                                   const double* edges, double* gW) {
  const SlkSpectrum sp = spec[blockIdx.x];
  const double mu = edges[(size_t)blockIdx.x * SLK_EDGE_FIELDS + E_MU];
  const int tid = threadIdx.x;
  const int tot = sp.nk * sp.nb;
  __shared__ double red[POST_THREADS / 64];
  double gs = 0.0, ps = 0.0;
  for (int q = tid; q < tot; q += blockDim.x) {
    const int k = q / sp.nb, b = q % sp.nb;
    const long long o = sp.w0 + (long long)k * sp.wld + b;
    const double x = (W[o] - mu) / kT;
    const double ex = exp(-fabs(x));
    const double f1 = ex / ((1.0 + ex) * (1.0 + ex));   // f (1 - f), symmetric in x
    gs += gW[o];
    ps += KW[sp.kw0 + k] * f1;
  }
  const double G = block_sum(gs, red);
  const double P = block_sum(ps, red);
  if (!(P > 0.0)) return;   // every level far from mu: mu does not move with them
  for (int q = tid; q < tot; q += blockDim.x) {
    const int k = q / sp.nb, b = q % sp.nb;
    const long long o = sp.w0 + (long long)k * sp.wld + b;
    const double x = (W[o] - mu) / kT;
    const double ex = exp(-fabs(x));
    const double f1 = ex / ((1.0 + ex) * (1.0 + ex));
    gW[o] -= G * KW[sp.kw0 + k] * f1 / P;
  }
}

// Atom-type projected DOS (after predict_slakonet.py:119-210): per state the
// weight of type t is sum over its orbitals of |psi_mu|^2 (coefficients of the
// S-orthonormal eigenvector, no overlap), Gaussians on E + shift, averaged
// over k with `scale` (1/nk).
#define PDOS_MAXT 8
__global__ __launch_bounds__(POST_THREADS) void pdos_weights_kernel(const double2* __restrict__ V, int n, int ld,
                                                                    const int* __restrict__ orb_type, int n_types,
                                                                    double* __restrict__ wts) {
  const int k = blockIdx.y;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  double acc[PDOS_MAXT];
#pragma unroll
  for (int t = 0; t < PDOS_MAXT; ++t) acc[t] = 0.0;
  const double2* vk = V + (size_t)k * ld * ld;
  for (int mu = 0; mu < n; ++mu) {
    const double2 c = vk[(size_t)mu * ld + b];
    const double p = c.x * c.x + c.y * c.y;
    const int t = orb_type[mu];
#pragma unroll
    for (int q = 0; q < PDOS_MAXT; ++q)
      if (q == t) acc[q] += p;
  }
  for (int t = 0; t < n_types; ++t) wts[((size_t)k * n + b) * n_types + t] = acc[t];
}

__global__ __launch_bounds__(POST_THREADS) void pdos_kernel(const double* __restrict__ W, int n, int wld, int nk,
                                                            const double* __restrict__ wts, int n_types,
                                                            const double* __restrict__ grid, int ne, double sigma,
                                                            double shift_ev, double scale, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const double E = (e < ne ? grid[e] : 0.0) + shift_ev;
  const double inv = 1.0 / sigma;
  const double norm = scale / (sigma * sqrt(2.0 * 3.141592653589793));
  __shared__ double se[DOS_CHUNK / 4];
  __shared__ double sw[DOS_CHUNK / 4][PDOS_MAXT];
  double acc[PDOS_MAXT];
#pragma unroll
  for (int t = 0; t < PDOS_MAXT; ++t) acc[t] = 0.0;
  const int tot = nk * n;
  for (int c0 = 0; c0 < tot; c0 += DOS_CHUNK / 4) {
    const int cn = min(DOS_CHUNK / 4, tot - c0);
    __syncthreads();
    for (int q = threadIdx.x; q < cn; q += blockDim.x) {
      const int g = c0 + q, k = g / n, b = g % n;
      se[q] = W[(size_t)k * wld + b] * H2E_OUT;
      for (int t = 0; t < n_types; ++t) sw[q][t] = wts[(size_t)g * n_types + t];
    }
    __syncthreads();
    for (int q = 0; q < cn; ++q) {
      const double x = (E - se[q]) * inv;
      const double gss = exp(-0.5 * x * x) * norm;
#pragma unroll
      for (int t = 0; t < PDOS_MAXT; ++t)
        if (t < n_types) acc[t] += sw[q][t] * gss;
    }
  }
  if (e < ne)
    for (int t = 0; t < n_types; ++t) out[(size_t)t * ne + e] = acc[t];
}

// dos stats: [dos_at_fermi, gap_center, trapz_total]
__global__ __launch_bounds__(POST_THREADS) void dos_stats_kernel(const SlkDosJob* jobs, const double* __restrict__ grid,
                                                                 int ne, const double* __restrict__ dos,
                                                                 double* stats) {
  const SlkDosJob jb = jobs[blockIdx.x];
  const double* d = dos + jb.out0;
  const int tid = threadIdx.x;
  __shared__ double red[POST_THREADS / 64];
  // argmin |E| (first), argmin dos over |E| < 2 (first), trapz
  double bestA = 1e300; int ia = 1 << 30;
  double bestB = 1e300; int ib = 1 << 30;
  double tr = 0.0;
  for (int i = tid; i < ne; i += blockDim.x) {
    const double a = fabs(grid[i]);
    if (a < bestA || (a == bestA && i < ia)) { bestA = a; ia = i; }
    if (a < 2.0) {
      if (d[i] < bestB || (d[i] == bestB && i < ib)) { bestB = d[i]; ib = i; }
    }
    if (i + 1 < ne) tr += (grid[i + 1] - grid[i]) * (d[i] + d[i + 1]) * 0.5;
  }
  __shared__ double sa[POST_THREADS], sb[POST_THREADS];
  __shared__ int sia[POST_THREADS], sib[POST_THREADS];
  sa[tid] = bestA; sia[tid] = ia; sb[tid] = bestB; sib[tid] = ib;
  const double tot = block_sum(tr, red);
  __syncthreads();
  if (tid == 0) {
    for (int i = 1; i < blockDim.x; ++i) {
      if (sa[i] < bestA || (sa[i] == bestA && sia[i] < ia)) { bestA = sa[i]; ia = sia[i]; }
      if (sb[i] < bestB || (sb[i] == bestB && sib[i] < ib)) { bestB = sb[i]; ib = sib[i]; }
    }
    stats[(size_t)blockIdx.x * 3 + 0] = ia < ne ? d[ia] : 0.0;
    stats[(size_t)blockIdx.x * 3 + 1] = ib < ne ? grid[ib] : 0.0;
    stats[(size_t)blockIdx.x * 3 + 2] = tot;
  }
}

extern "C" {

int slk_fermi_launch(const double* W, const double* KW, const SlkSpectrum* spec, int n_spec, double kT, int mode,
                     double tol, int max_iter, double* edges, hipStream_t stream) {
  if (n_spec <= 0) return 0;
  hipLaunchKernelGGL(fermi_kernel, dim3(n_spec), dim3(POST_THREADS), 0, stream, W, KW, spec, kT, mode, tol,
                     max_iter, edges);
  return (int)hipGetLastError();
}

int slk_edges_launch(const double* W, const SlkSpectrum* spec, int n_spec, double metal_ev, double* edges,
                     hipStream_t stream) {
  if (n_spec <= 0) return 0;
  hipLaunchKernelGGL(edges_kernel, dim3(n_spec), dim3(POST_THREADS), 0, stream, W, spec, metal_ev, edges);
  return (int)hipGetLastError();
}

int slk_dos_launch(const double* W, const double* KW, const SlkDosJob* jobs, int n_jobs, const double* grid, int ne,
                   double sigma, const double* edges, double* dos, double* stats, hipStream_t stream) {
  if (n_jobs <= 0) return 0;
  const dim3 g((ne + POST_THREADS - 1) / POST_THREADS, n_jobs);
  SampleSlot sl;
  if (sampler_enabled() && sampler_take(SLK_SAMPLE_DOS, -1.0, &sl))
    hipExtLaunchKernelGGL(dos_kernel, g, dim3(POST_THREADS), 0, stream, sl.e0, sl.e1, 0, W, KW, jobs, grid, ne, sigma,
                          edges, dos, sl.clk);
  else
    hipLaunchKernelGGL(dos_kernel, g, dim3(POST_THREADS), 0, stream, W, KW, jobs, grid, ne, sigma, edges, dos,
                       (unsigned long long*)nullptr);
  if (stats)
    hipLaunchKernelGGL(dos_stats_kernel, dim3(n_jobs), dim3(POST_THREADS), 0, stream, jobs, grid, ne, dos, stats);
  return (int)hipGetLastError();
}

int slk_fermi_chain_launch(const double* W, const double* KW, const SlkSpectrum* spec, int n_spec, double kT,
                           const double* edges, double* gW, hipStream_t stream) {
  if (n_spec <= 0) return 0;
  hipLaunchKernelGGL(fermi_chain_kernel, dim3(n_spec), dim3(POST_THREADS), 0, stream, W, KW, spec, kT, edges, gW);
  return (int)hipGetLastError();
}

int slk_dos_adjoint_launch(const double* W, const double* KW, const SlkDosJob* jobs, int n_jobs, int max_states,
                           const double* grid, int ne, double sigma, const double* edges, const double* gdos,
                           double* gW, hipStream_t stream) {
  if (n_jobs <= 0 || max_states <= 0) return 0;
  hipLaunchKernelGGL(dos_adjoint_kernel, dim3((max_states + POST_THREADS - 1) / POST_THREADS, n_jobs),
                     dim3(POST_THREADS), 0, stream, W, KW, jobs, grid, ne, sigma, edges, gdos, gW);
  return (int)hipGetLastError();
}

int slk_pdos_launch(const void* V, const double* W, int n, int ld, int nk, int wld, const int* orb_type, int n_types,
                    const double* grid, int ne, double sigma, double shift_ev, double scale, double* wts, double* out,
                    hipStream_t stream) {
  if (n_types < 1 || n_types > PDOS_MAXT) return 1;
  if (n <= 0 || nk <= 0 || ne <= 0) return 0;
  hipLaunchKernelGGL(pdos_weights_kernel, dim3((n + POST_THREADS - 1) / POST_THREADS, nk), dim3(POST_THREADS), 0,
                     stream, reinterpret_cast<const double2*>(V), n, ld, orb_type, n_types, wts);
  hipLaunchKernelGGL(pdos_kernel, dim3((ne + POST_THREADS - 1) / POST_THREADS), dim3(POST_THREADS), 0, stream, W, n,
                     wld, nk, wts, n_types, grid, ne, sigma, shift_ev, scale, out);
  return (int)hipGetLastError();
}

int slk_edge_fields(void) { return SLK_EDGE_FIELDS; }

}  // extern "C"
