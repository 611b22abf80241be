/* C-ABI of the SlaKoNet MI355X hot path (libslakonet_amd.so, gfx950).
 *
 * Plain device pointers, sizes and a HIP stream; no framework types. All
 * calls are stream-ordered and allocate nothing; callers own every buffer
 * (including the eigensolver workspace sized by slk_eigh_workspace_bytes).
 * Return value: 0 on success, a hipError_t value on launch failure, 2 for an
 * undersized workspace. Numerical failures are reported per matrix in `info`.
 *
 * The reference (atomgptlab/slakonet) has no FFI; each entry point below
 * replaces the Python operator cited beside it, and the ctypes binding in
 * slakonet_amd/_lib.py is what the reference-side integration would call.
 */
#ifndef SLAKONET_AMD_H
#define SLAKONET_AMD_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One structure inside a build batch. Its k-matrices are consecutive in the
 * output starting at matrix index mat0; each matrix is ld x ld row-major. */
typedef struct SlkStructDesc {
  int atom0, n_atoms, n_orb, ld;
  long long mat0;
  int k0, nk, t0, nt;
} SlkStructDesc;

/* One workgroup of the build: rows of atom i, columns of atoms [j0, j1)
 * (ncol <= 64 columns starting at orbital c0), k-points [kc0, kc0 + nkc). */
typedef struct SlkBuildWork {
  int s, i, j0, j1, c0, ncol, kc0, nkc;
} SlkBuildWork;

/* Device-resident SK tables: rows of 20 doubles (10 H then 10 S integrals in
 * the order ss-sigma, sp-sigma, sd-sigma, pp-sigma, pp-pi, pd-sigma, pd-pi,
 * dd-sigma, dd-pi, dd-delta); row g of a pair holds r = (g + 1) * step. */
typedef struct SlkTableSet {
  const double* tables;
  const int* pair_row0;
  const int* pair_ngrid;
  const double* pair_step;
  const int* pair_index;    /* [nspec * nspec] ordered pair -> pair id, -1 if absent */
  const int* spec_norb;     /* orbitals per species */
  const int* spec_slot;     /* [nspec * 9] local orbital -> spd slot (s,py,pz,px,dxy,dyz,dz2,dxz,dx2-y2) */
  const double* spec_onsite;/* [nspec * 3] eps_s, eps_p, eps_d (Hartree) */
  int nspec;
  double tail;              /* 1.0 Bohr, PolyInterpU default */
} SlkTableSet;

typedef struct SlkBuildArgs {
  SlkTableSet tab;
  const SlkStructDesc* structs;
  const SlkBuildWork* work;
  int n_work;
  const double* atom_pos;   /* [atoms][3] Bohr */
  const int* atom_spec;
  const int* atom_orb;
  const double* trans_cart; /* [T][3] Bohr */
  const int* trans_int;     /* [T][3] */
  const double* kfrac;      /* [K][3] fractional k */
  double cutoff;            /* 10.0 Bohr */
  void* H;
  void* S;
  int out_f32;              /* 0: complex128, 1: complex64 */
  /* Whole-matrix form (n_mat > 0): the build writes every element of the
   * batch's n_mat matrices of dimension ld -- a bandwidth-bound fill (zeros,
   * the on-site diagonal, identity padding of rows/columns >= n_orb with
   * H = pad_h on the padded diagonal and S = 1), then the blocks of the atom
   * pairs that have a pair-image within the cutoff, one workgroup per row
   * panel (the n_row_work items that follow the n_work column tiles in work:
   * one per structure, atom i and k chunk). n_mat = 0: one workgroup per
   * column tile writes the n_orb x n_orb part and the padding is left to
   * slk_pad_launch. */
  double pad_h;
  int n_struct, n_mat, ld; /* batch shape: structs entries, matrices, padded dimension */
  int n_row_work;
} SlkBuildArgs;

/* Host-side batch planner (no GPU needed). Replaces the per-structure image
 * and index bookkeeping of the reference (Periodic / get_cell_translations,
 * slakonet/atoms.py:48-243, and the basis index matrices of
 * slakonet/basis.py:66-153) with one pass over a batch: lattice translations
 * per structure (|T_a| <= floor(cutoff/d_a + spread_a) + 1, x slowest),
 * column tiles of <= 64 orbitals, build work items, all packed into one blob
 * for a single host-to-device copy. */
#define SLK_PLAN_SECTIONS 9   /* structs, work, atom_pos, atom_spec, atom_orb, trans_cart, trans_int, kfrac, kw */
typedef struct SlkPlanIn {
  int n_struct;
  const int* n_atoms;       /* [n_struct] */
  const int* spec;          /* [sum atoms] species index into spec_norb */
  const double* pos;        /* [sum atoms][3] Bohr */
  const double* cell;       /* [n_struct][3][3] Bohr, rows are lattice vectors */
  const int* nk;            /* [n_struct] */
  const double* kfrac;      /* [sum nk][3] */
  const double* kw;         /* [sum nk] */
  const int* spec_norb;     /* [nspec] */
  int nspec;
  double cutoff;            /* Bohr */
  int ld;                   /* padded dimension, 0: the largest n_orb */
} SlkPlanIn;
typedef struct SlkPlanOut {
  long long off[SLK_PLAN_SECTIONS];    /* byte offset of each section (256-aligned) */
  long long bytes[SLK_PLAN_SECTIONS];  /* used bytes of each section */
  long long total;
  int n_work, n_trans, n_mat, ld;
  int n_row_work;           /* row-panel items after the n_work column tiles in the work section */
} SlkPlanOut;
/* 0 on success, 1 bad input (empty structure, unknown species, singular
 * cell), 2 ld smaller than the largest structure. */
int slk_plan_size(const SlkPlanIn* in, SlkPlanOut* out);
/* Fills blob (out->total bytes); n_orb [n_struct], mat0 [n_struct]. */
int slk_plan_fill(const SlkPlanIn* in, const SlkPlanOut* out, void* blob, int* n_orb, long long* mat0);

/* Fused SK interpolation + rotation + Bloch sum + on-site into H(k), S(k).
 * Replaces hs_matrix(geometry, basis, sk_feed, cutoff=10.0)
 * (slakonet/slaterkoster.py:103-441), SkfFeed.off_site/on_site
 * (slakonet/skfeed.py:285-359) and PolyInterpU.__call__
 * (slakonet/interpolation.py:107-238). */
int slk_build_hs_launch(const SlkBuildArgs* args, hipStream_t stream);
/* Blocks of the whole-matrix form of slk_build_hs_launch: 0 (default) by
 * row panels for ld >= 448 and by column tiles below, 1 by column tiles
 * (every tile, zero blocks included; the fill then writes the padding
 * only). Both give the same bits. Process-wide; set
 * it before building, not while builds are in flight. Returns 1 for an
 * unknown mode. No reference counterpart (a measurement and test switch). */
int slk_set_build_mode(int mode);

/* Identity-padding of rows/columns [n_orb, ld) of a structure's matrices
 * (S = 1, H = pad_h on the padded diagonal). No reference counterpart: used to
 * bucket structures of uneven size into one batched solve. */
int slk_pad_launch(const SlkStructDesc* structs_dev, int s, int nk, int n_orb, int ld, void* H, void* S,
                   double pad_h, int out_f32, hipStream_t stream);

/* Generalised Hermitian eigenproblem H z = e S z for a batch of matrices.
 * Replaces eighb(h, s, scheme="chol") (slakonet/utils.py:585-816) and the
 * per-k loop of SimpleDftb.__call__ (slakonet/main.py:165-196).
 * w: [batch][wld] ascending; V (nullable): eigenvectors in columns;
 * info: [2 * batch], cleared by the call; info[b] > 0: S not positive definite at that pivot,
 * info[batch + b] = 2: the bulge chase of the two-stage tridiagonalisation
 * gave up a bounded wait (eigenvalues NaN). Eigenvalues come from bisection
 * on the FP64 tridiagonal; eigenvectors (V non-null) from inverse iteration on
 * it (LAPACK dstein method, Gram-Schmidt inside clusters of close
 * eigenvalues), then the reflectors and W^H. Any n <= 16384. */
size_t slk_eigh_workspace_bytes(int n, int ld, int batch, int vectors);
int slk_eigh_f64(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                 void* ws, size_t ws_bytes, int* info, hipStream_t stream);
/* Precision form, eigenvalues only: Cholesky of S and L^-1 in FP64, the
 * reduction C = L^-1 H L^-H (MFMA GEMMs) and the Householder
 * tridiagonalisation in FP32, Sturm bisection of T in FP64. Same arguments as
 * slk_eigh_f64 without V; workspace from slk_eigh_mixed_workspace_bytes. */
size_t slk_eigh_mixed_workspace_bytes(int n, int ld, int batch);
int slk_eigh_mixed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* ws,
                   size_t ws_bytes, int* info, hipStream_t stream);
int slk_eigh_mixed_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* ws,
                         size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms);
/* Both solves behind one entry point with flags: SLK_EIGH_MIXED selects the
 * precision form (V must then be null), SLK_EIGH_REAL declares that every H
 * and S of the batch is real symmetric, stored as complex with zero imaginary
 * parts (k at a point where every Bloch phase is +-1, e.g. Gamma or the
 * Monkhorst-Pack point (1/2, 1/2, 1/2) of a 1x1x1 mesh): the kernels then
 * skip the imaginary arithmetic. Results are those of the complex solve up to
 * rounding; with imaginary parts present under SLK_EIGH_REAL they are
 * undefined. stage_ms (nullable, 6 floats): per-stage times as in the _timed
 * variants (the call then waits for the solve). Returns 1 for unknown flags.
 * Replaces the same reference operators as slk_eigh_f64 (eighb,
 * slakonet/utils.py:585-816); the reference has no real-symmetric case. */
#define SLK_EIGH_MIXED 1
#define SLK_EIGH_REAL 2
int slk_eigh_ex(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V, void* ws,
                size_t ws_bytes, int* info, int flags, float* stage_ms, hipStream_t stream);
/* slk_eigh_mixed without a workspace argument and without any allocation:
 * the workspace is a static device array of slk_noalloc_scratch_bytes()
 * (64 MiB) per GPU, held from the library's load until the process ends.
 * Returns 4 when slk_eigh_mixed_workspace_bytes(n, ld, batch) exceeds it
 * (e.g. one 1024 x 1024 matrix or three 576 x 576 fit). Calls that overlap
 * on one GPU share the array and race: the caller serialises them. info is
 * cleared by the call. No counterpart in the reference; it is for callers
 * that cannot allocate device memory (HIP graph capture, allocation-free
 * serving loops). */
size_t slk_noalloc_scratch_bytes(void);
int slk_eigh_mixed_noalloc(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, int* info,
                           hipStream_t stream);
/* Tridiagonalisation used by the eigenvalue-only solves (slk_eigh_f64 with
 * V null, slk_eigh_mixed, and the matching slk_refine_edges_mixed): 0
 * automatic (two-stage: dense -> band of width 16 -> tridiagonal by bulge
 * chasing), 1 blocked one-stage Householder, 2 two-stage. Process-wide; set it
 * before solving, not while solves are in flight. Replaces the
 * tridiagonalisation inside torch.linalg.eigh called by eighb
 * (slakonet/utils.py:585-816); the reference exposes no such switch.
 * Returns 1 for an unknown mode. */
int slk_set_tridiag_mode(int mode);
/* Polls per wait of the bulge chase (two-stage tridiagonalisation) before a
 * wave gives up; a negative value gives up at the first wait. A matrix whose
 * chase gave up gets info[batch + b] = 2 and NaN eigenvalues; the kernel still
 * runs to its end. Process-wide; returns the previous limit (default 2^22).
 * No reference counterpart (the reference's eigensolver has no bounded wait);
 * it lets a caller or test exercise the failure path without a hang. */
int slk_set_chase_spin_limit(int polls);
/* Eigenvalues only for n <= 32: the whole solve of one matrix in LDS of one
 * wave64 workgroup, one launch per batch (slk_eigh_f64 dispatches here for
 * values-only solves of that size). info: [batch]. */
int slk_eigh_small_f64(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, int* info,
                       hipStream_t stream);
/* Same solve, plus elapsed milliseconds of its six stages (potrf, trinv,
 * reduce, hetrd, tridiagonal solve, back-transform) from HIP events recorded
 * on `stream`; returns after the last event completed. Measurement aid with
 * no reference counterpart. */
int slk_eigh_f64_timed(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, void* V,
                       void* ws, size_t ws_bytes, int* info, hipStream_t stream, float* stage_ms);

/* Batched complex GEMM on MFMA (complex128 / complex64, row-major):
 * C = alpha opA(A) opB(B) + beta C, op bit0 = transpose, bit1 = conjugate;
 * sA/sB/sC are batch strides in elements. Building block of the solver (no
 * reference counterpart), exposed for verification and reuse. */
int slk_zgemm(int batch, int M, int N, int K, double alpha, const void* A, int lda, long long sA, int opA,
              const void* B, int ldb, long long sB, int opB, double beta, void* C, int ldc, long long sC,
              hipStream_t stream);
int slk_cgemm(int batch, int M, int N, int K, float alpha, const void* A, int lda, long long sA, int opA,
              const void* B, int ldb, long long sB, int opB, float beta, void* C, int ldc, long long sC,
              hipStream_t stream);

/* Eigenvalue spectrum of one structure for the reductions below. */
typedef struct SlkSpectrum {
  long long w0;
  int nk, nb, wld;
  int kw0;
  double nelec;
} SlkSpectrum;

/* Fermi level per structure into edges[s * slk_edge_fields() + 0].
 * mode 0: bisection root of N(mu) = Ne; mode 1: the reference's half-step
 * Newton with its stopping rule. Replaces fermi_search
 * (slakonet/fermi.py:139-198). */
int slk_fermi_launch(const double* W, const double* KW, const SlkSpectrum* spec, int n_spec, double kT, int mode,
                     double tol, int max_iter, double* edges, hipStream_t stream);

/* Band edges / gap statistics given the Fermi level. Replaces
 * calculate_band_gap (slakonet/main.py:637-707),
 * calculate_band_structure_properties (slakonet/main.py:1539-1642) and the
 * get_fermi gap block (slakonet/optim.py:1011-1059). */
int slk_edges_launch(const double* W, const SlkSpectrum* spec, int n_spec, double metal_ev, double* edges,
                     hipStream_t stream);
int slk_edge_fields(void);

typedef struct SlkDosJob {
  long long w0;
  int nk, nb, wld, kw0;
  int edge;        /* structure index into edges for the Fermi shift, -1: none */
  long long out0;  /* offset of this structure's DOS row */
} SlkDosJob;

/* Gaussian DOS on a caller-supplied energy grid (eV) and its statistics
 * [dos_at_fermi, dos_gap_center, trapz total]. Replaces calculate_dos
 * (slakonet/main.py:709-797) and the DOS block of get_properties_dict
 * (slakonet/main.py:1735-1801). */
int slk_dos_launch(const double* W, const double* KW, const SlkDosJob* jobs, int n_jobs, const double* grid, int ne,
                   double sigma, const double* edges, double* dos, double* stats, hipStream_t stream);

/* Precision form, band-edge refinement (after slk_eigh_mixed, slk_fermi_launch
 * and slk_edges_launch on the same batch): for each structure's VBM and CBM
 * state, an FP64 Rayleigh quotient z^H H z / z^H S z with the FP64 H, S, where
 * z is the state's eigenvector from inverse iteration on the FP64 tridiagonal
 * T carried back through the FP32 reflectors and W = L^-1 (read from the
 * mixed solve's workspace `ws`, untouched since). Writes the refined energies
 * into w and into the edge rows (VBM/CBM, shifted and Hartree edges).
 * scratch: slk_refine_scratch_bytes(ld, n_spec). Returns 0 without work for
 * n <= 32 (solved in FP64). Replaces nothing in the reference, whose solve
 * is FP64 throughout (slakonet/utils.py:740-783); this is how the FP32 bulk
 * keeps band edges at FP64 accuracy. */
size_t slk_refine_scratch_bytes(int ld, int n_spec);
int slk_refine_edges_mixed(const void* H, const void* S, int n, int ld, int batch, const void* ws, size_t ws_bytes,
                           double* w, int wld, const SlkSpectrum* spec, int n_spec, double* edges, double* scratch,
                           size_t scratch_bytes, hipStream_t stream);

/* ---- Measurement --------------------------------------------------------- */

/* Live kernel sampling, no reference counterpart (the reference has no
 * tracing, only time.time() prints). While enabled, every launch of a
 * sampled kernel kind is counted with its algorithmic bytes, and every
 * `every`-th one (at most `capacity`) is started with a HIP start/stop event
 * pair that the runtime records at the kernel's own begin and end on its
 * launch stream, and records the launch's execution window (earliest
 * workgroup start, latest workgroup end on the GPU's constant-rate wall
 * clock). Kinds and their algorithmic amount: 0 = one-stage
 * tridiagonalisation mat-vec (bytes: one triangle of the trailing Hermitian
 * block per matrix), 1 = H/S build (bytes from slk_sampler_hint, per calling
 * thread: the H and S elements written), 2 = FP64 MFMA GEMM (flops of the
 * executed output tiles, 64x64 or 32x32; a real product counts a quarter),
 * 3 = band chase of complex Hermitian batches (flops of the steps of one
 * matrix, i.e. per workgroup), 4 = FP32 fused stage-1 update and next A22 V
 * of complex Hermitian batches (flops: 8 (m^2/2) 32 + 8 m^2 16 per matrix of
 * trailing size m), 5 = DOS (bytes from slk_sampler_hint: eigenvalues,
 * k-weights, grid read and DOS written; auxiliary amount: Gaussians of the
 * full sum, nk nb n_E), 6 = the fused update of real-symmetric batches
 * (SLK_EIGH_REAL; 2 (m^2/2) 32 + 2 m^2 16 flops per matrix), 7 = the chase of
 * real-symmetric batches (a quarter of the complex step flops).
 * enable = 0 disables and frees. Thread-safe. */
int slk_sampler_reset(int enable, int every, int capacity);
/* Amount (and auxiliary amount) of the calling thread's next launch of a
 * hinted kind (1, 5). */
int slk_sampler_hint(int kind, double bytes, double aux);
/* out[9]: out[0] = sampled launches, out[1] = their summed execution-window
 * milliseconds, out[2] = their summed algorithmic amount, out[3] = launches
 * counted, out[4] = amount of all counted launches, out[5] = summed event-pair
 * milliseconds of the sampled launches, out[6] = wall clock rate (kHz),
 * out[7] = summed auxiliary amount of the sampled launches, out[8] = of all
 * counted launches; waits for the sampled launches to complete. */
int slk_sampler_read(int kind, double* out);

/* Atom-type projected DOS from the eigenvectors of one structure (V: the
 * [nk][ld][ld] complex128 output of slk_eigh_f64, eigenvectors in columns;
 * W: eigenvalues [nk][wld], Hartree). Weight of type t for state (k, b):
 * sum of |V[k][mu][b]|^2 over orbitals mu with orb_type[mu] == t (no
 * overlap); out[t][e] = scale * sum_kb weight * Gaussian(grid[e] + shift_ev -
 * eps_kb * 27.211, sigma). n_types <= 8; wts: nk * n * n_types doubles of
 * scratch. Replaces compute_atom_projected_dos
 * (slakonet/predict_slakonet.py:119-210; calculate_projected_dos,
 * slakonet/main.py:799-840, is a placeholder in the reference). */
int slk_pdos_launch(const void* V, const double* W, int n, int ld, int nk, int wld, const int* orb_type, int n_types,
                    const double* grid, int ne, double sigma, double shift_ev, double scale, double* wts, double* out,
                    hipStream_t stream);

/* ---- Differentiation form (hand-written adjoints) ----------------------- */

/* Adjoint of slk_dos_launch at fixed Fermi level: gW[w0 + k*wld + b] =
 * dL/de_kb given gdos = dL/d dos (same layout as dos). Overwrites the
 * nk x nb entries of every job; max_states >= every job's nk*nb. Replaces
 * the autograd graph of calculate_dos (slakonet/main.py:709-797). */
int slk_dos_adjoint_launch(const double* W, const double* KW, const SlkDosJob* jobs, int n_jobs, int max_states,
                           const double* grid, int ne, double sigma, const double* edges, const double* gdos,
                           double* gW, hipStream_t stream);

/* Chain rule through the bisection Fermi level of slk_fermi_launch (mode 0):
 * gW (dL/de at fixed mu, layout of W) += (dL/dmu) dmu/de with dL/dmu =
 * -sum gW and dmu/de_kb = w_k f'_kb / sum w f' (implicit function theorem on
 * N(mu) = Ne, f' = f (1 - f) at kT). For losses of e - mu such as the
 * Fermi-shifted DOS. No reference counterpart: the reference's fermi_search
 * returns a detached leaf (slakonet/fermi.py:198), so its DOS gradient holds
 * mu fixed; this is the derivative of the loss as evaluated. */
int slk_fermi_chain_launch(const double* W, const double* KW, const SlkSpectrum* spec, int n_spec, double kT,
                           const double* edges, double* gW, hipStream_t stream);

/* Y[b] = C[b] diag(g[b]) (complex128 C [batch, ld, ld] row-major, g real
 * with row stride gld; columns >= n get 0). With Y C^H on slk_zgemm this
 * forms the weighted density Q = C diag(g) C^H of the eigenvalue adjoint
 * de = c^H (dH - e dS) c, which replaces _SymEigB.backward
 * (slakonet/utils.py:386-473) for eigenvalue gradients. */
int slk_scale_columns(const void* C, const double* g, int n, int ld, int batch, int gld, void* Y, hipStream_t stream);

/* Adjoint of slk_build_hs_launch with respect to the table values:
 * dtab[row*20 + q] += sum over k and pair-images of Re(Q(k)[nu][mu] e^{2 pi i k.T})
 * times the Slater-Koster coefficient and interpolation weight that carried
 * table entry (row, q) into H(k)[mu][nu] (QH) or S(k)[mu][nu] (QS).
 * QH, QS: complex128 [n_mat, ld, ld] as H/S. Work items sum over their k
 * range [kc0, kc0 + nkc). dtab must be zeroed by the caller; FP64 atomic
 * accumulation (order-dependent last bits). Replaces the autograd path
 * through hs_matrix / SKF interpolation (slakonet/skfeed.py,
 * slakonet/interpolation.py) used by the training loop
 * (slakonet/optim.py:1570-1790). */
int slk_adjoint_tables_launch(const SlkBuildArgs* args, const void* QH, const void* QS, double* dtab,
                              hipStream_t stream);

/* Same contraction with respect to the atom positions: dR[atom][c] +=
 * dL/dR_atom,c (Bohr), exact derivatives of the interpolated integrals and
 * the direction cosines (dual numbers). With QH, QS from the energy weights
 * dE/de = w_k (f + e f') this is the gradient of the band energy; forces are
 * its negative. dR [atoms][3] zeroed by the caller; FP64 atomics. Replaces the
 * autograd force path of SimpleDftb.get_forces (slakonet/main.py:400-585) and
 * its finite-difference fallback (slakonet/main.py:198-281). */
int slk_forces_launch(const SlkBuildArgs* args, const void* QH, const void* QS, double* dR, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SLAKONET_AMD_H */
