// Internal host launchers of the blocked dense kernels (slk_dense.hip).
#pragma once
#include <hip/hip_runtime.h>

// op flags for GEMM operands: bit0 transpose, bit1 conjugate.
enum { OP_N = 0, OP_T = 1, OP_CONJ = 2, OP_C = 3 };
// K-range limits for triangular operands (element (m,k) of opA, (k,n) of opB)
enum { KLIM_A_LOWER = 1, KLIM_A_UPPER = 2, KLIM_B_UPPER = 4, KLIM_B_LOWER = 8 };

// C = alpha opA(A) opB(B) + beta C for `batch` problems with strides sA/sB/sC.
// tri = 1 skips 64x64 output tiles strictly above the diagonal. re: A and B
// are real (zero imaginary parts); C gets zero imaginary parts, or with
// c_real the real values alone, element (i, j) of problem b at real index
// i * ldc + j of the complex slot C + b * sC (half the bytes).
void launch_zgemm(hipStream_t st, int batch, int M, int N, int K, double alpha, const double2* A, int lda,
                  long long sA, int opA, const double2* B, int ldb, long long sB, int opB, double beta, double2* C,
                  int ldc, long long sC, int klim, int tri, bool re = false, bool c_real = false);

void launch_cgemm(hipStream_t st, int batch, int M, int N, int K, float alpha, const float2* A, int lda,
                  long long sA, int opA, const float2* B, int ldb, long long sB, int opB, float beta, float2* C,
                  int ldc, long long sC, int klim, int tri, bool re = false, bool c_real = false);

// Zero-fill (no memset nodes so that the solve can be graph-captured).
void launch_zero(hipStream_t st, void* p, size_t bytes);

// Cholesky S = L L^H (into L, a copy of S; lower part) and W = L^-1 together
// (recursive halving, large GEMMs); info[b] = first failing pivot (1-based)
// or 0. T: temporary of ld * ld / 2 complex128 per matrix. re: S is real
// (zero imaginary parts), and so are L and W.
void launch_chol_inv(hipStream_t st, const double2* S, double2* L, double2* W, double2* T, int n, int ld, int batch,
                     int* info, bool re = false);

// Blocked Householder tridiagonalisation of the Hermitian C (in place): d, e
// (real tridiagonal), tau, reflectors in the strict lower part of C.
// Workspace: Vc, Wc [NB][ld]; Y [batch][2 ld] (y, then the next column);
// Yp [hetrd_partials_elems(ld)] and C12 [hetrd_c12_elems(ld)] per matrix. D, E are FP64 for both precisions. VT (FP32, nullable):
// reflector j is also written as row j of VT (contiguous, for single-vector
// back-transforms).
void launch_hetrd_blocked(hipStream_t st, double2* C, int n, int ld, int batch, double* D, double* E, double2* TAU,
                          double2* Vc, double2* Wc, double2* Y, double2* Yp, double2* C12);
void launch_hetrd_blocked_f32(hipStream_t st, float2* C, int n, int ld, int batch, double* D, double* E,
                              float2* TAU, float2* Vc, float2* Wc, float2* Y, float2* Yp, float2* C12,
                              float2* VT);
size_t hetrd_partials_elems(int ld);
size_t hetrd_c12_elems(int ld);

// complex128 -> complex64 copy of `count` elements.
void launch_to_c64(hipStream_t st, const double2* src, float2* dst, size_t count);

constexpr int SLK_NB = 32;   // panel width
constexpr int SLK_BW = 16;   // band width of the two-stage reduction

// Two-stage reduction of the Hermitian C (in place) for the eigenvalue path:
// dense -> band of width SLK_BW (16-column panels, block reflectors, trailing
// updates on the MFMA GEMM), then band -> real tridiagonal d, e by bulge
// chasing (one workgroup per matrix). Stage-1 reflectors: tau in TAU[j],
// reflector of column j as row j of VT (entries r >= j + SLK_BW, explicit 1
// at j + SLK_BW; VT nullable). Chase reflectors: row s of C's strict upper
// triangle holds sweep s's reflectors, tau in place of each leading 1.
// Workspace: Vc, Wc [2 SLK_BW][ld] and Yp [band_partials_elems(ld)] per
// matrix, band scratch [band_scratch_elems(ld)] per matrix.
// info (nullable, [batch]): set to 2 for a matrix whose bulge chase gave up a bounded wait.
// re: C is real symmetric, given as its real values alone (element (i, j) of
// matrix b at real index i * ld + j of the complex slot C + b * ld * ld, as
// the GEMMs write it with c_real); every output keeps zero imaginary parts.
void launch_band_tridiag(hipStream_t st, double2* C, int n, int ld, int batch, double* D, double* E, double2* TAU,
                         double2* Vc, double2* Wc, double2* Yp, double2* scratch, double2* VT, int* info,
                         bool re = false);
void launch_band_tridiag_f32(hipStream_t st, float2* C, int n, int ld, int batch, double* D, double* E, float2* TAU,
                             float2* Vc, float2* Wc, float2* Yp, float2* scratch, float2* VT, int* info,
                             bool re = false);
size_t band_partials_elems(int ld);
size_t band_scratch_elems(int ld);
// Panels of the band reduction per matrix, and the element offset inside the
// band scratch (of `batch` matrices) of the panels' block-reflector factors
// T ([batch][band_panels(ld)][SLK_BW][SLK_BW], row-major, Q_p = I - V_p T_p V_p^H).
int band_panels(int ld);
size_t band_tt_offset(int ld, int batch);
