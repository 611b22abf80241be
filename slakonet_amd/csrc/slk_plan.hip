// Host-side batch planner: lattice translations, column tiles, build work
// items and the packed upload blob for a batch of structures. Host code only
// (no kernels); it replaces the per-structure Python/numpy bookkeeping that
// otherwise serialises a serving loop on the interpreter lock.
//
// Blob sections (each 256-byte aligned, in this order): structs
// (SlkStructDesc), work (SlkBuildWork), atom_pos (double[3]), atom_spec (int),
// atom_orb (int), trans_cart (double[3]), trans_int (int[3]), kfrac
// (double[3]), kw (double).
#include <cmath>
#include <cstring>
#include <vector>

#include "slakonet_amd.h"

namespace {

constexpr int MAXCOL = 64;   // columns per build workgroup
constexpr int KC = 4;        // k-points per build workgroup

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Inverse of a 3x3 row-major matrix (cofactors / determinant).
bool inv3(const double* a, double* r) {
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (det == 0.0 || !std::isfinite(det)) return false;
  const double id = 1.0 / det;
  r[0] = c00 * id;
  r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  r[3] = c01 * id;
  r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  r[6] = c02 * id;
  r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  return true;
}

// Translation range per axis: |T_a| <= floor(cutoff / d_a + spread_a) + 1,
// with 1/d_a the norm of column a of cell^-1 and spread_a the extent of the
// fractional coordinates (any pair closer than the cutoff is covered).
bool trans_range(const double* cell, const double* pos, int na, double cutoff, int* rng) {
  double inv[9];
  if (!inv3(cell, inv)) return false;
  for (int a = 0; a < 3; ++a) {
    const double rn = std::sqrt(inv[a] * inv[a] + inv[3 + a] * inv[3 + a] + inv[6 + a] * inv[6 + a]);
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < na; ++i) {
      const double f = pos[3 * i] * inv[a] + pos[3 * i + 1] * inv[3 + a] + pos[3 * i + 2] * inv[6 + a];
      lo = std::fmin(lo, f);
      hi = std::fmax(hi, f);
    }
    const double spread = na > 0 ? hi - lo : 0.0;
    rng[a] = (int)std::floor(cutoff * rn + spread) + 1;
  }
  return true;
}

struct Tile {
  int j0, j1, c0, ncol;
};

void column_tiles(const int* norb_atom, int na, std::vector<Tile>* tiles) {
  tiles->clear();
  int j0 = 0, c0 = 0, acc = 0, off = 0;
  for (int j = 0; j < na; ++j) {
    const int no = norb_atom[j];
    if (acc + no > MAXCOL && acc > 0) {
      tiles->push_back({j0, j, c0, acc});
      j0 = j;
      c0 = off;
      acc = 0;
    }
    acc += no;
    off += no;
  }
  if (acc > 0) tiles->push_back({j0, na, c0, acc});
}

}  // namespace

extern "C" {

int slk_plan_size(const SlkPlanIn* in, SlkPlanOut* out) {
  std::memset(out, 0, sizeof(*out));
  long long atoms = 0, kpts = 0, trans = 0, work = 0, rows = 0;
  int maxn = 0;
  std::vector<int> norb;
  std::vector<Tile> tiles;
  bool rows_ok = true;
  for (int s = 0; s < in->n_struct; ++s) {
    const int na = in->n_atoms[s];
    const int nk = in->nk[s];
    if (na <= 0 || nk <= 0) return 1;
    norb.resize(na);
    int n = 0;
    for (int i = 0; i < na; ++i) {
      const int sp = in->spec[atoms + i];
      if (sp < 0 || sp >= in->nspec) return 1;
      norb[i] = in->spec_norb[sp];
      n += norb[i];
    }
    int rng[3];
    if (!trans_range(in->cell + 9 * s, in->pos + 3 * atoms, na, in->cutoff, rng)) return 1;
    trans += (long long)(2 * rng[0] + 1) * (2 * rng[1] + 1) * (2 * rng[2] + 1);
    column_tiles(norb.data(), na, &tiles);
    work += (long long)na * (long long)tiles.size() * ((nk + KC - 1) / KC);
    rows += (long long)na * ((nk + KC - 1) / KC);
    // a row panel screens na * nt candidates with 32-bit indices
    if ((long long)na * (2 * rng[0] + 1) * (2 * rng[1] + 1) * (2 * rng[2] + 1) >= (1LL << 30)) rows_ok = false;
    maxn = n > maxn ? n : maxn;
    atoms += na;
    kpts += nk;
  }
  const int ld = in->ld > 0 ? in->ld : maxn;
  if (ld < maxn) return 2;
  if (!rows_ok) rows = 0;
  if (trans >= (1LL << 31) || work + rows >= (1LL << 31)) return 1;
  const size_t sizes[SLK_PLAN_SECTIONS] = {
      (size_t)in->n_struct * sizeof(SlkStructDesc), (size_t)(work + rows) * sizeof(SlkBuildWork),
      (size_t)atoms * 3 * sizeof(double),           (size_t)atoms * sizeof(int),
      (size_t)atoms * sizeof(int),                  (size_t)trans * 3 * sizeof(double),
      (size_t)trans * 3 * sizeof(int),              (size_t)kpts * 3 * sizeof(double),
      (size_t)kpts * sizeof(double)};
  size_t off = 0;
  for (int q = 0; q < SLK_PLAN_SECTIONS; ++q) {
    out->off[q] = (long long)off;
    out->bytes[q] = (long long)sizes[q];
    off += align256(sizes[q]);
  }
  out->total = (long long)off;
  out->n_work = (int)work;
  out->n_row_work = (int)rows;
  out->n_trans = (int)trans;
  out->n_mat = (int)kpts;
  out->ld = ld;
  return 0;
}

int slk_plan_fill(const SlkPlanIn* in, const SlkPlanOut* out, void* blob, int* n_orb, long long* mat0) {
  char* b = reinterpret_cast<char*>(blob);
  std::memset(b, 0, (size_t)out->total);
  SlkStructDesc* sd = reinterpret_cast<SlkStructDesc*>(b + out->off[0]);
  SlkBuildWork* wk = reinterpret_cast<SlkBuildWork*>(b + out->off[1]);
  double* apos = reinterpret_cast<double*>(b + out->off[2]);
  int* aspec = reinterpret_cast<int*>(b + out->off[3]);
  int* aorb = reinterpret_cast<int*>(b + out->off[4]);
  double* tcart = reinterpret_cast<double*>(b + out->off[5]);
  int* tint = reinterpret_cast<int*>(b + out->off[6]);
  double* kf = reinterpret_cast<double*>(b + out->off[7]);
  double* kw = reinterpret_cast<double*>(b + out->off[8]);
  long long atom0 = 0, t0 = 0, k0 = 0, w0 = 0, r0 = out->n_work;
  std::vector<int> norb;
  std::vector<Tile> tiles;
  for (int s = 0; s < in->n_struct; ++s) {
    const int na = in->n_atoms[s];
    const int nk = in->nk[s];
    const double* cell = in->cell + 9 * s;
    const double* pos = in->pos + 3 * atom0;
    norb.resize(na);
    int n = 0;
    for (int i = 0; i < na; ++i) {
      const int sp = in->spec[atom0 + i];
      norb[i] = in->spec_norb[sp];
      aspec[atom0 + i] = sp;
      aorb[atom0 + i] = n;
      for (int c = 0; c < 3; ++c) apos[3 * (atom0 + i) + c] = pos[3 * i + c];
      n += norb[i];
    }
    int rng[3];
    if (!trans_range(cell, pos, na, in->cutoff, rng)) return 1;
    long long nt = 0;
    for (int x = -rng[0]; x <= rng[0]; ++x)
      for (int y = -rng[1]; y <= rng[1]; ++y)
        for (int z = -rng[2]; z <= rng[2]; ++z) {
          const long long t = t0 + nt++;
          tint[3 * t] = x;
          tint[3 * t + 1] = y;
          tint[3 * t + 2] = z;
          for (int c = 0; c < 3; ++c) tcart[3 * t + c] = x * cell[c] + y * cell[3 + c] + z * cell[6 + c];
        }
    for (int q = 0; q < nk; ++q) {
      for (int c = 0; c < 3; ++c) kf[3 * (k0 + q) + c] = in->kfrac[3 * (k0 + q) + c];
      kw[k0 + q] = in->kw[k0 + q];
    }
    column_tiles(norb.data(), na, &tiles);
    const int nkc = (nk + KC - 1) / KC;
    for (int i = 0; i < na; ++i)
      for (size_t t = 0; t < tiles.size(); ++t)
        for (int kc = 0; kc < nkc; ++kc) {
          SlkBuildWork& w = wk[w0++];
          w.s = s;
          w.i = i;
          w.j0 = tiles[t].j0;
          w.j1 = tiles[t].j1;
          w.c0 = tiles[t].c0;
          w.ncol = tiles[t].ncol;
          w.kc0 = kc * KC;
          w.nkc = nk - kc * KC < KC ? nk - kc * KC : KC;
        }
    if (out->n_row_work > 0)
      for (int i = 0; i < na; ++i)
        for (int kc = 0; kc < nkc; ++kc) {
          SlkBuildWork& w = wk[r0++];
          w.s = s;
          w.i = i;
          w.j0 = 0;
          w.j1 = na;
          w.c0 = 0;
          w.ncol = n;
          w.kc0 = kc * KC;
          w.nkc = nk - kc * KC < KC ? nk - kc * KC : KC;
        }
    SlkStructDesc& d = sd[s];
    d.atom0 = (int)atom0;
    d.n_atoms = na;
    d.n_orb = n;
    d.ld = out->ld;
    d.mat0 = k0;
    d.k0 = (int)k0;
    d.nk = nk;
    d.t0 = (int)t0;
    d.nt = (int)nt;
    n_orb[s] = n;
    mat0[s] = k0;
    atom0 += na;
    t0 += nt;
    k0 += nk;
  }
  return (w0 == out->n_work && t0 == out->n_trans && r0 == (long long)out->n_work + out->n_row_work) ? 0 : 1;
}

}  // extern "C"
