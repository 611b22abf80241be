// Per-structure result row shared by the Fermi / band-edge / DOS kernels and
// the band-edge refinement: edges[s * SLK_EDGE_FIELDS + f].
#pragma once

#define H2E_OUT 27.211   // Hartree -> eV for outputs

enum {
  E_MU = 0, E_METAL, E_VBM, E_CBM, E_VBMK, E_CBMK, E_SVBM, E_VBW, E_SCBM, E_CBW, E_MINDG, E_VB2, E_CB2,
  E_NITER, E_NVAL, E_NCON, SLK_EDGE_FIELDS
};
