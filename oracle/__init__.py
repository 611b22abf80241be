"""FP64 CPU oracle for the SlaKoNet hot path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithm (atomgptlab/slakonet) in numpy
so that the HIP path can be checked against it. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or as the timed CPU baseline. The product
package ``slakonet_amd`` never imports it and has no CPU fallback.

Modules and the reference code they follow:

* ``interp``  -- PolyInterpU / poly_interp / poly5_zero
                 (slakonet/interpolation.py:107-311)
* ``skrot``   -- sub_block_ref / sub_block_rot / _rot_{s,p,d}_{yz,xy}
                 (slakonet/slaterkoster.py:1013-1372)
* ``build``   -- hs_matrix + Periodic images + phases
                 (slakonet/slaterkoster.py:103-441, slakonet/atoms.py:165-558)
* ``solve``   -- eighb(scheme="chol") (slakonet/utils.py:585-816)
* ``post``    -- fermi_search, band gap, DOS, properties dict
                 (slakonet/fermi.py:139-215, slakonet/main.py:600-1819,
                 slakonet/optim.py:1011-1074)

How each part is pinned (tests/test_oracle_*.py):

* interp: exact reproduction of degree<=7 polynomials, exponential tables,
  the stencil clamp rule, tail end-point values.
* skrot: orthogonality of the rotation matrices and agreement with the
  textbook Slater-Koster (1954) two-centre table written independently in
  the tests.
* build: brute-force Hermiticity, H(k+G)=H(k), real at Gamma, S(k) > 0, and a
  hand-evaluated dimer block.
* solve: scipy.linalg.eigh(H, S) (LAPACK zhegvd) and residuals.
* post: DOS closed form / integral, bisection root of N(mu), hand-made
  eigenvalue sets for the gap rules; atom-projected DOS summed over types
  equals the total DOS for orthonormal eigenvectors. The exact stopped iterate of the
  reference's half-step Newton Fermi search is pinned by its invariants only
  (first iterate meeting |dL/dmu| < tol, start at mean eigenvalue): parity
  for fermi_energy_eV is therefore "partial" (see module docstring).
"""
