"""CPU oracle for the CP-ALS / MTTKRP path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use
this package, and only as the checker (or the timed CPU baseline).  The product
package ``tenrso_amd`` never imports it.

It wraps ``liboracle.so`` (built from ``cp_oracle.cpp``, a line-faithful C++
restatement of the reference Rust functions cited there) plus numpy helpers.

Parity status of each function (how it is pinned, see tests/test_oracle.py):
  mttkrp (faithful f64/f32, fast f64)  pinned: reference KATs (mttkrp.rs:1000-1043
                                       brute force, property_tests.rs:208-229 all-ones),
                                       numpy einsum brute force, rank-1 closed form,
                                       linearity, faithful == fast == blocked-parallel.
  khatri_rao                           pinned: khatri_rao.rs doc test + tests 479-554.
  gram / gram_hadamard                 pinned: cp/tests.rs:26-36 KAT, symmetry, einsum.
  solve_ls                             pinned by the normal equations (A G = M),
                                       numpy.linalg.lstsq, the closed form of the ridge
                                       retry on a singular diagonal system (for the f32
                                       and f64 epsilons) and LinalgError on systems that
                                       fail the retry; scirs2_linalg::lstsq itself is
                                       third-party and absent -> its rounding and its
                                       failure condition on borderline systems are
                                       parity unpinned.
  fit / recon_norm_sq / inner product  pinned: explicit dense 1-||X-Xhat||/||X||.
  cp_als (f64, f32, constrained)       pinned: exact low-rank recovery (fit -> 1),
                                       normal-equation residual of every update,
                                       convergence-rule unit cases, the oscillation
                                       rule against the fit history, LinalgError
                                       propagation.  x_lead (leading modes on another
                                       tensor) is pinned by x_lead = X reproducing the
                                       plain run bit for bit.  orthogonal=True uses
                                       Householder QR: parity unpinned.
  cp_reconstruct / extract_weights     pinned: numpy einsum, KAT outer.rs doc test.
  nmode_product / nmode_products_seq   pinned: nmode.rs tests (mode-0 exact values,
                                       identity, shapes, the error messages) and a
                                       brute-force loop on small shapes.
  cp_factor_grads                      pinned: FP64 central differences, identity
                                       dL/dA_n = A_n G_n - M_n.
  coo_mttkrp                           pinned: COO -> dense + dense MTTKRP.
  hicoo_from_coo                       pinned: hicoo.rs:414-575 unit cases (block
                                       counts, coordinates, 5-way, empty, errors),
                                       stable-order and round-trip invariants.
"""
from .oracle import *  # noqa: F401,F403
