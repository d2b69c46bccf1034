"""CPU oracle for the Hψ hot path -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch numpy restatement of the algorithms of LLNL/qball's
Hamiltonian-application path.  It exists to check the HIP product path; it is never the
thing measured or shipped.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.

Each function cites the reference file:line it follows (paths relative to the qball
source tree).  The reference itself cannot be built in this image (every source file
includes the autoconf-generated ``config.h``), so the oracle is pinned by:

* the reference-measured basis/grid counts (9139 / 72919 plane waves, 72^3 grid,
  423 rods / 845 z-columns / ntrans0 17 for the Si64 cell, and the Si512 / Al500 rows);
* closed forms (plane waves, constant potentials, single projectors);
* brute-force direct sums on tiny cells (DFT, projector contractions, density);
* invariants (Hermiticity, charge integral, norm conservation, Γ reality).

Functions whose numerics no such check pins are marked "parity unpinned" in their
docstrings.
"""
