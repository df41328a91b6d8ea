"""FP64 CPU oracle for the muspinsim spin-dynamics hot path.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy (complex128 /
float64) restatement of the reference algorithms, used as the checker for
the HIP path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``muspinsim_amd`` never does, and its GPU path fails loudly when the HIP
library is missing instead of falling back to anything here.

Every function cites the reference file:line it restates (paths relative to
the muspinsim source tree, v2.3.1).

Modules
-------
spin       spin matrices, Kronecker-structured operators, superoperators
physics    interaction-term tensors -> H, Zeeman, rho0, dissipators, L
dynamics   evolve / integrate_decaying / fast_evolve / Lindblad evolve
orient     quaternions, ZCW and Euler-grid orientation sets, get_xy
pipeline   ExperimentRunner.run_single routing and powder accumulation
baseline   multi-process CPU timing harness (mirrors mpi.py's split)

Pinning status (what checks each function; see tests/test_oracle_*.py)
---------------------------------------------------------------------
spin.*            reference known answers (tests/test_spinop.py) + identities
physics.*         reference known answers (tests/test_spinsys.py,
                  tests/test_experiment.py) + closed forms
dynamics.*        reference closed forms (tests/test_hamiltonian.py,
                  tests/test_lindbladian.py) + expm propagation + quadrature
orient.quat_*     reference axis-angle products (tests/test_config.py:214-286,
                  tests/test_utils.py:43-65)
orient.zcw_angles parity unpinned against soprano's ZCW point set (soprano is
                  absent); pinned only by the <3cos^2-1> = 0 sphere average
                  (tests/test_utils.py:67-73) and the published ZCW formula.
"""
