"""CPU FP64 restatement of NetworkSolvers.jl's sweep hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product
(``networksolvers_amd``) never does, and it fails loudly when its HIP library is
missing instead of falling back here.

Each function cites the reference file:line it restates (paths relative to the
NetworkSolvers.jl source tree).  Arithmetic that the reference delegates to
third-party Julia packages that are not available here (KrylovKit 0.6-0.8,
ITensors 0.7-0.9 / NDTensors, ITensorNetworks 0.11-0.13) is restated from their
published algorithms; every such function names the package and is pinned by
the checks listed in its docstring (ED, brute-force Kronecker matrices, closed
forms, invariants).  Functions with no such check say "parity unpinned".

Modules
  linalg  -- truncate! / factorize / qr / Lanczos eigsolve / Krylov exponentiate
  chain   -- MPS/MPO tensors, environments, H_eff apply, 2-site DMRG driver
  models  -- MPOs for the spin chains of the benchmark configurations
  plans   -- Euler-tour and post-order region plans (integer host logic)
  ed      -- exact diagonalisation in a fixed-Sz sector (scipy sparse Lanczos)
"""
