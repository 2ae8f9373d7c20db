"""CPU oracle for the tt-fort local-solver hot path (TEST INFRASTRUCTURE ONLY).

This package restates, in numpy, the reference algorithms of oseledets/tt-fort
that the HIP library replaces.  It exists to *check* the GPU product path and
to time the reference algorithm on host cores (the ``cpu_baseline`` leg of
``bench.py``).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import it; the product path never does.

Modules
-------
kernels    -- local operator kernels: bfun3 family, phi updates, full local
              matrix, chop rules, QR, block-Jacobi, duchol, Householder
              GMRES, and the Krylov exp(tA)v (ZGEXPV) with the (6,6) Pade
              on the Hessenberg (ZGPADM) and the thin zexp_mv layer.
operators  -- SDV file reader, MPO builders for the benchmark Hamiltonians,
              small-TT helpers (full vectors, random cores), and the exact
              tensor-DST solution of the Dirichlet Laplacian that pins the
              AMEn solves.
drivers    -- the sweep drivers tt_eigb and ztt_ksl, and tt_mvk4 (the
              approximate TT-matrix by TT-vector product).

Each function cites the reference file:line it follows.  Functions whose
results no golden vector, closed form or invariant pins are marked
"parity unpinned" in their docstring.
"""
