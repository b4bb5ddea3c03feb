"""CPU oracle for the DG transport operator and its solve.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy/scipy restatement of
the reference's assembled path (``src/rt/*.py``, ``src/dg/matrix/*.py``,
``src/dg/quadrature/*.py`` of jasonltorchinsky/radiative-transfer-amr).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import it, and only as the checker (or the timed CPU baseline).  The
product package never imports it.

Modules
  quad     - LGL / LG rules (eigenvalue route, independent of the product's
             Newton route) and the Lagrange basis with its [-1, 1] support rule
  emat     - the face-coupling matrices E^x, E^y, E^theta in the reference's
             four (level difference, order) branches
  meshq    - mesh queries restated from the reference: leaf ordering, DOF
             offsets, neighbour cells in a neighbour column, brute-force
             geometric face adjacency
  loops    - literal scalar-loop assembly mirroring the reference's loop nests
             and drop thresholds (tiny meshes only)
  assemble - vectorised assembly of the same matrices, optionally restricted
             to the rows of a subset of columns (large meshes)
  system   - interior/boundary split, right-hand side, exact and Krylov
             solves, per-cell block-Jacobi, MLP reference
  error    - analytic L2 error of a nodal solution (interpolation to
             n + 3 nodes per direction and weighted squared error)

Parity pins (what fixes each function) are listed in each module header.
Functions without an independent pin say "parity unpinned" there.
"""
