"""Steady monochromatic radiative transfer problem (interface of ``src/rt``).

``Problem(kappa, sigma, Phi, f, bcs_dirac)`` keeps the reference's
constructor and method names (``src/rt/class_Problem.py:11-32``).  The
operator and its solve run on the GPU through ``rtx``:

* ``solve(mesh, **kwargs)`` -> ``[Projection, convergence_info, matrix_info]``
  like ``src/rt/solve.py:21-149``, with a matrix-free device operator,
  device block-Jacobi and device GMRES instead of PETSc;
* the assembled-matrix methods return ``scipy.sparse.csr_matrix`` in the
  reference's DOF order (the "assembled scipy-sparse path"), their entries
  evaluated on the device (``rtx.assembly``);
* ``forcing_vector`` / ``boundary_conditions_vector`` return numpy vectors;
* ``operator(mesh)`` exposes the matrix-free device operator itself.
"""
from typing import Callable

from .problem import Problem  # noqa: F401

__all__ = ["Problem"]
