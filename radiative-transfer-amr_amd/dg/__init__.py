"""Discontinuous-Galerkin core: quadrature, hp mesh, nodal projections and
global DOF indexing (interface of the reference's ``src/dg``)."""
from . import quadrature, mesh, projection, matrix  # noqa: F401
