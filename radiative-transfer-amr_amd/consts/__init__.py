"""Numeric constants shared by the mesh, projection and transport packages.

Mirrors ``src/consts/__init__.py:8-15`` of the reference: integer and real
dtypes, pi, the tolerance ``EPS`` (numpy's ``finfo(float64).resolution``, i.e.
1e-15, not machine epsilon), the largest double and the root rank.
"""
import numpy as np

INT: type = np.int64
REAL: type = np.float64

PI: float = np.pi
EPS: float = float(np.finfo(REAL).resolution)   # 1e-15
INF: float = float(np.finfo(REAL).max)

COMM_ROOT: int = 0
