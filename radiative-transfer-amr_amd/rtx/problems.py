"""Workload definitions: meshes and coefficient callables of the benchmark
configurations (the analogue of the reference's
``experiments/*/problem.py`` and ``experiments/*/mesh.py``).

* ``henyey_greenstein(g)``: the reference's scattering kernel form
  (``experiments/two-strong-scatterers/problem.py:36-43``), renormalised on
  [0, 2 pi); the normalisation uses the closed form
  N_g = 4 E(m) / (1 - g), m = 4 g / (1 + g)^2 (complete elliptic integral of
  the second kind) instead of adaptive quadrature.
* ``uniform_mesh``: the reference's experiment meshes
  (``experiments/*/mesh.py``: ``nref_ang`` angular then ``nref_spt`` spatial
  uniform h-refinements).
* ``manufactured``: smooth manufactured solution with consistent forcing and
  inflow data.
* ``absorber_c2``: kappa = 1 + 0.5 sin(2 pi x) sin(2 pi y), sigma = 0.7 kappa,
  HG g = 0.8, unit isotropic inflow on the left wall, no forcing.
* ``two_scatterers``: two Gaussian strong-scattering blobs in a weakly
  scattering background.
* ``two_scatterers_c3``: the hp-mesh workload's fixed pair of blobs;
  ``two_scatterer_field(seed)``: a seeded member of the same family (blob
  centres, radius, strength and ratios drawn at random), the training
  fields of its learned coarse correction.
"""
import math

import numpy as np
from scipy import integrate, special

from dg.mesh import Mesh


def hg_norm(g: float) -> float:
    """int_0^{2 pi} (1 - g^2) / (1 + g^2 - 2 g cos t)^{3/2} dt."""
    if g == 0.0:
        return 2.0 * math.pi
    m = 4.0 * g / (1.0 + g) ** 2
    return 4.0 * special.ellipe(m) / (1.0 - g)


def henyey_greenstein(g: float):
    norm = hg_norm(g)

    def Phi(th, phi):
        return ((1.0 - g * g) / norm) / (1.0 + g * g - 2.0 * g * np.cos(th - phi)) ** 1.5
    Phi.g = g
    Phi.norm = norm
    return Phi


def hg_first_moment(g: float) -> float:
    """int_0^{2 pi} Phi(t, 0) cos t dt for the normalised HG kernel."""
    Phi = henyey_greenstein(g)
    v, _ = integrate.quad(lambda t: Phi(t, 0.0) * math.cos(t), 0.0, 2.0 * math.pi,
                          epsabs=1e-14, epsrel=1e-14, limit=400)
    return v


def uniform_mesh(Ls, ndofs, nref_ang, nref_spt, pbcs=(False, False)):
    mesh = Mesh(Ls=list(Ls), pbcs=list(pbcs), ndofs=list(ndofs), has_th=True)
    for _ in range(nref_ang):
        mesh.ref_mesh(kind="ang", form="h")
    for _ in range(nref_spt):
        mesh.ref_mesh(kind="spt", form="h")
    return mesh


class Workload:
    """kappa(x, y), sigma(x, y), Phi(th, phi), f(x, y, th), bcs(x, y, th)."""

    def __init__(self, name, kappa, sigma, Phi, f, bcs, u=None):
        self.name, self.kappa, self.sigma, self.Phi = name, kappa, sigma, Phi
        self.f, self.bcs, self.u = f, bcs, u

    @property
    def bcs_dirac(self):
        return [self.bcs, [None, None, None]]


def manufactured(kappa0=1.1, sigma0=0.5, g=0.6, width=0.1, c=(0.5, 0.5)):
    """u = exp(-|x - c|^2 / width) (1 + 0.5 cos th) with constant kappa, sigma
    and HG scattering; f = s . grad u + kappa u - sigma int Phi u."""
    Phi = henyey_greenstein(g)
    m1 = hg_first_moment(g)

    def U(x, y):
        return np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2) / width)

    def u(x, y, th):
        return U(x, y) * (1.0 + 0.5 * np.cos(th))

    def f(x, y, th):
        Ux = -2.0 * (x - c[0]) / width * U(x, y)
        Uy = -2.0 * (y - c[1]) / width * U(x, y)
        stream = (1.0 + 0.5 * np.cos(th)) * (np.cos(th) * Ux + np.sin(th) * Uy)
        return stream + kappa0 * u(x, y, th) - sigma0 * U(x, y) * (1.0 + 0.5 * m1 * np.cos(th))

    def kappa(x, y):
        return kappa0 + 0.0 * (x + y)

    def sigma(x, y):
        return sigma0 + 0.0 * (x + y)

    return Workload("manufactured", kappa, sigma, Phi, f, u, u=u)


def absorber_c2(g=0.8, ratio=0.7):
    Phi = henyey_greenstein(g)

    def kappa(x, y):
        return 1.0 + 0.5 * np.sin(2.0 * np.pi * x) * np.sin(2.0 * np.pi * y)

    def sigma(x, y):
        return ratio * kappa(x, y)

    def f(x, y, th):
        return 0.0 * (x + y + th)

    def bcs(x, y, th):
        return np.where(np.asarray(x) == 0.0, 1.0, 0.0) + 0.0 * (y + th)

    return Workload("absorber", kappa, sigma, Phi, f, bcs)


def two_scatterers(g=0.9, centres=((0.3, 0.6), (0.7, 0.4)), radius=0.1,
                   kappa_blob=10.0, ratio_blob=0.99, ratio_bg=0.1):
    Phi = henyey_greenstein(g)

    def blob(x, y):
        v = 0.0
        for cx, cy in centres:
            v = np.maximum(v, np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / radius ** 2))
        return v

    def kappa(x, y):
        return 1.0 + (kappa_blob - 1.0) * blob(x, y)

    def sigma(x, y):
        return (ratio_bg + (ratio_blob - ratio_bg) * blob(x, y)) * kappa(x, y)

    def f(x, y, th):
        return 0.0 * (x + y + th)

    def bcs(x, y, th):
        return np.where(np.asarray(x) == 0.0, 1.0, 0.0) + 0.0 * (y + th)

    return Workload("two_scatterers", kappa, sigma, Phi, f, bcs)


# ---------------------------------------------------------------- random fields
def grf(n, length_scale, seed, pad=2):
    """Zero-mean, unit-variance Gaussian random field on the cell centres of
    an n x n grid over [0, 1]^2 with squared-exponential covariance
    exp(-d^2 / (2 l^2)): white noise filtered in Fourier space on a grid
    padded ``pad`` times (to weaken the periodic wrap), then standardised."""
    m = n * pad
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((m, m))
    k = np.fft.fftfreq(m, d=1.0 / n)            # cycles per unit length
    kx, ky = np.meshgrid(k, k, indexing="ij")
    spec = np.exp(-2.0 * (np.pi * length_scale) ** 2 * (kx ** 2 + ky ** 2))
    field = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(spec)))[:n, :n]
    field -= field.mean()
    sd = field.std()
    return field / sd if sd > 0 else field


class GridField:
    """Bilinear interpolation of cell-centred values on an n x n grid over
    [0, Lx] x [0, Ly] (clamped at the walls); callable as f(x, y)."""

    def __init__(self, values, Ls=(1.0, 1.0)):
        self.v = np.asarray(values, dtype=np.float64)
        self.Ls = Ls

    def __call__(self, x, y):
        x, y = np.broadcast_arrays(np.asarray(x, dtype=float), np.asarray(y, dtype=float))
        nx, ny = self.v.shape
        u = np.clip(x / self.Ls[0] * nx - 0.5, 0.0, nx - 1.0)
        w = np.clip(y / self.Ls[1] * ny - 0.5, 0.0, ny - 1.0)
        i0 = np.minimum(np.floor(u).astype(int), nx - 2) if nx > 1 else np.zeros_like(u, int)
        j0 = np.minimum(np.floor(w).astype(int), ny - 2) if ny > 1 else np.zeros_like(w, int)
        fu, fw = u - i0, w - j0
        v = self.v
        i1, j1 = np.minimum(i0 + 1, nx - 1), np.minimum(j0 + 1, ny - 1)
        return ((1 - fu) * (1 - fw) * v[i0, j0] + fu * (1 - fw) * v[i1, j0]
                + (1 - fu) * fw * v[i0, j1] + fu * fw * v[i1, j1])


def grf_scatterer(seed, n=64, length_scale=0.15, ratio_range=(0.9, 0.999), g=0.5,
                  kappa_scale=1.0):
    """log kappa ~ GRF(length scale 0.15, mean log(kappa_scale), variance 1)
    on an n x n grid, sigma / kappa one scalar ~ U(ratio_range) (both from
    ``seed``), Henyey-Greenstein g, unit isotropic source, vacuum inflow."""
    logk = grf(n, length_scale, seed) + np.log(kappa_scale)
    ratio = float(np.random.default_rng(seed + 7919).uniform(*ratio_range))
    lk = GridField(logk)
    Phi = henyey_greenstein(g)

    def kappa(x, y):
        return np.exp(lk(x, y))

    def sigma(x, y):
        return ratio * kappa(x, y)

    def f(x, y, th):
        return 1.0 + 0.0 * (x + y + th)

    def bcs(x, y, th):
        return 0.0 * (x + y + th)

    wl = Workload("grf_scatterer", kappa, sigma, Phi, f, bcs)
    wl.seed, wl.ratio, wl.logk = seed, ratio, logk
    return wl


# ---------------------------------------------------------------- random hp mesh
def random_hp_mesh(seed=0, base_spt=4, base_ang=3, extra_spt=3, extra_ang=2, p_col=0.3,
                   p_cell=0.2, orders=(3, 9), Ls=(1.0, 1.0)):
    """Batched uneven-order hp mesh: a 2^base_spt x 2^base_spt x 2^base_ang
    base mesh, then (seeded) random h-refinement - every leaf column below
    base_spt + extra_spt levels is split with probability p_col per round,
    every leaf cell below base_ang + extra_ang with probability p_cell per
    round, extra_spt / extra_ang rounds, the reference's 2:1 balance
    refining neighbours as it goes (``Mesh.ref_col`` / ``ref_cell``) - and
    finally per-column (nx = ny) and per-cell (nth) node counts drawn
    uniformly from ``orders`` (inclusive).  Random orders do not keep the
    reference's |dn| <= 1 p-balance; the face couplings handle any orders.
    The defaults (16 x 16 x 8 base, p_col 0.3, p_cell 0.2, seed 0) give
    2,338 columns, 34,279 cells and 8,263,762 DOF."""
    rng = np.random.default_rng(seed)
    mesh = uniform_mesh(Ls, [orders[0]] * 3, base_ang, base_spt)
    for _ in range(extra_spt):
        for key in sorted(mesh.cols):
            col = mesh.cols.get(key)
            if col is None or not col.is_lf or col.lv >= base_spt + extra_spt:
                continue
            if rng.random() < p_col:
                mesh.ref_col(key, kind="spt", form="h")
    for _ in range(extra_ang):
        for key in sorted(mesh.cols):
            col = mesh.cols[key]
            for ck in sorted(col.cells):
                cell = col.cells.get(ck)
                if cell is None or not cell.is_lf or cell.lv >= base_ang + extra_ang:
                    continue
                if rng.random() < p_cell:
                    mesh.ref_cell(key, ck, form="h")
    for key in sorted(mesh.cols):
        col = mesh.cols[key]
        n = int(rng.integers(orders[0], orders[1] + 1))
        col.ndofs = [n, n]
        for ck in sorted(col.cells):
            col.cells[ck].ndofs = [int(rng.integers(orders[0], orders[1] + 1))]
    return mesh


def two_scatterers_c3(g=0.9):
    """kappa = 1 + 9 max(g1, g2), sigma / kappa = 0.1 + 0.89 max(g1, g2),
    g_k = exp(-|x - c_k|^2 / 0.1^2), c = (0.3, 0.6), (0.7, 0.4);
    Henyey-Greenstein g = 0.9; unit isotropic inflow on the left wall."""
    Phi = henyey_greenstein(g)

    def blob(x, y):
        return np.maximum(np.exp(-((x - 0.3) ** 2 + (y - 0.6) ** 2) / 0.01),
                          np.exp(-((x - 0.7) ** 2 + (y - 0.4) ** 2) / 0.01))

    def kappa(x, y):
        return 1.0 + 9.0 * blob(x, y)

    def sigma(x, y):
        return (0.1 + 0.89 * blob(x, y)) * kappa(x, y)

    def f(x, y, th):
        return 0.0 * (x + y + th)

    def bcs(x, y, th):
        return np.where(np.asarray(x) == 0.0, 1.0, 0.0) + 0.0 * (y + th)

    return Workload("two_scatterers_c3", kappa, sigma, Phi, f, bcs)


def two_scatterer_field(seed, g=0.9):
    """A seeded member of the two-strong-scatterer family: two Gaussian blobs
    with centres uniform in [0.15, 0.85]^2, a common radius U(0.07, 0.13),
    kappa = 1 + (k_b - 1) max(g1, g2) with k_b ~ U(5, 15), sigma / kappa =
    r_0 + (r_b - r_0) max(g1, g2) with r_b ~ U(0.95, 0.995) and r_0 ~
    U(0.05, 0.2); Henyey-Greenstein g; unit isotropic inflow on the left
    wall, no source (the form of ``two_scatterers_c3``)."""
    rng = np.random.default_rng(int(seed) + 7_777)
    cen = rng.uniform(0.15, 0.85, size=(2, 2))
    rad = float(rng.uniform(0.07, 0.13))
    kb = float(rng.uniform(5.0, 15.0))
    rb, r0 = float(rng.uniform(0.95, 0.995)), float(rng.uniform(0.05, 0.2))
    Phi = henyey_greenstein(g)

    def blob(x, y):
        return np.maximum(np.exp(-((x - cen[0, 0]) ** 2 + (y - cen[0, 1]) ** 2) / rad ** 2),
                          np.exp(-((x - cen[1, 0]) ** 2 + (y - cen[1, 1]) ** 2) / rad ** 2))

    def kappa(x, y):
        return 1.0 + (kb - 1.0) * blob(x, y)

    def sigma(x, y):
        return (r0 + (rb - r0) * blob(x, y)) * kappa(x, y)

    def f(x, y, th):
        return 0.0 * (x + y + th)

    def bcs(x, y, th):
        return np.where(np.asarray(x) == 0.0, 1.0, 0.0) + 0.0 * (y + th)

    wl = Workload("two_scatterer_field", kappa, sigma, Phi, f, bcs)
    wl.seed, wl.ratio = int(seed), rb
    return wl
