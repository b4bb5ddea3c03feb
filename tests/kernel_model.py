"""CPU model of the device kernels' index arithmetic (test-only).

Walks the plan tables exactly as ``csrc/rtx_apply.hip`` does (per-cell
transport with the link table, per-column scattering with the G^T pool) and
asserts every index stays inside its buffer.  Used to check the tables and
the indexing against the oracle on the CPU before any GPU run; it is never
imported by the product.
"""
import numpy as np

NM = 10


def _on_bdry(bd, i, j, nx, ny):
    return ((bd & 1) and i == nx - 1) or ((bd & 2) and j == ny - 1) or \
        ((bd & 4) and i == 0) or ((bd & 8) and j == 0)


def apply_model(T, kappa, sigma, G, G_off, x, restrict=False,
                ops=("M", "IC", "BC", "S")):
    ndof = T.ndof
    assert x.shape == (ndof,)
    y = np.zeros(ndof)
    for k in range(T.ncell):
        c = T.cell_col[k]
        nx, ny, nth = int(T.col_nx[c]), int(T.col_ny[c]), int(T.cell_nth[k])
        n = nx * ny * nth
        off = int(T.cell_dof_off[k])
        assert 0 <= off and off + n <= ndof
        bd = int(T.cell_bdry[k]) if restrict else 0
        x0, y0, x1, y1 = T.col_geom[c]
        dx, dy = x1 - x0, y1 - y0
        th0, th1 = T.cell_geom[k]
        dth = th1 - th0
        wx = T.lgl_w[nx * NM:nx * NM + nx]
        wy = T.lgl_w[ny * NM:ny * NM + ny]
        wt = T.lg_w[nth * NM:nth * NM + nth]
        xs = x[off:off + n].reshape(nx, ny, nth).copy()
        if bd:
            for i in range(nx):
                for j in range(ny):
                    if _on_bdry(bd, i, j, nx, ny):
                        xs[i, j, :] = 0
        th = th0 + dth * 0.5 * (T.lg_x[nth * NM:nth * NM + nth] + 1.0)
        ct, st = np.cos(th), np.sin(th)
        Dx = np.array([[T.lgl_D[(nx * NM + p) * NM + i] for i in range(nx)] for p in range(nx)])
        Dy = np.array([[T.lgl_D[(ny * NM + q) * NM + j] for j in range(ny)] for q in range(ny)])
        sp0 = int(T.col_sp_off[c])
        kap = kappa[sp0:sp0 + nx * ny].reshape(nx, ny)
        ys = np.zeros((nx, ny, nth))
        if "M" in ops:
            ys += (dx * dy * dth / 8) * (wx[:, None] * wy[None, :] * kap)[:, :, None] * wt * xs
        if "IC" in ops:
            sx = np.einsum("pi,i,iqr->pqr", Dx, wx, xs)
            sy = np.einsum("qj,j,pjr->pqr", Dy, wy, xs)
            ys -= (dy * dth / 4) * wy[None, :, None] * (wt * ct)[None, None, :] * sx
            ys -= (dx * dth / 4) * wx[:, None, None] * (wt * st)[None, None, :] * sy
        q = int(T.cell_quad[k])
        if "BC" in ops:
            if q >= 0:
                for F in (q, (q + 1) % 4):
                    tf = [ct, st, -ct, -st][F]
                    if F == 0:
                        ys[nx - 1] += (dy * dth / 4) * wy[:, None] * (wt * tf)[None, :] * xs[nx - 1]
                    elif F == 2:
                        ys[0] += (dy * dth / 4) * wy[:, None] * (wt * tf)[None, :] * xs[0]
                    elif F == 1:
                        ys[:, ny - 1] += (dx * dth / 4) * wx[:, None] * (wt * tf)[None, :] * xs[:, ny - 1]
                    else:
                        ys[:, 0] += (dx * dth / 4) * wx[:, None] * (wt * tf)[None, :] * xs[:, 0]
            for l in range(T.link_start[k], T.link_start[k + 1]):
                s, F = int(T.link_src[l]), int(T.link_face[l])
                c1 = T.cell_col[s]
                nx1, ny1, nt1 = int(T.col_nx[c1]), int(T.col_ny[c1]), int(T.cell_nth[s])
                off1 = int(T.cell_dof_off[s])
                assert off1 + nx1 * ny1 * nt1 <= ndof
                x1c = x[off1:off1 + nx1 * ny1 * nt1].reshape(nx1, ny1, nt1).copy()
                if restrict and T.cell_bdry[s]:
                    for i in range(nx1):
                        for j in range(ny1):
                            if _on_bdry(int(T.cell_bdry[s]), i, j, nx1, ny1):
                                x1c[i, j, :] = 0
                tr = {0: x1c[0], 2: x1c[nx1 - 1], 1: x1c[:, 0], 3: x1c[:, ny1 - 1]}[F]
                n1, n0 = tr.shape[0], (nx if F % 2 else ny)
                eo, to = int(T.link_Es_off[l]), int(T.link_Eth_off[l])
                assert eo + n1 * n0 <= T.Es_pool.size and to + nt1 * nth <= T.Eth_pool.size
                Es = T.Es_pool[eo:eo + n1 * n0].reshape(n1, n0)
                Et = T.Eth_pool[to:to + nt1 * nth].reshape(nt1, nth)
                cf = (dx if F % 2 else dy) * dth / 4
                contrib = cf * np.einsum("mo,ar,ma->or", Es, Et, tr)
                if F == 0:
                    ys[nx - 1] += contrib
                elif F == 2:
                    ys[0] += contrib
                elif F == 1:
                    ys[:, ny - 1] += contrib
                else:
                    ys[:, 0] += contrib
        y[off:off + n] = ys.ravel()
    if "S" in ops:
        for c in range(T.ncol):
            nx, ny = int(T.col_nx[c]), int(T.col_ny[c])
            na = int(T.col_nang[c])
            k0, k1 = T.col_cell_start[c], T.col_cell_start[c + 1]
            aoff, ant, aloc = [], [], []
            for k in range(k0, k1):
                for a in range(T.cell_nth[k]):
                    aoff.append(int(T.cell_dof_off[k]))
                    ant.append(int(T.cell_nth[k]))
                    aloc.append(a)
            assert len(aoff) == na
            g0 = int(G_off[c])
            assert g0 + na * na <= G.size
            Gt = G[g0:g0 + na * na].reshape(na, na)
            x0, y0, x1, y1 = T.col_geom[c]
            wx = T.lgl_w[nx * NM:nx * NM + nx]
            wy = T.lgl_w[ny * NM:ny * NM + ny]
            sp0 = int(T.col_sp_off[c])
            for s in range(nx * ny):
                i, j = divmod(s, ny)
                addr = np.array([aoff[A] + s * ant[A] + aloc[A] for A in range(na)])
                u = x[addr].copy()
                if restrict:
                    for A in range(na):
                        kk = k0 + np.searchsorted(np.cumsum([T.cell_nth[kq] for kq in range(k0, k1)]), A, side="right")
                        if T.cell_bdry[kk] and _on_bdry(int(T.cell_bdry[kk]), i, j, nx, ny):
                            u[A] = 0
                v = -(x1 - x0) * (y1 - y0) * wx[i] * wy[j] * sigma[sp0 + s] * (Gt.T @ u)
                y[addr] += v
    if restrict:
        for k in range(T.ncell):
            bd = int(T.cell_bdry[k])
            if not bd:
                continue
            c = T.cell_col[k]
            nx, ny, nth = int(T.col_nx[c]), int(T.col_ny[c]), int(T.cell_nth[k])
            off = int(T.cell_dof_off[k])
            blk = y[off:off + nx * ny * nth].reshape(nx, ny, nth)
            for i in range(nx):
                for j in range(ny):
                    if _on_bdry(bd, i, j, nx, ny):
                        blk[i, j, :] = 0
    return y
