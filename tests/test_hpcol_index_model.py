"""CPU model of the hp column kernel (``csrc/rtx_hpcol.hip``), test-only.

Restates the kernel's data flow with its own index arithmetic: the
node-major LDS copy of a column with the padded row stride, the per-link
angle-contracted traces at the column-local offsets the host builds, the
angle -> cell table, the matrix-core operand reads (K steps of 4 for FP64 /
FP32, 16 for BF16), the accumulator rows, the face-node mapping of each
upwind link and the output addresses.  Every index is checked against its
buffer (LDS tile, link block, E pools, G block, output vector), and the
model's result is compared with the oracle's assembled operator, so the
indexing is known to be in bounds and the formulas right before the kernel
runs on a GPU.
"""
import numpy as np
import pytest

from oracle import assemble as asm
from rtx import problems

NM = 10
HP_MAX_CELLS, HP_MAX_LINKS, HP_MT, HP_FV = 64, 192, 7, 8


def hp_stride(na, kstep, c, M):
    n = (na + kstep - 1) // kstep * kstep
    return n + ((c - n) % M + M) % M


def hp_layout(nsp, na, nin, ntr, ea, mixed, nes=0):
    """(S, NA, XSB, total bytes): the angle-major copy has NA rows (angles,
    padded to the K step) of S = nsp rounded up to odd entries, in a region
    that first holds the ntr neighbour-trace values; mixed adds a node-major
    BF16 copy of row stride XSB; the inn region holds nin = max(ninn, nffc)
    entries, the es region the column's nes distinct E^s values."""
    kstep = 16 if mixed else 4
    S, NA = nsp | 1, (na + kstep - 1) // kstep * kstep
    XSB = hp_stride(na, 16, 8, 16) if mixed else 0
    o = max(NA * S, ntr) * ea
    o = (o + 15) & ~15
    o += nsp * XSB * 2
    o = (o + 15) & ~15
    o += nin * ea + nes * ea + 3 * na * ea + na * 4
    return S, NA, XSB, (o + 15) & ~15


def face_tables(T):
    """cell_fo (column-local offset of each cell's inflow-face sums: one
    block of nx x nth per face that carries links) and col_ffc, as
    hpcol_build computes them."""
    cell_fo = np.zeros(max(T.ncell, 1), np.int64)
    col_ffc = np.zeros(T.ncol, np.int64)
    for c in range(T.ncol):
        fo = 0
        for k in range(T.col_cell_start[c], T.col_cell_start[c + 1]):
            lf = 0
            for l in range(T.link_start[k], T.link_start[k + 1]):
                lf |= 1 << int(T.link_face[l])
            cell_fo[k] = fo
            fo += bin(lf).count("1") * int(T.col_nx[c]) * int(T.cell_nth[k])
        col_ffc[c] = fo
    return cell_fo, col_ffc


def es_tables(T):
    """link_esl, col_es0, col_nes and the staged pool, as hpcol_build builds
    them: each column's distinct E^s blocks (by pool offset, n0 x n1 values)
    copied contiguously in link order."""
    link_esl = np.zeros(max(T.nlink, 1), np.int64)
    col_es0 = np.zeros(T.ncol, np.int64)
    col_nes = np.zeros(T.ncol, np.int64)
    pool = []
    for c in range(T.ncol):
        seen = {}
        col_es0[c] = len(pool)
        nes = 0
        for l in range(T.link_start[T.col_cell_start[c]], T.link_start[T.col_cell_start[c + 1]]):
            off = int(T.link_Es_off[l])
            if off not in seen:
                F = int(T.link_face[l])
                c1 = T.cell_col[int(T.link_src[l])]
                n0 = int(T.col_nx[c] if F & 1 else T.col_ny[c])
                n1 = int(T.col_nx[c1] if F & 1 else T.col_ny[c1])
                seen[off] = nes
                pool.extend(T.Es_pool[off:off + n0 * n1])
                nes += n0 * n1
            link_esl[l] = seen[off]
        col_nes[c] = nes
    return link_esl, col_es0, col_nes, np.array(pool + [0.0])


def admitted(T, c, col_ffc=None):
    """The host's rule for the column kernel: square columns of 3..10 nodes,
    at most 64 cells and 192 links, at most HP_FV x 512 inflow-face sums."""
    k0, k1 = T.col_cell_start[c], T.col_cell_start[c + 1]
    if col_ffc is None:
        col_ffc = face_tables(T)[1]
    return (T.col_nx[c] == T.col_ny[c] and T.col_nx[c] >= 3 and k1 - k0 <= HP_MAX_CELLS
            and T.link_start[k1] - T.link_start[k0] <= HP_MAX_LINKS
            and col_ffc[c] <= HP_FV * 512)


def test_bf16_stride_spreads_operand_reads_over_banks():
    """The BF16 matrix-core operand reads of a half-wave (16 rows of the
    node-major copy, 2 quads of k) fall on 64 distinct 4-byte banks."""
    for na in range(1, 300):
        XSB = hp_stride(na, 16, 8, 16)
        assert XSB >= (na + 15) // 16 * 16
        words = {(l16 * XSB // 2 + 2 * lg + w) % 64 for l16 in range(16) for lg in range(2)
                 for w in range(2)}
        assert len(words) == 64, na


def host_tables(T):
    """link_inn (column-local offset of each link's n1 x nth block) and
    col_inn, as hpcol_build computes them."""
    link_inn = np.zeros(max(T.nlink, 1), np.int64)
    col_inn = np.zeros(T.ncol, np.int64)
    for c in range(T.ncol):
        k0, k1 = T.col_cell_start[c], T.col_cell_start[c + 1]
        o = 0
        for l in range(T.link_start[k0], T.link_start[k1]):
            s, F = T.link_src[l], T.link_face[l]
            c1 = T.cell_col[s]
            n1 = T.col_nx[c1] if F & 1 else T.col_ny[c1]
            link_inn[l] = o
            o += n1 * T.cell_nth[T.link_target[l]]
        col_inn[c] = o
    return link_inn, col_inn


def _on_bdry(bd, i, j, nx, ny):
    return bool(((bd & 1) and i == nx - 1) or ((bd & 2) and j == ny - 1) or
                ((bd & 4) and i == 0) or ((bd & 8) and j == 0))


def hpcol_model(T, kappa, sigma, G, G_off, x, restrict=True, lds_cap=160 * 1024):
    link_inn, col_inn = host_tables(T)
    cell_fo, col_ffc = face_tables(T)
    link_esl, col_es0, col_nes, es_pool = es_tables(T)
    y = np.full(T.ndof, np.nan)
    for c in range(T.ncol):
        assert admitted(T, c, col_ffc)
        k0, k1 = int(T.col_cell_start[c]), int(T.col_cell_start[c + 1])
        ncells = k1 - k0
        nx, ny = int(T.col_nx[c]), int(T.col_ny[c])
        nsp, na = nx * ny, int(T.col_nang[c])
        ncd = nsp * na
        cdof = int(T.cell_dof_off[k0])
        L0, nl = int(T.link_start[k0]), int(T.link_start[k1] - T.link_start[k0])
        ninn, nffc = int(col_inn[c]), int(col_ffc[c])
        nin, nes, es0 = max(ninn, nffc), int(col_nes[c]), int(col_es0[c])
        es = es_pool[es0:es0 + nes]                  # the LDS copy
        assert ncells <= HP_MAX_CELLS and nl <= HP_MAX_LINKS and nsp <= HP_MT * 16
        ntr = int(T.col_ntrace[c])
        for ea, mixed in ((8, False), (4, False), (4, True)):
            assert hp_layout(nsp, na, nin, ntr, ea, mixed, nes)[3] <= lds_cap
        S, NA, _, _ = hp_layout(nsp, na, nin, ntr, 8, False, nes)
        XSB = hp_layout(nsp, na, nin, ntr, 4, True, nes)[2]
        region = max(NA * S, ntr)
        x0, y0, x1, y1 = T.col_geom[c]
        dx, dy = x1 - x0, y1 - y0
        c_dof = [int(T.cell_dof_off[k]) - cdof for k in range(k0, k1)] + [ncd]
        c_nth = [int(T.cell_nth[k]) for k in range(k0, k1)]
        c_bd = [int(T.cell_bdry[k]) if restrict else 0 for k in range(k0, k1)]
        c_ao = [int(T.cell_ang_off[k]) for k in range(k0, k1)] + [na]
        c_l0 = [int(T.link_start[k]) - L0 for k in range(k0, k1)] + [nl]
        # phase 1: angle-major copy (xs[A * S + s]), padding rows zero
        xs = np.full(NA * S, np.nan)
        xs_b = np.full(nsp * XSB, np.nan)
        acell = []
        for k in range(ncells):
            acell += [k | (a << 8) for a in range(c_nth[k])]
        assert len(acell) == na
        for t in range(ncd):
            # the kernel's lookup: DOF t lies in the cell of angle t // nsp
            kk = acell[int((t + 0.5) / nsp)] & 255
            assert kk == int(np.searchsorted(c_dof[:ncells], t, side="right") - 1)
            assert c_dof[kk] == nsp * c_ao[kk]
            nt, lt = c_nth[kk], t - c_dof[kk]
            s = int((lt + 0.5) / nt)
            a = lt - s * nt
            assert s == lt // nt and 0 <= a < nt
            v = x[cdof + t]
            i = int((s + 0.5) / ny)
            if c_bd[kk] and _on_bdry(c_bd[kk], i, s - i * ny, nx, ny):
                v = 0.0
            col = c_ao[kk] + a
            assert 0 <= col * S + s < na * S and col < na
            xs[col * S + s] = v
            xs_b[s * XSB + col] = v
        xs[na * S:] = 0.0
        for s in range(nsp):
            xs_b[s * XSB + na:(s + 1) * XSB] = 0.0
        # entries S - 1 of each row are padding when nsp is even (never read)
        used = np.zeros(NA * S, bool)
        for A_ in range(NA):
            used[A_ * S:A_ * S + nsp] = True
        assert not np.isnan(xs[used]).any() and not np.isnan(xs_b).any()
        # phase 1: neighbour traces gathered into the x region at the
        # column-local trace offsets, then contracted with E^theta
        l_tr = [int(T.link_meta[4 * (L0 + li) + 2]) for li in range(nl)] + [ntr]
        tr = np.full(region, np.nan)
        for e in range(ntr):
            li = int(np.searchsorted(l_tr[:nl], e, side="right") - 1)
            l = L0 + li
            w1 = int(T.link_meta[4 * l + 1])
            F, nx1, ny1 = w1 & 3, (w1 >> 2) & 15, (w1 >> 6) & 15
            nt1, bd1 = (w1 >> 10) & 15, ((w1 >> 14) & 15) if restrict else 0
            le = e - l_tr[li]
            m1 = int((le + 0.5) / nt1)
            a = le - m1 * nt1
            n1 = nx1 if F & 1 else ny1
            assert 0 <= m1 < n1 and 0 <= a < nt1 and e < region
            i1, j1 = {0: (0, m1), 2: (nx1 - 1, m1), 1: (m1, 0), 3: (m1, ny1 - 1)}[F]
            v = 0.0
            if not (bd1 and _on_bdry(bd1, i1, j1, nx1, ny1)):
                addr = int(T.link_meta[4 * l]) + (i1 * ny1 + j1) * nt1 + a
                assert 0 <= addr < T.ndof
                v = x[addr]
            tr[e] = v
        inn = np.full(max(nin, 1), np.nan)
        l_inn = [int(link_inn[L0 + li]) for li in range(nl)] + [ninn]
        for e in range(ninn):
            li = int(np.searchsorted(l_inn[:nl], e, side="right") - 1)
            l = L0 + li
            w1 = int(T.link_meta[4 * l + 1])
            F, nt1, kk = w1 & 3, (w1 >> 10) & 15, w1 >> 18
            assert kk == T.link_target[l] - k0 and F == T.link_face[l]
            nt = c_nth[kk]
            le = e - l_inn[li]
            m1 = int((le + 0.5) / nt)
            r = le - m1 * nt
            assert 0 <= r < nt
            et = int(T.link_Eth_off[l]) + r
            acc = 0.0
            for a in range(nt1):
                ti = l_tr[li] + m1 * nt1 + a
                assert et + a * nt < T.Eth_pool.size and ti < l_tr[li + 1]
                acc += T.Eth_pool[et + a * nt] * tr[ti]
            inn[e] = acc
        # phase 1: inflow-face sums, entry e -> (cell kk, h-th link face F,
        # own face node m0, angle r), computed from inn and then written over it
        c_fo = [int(cell_fo[k]) for k in range(k0, k1)] + [nffc]
        c_lf = []
        for k in range(k0, k1):
            lf = 0
            for l in range(T.link_start[k], T.link_start[k + 1]):
                lf |= 1 << int(T.link_face[l])
            c_lf.append(lf)
        fv = np.full(max(nffc, 1), np.nan)
        for e in range(nffc):
            # the kernel's search: the largest kk with c_fo[kk] <= e
            kk = int(np.searchsorted(c_fo[:ncells], e, side="right") - 1)
            assert c_fo[kk] <= e < c_fo[kk + 1]
            nt, lf = c_nth[kk], c_lf[kk]
            le = e - c_fo[kk]
            row = int((le + 0.5) / nt)
            r = le - row * nt
            hs = int((row + 0.5) / nx)
            m0 = row - hs * nx
            assert row == le // nt and hs == row // nx and 0 <= r < nt and 0 <= m0 < nx
            faces = [F for F in range(4) if (lf >> F) & 1]
            assert hs < len(faces)
            F = faces[hs]
            dth = T.cell_geom[k0 + kk, 1] - T.cell_geom[k0 + kk, 0]
            fxy = ((dx if F & 1 else dy) * 0.25) * dth
            acc = 0.0
            for li in range(c_l0[kk], c_l0[kk + 1]):
                l = L0 + li
                if int(T.link_face[l]) != F:
                    continue
                c1 = T.cell_col[int(T.link_src[l])]
                n1 = int(T.col_nx[c1] if F & 1 else T.col_ny[c1])
                eo = int(link_esl[l]) + m0
                f = 0.0
                for m1 in range(n1):
                    ii = l_inn[li] + m1 * nt + r
                    assert eo + m1 * nx < nes and ii < ninn
                    assert es[eo + m1 * nx] == T.Es_pool[int(T.link_Es_off[l]) + m0 + m1 * nx]
                    f += es[eo + m1 * nx] * inn[ii]
                acc += fxy * f
            fv[e] = acc
        inn[:nffc] = fv[:nffc]
        # phase 2: matrix-core operands (FP64 K step 4, BF16 K step 16)
        g0 = int(G_off[c])
        assert g0 + na * na <= G.size
        Gt = G[g0:g0 + na * na].reshape(na, na)
        for kb in range(0, na, 4):
            for lg in range(4):
                Ak = kb + lg
                assert Ak < NA                      # padding row, zero beyond na
        for kb in range(0, na, 16):
            assert kb + 16 <= XSB
        X = xs.reshape(NA, S)[:na, :nsp].T
        scat = X @ Gt                               # [s, R]
        for a in range(na):
            kk = int(np.searchsorted(c_ao[:ncells], a, side="right") - 1)
            assert acell[a] == kk | ((a - c_ao[kk]) << 8)
        wx = T.lgl_w[nx * NM:nx * NM + nx]
        wy = T.lgl_w[ny * NM:ny * NM + ny]
        WDx = np.array([[wx[i] * T.lgl_D[(nx * NM + p) * NM + i] for i in range(nx)] for p in range(nx)])
        WDy = np.array([[wy[j] * T.lgl_D[(ny * NM + q) * NM + j] for j in range(ny)] for q in range(ny)])
        sp0 = int(T.col_sp_off[c])
        ga0 = int(T.col_ang_start[c])
        for R in range(na):
            ac = acell[R]
            kk, r = ac & 255, ac >> 8
            nt, dof, bd = c_nth[kk], c_dof[kk], c_bd[kk]
            quad = int(T.cell_quad[k0 + kk])
            dth = T.cell_geom[k0 + kk, 1] - T.cell_geom[k0 + kk, 0]
            ct, sn, w_r = T.ang_trig[3 * (ga0 + R):3 * (ga0 + R) + 3]
            vol = dx * dy * 0.125 * dth * w_r
            cx, cy = dy * 0.25 * dth * w_r * ct, dx * 0.25 * dth * w_r * sn
            for s in range(nsp):
                p = int((s + 0.5) / ny)
                q = s - p * ny
                xv = xs[R * S + s]
                val = vol * wx[p] * wy[q] * kappa[sp0 + s] * xv
                sx = sum(WDx[p, i] * xs[R * S + i * ny + q] for i in range(nx))
                sy = sum(WDy[q, j] * xs[R * S + p * ny + j] for j in range(ny))
                val -= cx * wy[q] * sx + cy * wx[p] * sy
                if quad >= 0:
                    for hh in range(2):
                        F = (quad + hh) & 3
                        on = (F == 0 and p == nx - 1) or (F == 2 and p == 0) or \
                             (F == 1 and q == ny - 1) or (F == 3 and q == 0)
                        if on:
                            tf = [ct, sn, -ct, -sn][F]
                            val += ((dx if F & 1 else dy) * 0.25) * dth * \
                                (wx[p] if F & 1 else wy[q]) * w_r * tf * xv
                slot = 0
                for F in range(4):
                    if not (c_lf[kk] >> F) & 1:
                        continue
                    m0 = {0: q if p == nx - 1 else -1, 2: q if p == 0 else -1,
                          1: p if q == ny - 1 else -1, 3: p if q == 0 else -1}[F]
                    if m0 >= 0:
                        ii = c_fo[kk] + (slot * nx + m0) * nt + r
                        assert c_fo[kk] <= ii < c_fo[kk + 1]
                        val += inn[ii]
                    slot += 1
                sgw = dx * dy * sigma[sp0 + s] * wx[p] * wy[q]
                val -= sgw * scat[s, R]
                masked = bool(bd) and _on_bdry(bd, p, q, nx, ny)
                addr = cdof + dof + s * nt + r
                assert 0 <= addr < T.ndof and np.isnan(y[addr])
                y[addr] = 0.0 if masked else val
    assert not np.isnan(y).any()                     # every output written once
    return y


def _small_random_hp(seed):
    return problems.random_hp_mesh(seed, base_spt=1, base_ang=2, extra_spt=1, extra_ang=1,
                                   p_col=0.5, p_cell=0.5)


@pytest.mark.parametrize("case", ["random_hp_1", "random_hp_2", "periodic_x"])
def test_hpcol_model_matches_oracle(case):
    from rtx.plan import MeshTables
    if case == "periodic_x":
        mesh = problems.uniform_mesh([1.0, 1.0], [4, 4, 3], 2, 2, pbcs=(True, False))
        mesh.ref_col(sorted(mesh.cols)[0], kind="spt", form="h")
    else:
        mesh = _small_random_hp(int(case[-1]))
    wl = problems.two_scatterers()
    T = MeshTables(mesh)
    G, G_off = T.G_pool(wl.Phi)
    kappa, sigma = T.eval_spatial(wl.kappa), T.eval_spatial(wl.sigma)
    x = np.random.default_rng(3).uniform(-1, 1, T.ndof)
    y = hpcol_model(T, kappa, sigma, G, G_off, x)
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi, drop=False)
    A = (V["BC"] - V["IC"] + V["M"] + V["S"]).tocsr()
    mask = asm.intr_mask(mesh)
    ref = mask * (A @ (mask * x))
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    assert err < 1e-12, err


def test_hpcol_host_tables_fit_c3_and_c4():
    """Every column of the C3 and C4 meshes fits the kernel's static tables
    and the device's LDS in all three precisions (no generic fallback)."""
    from rtx.plan import MeshTables
    for p_col in (0.3, 0.4):
        T = MeshTables(problems.random_hp_mesh(0, p_col=p_col))
        _, col_inn = host_tables(T)
        _, col_ffc = face_tables(T)
        _, _, col_nes, _ = es_tables(T)
        for c in range(T.ncol):
            k0, k1 = T.col_cell_start[c], T.col_cell_start[c + 1]
            nsp, na = int(T.col_nx[c] * T.col_ny[c]), int(T.col_nang[c])
            assert k1 - k0 <= HP_MAX_CELLS
            assert T.link_start[k1] - T.link_start[k0] <= HP_MAX_LINKS
            assert admitted(T, c, col_ffc)
            nin = max(int(col_inn[c]), int(col_ffc[c]))
            for ea, mixed in ((8, False), (4, False), (4, True)):
                assert hp_layout(nsp, na, nin, int(T.col_ntrace[c]), ea, mixed,
                                 int(col_nes[c]))[3] <= 148 * 1024


def _hp_find(starts, hi, t, log):
    """The kernel's fixed-step search, step for step."""
    lo, s = 0, 1 << (log - 1)
    while s > 0:
        mid = lo + s
        if mid < hi and starts[mid] <= t:
            lo = mid
        s >>= 1
    return lo


def test_fixed_step_search_is_the_largest_start_at_or_below():
    """Largest i < hi with starts[i] <= t, for ascending starts with repeats
    (cells without face sums repeat their offset), at every hi up to 2^log."""
    rng = np.random.default_rng(7)
    for log, hmax in ((8, 192), (6, 64)):
        for _ in range(200):
            hi = int(rng.integers(1, hmax + 1))
            starts = np.concatenate([[0], np.cumsum(rng.integers(0, 4, hi - 1))]).astype(int)
            for t in range(int(starts[-1]) + 3):
                ref = int(np.searchsorted(starts, t, side="right") - 1)
                assert _hp_find(starts, hi, t, log) == ref
