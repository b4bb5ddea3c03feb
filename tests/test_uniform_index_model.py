"""CPU model of the uniform-bin apply kernel's index arithmetic (test-only).

Restates the compile-time shapes of ``csrc/rtx_uniform.hip`` (threads per
workgroup, LDS row strides, whole-cell DOF passes, trace and flux-slot
passes, phase-2 line mapping) for every compiled order bin and part count,
and checks the properties the kernel relies on: every DOF, trace slot and
line is visited exactly once, every LDS index stays inside its tile, and
lanes walking a cell's DOFs in order hit distinct LDS banks (32 banks of
4 bytes; 8-byte elements).
"""
import pytest

BINS = [(5, 5, 16), (4, 4, 16), (4, 4, 4), (3, 3, 8), (5, 5, 8), (4, 4, 8), (3, 3, 4)]


def u_threads(nth, nc, h):
    return 512 if ((nc // h) * nth + 15) // 16 > 4 else 256


def u_pad(na, nth):
    if nth % 2 == 0:
        return na | 1
    v = na
    while v % 16 != nth % 16:
        v += 1
    return v


def u_valid(nth, nc, h):
    # whole cells per part, at most eight 16-angle output tiles (one per
    # wavefront of a 512-thread workgroup)
    return nc % h == 0 and ((nc // h) * nth + 15) // 16 <= 512 // 64


CASES = [(n, nth, nc, h) for (n, nth, nc) in BINS for h in (1, 2, 4) if u_valid(nth, nc, h)]


@pytest.mark.parametrize("n,nth,nc,h", CASES)
def test_passes_cover_each_index_once(n, nth, nc, h):
    ut = u_threads(nth, nc, h)
    nsp, na = n * n, nc * nth
    ncd, cd, oc = nsp * na, nsp * nth, nc // h
    noa = oc * nth
    nap, noap = u_pad(na, nth), u_pad(noa, nth)
    assert nap >= na and noap >= noa and nap % 2 == 1 and noap % 2 == 1
    cpp = ut // cd if cd <= ut else 0
    xs_ = cpp * cd if cpp else ut
    xu = (ncd + xs_ - 1) // xs_
    # phase 1: column DOFs t = t0 + u XS, t0 = tid (tid < XS) else NCD
    seen, lds = [], set()
    for tid in range(ut):
        t0 = tid if tid < xs_ else ncd
        for u in range(xu):
            t = t0 + u * xs_
            if t >= ncd:
                continue
            kk = t0 // cd + u * cpp if cpp else t // cd
            lt = t0 % cd if cpp else t % cd
            assert kk * cd + lt == t
            sn, a = lt // nth, lt % nth
            li = sn * nap + kk * nth + a
            assert 0 <= li < nsp * nap
            seen.append(t)
            lds.add(li)
    assert sorted(seen) == list(range(ncd)) and len(lds) == ncd
    # phase 5: own DOFs (kl, lt) of the part, kl = tid / CD + k CPP
    own = []
    for tid in range(ut):
        if cpp:
            if tid >= xs_:
                continue
            lt = tid % cd
            kl = tid // cd
            while kl < oc:
                own.append(kl * cd + lt)
                assert (lt // nth) * noap + kl * nth + lt % nth < nsp * noap
                kl += cpp
        else:
            own.extend(range(tid, oc * cd, ut))
    assert sorted(own) == list(range(oc * cd))
    # traces and flux slots: (link li, slot le), li = tid / TRL + u LPP
    trl, maxl = n * nth, 4 * oc
    lpp = ut // trl
    ts = lpp * trl
    tu = (maxl * trl + ts - 1) // ts
    slots = []
    for tid in range(ut):
        e0 = tid if tid < ts else maxl * trl
        for u in range(tu):
            li = e0 // trl + u * lpp
            if li < maxl:
                slots.append(li * trl + e0 % trl)
    assert sorted(slots) == list(range(maxl * trl))
    # phase-2 lines (q or p, part angle)
    nline = oc * n * nth
    lines = sorted((e // noa, e % noa) for e in range(nline))
    assert lines == [(q, o) for q in range(n) for o in range(noa)]


@pytest.mark.parametrize("n,nth,nc", [b for b in BINS if b[1] % 2 == 1])
def test_cell_walk_hits_distinct_banks(n, nth, nc):
    """16 consecutive DOFs of one cell (node-major, angle fastest) land on
    16 distinct 8-byte bank pairs with the padded stride; with the plain odd
    stride na | 1 they do not (so the padding is what removes the conflict)."""
    nsp, na = n * n, nc * nth
    cd = nsp * nth

    def worst(stride):
        w = 0
        for start in range(0, cd - 15):
            banks = [((lt // nth) * stride + lt % nth) % 16 for lt in range(start, start + 16)]
            w = max(w, 16 - len(set(banks)) + 1)
        return w

    assert worst(u_pad(na, nth)) == 1
    assert worst(na | 1) > 1


@pytest.mark.parametrize("n,nth,nc,h", CASES)
def test_lds_fits_three_workgroups_for_the_c2_bin(n, nth, nc, h):
    """Dynamic LDS of an FP64 launch at the largest link count a part may
    have (4 per cell), plus about 4.2 KB of static arrays."""
    nsp, na, oc = n * n, nc * nth, nc // h
    lds = (nsp * u_pad(na, nth) + nsp * u_pad(oc * nth, nth) + 2 * 4 * oc * n * nth) * 8 + 4400
    assert lds <= 160 * 1024
    if (n, nth, nc, h) == (5, 5, 16, 1):
        # the C2 bin has at most 2 links per cell: 3 workgroups per CU
        lds2 = (nsp * u_pad(na, nth) * 2 + 2 * 2 * oc * n * nth) * 8 + 4400
        assert 3 * lds2 <= 160 * 1024


@pytest.mark.parametrize("n,nth,nc", BINS)
def test_bf16_tile_padding_and_operands(n, nth, nc):
    """The mixed path's BF16 column tile: phase 1 and the padding loops
    write every element the MFMA operand reads exactly once, the padding is
    exactly the region past NA angles or NSP nodes, and every operand read
    (four consecutive angles) is 8-byte aligned inside the tile."""
    nsp, na = n * n, nc * nth
    mt, nkb = (nsp + 15) // 16, (na + 15) // 16
    napb = (na + 15) // 16 * 16 + 4
    assert napb % 4 == 0 and napb >= nkb * 16
    written = {}
    for sn in range(nsp):
        for col in range(na):
            written[(sn, col)] = written.get((sn, col), 0) + 1
    pc, pr = nkb * 16 - na, (mt * 16 - nsp) * nkb * 16
    for e in range(nsp * pc):
        key = (e // pc, na + e % pc)
        written[key] = written.get(key, 0) + 1
    for e in range(pr):
        key = (nsp + e // (nkb * 16), e % (nkb * 16))
        written[key] = written.get(key, 0) + 1
    assert all(v == 1 for v in written.values())
    assert set(written) == {(r, c) for r in range(mt * 16) for c in range(nkb * 16)}
    for t in range(mt):
        for ks in range(nkb):
            for lane in range(64):
                s, lg = t * 16 + lane % 16, lane // 16
                first = s * napb + ks * 16 + 4 * lg
                assert (first * 2) % 8 == 0
                assert all((s, ks * 16 + 4 * lg + u) in written for u in range(4))


def ue_waves(nth, nc):
    return max((nc * nth + 15) // 16, 4)


@pytest.mark.parametrize("n,nth,nc", BINS)
def test_output_layout_kernel_indices(n, nth, nc):
    """k_uniform_e: lane (l16, lg) of wave w < NTILE owns output angle
    R = 16 w + l16 and node rows s = 16 t + lg + 4 v (FP64 MFMA layout) or
    16 t + 4 lg + v (FP32 / BF16); its outputs, staged in the column tile at
    (cell, node, angle) = kk CD + s nth + r, cover the column's DOFs exactly
    once and fit the tile; the derivative reads of the x- and y-lines stay
    inside it; the first-round rule fits the LDS of four workgroups per CU
    (C2 shape, FP64: column tile + traces and flux slots of 2 NC links)."""
    nsp, na = n * n, nc * nth
    ncd, cd, nap = nsp * na, nsp * nth, u_pad(na, nth)
    mt, ntile = (nsp + 15) // 16, (na + 15) // 16
    waves = ue_waves(nth, nc)
    assert ntile <= waves and nsp * nap >= ncd
    for f64 in (True, False):
        staged = []
        for w in range(ntile):
            for lane in range(64):
                l16, lg = lane % 16, lane // 16
                R = 16 * w + l16
                if R >= na:
                    continue
                kk, r = R // nth, R % nth
                for t in range(mt):
                    for v in range(4):
                        s = 16 * t + ((lg + 4 * v) if f64 else (4 * lg + v))
                        if s >= nsp:
                            continue
                        p, q = s // n, s % n
                        for i in range(n):
                            assert 0 <= (i * n + q) * nap + R < nsp * nap
                            assert 0 <= (p * n + i) * nap + R < nsp * nap
                        staged.append(kk * cd + s * nth + r)
        assert sorted(staged) == list(range(ncd))
    if (n, nth, nc) == (5, 5, 16):
        trl, links = n * nth, 2 * nc
        dyn = nsp * nap * 8 + 2 * links * trl * 8
        static = 4112                    # the FP64 instance's static LDS (hipcc resource report)
        assert 4 * (dyn + static) <= 160 * 1024
