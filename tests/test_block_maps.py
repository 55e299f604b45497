"""CPU replicas of the device index maps, checked for what the kernels
assume of them (no GPU needed):

- the colour enumeration (color_cell, nim_stencil.h) and the skewed k-slab
  schedule (build_schedule, nim_api.hip): every cell exactly once, in its
  colour, launches that share a launch at least two planes apart, and
  every colour phase after the lower colours on its slab and the slabs next
  to it (the causality that makes the slab schedule bitwise the colour
  order);
- the radial-slab block maps (cell_of_block with rslab, nim_kernels.h;
  sched_cell with Sched.rsb, nim_kernels.hip): bijections onto the cells
  (and onto each colour's cells), every block of XCD x inside rows
  [x n_r / 8, (x + 1) n_r / 8), under the round-robin dispatch that
  xcd_block() renumbers;
- the LDS tiles of the energy sources (load_tile / cross_diffusion_tile,
  nim_stencil.h): every LDS entry a cell's derivatives read was staged;
- the two-cells-per-thread source kernels (pair_of_block, nim_kernels.h;
  cell_pair / other3, nim_stencil.h): every cell exactly once, the two cells
  of a thread at azimuthal indices i and i + 2 of one row, every pair load
  aligned and inside its row, and the values picked from the two
  other-half pairs those at i - 1, i + 1 and i + 3.
"""
import itertools

import pytest

NXCD = 8


def xcd_block(b, nb):
    q, rem = nb // NXCD, nb % NXCD
    x, r = b % NXCD, b // NXCD
    return x * q + (x if x < rem else rem) + r


def color_cell(NT, NR, color, k0, k1, tid):
    nth, nrh = NT >> 1, (NR + 1) >> 1
    ih, rest = tid % nth, tid // nth
    jj, k = rest % nrh, k0 + rest // nrh
    if k >= k1:
        return None
    j = ((color + k) & 1) + 2 * jj
    if j >= NR:
        return None
    r = (color - j - 3 * k) & 3
    return (2 * ih + (r >> 1), j, k)


def build_schedule(NT, NR, NZ, ks, cpb):
    S = (NZ + ks - 1) // ks
    per_plane = (NT // 2) * ((NR + 1) // 2)
    out = []
    for L in range(S + 6):
        items, blocks = [], 0
        for col in range(4):
            sl = L - 2 * col
            if 0 <= sl < S:
                k0, k1 = sl * ks, min(NZ, (sl + 1) * ks)
                items.append((col, k0, k1, blocks))
                blocks += (per_plane * (k1 - k0) + cpb - 1) // cpb
        if items:
            out.append((items, blocks))
    return out


@pytest.mark.parametrize("dims", [(36, 6, 12), (48, 24, 16), (16, 5, 7)])
@pytest.mark.parametrize("ks", [2, 3, "all"])
def test_slab_schedule_covers_and_orders(dims, ks):
    NT, NR, NZ = dims
    ks = NZ if ks == "all" else ks
    for cpb in (64, 256):
        seen, when = {}, {}
        sch = build_schedule(NT, NR, NZ, ks, cpb)
        for L, (items, blocks) in enumerate(sch):
            for lb in range(blocks):
                q = max(t for t in range(len(items)) if lb >= items[t][3])
                col, k0, k1, bs = items[q]
                for tx in range(cpb):
                    c = color_cell(NT, NR, col, k0, k1, (lb - bs) * cpb + tx)
                    if c is None:
                        continue
                    assert c not in seen
                    assert (2 * c[0] + c[1] + 3 * c[2]) % 4 == col
                    seen[c], when[c] = col, L
            for a, b in itertools.combinations([(it[1], it[2]) for it in items], 2):
                assert max(a[0], b[0]) - min(a[1], b[1]) >= 2, (a, b)
        assert len(seen) == NT * NR * NZ
        phase = {}
        for (i, j, k), col in seen.items():
            phase.setdefault((k, col), []).append(when[(i, j, k)])
        for (k, col), ws in phase.items():
            for k2 in range(max(0, k - 2), min(NZ, k + 3)):
                for col2 in range(4):
                    if (k2, col2) not in phase:
                        continue
                    if col2 < col:
                        assert max(phase[(k2, col2)]) < min(ws)
                    if col2 > col:
                        assert min(phase[(k2, col2)]) > max(ws)


@pytest.mark.parametrize("dims", [(256, 128, 2), (128, 64, 3), (512, 256, 1)])
def test_radial_slab_cell_map(dims):
    """cell_of_block with rslab (set where n_r % 8 == 0 and the slab-plane
    n_theta n_r / 8 is a multiple of the block size) for 64- and 256-cell
    blocks."""
    NT, NR, NZ = dims
    N = NT * NR * NZ
    for bs in (64, 256):
        sp = NT * (NR // NXCD)
        assert NR % NXCD == 0 and sp % bs == 0
        nb, bps = N // bs, sp // bs
        q = NZ * bps
        cells = []
        for b in range(nb):
            lb = xcd_block(b, nb)
            x, r = lb // q, lb % q
            k = r // bps
            base = k * NT * NR + x * sp + (r - k * bps) * bs
            blk = range(base, base + bs)
            assert all(((c // NT) % NR) // (NR // NXCD) == b % NXCD for c in blk)
            cells.extend(blk)
        assert sorted(cells) == list(range(N))


@pytest.mark.parametrize("dims", [(256, 128, 2), (128, 64, 3), (64, 32, 3)])
def test_radial_slab_colour_map(dims):
    """sched_cell with Sched.rsb: each colour's cells once, XCD x on its
    radial slab."""
    NT, NR, NZ = dims
    nth, nrh = NT // 2, (NR + 1) // 2
    for bs in (64, 256):
        if nrh % NXCD or (nth * (nrh // NXCD)) % bs:
            continue
        rsb = nth * (nrh // NXCD) // bs
        nb = (nth * nrh * NZ) // bs
        q = nb // NXCD
        for color in range(4):
            seen = set()
            for b in range(nb):
                lb = xcd_block(b, nb)
                x, r = lb // q, lb % q
                kk = r // rsb
                for t in range(bs):
                    tid = kk * (nth * nrh) + x * (nth * (nrh // NXCD)) + (r - kk * rsb) * bs + t
                    c = color_cell(NT, NR, color, 0, NZ, tid)
                    if c is None:
                        continue
                    assert (2 * c[0] + c[1] + 3 * c[2]) % 4 == color
                    assert c[1] // (NR // NXCD) == b % NXCD
                    assert c not in seen
                    seen.add(c)
            want = {(i, j, k) for i in range(NT) for j in range(NR) for k in range(NZ)
                    if (2 * i + j + 3 * k) % 4 == color}
            assert seen == want


@pytest.mark.parametrize("dims,TI", [((128, 64, 3), 64), ((64, 5, 3), 64), ((64, 3, 3), 64), ((96, 7, 4), 32),
                                     ((32, 9, 3), 32)])
def test_tile_reads_hit_staged_entries(dims, TI):
    NT, NR, NZ = dims
    TJ = 256 // TI
    for k in range(NZ):
        for tj in range((NR + TJ - 1) // TJ):
            for ti in range(NT // TI):
                i0, j0 = ti * TI, tj * TJ
                R = {(jj, li) for jj in range(TJ + 1) for li in range(TI + 2) if j0 + jj <= NR}
                Tf = {(jj, li) for jj in range(TJ + 4) for li in range(TI + 1) if 0 <= j0 - 2 + jj < NR}
                Z = {(kk, jj, li) for kk in range(2) for jj in range(TJ + 4) for li in range(TI + 2)
                     if 0 <= j0 - 2 + jj < NR}
                for t in range(256):
                    li, jr = t % TI, t // TI
                    j = j0 + jr
                    if j >= NR:
                        continue
                    lo, hi = j == 0, j == NR - 1
                    ja = (0 if lo else j if hi else j + 1) - j0 + 2
                    jb = (1 if lo else j - 1) - j0 + 2
                    jc = (2 if lo else j - 2 if hi else j - 1) - j0 + 2
                    jm = jr + 2
                    for e in [(jr + 1, li + 2), (jr + 1, li), (jr, li + 2), (jr, li), (jr + 1, li + 1), (jr, li + 1)]:
                        assert e in R
                    for r_ in (ja, jb, jc, jm):
                        assert (r_, li) in Tf and (r_, li + 1) in Tf
                    for kk in (0, 1):
                        for e in [(jm, li + 2), (jm, li), (ja, li + 1), (jb, li + 1), (jc, li + 1), (jm, li + 1)]:
                            assert (kk,) + e in Z


def ipos(NT, i):
    return (i & 1) * (NT >> 1) + (i >> 1)


def iof(NT, p):
    h = NT >> 1
    return 2 * p if p < h else 2 * (p - h) + 1


def pair_of_block(NT, NR, NZ, lb, tid, bs, rslab, krev):
    bc = 2 * bs
    if rslab:
        sp = NT * (NR // NXCD)
        bps = sp // bc
        q = NZ * bps
        x, r = lb // q, lb % q
        kk = r // bps
        k = NZ - 1 - kk if krev else kk
        return k * NT * NR + x * sp + (r - kk * bps) * bc + 2 * tid
    return lb * bc + 2 * tid


def cell_pair(NT, NR, c):
    h = NT >> 1
    q = c % NT
    odd = q >= h
    qq = q - h if odd else q
    ob = 0 if odd else h
    lo = ob + (qq if odd else (qq - 2 + h) % h)
    hi = ob + ((qq + 2) % h if odd else qq)
    return iof(NT, q), (c // NT) % NR, c // (NT * NR), q, lo, hi, odd


@pytest.mark.parametrize("dims", [(256, 128, 2), (48, 24, 3), (36, 6, 4), (12, 4, 4), (128, 64, 3), (20, 5, 3), (4, 3, 4),
                                  (8, 8, 4), (512, 64, 1)])
@pytest.mark.parametrize("krev", [False, True])
def test_pair_map_and_neighbours(dims, krev):
    """pair_of_block under the round-robin dispatch, with radial slabs where
    the host enables them (n_r % 8 == 0 and slab-planes of whole 512-cell
    blocks) and in storage order otherwise; cell_pair's positions."""
    NT, NR, NZ = dims
    assert NT % 4 == 0
    N, bs, h = NT * NR * NZ, 256, NT >> 1
    rslab = NR % NXCD == 0 and (NT * (NR // NXCD)) % (2 * bs) == 0
    nb = (N + 2 * bs - 1) // (2 * bs)
    seen = []
    for b in range(nb):
        lb = xcd_block(b, nb) if rslab else b
        for t in range(bs):
            c = pair_of_block(NT, NR, NZ, lb, t, bs, rslab, krev and rslab)
            if c >= N:
                continue
            assert c % 2 == 0 and c + 1 < N
            i, j, k, q, lo, hi, odd = cell_pair(NT, NR, c)
            # both cells in one row, azimuthal indices i and i + 2
            assert (c + 1) // NT == c // NT and iof(NT, q + 1) == i + 2 and (q < h) == (q + 1 < h)
            # aligned pairs of the other half-row
            for x in (lo, hi):
                assert x % 2 == 0 and 0 <= x and x + 1 < NT and (x < h) == odd
            la, lb_, ha, hb = lo, lo + 1, hi, hi + 1
            m1, p1, p3 = (la, lb_, ha) if odd else (lb_, ha, hb)
            assert m1 == ipos(NT, (i - 1) % NT)
            assert p1 == ipos(NT, (i + 1) % NT)
            assert p3 == ipos(NT, (i + 3) % NT)
            seen.extend((c, c + 1))
    assert sorted(seen) == list(range(N))
