"""The x-merge bookkeeping of the FP32 tent kernels, modelled on the CPU.

The cells kernels (hpc-mpm-code_amd/csrc/fast32.hip, xmerge_store) work on
groups of G consecutive non-empty cells; a cell whose left neighbour is the
previous cell of its group adds that cell's right-corner partials into its
own left-corner ones, and stores its right-corner partials only when its right
neighbour does not take them. The node passes (node_rows4) decide from the
occupancy bits and the binning's group-start bits which stored values a node
adds. This model applies both rules to random occupancies, grid rows that do
and do not fill a group, and every group size, and checks that each node gets
every corner contribution of its (up to 8) cells exactly once and that no
value is read that was not stored. Integer contributions make the sums exact.
"""
import random


def run(nex, ney, nez, G, p_occ, seed):
    rnd = random.Random(seed)
    nel = nex*ney*nez
    occ = [rnd.random() < p_occ for _ in range(nel)]
    active = [c for c in range(nel) if occ[c]]
    gs = [False]*nel
    for a, c in enumerate(active):
        if a % G == 0: gs[c] = True
    # per cell corner values: C[c][(ii,jj,kk)] random ints (exact sums)
    C = {c: {(ii,jj,kk): rnd.randint(1, 1000) for ii in (0,1) for jj in (0,1) for kk in (0,1)} for c in active}
    stored = {}  # (c, ii, jj, kk) -> value
    for g0 in range(0, len(active), G):
        grp = active[g0:g0+G]
        for p, c in enumerate(grp):
            ix = c % nex
            pc = grp[p-1] if p > 0 else None
            nc = grp[p+1] if p + 1 < len(grp) else None
            merge = p > 0 and pc == c - 1 and ix > 0
            keep_right = not (nc is not None and nc == c + 1 and ix + 1 < nex)
            for jj in (0,1):
                for kk in (0,1):
                    v = C[c][(0,jj,kk)] + (C[pc][(1,jj,kk)] if merge else 0)
                    stored[(c,0,jj,kk)] = v
                    if keep_right: stored[(c,1,jj,kk)] = C[c][(1,jj,kk)]
    reads = 0
    for k in range(nez+1):
        for j in range(ney+1):
            for i in range(nex+1):
                brute = 0
                for dk in (0,1):
                    for dj in (0,1):
                        for di in (0,1):
                            ci, cj, ck = i-1+di, j-1+dj, k-1+dk
                            if 0 <= ci < nex and 0 <= cj < ney and 0 <= ck < nez:
                                c = ci + nex*(cj + ney*ck)
                                if occ[c]: brute += C[c][(1-di,1-dj,1-dk)]
                tot = 0
                for r in range(4):
                    cj, ck = j-1+(r&1), k-1+(r>>1)
                    row = 0 <= cj < ney and 0 <= ck < nez
                    base = (cj + ck*ney)*nex if row else 0
                    in_l = row and i >= 1
                    in_r = row and i < nex
                    cl = base + i - 1 if in_l else 0
                    cr = base + i if in_r else 0
                    on_l = in_l and occ[cl]
                    on_r = in_r and occ[cr]
                    merged = on_l and on_r and not gs[cr]
                    jj, kk = 1-(r&1), 1-(r>>1)
                    if on_l and not merged:
                        tot += stored[(cl,1,jj,kk)]; reads += 1
                    if on_r:
                        tot += stored[(cr,0,jj,kk)]; reads += 1
                assert tot == brute, (i,j,k,tot,brute)
    nst = len(stored)
    return nst / max(1, len(active)), reads


def test_merge_and_read_rules_cover_every_corner_once():
    for seed in range(40):
        nex = random.Random(seed).choice([1, 2, 3, 5, 7, 16, 33])
        for G in (64, 32, 16, 8, 1):   # G = 1: every cell starts a group (no merge)
            for p in (0.2, 0.7, 1.0):
                run(nex, 3, 2, G, p, seed)


def test_full_rows_store_about_half():
    # 64-cell rows in groups of 32: left runs always, right runs at group ends
    stored_per_cell, _ = run(64, 4, 4, 32, 1.0, 1)
    assert stored_per_cell == 4 + 4 / 32


# ---------------------------------------------------------------- stencils
# The stencil kernels (hpc-mpm-code_amd/csrc/stencil32.hip: x_links,
# x_merge_row, row_reads) apply the same rule to NX corner columns
# (NX = 2 or 3, the quadratic B-spline) with any group size (21 cells in the
# quadratic P2G, one z-layer of the stencil per lane): S[NX-1] = col[NX-1],
# S[ii] = col[ii] + S[ii+1] of a linked left neighbour; S[0] always stored,
# S[ii >= 1] only when the right neighbour does not take them; a control
# point adds S[d] of the cell d to its left when the cell right of that one
# did not take it.
def run_stencil(nex, ney, nez, NX, G, p_occ, seed):
    rnd = random.Random(seed)
    nel = nex * ney * nez
    occ = [rnd.random() < p_occ for _ in range(nel)]
    active = [c for c in range(nel) if occ[c]]
    gs = [False] * nel
    for a, c in enumerate(active):
        if a % G == 0:
            gs[c] = True
    # one (y, z) corner row per cell suffices: the rule acts along x only
    col = {c: [rnd.randint(1, 1000) for _ in range(NX)] for c in active}
    stored = {}
    for g0 in range(0, len(active), G):
        grp = active[g0:g0 + G]
        S_prev = None
        for u, c in enumerate(grp):
            ix = c % nex
            pc = grp[u - 1] if u > 0 else None
            nc = grp[u + 1] if u + 1 < len(grp) else None
            link = u >= 1 and pc == c - 1 and ix > 0
            keep_right = not (u + 1 < G and nc is not None and nc == c + 1 and ix + 1 < nex)
            S = list(col[c])
            for ii in range(NX - 2, -1, -1):
                if link:
                    S[ii] += S_prev[ii + 1]
            for ii in range(NX):
                if ii == 0 or keep_right:
                    stored[(c, ii)] = S[ii]
            S_prev = S
    reads = 0
    for cj in range(ney):
        for ck in range(nez):
            base = (cj + ck * ney) * nex
            for i in range(nex + NX - 1):
                brute = sum(col[base + i - d][d] for d in range(NX)
                            if 0 <= i - d < nex and occ[base + i - d])
                xlo, xhi = max(i - (NX - 1), 0), min(i, nex - 1)
                tot = 0
                for d in range(NX - 1, -1, -1):
                    cx = i - d
                    if not (xlo <= cx <= xhi) or not occ[base + cx]:
                        continue
                    taken = d > 0 and cx + 1 <= xhi and occ[base + cx + 1] and not gs[base + cx + 1]
                    if not taken:
                        tot += stored[(base + cx, d)]
                        reads += 1
                assert tot == brute, (i, cj, ck, tot, brute)
    return len(stored) / max(1, len(active)), reads


def test_stencil_merge_and_read_rules_cover_every_column_once():
    for seed in range(30):
        nex = random.Random(seed).choice([1, 2, 3, 4, 7, 21, 22, 64])
        for NX in (2, 3):
            for G in (64, 21, 16, 3, 2, 1):
                for p in (0.15, 0.6, 1.0):
                    run_stencil(nex, 3, 2, NX, G, p, seed)


def test_stencil_full_rows_store_about_a_third():
    # 64-cell rows, 21-cell groups: column 0 always, columns 1 and 2 at run ends
    stored_per_cell, _ = run_stencil(64, 4, 4, 3, 21, 1.0, 1)
    assert stored_per_cell < 1.25
