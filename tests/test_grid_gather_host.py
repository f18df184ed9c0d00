"""The index arithmetic of the structured (grid) assembly, checked on the CPU
against the CSR pattern and the element DOF lists of iron's generated mesh.

grid_gather_kernel (iron_amd/csrc/mife_assemble.hip) writes the CSR rows of
one x-line of grid nodes per wavefront from the staged element matrices: each
lane takes one (node, x-neighbour) pair, the y/z neighbour blocks are walked in
order, and the column of an entry, the elements that meet there and their
local indices all follow from the grid indices.  This module restates that arithmetic line by line in
Python and checks, on small Q1 and Q2 grids with one and three components,
that
  * every row's entry count equals its CSR length (no entry is skipped or
    written past the row), and the column of entry p is the pattern's;
  * the (element, local row, local column) contributions of every entry are
    exactly the element entries that the position map sends there (the
    coloured scatter's), each once, in increasing colour order;
  * every staged index lies inside the stage buffer.
"""
import numpy as np
import pytest

from oracle import fe_oracle as fo


def elem_range(i, P, n):
    lo = (i - 1) // P if i > 0 else 0
    hi = min(i // P, n - 1)
    return lo, hi


def elem_of_parity(lo, hi, bit):
    if (lo & 1) == bit:
        return lo
    if hi != lo and (hi & 1) == bit:
        return hi
    return -1


def x_pair(t, P):
    per = P * P + 2 * P
    tp = t + P
    e, r = tp // per, tp % per
    if r < 2 * P + 1:
        ix = P * e
        return ix, ix - P + r
    r2 = r - (2 * P + 1)
    k = r2 // (P + 1)
    return P * e + 1 + k, P * e + (r2 - k * (P + 1))


def x_pair_start(ix, P, nx):
    if ix > P * nx:
        return nx * (P * P + 2 * P) + 1
    e, k = ix // P, ix % P
    tp = e * (P * P + 2 * P) + ((2 * P + 1 + (k - 1) * (P + 1)) if k else 0)
    return max(tp - P, 0)


def stage_pos(e, m, n, ne, P, NC):
    """grid_stage_pos: [ez][ey][c][az][ay][c2][bz][by][ex][ax][bx]."""
    nx, ny, _ = ne
    NB = P + 1
    NN = NB ** 3
    c, a = divmod(m, NN)
    c2, b = divmod(n, NN)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    U = ((((((ez * ny + ey) * NC + c) * NB + a // (NB * NB)) * NB + (a // NB) % NB) * NC + c2) * NB
         + b // (NB * NB)) * NB + (b // NB) % NB
    return (U * nx + ex) * NB * NB + (a % NB) * NB + b % NB


def gather_plan(ne, P, NC, rp, lanes=64):
    """[(row, p, column, [(e, a_local_dof, b_local_dof), ...])] as the kernel computes them:
    one wavefront per (component, x-line, run of R nodes), lanes over the run's
    x-pairs, the y/z neighbour blocks walked in order, the stage index formed as
    U * stripe + X.  Each (c2, qz) round's sums go to LDS slots and the store plan
    sends slot k to the CSR; the plan checks that the two agree."""
    nx, ny, nz = ne
    NB = P + 1
    NN = NB ** 3
    ND = NC * NN
    E = nx * ny * nz
    Nx, Ny, Nz = P * nx + 1, P * ny + 1, P * nz + 1
    Nn = Nx * Ny * Nz
    PER = P * P + 2 * P
    R = P * (lanes // PER)
    Q = 2 * P + 1
    SU = NC * NC * NB ** 4
    runs = -(-Nx // R)
    stripe = nx * NB * NB
    inverse = {stage_pos(e, m, n, ne, P, NC): (e, m, n) for e in range(E) for m in range(ND) for n in range(ND)}
    assert sorted(inverse) == list(range(E * ND * ND))          # the layout is a permutation
    out = []
    for unit in range(NC * Ny * Nz * runs):
        run, line = unit % runs, unit // runs
        iy, iz, c = line % Ny, (line // Ny) % Nz, line // (Ny * Nz)
        r0, r1 = run * R, min(run * R + R, Nx)
        rowline = (c * Nz + iz) * Ny * Nx + iy * Nx
        t0 = x_pair_start(r0, P, nx)
        npairs = x_pair_start(r1, P, nx) - t0
        assert npairs <= lanes
        ly, hy = elem_range(iy, P, ny)
        lz, hz = elem_range(iz, P, nz)
        cy, cz = P * (hy - ly) + P + 1, P * (hz - lz) + P + 1
        # store plan: slot k -> CSR position at round (c2, qz)
        fplan = []
        for k in range(Q * lanes):
            if k >= cy * npairs:
                fplan.append(None)
                continue
            fix, _ = x_pair(t0 + k // cy, P)
            flx, fhx = elem_range(fix, P, nx)
            fplan.append((rp[rowline + fix] + k - cy * (x_pair_start(fix, P, nx) - t0), P * (fhx - flx) + P + 1,
                          rp[rowline + fix + 1]))
        for c2 in range(NC):
            for qz in range(cz):
                jz = P * lz + qz
                ljz, hjz = elem_range(jz, P, nz)
                az, bz = max(lz, ljz), min(hz, hjz)
                slots = {}
                for lane in range(npairs):
                    ix, jx = x_pair(t0 + lane, P)
                    assert r0 <= ix < r1
                    lx, hx = elem_range(ix, P, nx)
                    ljx, hjx = elem_range(jx, P, nx)
                    cx, qx = P * (hx - lx) + P + 1, jx - P * lx
                    ax, bx = max(lx, ljx), min(hx, hjx)
                    X = []
                    for s in range(2):
                        ex = elem_of_parity(ax, bx, s)
                        X.append(None if ex < 0 else (ex * NB + (ix - P * ex)) * NB + (jx - P * ex))
                    row = rowline + ix
                    wslot = cy * (x_pair_start(ix, P, nx) - t0) + qx
                    for qy in range(cy):
                        jy = P * ly + qy
                        ljy, hjy = elem_range(jy, P, ny)
                        ay, by = max(ly, ljy), min(hy, hjy)
                        contrib = []
                        for sz in range(2):
                            ez = elem_of_parity(az, bz, sz)
                            if ez < 0:
                                continue
                            for sy in range(2):
                                ey = elem_of_parity(ay, by, sy)
                                if ey < 0:
                                    continue
                                U = ((ez * ny + ey) * SU + (((((c * NB + (iz - P * ez)) * NB + (iy - P * ey)) * NC
                                                              + c2) * NB + (jz - P * ez)) * NB + (jy - P * ey)))
                                for sx in range(2):
                                    if X[sx] is not None:
                                        contrib.append(inverse[U * stripe + X[sx]])
                        p = ((c2 * cz + qz) * cy + qy) * cx + qx
                        slot = wslot + qy * cx
                        assert slot not in slots and slot < Q * lanes
                        slots[slot] = rp[row] + p
                        out.append((row, p, c2 * Nn + jx + Nx * (jy + Ny * jz), contrib))
                # the store plan sends every filled slot where its entry belongs, and no other
                blk = (c2 * cz + qz) * cy
                stored = {k: f[0] + blk * f[1] for k, f in enumerate(fplan) if f is not None and f[0] + blk * f[1] < f[2]}
                assert stored == slots
    return out


@pytest.mark.parametrize("ne,P,NC,lanes", [((3, 2, 2), 1, 1, 64), ((2, 3, 2), 2, 1, 64), ((3, 2, 2), 1, 3, 64),
                                           ((1, 1, 2), 2, 1, 64), ((1, 2, 1), 1, 3, 64), ((5, 2, 1), 2, 1, 16),
                                           ((4, 1, 2), 1, 3, 6), ((9, 1, 1), 2, 1, 64), ((23, 1, 1), 1, 3, 64)])
def test_grid_gather_index_arithmetic(ne, P, NC, lanes):
    ob = fo.Basis(3, P, P + 1)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0] * NC)
    rp, col = fo.csr_pattern(od)
    ed = od.element_dofs()                       # [E][NC NN], component-major
    E, ND = ed.shape
    # the coloured scatter: colour = parity of the element index per direction
    nx, ny, nz = ne
    eidx = np.arange(E)
    ex, ey, ez = eidx % nx, (eidx // nx) % ny, eidx // (nx * ny)
    colour = (ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2)
    scatter = {}
    for e in np.argsort(colour, kind="stable"):
        for i in range(ND):
            for j in range(ND):
                scatter.setdefault((int(ed[e, i]), int(ed[e, j])), []).append((int(e), i, j))
    plan = gather_plan(ne, P, NC, rp, lanes)
    n_rows = len(rp) - 1
    assert n_rows == NC * int(np.prod([P * n + 1 for n in ne]))
    seen = 0
    lengths = np.diff(rp)
    count = np.zeros(n_rows, dtype=np.int64)
    for row, p, column, contrib in plan:
        count[row] += 1
        assert p < lengths[row]
        assert col[rp[row] + p] == column
        assert contrib == scatter[(row, column)]     # same elements, same order
        seen += len(contrib)
    assert np.array_equal(count, lengths)
    assert seen == E * ND * ND                       # every element entry lands once
    assert len({(r, p) for r, p, _, _ in plan}) == len(plan)   # no entry written twice


@pytest.mark.parametrize("P", [1, 2])
@pytest.mark.parametrize("nx", [1, 2, 3, 7])
def test_x_pairs_enumerate_each_coupling_once(P, nx):
    """x_pair(t) over t < nx (P^2 + 2P) + 1 lists every coupled (ix, jx) of an x-line
    once, ordered by ix then jx."""
    Nx = P * nx + 1
    want = []
    for ix in range(Nx):
        lx, hx = elem_range(ix, P, nx)
        want += [(ix, jx) for jx in range(P * lx, P * hx + P + 1)]
    got = [x_pair(t, P) for t in range(nx * (P * P + 2 * P) + 1)]
    assert got == want
    # x_pair_start(ix) is the first pair of node ix; one past the last node, the count
    starts = [x_pair_start(ix, P, nx) for ix in range(Nx + 1)]
    assert starts == [next(t for t, (i, _) in enumerate(got) if i == ix) for ix in range(Nx)] + [len(got)]


@pytest.mark.parametrize("ne,P,NC", [((3, 2, 2), 1, 1), ((2, 3, 2), 2, 1), ((3, 2, 4), 1, 3)])
def test_grid_element_dofs_are_the_generated_meshs(ne, P, NC):
    """The host check that admits a structured assembly rebuilds the generated mesh's
    element DOFs; they must be the oracle's DofMap of the same regular mesh."""
    from iron_amd import fem
    ob = fo.Basis(3, P, P + 1)
    om = fo.regular_mesh(ne, [ob], (1.0, 1.0, 1.0))
    od = fo.DofMap(om, [0] * NC)
    assert np.array_equal(fem.grid_element_dofs(ne, P, NC), od.element_dofs())


@pytest.mark.parametrize("ne,P,NC", [((3, 2, 2), 1, 1), ((2, 3, 2), 2, 1), ((3, 2, 4), 1, 3), ((5, 1, 3), 2, 1)])
def test_element_kernels_stripe_store_is_the_stage_layout(ne, P, NC):
    """GridStripe: the element kernels put entry (m, n) of wavefront w's element at LDS
    slot (u WPB + w) NB^2 + ax NB + bx and the block store sends slot f to
    base(e) + u nx NB^2 + j; the composition must be the stage layout."""
    nx, ny, nz = ne
    NB = P + 1
    NN = NB ** 3
    ND = NC * NN
    B2 = NB * NB
    SU = NC * NC * B2 * B2
    WPB = 4
    for e in range(nx * ny * nz):
        w = e % WPB
        base = ((e // nx) * SU * nx + e % nx) * B2
        for m in range(ND):
            for n in range(ND):
                c, a = divmod(m, NN)
                c2, b = divmod(n, NN)
                u = ((((c * NB + a // B2) * NB + (a // NB) % NB) * NC + c2) * NB + b // B2) * NB + (b // NB) % NB
                f = (u * WPB + w) * B2 + (a % NB) * NB + b % NB
                fu, r = divmod(f, WPB * B2)
                fw, j = divmod(r, B2)
                assert (fu, fw) == (u, w)
                assert base + fu * nx * B2 + j == stage_pos(e, m, n, ne, P, NC)
