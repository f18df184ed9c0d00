"""The index arithmetic of the structured (grid) assembly, checked on the CPU
against the CSR pattern and the element DOF lists of iron's generated mesh.

grid_gather_kernel (iron_amd/csrc/mife_assemble.hip) writes CSR row `row`
entry by entry from the staged element matrices, with the column of entry p,
the elements that meet there and their local indices all computed from the
node's grid index.  This module restates that arithmetic line by line in
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


def gather_plan(ne, P, NC):
    """[(row, p, column, [(e, a_local_dof, b_local_dof), ...])] as the kernel computes them."""
    nx, ny, nz = ne
    NB = P + 1
    NN = NB ** 3
    Nx, Ny, Nz = P * nx + 1, P * ny + 1, P * nz + 1
    Nn = Nx * Ny * Nz
    out = []
    for row in range(NC * Nn):
        c = row // Nn
        node = row - c * Nn
        ix, iy, iz = node % Nx, (node // Nx) % Ny, node // (Nx * Ny)
        lx, hx = elem_range(ix, P, nx)
        ly, hy = elem_range(iy, P, ny)
        lz, hz = elem_range(iz, P, nz)
        jx0, jy0, jz0 = P * lx, P * ly, P * lz
        cx, cy, cz = P * hx + P - jx0 + 1, P * hy + P - jy0 + 1, P * hz + P - jz0 + 1
        cnt = cx * cy * cz
        for p in range(NC * cnt):
            c2, q = p // cnt, p % cnt
            jx, jy, jz = jx0 + q % cx, jy0 + (q // cx) % cy, jz0 + q // (cx * cy)
            ljx, hjx = elem_range(jx, P, nx)
            ljy, hjy = elem_range(jy, P, ny)
            ljz, hjz = elem_range(jz, P, nz)
            ax, bx = max(lx, ljx), min(hx, hjx)
            ay, by = max(ly, ljy), min(hy, hjy)
            az, bz = max(lz, ljz), min(hz, hjz)
            contrib = []
            for col in range(8):
                ex = elem_of_parity(ax, bx, col & 1)
                ey = elem_of_parity(ay, by, (col >> 1) & 1)
                ez = elem_of_parity(az, bz, col >> 2)
                if ex < 0 or ey < 0 or ez < 0:
                    continue
                e = ex + nx * (ey + ny * ez)
                a = (ix - P * ex) + NB * ((iy - P * ey) + NB * (iz - P * ez))
                b = (jx - P * ex) + NB * ((jy - P * ey) + NB * (jz - P * ez))
                contrib.append((e, c * NN + a, c2 * NN + b))
            out.append((row, p, c2 * Nn + jx + Nx * (jy + Ny * jz), contrib))
    return out


@pytest.mark.parametrize("ne,P,NC", [((3, 2, 2), 1, 1), ((2, 3, 2), 2, 1), ((3, 2, 2), 1, 3), ((1, 1, 2), 2, 1),
                                     ((1, 2, 1), 1, 3)])
def test_grid_gather_index_arithmetic(ne, P, NC):
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
    plan = gather_plan(ne, P, NC)
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
        for e, a, b in contrib:
            assert 0 <= e < E and 0 <= a < ND and 0 <= b < ND
            assert e * ND * ND + a * ND + b < E * ND * ND
        seen += len(contrib)
    assert np.array_equal(count, lengths)
    assert seen == E * ND * ND                       # every element entry lands once
