"""Seeded synthetic meshes and the gmsh MSH 4.1 ASCII writer.

Every workload mesh of this project comes from these generators, so the
reference reader (fvhyper/src/mesh.cpp:264-617) and this library's reader see
the same file.  A generator returns a :class:`MeshArrays` whose arrays are in
*file order*: nodes, cells and boundary elements exactly as
:func:`write_msh` writes them, which is the order the reference numbers them
in (mesh.cpp:356-413, 418-483, 487-570).

Random numbers come from an in-repo splitmix64 stream with
``u = (x >> 11) * 2**-53`` so that the same seed gives the same mesh on every
machine (no implementation-defined library distributions).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """The first ``n`` outputs of splitmix64 started at ``seed`` (uint64)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + _GOLDEN * (np.arange(1, n + 1, dtype=np.uint64))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """``n`` doubles in [0, 1) from :func:`splitmix64` (53-bit mantissas)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclasses.dataclass
class MeshArrays:
    """A 2-D mesh in file order.

    x, y        node coordinates (float64, length N)
    cells       (C, 4) int32 node indices, 0-based; column 3 is -1 for triangles
    bnd         (B, 2) int32 node indices of the boundary line elements
    bnd_tag     (B,) int32 index into ``names`` of each boundary element
    names       physical names of the boundary groups (BC keys)
    """

    x: np.ndarray
    y: np.ndarray
    cells: np.ndarray
    bnd: np.ndarray
    bnd_tag: np.ndarray
    names: List[str]

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])

    @property
    def n_bnd(self) -> int:
        return int(self.bnd.shape[0])


def _split_quads(q00, q10, q11, q01, diag_ac):
    """Two triangles per quad; ``diag_ac`` picks the 00-11 diagonal."""
    n = q00.shape[0]
    t = np.empty((2 * n, 4), dtype=np.int32)
    t[:, 3] = -1
    t[0::2, 0] = q00
    t[0::2, 1] = q10
    t[0::2, 2] = np.where(diag_ac, q11, q01)
    t[1::2, 0] = np.where(diag_ac, q00, q10)
    t[1::2, 1] = q11
    t[1::2, 2] = q01
    return t


def _incircle(ax, ay, bx, by, cx, cy, dx, dy):
    """> 0 when d lies inside the circumcircle of the ccw triangle abc."""
    adx, ady = ax - dx, ay - dy
    bdx, bdy = bx - dx, by - dy
    cdx, cdy = cx - dx, cy - dy
    ad = adx * adx + ady * ady
    bd = bdx * bdx + bdy * bdy
    cd = cdx * cdx + cdy * cdy
    return (adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx))


def _grid_mesh(nx, ny, x0, y0, hx, hy, keep, jitter, seed, delaunay, boundary_name):
    """Structured (nx x ny) grid of squares, masked by ``keep(ic, jc)``.

    Interior nodes move by U(-jitter, +jitter)*h per coordinate; nodes on the
    domain boundary stay put.  Each square becomes two triangles: along the
    00-11 diagonal, or, with ``delaunay``, along whichever diagonal passes the
    empty-circumcircle test for the jittered corners.  ``boundary_name(mx, my,
    nx_out, ny_out)`` names a boundary edge from its midpoint and outward
    normal.
    """
    ic, jc = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    cell_mask = keep(ic, jc)                      # (ny, nx)
    # nodes used by kept squares
    node_used = np.zeros((ny + 1, nx + 1), dtype=bool)
    node_used[:-1, :-1] |= cell_mask
    node_used[:-1, 1:] |= cell_mask
    node_used[1:, :-1] |= cell_mask
    node_used[1:, 1:] |= cell_mask
    # boundary nodes: a used node that touches a missing square (or the frame)
    padded = np.zeros((ny + 2, nx + 2), dtype=bool)
    padded[1:-1, 1:-1] = cell_mask
    touch_missing = ~(padded[:-1, :-1] & padded[:-1, 1:] & padded[1:, :-1] & padded[1:, 1:])
    on_bnd = node_used & touch_missing
    node_id = -np.ones((ny + 1, nx + 1), dtype=np.int64)
    order = np.flatnonzero(node_used.ravel())
    node_id.ravel()[order] = np.arange(order.size)
    jj, ii = np.divmod(order, nx + 1)
    x = x0 + ii * hx
    y = y0 + jj * hy
    x = x.astype(np.float64)
    y = y.astype(np.float64)
    if jitter > 0.0:
        u = uniform01(seed, 2 * order.size).reshape(order.size, 2)
        interior = ~on_bnd.ravel()[order]
        x = np.where(interior, x + (2.0 * u[:, 0] - 1.0) * jitter * hx, x)
        y = np.where(interior, y + (2.0 * u[:, 1] - 1.0) * jitter * hy, y)
    # cells, row-major over kept squares
    kj, ki = np.nonzero(cell_mask)
    q00 = node_id[kj, ki]
    q10 = node_id[kj, ki + 1]
    q11 = node_id[kj + 1, ki + 1]
    q01 = node_id[kj + 1, ki]
    if delaunay:
        # ccw triangle (00, 10, 11); 01 inside its circumcircle -> use 10-01
        inside = _incircle(x[q00], y[q00], x[q10], y[q10], x[q11], y[q11], x[q01], y[q01]) > 0
        diag_ac = ~inside
    else:
        diag_ac = np.ones(q00.shape[0], dtype=bool)
    cells = _split_quads(q00, q10, q11, q01, diag_ac)
    # boundary edges: square sides whose other side is missing
    edges = []
    tags = []
    names: List[str] = []
    sides = [  # (di, dj) of the neighbour square, node pair of the shared side, outward normal
        (0, -1, (q00, q10), (0.0, -1.0)),
        (1, 0, (q10, q11), (1.0, 0.0)),
        (0, 1, (q11, q01), (0.0, 1.0)),
        (-1, 0, (q01, q00), (-1.0, 0.0)),
    ]
    for di, dj, (na, nb), (ox, oy) in sides:
        ni, nj = ki + di, kj + dj
        outside = (ni < 0) | (ni >= nx) | (nj < 0) | (nj >= ny)
        nbr_kept = np.zeros_like(outside)
        inside_idx = ~outside
        nbr_kept[inside_idx] = cell_mask[nj[inside_idx], ni[inside_idx]]
        is_b = ~nbr_kept
        if not np.any(is_b):
            continue
        a, b = na[is_b], nb[is_b]
        # midpoint of the undisplaced side (boundary nodes are not jittered)
        mx = x0 + (ki[is_b] + (0.5 if dj != 0 else (1.0 if di > 0 else 0.0))) * hx
        my = y0 + (kj[is_b] + (0.5 if di != 0 else (1.0 if dj > 0 else 0.0))) * hy
        for k in range(a.size):
            nm = boundary_name(float(mx[k]), float(my[k]), ox, oy)
            if nm not in names:
                names.append(nm)
            edges.append((int(a[k]), int(b[k])))
            tags.append(names.index(nm))
    bnd = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
    bnd_tag = np.asarray(tags, dtype=np.int32)
    # group boundary elements by name (one line-element block per group)
    order_b = np.argsort(bnd_tag, kind="stable")
    return MeshArrays(x, y, cells, bnd[order_b], bnd_tag[order_b], names)


def sod_tube(nx: int = 400, ny: int = 40, length: float = 1.0, height: float = 0.1,
             jitter: float = 0.2, seed: int = 20240601) -> MeshArrays:
    """Shock-tube channel [0,length]x[0,height]: nx x ny squares split into
    two triangles each along a fixed diagonal, interior nodes jittered by
    U(+-jitter*h).  Top/bottom are ``wall`` (slip), ends ``transmissive``.
    The 400 x 40 default is the 32,000-cell Sod workload."""
    hx, hy = length / nx, height / ny

    def name(mx, my, ox, oy):
        return "transmissive" if ox != 0.0 else "wall"

    return _grid_mesh(nx, ny, 0.0, 0.0, hx, hy, lambda i, j: np.ones_like(i, dtype=bool),
                      jitter, seed, False, name)


def forward_step(n_per_unit: int = 440, jitter: float = 0.2, seed: int = 20240602,
                 step_x: float = 0.6, step_h: float = 0.2, length: float = 3.0,
                 height: float = 1.0, step_name: str = "wall") -> MeshArrays:
    """Mach-3 forward-facing step channel [0,3]x[0,1] minus [0.6,3]x[0,0.2].

    A square lattice of spacing 1/n_per_unit, interior nodes jittered by
    U(+-jitter*h), each square split along its Delaunay diagonal.  With the
    default 440 per unit the mean edge length is ~1/400 and the mesh has
    975,744 triangles.  Boundaries: ``inlet`` (x=0), ``outlet`` (x=3), the
    step face x=0.6 named ``step_name``, every other side ``wall``.
    """
    h = 1.0 / n_per_unit
    nx = int(round(length * n_per_unit))
    ny = int(round(height * n_per_unit))
    sx = int(round(step_x * n_per_unit))
    sy = int(round(step_h * n_per_unit))
    if abs(sx * h - step_x) > 1e-12 or abs(sy * h - step_h) > 1e-12:
        raise ValueError("step corner must fall on the lattice")

    def keep(i, j):
        return ~((i >= sx) & (j < sy))

    def name(mx, my, ox, oy):
        if ox < 0 and abs(mx) < 0.5 * h:
            return "inlet"
        if ox > 0 and abs(mx - length) < 0.5 * h:
            return "outlet"
        if ox > 0 and abs(mx - step_x) < 0.5 * h:
            return step_name
        return "wall"

    return _grid_mesh(nx, ny, 0.0, 0.0, h, h, keep, jitter, seed, True, name)


def flat_plate(nx: int = 60, ny: int = 20, length: float = 1.5, height: float = 0.5, plate_x: float = 0.25,
               jitter: float = 0.15, seed: int = 11) -> MeshArrays:
    """Flat-plate box [0,length]x[0,height] with the group names of the
    reference's flatPlate example: ``left``, ``right``, ``top``, and the
    bottom split at x = plate_x into ``bot0`` (slip, upstream) and ``bot1``
    (the plate).  Squares split into triangles along a fixed diagonal."""
    hx, hy = length / nx, height / ny

    def name(mx, my, ox, oy):
        if ox < 0:
            return "left"
        if ox > 0:
            return "right"
        if oy > 0:
            return "top"
        return "bot0" if mx < plate_x else "bot1"

    return _grid_mesh(nx, ny, 0.0, 0.0, hx, hy, lambda i, j: np.ones_like(i, dtype=bool), jitter, seed, False, name)


def mixed_channel(nx: int = 12, ny: int = 6, jitter: float = 0.15, seed: int = 7,
                  tri_fraction: float = 0.5, names=("wall", "transmissive")) -> MeshArrays:
    """A small jittered channel [0,1]x[0,0.5] of quads, a seeded fraction of
    them split into triangles: exercises 3- and 4-edge cells together."""
    m = _grid_mesh(nx, ny, 0.0, 0.0, 1.0 / nx, 0.5 / ny,
                   lambda i, j: np.ones_like(i, dtype=bool), jitter, seed, False,
                   lambda mx, my, ox, oy: names[1] if ox != 0.0 else names[0])
    tri = m.cells
    quads = np.empty((tri.shape[0] // 2, 4), dtype=np.int32)
    quads[:, 0] = tri[0::2, 0]
    quads[:, 1] = tri[0::2, 1]
    quads[:, 2] = tri[0::2, 2]
    quads[:, 3] = tri[1::2, 2]
    split = uniform01(seed + 1, quads.shape[0]) < tri_fraction
    out = []
    for k in range(quads.shape[0]):
        if split[k]:
            out.append(tri[2 * k])
            out.append(tri[2 * k + 1])
        else:
            out.append(quads[k])
    m.cells = np.asarray(out, dtype=np.int32)
    return m


# ---------------------------------------------------------------------------
# MSH 4.1 ASCII writer: exactly the subset the reference reader parses
# (mesh.cpp:264-617): single-space tokens, 1-based tags, curve entities with
# one physical tag at field 8, one node block, line blocks then cell blocks.
# ---------------------------------------------------------------------------

def write_msh(path: str, m: MeshArrays) -> None:
    nb_names = len(m.names)
    n_nodes = m.x.size
    lines: List[str] = []
    lines += ["$MeshFormat", "4.1 0 8", "$EndMeshFormat"]
    lines += ["$PhysicalNames", str(nb_names + 1)]
    for k, nm in enumerate(m.names):
        if " " in nm or '"' in nm:
            raise ValueError("physical names may not hold spaces or quotes")
        lines.append('1 %d "%s"' % (k + 1, nm))
    lines.append('2 %d "internal"' % (nb_names + 1))
    lines.append("$EndPhysicalNames")
    lines += ["$Entities", "0 %d 1 0" % nb_names]
    for k in range(nb_names):
        lines.append("%d 0 0 0 0 0 0 1 %d 0" % (k + 1, k + 1))
    lines.append("1 0 0 0 0 0 0 1 %d 0" % (nb_names + 1))
    lines.append("$EndEntities")
    lines += ["$Nodes", "1 %d 1 %d" % (n_nodes, n_nodes), "2 1 0 %d" % n_nodes]
    body = [("\n".join(str(i) for i in range(1, n_nodes + 1)))] if n_nodes else []
    coords = "\n".join("%.17g %.17g 0" % (a, b) for a, b in zip(m.x.tolist(), m.y.tolist()))
    # element blocks: one per contiguous run of boundary tag, then cells by type run
    blocks = []
    tag = 1
    b = 0
    while b < m.n_bnd:
        e = b
        while e < m.n_bnd and m.bnd_tag[e] == m.bnd_tag[b]:
            e += 1
        rows = ["%d %d %d" % (tag + r, m.bnd[b + r, 0] + 1, m.bnd[b + r, 1] + 1) for r in range(e - b)]
        blocks.append("1 %d 1 %d\n" % (int(m.bnd_tag[b]) + 1, e - b) + "\n".join(rows))
        tag += e - b
        b = e
    c = 0
    is_tri = m.cells[:, 3] < 0
    while c < m.n_cells:
        e = c
        while e < m.n_cells and is_tri[e] == is_tri[c]:
            e += 1
        if is_tri[c]:
            rows = ["%d %d %d %d" % (tag + r, m.cells[c + r, 0] + 1, m.cells[c + r, 1] + 1,
                                     m.cells[c + r, 2] + 1) for r in range(e - c)]
            blocks.append("2 1 2 %d\n" % (e - c) + "\n".join(rows))
        else:
            rows = ["%d %d %d %d %d" % (tag + r, m.cells[c + r, 0] + 1, m.cells[c + r, 1] + 1,
                                        m.cells[c + r, 2] + 1, m.cells[c + r, 3] + 1)
                    for r in range(e - c)]
            blocks.append("2 1 3 %d\n" % (e - c) + "\n".join(rows))
        tag += e - c
        c = e
    n_el = m.n_bnd + m.n_cells
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
        if n_nodes:
            f.write(body[0] + "\n" + coords + "\n")
        f.write("$EndNodes\n")
        f.write("$Elements\n%d %d 1 %d\n" % (len(blocks), n_el, n_el))
        for blk in blocks:
            f.write(blk + "\n")
        f.write("$EndElements\n")


# ---------------------------------------------------------------------------
# Partitioned MSH 4.1 for several ranks, in the layout the reference reads
# (mesh.cpp:316-340, 574-617, 803-855): rank r reads <name>_<r+1>.msh with its
# own cells, one layer of neighbour cells marked in $GhostElements with their
# owner partition, the boundary elements of its own cells, and the nodes of
# all of them.  Element and node tags are the single-file writer's (global),
# which make_comms (mesh.cpp:673-799) relies on to pair the ranks' cells.
# ---------------------------------------------------------------------------

def cell_centroids(m: MeshArrays):
    tri = m.cells[:, 3] < 0
    idx = np.where(m.cells < 0, m.cells[:, :1], m.cells)
    n = np.where(tri, 3.0, 4.0)
    cx = (m.x[idx].sum(axis=1) - np.where(tri, m.x[idx[:, 3]], 0.0)) / n
    cy = (m.y[idx].sum(axis=1) - np.where(tri, m.y[idx[:, 3]], 0.0)) / n
    return cx, cy


def partition_rcb(m: MeshArrays, parts: int) -> np.ndarray:
    """Recursive coordinate bisection of the cell centroids into `parts`
    parts of equal size (within one cell); returns the part of each cell."""
    cx, cy = cell_centroids(m)
    out = np.zeros(m.n_cells, dtype=np.int32)

    def rec(ids, p0, np_):
        if np_ == 1:
            out[ids] = p0
            return
        a = np_ // 2
        c = cx[ids] if np.ptp(cx[ids]) >= np.ptp(cy[ids]) else cy[ids]
        order = ids[np.argsort(c, kind="stable")]
        k = (len(ids) * a) // np_
        rec(order[:k], p0, a)
        rec(order[k:], p0 + a, np_ - a)
    rec(np.arange(m.n_cells), 0, parts)
    return out


def _cell_edge_keys(m: MeshArrays):
    """Sorted node-pair keys of every cell side and the side's cell."""
    tri = m.cells[:, 3] < 0
    a = []
    b = []
    for i in range(4):
        j = (i + 1) % 4
        n0 = m.cells[:, i]
        n1 = np.where(tri & (i == 2), m.cells[:, 0], m.cells[:, j])
        ok = ~(tri & (i == 3))
        a.append(np.stack([np.minimum(n0, n1)[ok], np.maximum(n0, n1)[ok]], 1))
        b.append(np.nonzero(ok)[0])
    keys = np.concatenate(a).astype(np.int64)
    owner = np.concatenate(b)
    k = keys[:, 0] * (m.x.size + 1) + keys[:, 1]
    order = np.argsort(k, kind="stable")
    return k[order], owner[order]


def cell_neighbours(m: MeshArrays):
    """Face neighbours: (cell, neighbour) pairs over shared edges."""
    ks, os_ = _cell_edge_keys(m)
    same = np.nonzero(ks[1:] == ks[:-1])[0]
    return np.concatenate([os_[same], os_[same + 1]]), np.concatenate([os_[same + 1], os_[same]])


def write_partitioned_msh(directory: str, name: str, m: MeshArrays, part: np.ndarray) -> int:
    """Write <name>_<r+1>.msh for every part r of `part` (one per cell);
    returns the number of parts."""
    import os
    P = int(part.max()) + 1
    src, dst = cell_neighbours(m)
    # the boundary element of each edge belongs to the cell holding both nodes
    ks, os_ = _cell_edge_keys(m)
    bk = np.minimum(m.bnd[:, 0], m.bnd[:, 1]).astype(np.int64) * (m.x.size + 1) + np.maximum(m.bnd[:, 0], m.bnd[:, 1])
    pos = np.searchsorted(ks, bk)
    if np.any(pos >= ks.size) or np.any(ks[np.minimum(pos, ks.size - 1)] != bk):
        raise ValueError("a boundary element is not a side of any cell")
    bnd_cell = os_[pos]
    nb_names = len(m.names)
    head = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", str(nb_names + 1)]
    head += ['1 %d "%s"' % (k + 1, nm) for k, nm in enumerate(m.names)]
    head += ['2 %d "internal"' % (nb_names + 1), "$EndPhysicalNames", "$Entities", "0 %d 1 0" % nb_names]
    head += ["%d 0 0 0 0 0 0 1 %d 0" % (k + 1, k + 1) for k in range(nb_names)]
    head += ["1 0 0 0 0 0 0 1 %d 0" % (nb_names + 1), "$EndEntities"]
    for r in range(P):
        own = np.nonzero(part == r)[0]
        gh = np.unique(dst[(part[src] == r) & (part[dst] != r)])
        cells = np.concatenate([own, gh])
        bsel = np.nonzero(part[bnd_cell] == r)[0]
        used = np.unique(np.concatenate([m.cells[cells][m.cells[cells] >= 0], m.bnd[bsel].ravel()]))
        lines = list(head)
        lines += ["$Nodes", "1 %d 1 %d" % (used.size, m.x.size), "2 1 0 %d" % used.size]
        lines += [str(int(v) + 1) for v in used]
        lines += ["%.17g %.17g 0" % (m.x[v], m.y[v]) for v in used]
        lines.append("$EndNodes")
        blocks = []
        for t in range(nb_names):
            bs = bsel[m.bnd_tag[bsel] == t]
            if bs.size:
                blocks.append(["1 %d 1 %d" % (t + 1, bs.size)] +
                              ["%d %d %d" % (b + 1, m.bnd[b, 0] + 1, m.bnd[b, 1] + 1) for b in bs])
        for want_tri in (True, False):
            cs = cells[(m.cells[cells, 3] < 0) == want_tri]
            if cs.size == 0:
                continue
            if want_tri:
                rows = ["%d %d %d %d" % (m.n_bnd + c + 1, m.cells[c, 0] + 1, m.cells[c, 1] + 1, m.cells[c, 2] + 1)
                        for c in cs]
                blocks.append(["2 1 2 %d" % cs.size] + rows)
            else:
                rows = ["%d %d %d %d %d" % (m.n_bnd + c + 1, m.cells[c, 0] + 1, m.cells[c, 1] + 1,
                                            m.cells[c, 2] + 1, m.cells[c, 3] + 1) for c in cs]
                blocks.append(["2 1 3 %d" % cs.size] + rows)
        n_el = sum(len(b_) - 1 for b_ in blocks)
        lines += ["$Elements", "%d %d 1 %d" % (len(blocks), n_el, m.n_bnd + m.n_cells)]
        for b_ in blocks:
            lines += b_
        lines.append("$EndElements")
        lines += ["$GhostElements", str(gh.size)]
        lines += ["%d %d 1 %d" % (m.n_bnd + c + 1, part[c] + 1, r + 1) for c in gh]
        lines.append("$EndGhostElements")
        with open(os.path.join(directory, "%s_%d.msh" % (name, r + 1)), "w") as f:
            f.write("\n".join(lines) + "\n")
    return P


# ---------------------------------------------------------------------------
# NACA 0012 O-mesh: quad boundary band + triangulated rings to the farfield
# ---------------------------------------------------------------------------

# ---------------------------------------------------------------------------
# SU2 ASCII meshes (2-D): NDIME, NELEM with element lines "type nodes... [index]"
# (5 triangle, 9 quadrilateral), NPOIN with "x y [index]", NMARK with
# MARKER_TAG / MARKER_ELEMS and line elements "3 a b".  Cells, nodes and
# boundary elements keep file order; marker tags are the physical names.
# ---------------------------------------------------------------------------

def read_su2(path: str) -> MeshArrays:
    with open(path) as f:
        tok = [ln.split("%")[0].strip() for ln in f]
    tok = [t for t in tok if t]
    i = 0
    cells, x, y, bnd, tags, names = [], [], [], [], [], []

    def key(t, k):
        return t.replace(" ", "").upper().startswith(k + "=")

    def val(t):
        return t.split("=", 1)[1].strip()
    while i < len(tok):
        t = tok[i]
        if key(t, "NDIME"):
            if int(val(t)) != 2:
                raise ValueError("SU2: only 2-D meshes")
            i += 1
        elif key(t, "NELEM"):
            n = int(val(t))
            for k in range(n):
                f = tok[i + 1 + k].split()
                kind = int(f[0])
                if kind == 5:
                    cells.append([int(f[1]), int(f[2]), int(f[3]), -1])
                elif kind == 9:
                    cells.append([int(f[1]), int(f[2]), int(f[3]), int(f[4])])
                else:
                    raise ValueError("SU2: element type %d is not a triangle or quadrilateral" % kind)
            i += 1 + n
        elif key(t, "NPOIN"):
            n = int(val(t).split()[0])
            for k in range(n):
                f = tok[i + 1 + k].split()
                x.append(float(f[0]))
                y.append(float(f[1]))
            i += 1 + n
        elif key(t, "NMARK"):
            i += 1
        elif key(t, "MARKER_TAG"):
            names.append(val(t))
            if not key(tok[i + 1], "MARKER_ELEMS"):
                raise ValueError("SU2: MARKER_TAG without MARKER_ELEMS")
            n = int(val(tok[i + 1]))
            for k in range(n):
                f = tok[i + 2 + k].split()
                if int(f[0]) != 3:
                    raise ValueError("SU2: marker element type %s is not a line" % f[0])
                bnd.append([int(f[1]), int(f[2])])
                tags.append(len(names) - 1)
            i += 2 + n
        else:
            i += 1
    return MeshArrays(np.array(x, dtype=np.float64), np.array(y, dtype=np.float64),
                      np.array(cells, dtype=np.int32).reshape(-1, 4), np.array(bnd, dtype=np.int32).reshape(-1, 2),
                      np.array(tags, dtype=np.int32), names)


def write_su2(path: str, m: MeshArrays) -> None:
    """The inverse of read_su2 (full precision coordinates)."""
    lines = ["NDIME= 2", "NELEM= %d" % m.cells.shape[0]]
    for k, c in enumerate(m.cells):
        lines.append(("5 %d %d %d %d" % (c[0], c[1], c[2], k)) if c[3] < 0 else
                     ("9 %d %d %d %d %d" % (c[0], c[1], c[2], c[3], k)))
    lines.append("NPOIN= %d" % m.x.size)
    for k in range(m.x.size):
        lines.append("%s %s %d" % (repr(float(m.x[k])), repr(float(m.y[k])), k))
    lines.append("NMARK= %d" % len(m.names))
    for t, nm in enumerate(m.names):
        sel = np.nonzero(np.asarray(m.bnd_tag) == t)[0]
        lines.append("MARKER_TAG= %s" % nm)
        lines.append("MARKER_ELEMS= %d" % sel.size)
        lines += ["3 %d %d" % (m.bnd[b, 0], m.bnd[b, 1]) for b in sel]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def naca4_thickness(x, t=0.12):
    """Half thickness of a symmetric NACA 4-digit section, closed trailing edge."""
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def _zipper(a_idx, a_ang, b_idx, b_ang):
    """Counter-clockwise triangles between an inner ring A and an outer ring B,
    given node indices and polar angles (each ring ascending, one turn)."""
    na, nb = a_idx.size, b_idx.size
    # unwrap both to start near a_ang[0]
    base = a_ang[0]
    ua = base + np.mod(a_ang - base, 2 * np.pi)
    ub = base + np.mod(b_ang - base, 2 * np.pi)
    jb = int(np.argmin(ub))          # ring B starts at its first node past A's start
    b_idx = np.roll(b_idx, -jb)
    ub = np.roll(ub, -jb)
    ua_next = np.append(ua[1:], ua[0] + 2 * np.pi)
    ub_next = np.append(ub[1:], ub[0] + 2 * np.pi)
    ev_ang = np.concatenate([ua_next, ub_next])
    ev_src = np.concatenate([np.zeros(na, dtype=np.int8), np.ones(nb, dtype=np.int8)])
    order = np.lexsort((ev_src, ev_ang))
    src = ev_src[order]
    ia = np.concatenate([[0], np.cumsum(src == 0)])[:-1]
    ib = np.concatenate([[0], np.cumsum(src == 1)])[:-1]
    tri = np.empty((na + nb, 4), dtype=np.int32)
    tri[:, 3] = -1
    isa = src == 0
    tri[isa, 0] = a_idx[ia[isa] % na]
    tri[isa, 1] = b_idx[ib[isa] % nb]
    tri[isa, 2] = a_idx[(ia[isa] + 1) % na]
    tri[~isa, 0] = a_idx[ia[~isa] % na]
    tri[~isa, 1] = b_idx[ib[~isa] % nb]
    tri[~isa, 2] = b_idx[(ib[~isa] + 1) % nb]
    return tri



def _naca_layout(n_surf, n_layers, first, growth, r_far, h):
    """Node rings of the O-mesh.

    Surface node k sits at parameter beta_k = 2 pi k / n_surf,
    x = (1 + cos beta)/2, upper surface for beta < pi.  Band layer j lies at
    wall distance d_j = first (growth^j - 1)/(growth - 1) along a direction
    that turns from the surface normal to the straight line towards C_k, the
    point of polar angle beta_k on the circle of radius R_b = d_L + 1/2 about
    mid-chord; the last layer is that circle.  Outside it, concentric circles
    at radial spacing ~h carry max(previous, 2 pi r / h) uniformly spaced
    nodes each, the last at r_far.  Returns (band [L+1, n_surf, 2], rings)."""
    k = np.arange(n_surf)
    beta = 2.0 * np.pi * k / n_surf
    xs = 0.5 * (1.0 + np.cos(beta))
    ys = np.where(beta < np.pi, 1.0, -1.0) * naca4_thickness(xs)
    ys[0] = 0.0
    if n_surf % 2 == 0:
        ys[n_surf // 2] = 0.0
    P = np.stack([xs, ys], 1)
    tang = np.roll(P, -1, 0) - np.roll(P, 1, 0)
    nrm = np.stack([tang[:, 1], -tang[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = first * (growth ** np.arange(n_layers + 1) - 1.0) / (growth - 1.0)
    r_b = d[-1] + 0.5
    centre = np.array([0.5, 0.0])
    C = centre + r_b * np.stack([np.cos(beta), np.sin(beta)], 1)
    band = np.empty((n_layers + 1, n_surf, 2))
    for j in range(n_layers + 1):
        s = d[j] / d[-1]
        w = s * s
        band[j] = P + (1.0 - w) * d[j] * nrm + w * s * (C - P)
    n_rings = max(1, int(round((r_far - r_b) / h)))
    rings = []
    n_prev = n_surf
    for j in range(1, n_rings + 1):
        r = r_b + (r_far - r_b) * j / n_rings
        n_j = max(n_prev, int(round(2.0 * np.pi * r / h)))
        th = 2.0 * np.pi * (np.arange(n_j) + 0.5 * (j % 2)) / n_j
        rings.append(centre + r * np.stack([np.cos(th), np.sin(th)], 1))
        n_prev = n_j
    return band, rings


def _naca_count(n_surf, n_layers, first, growth, r_far, h):
    """Cell count of the O-mesh for ring spacing h, from the ring node counts
    alone (the same rule as _naca_layout, without coordinates)."""
    d_last = first * (growth ** n_layers - 1.0) / (growth - 1.0)
    r_b = d_last + 0.5
    n_rings = max(1, int(round((r_far - r_b) / h)))
    total = n_surf * n_layers
    n_prev = n_surf
    for j in range(1, n_rings + 1):
        r = r_b + (r_far - r_b) * j / n_rings
        n_j = max(n_prev, int(round(2.0 * np.pi * r / h)))
        total += n_prev + n_j
        n_prev = n_j
    return total


def naca0012_omesh(n_surf: int = 512, n_layers: int = 64, first: float = 1e-3, growth: float = 1.1,
                   r_far: float = 50.0, target_cells: int = 2_000_000, h: float = None) -> MeshArrays:
    """NACA 0012 O-mesh: the analytic 4-digit section (closed trailing edge)
    with n_surf cosine-clustered surface points; n_layers quad layers of
    geometric thickness (first spacing ``first``, ratio ``growth``: 64 layers
    of 1e-3 x 1.1 reach ~4.45 chords); then triangles between concentric
    circles out to a farfield of radius r_far, with the ring spacing h chosen
    so that the mesh has about target_cells cells.  Boundaries ``airfoil``
    and ``farfield``.  Cells in file order: quads layer by layer from the
    wall, then triangles ring by ring, all counter-clockwise."""
    if h is None:
        lo, hi = 1e-3, 10.0
        for _ in range(40):
            mid = math.sqrt(lo * hi)
            if _naca_count(n_surf, n_layers, first, growth, r_far, mid) > target_cells:
                lo = mid
            else:
                hi = mid
        h = hi
    band, rings = _naca_layout(n_surf, n_layers, first, growth, r_far, h)
    xy = np.concatenate([band.reshape(-1, 2)] + rings, 0)
    L, S = n_layers, n_surf
    jj, ss = np.meshgrid(np.arange(L), np.arange(S), indexing="ij")
    a = jj * S + ss
    b = jj * S + (ss + 1) % S
    cells = [np.stack([a, a + S, b + S, b], -1).reshape(-1, 4).astype(np.int32)]
    prev_idx = L * S + np.arange(S)
    prev_ang = 2.0 * np.pi * np.arange(S) / S
    start = (L + 1) * S
    for j, r in enumerate(rings, 1):
        n_j = r.shape[0]
        idx = start + np.arange(n_j)
        ang = 2.0 * np.pi * (np.arange(n_j) + 0.5 * (j % 2)) / n_j
        cells.append(_zipper(prev_idx, prev_ang, idx, ang))
        prev_idx, prev_ang, start = idx, ang, start + n_j
    cells = np.concatenate(cells, 0)
    surf = np.stack([np.arange(S), (np.arange(S) + 1) % S], 1)
    far = np.stack([prev_idx, np.roll(prev_idx, -1)], 1)
    bnd = np.concatenate([surf, far], 0).astype(np.int32)
    tag = np.concatenate([np.zeros(S, dtype=np.int32), np.ones(far.shape[0], dtype=np.int32)])
    return MeshArrays(xy[:, 0].copy(), xy[:, 1].copy(), cells, bnd, tag, ["airfoil", "farfield"])


def naca0012_sized(total_cells: int, r_far: float = 50.0) -> MeshArrays:
    """A coarsened NACA 0012 O-mesh of about total_cells cells: surface points
    (a multiple of 8) and band layers scale with sqrt(total/2M) from 512 and
    64; the first spacing and growth keep the band ~4.45 chords thick."""
    s = math.sqrt(total_cells / 2.0e6)
    n_surf = max(48, 8 * int(round(64 * s)))
    n_layers = max(8, int(round(64 * s)))
    first = 1e-3 / s
    lo, hi = 1.0001, 3.0
    for _ in range(60):
        g = 0.5 * (lo + hi)
        if first * (g ** n_layers - 1.0) / (g - 1.0) > 4.45:
            hi = g
        else:
            lo = g
    return naca0012_omesh(n_surf=n_surf, n_layers=n_layers, first=first, growth=0.5 * (lo + hi),
                          r_far=r_far, target_cells=total_cells)
