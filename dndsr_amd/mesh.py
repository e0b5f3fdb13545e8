"""Synthetic structured meshes handed to the solver as unstructured connectivity.

The reference reads CGNS meshes (src/Geom/Mesh/Mesh_Serial_ReadFromCGNS.cpp); its
periodic test meshes (IV10_10, Uniform32_3D_Periodic) are uniform quad / hex grids,
which are rebuilt here directly.  Periodic faces are stored once: the face geometry
lives in the frame of its back cell (``face2cell[:, 0]``), and ``faceShift`` is the
translation that carries a face point into the frame of the front cell -- the role
of ``periodicInfo.TransCoord/TransCoordBack`` in src/CFV/FiniteVolume.hpp:283-317.

Node order and face-normal conventions:
  * Quad4 / Hex8 nodes follow the CGNS order.
  * 2-D face normal = (t_y, -t_x) for the edge tangent t (FacialJacobianToNormVec,
    src/Geom/Geometric.hpp:113-122); 3-D normal = d x/d xi  x  d x/d eta.
    Faces are emitted so that this normal points from the back cell to the front cell.
  * cell2face follows the element face order: Quad4 (y-, x+, y+, x-),
    Hex8 (z-, y-, x+, y+, x-, z+).
Boundary faces (``box_quad_grid``, ``box_hex_grid``): a direction that is not
periodic ends in boundary faces with ``face2cell[f] = (cell, -1)``, the normal
pointing out of the domain, and a zone index ``faceZone[f]`` (the mesh's face zone,
UnstructuredMesh::GetFaceZone) per side.
Cells are numbered lexicographically (x fastest).  Faces are numbered by blocks of
64 consecutive owner cells: within a block, the 64 +x faces, then the 64 +y faces
(then +z), so a run of 64 faces is one direction over 64 consecutive cells (one
coalesced row per wavefront) and the faces of a block of cells are processed back
to back (the cells' data is reused from cache across the directions).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

LINE2, QUAD4, HEX8, TRI3, TET4 = 1, 2, 3, 4, 5


FACE_BLOCK = 64


def _by_cell_block(dim, N, face2node, face2cell, faceShift, cell2face, B=FACE_BLOCK):
    """Renumber direction-major faces (d N + c) block-direction-major: cell c lies in block
    k = c // B of size r (B, or the remainder for the last block); its +d face becomes
    k B dim + d r + c % B."""
    f = np.arange(dim * N)
    c, d = f % N, f // N
    k = c // B
    r = np.minimum(B, N - k * B)
    new = k * B * dim + d * r + (c - k * B)
    inv = np.empty_like(new)
    inv[new] = f
    return face2node[inv], face2cell[inv], faceShift[inv], new[cell2face]


@dataclass
class Mesh:
    dim: int
    coords: np.ndarray        # [nNode, 3] float64
    cellType: np.ndarray      # [nCell] int32
    cell2nodePtr: np.ndarray  # [nCell+1] int64
    cell2node: np.ndarray     # int64
    cell2facePtr: np.ndarray  # [nCell+1] int64
    cell2face: np.ndarray     # int64
    faceType: np.ndarray      # [nFace] int32
    face2nodePtr: np.ndarray  # [nFace+1] int64
    face2node: np.ndarray     # int64
    face2cell: np.ndarray     # [nFace, 2] int64, -1 = boundary
    faceShift: np.ndarray     # [nFace, 3] float64
    shape: tuple = ()
    faceZone: np.ndarray | None = None  # [nFace] int32: boundary zone of faces with face2cell[:, 1] = -1
    nBCZone: int = 0

    @property
    def nCell(self) -> int:
        return int(self.cellType.shape[0])

    @property
    def nFace(self) -> int:
        return int(self.faceType.shape[0])

    @property
    def nNode(self) -> int:
        return int(self.coords.shape[0])


def _jitter_nodes(X: np.ndarray, h: np.ndarray, jitter: float, seed: int) -> np.ndarray:
    """Periodic-consistent node perturbation: node (n) and node (n + N_d) move together."""
    if jitter == 0.0:
        return X
    rng = np.random.default_rng(seed)
    sh = X.shape[:-1]
    dim = len(sh)
    core = tuple(s - 1 for s in sh)
    d = rng.uniform(-jitter, jitter, size=core + (3,)) * np.concatenate([h, np.zeros(3 - dim)])
    if dim == 2:
        d[..., 2] = 0.0
    idx = np.meshgrid(*[np.arange(s) % c for s, c in zip(sh, core)], indexing="ij")
    return X + d[tuple(idx)]


def periodic_quad_grid(nx: int, ny: int, Lx: float = 10.0, Ly: float = 10.0,
                       x0: float = 0.0, y0: float = 0.0, jitter: float = 0.0, seed: int = 0) -> Mesh:
    """Doubly periodic nx x ny Quad4 grid on [x0, x0+Lx] x [y0, y0+Ly]."""
    assert nx >= 2 and ny >= 2
    hx, hy = Lx / nx, Ly / ny
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    X = np.stack([x0 + I * hx, y0 + J * hy, np.zeros_like(I, dtype=float)], axis=-1).astype(np.float64)
    X = _jitter_nodes(X, np.array([hx, hy]), jitter, seed)
    nid = lambda i, j: i * (ny + 1) + j  # noqa: E731
    coords = X.reshape(-1, 3)
    N = nx * ny
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ci, cj = ci.T.ravel(), cj.T.ravel()          # cell c = i + nx*j
    cell2node = np.stack([nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)], 1)
    c = np.arange(N)
    # x-faces (+x of cell c), tangent (0,+1) -> normal +x
    fx_nodes = np.stack([nid(ci + 1, cj), nid(ci + 1, cj + 1)], 1)
    fx_R = ((ci + 1) % nx) + nx * cj
    fx_shift = np.zeros((N, 3))
    fx_shift[ci == nx - 1, 0] = -Lx
    # y-faces (+y of cell c), tangent (-1,0) -> normal +y
    fy_nodes = np.stack([nid(ci + 1, cj + 1), nid(ci, cj + 1)], 1)
    fy_R = ci + nx * ((cj + 1) % ny)
    fy_shift = np.zeros((N, 3))
    fy_shift[cj == ny - 1, 1] = -Ly
    face2node = np.concatenate([fx_nodes, fy_nodes])
    face2cell = np.stack([np.concatenate([c, c]), np.concatenate([fx_R, fy_R])], 1)
    faceShift = np.concatenate([fx_shift, fy_shift])
    xm = ((ci - 1) % nx) + nx * cj
    ym = ci + nx * ((cj - 1) % ny)
    cell2face = np.stack([N + ym, c, N + c, xm], 1)   # y-, x+, y+, x-
    face2node, face2cell, faceShift, cell2face = _by_cell_block(2, N, face2node, face2cell, faceShift, cell2face)
    return Mesh(2, coords, np.full(N, QUAD4, np.int32), np.arange(N + 1, dtype=np.int64) * 4,
                cell2node.ravel().astype(np.int64), np.arange(N + 1, dtype=np.int64) * 4,
                cell2face.ravel().astype(np.int64), np.full(2 * N, LINE2, np.int32),
                np.arange(2 * N + 1, dtype=np.int64) * 2, face2node.ravel().astype(np.int64),
                face2cell.astype(np.int64), faceShift.astype(np.float64), (nx, ny))


def periodic_hex_grid(nx: int, ny: int, nz: int, Lx: float = 1.0, Ly: float = 1.0, Lz: float = 1.0,
                      x0: float = 0.0, y0: float = 0.0, z0: float = 0.0,
                      jitter: float = 0.0, seed: int = 0) -> Mesh:
    """Triply periodic nx x ny x nz Hex8 grid on [x0,x0+Lx] x [y0,y0+Ly] x [z0,z0+Lz]."""
    assert nx >= 2 and ny >= 2 and nz >= 2
    hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    X = np.stack([x0 + I * hx, y0 + J * hy, z0 + K * hz], axis=-1).astype(np.float64)
    X = _jitter_nodes(X, np.array([hx, hy, hz]), jitter, seed)
    coords = X.reshape(-1, 3)
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k  # noqa: E731
    N = nx * ny * nz
    c = np.arange(N)
    ci = c % nx
    cj = (c // nx) % ny
    ck = c // (nx * ny)
    cell2node = np.stack([nid(ci, cj, ck), nid(ci + 1, cj, ck), nid(ci + 1, cj + 1, ck), nid(ci, cj + 1, ck),
                          nid(ci, cj, ck + 1), nid(ci + 1, cj, ck + 1), nid(ci + 1, cj + 1, ck + 1),
                          nid(ci, cj + 1, ck + 1)], 1)
    cid = lambda i, j, k: (i % nx) + nx * ((j % ny) + ny * (k % nz))  # noqa: E731
    # +x face: a=(x1,y0,z0) b=(x1,y1,z0) c=(x1,y1,z1) d=(x1,y0,z1): (+y) x (+z) = +x
    fx = np.stack([nid(ci + 1, cj, ck), nid(ci + 1, cj + 1, ck), nid(ci + 1, cj + 1, ck + 1), nid(ci + 1, cj, ck + 1)], 1)
    # +y face: a=(x0,y1,z0) b=(x0,y1,z1) c=(x1,y1,z1) d=(x1,y1,z0): (+z) x (+x) = +y
    fy = np.stack([nid(ci, cj + 1, ck), nid(ci, cj + 1, ck + 1), nid(ci + 1, cj + 1, ck + 1), nid(ci + 1, cj + 1, ck)], 1)
    # +z face: a=(x0,y0,z1) b=(x1,y0,z1) c=(x1,y1,z1) d=(x0,y1,z1): (+x) x (+y) = +z
    fz = np.stack([nid(ci, cj, ck + 1), nid(ci + 1, cj, ck + 1), nid(ci + 1, cj + 1, ck + 1), nid(ci, cj + 1, ck + 1)], 1)
    sx = np.zeros((N, 3)); sx[ci == nx - 1, 0] = -Lx
    sy = np.zeros((N, 3)); sy[cj == ny - 1, 1] = -Ly
    sz = np.zeros((N, 3)); sz[ck == nz - 1, 2] = -Lz
    face2node = np.concatenate([fx, fy, fz])
    R = np.concatenate([cid(ci + 1, cj, ck), cid(ci, cj + 1, ck), cid(ci, cj, ck + 1)])
    face2cell = np.stack([np.concatenate([c, c, c]), R], 1)
    faceShift = np.concatenate([sx, sy, sz])
    xm, ym, zm = cid(ci - 1, cj, ck), cid(ci, cj - 1, ck), cid(ci, cj, ck - 1)
    cell2face = np.stack([2 * N + zm, N + ym, c, N + c, xm, 2 * N + c], 1)  # z-, y-, x+, y+, x-, z+
    face2node, face2cell, faceShift, cell2face = _by_cell_block(3, N, face2node, face2cell, faceShift, cell2face)
    F = 3 * N
    return Mesh(3, coords, np.full(N, HEX8, np.int32), np.arange(N + 1, dtype=np.int64) * 8,
                cell2node.ravel().astype(np.int64), np.arange(N + 1, dtype=np.int64) * 6,
                cell2face.ravel().astype(np.int64), np.full(F, QUAD4, np.int32),
                np.arange(F + 1, dtype=np.int64) * 4, face2node.ravel().astype(np.int64),
                face2cell.astype(np.int64), faceShift.astype(np.float64), (nx, ny, nz))


def _jitter_nodes_box(X: np.ndarray, h: np.ndarray, jitter: float, seed: int, periodic) -> np.ndarray:
    """Node perturbation for partly periodic grids: periodic directions move node n and
    n + N_d together; in a bounded direction the nodes on its two end planes keep that
    coordinate (the domain keeps its shape)."""
    if jitter == 0.0:
        return X
    rng = np.random.default_rng(seed)
    sh = X.shape[:-1]
    dim = len(sh)
    d = rng.uniform(-jitter, jitter, size=sh + (3,)) * np.concatenate([h, np.zeros(3 - dim)])
    if dim == 2:
        d[..., 2] = 0.0
    for a in range(dim):
        if periodic[a]:
            src = [slice(None)] * dim
            dst = [slice(None)] * dim
            src[a], dst[a] = 0, sh[a] - 1
            d[tuple(dst)] = d[tuple(src)]
        else:
            for end in (0, sh[a] - 1):
                sl = [slice(None)] * dim
                sl[a] = end
                d[tuple(sl) + (a,)] = 0.0
    return X + d


def _zone_of(zones, key):
    if zones is None or key not in zones:
        raise ValueError(f"a bounded direction needs a zone for side {key!r}")
    return int(zones[key])


def box_quad_grid(nx: int, ny: int, Lx: float = 1.0, Ly: float = 1.0, x0: float = 0.0, y0: float = 0.0,
                  periodic=(False, False), zones: dict | None = None, jitter: float = 0.0, seed: int = 0,
                  node_map=None) -> Mesh:
    """nx x ny Quad4 grid on [x0, x0+Lx] x [y0, y0+Ly]; each direction periodic or bounded.
    zones: {'xmin', 'xmax', 'ymin', 'ymax'} -> zone index for the bounded sides.
    node_map: optional f(coords[nNode, 3]) -> coords applied after the jitter."""
    assert nx >= 2 and ny >= 2
    px, py = bool(periodic[0]), bool(periodic[1])
    hx, hy = Lx / nx, Ly / ny
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    X = np.stack([x0 + I * hx, y0 + J * hy, np.zeros_like(I, dtype=float)], axis=-1).astype(np.float64)
    X = _jitter_nodes_box(X, np.array([hx, hy]), jitter, seed, (px, py))
    coords = X.reshape(-1, 3)
    if node_map is not None:
        coords = np.ascontiguousarray(node_map(coords.copy()), np.float64)
    nid = lambda i, j: i * (ny + 1) + j  # noqa: E731
    N = nx * ny
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ci, cj = ci.T.ravel(), cj.T.ravel()
    c = np.arange(N)
    fx_nodes = np.stack([nid(ci + 1, cj), nid(ci + 1, cj + 1)], 1)
    fx_R = ((ci + 1) % nx) + nx * cj
    fx_shift = np.zeros((N, 3))
    fy_nodes = np.stack([nid(ci + 1, cj + 1), nid(ci, cj + 1)], 1)
    fy_R = ci + nx * ((cj + 1) % ny)
    fy_shift = np.zeros((N, 3))
    zone = np.full(2 * N, -1, np.int32)
    if px:
        fx_shift[ci == nx - 1, 0] = -Lx
    else:
        fx_R = np.where(ci == nx - 1, -1, fx_R)
        zone[:N][ci == nx - 1] = _zone_of(zones, "xmax")
    if py:
        fy_shift[cj == ny - 1, 1] = -Ly
    else:
        fy_R = np.where(cj == ny - 1, -1, fy_R)
        zone[N:][cj == ny - 1] = _zone_of(zones, "ymax")
    face2node = np.concatenate([fx_nodes, fy_nodes])
    face2cell = np.stack([np.concatenate([c, c]), np.concatenate([fx_R, fy_R])], 1)
    faceShift = np.concatenate([fx_shift, fy_shift])
    xm = ((ci - 1) % nx) + nx * cj
    ym = ci + nx * ((cj - 1) % ny)
    cell2face = np.stack([N + ym, c, N + c, xm], 1)   # y-, x+, y+, x-
    face2node, face2cell, faceShift, cell2face = _by_cell_block(2, N, face2node, face2cell, faceShift, cell2face)
    zone = zone[np.argsort(_block_perm(2, N))]
    extra_n, extra_c, extra_z = [], [], []
    F = 2 * N
    if not px:  # x- boundary faces, normal -x
        cells = c[ci == 0]
        extra_n.append(np.stack([nid(0 * cells, cells // nx + 1), nid(0 * cells, cells // nx)], 1))
        extra_c.append(cells)
        extra_z.append(np.full(len(cells), _zone_of(zones, "xmin"), np.int32))
        cell2face[cells, 3] = F + np.arange(len(cells))
        F += len(cells)
    if not py:  # y- boundary faces, normal -y
        cells = c[cj == 0]
        extra_n.append(np.stack([nid(cells % nx, 0 * cells), nid(cells % nx + 1, 0 * cells)], 1))
        extra_c.append(cells)
        extra_z.append(np.full(len(cells), _zone_of(zones, "ymin"), np.int32))
        cell2face[cells, 0] = F + np.arange(len(cells))
        F += len(cells)
    if extra_n:
        en = np.concatenate(extra_n)
        ec = np.concatenate(extra_c)
        face2node = np.concatenate([face2node, en])
        face2cell = np.concatenate([face2cell, np.stack([ec, -np.ones_like(ec)], 1)])
        faceShift = np.concatenate([faceShift, np.zeros((len(ec), 3))])
        zone = np.concatenate([zone] + extra_z)
    nz = int(zone.max()) + 1 if (zone >= 0).any() else 0
    return Mesh(2, coords, np.full(N, QUAD4, np.int32), np.arange(N + 1, dtype=np.int64) * 4,
                np.stack([nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)], 1).ravel().astype(np.int64),
                np.arange(N + 1, dtype=np.int64) * 4, cell2face.ravel().astype(np.int64),
                np.full(F, LINE2, np.int32), np.arange(F + 1, dtype=np.int64) * 2, face2node.ravel().astype(np.int64),
                face2cell.astype(np.int64), faceShift.astype(np.float64), (nx, ny),
                None if nz == 0 else zone.astype(np.int32), nz)


def _block_perm(dim, N, B=FACE_BLOCK):
    """New index of direction-major face d N + c under the block-direction-major numbering."""
    f = np.arange(dim * N)
    c, d = f % N, f // N
    k = c // B
    r = np.minimum(B, N - k * B)
    return k * B * dim + d * r + (c - k * B)


def box_hex_grid(nx: int, ny: int, nz: int, Lx: float = 1.0, Ly: float = 1.0, Lz: float = 1.0,
                 x0: float = 0.0, y0: float = 0.0, z0: float = 0.0, periodic=(False, False, False),
                 zones: dict | None = None, jitter: float = 0.0, seed: int = 0, node_map=None) -> Mesh:
    """nx x ny x nz Hex8 grid; each direction periodic or bounded.  zones: {'xmin', 'xmax',
    'ymin', 'ymax', 'zmin', 'zmax'} -> zone index for the bounded sides."""
    assert nx >= 2 and ny >= 2 and nz >= 2
    per = [bool(p) for p in periodic]
    hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    X = np.stack([x0 + I * hx, y0 + J * hy, z0 + K * hz], axis=-1).astype(np.float64)
    X = _jitter_nodes_box(X, np.array([hx, hy, hz]), jitter, seed, per)
    coords = X.reshape(-1, 3)
    if node_map is not None:
        coords = np.ascontiguousarray(node_map(coords.copy()), np.float64)
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k  # noqa: E731
    N = nx * ny * nz
    c = np.arange(N)
    ci, cj, ck = c % nx, (c // nx) % ny, c // (nx * ny)
    cell2node = np.stack([nid(ci, cj, ck), nid(ci + 1, cj, ck), nid(ci + 1, cj + 1, ck), nid(ci, cj + 1, ck),
                          nid(ci, cj, ck + 1), nid(ci + 1, cj, ck + 1), nid(ci + 1, cj + 1, ck + 1),
                          nid(ci, cj + 1, ck + 1)], 1)
    cid = lambda i, j, k: (i % nx) + nx * ((j % ny) + ny * (k % nz))  # noqa: E731
    fx = np.stack([nid(ci + 1, cj, ck), nid(ci + 1, cj + 1, ck), nid(ci + 1, cj + 1, ck + 1), nid(ci + 1, cj, ck + 1)], 1)
    fy = np.stack([nid(ci, cj + 1, ck), nid(ci, cj + 1, ck + 1), nid(ci + 1, cj + 1, ck + 1), nid(ci + 1, cj + 1, ck)], 1)
    fz = np.stack([nid(ci, cj, ck + 1), nid(ci + 1, cj, ck + 1), nid(ci + 1, cj + 1, ck + 1), nid(ci, cj + 1, ck + 1)], 1)
    idxs = (ci, cj, ck)
    ns = (nx, ny, nz)
    Ls = (Lx, Ly, Lz)
    R = [cid(ci + 1, cj, ck), cid(ci, cj + 1, ck), cid(ci, cj, ck + 1)]
    sh = [np.zeros((N, 3)) for _ in range(3)]
    zone = np.full(3 * N, -1, np.int32)
    for a, key in enumerate(("xmax", "ymax", "zmax")):
        last = idxs[a] == ns[a] - 1
        if per[a]:
            sh[a][last, a] = -Ls[a]
        else:
            R[a] = np.where(last, -1, R[a])
            zone[a * N:(a + 1) * N][last] = _zone_of(zones, key)
    face2node = np.concatenate([fx, fy, fz])
    face2cell = np.stack([np.concatenate([c, c, c]), np.concatenate(R)], 1)
    faceShift = np.concatenate(sh)
    xm, ym, zm = cid(ci - 1, cj, ck), cid(ci, cj - 1, ck), cid(ci, cj, ck - 1)
    cell2face = np.stack([2 * N + zm, N + ym, c, N + c, xm, 2 * N + c], 1)  # z-, y-, x+, y+, x-, z+
    face2node, face2cell, faceShift, cell2face = _by_cell_block(3, N, face2node, face2cell, faceShift, cell2face)
    zone = zone[np.argsort(_block_perm(3, N))]
    F = 3 * N
    extra_n, extra_c, extra_z = [], [], []
    # minus-side boundary faces with outward normals; (axis, zone key, cell2face slot, node builder)
    minus = [
        (0, "xmin", 4, lambda i, j, k: np.stack([nid(i, j, k), nid(i, j, k + 1), nid(i, j + 1, k + 1), nid(i, j + 1, k)], 1)),
        (1, "ymin", 1, lambda i, j, k: np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j, k + 1), nid(i, j, k + 1)], 1)),
        (2, "zmin", 0, lambda i, j, k: np.stack([nid(i, j, k), nid(i, j + 1, k), nid(i + 1, j + 1, k), nid(i + 1, j, k)], 1)),
    ]
    for a, key, slot, nodes in minus:
        if per[a]:
            continue
        sel = idxs[a] == 0
        cells = c[sel]
        extra_n.append(nodes(ci[sel], cj[sel], ck[sel]))
        extra_c.append(cells)
        extra_z.append(np.full(len(cells), _zone_of(zones, key), np.int32))
        cell2face[cells, slot] = F + np.arange(len(cells))
        F += len(cells)
    if extra_n:
        ec = np.concatenate(extra_c)
        face2node = np.concatenate([face2node] + extra_n)
        face2cell = np.concatenate([face2cell, np.stack([ec, -np.ones_like(ec)], 1)])
        faceShift = np.concatenate([faceShift, np.zeros((len(ec), 3))])
        zone = np.concatenate([zone] + extra_z)
    nzone = int(zone.max()) + 1 if (zone >= 0).any() else 0
    return Mesh(3, coords, np.full(N, HEX8, np.int32), np.arange(N + 1, dtype=np.int64) * 8,
                cell2node.ravel().astype(np.int64), np.arange(N + 1, dtype=np.int64) * 6,
                cell2face.ravel().astype(np.int64), np.full(F, QUAD4, np.int32),
                np.arange(F + 1, dtype=np.int64) * 4, face2node.ravel().astype(np.int64),
                face2cell.astype(np.int64), faceShift.astype(np.float64), (nx, ny, nz),
                None if nzone == 0 else zone.astype(np.int32), nzone)


def bump_channel_hex(nx: int, ny: int, nz: int, L: float = 3.0, H: float = 1.0, W: float = 0.5,
                     amp: float = 0.1, jitter: float = 0.0, seed: int = 0, stretch: float = 1.0) -> Mesh:
    """Channel [0, L] x [0, H] x [0, W] with a sinusoidal bump of height amp * H on the lower
    wall over x in [L/3, 2L/3] (y_wall = amp H sin^2(3 pi x / L - pi)); the grid lines follow
    the wall, y' = y + y_wall(x) (1 - y / H), and stretch > 1 clusters them toward the lower
    wall (y -> H (stretch^(y/H) - 1) / (stretch - 1)).  Spanwise (z) periodic.  Zones: 0 inlet
    (x = 0), 1 outlet (x = L), 2 lower wall, 3 upper boundary."""
    def bump(P):
        x, y = P[:, 0], P[:, 1]
        if stretch != 1.0:
            y = H * (stretch ** (y / H) - 1) / (stretch - 1)
        s = np.where((x > L / 3) & (x < 2 * L / 3), np.sin(3 * np.pi * x / L - np.pi) ** 2, 0.0)
        Q = P.copy()
        Q[:, 1] = y + amp * H * s * (1 - y / H)
        return Q

    return box_hex_grid(nx, ny, nz, L, H, W, periodic=(False, False, True),
                        zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=jitter, seed=seed, node_map=bump)


# ---------------------------------------------------------------------------- simplex meshes
# Element face tables with outward normals for positively oriented cells:
# Tri3 edges (src/Geom/Elements/Tri3.hpp:83-85), Tet4 faces (Tet4.hpp:92-95).
TRI3_FACES = np.array([[0, 1], [1, 2], [2, 0]])
TET4_FACES = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]])


def simplex_mesh(dim: int, coords: np.ndarray, cells: np.ndarray, canon: np.ndarray | None = None,
                 bnd_zone=None, shape: tuple = ()) -> Mesh:
    """Tri3 (dim 2) or Tet4 (dim 3) mesh from cells [N, dim + 1] of node ids.

    Cells are reoriented to a positive Jacobian.  Faces are numbered by their first
    occurrence in (cell, element face) order; that cell is the face's back cell and the
    face keeps its node order, so the normal points out of it.  canon[n] is the node id
    after periodic identification: two cell faces with the same canonical node set are
    one face, and faceShift is the translation from the back cell's copy of the face to
    the front cell's.  A face seen once is a boundary face; bnd_zone(face2node [nb, dim])
    returns its zone."""
    coords = np.ascontiguousarray(coords, np.float64)
    cells = np.array(cells, np.int64)
    N, nNode = len(cells), len(coords)
    P = coords[cells]
    if dim == 2:
        e1, e2 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
        det = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    else:
        det = np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])
    if np.any(det == 0):
        raise ValueError("degenerate cell")
    neg = det < 0
    cells[neg, 1], cells[neg, 2] = cells[neg, 2].copy(), cells[neg, 1].copy()
    FT = TRI3_FACES if dim == 2 else TET4_FACES
    nf = len(FT)
    occ = cells[:, FT].reshape(N * nf, dim)              # occurrence o = c nf + k
    canon = np.arange(nNode, dtype=np.int64) if canon is None else np.asarray(canon, np.int64)
    cn = np.sort(canon[occ], axis=1)
    base = int(canon.max()) + 1
    if base ** dim >= 2 ** 62:
        raise ValueError("too many nodes for the face key")
    key = cn[:, 0].copy()
    for a in range(1, dim):
        key = key * base + cn[:, a]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    start = np.r_[True, ks[1:] != ks[:-1]]
    gid = np.cumsum(start) - 1
    counts = np.bincount(gid)
    if counts.max() > 2:
        raise ValueError("a face is shared by more than two cells")
    pos = np.flatnonzero(start)
    first = order[pos]
    second = np.where(counts == 2, order[np.minimum(pos + 1, len(order) - 1)], -1)
    fo = np.argsort(first, kind="stable")
    first, second = first[fo], second[fo]
    F = len(first)
    face_of_group = np.empty(F, np.int64)
    face_of_group[fo] = np.arange(F)
    occ_face = np.empty(N * nf, np.int64)
    occ_face[order] = face_of_group[gid]
    cell2face = occ_face.reshape(N, nf)
    face2node = occ[first]
    has2 = second >= 0
    face2cell = np.stack([first // nf, np.where(has2, second // nf, -1)], 1)
    faceShift = np.zeros((F, 3))
    if has2.any():
        # translation between the two copies: the node of smallest canonical id in each
        a, b = occ[first[has2]], occ[second[has2]]
        ia = np.argmin(canon[a], axis=1)
        ib = np.argmin(canon[b], axis=1)
        r = np.arange(len(a))
        faceShift[has2] = coords[b[r, ib]] - coords[a[r, ia]]
    zone = None
    nzone = 0
    if (~has2).any():
        if bnd_zone is None:
            raise ValueError("the mesh has boundary faces: bnd_zone is required")
        zone = np.full(F, -1, np.int32)
        zone[~has2] = np.asarray(bnd_zone(face2node[~has2]), np.int32)
        nzone = int(zone.max()) + 1
    ct, ft = (TRI3, LINE2) if dim == 2 else (TET4, TRI3)
    return Mesh(dim, coords, np.full(N, ct, np.int32), np.arange(N + 1, dtype=np.int64) * (dim + 1),
                cells.ravel(), np.arange(N + 1, dtype=np.int64) * nf, cell2face.ravel(),
                np.full(F, ft, np.int32), np.arange(F + 1, dtype=np.int64) * dim, face2node.ravel().astype(np.int64),
                face2cell.astype(np.int64), faceShift, shape, zone, nzone)


def _lattice_zone(lat, ns, per, zones):
    """Zone of a boundary face from the lattice indices lat[node] = (i, j, k) of its nodes."""
    keys = (("xmin", "xmax"), ("ymin", "ymax"), ("zmin", "zmax"))

    def f(fn):
        L = lat[fn]                                   # [nb, dim, 3]
        z = np.full(len(fn), -1, np.int32)
        for a in range(len(ns)):
            if per[a]:
                continue
            for end, key in ((0, keys[a][0]), (ns[a], keys[a][1])):
                on = np.all(L[:, :, a] == end, axis=1)
                z[on] = _zone_of(zones, key)
        if np.any(z < 0):
            raise ValueError("boundary face off the box sides")
        return z
    return f


def tri_grid(nx: int, ny: int, Lx: float = 1.0, Ly: float = 1.0, x0: float = 0.0, y0: float = 0.0,
             periodic=(False, False), zones: dict | None = None, jitter: float = 0.0, seed: int = 0,
             node_map=None) -> Mesh:
    """nx x ny lattice on [x0, x0+Lx] x [y0, y0+Ly], each square split into two Tri3 cells
    along its (i, j)-(i+1, j+1) diagonal; directions periodic or bounded as box_quad_grid."""
    assert nx >= 2 and ny >= 2
    per = [bool(p) for p in periodic]
    hx, hy = Lx / nx, Ly / ny
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    X = np.stack([x0 + I * hx, y0 + J * hy, np.zeros_like(I, dtype=float)], axis=-1).astype(np.float64)
    X = _jitter_nodes_box(X, np.array([hx, hy]), jitter, seed, per)
    coords = X.reshape(-1, 3)
    if node_map is not None:
        coords = np.ascontiguousarray(node_map(coords.copy()), np.float64)
    nid = lambda i, j: i * (ny + 1) + j  # noqa: E731
    lat = np.stack([I.ravel(), J.ravel(), 0 * I.ravel()], 1)
    canon = nid(I % nx if per[0] else I, J % ny if per[1] else J).ravel()
    c = np.arange(nx * ny)
    ci, cj = c % nx, c // nx
    a, b, cc, d = nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)
    cells = np.stack([np.stack([a, b, cc], 1), np.stack([a, cc, d], 1)], 1).reshape(-1, 3)
    return simplex_mesh(2, coords, cells, canon, _lattice_zone(lat, (nx, ny), per, zones), (nx, ny))


def tet_grid(nx: int, ny: int, nz: int, Lx: float = 1.0, Ly: float = 1.0, Lz: float = 1.0,
             x0: float = 0.0, y0: float = 0.0, z0: float = 0.0, periodic=(False, False, False),
             zones: dict | None = None, jitter: float = 0.0, seed: int = 0, node_map=None) -> Mesh:
    """nx x ny x nz lattice, each cube split into six Tet4 cells around its (0,0,0)-(1,1,1)
    diagonal (the Kuhn subdivision: one tetrahedron per ordering of the axes, so that
    neighbouring cubes share face triangulations); directions periodic or bounded as
    box_hex_grid.  With jittered nodes this is the jittered-lattice Delaunay-type
    tetrahedralisation of the box."""
    assert nx >= 2 and ny >= 2 and nz >= 2
    per = [bool(p) for p in periodic]
    hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    X = np.stack([x0 + I * hx, y0 + J * hy, z0 + K * hz], axis=-1).astype(np.float64)
    X = _jitter_nodes_box(X, np.array([hx, hy, hz]), jitter, seed, per)
    coords = X.reshape(-1, 3)
    if node_map is not None:
        coords = np.ascontiguousarray(node_map(coords.copy()), np.float64)
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k  # noqa: E731
    lat = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    canon = nid(I % nx if per[0] else I, J % ny if per[1] else J, K % nz if per[2] else K).ravel()
    c = np.arange(nx * ny * nz)
    ci, cj, ck = c % nx, (c // nx) % ny, c // (nx * ny)
    tets = []
    for p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        o = np.zeros((3, len(c)), np.int64)
        path = [nid(ci, cj, ck)]
        for ax in p:
            o[ax] += 1
            path.append(nid(ci + o[0], cj + o[1], ck + o[2]))
        tets.append(np.stack(path, 1))
    cells = np.stack(tets, 1).reshape(-1, 4)
    return simplex_mesh(3, coords, cells, canon, _lattice_zone(lat, (nx, ny, nz), per, zones), (nx, ny, nz))


def bump_channel_tet(nx: int, ny: int, nz: int, L: float = 3.0, H: float = 1.0, W: float = 0.5,
                     amp: float = 0.1, jitter: float = 0.0, seed: int = 0) -> Mesh:
    """The bump channel of bump_channel_hex (same wall-following grid lines, zones 0 inlet,
    1 outlet, 2 lower wall, 3 upper boundary, spanwise periodic) on tetrahedra: each
    hexahedron of the nx x ny x nz grid split into six Tet4 cells (6 nx ny nz cells)."""
    def bump(P):
        x, y = P[:, 0], P[:, 1]
        s = np.where((x > L / 3) & (x < 2 * L / 3), np.sin(3 * np.pi * x / L - np.pi) ** 2, 0.0)
        Q = P.copy()
        Q[:, 1] = y + amp * H * s * (1 - y / H)
        return Q

    return tet_grid(nx, ny, nz, L, H, W, periodic=(False, False, True),
                    zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=jitter, seed=seed, node_map=bump)


def permute_cells(m: Mesh, order: np.ndarray) -> Mesh:
    """The same mesh with cell order[k] renumbered k.  Faces are reordered by their back
    cell's new index (stable), keeping each face's orientation, node order, shift and
    zone; cell2face keeps each cell's element face order.  A field u of the original
    mesh corresponds to u[order] on the new one."""
    order = np.asarray(order, np.int64)
    N = m.nCell
    if order.shape != (N,) or not np.array_equal(np.sort(order), np.arange(N)):
        raise ValueError("order must be a permutation of the cells")
    new_of = np.empty(N, np.int64)
    new_of[order] = np.arange(N)
    f2c = m.face2cell
    f2c_new = np.where(f2c >= 0, new_of[np.maximum(f2c, 0)], -1)
    forder = np.argsort(f2c_new[:, 0], kind="stable")
    fnew_of = np.empty(m.nFace, np.int64)
    fnew_of[forder] = np.arange(m.nFace)
    cptr = m.cell2nodePtr
    nper = np.diff(cptr)
    fptr = m.cell2facePtr
    fper = np.diff(fptr)
    if np.any(nper != nper[0]) or np.any(fper != fper[0]):
        raise ValueError("permute_cells expects one cell type")
    c2n = m.cell2node.reshape(N, nper[0])[order].ravel()
    c2f = fnew_of[m.cell2face.reshape(N, fper[0])[order]].ravel()
    fn = np.diff(m.face2nodePtr)
    if np.any(fn != fn[0]):
        raise ValueError("permute_cells expects one face type")
    f2n = m.face2node.reshape(m.nFace, fn[0])[forder].ravel()
    return Mesh(m.dim, m.coords, m.cellType[order], m.cell2nodePtr.copy(), c2n, m.cell2facePtr.copy(), c2f,
                m.faceType[forder], m.face2nodePtr.copy(), f2n, f2c_new[forder], m.faceShift[forder], m.shape,
                None if m.faceZone is None else m.faceZone[forder], m.nBCZone)


def colour_sorted(m: Mesh, colors: np.ndarray) -> Mesh:
    """The mesh with its cells grouped by colour (stable within a colour), colors being
    the library's multicolouring of m (VariationalReconstruction.colors()).  The library
    colours greedily in cell order, so on the new mesh it finds the same colour classes,
    now contiguous index ranges: each Gauss-Seidel colour pass then reads its own cells'
    rows coalesced instead of every few elements."""
    return permute_cells(m, np.argsort(np.asarray(colors), kind="stable"))


def bump_box_tet(n: int = 69, amp: float = 0.1, jitter: float = 0.1, seed: int = 20261202) -> Mesh:
    """Unit box [0,1]^3 on an n^3 lattice (n + 1 points per direction, vertices jittered
    uniformly by +-jitter of the edge with numpy's PCG64 seeded by seed, boundary
    vertices only within their side), each cube split into six Tet4 cells (6 n^3 cells;
    n = 69: 1 971 054), the lower wall raised by a bump y_w = amp sin^2(pi x) with the
    grid lines following it (y' = y + y_w (1 - y)).  z periodic.  Zones: 0 inlet (x = 0),
    1 outlet (x = 1), 2 lower wall, 3 upper boundary."""
    def bump(P):
        Q = P.copy()
        Q[:, 1] = P[:, 1] + amp * np.sin(np.pi * P[:, 0]) ** 2 * (1 - P[:, 1])
        return Q

    return tet_grid(n, n, n, 1.0, 1.0, 1.0, periodic=(False, False, True),
                    zones=dict(xmin=0, xmax=1, ymin=2, ymax=3), jitter=jitter, seed=seed, node_map=bump)


def bump_box_delaunay(n: int = 66, amp: float = 0.1, jitter: float = 0.1, seed: int = 20261202) -> Mesh:
    """Unit box [0,1]^3: the n^3 lattice's (n + 1)^3 vertices jittered uniformly by +-jitter of
    the edge (numpy PCG64 seeded by seed; boundary vertices move only within their side) and
    Delaunay-tetrahedralised (scipy.spatial.Delaunay, Qhull), then the lower wall raised by the
    bump y_w = amp sin^2(pi x) with the grid following it (y' = y + y_w (1 - y)).  About
    6.95 n^3 tetrahedra (n = 66: 1.998M), with the slivers a Delaunay mesh of jittered points
    carries.  Cells come in Qhull's order (see reorder_cells).  Zones: 0 inlet (x = 0), 1 outlet
    (x = 1), 2 lower wall, 3 upper boundary, 4 the spanwise sides (z = 0, 1)."""
    from scipy.spatial import Delaunay
    h = 1.0 / n
    I, J, K = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    X = np.stack([I * h, J * h, K * h], axis=-1).astype(np.float64)
    X = _jitter_nodes_box(X, np.array([h, h, h]), jitter, seed, (False, False, False))
    P = X.reshape(-1, 3)
    cells = Delaunay(P).simplices.astype(np.int64)
    Q = P.copy()
    Q[:, 1] = P[:, 1] + amp * np.sin(np.pi * P[:, 0]) ** 2 * (1 - P[:, 1])
    lat = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    zones = dict(xmin=0, xmax=1, ymin=2, ymax=3, zmin=4, zmax=4)
    return simplex_mesh(3, Q, cells, None, _lattice_zone(lat, (n, n, n), (False, False, False), zones), (n, n, n))


def reorder_cells(m: Mesh, method: str = "rcm") -> Mesh:
    """Locality renumbering of the cells of an unstructured mesh (the role of
    UnstructuredMesh::ReorderLocalCells, src/Geom/Mesh/Mesh_Reorder.cpp:662-700, whose
    permutation is a Metis partition with a reverse Cuthill-McKee order inside each part,
    Mesh_CellPermutation.hpp:42-130; one GPU holds one part).  "rcm": reverse Cuthill-McKee on
    the face-adjacency graph (scipy.sparse.csgraph), which bounds the index distance between
    face neighbours by the graph bandwidth; "morton": cells sorted along the Z-order curve of
    their node-mean points.  Faces follow their back cell (permute_cells)."""
    N = m.nCell
    if method == "rcm":
        import scipy.sparse as sp
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        f2c = m.face2cell
        inner = (f2c[:, 1] >= 0) & (f2c[:, 0] != f2c[:, 1])
        a, b = f2c[inner, 0], f2c[inner, 1]
        g = sp.coo_matrix((np.ones(2 * len(a), np.int8), (np.r_[a, b], np.r_[b, a])), shape=(N, N)).tocsr()
        order = np.asarray(reverse_cuthill_mckee(g, symmetric_mode=True), np.int64)
    elif method == "morton":
        ptr = m.cell2nodePtr
        cnt = np.diff(ptr)
        c = np.add.reduceat(m.coords[m.cell2node], ptr[:-1], axis=0) / cnt[:, None]
        lo, hi = c.min(0), c.max(0)
        q = np.clip(((c - lo) / np.maximum(hi - lo, 1e-300) * (1 << 20)).astype(np.int64), 0, (1 << 20) - 1)
        key = np.zeros(N, np.int64)
        for bit in range(20):
            for ax in range(3):
                key |= ((q[:, ax] >> bit) & 1) << (3 * bit + ax)
        order = np.argsort(key, kind="stable")
    else:
        raise ValueError(f"unknown reorder method {method!r}")
    return permute_cells(m, order)


def bandwidth(m: Mesh) -> int:
    """Largest index distance between the two cells of an interior face."""
    f2c = m.face2cell
    inner = f2c[:, 1] >= 0
    return int(np.abs(f2c[inner, 0] - f2c[inner, 1]).max()) if inner.any() else 0


QUAD4_FACES = np.array([[0, 1], [1, 2], [2, 3], [3, 0]])


def mixed2d_mesh(coords: np.ndarray, cells: list, canon: np.ndarray | None = None, bnd_zone=None,
                 shape: tuple = ()) -> Mesh:
    """2-D mesh of Tri3 and Quad4 cells (cells: a list of node-id sequences of length 3
    or 4, counter-clockwise).  Faces (Line2) are matched as in simplex_mesh: numbered by
    first occurrence in (cell, element face) order, the back cell's edge orientation
    kept (normal out of it), periodic copies identified through canon, a face seen
    once is a boundary face with zone bnd_zone(face2node [nb, 2])."""
    coords = np.ascontiguousarray(coords, np.float64)
    nNode = len(coords)
    ncn = np.array([len(c) for c in cells], np.int64)
    if np.any((ncn != 3) & (ncn != 4)):
        raise ValueError("cells must have 3 or 4 nodes")
    N = len(cells)
    c2nPtr = np.concatenate([[0], np.cumsum(ncn)])
    c2n = np.concatenate([np.asarray(c, np.int64) for c in cells])
    # signed area (shoelace) must be positive
    x, y = coords[c2n, 0], coords[c2n, 1]
    nxt = np.arange(len(c2n)) + 1
    nxt[c2nPtr[1:] - 1] = c2nPtr[:-1]
    area2 = np.add.reduceat(x * y[nxt] - x[nxt] * y, c2nPtr[:-1])
    if np.any(area2 <= 0):
        raise ValueError("cells must be counter-clockwise and non-degenerate")
    # one occurrence per cell edge: node k -> node k+1 (Tri3 / Quad4 face tables)
    occ = np.stack([c2n, c2n[nxt]], 1)
    occ_cell = np.repeat(np.arange(N), ncn)
    canon = np.arange(nNode, dtype=np.int64) if canon is None else np.asarray(canon, np.int64)
    cn = np.sort(canon[occ], axis=1)
    base = int(canon.max()) + 1
    key = cn[:, 0] * base + cn[:, 1]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    start = np.r_[True, ks[1:] != ks[:-1]]
    gid = np.cumsum(start) - 1
    counts = np.bincount(gid)
    if counts.max() > 2:
        raise ValueError("a face is shared by more than two cells")
    pos = np.flatnonzero(start)
    first = order[pos]
    second = np.where(counts == 2, order[np.minimum(pos + 1, len(order) - 1)], -1)
    fo = np.argsort(first, kind="stable")
    first, second = first[fo], second[fo]
    F = len(first)
    face_of_group = np.empty(F, np.int64)
    face_of_group[fo] = np.arange(F)
    occ_face = np.empty(len(occ), np.int64)
    occ_face[order] = face_of_group[gid]
    face2node = occ[first]
    has2 = second >= 0
    face2cell = np.stack([occ_cell[first], np.where(has2, occ_cell[np.maximum(second, 0)], -1)], 1)
    faceShift = np.zeros((F, 3))
    if has2.any():
        a, b = occ[first[has2]], occ[second[has2]]
        ia, ib = np.argmin(canon[a], axis=1), np.argmin(canon[b], axis=1)
        r = np.arange(len(a))
        faceShift[has2] = coords[b[r, ib]] - coords[a[r, ia]]
    zone, nzone = None, 0
    if (~has2).any():
        if bnd_zone is None:
            raise ValueError("the mesh has boundary faces: bnd_zone is required")
        zone = np.full(F, -1, np.int32)
        zone[~has2] = np.asarray(bnd_zone(face2node[~has2]), np.int32)
        nzone = int(zone.max()) + 1
    ctype = np.where(ncn == 3, TRI3, QUAD4).astype(np.int32)
    return Mesh(2, coords, ctype, c2nPtr, c2n, c2nPtr.copy(), occ_face, np.full(F, LINE2, np.int32),
                np.arange(F + 1, dtype=np.int64) * 2, face2node.ravel().astype(np.int64), face2cell.astype(np.int64),
                faceShift, shape, zone, nzone)


def mixed_grid(nx: int, ny: int, Lx: float = 1.0, Ly: float = 1.0, x0: float = 0.0, y0: float = 0.0,
               periodic=(False, False), zones: dict | None = None, jitter: float = 0.0, seed: int = 0,
               node_map=None, split=None) -> Mesh:
    """nx x ny lattice of Quad4 cells in which the squares (i, j) with split(i, j) true
    are cut into two Tri3 cells along their (i, j)-(i+1, j+1) diagonal (default: the
    right half of the lattice); directions periodic or bounded as box_quad_grid."""
    assert nx >= 2 and ny >= 2
    per = [bool(p) for p in periodic]
    hx, hy = Lx / nx, Ly / ny
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    X = np.stack([x0 + I * hx, y0 + J * hy, np.zeros_like(I, dtype=float)], axis=-1).astype(np.float64)
    X = _jitter_nodes_box(X, np.array([hx, hy]), jitter, seed, per)
    coords = X.reshape(-1, 3)
    if node_map is not None:
        coords = np.ascontiguousarray(node_map(coords.copy()), np.float64)
    nid = lambda i, j: i * (ny + 1) + j  # noqa: E731
    lat = np.stack([I.ravel(), J.ravel(), 0 * I.ravel()], 1)
    canon = nid(I % nx if per[0] else I, J % ny if per[1] else J).ravel()
    split = split if split is not None else (lambda i, j: i >= nx // 2)
    cells = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if split(i, j):
                cells += [(a, b, c), (a, c, d)]
            else:
                cells.append((a, b, c, d))
    return mixed2d_mesh(coords, cells, canon, _lattice_zone(lat, (nx, ny), per, zones), (nx, ny))
