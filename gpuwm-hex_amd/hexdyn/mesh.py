"""Quasi-uniform icosahedral Voronoi (SCVT) meshes with MPAS connectivity.

The reference reads MPAS ``x1.N.grid.nc`` files (``mesh.py:1374``
``Mesh.from_netcdf``) and consumes the tables listed in
``tools/mpas_operator_oracle/operator_oracle.F90:61-147``.  This module
produces the same tables for a spherical centroidal Voronoi tessellation
built here: an icosahedron subdivided ``level`` times, a seeded jitter of the
generators, and Lloyd iterations on the sphere.  Every table uses 0-based
indices; missing slots hold -1.

Orientation conventions (the ones the operators rely on):

* ``cellsOnEdge[e] = (c1, c2)``: the edge normal points from c1 to c2;
* ``verticesOnEdge[e] = (v1, v2)``: v2 lies on the left of the normal
  (k x n), so the circulation around v2 counts +u;
* ``edgesOnCell``/``cellsOnCell``/``verticesOnCell`` run counter-clockwise;
  edge i joins vertex i and vertex i+1 (the "edge-start" slot convention of
  the NCAR x1 meshes, ``mesh.py:840-848``), so vertex i lies between edge
  i-1 and edge i;
* ``cellsOnVertex`` runs counter-clockwise, ``edgesOnVertex[v, j]`` joins
  ``cellsOnVertex[v, j]`` and ``cellsOnVertex[v, j+1]`` and
  ``kiteAreasOnVertex[v, j]`` is the part of triangle v inside cell
  ``cellsOnVertex[v, j]``;
* ``weightsOnEdge``/``edgesOnEdge`` are the TRiSK tangential weights
  (Thuburn et al. 2009, Ringler et al. 2010) with ``dvEdge/dcEdge`` folded in.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict

import numpy as np
from scipy.spatial import ConvexHull

EARTH_RADIUS = 6371229.0
OMEGA = 7.29212e-5


def _normalize(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _icosahedron() -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = np.array(
        [
            [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
            [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
            [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
        ],
        dtype=np.float64,
    )
    faces = np.array(
        [
            [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
            [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
            [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
            [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
        ],
        dtype=np.int64,
    )
    return _normalize(verts), faces


def _subdivide(verts: np.ndarray, faces: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    pairs = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    pairs.sort(axis=1)
    unique, inverse = np.unique(pairs, axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)
    mids = _normalize(0.5 * (verts[unique[:, 0]] + verts[unique[:, 1]]))
    base = verts.shape[0]
    nf = faces.shape[0]
    m01 = base + inverse[:nf]
    m12 = base + inverse[nf : 2 * nf]
    m20 = base + inverse[2 * nf :]
    a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
    new_faces = np.concatenate(
        [
            np.stack([a, m01, m20], axis=1),
            np.stack([b, m12, m01], axis=1),
            np.stack([c, m20, m12], axis=1),
            np.stack([m01, m12, m20], axis=1),
        ]
    )
    return np.concatenate([verts, mids]), new_faces


def _oriented_triangles(points: np.ndarray) -> np.ndarray:
    tri = ConvexHull(points).simplices.astype(np.int64)
    a, b, c = points[tri[:, 0]], points[tri[:, 1]], points[tri[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a) < 0.0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return tri


def _spherical_triangle_area(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """Signed area of unit-sphere triangles (Van Oosterom & Strackee)."""
    num = np.einsum("ij,ij->i", a, np.cross(b, c))
    den = 1.0 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + np.einsum("ij,ij->i", c, a)
    return 2.0 * np.arctan2(num, den)


def _circumcenters(points: np.ndarray, tri: np.ndarray) -> np.ndarray:
    a, b, c = points[tri[:, 0]], points[tri[:, 1]], points[tri[:, 2]]
    return _normalize(np.cross(b - a, c - a))


def _kite_parts(points: np.ndarray, tri: np.ndarray, cc: np.ndarray):
    """Areas and area-weighted centroids of the 3 kites of every triangle."""
    areas = np.empty((tri.shape[0], 3))
    moments = np.empty((tri.shape[0], 3, 3))
    for j in range(3):
        p = points[tri[:, j]]
        q = points[tri[:, (j + 1) % 3]]
        r = points[tri[:, (j + 2) % 3]]
        m_pq = _normalize(p + q)
        m_pr = _normalize(p + r)
        a1 = _spherical_triangle_area(p, m_pq, cc)
        a2 = _spherical_triangle_area(p, cc, m_pr)
        areas[:, j] = a1 + a2
        moments[:, j, :] = a1[:, None] * (p + m_pq + cc) / 3.0 + a2[:, None] * (p + cc + m_pr) / 3.0
    return areas, moments


def generate_points(level: int, *, seed: int = 1, lloyd_iterations: int = 20, jitter: float = 0.02) -> np.ndarray:
    """Icosahedral generators, a seeded jitter, then spherical Lloyd iterations."""
    verts, faces = _icosahedron()
    for _ in range(level):
        verts, faces = _subdivide(verts, faces)
    rng = np.random.default_rng(seed)
    spacing = np.sqrt(4.0 * np.pi / verts.shape[0])
    points = _normalize(verts + jitter * spacing * rng.standard_normal(verts.shape))
    return _lloyd(points, lloyd_iterations)


def random_points(n: int, *, seed: int = 1, lloyd_iterations: int = 31) -> np.ndarray:
    """Seeded uniform random generators on the sphere relaxed by a few
    spherical Lloyd iterations: a ragged Voronoi mesh of pentagons, hexagons
    and heptagons (maxEdges 7 for n = 642, seed 1, 31 iterations)."""
    rng = np.random.default_rng(seed)
    return _lloyd(_normalize(rng.standard_normal((n, 3))), lloyd_iterations)


def _lloyd(points: np.ndarray, iterations: int) -> np.ndarray:
    for _ in range(iterations):
        tri = _oriented_triangles(points)
        cc = _circumcenters(points, tri)
        areas, moments = _kite_parts(points, tri, cc)
        acc = np.zeros_like(points)
        for j in range(3):
            np.add.at(acc, tri[:, j], moments[:, j, :])
        points = _normalize(acc)
    return points


@dataclass
class Mesh:
    """MPAS mesh tables (0-based), geometry in metres on a sphere of radius R."""

    arrays: Dict[str, np.ndarray] = field(default_factory=dict)
    sphere_radius: float = EARTH_RADIUS

    def __getattr__(self, name: str) -> np.ndarray:
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    @property
    def n_cells(self) -> int:
        return int(self.arrays["areaCell"].size)

    @property
    def n_edges(self) -> int:
        return int(self.arrays["dcEdge"].size)

    @property
    def n_vertices(self) -> int:
        return int(self.arrays["areaTriangle"].size)

    @property
    def max_edges(self) -> int:
        return int(self.arrays["edgesOnCell"].shape[1])


def _lat_lon(xyz: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    lat = np.arcsin(np.clip(xyz[:, 2], -1.0, 1.0))
    lon = np.mod(np.arctan2(xyz[:, 1], xyz[:, 0]), 2.0 * np.pi)
    return lat, lon


def _arc(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.einsum("ij,ij->i", a, b))


def build_mesh(points: np.ndarray, *, radius: float = EARTH_RADIUS) -> Mesh:
    """Build every MPAS table for the Voronoi dual of ``points`` (unit vectors)."""
    points = _normalize(np.asarray(points, dtype=np.float64))
    n_cells = points.shape[0]
    tri = _oriented_triangles(points)
    n_vertices = tri.shape[0]
    cc = _circumcenters(points, tri)
    kite, _ = _kite_parts(points, tri, cc)

    # Edges: unique Delaunay edges with both adjacent triangles.
    pairs = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    owner_tri = np.concatenate([np.arange(n_vertices)] * 3)
    key = np.sort(pairs, axis=1)
    edges, inverse = np.unique(key, axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)
    n_edges = edges.shape[0]
    # In a CCW triangle (a, b, c) the directed edge a->b has the triangle on its
    # left.  With the normal pointing c1 -> c2, the left vertex is v2.
    directed_same = pairs[:, 0] == edges[inverse, 0]
    left_tri = np.full(n_edges, -1)
    right_tri = np.full(n_edges, -1)
    left_tri[inverse[directed_same]] = owner_tri[directed_same]
    right_tri[inverse[~directed_same]] = owner_tri[~directed_same]
    if np.any(left_tri < 0) or np.any(right_tri < 0):
        raise ValueError("non-manifold Delaunay triangulation")
    cells_on_edge = edges.copy()
    vertices_on_edge = np.stack([right_tri, left_tri], axis=1)

    # Per-vertex (triangle) tables, counter-clockwise.  Pairs are looked up
    # through sorted int64 keys a*n + b.
    cells_on_vertex = tri.copy()
    n = np.int64(n_cells)
    edge_keys = edges[:, 0].astype(np.int64) * n + edges[:, 1]  # sorted (np.unique order)
    edges_on_vertex = np.empty((n_vertices, 3), dtype=np.int64)
    for j in range(3):
        a = tri[:, j]
        b = tri[:, (j + 1) % 3]
        k = np.minimum(a, b).astype(np.int64) * n + np.maximum(a, b)
        pos = np.searchsorted(edge_keys, k)
        if np.any(edge_keys[np.minimum(pos, n_edges - 1)] != k):
            raise ValueError("triangle side without an edge")
        edges_on_vertex[:, j] = pos
    kite_areas = kite * radius * radius
    area_triangle = _spherical_triangle_area(points[tri[:, 0]], points[tri[:, 1]], points[tri[:, 2]]) * radius * radius

    # Per-cell tables, counter-clockwise around the generator: every cell's
    # (neighbour, edge) entries, ordered by the angle of the neighbour in the
    # cell's tangent plane (east-north frame; x-based at the poles).
    owner = np.concatenate([edges[:, 0], edges[:, 1]]).astype(np.int64)
    other_cell = np.concatenate([edges[:, 1], edges[:, 0]]).astype(np.int64)
    edge_id = np.concatenate([np.arange(n_edges), np.arange(n_edges)]).astype(np.int64)
    order = np.lexsort((edge_id, owner))
    owner, other_cell, edge_id = owner[order], other_cell[order], edge_id[order]
    n_edges_on_cell = np.bincount(owner, minlength=n_cells).astype(np.int64)
    max_edges = int(n_edges_on_cell.max())
    start = np.concatenate([[0], np.cumsum(n_edges_on_cell)[:-1]])
    slot = np.arange(owner.size) - start[owner]
    nb_cell = np.full((n_cells, max_edges), -1, dtype=np.int64)
    nb_edge = np.full((n_cells, max_edges), -1, dtype=np.int64)
    nb_cell[owner, slot] = other_cell
    nb_edge[owner, slot] = edge_id
    north = np.array([0.0, 0.0, 1.0])
    ref = north[None, :] - points * points[:, 2:3]
    polar = np.linalg.norm(ref, axis=1) < 1e-8
    ref[polar] = np.array([1.0, 0.0, 0.0])[None, :] - points[polar] * points[polar, 0:1]
    ref = ref / np.linalg.norm(ref, axis=1, keepdims=True)
    other = np.cross(points, ref)
    valid = nb_cell >= 0
    d = points[np.where(valid, nb_cell, 0)] - points[:, None, :]
    ang = np.arctan2(np.einsum("csk,ck->cs", d, other), np.einsum("csk,ck->cs", d, ref))
    ang = np.where(valid, ang, np.inf)
    rank = np.argsort(ang, axis=1, kind="stable")
    cells_on_cell = np.take_along_axis(nb_cell, rank, axis=1)
    edges_on_cell = np.take_along_axis(nb_edge, rank, axis=1)
    # vertex between edges i-1 and i: the CCW triangle (c, n_{i-1}, n_i)
    tri_keys = np.concatenate([tri[:, 0].astype(np.int64) * n + tri[:, 1], tri[:, 1].astype(np.int64) * n + tri[:, 2],
                               tri[:, 2].astype(np.int64) * n + tri[:, 0]])
    tri_ids = np.concatenate([np.arange(n_vertices)] * 3).astype(np.int64)
    korder = np.argsort(tri_keys)
    tri_keys, tri_ids = tri_keys[korder], tri_ids[korder]

    def tri_lookup(a, b):
        k = a.astype(np.int64) * n + b
        pos = np.minimum(np.searchsorted(tri_keys, k), tri_keys.size - 1)
        if np.any(tri_keys[pos] != k):
            raise ValueError("cell neighbourhood is not a fan")
        return tri_ids[pos]

    cidx = np.repeat(np.arange(n_cells), max_edges).reshape(n_cells, max_edges)
    sl = np.arange(max_edges)[None, :]
    prev_slot = (sl - 1) % n_edges_on_cell[:, None]
    prv = np.take_along_axis(cells_on_cell, np.where(sl < n_edges_on_cell[:, None], prev_slot, 0), axis=1)
    vertices_on_cell = np.full((n_cells, max_edges), -1, dtype=np.int64)
    used = sl < n_edges_on_cell[:, None]
    v_a = tri_lookup(cidx[used], prv[used])
    v_b = tri_lookup(prv[used], cells_on_cell[used])
    if np.any(v_a != v_b):
        raise ValueError("cell neighbourhood is not a fan")
    vertices_on_cell[used] = v_a

    xyz_cell = points * radius
    xyz_vertex = cc * radius
    mid = _normalize(points[edges[:, 0]] + points[edges[:, 1]])
    xyz_edge = mid * radius
    dc_edge = _arc(points[edges[:, 0]], points[edges[:, 1]]) * radius
    dv_edge = _arc(cc[vertices_on_edge[:, 0]], cc[vertices_on_edge[:, 1]]) * radius
    area_cell = np.zeros(n_cells)
    for j in range(3):
        np.add.at(area_cell, tri[:, j], kite_areas[:, j])

    lat_cell, lon_cell = _lat_lon(points)
    lat_vertex, lon_vertex = _lat_lon(cc)
    lat_edge, lon_edge = _lat_lon(mid)
    normal = points[edges[:, 1]] - points[edges[:, 0]]
    normal = normal - mid * np.einsum("ij,ij->i", normal, mid)[:, None]
    normal = _normalize(normal)
    east = np.stack([-mid[:, 1], mid[:, 0], np.zeros(n_edges)], axis=1)
    east_norm = np.linalg.norm(east, axis=1, keepdims=True)
    east = np.where(east_norm > 1e-12, east / np.maximum(east_norm, 1e-300), np.array([[1.0, 0.0, 0.0]]))
    northv = np.cross(mid, east)
    angle_edge = np.arctan2(np.einsum("ij,ij->i", normal, northv), np.einsum("ij,ij->i", normal, east))

    # kiteForCell: slot of the cell in cellsOnVertex for each verticesOnCell entry
    arrays: Dict[str, np.ndarray] = {
        "xCell": xyz_cell[:, 0], "yCell": xyz_cell[:, 1], "zCell": xyz_cell[:, 2],
        "latCell": lat_cell, "lonCell": lon_cell,
        "xVertex": xyz_vertex[:, 0], "yVertex": xyz_vertex[:, 1], "zVertex": xyz_vertex[:, 2],
        "latVertex": lat_vertex, "lonVertex": lon_vertex,
        "xEdge": xyz_edge[:, 0], "yEdge": xyz_edge[:, 1], "zEdge": xyz_edge[:, 2],
        "latEdge": lat_edge, "lonEdge": lon_edge,
        "cellsOnEdge": cells_on_edge, "verticesOnEdge": vertices_on_edge,
        "edgesOnCell": edges_on_cell, "cellsOnCell": cells_on_cell,
        "verticesOnCell": vertices_on_cell, "nEdgesOnCell": n_edges_on_cell,
        "cellsOnVertex": cells_on_vertex, "edgesOnVertex": edges_on_vertex,
        "kiteAreasOnVertex": kite_areas, "areaTriangle": area_triangle,
        "areaCell": area_cell, "dcEdge": dc_edge, "dvEdge": dv_edge,
        "angleEdge": angle_edge,
        "fCell": 2.0 * OMEGA * np.sin(lat_cell),
        "fEdge": 2.0 * OMEGA * np.sin(lat_edge),
        "fVertex": 2.0 * OMEGA * np.sin(lat_vertex),
    }
    mesh = Mesh(arrays=arrays, sphere_radius=radius)
    _build_trisk_weights(mesh)
    arrays["kiteForCell"] = _kite_for_cell(mesh)
    arrays["nominalMinDc"] = np.array([dc_edge.min()])
    return mesh


def _kite_for_cell(mesh: Mesh) -> np.ndarray:
    voc = mesh.verticesOnCell
    cov = mesh.cellsOnVertex
    result = np.full(voc.shape, -1, dtype=np.int64)
    for slot in range(voc.shape[1]):
        active = mesh.nEdgesOnCell > slot
        cells = np.flatnonzero(active)
        v = voc[cells, slot]
        match = cov[v] == cells[:, None]
        if np.any(match.sum(axis=1) != 1):
            raise ValueError("cell does not appear exactly once at its vertex")
        result[cells, slot] = np.argmax(match, axis=1)
    return result


def _build_trisk_weights(mesh: Mesh) -> None:
    """TRiSK weights: v_e = sum_j weightsOnEdge[e, j] * u[edgesOnEdge[e, j]].

    For cell i of edge e, walk counter-clockwise from e over the other edges
    e' of i, accumulating R = sum of kiteArea(i, v)/areaCell(i) over the
    vertices passed.  The weight is (1/2 - R) * s * dvEdge(e')/dcEdge(e), with
    s the product of the outward signs of e and e' on cell i.
    """
    a = mesh.arrays
    coe = a["cellsOnEdge"]
    eoc = a["edgesOnCell"]
    voc = a["verticesOnCell"]
    neoc = a["nEdgesOnCell"]
    kite = a["kiteAreasOnVertex"]
    area = a["areaCell"]
    dc = a["dcEdge"]
    dv = a["dvEdge"]
    n_edges = coe.shape[0]
    max_edges = eoc.shape[1]
    max_eoe = 2 * max_edges
    kfc = _kite_for_cell(mesh)
    eoe = np.full((n_edges, max_eoe), -1, dtype=np.int64)
    weights = np.zeros((n_edges, max_eoe))
    count = np.zeros(n_edges, dtype=np.int64)
    for e in range(n_edges):
        for side in range(2):
            cell = coe[e, side]
            n = neoc[cell]
            slots = list(eoc[cell, :n])
            start = slots.index(e)
            sign_e = 1.0 if coe[e, 0] == cell else -1.0
            r = 0.0
            for step in range(1, n):
                j = (start + step) % n
                vertex_slot = (start + step) % n  # the vertex between slots j-1 and j
                vertex = voc[cell, vertex_slot]
                r += kite[vertex, kfc[cell, vertex_slot]] / area[cell]
                ep = eoc[cell, j]
                sign_ep = 1.0 if coe[ep, 0] == cell else -1.0
                eoe[e, count[e]] = ep
                weights[e, count[e]] = (0.5 - r) * sign_e * sign_ep * dv[ep] / dc[e]
                count[e] += 1
    width = int(count.max())
    a["edgesOnEdge"] = eoe[:, :width]
    a["weightsOnEdge"] = weights[:, :width]
    a["nEdgesOnEdge"] = count


def icosahedral_mesh(level: int, *, seed: int = 1, lloyd_iterations: int = 20, radius: float = EARTH_RADIUS) -> Mesh:
    """x1.(10*4**level+2): 2562 cells at level 4, 10242 at 5, 40962 at 6."""
    return build_mesh(generate_points(level, seed=seed, lloyd_iterations=lloyd_iterations), radius=radius)


def edge_signs_on_cells(mesh: Mesh) -> np.ndarray:
    """+1 where the cell is the edge's first cell, -1 otherwise, 0 past the count."""
    eoc = mesh.edgesOnCell
    signs = np.zeros(eoc.shape)
    for slot in range(eoc.shape[1]):
        cells = np.flatnonzero(mesh.nEdgesOnCell > slot)
        e = eoc[cells, slot]
        signs[cells, slot] = np.where(mesh.cellsOnEdge[e, 0] == cells, 1.0, -1.0)
    return signs


def edge_signs_on_vertices(mesh: Mesh) -> np.ndarray:
    """+1 where the vertex is the edge's second vertex, -1 otherwise."""
    eov = mesh.edgesOnVertex
    owner = np.arange(eov.shape[0])[:, None]
    return np.where(mesh.verticesOnEdge[eov, 1] == owner, 1.0, -1.0)
