"""Mesh generator: MPAS table invariants and TRiSK/discrete-operator identities."""
import numpy as np
import pytest

from hexdyn.mesh import build_mesh, edge_signs_on_cells, edge_signs_on_vertices, icosahedral_mesh, random_points


@pytest.fixture(scope="module", params=["icosahedral", "ragged"])
def mesh(request):
    """The quasi-uniform x1.642 mesh, and a ragged one from 642 random
    generators (pentagons, hexagons and heptagons, maxEdges 7)."""
    if request.param == "icosahedral":
        return icosahedral_mesh(3)
    return build_mesh(random_points(642, seed=1, lloyd_iterations=31))


def _accuracy_factor(mesh):
    """Discretization-error bounds (not the exact identities) are looser on
    the ragged mesh, whose cells are far from centroidal and regular."""
    return 1.0 if mesh.max_edges == 6 else 5.0


def test_counts_and_euler(mesh):
    assert mesh.n_cells == 642
    assert mesh.n_cells - mesh.n_edges + mesh.n_vertices == 2
    counts = np.bincount(mesh.nEdgesOnCell, minlength=8)
    if mesh.max_edges == 6:
        assert counts[5] == 12
    else:
        assert mesh.max_edges == 7 and counts[7] > 0
    # Euler on the sphere: sum over cells of (6 - sides) = 12
    assert int((6 - mesh.nEdgesOnCell).sum()) == 12


def test_areas_partition(mesh):
    R = mesh.sphere_radius
    assert abs(mesh.areaCell.sum() / (4 * np.pi * R * R) - 1) < 1e-12
    assert np.allclose(mesh.kiteAreasOnVertex.sum(1), mesh.areaTriangle, rtol=1e-12)


def test_orientation_conventions(mesh):
    # each edge appears once with +1 and once with -1 among its two cells
    s = edge_signs_on_cells(mesh)
    acc = np.zeros(mesh.n_edges)
    for j in range(mesh.max_edges):
        c = np.flatnonzero(mesh.nEdgesOnCell > j)
        np.add.at(acc, mesh.edgesOnCell[c, j], s[c, j])
    assert np.all(acc == 0)
    sv = edge_signs_on_vertices(mesh)
    acc = np.zeros(mesh.n_edges)
    np.add.at(acc, mesh.edgesOnVertex.ravel(), sv.ravel())
    assert np.all(acc == 0)


def _solid_body(mesh, axis=(0.0, 0.0, 1.0), speed=20.0):
    R = mesh.sphere_radius
    xe = np.stack([mesh.xEdge, mesh.yEdge, mesh.zEdge], 1)
    vel = np.cross(np.array(axis), xe) / R * speed
    pc = np.stack([mesh.xCell, mesh.yCell, mesh.zCell], 1)
    n = pc[mesh.cellsOnEdge[:, 1]] - pc[mesh.cellsOnEdge[:, 0]]
    n -= xe * (np.einsum("ij,ij->i", n, xe) / R / R)[:, None]
    n /= np.linalg.norm(n, axis=1)[:, None]
    t = np.cross(xe / R, n)
    return np.einsum("ij,ij->i", vel, n), np.einsum("ij,ij->i", vel, t)


def test_trisk_antisymmetry(mesh):
    W = {}
    for e in range(mesh.n_edges):
        for j in range(mesh.nEdgesOnEdge[e]):
            n = mesh.edgesOnEdge[e, j]
            W[(e, n)] = mesh.weightsOnEdge[e, j] * mesh.dcEdge[e] / mesh.dvEdge[n]
    assert max(abs(W[(a, b)] + W[(b, a)]) for (a, b) in W) < 1e-12


@pytest.mark.parametrize("axis", [(0, 0, 1.0), (1.0, 0, 0), (0.3, -0.5, 0.8)])
def test_tangential_reconstruction(mesh, axis):
    axis = np.array(axis) / np.linalg.norm(axis)
    u, vt = _solid_body(mesh, axis)
    v = np.zeros(mesh.n_edges)
    for j in range(mesh.edgesOnEdge.shape[1]):
        a = mesh.nEdgesOnEdge > j
        v[a] += mesh.weightsOnEdge[a, j] * u[mesh.edgesOnEdge[a, j]]
    f = _accuracy_factor(mesh)
    assert np.sqrt(np.mean((v - vt) ** 2)) < 0.01 * 20.0 * f
    assert np.abs(v - vt).max() < 0.03 * 20.0 * f


def test_divergence_and_curl_identities(mesh):
    u, _ = _solid_body(mesh)
    s = edge_signs_on_cells(mesh)
    div = np.zeros(mesh.n_cells)
    for j in range(mesh.max_edges):
        c = np.flatnonzero(mesh.nEdgesOnCell > j)
        e = mesh.edgesOnCell[c, j]
        div[c] += s[c, j] * mesh.dvEdge[e] * u[e]
    # point-sampled flux of a divergence-free flow: zero up to quadrature error
    f = _accuracy_factor(mesh)
    assert np.abs(div / mesh.areaCell).max() < 1e-3 * 20.0 / mesh.dcEdge.mean() * f
    # curl of a gradient vanishes: u = grad(phi) along edges
    rng = np.random.default_rng(0)
    phi = rng.standard_normal(mesh.n_cells)
    g = (phi[mesh.cellsOnEdge[:, 1]] - phi[mesh.cellsOnEdge[:, 0]]) / mesh.dcEdge
    sv = edge_signs_on_vertices(mesh)
    circ = (sv * mesh.dcEdge[mesh.edgesOnVertex] * g[mesh.edgesOnVertex]).sum(1)
    assert np.abs(circ).max() < 1e-12
    # solid-body vorticity approximates 2*omega*sin(lat)
    vort = (sv * mesh.dcEdge[mesh.edgesOnVertex] * u[mesh.edgesOnVertex]).sum(1) / mesh.areaTriangle
    exact = 2 * 20.0 / mesh.sphere_radius * np.sin(mesh.latVertex)
    assert np.abs(vort - exact).max() < 0.02 * np.abs(exact).max() * f
