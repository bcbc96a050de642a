"""Set-up coefficient builders against closed forms."""
import numpy as np
import pytest

from hexdyn import setup as S
from hexdyn.mesh import icosahedral_mesh


@pytest.fixture(scope="module")
def mesh():
    return icosahedral_mesh(3)


def test_offcentering_closed_form():
    # heights below the 30 km transition: every epsilon is the 0.1 minimum
    rdzw = (1.0 / np.full(10, 1000.0)).astype(np.float32)
    etp, etm, ewp, ewm = S.acoustic_offcentering(rdzw)
    assert np.all(etp == np.float32(0.5) * (np.float32(1) + np.float32(0.1)))
    assert np.all(etm == np.float32(0.5) * (np.float32(1) - np.float32(0.1)))
    assert np.all(ewp == etp[0]) and np.all(ewm == etm[0])
    # inside the transition the profile rises as sin^2
    rdzw = (1.0 / np.full(40, 1000.0)).astype(np.float32)
    etp, etm, ewp, ewm = S.acoustic_offcentering(rdzw)
    top = 40000.0
    eps = 0.1 + np.sin(0.5 * np.pi * (top - 30000.0) / 20000.0) ** 2 * 0.4
    assert abs(float(ewp[-1]) - 0.5 * (1 + eps)) < 1e-6


def test_damping_closed_form():
    z = np.linspace(0, 30000, 31, dtype=np.float64)[:, None].repeat(3, 1).astype(np.float32)
    dss = S.vertical_velocity_damping(z, xnutr=0.2, zd=22000.0)
    mid = 0.5 * (z[:-1, 0] + z[1:, 0])
    exact = np.where(mid > 22000.0, 0.2 * np.sin(0.5 * np.pi * (mid - 22000.0) / 8000.0) ** 2, 0.0)
    assert np.allclose(dss[:, 0], exact, atol=1e-6)
    assert np.all(S.vertical_velocity_damping(z, xnutr=0.0, zd=22000.0) == 0)


def test_deriv_two_exact_on_quadratics(mesh):
    d2 = S.deriv_two(mesh)
    # second derivative along the direction to the edge of a quadratic in the local tangent plane
    R = mesh.sphere_radius
    pc = np.stack([mesh.xCell, mesh.yCell, mesh.zCell], 1) / R
    e1, e2 = S._tangent_basis(pc)
    rng = np.random.default_rng(3)
    a3, a4, a5 = rng.standard_normal(3) / R ** 2
    worst = 0.0
    for c in range(0, mesh.n_cells, 37):
        n = mesh.nEdgesOnCell[c]
        nb = mesh.cellsOnCell[c, :n]
        pts = np.concatenate([[c], nb])
        p0 = pc[c]

        def local(p):
            d = p - (p @ p0)[:, None] * p0[None, :]
            ang = np.arctan2(d @ e2[c], d @ e1[c])
            dist = np.arctan2(np.linalg.norm(np.cross(p0[None, :], p), axis=-1), p @ p0) * R
            return np.cos(ang) * dist, np.sin(ang) * dist

        x, y = local(pc[pts])
        f = a3 * x * x + a4 * x * y + a5 * y * y + 2.0 * x - y + 5.0
        for s in range(n):
            e = mesh.edgesOnCell[c, s]
            side = 0 if mesh.cellsOnEdge[e, 0] == c else 1
            pv = np.stack([mesh.xVertex, mesh.yVertex, mesh.zVertex], 1)[mesh.verticesOnEdge[e]] / R
            mid = pv.sum(0) / np.linalg.norm(pv.sum(0))
            mx, my = local(mid[None, :])
            th = np.arctan2(my[0], mx[0])
            exact = 2 * a3 * np.cos(th) ** 2 + 2 * a4 * np.cos(th) * np.sin(th) + 2 * a5 * np.sin(th) ** 2
            got = d2[e, side, : n + 1] @ f
            worst = max(worst, abs(got - exact) / (abs(exact) + 1e-30))
    # normal equations in metres (as MPAS builds them) carry ~1e-6 relative conditioning error
    assert worst < 1e-5


def test_advection_coefficient_properties(mesh):
    d2 = S.deriv_two(mesh)
    adv, adv3, nadv, cells = S.advection_coefficients(mesh, d2)
    dv = mesh.dvEdge.astype(np.float32)
    # a constant field has edge value 1 (times dvEdge) and no upwind correction
    assert np.allclose(adv.sum(1) / dv, 1.0, rtol=1e-5)
    assert np.abs(adv3.sum(1) / dv).max() < 1e-5
    # native insertion order: the two owners come first
    assert np.all(cells[:, 0] == mesh.cellsOnEdge[:, 0]) and np.all(cells[:, 1] == mesh.cellsOnEdge[:, 1])
    assert set(np.unique(nadv)) <= {9, 10}
    # a smooth field (a degree-1 harmonic) is interpolated to the edge point to 4th order:
    # far closer than the plain two-cell average, whose curvature error is O(dc^2)
    R = mesh.sphere_radius
    pc = np.stack([mesh.xCell, mesh.yCell, mesh.zCell], 1) / R
    g = np.array([0.2, -0.4, 0.9])
    f = pc @ g
    val = (adv * np.where(cells >= 0, f[np.maximum(cells, 0)], 0.0)).sum(1) / dv
    p1, p2 = pc[mesh.cellsOnEdge[:, 0]], pc[mesh.cellsOnEdge[:, 1]]
    mid = (p1 + p2) / np.linalg.norm(p1 + p2, axis=1, keepdims=True)
    exact = mid @ g
    avg = 0.5 * (f[mesh.cellsOnEdge[:, 0]] + f[mesh.cellsOnEdge[:, 1]])
    assert np.abs(val - exact).max() < 0.5 * np.abs(avg - exact).max()


def _edge_winds(mesh, east, north):
    """Normal and tangential edge components of an (east, north) wind."""
    ang = mesh.angleEdge
    return east * np.cos(ang) + north * np.sin(ang), -east * np.sin(ang) + north * np.cos(ang)


def _deformation(mesh, weights, un, vt):
    c2, s2, cs = (w.astype(np.float64) for w in weights)
    out = [np.zeros(mesh.n_cells) for _ in range(4)]
    for s in range(mesh.max_edges):
        ok = s < mesh.nEdgesOnCell
        e = mesh.edgesOnCell[:, s]
        u, v = np.where(ok, un[e], 0.0), np.where(ok, vt[e], 0.0)
        out[0] += c2[:, s] * u - cs[:, s] * v      # du/dx
        out[1] += cs[:, s] * u - s2[:, s] * v      # du/dy
        out[2] += cs[:, s] * u + c2[:, s] * v      # dv/dx
        out[3] += s2[:, s] * u + cs[:, s] * v      # dv/dy
    return out


def test_deformation_weights_differentiate_linear_winds():
    # The weights are the divergence-theorem gradient on the cell's tangent
    # plane (x east, y north): a linear shear u = a*y must give du/dy = a and
    # zero elsewhere; a pure stretch v = b*y gives dv/dy = b.
    m = icosahedral_mesh(4)
    w = S.deformation_weights(m)
    R = m.sphere_radius
    lat_e = np.arcsin(m.zEdge / R)
    near_eq = np.abs(np.arcsin(m.zCell / R)) < 0.2
    a = 1e-5
    un, vt = _edge_winds(m, a * R * lat_e, 0.0 * lat_e)
    dudx, dudy, dvdx, dvdy = _deformation(m, w, un, vt)
    assert np.median(np.abs(dudy[near_eq] / a - 1.0)) < 0.01
    for g in (dudx, dvdx, dvdy):
        assert np.median(np.abs(g[near_eq])) < 0.02 * a
    un, vt = _edge_winds(m, 0.0 * lat_e, a * R * lat_e)
    dudx, dudy, dvdx, dvdy = _deformation(m, w, un, vt)
    assert np.median(np.abs(dvdy[near_eq] / a - 1.0)) < 0.01
    assert np.median(np.abs(dudx[near_eq])) < 0.02 * a


def test_deformation_weights_sign_follows_edge_orientation():
    # times the cell's outward sign the weights are orientation free: over a
    # cell's sides c2 + s2 = dl/A sums to the tangent-plane perimeter over
    # area, which is the spherical perimeter over area to O(dc^2/R^2)
    m = icosahedral_mesh(3)
    c2, s2, cs = (x.astype(np.float64) for x in S.deformation_weights(m))
    out_sign = np.where(m.cellsOnEdge[m.edgesOnCell, 0] == np.arange(m.n_cells)[:, None], 1.0, -1.0)
    used = np.arange(m.max_edges)[None, :] < m.nEdgesOnCell[:, None]
    per_over_area = np.where(used, (c2 + s2) * out_sign, 0.0).sum(axis=1)
    dl_sum = np.where(used, m.dvEdge[m.edgesOnCell], 0.0).sum(axis=1)
    assert np.allclose(per_over_area, dl_sum / m.areaCell, rtol=2e-2)
    # a flipped sign on any one side would move that cell's sum by 2 dl/A
    assert np.all(per_over_area > 0.5 * dl_sum / m.areaCell)
