"""Set-up-time coefficient builders for the dycore step.

Every array here is built once on the host and then read by every step, by
the HIP library and by the oracle alike, so both consume identical bits.

* RK schedule: ``integration.py:110-160`` (``RKSchedule.from_mpas``).
* Off-centring ``etp/etm/ewp/ewm``: ``offcentering_v841.py:45-174``.
* Rayleigh w damping ``dss``: ``damping_v841.py:23-107`` (the binary32 path, host sinf).
* Quadratic-fit second derivatives ``deriv_two``: ``transport.py:610``
  (``initialize_deriv_two``), the MPAS ``helpers.F`` least-squares method.
* Advection coefficient compression in native insertion order:
  ``transport.py:2277`` and ``transport.py:2413`` (core.F:1475-1529).
* Stored mesh inverses: ``dynamics_v841.py:28`` (``precomputed_mesh_inverse_v841``).
* Smagorinsky deformation weights ``coef_c2/s2/cs``: ``mixing_v841.py:372-455``
  (MPAS ``atm_initialize_deformation_weights``, spherical branch), evaluated
  here in binary64 and rounded once.
* Mesh scaling of del2/del4: ``mixing.py:271-318`` (``compute_mesh_mixing_scaling``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np

from .mesh import Mesh

F32 = np.float32


@dataclass(frozen=True)
class RKStage:
    stage: int
    large_timestep: float
    acoustic_timestep: float
    acoustic_steps: int


def rk_schedule(dt: float, *, order: int = 3, acoustic_substeps: int = 2, dynamics_splits: int = 1) -> Tuple[RKStage, ...]:
    """Stage timesteps and acoustic counts of ``integration.py:110-160``.

    The divisions are evaluated in binary64 exactly as the reference's Python
    does; consumers round each value to binary32 once.
    """
    if dt <= 0.0:
        raise ValueError("config_dt must be positive")
    if dynamics_splits < 1 or acoustic_substeps < 1:
        raise ValueError("splits and substeps must be positive")
    dt_dyn = dt / dynamics_splits
    if order == 3:
        large = (dt_dyn / 3.0, dt_dyn / 2.0, dt_dyn)
        acoustic = (dt_dyn / 3.0, dt_dyn / acoustic_substeps, dt_dyn / acoustic_substeps)
        counts = (1, max(1, acoustic_substeps // 2), acoustic_substeps)
    elif order == 2:
        large = (dt_dyn / 2.0, dt_dyn / 2.0, dt_dyn)
        acoustic = (dt_dyn / acoustic_substeps,) * 3
        counts = (max(1, acoustic_substeps // 2), max(1, acoustic_substeps // 2), acoustic_substeps)
    else:
        raise ValueError("config_time_integration_order must be 2 or 3")
    return tuple(RKStage(i + 1, large[i], acoustic[i], counts[i]) for i in range(3))


def acoustic_offcentering(rdzw: np.ndarray, *, minimum: float = 0.1, maximum: float = 0.5,
                          transition_bottom_z: float = 30000.0, transition_top_z: float = 50000.0):
    """``etp, etm`` (L) and ``ewp, ewm`` (L+1), binary32 source order.

    Heights are the cumulative sums of ``1/rdzw`` in binary32; the sine of the
    transition phase is taken in binary64 and rounded once.
    """
    rdzw = np.asarray(rdzw, dtype=F32)
    nlev = rdzw.size
    one, half = F32(1.0), F32(0.5)
    lo, hi = F32(transition_bottom_z), F32(transition_top_z)
    emin, emax = F32(minimum), F32(maximum)
    height_w = np.zeros(nlev + 1, dtype=F32)
    height_u = np.zeros(nlev, dtype=F32)
    for k in range(nlev):
        height_w[k + 1] = height_w[k] + one / rdzw[k]
        height_u[k] = half * (height_w[k + 1] + height_w[k])
    width = hi - lo
    pi = F32(np.pi)

    def eps(height: np.float32) -> np.float32:
        if height <= lo:
            return emin
        if height >= hi:
            return emax
        z = (height - lo) / width
        s = F32(np.sin(np.float64(half * pi * z)))
        return F32(emin + (s * s) * (emax - emin))

    etp = np.array([half * (one + eps(h)) for h in height_u], dtype=F32)
    etm = np.array([half * (one - eps(h)) for h in height_u], dtype=F32)
    ewp = np.array([half * (one + eps(h)) for h in height_w], dtype=F32)
    ewm = np.array([half * (one - eps(h)) for h in height_w], dtype=F32)
    return etp, etm, ewp, ewm


def _libm_f32(name: str, nargs: int = 1):
    """The host C library's binary32 function `name` (sinf, cosf, powf)."""
    import ctypes
    import ctypes.util

    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    fn = getattr(libm, name)
    fn.argtypes = [ctypes.c_float] * nargs
    fn.restype = ctypes.c_float
    return fn


def _host_sinf(x: np.ndarray) -> np.ndarray:
    """binary32 sine by the host C library's sinf, elementwise."""
    fn = _libm_f32("sinf")
    flat = np.asarray(x, dtype=F32).ravel()
    return np.fromiter((fn(float(v)) for v in flat), dtype=F32, count=flat.size).reshape(np.shape(x))


def _host_cosf(x: np.ndarray) -> np.ndarray:
    """binary32 cosine by the host C library's cosf, elementwise."""
    fn = _libm_f32("cosf")
    flat = np.asarray(x, dtype=F32).ravel()
    return np.fromiter((fn(float(v)) for v in flat), dtype=F32, count=flat.size).reshape(np.shape(x))


def angle_cos_sin(angle_edge: np.ndarray):
    """cos and sin of the binary32 angleEdge for the perturbation-Coriolis
    reference wind u_init*cos(angle) - v_init*sin(angle).  Native MPAS
    (gfortran, RKIND single) calls the C library's cosf/sinf; the reference's
    NumPy restatement takes np.cos/np.sin of each float32 angle as a scalar
    (dynamics_v841.py:160-165), which is the same library call.  Both are the
    host cosf/sinf here, evaluated once per edge."""
    a = np.asarray(angle_edge, dtype=F32)
    return _host_cosf(a), _host_sinf(a)


def vertical_velocity_damping(zgrid: np.ndarray, *, xnutr: float, zd: float) -> np.ndarray:
    """``dss`` (L, C): xnutr*sin^2(pi/2*(zmid-zd)/(ztop-zd)) above ``zd``,
    recomputed from zgrid as the v8.4.1 core does at start-up
    (damping_v841.py:23-107): binary32 phase, the host C library's sinf
    (the GNU sinf of the native build), then one product for the square."""
    z = np.asarray(zgrid, dtype=F32)
    nlev = z.shape[0] - 1
    dss = np.zeros((nlev, z.shape[1]), dtype=F32)
    if xnutr == 0.0:
        return dss
    half = F32(0.5)
    start = F32(zd)
    pi = F32(np.pi)
    mid = half * (z[:-1] + z[1:])
    top = z[nlev][None, :].repeat(nlev, axis=0)
    active = mid > start
    phase = (half * pi) * (mid[active] - start) / (top[active] - start)
    s = _host_sinf(phase)
    dss[active] = F32(xnutr) * (s * s)
    return dss


def _tangent_basis(p: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    ref = np.where(np.abs(p[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    e1 = ref - p * np.einsum("ij,ij->i", ref, p)[:, None]
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(p, e1)
    return e1, e2


def deriv_two(mesh: Mesh) -> np.ndarray:
    """Second derivative along each edge normal of the quadratic least-squares
    fit over a cell and its neighbours: shape (E, 2, maxEdges+1).

    Local coordinates are tangent-plane polar coordinates (great-circle
    distance, azimuth) around the cell; the second derivative along a
    direction is invariant to the choice of azimuth origin, so any tangent
    basis gives the same coefficients.
    """
    R = mesh.sphere_radius
    pc = np.stack([mesh.xCell, mesh.yCell, mesh.zCell], axis=1) / R
    pv = np.stack([mesh.xVertex, mesh.yVertex, mesh.zVertex], axis=1) / R
    counts = mesh.nEdgesOnCell
    max_edges = mesh.max_edges
    width = max_edges + 1
    result = np.zeros((mesh.n_edges, 2, width))
    e1, e2 = _tangent_basis(pc)

    def azimuth_distance(cells: np.ndarray, targets: np.ndarray):
        p0 = pc[cells][:, None, :]
        d = targets - p0 * np.einsum("ijk,ijk->ij", targets, p0)[:, :, None]
        ang = np.arctan2(np.einsum("ijk,ik->ij", d, e2[cells]), np.einsum("ijk,ik->ij", d, e1[cells]))
        dist = np.arctan2(np.linalg.norm(np.cross(p0, targets), axis=2), np.einsum("ijk,ijk->ij", p0, targets)) * R
        return ang, dist

    for n in np.unique(counts):
        cells = np.flatnonzero(counts == n)
        nb = mesh.cellsOnCell[cells, :n]
        ang, dist = azimuth_distance(cells, pc[nb])
        xp = np.cos(ang) * dist
        yp = np.sin(ang) * dist
        design = np.zeros((cells.size, n + 1, 6))
        design[:, :, 0] = 1.0
        design[:, 1:, 1] = xp
        design[:, 1:, 2] = yp
        design[:, 1:, 3] = xp * xp
        design[:, 1:, 4] = xp * yp
        design[:, 1:, 5] = yp * yp
        dt = np.transpose(design, (0, 2, 1))
        fit = np.linalg.solve(dt @ design, dt)  # (cells, 6, n+1)
        edges = mesh.edgesOnCell[cells, :n]
        ve = mesh.verticesOnEdge[edges]
        mid = pv[ve[:, :, 0]] + pv[ve[:, :, 1]]
        mid /= np.linalg.norm(mid, axis=2, keepdims=True)
        eang, _ = azimuth_distance(cells, mid)
        c, s = np.cos(eang), np.sin(eang)
        coef = (2.0 * (c * c)[:, :, None] * fit[:, None, 3, :]
                + 2.0 * (c * s)[:, :, None] * fit[:, None, 4, :]
                + 2.0 * (s * s)[:, :, None] * fit[:, None, 5, :])  # (cells, n, n+1)
        side = (mesh.cellsOnEdge[edges, 0] != cells[:, None]).astype(np.int64)
        for slot in range(n):
            result[edges[:, slot], side[:, slot], : n + 1] = coef[:, slot, :]
    return result


def advection_coefficients(mesh: Mesh, deriv: np.ndarray):
    """``adv_coefs``, ``adv_coefs_3rd`` (E, W), counts and cells, binary32.

    Stencil order is native insertion order: the two owners, then the first
    owner's unseen neighbours in ``cellsOnCell`` order, then the second's.
    Accumulation, scaling by ``-dc*dc/12``, the two ``+1/2`` owner terms and
    the final ``*dvEdge`` follow ``transport.py:2413-2443``.
    """
    coe = mesh.cellsOnEdge
    coc = mesh.cellsOnCell
    counts = mesh.nEdgesOnCell
    n_edges = coe.shape[0]
    width = 2 * mesh.max_edges - 2
    deriv32 = np.asarray(deriv, dtype=F32)
    dc = mesh.dcEdge.astype(F32)
    dv = mesh.dvEdge.astype(F32)
    cells = np.full((n_edges, width), -1, dtype=np.int64)
    nadv = np.zeros(n_edges, dtype=np.int64)
    adv = np.zeros((n_edges, width), dtype=F32)
    adv3 = np.zeros((n_edges, width), dtype=F32)
    for e in range(n_edges):
        c1, c2 = int(coe[e, 0]), int(coe[e, 1])
        ordered = [c1, c2]
        for c in (c1, c2):
            for s in range(counts[c]):
                nb = int(coc[c, s])
                if nb not in ordered:
                    ordered.append(nb)
        if len(ordered) > width:
            raise ValueError("advection stencil wider than 2*maxEdges-2")
        nadv[e] = len(ordered)
        cells[e, : len(ordered)] = ordered
        pos = {c: i for i, c in enumerate(ordered)}
        for side, owner in ((0, c1), (1, c2)):
            members = [owner] + [int(coc[owner, s]) for s in range(counts[owner])]
            for slot, cell in enumerate(members):
                value = deriv32[e, side, slot]
                p = pos[cell]
                adv[e, p] = adv[e, p] + value
                adv3[e, p] = adv3[e, p] + value if side == 0 else adv3[e, p] - value
        n = len(ordered)
        dc2 = dc[e] * dc[e]
        adv[e, :n] = (-dc2 * adv[e, :n]) / F32(12.0)
        adv3[e, :n] = (-dc2 * adv3[e, :n]) / F32(12.0)
        adv[e, pos[c1]] = adv[e, pos[c1]] + F32(0.5)
        adv[e, pos[c2]] = adv[e, pos[c2]] + F32(0.5)
        adv[e, :n] = adv[e, :n] * dv[e]
        adv3[e, :n] = adv3[e, :n] * dv[e]
    return adv, adv3, nadv, cells


def stored_inverse(values: np.ndarray) -> np.ndarray:
    """Round to binary32 first, reciprocal second (``dynamics_v841.py:28-46``)."""
    v = np.asarray(values, dtype=F32)
    if np.any(v == 0):
        raise ValueError("zero geometry value")
    return (F32(1.0) / v).astype(F32)


def _arc_length(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Great-circle length between points a and b on the sphere through a (MPAS ``mpas_arc_length``)."""
    r = np.sqrt(np.sum(a * a, axis=-1))
    c = np.sqrt(np.sum((b - a) ** 2, axis=-1))
    return r * 2.0 * np.arcsin(c / (2.0 * r))


def _sphere_angle(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """Signed spherical angle at a between the arcs a-b and a-c (MPAS ``mpas_sphere_angle``).

    Half-angle form of the spherical law of cosines; positive when
    ((b-a) x (c-a)) . a >= 0.
    """
    sa = _arc_length(b, c)
    sb = _arc_length(a, c)
    sc = _arc_length(a, b)
    d = np.cross(b - a, c - a)
    s = 0.5 * (sa + sb + sc)
    ratio = (np.sin(s - sb) * np.sin(s - sc)) / (np.sin(sb) * np.sin(sc))
    sin_half = np.sqrt(np.minimum(1.0, np.maximum(0.0, ratio)))
    mag = 2.0 * np.arcsin(np.clip(sin_half, -1.0, 1.0))
    return np.where(np.sum(d * a, axis=-1) >= 0.0, mag, -mag)


def deformation_weights(mesh: Mesh):
    """Smagorinsky deformation weights ``(coef_c2, coef_s2, coef_cs)``, each (C, maxEdges), binary32.

    Per cell (``mixing_v841.py:372-455``): the cell's vertices are laid out on
    the tangent plane at polar angles accumulated from the angle to north
    and at the great-circle distance from the centre; the polygon area comes
    from the shoelace sum and each side's outward normal angle from atan2 of
    the side.  For side j with length dl and normal angle t:
    c2 = dl cos^2 t / A, s2 = dl sin^2 t / A, cs = dl sin t cos t / A, all
    negated where the cell is not the edge's first cell.
    """
    radius = float(mesh.sphere_radius)
    cell_xyz = np.stack([mesh.xCell, mesh.yCell, mesh.zCell], axis=1) / radius
    vert_xyz = np.stack([mesh.xVertex, mesh.yVertex, mesh.zVertex], axis=1) / radius
    counts = mesh.nEdgesOnCell
    me = mesh.max_edges
    n_cells = counts.size
    c2 = np.zeros((n_cells, me), dtype=np.float64)
    s2 = np.zeros_like(c2)
    cs = np.zeros_like(c2)
    north = np.array([0.0, 0.0, 1.0])
    for n in np.unique(counts):
        cells = np.flatnonzero(counts == n)
        p = cell_xyz[cells]                                 # (m, 3)
        v = vert_xyz[mesh.verticesOnCell[cells, :n]]        # (m, n, 3)
        polar = p[:, 2] == 1.0
        theta_abs = np.where(polar, 0.5 * np.pi,
                             0.5 * np.pi - _sphere_angle(p, v[:, 0], np.broadcast_to(north, p.shape)))
        vn = np.roll(v, -1, axis=1)
        pp = np.broadcast_to(p[:, None, :], v.shape)
        thetav = _sphere_angle(pp, v, vn)                   # (m, n)
        dl_sphere = radius * _arc_length(pp, v)
        thetat = np.empty_like(thetav)
        thetat[:, 0] = theta_abs
        for j in range(1, n):
            thetat[:, j] = thetat[:, j - 1] + thetav[:, j - 1]
        xp = np.cos(thetat) * dl_sphere
        yp = np.sin(thetat) * dl_sphere
        xq, yq = np.roll(xp, -1, axis=1), np.roll(yp, -1, axis=1)
        area = np.zeros(cells.size)
        for j in range(n):
            area = (area + 0.25 * (xp[:, j] + xq[:, j]) * (yq[:, j] - yp[:, j])
                    - 0.25 * (yp[:, j] + yq[:, j]) * (xq[:, j] - xp[:, j]))
        dx, dy = xq - xp, yq - yp
        t = np.arctan2(dy, dx) - 0.5 * np.pi
        dl = np.sqrt(dx * dx + dy * dy)
        st, ct = np.sin(t), np.cos(t)
        sign = np.where(mesh.cellsOnEdge[mesh.edgesOnCell[cells, :n], 0] == cells[:, None], 1.0, -1.0)
        c2[cells, :n] = sign * dl * ct * ct / area[:, None]
        s2[cells, :n] = sign * dl * st * st / area[:, None]
        cs[cells, :n] = sign * dl * st * ct / area[:, None]
    return c2.astype(F32), s2.astype(F32), cs.astype(F32)


def mesh_mixing_scaling(mesh: Mesh, *, scale_with_mesh: bool = True):
    """``meshScalingDel2/Del4`` per edge (``mixing.py:271-318``): 1/rho_e^(1/4) and 1/rho_e^(3/4) of the
    binary32 mean cell mesh density, or ones.  The power is binary64 pow of the binary32 mean, rounded
    once (the reference's ``rkind_libm.powf_rkind``, the evaluation hexdyn_create makes).  On a
    quasi-uniform mesh (density 1) both are exactly 1."""
    n_edges = mesh.cellsOnEdge.shape[0]
    if not scale_with_mesh:
        return np.ones(n_edges, F32), np.ones(n_edges, F32)
    density = np.asarray(mesh.arrays.get("meshDensity", np.ones(mesh.nEdgesOnCell.size)), dtype=F32)
    mean = (F32(0.5) * (density[mesh.cellsOnEdge[:, 0]] + density[mesh.cellsOnEdge[:, 1]])).astype(F32)
    one = F32(1.0)
    del2 = (one / (mean.astype(np.float64) ** np.float64(F32(0.25))).astype(F32)).astype(F32)
    del4 = (one / (mean.astype(np.float64) ** np.float64(F32(0.75))).astype(F32)).astype(F32)
    return del2, del4
