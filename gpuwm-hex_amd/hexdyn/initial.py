"""Vertical grid, terrain, reference state and synthetic initial states.

The reference takes these from MPAS ``init_atmosphere`` output (``driver.py``
loaders, ``vertical.py:510`` via an absent engine).  Here they are generated:

* a smoothed terrain-following height coordinate to ``ztop`` (30 km), with
  ``zz``, ``rdzw``, ``rdzu``, ``fzm``, ``fzp``, ``cf1..3``, ``zxu`` and the
  ``zb``/``zb3`` slope metrics in the MPAS definitions (``terrain.py:182``,
  ``VerticalGrid`` at ``vertical.py:83-112``);
* seeded random spherical-harmonic terrain up to degree 16, clipped to 0-4 km;
* the Jablonowski-Williamson (2006) baroclinic jet, put into discrete
  hydrostatic balance column by column in the model's own vertical operator;
* a seeded Gaussian theta perturbation, relative humidity and log-normal
  condensates.

Everything is computed in binary64 and rounded once to binary32.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, Optional

import os

import numpy as np
from scipy.special import sph_harm_y

from .mesh import Mesh, edge_signs_on_cells
from . import setup as hsetup

RGAS = 287.0
CP = 1004.5
CV = CP - RGAS
P0 = 100000.0
GRAVITY = 9.80616
RV = 461.6
RVORD = RV / RGAS

F32 = np.float32


@dataclass
class VerticalGrid:
    """Height coordinate and metrics; per-level arrays are binary32."""

    zw: np.ndarray          # (L+1,) computational interface heights
    zgrid: np.ndarray       # (L+1, C) physical interface heights
    zz: np.ndarray          # (L, C)
    zxu: np.ndarray         # (L, E)
    rdzw: np.ndarray        # (L,)
    rdzu: np.ndarray        # (L,)  index 0 unused
    fzm: np.ndarray         # (L,)
    fzp: np.ndarray         # (L,)
    cf1: float
    cf2: float
    cf3: float
    zb_cell: np.ndarray     # (L+1, C, maxEdges)
    zb3_cell: np.ndarray    # (L+1, C, maxEdges), scaled by coef_3rd_order
    dzw: np.ndarray         # (L,) binary64 layer depths
    zb_edge: Optional[np.ndarray] = None   # (L+1, 2, E) MPAS zb as an init file holds it
    zb3_edge: Optional[np.ndarray] = None  # (L+1, 2, E) MPAS zb3, unscaled
    dzu: Optional[np.ndarray] = None       # (L,) binary32, index 0 unused


def terrain_slot_metrics(mesh: Mesh, zb_edge: np.ndarray, zb3_edge: np.ndarray, coef_3rd_order: float):
    """Per-cell-slot zb, zb3 from MPAS's per-edge (level, side, edge) arrays.

    Slot s of cell c takes side 0 of its edge when c is the edge's first cell,
    side 1 otherwise; zb3 is scaled by coef_3rd_order once, in binary32, as the
    model does when it reads an init file (MPAS core.F, atm_mpas_init_block).
    """
    nlev1 = zb_edge.shape[0]
    zb_cell = np.zeros((nlev1, mesh.n_cells, mesh.max_edges), dtype=F32)
    zb3_cell = np.zeros_like(zb_cell)
    coef = F32(coef_3rd_order)
    for s in range(mesh.max_edges):
        cells = np.flatnonzero(mesh.nEdgesOnCell > s)
        e = mesh.edgesOnCell[cells, s]
        side = np.where(mesh.cellsOnEdge[e, 0] == cells, 0, 1)
        zb_cell[:, cells, s] = zb_edge[:, side, e]
        zb3_cell[:, cells, s] = coef * zb3_edge[:, side, e].astype(F32)
    return zb_cell, zb3_cell


@dataclass
class ReferenceState:
    rho_base: np.ndarray
    rtheta_base: np.ndarray
    exner_base: np.ndarray
    theta_base: np.ndarray


@dataclass
class InitialState:
    rho_zz: np.ndarray
    rtheta_m: np.ndarray
    ru: np.ndarray
    rw: np.ndarray
    scalars: np.ndarray      # (T, L, C)
    extras: Dict[str, np.ndarray] = field(default_factory=dict)


def random_terrain(mesh: Mesh, *, seed: int = 2, lmax: int = 16, hmax: float = 4000.0) -> np.ndarray:
    """Seeded real spherical-harmonic field up to ``lmax``, clipped to [0, hmax]."""
    rng = np.random.default_rng(seed)
    theta = np.pi / 2.0 - mesh.latCell
    phi = mesh.lonCell
    field = np.zeros(mesh.n_cells)
    for l in range(1, lmax + 1):
        amp = 1.0 / (l + 1.0)
        for m in range(0, l + 1):
            y = sph_harm_y(l, m, theta, phi)
            a, b = rng.standard_normal(2) * amp
            field += a * y.real + (b * y.imag if m > 0 else 0.0)
    field = (field - field.mean()) / field.std()
    return np.clip(1200.0 + 1400.0 * field, 0.0, hmax)


def _smooth(mesh: Mesh, h: np.ndarray, passes: int) -> np.ndarray:
    counts = mesh.nEdgesOnCell
    out = h.copy()
    for _ in range(passes):
        acc = np.zeros_like(out)
        for s in range(mesh.max_edges):
            cells = np.flatnonzero(counts > s)
            acc[cells] += out[mesh.cellsOnCell[cells, s]]
        out = 0.5 * out + 0.5 * acc / counts
    return out


def build_vertical_grid(mesh: Mesh, terrain: np.ndarray, *, nlev: int = 55, ztop: float = 30000.0,
                        stretch: float = 0.3, coef_3rd_order: float = 0.25,
                        deriv: Optional[np.ndarray] = None) -> VerticalGrid:
    s = np.linspace(0.0, 1.0, nlev + 1)
    zw = ztop * (stretch * s + (1.0 - stretch) * s * s)
    zgrid = np.zeros((nlev + 1, mesh.n_cells))
    hs = _smooth(mesh, terrain, passes=2)
    for k in range(nlev + 1):
        # terrain influence decays quadratically to a flat top
        decay = (1.0 - zw[k] / ztop) ** 2
        zgrid[k] = zw[k] + decay * hs
    zgrid[nlev] = ztop
    dzw = np.diff(zw)
    zz = (dzw[:, None]) / np.diff(zgrid, axis=0)
    if np.any(zz <= 0):
        raise ValueError("terrain-following coordinate folded")
    dzu = np.zeros(nlev)
    dzu[1:] = 0.5 * (dzw[1:] + dzw[:-1])
    rdzu = np.zeros(nlev)
    rdzu[1:] = 1.0 / dzu[1:]
    fzm = np.zeros(nlev)
    fzp = np.zeros(nlev)
    fzm[1:] = 0.5 * dzw[:-1] / dzu[1:]
    fzp[1:] = 0.5 * dzw[1:] / dzu[1:]
    d1 = 0.5 * dzw[0]
    d2 = dzw[0] + 0.5 * dzw[1]
    d3 = dzw[0] + dzw[1] + 0.5 * dzw[2]
    den = d2 * d3 * (d3 - d2) + d1 * d3 * (d1 - d3) + d1 * d2 * (d2 - d1)
    cf1 = d2 * d3 * (d3 - d2) / den
    cf2 = d1 * d3 * (d1 - d3) / den
    cf3 = d1 * d2 * (d2 - d1) / den
    c1 = mesh.cellsOnEdge[:, 0]
    c2 = mesh.cellsOnEdge[:, 1]
    zxu = 0.5 * (zgrid[:-1, c2] - zgrid[:-1, c1] + zgrid[1:, c2] - zgrid[1:, c1]) / mesh.dcEdge[None, :]
    # zb, zb3 (L+1, 2, E): MPAS init_atmosphere slope metrics.
    if deriv is None:
        deriv = hsetup.deriv_two(mesh)
    zb = np.zeros((nlev + 1, 2, mesh.n_edges))
    zb3 = np.zeros_like(zb)
    d2c1 = deriv[:, 0, 0][None, :] * zgrid[:, c1]
    d2c2 = deriv[:, 1, 0][None, :] * zgrid[:, c2]
    for s in range(mesh.max_edges):
        nb1 = mesh.cellsOnCell[c1, s]
        nb2 = mesh.cellsOnCell[c2, s]
        on1 = mesh.nEdgesOnCell[c1] > s
        on2 = mesh.nEdgesOnCell[c2] > s
        d2c1 += np.where(on1, deriv[:, 0, s + 1], 0.0)[None, :] * zgrid[:, np.where(on1, nb1, 0)]
        d2c2 += np.where(on2, deriv[:, 1, s + 1], 0.0)[None, :] * zgrid[:, np.where(on2, nb2, 0)]
    dc2 = (mesh.dcEdge ** 2)[None, :]
    z_edge = 0.5 * (zgrid[:, c1] + zgrid[:, c2]) - dc2 * (d2c1 + d2c2) / 12.0
    z_edge3 = -dc2 * (d2c1 - d2c2) / 12.0
    zb[:, 0, :] = (z_edge - zgrid[:, c1]) * mesh.dvEdge / mesh.areaCell[c1]
    zb[:, 1, :] = (z_edge - zgrid[:, c2]) * mesh.dvEdge / mesh.areaCell[c2]
    zb3[:, 0, :] = z_edge3 * mesh.dvEdge / mesh.areaCell[c1]
    zb3[:, 1, :] = z_edge3 * mesh.dvEdge / mesh.areaCell[c2]
    zb_edge, zb3_edge = zb.astype(F32), zb3.astype(F32)
    zb_cell, zb3_cell = terrain_slot_metrics(mesh, zb_edge, zb3_edge, coef_3rd_order)
    return VerticalGrid(
        zw=zw, zgrid=zgrid.astype(F32), zz=zz.astype(F32), zxu=zxu.astype(F32),
        rdzw=(1.0 / dzw).astype(F32), rdzu=rdzu.astype(F32), fzm=fzm.astype(F32), fzp=fzp.astype(F32),
        cf1=float(F32(cf1)), cf2=float(F32(cf2)), cf3=float(F32(cf3)),
        zb_cell=zb_cell, zb3_cell=zb3_cell, dzw=dzw, zb_edge=zb_edge, zb3_edge=zb3_edge, dzu=dzu.astype(F32),
    )


# Jablonowski-Williamson (2006) steady jet, eqs. 2-7.
_JW_U0 = 35.0
_JW_ETA0 = 0.252
_JW_ETAT = 0.2
_JW_T0 = 288.0
_JW_GAMMA = 0.005
_JW_DT = 4.8e5
_OMEGA = 7.29212e-5


def _jw_fields(eta: np.ndarray, lat: np.ndarray, radius: float):
    etav = (eta - _JW_ETA0) * np.pi / 2.0
    cv = np.cos(etav)
    sv = np.sin(etav)
    ex = RGAS * _JW_GAMMA / GRAVITY
    tbar = _JW_T0 * eta ** ex
    phibar = _JW_T0 * GRAVITY / _JW_GAMMA * (1.0 - eta ** ex)
    strat = eta < _JW_ETAT
    et = _JW_ETAT
    tbar = np.where(strat, tbar + _JW_DT * (et - eta) ** 5, tbar)
    e = np.minimum(eta, et)
    logterm = np.log(np.where(strat, eta / et, 1.0))
    corr = RGAS * _JW_DT * ((logterm + 137.0 / 60.0) * et ** 5 - 5.0 * et ** 4 * e + 5.0 * et ** 3 * e ** 2
                            - (10.0 / 3.0) * et ** 2 * e ** 3 + (5.0 / 4.0) * et * e ** 4 - 0.2 * e ** 5)
    phibar = np.where(strat, phibar - corr, phibar)
    sphi, cphi = np.sin(lat), np.cos(lat)
    a1 = -2.0 * sphi ** 6 * (cphi ** 2 + 1.0 / 3.0) + 10.0 / 63.0
    a2 = 8.0 / 5.0 * cphi ** 3 * (sphi ** 2 + 2.0 / 3.0) - np.pi / 4.0
    c32 = np.abs(cv) ** 1.5
    temp = tbar + 0.75 * eta * np.pi * _JW_U0 / RGAS * sv * np.sqrt(np.abs(cv)) * (
        a1 * 2.0 * _JW_U0 * c32 + a2 * radius * _OMEGA)
    phi = phibar + _JW_U0 * c32 * (a1 * _JW_U0 * c32 + a2 * radius * _OMEGA)
    u = _JW_U0 * c32 * np.sin(2.0 * lat) ** 2
    return temp, phi, u


def _newton_update(eta: np.ndarray, z: np.ndarray, lat: np.ndarray, radius: float):
    """One Newton update of phi(eta) = g*z, elementwise; returns the new eta
    and max |step/eta|."""
    temp, phi, _ = _jw_fields(eta, lat, radius)
    f = phi - GRAVITY * z
    dfdeta = -RGAS * temp / eta
    step = f / dfdeta
    new = np.clip(eta - step, 1e-5, 1.3)
    return new, float(np.max(np.abs(step / new))) if new.size else 0.0


def _eta_from_height(z: np.ndarray, lat: np.ndarray, radius: float) -> np.ndarray:
    """Newton iteration for the JW vertical coordinate eta at heights z, all
    points iterated until the largest relative step is below 1e-13.  Large
    arrays are split into column chunks updated on threads (numpy releases
    the GIL in its elementwise loops); every element sees the same operations
    and the same iteration count as one whole-array update."""
    eta = np.clip(np.exp(-z / 7500.0), 1e-4, 1.2)
    z = np.broadcast_to(z, eta.shape)
    lat = np.broadcast_to(lat, eta.shape)
    cols = eta.shape[-1] if eta.ndim else 1
    nchunk = 1
    if eta.size >= (1 << 20) and eta.ndim == 2:
        nchunk = max(1, min(16, os.cpu_count() or 1))
    if nchunk == 1:
        for _ in range(60):
            eta, rel = _newton_update(eta, z, lat, radius)
            if rel < 1e-13:
                break
        return eta
    from concurrent.futures import ThreadPoolExecutor

    bounds = np.linspace(0, cols, nchunk + 1, dtype=np.int64)
    parts = [(slice(None), slice(int(bounds[i]), int(bounds[i + 1]))) for i in range(nchunk)]
    eta = np.ascontiguousarray(eta)
    with ThreadPoolExecutor(max_workers=nchunk) as pool:
        for _ in range(60):
            def run(sl):
                new, rel = _newton_update(eta[sl], z[sl], lat[sl], radius)
                return sl, new, rel
            results = list(pool.map(run, parts))
            out = np.empty_like(eta)
            rel = 0.0
            for sl, new, r in results:
                out[sl] = new
                rel = max(rel, r)
            eta = out
            if rel < 1e-13:
                break
    return eta


def _hydrostatic_column(p_bottom, theta_m, qtot, zz, grid: VerticalGrid) -> np.ndarray:
    """Pressure at layer centres satisfying the model's discrete balance
    rdzu(k)*(p(k)-p(k-1)) = -g*(fzm(k)*rz(k)*(1+qt(k)) + fzp(k)*rz(k-1)*(1+qt(k-1))),
    with rz the zz-scaled dry density from the equation of state."""
    nlev = theta_m.shape[0]
    dzw = grid.dzw
    dzu = 0.5 * (dzw[1:] + dzw[:-1])
    fzm = 0.5 * dzw[:-1] / dzu
    fzp = 0.5 * dzw[1:] / dzu

    def rz(p, k):
        return (P0 / (RGAS * theta_m[k] * zz[k])) * (p / P0) ** (CV / CP)

    p = np.zeros_like(theta_m)
    p[0] = p_bottom
    for k in range(1, nlev):
        below = fzp[k - 1] * rz(p[k - 1], k - 1) * (1.0 + qtot[k - 1])
        guess = p[k - 1] * np.exp(-dzu[k - 1] / 8000.0)
        x = guess
        for _ in range(30):
            r = rz(x, k)
            f = (x - p[k - 1]) / dzu[k - 1] + GRAVITY * (fzm[k - 1] * r * (1.0 + qtot[k]) + below)
            df = 1.0 / dzu[k - 1] + GRAVITY * fzm[k - 1] * (1.0 + qtot[k]) * r * (CV / CP) / x
            dx = f / df
            x = x - dx
            if np.max(np.abs(dx / x)) < 1e-15:
                break
        p[k] = x
    return p


def _qsat(temp: np.ndarray, pres: np.ndarray) -> np.ndarray:
    es = 610.78 * np.exp(17.27 * (temp - 273.16) / (temp - 35.86))
    es = np.minimum(es, 0.5 * pres)
    return 0.622 * es / (pres - es)


def build_reference_and_state(mesh: Mesh, grid: VerticalGrid, *, moist: bool, n_scalars: int,
                              seed_theta: int = 3, seed_rh: int = 4, seed_cond: int = 5,
                              theta_amplitude: float = 0.5, theta_radius: float = 600.0e3):
    """JW jet + seeded theta bubble, discretely hydrostatic; reference state from
    the jet-free JW profile.  Returns (ReferenceState, InitialState)."""
    R = mesh.sphere_radius
    nlev = grid.zz.shape[0]
    zgrid = grid.zgrid.astype(np.float64)
    zmid = 0.5 * (zgrid[:-1] + zgrid[1:])
    zz = grid.zz.astype(np.float64)
    lat = mesh.latCell[None, :].repeat(nlev, axis=0)

    # Base state: the latitude-independent JW mean profile.
    eta_b = _eta_from_height(zmid, np.zeros_like(zmid) + np.arcsin(np.sqrt(0.5)), R)
    temp_b, _, _ = _jw_fields(eta_b, np.zeros_like(eta_b) + np.arcsin(np.sqrt(0.5)), R)
    p_b_analytic = eta_b * P0
    theta_b = temp_b * (P0 / p_b_analytic) ** (RGAS / CP)
    p_b = _hydrostatic_column(p_b_analytic[0], theta_b, np.zeros_like(theta_b), zz, grid)
    rho_b = (P0 / (RGAS * theta_b * zz)) * (p_b / P0) ** (CV / CP)
    exner_b = (p_b / P0) ** (RGAS / CP)

    # Full state.
    eta = _eta_from_height(zmid, lat, R)
    temp, _, _ = _jw_fields(eta, lat, R)
    p_a = eta * P0
    theta = temp * (P0 / p_a) ** (RGAS / CP)
    rng = np.random.default_rng(seed_theta)
    lon0 = rng.uniform(0.0, 2.0 * np.pi)
    lat0 = np.deg2rad(rng.uniform(35.0, 45.0))
    cc = np.sin(lat0) * np.sin(mesh.latCell) + np.cos(lat0) * np.cos(mesh.latCell) * np.cos(mesh.lonCell - lon0)
    dist = R * np.arccos(np.clip(cc, -1.0, 1.0))
    bubble = theta_amplitude * np.exp(-(dist / theta_radius) ** 2)
    theta = theta + bubble[None, :] * np.exp(-((zmid - 5000.0) / 6000.0) ** 2)

    scalars = np.zeros((n_scalars, nlev, mesh.n_cells))
    qtot = np.zeros_like(theta)
    if moist and n_scalars > 0:
        rh = np.random.default_rng(seed_rh).uniform(0.3, 0.9, size=theta.shape)
        qv = rh * _qsat(temp, p_a)
        qv = np.where(zmid > 15000.0, np.minimum(qv, 5e-6), qv)
        scalars[0] = qv
        rng_c = np.random.default_rng(seed_cond)
        for i in range(1, n_scalars):
            draw = 1.0e-5 * np.exp(rng_c.standard_normal(theta.shape))
            scalars[i] = np.where(rh > 0.8, draw, 0.0)
        qtot = scalars.sum(axis=0)
        theta_m = theta * (1.0 + RVORD * scalars[0])
    else:
        theta_m = theta
    p = _hydrostatic_column(p_a[0], theta_m, qtot, zz, grid)
    rho_zz = (P0 / (RGAS * theta_m * zz)) * (p / P0) ** (CV / CP)
    rtheta = rho_zz * theta_m

    # Normal momentum at edges from the JW zonal jet.
    c1, c2 = mesh.cellsOnEdge[:, 0], mesh.cellsOnEdge[:, 1]
    lat_e = mesh.latEdge[None, :].repeat(nlev, axis=0)
    # eta at the edge from its two cells (the coordinate is smooth on the mesh scale)
    eta_e = 0.5 * (eta[:, c1] + eta[:, c2])
    _, _, u_zonal = _jw_fields(eta_e, lat_e, R)
    u = u_zonal * np.cos(mesh.angleEdge)[None, :]
    ru = 0.5 * (rho_zz[:, c1] + rho_zz[:, c2]) * u

    ref = ReferenceState(rho_base=rho_b.astype(F32), rtheta_base=(rho_b * theta_b).astype(F32),
                         exner_base=exner_b.astype(F32), theta_base=theta_b.astype(F32))
    state = InitialState(rho_zz=rho_zz.astype(F32), rtheta_m=rtheta.astype(F32), ru=ru.astype(F32),
                         rw=np.zeros((nlev + 1, mesh.n_cells), dtype=F32), scalars=scalars.astype(F32))
    state.rw[:] = _rw_for_zero_w(mesh, grid, state.ru)
    return ref, state


def _rw_for_zero_w(mesh: Mesh, grid: VerticalGrid, ru: np.ndarray) -> np.ndarray:
    """rw on interior interfaces such that the recovered w is zero."""
    nlev = ru.shape[0]
    signs = edge_signs_on_cells(mesh)
    fzm = grid.fzm.astype(np.float64)
    fzp = grid.fzp.astype(np.float64)
    zz = grid.zz.astype(np.float64)
    ru64 = ru.astype(np.float64)
    acc = np.zeros((nlev + 1, mesh.n_cells))
    for s in range(mesh.max_edges):
        cells = np.flatnonzero(mesh.nEdgesOnCell > s)
        e = mesh.edgesOnCell[cells, s]
        flux = fzm[1:nlev, None] * ru64[1:nlev, e] + fzp[1:nlev, None] * ru64[: nlev - 1, e]
        acc[1:nlev, cells] += signs[cells, s] * (grid.zb_cell[1:nlev, cells, s]
                                                + np.sign(flux) * grid.zb3_cell[1:nlev, cells, s]) * flux
    rw = np.zeros((nlev + 1, mesh.n_cells))
    zint = fzm[1:nlev, None] * zz[1:nlev] + fzp[1:nlev, None] * zz[: nlev - 1]
    rw[1:nlev] = -zint * acc[1:nlev]
    return rw.astype(F32)
