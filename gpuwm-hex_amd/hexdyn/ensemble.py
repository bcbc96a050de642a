"""Ensemble members advanced together: one dycore instance per mesh group.

The reference runs one process per ensemble member (``docs/member-batch.md``:
a member batch is a set of independent forecast processes).  Here the
members that share a mesh, vertical grid and timestep form one group, and a
group is ONE dycore instance on the disjoint union of its members' meshes:
every kernel launch of a large step (and the hipGraph that replays the
step) covers all members of the group at once, so a group of n members costs
one launch sequence instead of n.  Each entity's arithmetic reads only its
own stencil, so every member's result is bit for bit the result of running
it alone.  Groups on meshes of different size (and timestep) advance in
turn on one stream, each by the number of its own steps that spans the
common interval.

What is shared across a group: the validation flag (a non-finite member
refuses the whole group's step) and the handle's FP64 integrals (sums over
all members).
"""

from __future__ import annotations

import time
from dataclasses import replace
from typing import Dict, List, Optional, Sequence

import numpy as np

from .case import Case, initial_state_arrays
from .mesh import Mesh

# index tables of an MPAS mesh and the entity they point to
_INDEX_TABLES = {
    "cellsOnEdge": "C", "verticesOnEdge": "V", "edgesOnCell": "E", "cellsOnCell": "C", "verticesOnCell": "V",
    "edgesOnVertex": "E", "cellsOnVertex": "C", "edgesOnEdge": "E",
}


def _offset(a: np.ndarray, off: int) -> np.ndarray:
    a = np.asarray(a)
    return np.where(a >= 0, a + off, a).astype(a.dtype)


def union_mesh(mesh: Mesh, n: int) -> Mesh:
    """n disjoint copies of a mesh (entity numbers of copy k offset by k
    times the entity counts; padding -1 kept)."""
    C, E, V = mesh.n_cells, mesh.n_edges, mesh.n_vertices
    counts = {"C": C, "E": E, "V": V}
    out = {}
    for name, a in mesh.arrays.items():
        a = np.asarray(a)
        if name in _INDEX_TABLES:
            step = counts[_INDEX_TABLES[name]]
            out[name] = np.concatenate([_offset(a, k * step) for k in range(n)], axis=0)
        elif a.ndim >= 1 and a.shape[0] in (C, E, V) and a.shape[0] > 1:
            out[name] = np.concatenate([a] * n, axis=0)
        else:
            out[name] = a
    return Mesh(arrays=out, sphere_radius=mesh.sphere_radius)


def union_case(case: Case, n: int) -> Case:
    """The group instance's case: n copies of `case` side by side (the
    initial state of every copy is the case's; members set their own)."""
    g, r, st = case.grid, case.reference, case.state
    cat1 = lambda a: np.concatenate([a] * n, axis=1)  # noqa: E731  [level][entity]
    cat0 = lambda a: np.concatenate([a] * n, axis=0)  # noqa: E731  [entity]...
    grid = replace(g, zgrid=cat1(g.zgrid), zz=cat1(g.zz), zxu=cat1(g.zxu), zb_cell=cat1(g.zb_cell),
                   zb3_cell=cat1(g.zb3_cell),
                   zb_edge=None if g.zb_edge is None else np.concatenate([g.zb_edge] * n, axis=2),
                   zb3_edge=None if g.zb3_edge is None else np.concatenate([g.zb3_edge] * n, axis=2))
    ref = replace(r, rho_base=cat1(r.rho_base), rtheta_base=cat1(r.rtheta_base), exner_base=cat1(r.exner_base),
                  theta_base=cat1(r.theta_base))
    state = replace(st, rho_zz=cat1(st.rho_zz), rtheta_m=cat1(st.rtheta_m), ru=cat1(st.ru), rw=cat1(st.rw),
                    scalars=np.concatenate([st.scalars] * n, axis=-1), extras={})
    C = case.mesh.n_cells
    adv_cells = np.concatenate([_offset(case.adv_cells, k * C) for k in range(n)], axis=0)
    return replace(case, name=f"{case.name}x{n}", mesh=union_mesh(case.mesh, n), grid=grid, reference=ref,
                   state=state, adv_coefs=cat0(case.adv_coefs), adv_coefs_3rd=cat0(case.adv_coefs_3rd),
                   n_adv=cat0(case.n_adv), adv_cells=adv_cells, inv_area_cell=cat0(case.inv_area_cell),
                   inv_area_triangle=cat0(case.inv_area_triangle), inv_dc_edge=cat0(case.inv_dc_edge),
                   inv_dv_edge=cat0(case.inv_dv_edge), cos_angle_edge=cat0(case.cos_angle_edge),
                   sin_angle_edge=cat0(case.sin_angle_edge), dss=cat1(case.dss))


def member_state(case: Case, member: int, sigma: float = 0.1, rings: int = 3) -> Dict[str, np.ndarray]:
    """The case's initial state with a seeded theta perturbation (seed
    100 + member): Gaussian noise per cell and level, smoothed over `rings`
    neighbour rings (each pass averages a cell with its cellsOnCell), then
    rescaled to standard deviation `sigma` (K), added to theta_m at constant
    density."""
    st = {k: np.array(v, copy=True) for k, v in initial_state_arrays(case).items()}
    m = case.mesh
    rng = np.random.default_rng(100 + member)
    L, C = st["rho_zz"].shape
    x = rng.standard_normal((L, C))
    coc = np.asarray(m.cellsOnCell)
    n = np.asarray(m.nEdgesOnCell)
    mask = np.arange(coc.shape[1])[None, :] < n[:, None]
    idx = np.where(mask, coc, 0)
    for _ in range(rings):
        x = (x + (x[:, idx] * mask[None]).sum(axis=2)) / (1.0 + n[None, :])
    x *= sigma / x.std()
    theta = st["rtheta_m"].astype(np.float64) / st["rho_zz"].astype(np.float64)
    st["rtheta_m"] = (st["rho_zz"] * (theta + x)).astype(np.float32)
    return st


def join_states(states: Sequence[Dict[str, np.ndarray]]) -> Dict[str, np.ndarray]:
    """Member states (logical layout) side by side along the entity axis."""
    return {k: np.concatenate([s[k] for s in states], axis=-1) for k in states[0]}


def split_state(state: Dict[str, np.ndarray], n: int) -> List[Dict[str, np.ndarray]]:
    return [{k: np.split(np.asarray(v), n, axis=-1)[i] for k, v in state.items()} for i in range(n)]


class EnsembleGroup:
    """Members of one mesh and timestep in one dycore instance."""

    def __init__(self, case: Case, states: Sequence[Dict[str, np.ndarray]]) -> None:
        from .driver import HexDycoreDriver

        self.case = case
        self.n = len(states)
        self.union = union_case(case, self.n)
        self.driver = HexDycoreDriver.from_host(self.union, state=join_states(states))

    @property
    def dt(self) -> float:
        return float(self.case.config.config_dt)

    def step(self, n: int = 1) -> None:
        self.driver.step_device(n)

    def member_states(self) -> List[Dict[str, np.ndarray]]:
        return split_state(self.driver.get_state(), self.n)

    def close(self) -> None:
        self.driver.close()


class Ensemble:
    """Groups on meshes of different size advanced together over common
    intervals (each interval a whole number of every group's steps)."""

    def __init__(self, groups: Sequence[EnsembleGroup]) -> None:
        self.groups = list(groups)
        self.interval = float(np.lcm.reduce([int(round(g.dt)) for g in self.groups]))

    def advance(self, seconds: float) -> float:
        """Advance every member by `seconds` (a multiple of the common
        interval); returns the wall time."""
        import torch

        k = seconds / self.interval
        if abs(k - round(k)) > 1e-9:
            raise ValueError(f"{seconds} s is not a multiple of the common interval {self.interval} s")
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        for _ in range(int(round(k))):
            for g in self.groups:
                g.step(int(round(self.interval / g.dt)))
        torch.cuda.synchronize()
        return time.perf_counter() - w0

    def close(self) -> None:
        for g in self.groups:
            g.close()


def config3(groups: Optional[Sequence[tuple]] = None, mixing: str = "2d_smagorinsky") -> Ensemble:
    """The batched ensemble workload: 8 members on x1.40962 (dt 720 s) and 24
    on x1.10242 (dt 1440 s), each with its seeded theta perturbation."""
    from .case import WORKLOADS, build_case
    from .config import DycoreConfig

    groups = groups or (("x1.40962-moist", 8), ("x1.10242-moist", 24))
    out, member = [], 0
    for name, n in groups:
        case = build_case(name, config=DycoreConfig(config_dt=WORKLOADS[name]["dt"], config_horiz_mixing=mixing))
        states = [member_state(case, member + k) for k in range(n)]
        member += n
        out.append(EnsembleGroup(case, states))
    return Ensemble(out)
