"""Ensemble groups: members of one mesh advanced as one instance on the
disjoint union of their meshes equal each member run alone, bit for bit
(the CPU restatement on the union case; the HIP library through its C ABI,
with groups on meshes of different size and timestep)."""

import numpy as np
import pytest

from hexdyn import ensemble
from hexdyn.case import WORKLOADS
from hexdyn.config import DycoreConfig


def _bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_union_mesh_tables(get_case):
    case = get_case("x1.162-moist")
    m = case.mesh
    u = ensemble.union_mesh(m, 3)
    C, E, V = m.n_cells, m.n_edges, m.n_vertices
    assert (u.n_cells, u.n_edges, u.n_vertices) == (3 * C, 3 * E, 3 * V)
    for k in range(3):
        coe = u.cellsOnEdge[k * E:(k + 1) * E]
        assert coe.min() >= k * C and coe.max() < (k + 1) * C
        eoc = u.edgesOnCell[k * C:(k + 1) * C]
        valid = eoc[eoc >= 0]
        assert valid.min() >= k * E and valid.max() < (k + 1) * E
        np.testing.assert_array_equal(u.kiteForCell[k * C:(k + 1) * C], m.kiteForCell)
    # Euler characteristic of each copy: V - E + C = 2
    assert u.n_vertices - u.n_edges + u.n_cells == 3 * 2


def test_member_perturbation_is_seeded_and_smooth(get_case):
    case = get_case("x1.162-moist")
    a = ensemble.member_state(case, 0)
    b = ensemble.member_state(case, 0)
    c = ensemble.member_state(case, 1)
    assert _bits(a["rtheta_m"], b["rtheta_m"]) and not _bits(a["rtheta_m"], c["rtheta_m"])
    base = case.state.rtheta_m.astype(np.float64) / case.state.rho_zz
    d = a["rtheta_m"].astype(np.float64) / a["rho_zz"] - base
    assert abs(d.std() - 0.1) < 1e-3
    # smoothed: neighbouring cells are correlated
    coc = case.mesh.cellsOnCell[:, 0]
    corr = np.corrcoef(d[10], d[10][coc])[0, 1]
    assert corr > 0.3


def test_union_case_steps_each_member_bitwise_on_the_oracle(get_case):
    from oracle.oracle import Oracle

    case = get_case("x1.162-moist")
    states = [ensemble.member_state(case, k) for k in range(3)]
    o = Oracle(ensemble.union_case(case, 3), 32)
    o.set_state(ensemble.join_states(states), rebuild=True)
    assert o.step(2) == 0
    got = ensemble.split_state(o.get_state(), 3)
    for k, st in enumerate(states):
        s = Oracle(case, 32)
        s.set_state(st, rebuild=True)
        assert s.step(2) == 0
        want = s.get_state()
        for f in want:
            assert _bits(got[k][f], want[f]), (k, f)


@pytest.mark.gpu
def test_groups_of_uneven_meshes_equal_single_member_runs(get_case):
    """3 members on x1.642 (dt 3600 s) and 2 on x1.162 (dt 7200 s) advanced
    4 h together: every member equals its own single-instance run."""
    import torch

    from hexdyn.driver import HexDycoreDriver

    print("[gpu] ensemble groups x1.642 x3 + x1.162 x2", flush=True)
    assert torch.cuda.is_available()
    groups, singles = [], []
    member = 0
    for name, n in (("x1.642-moist", 3), ("x1.162-moist", 2)):
        case = get_case(name, config=DycoreConfig(config_dt=WORKLOADS[name]["dt"], config_horiz_mixing="2d_smagorinsky"))
        states = [ensemble.member_state(case, member + k) for k in range(n)]
        member += n
        groups.append(ensemble.EnsembleGroup(case, states))
        singles.append((case, states))
    ens = ensemble.Ensemble(groups)
    assert ens.interval == 7200.0
    ens.advance(4 * 3600.0)
    for g, (case, states) in zip(groups, singles):
        got = g.member_states()
        nsteps = int(4 * 3600 / case.config.config_dt)
        for k, st in enumerate(states):
            d = HexDycoreDriver.from_host(case, state=st)
            d.step_device(nsteps)
            want = d.get_state()
            d.close()
            for f in want:
                assert _bits(got[k][f], want[f]), (case.name, k, f)
    ens.close()
