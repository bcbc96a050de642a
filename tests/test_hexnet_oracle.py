"""The binary64 statement of the corrector network pinned by mathematics:
central finite differences of its own loss for every parameter block, the
dot-product test of the neighbour-mean adjoint, the closed form of the first
Adam step, bf16 rounding of known values and the exact least-squares wind of
a uniform flow.  Each check is shown to fail on a planted mistake."""

import numpy as np
import pytest

from oracle import hexnet_oracle as hn


@pytest.fixture(scope="module")
def graph(get_case):
    case = get_case("x1.162-moist")
    m = case.mesh
    g = hn.Graph(m.cellsOnCell, m.nEdgesOnCell, m.edgesOnCell, m.cellsOnEdge, case.cos_angle_edge,
                 case.sin_angle_edge)
    return case, g


def _problem(g, layers=2, L=6, seed=0):
    rng = np.random.default_rng(seed)
    from hexdyn.learned import init_params

    flat = init_params(layers, seed=seed, out_gain=0.5)
    # non-zero biases so every ReLU boundary is exercised
    p = hn.unpack(flat, layers, round_weights=False)
    for name in p:
        if name.startswith("b_"):
            p[name] = rng.standard_normal(p[name].shape) * 0.1
    X = rng.standard_normal((g.C, hn.N_IN))
    Tc = rng.standard_normal((g.C, 3, hn.ROW))
    Te = rng.standard_normal((g.E, hn.ROW))
    return p, X, Tc, Te, L


def _loss(p, X, Tc, Te, g, layers, L):
    cache = hn.forward(p, X, g, layers)
    parts, dY = hn.loss(cache["Y"], Tc, Te, g, L)
    return parts["total"], cache, dY


@pytest.mark.parametrize("block", ["W_in", "b_in", "W_0", "b_0", "W_1", "b_1", "W_out", "b_out"])
def test_gradients_equal_central_differences(graph, block):
    case, g = graph
    layers = 2
    p, X, Tc, Te, L = _problem(g, layers)
    total, cache, dY = _loss(p, X, Tc, Te, g, layers, L)
    grads = hn.backward(p, cache, dY, g, layers)
    rng = np.random.default_rng(1)
    flat_idx = rng.choice(p[block].size, size=min(12, p[block].size), replace=False)
    for i in flat_idx:
        idx = np.unravel_index(i, p[block].shape)
        h = 1e-6 * max(1.0, abs(p[block][idx]))
        q = {k: v.copy() for k, v in p.items()}
        q[block][idx] += h
        lp = _loss(q, X, Tc, Te, g, layers, L)[0]
        q[block][idx] -= 2 * h
        lm = _loss(q, X, Tc, Te, g, layers, L)[0]
        fd = (lp - lm) / (2 * h)
        an = grads[block][idx]
        assert abs(fd - an) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (block, idx, fd, an)


def test_planted_transpose_fails_the_difference_check(graph):
    """A decoder gradient computed with dY^T H instead of H^T dY (transposed
    back) is caught by the same finite differences."""
    case, g = graph
    layers = 1
    p, X, Tc, Te, L = _problem(g, layers)
    total, cache, dY = _loss(p, X, Tc, Te, g, layers, L)
    wrong = (dY.T @ cache["H"][layers]).T * 0.5 + 0.5 * (cache["H"][layers].T @ dY)[::-1]
    idx = (3, 5)
    h = 1e-6
    q = {k: v.copy() for k, v in p.items()}
    q["W_out"][idx] += h
    lp = _loss(q, X, Tc, Te, g, layers, L)[0]
    q["W_out"][idx] -= 2 * h
    lm = _loss(q, X, Tc, Te, g, layers, L)[0]
    fd = (lp - lm) / (2 * h)
    assert abs(fd - wrong[idx]) > 1e-3 * abs(fd)


def test_neighbour_mean_adjoint_dot_product(graph):
    case, g = graph
    rng = np.random.default_rng(2)
    H = rng.standard_normal((g.C, 8))
    dA = rng.standard_normal((g.C, 8))
    lhs = (g.mean_neighbours(H) * dA).sum()
    rhs = (H * g.mean_neighbours_adjoint(dA)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # a planted mistake (dividing by the receiving cell's count) breaks it on a
    # mesh with pentagons and hexagons
    wrong = np.zeros_like(dA)
    for s in range(g.coc.shape[1]):
        rows = np.flatnonzero(g.mask[:, s])
        np.add.at(wrong, g.coc[rows, s], dA[rows])
    wrong /= g.neoc[:, None]
    assert abs((H * wrong).sum() - lhs) > 1e-6 * abs(lhs)


def test_first_adam_step_closed_form():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(1000)
    gr = rng.standard_normal(1000)
    z = np.zeros(1000)
    new, m, v = hn.adam_step(p, gr, z, z, 1, lr=1e-3)
    # m_hat = g and v_hat = g^2 after one step: the update is lr * g / (|g| + eps)
    np.testing.assert_allclose(new, p - 1e-3 * gr / (np.abs(gr) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(p - new, 1e-3 * np.sign(gr), atol=1e-3 * 1e-8 / np.abs(gr).min() + 1e-15)


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, -2.5, 65504.0],
                 dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, 65536.0])
    np.testing.assert_array_equal(hn.bf16(x), want)


def test_uniform_flow_wind_reconstruction_is_exact(graph):
    """u_e = U . n_e for a uniform (east, north) wind: the least-squares cell
    wind returns U exactly (up to rounding)."""
    case, g = graph
    L = 4
    ue, un = 7.5, -3.25
    u = (ue * g.cos + un * g.sin)[None, :].repeat(L, axis=0)
    state = {"u": u, "w": np.zeros((L + 1, g.C)), "theta_m": np.ones((L, g.C)), "rho_zz": np.ones((L, g.C)),
             "scalars": np.zeros((1, L, g.C))}
    X = hn.features(state, g, np.zeros(g.E), L, moist=False).reshape(g.C, 6, hn.ROW)
    np.testing.assert_allclose(X[:, 0, :L], ue, rtol=0, atol=1e-12)
    np.testing.assert_allclose(X[:, 1, :L], un, rtol=0, atol=1e-12)
