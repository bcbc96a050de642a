"""Learned corrector on the GPU (libhexdyn.so hexdyn_net_* through the C ABI)
against the binary64 statement of the same network (oracle/hexnet_oracle.py).

Tolerances: the device keeps activations and gradients in bf16 (8-bit
significand, unit roundoff 2^-9 ~ 2e-3) with fp32 accumulation, against a
binary64 statement fed the same bf16 weights and inputs; relative L2 of the
outputs <= 2e-2, of each weight gradient <= 5e-2.  Adam runs in fp32:
relative 1e-6.  State copies and fallbacks are bitwise.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import hexnet_oracle as hn  # noqa: E402


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _graph(case):
    m = case.mesh
    return hn.Graph(m.cellsOnCell, m.nEdgesOnCell, m.edgesOnCell, m.cellsOnEdge, case.cos_angle_edge,
                    case.sin_angle_edge)


@pytest.fixture(scope="module")
def pair(get_case):
    from hexdyn import learned

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    case = get_case("x1.642-moist")
    fine, coarse = learned.make_pair_drivers(case)
    yield case, fine, coarse
    fine.close()
    coarse.close()


def _stepped_pair(case, fine, coarse, seed=7):
    from hexdyn import learned

    fine.set_state(learned.perturbed_state(case, seed=seed))
    fine.step_device(2)
    learned.copy_state(coarse, fine)
    coarse.step_device(1)
    fine.step_device(2)


def test_copy_state_is_bitwise(pair):
    from hexdyn import learned

    print("[gpu] hexdyn_copy_state x1.642-moist", flush=True)
    case, fine, coarse = pair
    fine.set_state(learned.perturbed_state(case, seed=3))
    fine.step_device(1)
    learned.copy_state(coarse, fine)
    a, b = fine.get_state(), coarse.get_state()
    for k in a:
        assert np.array_equal(np.asarray(a[k]).view(np.uint32), np.asarray(b[k]).view(np.uint32)), k
    assert coarse.model_time() == fine.model_time()


def test_features_match_statement(pair):
    from hexdyn import learned

    print("[gpu] corrector features x1.642-moist", flush=True)
    case, fine, coarse = pair
    _stepped_pair(case, fine, coarse)
    net = learned.CorrectorNet(coarse, 4)
    got = net.raw_features(coarse).cpu().numpy()
    st = coarse.get_state()
    want = hn.features(st, _graph(case), case.mesh.arrays['fEdge'], case.n_levels, moist=case.n_moist > 0)
    scale = np.abs(want.reshape(-1, 6, 64)).max(axis=(0, 2))  # per channel
    err = np.abs(got - want).reshape(-1, 6, 64).max(axis=(0, 2))
    assert np.all(err <= 1e-5 * scale + 1e-30), (err, scale)
    net.close()


@pytest.fixture(scope="module")
def trained_inputs(pair):
    """A network with random weights, scales from the pair and one forward,
    loss and backward pass; everything the statement needs exported."""
    from hexdyn import learned

    case, fine, coarse = pair
    _stepped_pair(case, fine, coarse, seed=11)
    L = case.n_levels
    net = learned.CorrectorNet(coarse, 4)
    flat = learned.init_params(4, seed=5, out_gain=0.5)
    rng = np.random.default_rng(6)
    parts = learned.unpack(flat, 4)
    for name, a in parts.items():
        if name.startswith("b_"):
            a[...] = (rng.standard_normal(a.shape) * 0.05).astype(np.float32)
    net.set_params(flat)
    stats = learned.ScaleStats()
    stats.add(net, coarse, fine)
    net.set_scales(*stats.scales(L))
    net.forward(coarse, keep=True)
    net.set_target(coarse, fine)
    lo = net.loss()
    net.backward()
    out = {"flat": flat, "loss": lo, "grads": net.get_grads(), "X": net.export("X").cpu().numpy(),
           "Y": net.export("Y").cpu().numpy(), "Tc": net.export("target_cells").cpu().numpy(),
           "Te": net.export("target_edges").cpu().numpy(), "net": net, "L": L}
    yield out
    net.close()


def test_forward_and_loss_match_fp64_statement(pair, trained_inputs):
    print("[gpu] corrector forward vs FP64 x1.642-moist", flush=True)
    case = pair[0]
    t = trained_inputs
    g = _graph(case)
    p = hn.unpack(t["flat"], 4)
    cache = hn.forward(p, t["X"], g, 4)
    assert _rel(t["Y"], cache["Y"]) <= 2e-2
    parts, _ = hn.loss(cache["Y"], t["Tc"], t["Te"], g, t["L"])
    for k in ("total", "u", "theta_m", "rho_zz", "w"):
        assert abs(t["loss"][k] - parts[k]) <= 2e-2 * parts[k], (k, t["loss"][k], parts[k])


def test_gradients_match_fp64_statement(pair, trained_inputs):
    print("[gpu] corrector gradients vs FP64 x1.642-moist", flush=True)
    case = pair[0]
    t = trained_inputs
    g = _graph(case)
    p = hn.unpack(t["flat"], 4)
    cache = hn.forward(p, t["X"], g, 4)
    _, dY = hn.loss(cache["Y"], t["Tc"], t["Te"], g, t["L"])
    want = hn.backward(p, cache, dY, g, 4)
    from hexdyn import learned

    got = learned.unpack(t["grads"], 4)
    bad = []
    for name, _ in hn.layout(4):
        r = _rel(got[name], want[name])
        if not r <= 5e-2:
            bad.append((name, r))
    assert not bad, bad


def test_adam_step_matches_statement(trained_inputs):
    print("[gpu] corrector Adam step", flush=True)
    t = trained_inputs
    net = t["net"]
    p0 = net.get_params().astype(np.float64)
    g = t["grads"].astype(np.float64)
    net.adam(1e-3, 1)
    p1 = net.get_params().astype(np.float64)
    want, _, _ = hn.adam_step(p0, g, np.zeros_like(g), np.zeros_like(g), 1, lr=1e-3)
    np.testing.assert_allclose(p1, want, rtol=1e-6, atol=1e-9)


def test_apply_keeps_dry_mass(pair, trained_inputs):
    print("[gpu] corrector apply conserves dry mass", flush=True)
    case, fine, coarse = pair
    net = trained_inputs["net"]
    m0 = coarse.integrals()["dry_mass"]
    net.forward(coarse)
    before = coarse.get_state()
    net.apply(coarse)
    after = coarse.get_state()
    m1 = coarse.integrals()["dry_mass"]
    assert abs(m1 - m0) <= 1e-9 * abs(m0)
    # the checks' own binary64 dry mass is the same sum in another order
    ch = net.checks(coarse)
    assert abs(ch[3] - m1) <= 1e-12 * abs(m1) and all(np.isfinite(ch[:3]))
    assert not np.array_equal(before["u"], after["u"])  # the correction is not empty


def test_training_reduces_loss_and_gains_over_the_coarse_step(get_case):
    """Train on pairs of the x1.642 moist case generated by the dycore; on
    held-out pairs (other seeds and spin-ups) the trained network's loss is
    below the untrained one's, which is the error of the plain coarse step
    (the untrained correction is about 1 % of the target scale)."""
    from hexdyn import learned

    print("[gpu] corrector training x1.642-moist", flush=True)
    case = get_case("x1.642-moist")
    fine, coarse = learned.make_pair_drivers(case)
    tc = learned.TrainConfig(pairs=240, trajectories=4, layers=4, lr=1e-3, lr_final=1e-4, seed=0, spinup_min=2,
                             spinup_max=6)
    net = learned.CorrectorNet(coarse, tc.layers)
    net.set_params(learned.init_params(tc.layers, tc.seed))
    learned.fit_scales(net, case, fine, coarse, tc)
    seeds, spins = (901, 902, 903), (10, 40, 90)
    before = learned.validation_loss(net, case, fine, coarse, seeds, spins)
    net, hist = learned.train(case, tc, log=lambda m: print(m, flush=True), drivers=(fine, coarse), net=net,
                              fit=False)
    after = learned.validation_loss(net, case, fine, coarse, seeds, spins)
    ratios = [a / b for a, b in zip(after, before)]
    print(f"[gpu] held-out loss before {before} after {after} ratios {ratios}", flush=True)
    assert all(np.isfinite(after))
    assert max(ratios) < 0.9
    net.close()
    fine.close()
    coarse.close()


def test_fallback_on_tripped_checks_equals_plain_steps(get_case):
    """Checks forced to trip from the host (a zero CFL limit, a zero mass
    tolerance, a zero divergence limit): every step falls back, and the
    state equals two plain fine steps per step, bit for bit."""
    from hexdyn import learned
    from hexdyn.driver import HexDycoreDriver

    print("[gpu] corrector fallback x1.642-moist", flush=True)
    case = get_case("x1.642-moist")
    state = learned.perturbed_state(case, seed=0)
    ref = HexDycoreDriver.from_host(case, state=state)
    ref.step_device(6)
    want = ref.get_state()
    ref.close()
    flat = learned.init_params(4, seed=1)
    inf = float("inf")
    for limits, why in ((learned.Limits(cfl_h=0.0, mass_rel=inf), "cfl_h"),
                        (learned.Limits(cfl_h=inf, div=0.0, mass_rel=inf), "div"),
                        (learned.Limits(cfl_h=inf, mass_rel=0.0), "mass")):
        st = learned.LearnedStepper(case, limits=limits)
        st.net.set_params(flat)
        st.net.set_scales(np.zeros(learned.N_IN, np.float32), np.full(learned.N_IN, 1e-3, np.float32),
                          np.full(learned.N_OUT, 1e-3, np.float32))
        st.set_state(state)
        st.step(3)
        assert st.fallbacks == 3 and set(st.trips) == {why}, (st.trips, why)
        got = st.get_state()
        for k in want:
            assert np.array_equal(np.asarray(got[k]).view(np.uint32), np.asarray(want[k]).view(np.uint32)), (why, k)
        st.close()


def test_ragged_mesh_forward_and_gradients_match_statement(get_case):
    """The maxEdges-7 instantiations (features, neighbour mean and its
    adjoint, loss gather, checks) on the r642 mesh with heptagons, against
    the binary64 statement."""
    from hexdyn import learned

    print("[gpu] corrector on the ragged r642 mesh (maxEdges 7)", flush=True)
    case = get_case("r642-moist")
    assert case.mesh.max_edges == 7
    fine, coarse = learned.make_pair_drivers(case)
    _stepped_pair(case, fine, coarse, seed=5)
    net = learned.CorrectorNet(coarse, 2)
    flat = learned.init_params(2, seed=9, out_gain=0.5)
    net.set_params(flat)
    stats = learned.ScaleStats()
    stats.add(net, coarse, fine)
    net.set_scales(*stats.scales(case.n_levels))
    g = _graph(case)
    got_raw = net.raw_features(coarse).cpu().numpy()
    want_raw = hn.features(coarse.get_state(), g, case.mesh.arrays["fEdge"], case.n_levels, moist=True)
    scale = np.abs(want_raw.reshape(-1, 6, 64)).max(axis=(0, 2))
    assert np.all(np.abs(got_raw - want_raw).reshape(-1, 6, 64).max(axis=(0, 2)) <= 1e-5 * scale + 1e-30)
    net.forward(coarse, keep=True)
    net.set_target(coarse, fine)
    lo = net.loss()
    net.backward()
    X, Y = net.export("X").cpu().numpy(), net.export("Y").cpu().numpy()
    Tc, Te = net.export("target_cells").cpu().numpy(), net.export("target_edges").cpu().numpy()
    p = hn.unpack(flat, 2)
    cache = hn.forward(p, X, g, 2)
    assert _rel(Y, cache["Y"]) <= 2e-2
    parts, dY = hn.loss(cache["Y"], Tc, Te, g, case.n_levels)
    assert abs(lo["total"] - parts["total"]) <= 2e-2 * parts["total"]
    want = hn.backward(p, cache, dY, g, 2)
    got = learned.unpack(net.get_grads(), 2)
    bad = [(n, _rel(got[n], want[n])) for n, _ in hn.layout(2) if not _rel(got[n], want[n]) <= 5e-2]
    assert not bad, bad
    ch = net.checks(coarse)
    assert all(np.isfinite(ch)) and ch[3] > 0
    net.close()
    fine.close()
    coarse.close()
