"""Launch shapes of the gather kernels (wavefronts per block) and the
autotuner: every shape gives the same bits; the tuner's choice is no slower
than the default over whole steps; refusals."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_launch_shapes_give_the_same_bits(get_case):
    from hexdyn import autotune
    from hexdyn.config import DycoreConfig
    from hexdyn.driver import HexDycoreDriver

    print("[gpu] launch shapes x1.2562-moist", flush=True)
    case = get_case("x1.2562-moist", config=DycoreConfig(config_dt=1800.0, config_horiz_mixing="2d_smagorinsky"))
    runs = []
    # the last table sets the edge-flux pass's gather form explicitly (the
    # others run its edge-tiled form): both give the same bits
    for table in ({"*": 4}, {"*": 1}, {"*": 2}, {"*": 8}, {"*": 16},
                  {"*": 4, "k_diag_edge_flux": 16, "k_tend_mom_ac": 2, "k_sc_edge": 8}):
        d = HexDycoreDriver.from_host(case)
        for k, w in table.items():
            autotune.set_launch(d, k, w)
        d.step_device(2)
        runs.append(d.get_state())
        assert autotune.get_launch(d, "k_rec_cell") == table.get("k_rec_cell", table["*"])
        d.close()
    for r in runs[1:]:
        for k in runs[0]:
            assert _bits(r[k], runs[0][k]), k


def test_refusals(get_case):
    from hexdyn import autotune
    from hexdyn.driver import HexDycoreDriver

    d = HexDycoreDriver.from_host(get_case("x1.642-moist"))
    for k, w in (("k_vcoef", 2), ("k_ac_cell", 1), ("k_diag_edge_flux_t", 4), ("k_rec_cell", 3), ("k_rec_cell", 32),
                 ("notakernel", 2)):
        with pytest.raises(ValueError):
            autotune.set_launch(d, k, w)
    d.close()


def test_tuner_is_not_slower_and_keeps_bits(get_case, tmp_path, monkeypatch):
    from hexdyn import autotune
    from hexdyn.driver import HexDycoreDriver

    print("[gpu] autotune x1.2562-moist", flush=True)
    monkeypatch.setenv("HEXDYN_CACHE", str(tmp_path))
    case = get_case("x1.2562-moist")
    ref = HexDycoreDriver.from_host(case)
    ref.step_device(3)
    want = ref.get_state()
    ref.close()
    d = HexDycoreDriver.from_host(case)
    rec = autotune.tune(d, steps=1, check_steps=4)
    assert rec["tuned_ms_per_step"] <= rec["default_ms_per_step"] or rec["table"] == {}
    path = autotune.save(rec)
    assert path.exists()
    d.close()
    d2 = HexDycoreDriver.from_host(case)  # picks the stored table up at create
    assert d2.tuning is not None and d2.tuning["table"] == rec["table"]
    for k, w in rec["table"].items():
        assert autotune.get_launch(d2, k) == w
    d2.step_device(3)
    got = d2.get_state()
    d2.close()
    for k in want:
        assert _bits(got[k], want[k]), k
