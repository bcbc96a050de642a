from hexdyn.bytes_model import floats_per_cell_level, kernel_bytes, step_bytes, KNOWN_KERNELS


def test_phase_table_headline_value():
    # x1.40962 L55, T=6, maxEdges 6, 4 acoustic substeps per subcycle (1, 1, 2), mixing on
    f = floats_per_cell_level(T=6, max_edges=6, n_ac=4, mix=True)
    assert 2150 <= f <= 2250
    b = step_bytes(C=40962, L=55, T=6, max_edges=6, n_ac=4, mix=True)
    assert 19.0e9 < b < 20.5e9


def test_kernel_bytes_defined():
    for k in KNOWN_KERNELS:
        assert kernel_bytes(k, C=2562, E=7680, V=5120, L=55, T=6) > 0


def test_every_profiled_kernel_has_an_array_list():
    """The committed kernel profile's rows (every kernel a headline step
    launches) all have a per-launch array list, so the bench can price
    whichever of them dominates."""
    import re
    from pathlib import Path

    text = (Path(__file__).resolve().parents[1] / "profiles" / "kernels_x1.40962-moist.md").read_text()
    names = re.findall(r"^\| (k_\w+) \|", text, re.M)
    assert names
    assert [n for n in names if n not in KNOWN_KERNELS] == []
