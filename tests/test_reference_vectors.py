"""The oracle against golden vectors of the reference's own NumPy operators.

`tests/golden/reference_<name>.npz` holds the reference's outputs for the
functions only its own code pins bit for bit, written by
`tests/golden/make_reference_vectors.py` from diagnostics.py,
acoustic_v841.py, mixing.py, integration.py and transport.py on operands taken from the
oracle at fixed checkpoints of one large step (`tests/golden/reference_walk.py`).
Each test walks the oracle to the same checkpoint, checks that it holds the
same operands (sha256), runs the phase that restates the function and
compares every output word.  The case is this package's seeded x1.162 moist
case at 10 levels; its digest is checked first, so a changed case generator
or walk cannot pass silently.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle import Oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))

import reference_walk as rw  # noqa: E402

NAMES = ["solve_diagnostics", "vertical_coefficients", "acoustic", "divergence_damping", "recovery", "advection",
         "momentum", "omega"]


def _load(name):
    return np.load(GOLDEN / f"reference_{name}.npz", allow_pickle=False)


@pytest.fixture(scope="module")
def case(get_case):
    return get_case(rw.CASE, nlev=rw.NLEV)


def _same_bits(got, want, what):
    got = np.ascontiguousarray(got, dtype=np.float32).ravel()
    want = np.ascontiguousarray(want, dtype=np.float32).ravel()
    assert got.size == want.size, f"{what}: size {got.size} != {want.size}"
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (f"{what}: {bad.size} of {got.size} words differ, first at {bad[0]}: "
                           f"{got[bad[0]]!r} vs reference {want[bad[0]]!r}")


def _walk_to(case, label):
    for name, o in rw.walk(case):
        if name == label:
            return o, name
    raise KeyError(label)


def _checkpoints(case, z, table_for):
    """Walk once, yielding (checkpoint, oracle) at every checkpoint that has
    vectors in z, after checking the operands there."""
    for name, o in rw.walk(case):
        key = f"{name}.operands"
        if key in z.files:
            ops = rw.read_all(o, case, table_for(name))
            extra = []
            if name.startswith("acoustic_"):
                b64, c64 = o.btri_ctri_f64()
                extra = [rw.read(o, case, src, shp) for _, src, shp in rw.COEFFICIENTS]
                extra += [b64.astype(np.float32), c64.astype(np.float32), np.asarray(case.grid.rdzw, np.float32)]
            assert bytes(z[key]).decode() == bytes(rw.operand_digest(list(ops.values()) + extra)).decode(), \
                f"{name}: the walk reached different operands than the vectors were made from"
        yield name, o


@pytest.mark.parametrize("name", NAMES)
def test_vectors_were_made_on_this_case(case, name):
    assert bytes(_load(name)["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()


def test_solve_diagnostics_matches_reference_bits(case):
    """compute_solve_diagnostics (diagnostics.py:95-306), APVM on, stored
    inverses: rk step 3 reconstructs v, rk step 1 keeps the cached v."""
    z = _load("solve_diagnostics")
    o = Oracle(case, 32)
    from hexdyn.case import initial_state_arrays

    o.set_state(initial_state_arrays(case), rebuild=True)
    o.import_field("U", z["u"])
    o.import_field("RHO_ZZ", z["rho_zz"])
    for rk in (3, 1):
        o.import_field("V", z["v_cached"])
        o.solve_diagnostics(rk)
        names = ["V", "PV_EDGE"] + (["KE", "DIVERGENCE", "VORTICITY"] if rk == 3 else [])
        for f in names:
            _same_bits(o.export_field(f), z[f"rk{rk}_{f}"], f"rk{rk} {f}")


def test_vertical_coefficients_match_reference_bits(case):
    """compute_vertical_implicit_coefficients_v841 (acoustic_v841.py:81-258)
    at stages 1 and 2: cofwr/cofwz/coftz/cofwt, a/b/c and the Thomas
    factors alpha/gamma."""
    z = _load("vertical_coefficients")
    seen = 0
    for name, o in _checkpoints(case, z, lambda n: rw.VCOEF[int(n[-1])]):
        if f"{name}.operands" not in z.files:
            continue
        seen += 1
        for out, src, shp in rw.VCOEF_OUT:
            _same_bits(rw.read(o, case, src, shp), z[f"{name}.{out}"], f"{name} {out}")
        b64, c64 = o.btri_ctri_f64()
        _same_bits(b64.astype(np.float32), z[f"{name}.b_tri"], f"{name} b_tri")
        _same_bits(c64.astype(np.float32), z[f"{name}.c_tri"], f"{name} c_tri")
    assert seen == 2


def test_acoustic_substeps_match_reference_bits(case):
    """advance_acoustic_step_v841 (acoustic_v841.py:299-560): the stage-2
    substep and both stage-3 substeps (small_step 1 and 2)."""
    z = _load("acoustic")
    seen = 0
    pending = None
    for name, o in _checkpoints(case, z, lambda n: rw.ACOUSTIC_STATE + rw.ACOUSTIC_FORCING):
        if pending and name == f"{pending}_out":
            for out, src, shp in rw.ACOUSTIC_STATE:
                _same_bits(rw.read(o, case, src, shp), z[f"{pending}.{out}"], f"{pending} {out}")
            seen += 1
            pending = None
        if f"{name}.operands" in z.files:
            pending = name
    assert seen == 3


def test_divergence_damping_matches_reference_bits(case):
    """divergence_damping_3d (mixing.py:851-932) with the stored inverse of
    dts, at stages 2 and 3."""
    z = _load("divergence_damping")
    seen = 0
    pending = None
    for name, o in _checkpoints(case, z, lambda n: rw.DAMPING):
        if pending and name == f"{pending}_out":
            _same_bits(rw.read(o, case, "F:RU_P", "edge"), z[f"{pending}.ru_p"], f"{pending} ru_p")
            seen += 1
            pending = None
        if f"{name}.operands" in z.files:
            pending = name
    assert seen == 2


def test_recovery_matches_reference_bits(case):
    """recover_large_step_variables (integration.py:198-327) at stages 2 and
    3: densities, theta_m, rw and w with the terrain terms, u, the time
    averages, and at stage 3 exner (binary32 powf) and pressure."""
    z = _load("recovery")
    seen = 0
    pending = None
    for name, o in _checkpoints(case, z, lambda n: rw.RECOVERY_STATE):
        if pending and name == f"{pending}_out":
            for out, src, shp in rw.RECOVERY_OUT:
                _same_bits(rw.read(o, case, src, shp), z[f"{pending}.{out}"], f"{pending} {out}")
            seen += 1
            pending = None
        if f"{name}.operands" in z.files:
            pending = name
    assert seen == 2


def test_advection_stencils_match_reference_bits(case):
    """The v8.4.1 advection-stencil compression in native insertion order
    (transport.py _native_advection_order :2264-2403 and
    _native_advection_compression :2405-2443) from the same deriv_two:
    cells, adv_coefs and adv_coefs_3rd word for word.  The quadratic-fit
    deriv_two itself (initialize_deriv_two, :610-738) is a different
    evaluation of the same least-squares problem here: within 1e-10
    relative in binary64 (a file's deriv_two is used as given)."""
    from hexdyn import setup as hsetup

    z = _load("advection")
    d2 = hsetup.deriv_two(case.mesh)
    assert np.array_equal(d2.astype(np.float32).view(np.uint32), z["deriv_two"].view(np.uint32))
    adv, adv3, nadv, cells = hsetup.advection_coefficients(case.mesh, z["deriv_two"])
    assert np.array_equal(cells, z["adv_cells"])
    _same_bits(adv, z["adv_coefs"], "adv_coefs")
    _same_bits(adv3, z["adv_coefs_3rd"], "adv_coefs_3rd")
    ref = z["deriv_two_reference"]
    assert np.abs(d2 - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("vectors", ["transport", "transport_source"])
def test_transport_stages_match_reference_bits(case, vectors):
    """advance_scalars (transport.py:1649-1826) at stages 1 and 2 and
    advance_scalars_monotonic (:1829-2175, Zalesak FCT) at stage 3, all six
    species, split-transport densities: every mixing ratio word for word,
    without and with a held scalar tendency (scalar_tendency)."""
    z = _load(vectors)
    assert bytes(z["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()
    source = rw.scalar_tendency(case) if vectors == "transport_source" else None
    held, seen = {}, 0
    for name, o in rw.walk_transport(case, source):
        if not name.startswith("scalars_"):
            continue
        st = int(name.split("_")[1])
        if not name.endswith("_out"):
            ops = rw.read_all(o, case, rw.TRANSPORT)
            ops["scalar_stage"] = ops["scalar_old"] if st == 1 else rw.read(o, case, "F:SCALARS_STAGE", "tracer")
            held[st] = ops
            continue
        if st == 1:
            held.update(rw.read_all(o, case, rw.TRANSPORT_FLUX))
        ops = held[st]
        digest = rw.operand_digest([ops["scalar_old"], ops["scalar_stage"], ops["rho_zz_old"], ops["rho_zz_new"],
                                    held["uh_avg"], held["ww_avg"]])
        assert bytes(z[f"scalars_{st}.operands"]).decode() == bytes(digest).decode(), f"stage {st} operands"
        _same_bits(rw.read(o, case, "F:SCALARS_STAGE", "tracer"), z[f"scalars_{st}.scalars"], f"stage {st} scalars")
        seen += 1
    assert seen == 3


def test_subcycle_pools_match_reference_bits(case):
    """density_tendency (dynamics.py:122-148) and
    pressure_gradient_euler_tendency (:188-224) at the first subcycle start.
    The density tendency follows the reference's device kernel
    (cuda_horizontal.py:412-433, `0 - div`); the NumPy form `-div` differs
    from it only in the sign of an exactly zero divergence, which these
    operands do not hold."""
    z = _load("pools")
    assert bytes(z["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()
    for name, o in rw.walk(case):
        if name == "vcoef_1":
            ops = rw.read_all(o, case, rw.POOLS)
            assert bytes(z["pools_1.operands"]).decode() == bytes(rw.operand_digest(ops.values())).decode()
            assert np.count_nonzero(z["pools_1.tend_rho"] == 0) == 0
            _same_bits(rw.read(o, case, "F:TEND_RHO", "cell"), z["pools_1.tend_rho"], "tend_rho")
            _same_bits(rw.read(o, case, "F:EULER_RU", "edge"), z["pools_1.euler_ru"], "euler_ru")
            return
    raise AssertionError("checkpoint not reached")


def test_damping_profile_matches_reference_bits():
    """build_v841_vertical_velocity_damping (damping_v841.py:23-107): the
    Rayleigh w damping recomputed from zgrid, word for word."""
    from hexdyn import setup as hsetup

    z = _load("damping_profile")
    xnutr, zd = (float(v) for v in z["xnutr_zd"])
    _same_bits(hsetup.vertical_velocity_damping(z["zgrid"], xnutr=xnutr, zd=zd), z["dss"], "dss")


def test_offcentering_matches_reference_bits():
    """build_v841_acoustic_offcentering (offcentering_v841.py:45-174) on a
    column whose top reaches into the 30-50 km sine transition."""
    from hexdyn import setup as hsetup

    z = _load("offcentering")
    got = hsetup.acoustic_offcentering(z["rdzw"])
    for name, val in zip(("etp", "etm", "ewp", "ewm"), got):
        _same_bits(val, z[name], name)
    assert len(set(z["etp"].tolist())) > 5  # the transition is exercised


def test_vertical_transport_of_momentum_matches_reference_bits(case):
    """vertical_transport_u (dynamics.py:227-270): the first term of the
    momentum tendency, as the oracle's dyn_tend computes it, from the stage-1
    recovered u and rw."""
    z = _load("vertical_transport")
    assert bytes(z["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()
    for name, o in rw.walk(case):
        if name == "vcoef_2":
            ops = rw.read_all(o, case, [("normal_velocity", "I:u", "edge"), ("rho_w", "I:c_rw", "w")])
            assert bytes(z["vtu_2.operands"]).decode() == bytes(rw.operand_digest(ops.values())).decode()
            _same_bits(o.vertical_transport_u(), z["vtu_2.result"], "vertical transport of u")
            return
    raise AssertionError("checkpoint not reached")


def test_split_flux_average_matches_reference_bits(case):
    """accumulate_split_flux then finish_split_flux (integration.py:27-82):
    the subcycles' mass fluxes summed in source order and scaled by the
    stored 1/3, the transport's uh_avg and ww_avg."""
    z = _load("split_flux")
    assert bytes(z["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()
    fluxes = []
    for name, o in rw.walk_transport(case):
        if name.startswith("subcycle_"):
            fluxes += [rw.read(o, case, "F:FLUX_U", "edge"), rw.read(o, case, "F:FLUX_W", "w")]
        elif name == "scalars_1_out":
            assert bytes(z["split.operands"]).decode() == bytes(rw.operand_digest(fluxes)).decode()
            _same_bits(rw.read(o, case, "I:split_u", "edge"), z["split.uh_avg"], "uh_avg")
            _same_bits(rw.read(o, case, "I:split_w", "w"), z["split.ww_avg"], "ww_avg")
            return
    raise AssertionError("checkpoint not reached")


def test_moist_coefficients_match_reference_bits(case):
    """compute_moist_coefficients (regional_v841.py:1180-1219): qtot, cqw and
    cqu of the step-start scalars, six loading species."""
    z = _load("moist_coefficients")
    assert bytes(z["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()
    for name, o in rw.walk(case):
        if name == "vcoef_1":
            q = rw.read(o, case, "I:o_q", "tracer")
            assert bytes(z["moist.operands"]).decode() == bytes(rw.operand_digest([q])).decode()
            _same_bits(rw.read(o, case, "F:QTOT", "cell"), z["moist.qtot"], "qtot")
            _same_bits(rw.read(o, case, "F:CQW", "cell"), z["moist.cqw"], "cqw")
            _same_bits(rw.read(o, case, "F:CQU", "edge"), z["moist.cqu"], "cqu")
            return
    raise AssertionError("checkpoint not reached")


def _walk_checked(case, z, key, table, label):
    for name, o in rw.walk(case):
        if name == label:
            ops = rw.read_all(o, case, table)
            assert bytes(z[key]).decode() == bytes(rw.operand_digest(ops.values())).decode(), \
                f"{label}: the walk reached different operands than the vectors were made from"
            return o, ops
    raise AssertionError("checkpoint not reached")


def test_momentum_terms_match_reference_bits(case):
    """vector_invariant_momentum_tendency (dynamics.py:273-305) fed the
    reference's mass_flux_divergence (:86-121): the oracle's momentum terms
    (the code of every stage's tendency, v8.4.1 form with zero reference
    wind) and its mass-flux divergence equal the reference word for word."""
    z = _load("momentum")
    assert bytes(z["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()
    o, _ = _walk_checked(case, z, "momentum.operands", rw.MOMENTUM, "vcoef_2")
    got = o.momentum_terms()
    _same_bits(o.export_internal("hdiv"), z["momentum.hdiv"], "mass-flux divergence")
    _same_bits(got, z["momentum.result"], "momentum terms")


def test_omega_conversion_matches_reference_bits(case):
    """convert_w_tendency_to_omega (acoustic.py:246-281): the terrain
    zb/zb3 correction with the upwind sign of the momentum tendency and the
    interface zz, on the stage's Euler w pool and momentum tendency."""
    z = _load("omega")
    assert bytes(z["case_digest"]).decode() == bytes(rw.case_digest(case)).decode()
    o, ops = _walk_checked(case, z, "omega.operands", rw.OMEGA, "vcoef_2")
    _same_bits(o.omega_conversion(ops["w_tendency"], ops["u_tendency"]), z["omega.result"], "omega tendency")


def test_omega_vectors_catch_a_flipped_upwind_sign(case):
    """The same comparison fails when the oracle's upwind sign is planted
    reversed (the check is not blind to that mistake)."""
    from oracle.oracle import PLANT_OMEGA_SIGN

    z = _load("omega")
    o, ops = _walk_checked(case, z, "omega.operands", rw.OMEGA, "vcoef_2")
    o.set_plant(PLANT_OMEGA_SIGN)
    got = o.omega_conversion(ops["w_tendency"], ops["u_tendency"])
    assert np.count_nonzero(got.view(np.uint32) != z["omega.result"].view(np.uint32)) > 0


@pytest.fixture(scope="module")
def mixing_case():
    return rw.build_mixing_case()


def test_mesh_scaling_matches_reference_bits(mixing_case):
    """compute_mesh_mixing_scaling (mixing.py:271-323) on a variable mesh
    density: setup.mesh_mixing_scaling (the host evaluation hexdyn_create and
    the oracle make) gives the same del2/del4 words."""
    from hexdyn.setup import mesh_mixing_scaling

    z = _load("mixing")
    assert np.array_equal(z["mixing_density"], mixing_case.mesh.arrays["meshDensity"])
    del2, del4 = mesh_mixing_scaling(mixing_case.mesh)
    assert np.count_nonzero(del2 != np.float32(1)) > 0
    _same_bits(del2, z["scaling.del2"], "meshScalingDel2")
    _same_bits(del4, z["scaling.del4"], "meshScalingDel4")


def test_mixing_filters_match_reference_bits(mixing_case):
    """The del2/del4 filters of the v8.4.1 mixing on the oracle's kdiff
    (mixing.py momentum_horizontal_filter_tendency :391-530 onto the u pool,
    vertical_momentum_horizontal_filter_tendency :533-635 added to the w pool,
    theta_horizontal_filter_tendency :638-733), variable mesh density,
    visc4 0.05: every word equal."""
    z = _load("mixing")
    assert bytes(z["mixing_case_digest"]).decode() == bytes(rw.case_digest(mixing_case)).decode()
    ops = None
    for name, o in rw.walk_mixing(mixing_case):
        if name == "mixing":
            ops = rw.read_all(o, mixing_case, rw.MIXING)
            continue
        kdiff = rw.read(o, mixing_case, "I:kdiff", "cell")
        assert bytes(z["mixing.operands"]).decode() == bytes(rw.operand_digest(list(ops.values()) + [kdiff])).decode()
        _same_bits(rw.read(o, mixing_case, "I:euler_ru", "edge"), z["mixing.euler_ru"], "u pool with del2/del4 u")
        _same_bits(rw.read(o, mixing_case, "I:euler_theta", "cell"), z["mixing.theta_tendency"], "theta del2/del4")
        want_w = (ops["euler_rw"] + z["mixing.w_tendency"]).astype(np.float32)
        _same_bits(rw.read(o, mixing_case, "I:euler_rw", "w"), want_w, "w pool with del2/del4 w")
        return
    raise AssertionError("checkpoint not reached")
