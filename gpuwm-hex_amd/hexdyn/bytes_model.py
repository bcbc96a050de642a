"""Algorithmic HBM bytes of the dycore large step.

Two models, both in binary32 words:

* ``step_bytes``: the phase-level count.  Each named array a phase reads or
  writes is counted once at its declared shape; in-phase intermediates are
  not counted (except the FCT scale factors, which need a grid-wide
  synchronisation); level-independent geometry is excluded.  Floats per
  (cell, level), with E = 3C and V = 2C:

      solve_diagnostics        20           per call, 4 calls per subcycle
      implicit coefficients    18           stages 1 and 2
      subcycle pools           18           once per subcycle
      mixing                   19           once per subcycle, if on
      stage tendencies         33 + 2*ME    per stage
      acoustic substep         50           per substep
      divergence damping        9           per substep
      recovery                 38 + 2*ME    per stage
      split flux               10           per subcycle
      moist coefficients       T + 5        per step
      transport stages 1-2     4T + 13      per stage
      transport stage 3 (FCT)  8T + 13      once

* ``kernel_bytes``: per launch of one HIP kernel, every distinct array the
  kernel reads or writes counted once at its shape (neighbour gathers of the
  same array are not counted again).  This is the numerator of a kernel's
  achieved-bandwidth roofline.
"""

from __future__ import annotations

from typing import Dict, List, Tuple

WORD = 4
HBM_PEAK_GBS = 8000.0


def floats_per_cell_level(*, T: int, max_edges: int, n_ac: int, mix: bool, splits: int = 3) -> float:
    per_sub = (4 * 20 + 2 * 18 + 18 + (19 if mix else 0) + 3 * (33 + 2 * max_edges) + n_ac * (50 + 9)
               + 3 * (38 + 2 * max_edges) + 10)
    transport = (2 * (4 * T + 13) + (8 * T + 13)) if T > 0 else 0
    return splits * per_sub + (T + 5 if T > 0 else 0) + transport


def step_bytes(*, C: int, L: int, T: int, max_edges: int, n_ac: int, mix: bool = False, splits: int = 3) -> float:
    """Algorithmic bytes of one large step (phase-level count)."""
    return WORD * C * L * floats_per_cell_level(T=T, max_edges=max_edges, n_ac=n_ac, mix=mix, splits=splits)


# Per-kernel array lists: (entity, levels) with entity in {"C", "E", "V"} and
# levels "L" (layer) or "W" (interface, L+1).  "S" prefix: one row per scalar.
_K: Dict[str, List[Tuple[str, str]]] = {
    "k_diag_vertex": [("E", "L"), ("V", "L"), ("V", "L"), ("V", "L")],
    "k_diag_cell": [("E", "L"), ("V", "L"), ("V", "L"), ("C", "L"), ("C", "L"), ("C", "L")],
    "k_diag_edge": [("E", "L"), ("V", "L"), ("C", "L"), ("E", "L"), ("E", "L")],
    # k_diag_edge's arrays, then k_tend_flux's (the whole-step form of the pass)
    "k_diag_edge_flux": [("E", "L"), ("V", "L"), ("C", "L"), ("E", "L"), ("E", "L"),
                         ("E", "L"), ("C", "L"), ("C", "W"), ("E", "L"), ("C", "L"),
                         ("E", "L"), ("E", "L"), ("E", "W")],
    "k_moist_cell": [("SC", "L"), ("C", "L"), ("C", "L")],
    "k_moist_edge": [("SC", "L"), ("E", "L")],
    # zz, exner, theta, cqw, qtot, rtheta_p, rb, rtb, exb | cofwz, cofwt, atri, alpha, gamma, coftz
    "k_vcoef": [("C", "L")] * 14 + [("C", "W")],
    # with 2-D Smagorinsky mixing (the default): + u, v | kdiff
    "k_sub_cell": [("C", "L")] * 12 + [("C", "W")] * 3 + [("E", "L")] * 3 + [("C", "L")],
    # with mixing: + divergence, vorticity, kdiff, c_rho | delsq_u
    "k_sub_edge": [("E", "L")] * 5 + [("C", "L")] * 3 + [("C", "L"), ("V", "L"), ("C", "L"), ("C", "L"),
                                                          ("E", "L")],
    # u, c_rw, pv_edge, c_rho, ke, hdiv, euler_ru | tend_ru
    "k_tend_mom": [("E", "L"), ("C", "W"), ("E", "L"), ("C", "L"), ("C", "L"), ("C", "L"), ("E", "L"),
                   ("E", "L")],
    # k_tend_mom's arrays and the first acoustic substep's ru_p (ru_avg only with > 2 substeps)
    "k_tend_mom_ac": [("E", "L"), ("C", "W"), ("E", "L"), ("C", "L"), ("C", "L"), ("C", "L"), ("E", "L"),
                      ("E", "L"), ("E", "L")],
    # c_ru, theta, w, s_ru, s_theta | fth, fcorr, fw   (s_ru, s_theta, fcorr at stages 2-3)
    "k_tend_flux": [("E", "L"), ("C", "L"), ("C", "W"), ("E", "L"), ("C", "L"),
                    ("E", "L"), ("E", "L"), ("E", "W")],
    # theta, s_theta, c_rw, s_rw, w, fth, fcorr, fw, euler_rw, tend_ru, zb, zb3, zz | tend_rt, tend_rw
    "k_tend_cell": [("C", "L"), ("C", "L"), ("C", "W"), ("C", "W"), ("C", "W"), ("E", "L"), ("E", "L"),
                    ("E", "W"), ("C", "W"), ("E", "L"), ("CM", "W"), ("CM", "W"), ("C", "L"),
                    ("C", "L"), ("C", "W")],
    # the whole-step form (a second substep: the first one's damping applied, ru_avg formed from ru_p):
    # tend_ru, ru_p, rtheta_pp, rtheta_pp_old, s_theta, zz, cqu, s_exner, zxu, rho_pp | ru_p, ru_avg
    "k_ac_edge": [("E", "L"), ("E", "L"), ("C", "L"), ("C", "L"), ("C", "L"), ("C", "L"), ("E", "L"),
                  ("C", "L"), ("E", "L"), ("C", "L"), ("E", "L"), ("E", "L")],
    # rho_pp, rtheta_pp, rw_p, ww_avg, ru_p, s_theta, tend_rho, tend_rt, tend_rw, coftz, cofwz,
    # cofwt, alpha, gamma, zz, dss, s_rw, c_rw, c_rho, w | rw_p, ww_avg, rho_pp, rtheta_pp, rtheta_pp_old
    # (cofwr and the sub-diagonal are formed from zz and the coefficients read)
    "k_ac_cell": [("C", "L"), ("C", "L"), ("C", "W"), ("C", "W"), ("E", "L"), ("C", "L"), ("C", "L"),
                  ("C", "L"), ("C", "W"), ("C", "W"), ("C", "L"), ("C", "L"),
                  ("C", "L"), ("C", "L"), ("C", "L"), ("C", "L"), ("C", "W"), ("C", "W"), ("C", "L"),
                  ("C", "W"), ("C", "W"), ("C", "W"), ("C", "L"), ("C", "L"), ("C", "L")],
    "k_divdamp": [("C", "L"), ("C", "L"), ("C", "L"), ("E", "L"), ("E", "L")],
    # s_rho_p, rho_pp, s_rtheta_p, rtheta_pp, rb, rtb, s_rw, rw_p, ww_avg, s_ru, ru_p, zz, zb, zb3, exb |
    # rho_p, rtheta_p, c_rho, c_rt, theta, c_rw, flux_w, w, (exner, pp)
    "k_rec_cell": [("C", "L")] * 6 + [("C", "W")] * 3 + [("E", "L")] * 2 + [("C", "L"), ("CM", "W"), ("CM", "W"),
                   ("C", "L")] + [("C", "L")] * 5 + [("C", "W")] * 3 + [("C", "L")] * 2,
    # the whole-step form after one damped substep: s_ru, ru_p, s_rho_p, rho_pp, rb, rtheta_pp,
    # rtheta_pp_old, s_theta | ru_p, c_ru, u
    "k_rec_edge": [("E", "L")] * 2 + [("C", "L")] * 6 + [("E", "L")] * 3,
    "k_sc_edge": [("SC", "L"), ("E", "L"), ("SE", "L")],
    "k_sc_cell": [("SC", "L"), ("C", "W"), ("E", "L"), ("SE", "L"), ("C", "L"), ("C", "L"), ("SC", "L"),
                  ("SC", "L")],
    # qs, qold, split_u | hres
    "k_fct_edge1": [("SC", "L"), ("SC", "L"), ("E", "L"), ("SE", "L")],
    # qs, qold, rho, split_w, split_u, hres | target, mass, sin, sout, vres
    "k_fct_cell1": [("SC", "L"), ("SC", "L"), ("C", "L"), ("C", "W"), ("E", "L"), ("SE", "L"), ("C", "L"),
                    ("SC", "L"), ("SC", "L"), ("SC", "L"), ("SC", "W")],
    "k_fct_edge2": [("SC", "L"), ("SC", "L"), ("SE", "L"), ("SE", "L")],
    "k_fct_cell2": [("SC", "L"), ("SC", "L"), ("SC", "W"), ("SC", "L"), ("SE", "L"), ("C", "L"), ("SC", "L")],
    # horizontal mixing, once per subcycle
    "k_mix_lapdiv": [("E", "L"), ("C", "L")],
    "k_mix_lapvort": [("E", "L"), ("V", "L")],
    # delsq_div, delsq_vort, c_rho, euler_ru | euler_ru
    "k_mix_edge4": [("C", "L"), ("V", "L"), ("C", "L"), ("E", "L"), ("E", "L")],
    # theta, c_rho, w, kdiff | delsq_theta, euler_theta, delsq_w, tend_w
    "k_mix_cell": [("C", "L"), ("C", "L"), ("C", "W"), ("C", "L"), ("C", "L"), ("C", "L"), ("C", "L"), ("C", "W")],
    # delsq_theta, delsq_w, euler_theta, tend_w, euler_rw | euler_theta, euler_rw
    "k_mix_cell4": [("C", "L"), ("C", "L"), ("C", "L"), ("C", "W"), ("C", "W"), ("C", "L"), ("C", "W")],
}
# state rebuild (after a state is set): rho, rtheta_m, zz, rtb, rb, exb, rw, ru, zb, zb3 |
# theta, exner, rho_p, rtheta_p, p', w
_K["k_rebuild_cell"] = ([("C", "L")] * 6 + [("C", "W"), ("E", "L"), ("CM", "W"), ("CM", "W")]
                        + [("C", "L")] * 5 + [("C", "W")])
# ru, rho | u
_K["k_rebuild_edge"] = [("E", "L"), ("C", "L"), ("E", "L")]
_K["k_diag_edge_flux_t"] = _K["k_diag_edge_flux"]  # the edge-tiled form moves the same arrays


def kernel_bytes(kernel: str, *, C: int, E: int, V: int, L: int, T: int, max_edges: int = 6) -> float:
    """Algorithmic bytes of one launch of ``kernel`` (scalar kernels: all species of one launch)."""
    if kernel not in _K:
        raise KeyError(kernel)
    total = 0
    for ent, lev in _K[kernel]:
        nl = L if lev == "L" else L + 1
        if ent == "C":
            rows = C
        elif ent == "E":
            rows = E
        elif ent == "V":
            rows = V
        elif ent == "CM":
            rows = C * max_edges
        elif ent == "SC":
            rows = C * T
        elif ent == "SE":
            rows = E * T
        else:
            rows = C
        total += rows * nl
    return float(WORD * total)


KNOWN_KERNELS = tuple(_K)
