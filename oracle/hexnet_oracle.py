"""TEST INFRASTRUCTURE — binary64 NumPy statement of the learned corrector's
network (forward pass, loss, backward pass and one Adam step).

Only tests/ use this module, as the checker of the device network in
libhexdyn.so (hexdyn_net_*).  It follows the network the device kernels
compute (csrc/hexdyn_net.hip) with every operation in binary64: the same
bf16-rounded weight matrices and the same bf16 inputs, no rounding of the
activations or gradients.  Its gradients are pinned by central finite
differences of its own loss (tests/test_hexnet_oracle.py), and its Adam step
by the closed form of the first step.

Shapes (logical cell and edge order):
  X [C][384], H [C][128], Y [C][320] = (east wind, north wind, theta_m,
  rho_zz, w) x 64 rows; edge prediction of u [E][64];
  targets Tc [C][3][64] (theta_m, rho_zz, w) and Te [E][64].
"""

from __future__ import annotations

from typing import Dict, List

import numpy as np

HID = 128
N_IN = 384
N_OUT = 320
ROW = 64


def bf16(x) -> np.ndarray:
    """Round binary32 values to bf16 (nearest, ties to even), as binary64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(layers: int) -> List[tuple]:
    out = [("W_in", (N_IN, HID)), ("b_in", (HID,))]
    for l in range(layers):
        out += [(f"W_{l}", (2 * HID, HID)), (f"b_{l}", (HID,))]
    out += [("W_out", (HID, N_OUT)), ("b_out", (N_OUT,))]
    return out


def unpack(flat: np.ndarray, layers: int, round_weights: bool = True) -> Dict[str, np.ndarray]:
    """Parameter blocks in binary64; matrices rounded to bf16 as the device
    multiplies them, biases kept as given (the device adds them in fp32)."""
    p, off = {}, 0
    for name, shape in layout(layers):
        n = int(np.prod(shape))
        a = np.asarray(flat[off:off + n], dtype=np.float32).reshape(shape)
        p[name] = bf16(a) if (round_weights and len(shape) == 2) else a.astype(np.float64)
        off += n
    return p


def pack(grads: Dict[str, np.ndarray], layers: int) -> np.ndarray:
    return np.concatenate([grads[name].ravel() for name, _ in layout(layers)])


class Graph:
    """Mesh tables the network reads (0-based, logical order)."""

    def __init__(self, cells_on_cell, n_edges_on_cell, edges_on_cell, cells_on_edge, cos_edge, sin_edge) -> None:
        self.coc = np.asarray(cells_on_cell)
        self.neoc = np.asarray(n_edges_on_cell)
        self.eoc = np.asarray(edges_on_cell)
        self.coe = np.asarray(cells_on_edge)
        self.cos = np.asarray(cos_edge, dtype=np.float64)
        self.sin = np.asarray(sin_edge, dtype=np.float64)
        self.C = self.neoc.size
        self.E = self.coe.shape[0]
        me = self.coc.shape[1]
        self.mask = np.arange(me)[None, :] < self.neoc[:, None]

    def mean_neighbours(self, H: np.ndarray) -> np.ndarray:
        idx = np.where(self.mask, self.coc, 0)
        return (H[idx] * self.mask[:, :, None]).sum(axis=1) / self.neoc[:, None]

    def mean_neighbours_adjoint(self, dA: np.ndarray) -> np.ndarray:
        """Adjoint of mean_neighbours: dH_j = sum over cells c having j as a
        neighbour of dA_c / n_c."""
        out = np.zeros_like(dA)
        w = dA / self.neoc[:, None]
        for s in range(self.coc.shape[1]):
            rows = np.flatnonzero(self.mask[:, s])
            np.add.at(out, self.coc[rows, s], w[rows])
        return out


def relu(x):
    return np.maximum(x, 0.0)


def forward(p: Dict[str, np.ndarray], X: np.ndarray, g: Graph, layers: int) -> Dict[str, object]:
    X = np.asarray(X, dtype=np.float64)
    P0 = X @ p["W_in"] + p["b_in"]
    H = [relu(P0)]
    A, Pl = [], []
    for l in range(layers):
        a = g.mean_neighbours(H[l])
        z = np.concatenate([H[l], a], axis=1)
        pre = z @ p[f"W_{l}"] + p[f"b_{l}"]
        A.append(a)
        Pl.append(pre)
        H.append(H[l] + relu(pre))
    Y = H[layers] @ p["W_out"] + p["b_out"]
    return {"X": X, "P0": P0, "H": H, "A": A, "P": Pl, "Y": Y}


def edge_prediction(Y: np.ndarray, g: Graph) -> np.ndarray:
    c0, c1 = g.coe[:, 0], g.coe[:, 1]
    ue = 0.5 * (Y[c0, 0:ROW] + Y[c1, 0:ROW])
    un = 0.5 * (Y[c0, ROW:2 * ROW] + Y[c1, ROW:2 * ROW])
    return g.cos[:, None] * ue + g.sin[:, None] * un


def loss(Y: np.ndarray, Tc: np.ndarray, Te: np.ndarray, g: Graph, L: int):
    """Loss parts {total, u, theta_m, rho_zz, w} and dLoss/dY."""
    C, E = g.C, g.E
    du = edge_prediction(Y, g)[:, :L] - Te[:, :L]
    dth = Y[:, 2 * ROW:2 * ROW + L] - Tc[:, 0, :L]
    drh = Y[:, 3 * ROW:3 * ROW + L] - Tc[:, 1, :L]
    dw = Y[:, 4 * ROW + 1:4 * ROW + L] - Tc[:, 2, 1:L]
    parts = {"u": (du ** 2).sum() / (E * L), "theta_m": (dth ** 2).sum() / (C * L),
             "rho_zz": (drh ** 2).sum() / (C * L), "w": (dw ** 2).sum() / (C * (L - 1))}
    parts["total"] = parts["u"] + parts["theta_m"] + parts["rho_zz"] + parts["w"]
    dY = np.zeros_like(Y)
    ge = 2.0 * du / (E * L)                     # dLoss/du_e
    c0, c1 = g.coe[:, 0], g.coe[:, 1]
    for cells in (c0, c1):
        np.add.at(dY[:, 0:L], cells, 0.5 * g.cos[:, None] * ge)
        np.add.at(dY[:, ROW:ROW + L], cells, 0.5 * g.sin[:, None] * ge)
    dY[:, 2 * ROW:2 * ROW + L] = 2.0 * dth / (C * L)
    dY[:, 3 * ROW:3 * ROW + L] = 2.0 * drh / (C * L)
    dY[:, 4 * ROW + 1:4 * ROW + L] = 2.0 * dw / (C * (L - 1))
    return parts, dY


def backward(p: Dict[str, np.ndarray], cache, dY: np.ndarray, g: Graph, layers: int) -> Dict[str, np.ndarray]:
    H, A, Pl = cache["H"], cache["A"], cache["P"]
    gr = {"W_out": H[layers].T @ dY, "b_out": dY.sum(axis=0)}
    dH = dY @ p["W_out"].T
    for l in range(layers - 1, -1, -1):
        dP = dH * (Pl[l] > 0)
        z = np.concatenate([H[l], A[l]], axis=1)
        gr[f"W_{l}"] = z.T @ dP
        gr[f"b_{l}"] = dP.sum(axis=0)
        dZ = dP @ p[f"W_{l}"].T
        dH = dH + dZ[:, :HID] + g.mean_neighbours_adjoint(dZ[:, HID:])
    dP0 = dH * (cache["P0"] > 0)
    gr["W_in"] = cache["X"].T @ dP0
    gr["b_in"] = dP0.sum(axis=0)
    return gr


def adam_step(params, grads, m, v, t: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam step (t counts from 1); returns (params, m, v)."""
    m = b1 * m + (1.0 - b1) * grads
    v = b2 * v + (1.0 - b2) * grads * grads
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def features(state: Dict[str, np.ndarray], g: Graph, f_edge: np.ndarray, L: int, moist: bool) -> np.ndarray:
    """Raw feature columns [C][384] of a logical-layout state: least-squares
    east/north cell wind from the edge normal winds, w, theta_m, rho_zz, qv,
    and the mean Coriolis parameter of the cell's edges in row L of the east
    channel."""
    C = g.C
    u = np.asarray(state["u"], dtype=np.float64)       # [L][E]
    X = np.zeros((C, 6, ROW))
    idx = np.where(g.mask, g.eoc, 0)
    cs = np.where(g.mask, g.cos[idx], 0.0)
    sn = np.where(g.mask, g.sin[idx], 0.0)
    m11, m12, m22 = (cs * cs).sum(1), (cs * sn).sum(1), (sn * sn).sum(1)
    det = m11 * m22 - m12 * m12
    ue = u[:, idx]                                      # [L][C][ME]
    b1 = (cs[None] * ue).sum(2)
    b2 = (sn[None] * ue).sum(2)
    X[:, 0, :L] = ((m22 * b1 - m12 * b2) / det).T
    X[:, 1, :L] = ((m11 * b2 - m12 * b1) / det).T
    X[:, 0, L] = np.where(g.mask, np.asarray(f_edge, dtype=np.float64)[idx], 0.0).sum(1) / g.neoc
    X[:, 2, :L + 1] = np.asarray(state["w"], dtype=np.float64).T
    X[:, 3, :L] = np.asarray(state["theta_m"], dtype=np.float64).T
    X[:, 4, :L] = np.asarray(state["rho_zz"], dtype=np.float64).T
    if moist:
        X[:, 5, :L] = np.asarray(state["scalars"][0], dtype=np.float64).T
    return X.reshape(C, N_IN)
