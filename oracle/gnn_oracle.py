"""CPU restatement (numpy) of the edge-message graph network of the learned DEM surrogate.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the HIP forward and backward kernels.

The network has no counterpart in the reference (exaDEM has no learned part); this file defines it.
Two modes:
  exact=True   float64 everywhere, no rounding: its backward is pinned by central finite differences
               of its forward (tests/test_gnn_oracle.py);
  exact=False  float32 arithmetic with the values rounded to bfloat16 where the device stores them
               (weights, edge features, node inputs, hidden states, P|Q, messages, update
               pre-activations, dU, dP|dQ, dM), the checker of the device kernels.

Network (H = 128 hidden, L = 3 message-passing layers, FN = 8 node inputs, FE = 16 edge inputs,
FO = 6 outputs):
  h0      = relu(x W_in^T + b_in)
  layer l: [P | Q] = h_l W_AB^T                       (W_AB: 2H x H; P = first H columns)
           z_ij    = P_i + Q_j + e_ij W_E^T + b_E      (one z per directed neighbour entry i -> j)
           m_i     = s * sum_j relu(z_ij)              (s = 1/8)
           U       = [h_l | m] W_U^T + b_U             (W_U: H x 2H)
           h_{l+1} = h_l + relu(U)
  out     = h_L W_out^T + b_out
  loss    = mean over nodes and outputs of (out - target)^2
"""
import numpy as np

H, L, FN, FE, FO = 128, 3, 8, 16, 6
MSG_SCALE = 0.125


def param_layout():
    """(name, shape) in the order of the flat parameter vector shared with the device."""
    out = [("W_in", (H, FN)), ("b_in", (H,))]
    for l in range(L):
        out += [(f"W_AB{l}", (2 * H, H)), (f"W_E{l}", (H, FE)), (f"b_E{l}", (H,)), (f"W_U{l}", (H, 2 * H)),
                (f"b_U{l}", (H,))]
    out += [("W_out", (FO, H)), ("b_out", (FO,))]
    return out


def n_params():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unflatten(flat):
    p, o = {}, 0
    for name, shape in param_layout():
        k = int(np.prod(shape))
        p[name] = np.asarray(flat[o:o + k]).reshape(shape)
        o += k
    return p


def flatten(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in param_layout()])


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.reshape(np.shape(x))


class Net:
    """Forward and backward of the network over a fixed directed edge list.

    x: (N, FN) node inputs; src, dst: (E,) edge i -> j; e: (E, FE) edge inputs; mirror: (E,) index of
    the reverse edge j -> i (used only to state the device's dQ order, not needed for the math)."""

    def __init__(self, flat_params, exact=True):
        self.exact = exact
        self.dt = np.float64 if exact else np.float32
        p = unflatten(np.asarray(flat_params, np.float64))
        self.p = {k: self.r(v.astype(self.dt)) for k, v in p.items()}

    def r(self, a):
        return a if self.exact else bf16(a.astype(np.float32))

    def forward(self, x, src, dst, e):
        p, r, dt = self.p, self.r, self.dt
        n = x.shape[0]
        x = r(np.asarray(x, dt))
        e = r(np.asarray(e, dt))
        c = {"x": x, "e": e, "src": src, "dst": dst, "n": n}
        h = r(np.maximum(x @ p["W_in"].T + p["b_in"], 0))
        c["h0"] = h
        for l in range(L):
            pq = r(h @ p[f"W_AB{l}"].T)
            z = pq[src, :H] + pq[dst, H:] + (e @ p[f"W_E{l}"].T).astype(dt) + p[f"b_E{l}"]
            m = np.zeros((n, H), dt)
            np.add.at(m, src, np.maximum(z, 0))
            m = r(m * dt(MSG_SCALE))
            u = r(np.concatenate([h, m], axis=1) @ p[f"W_U{l}"].T + p[f"b_U{l}"])
            hn = r(h + np.maximum(u, 0))
            c[f"pq{l}"], c[f"m{l}"], c[f"u{l}"], c[f"h{l + 1}"] = pq, m, u, hn
            h = hn
        out = h @ p["W_out"].T + p["b_out"]
        c["out"] = out
        return out, c

    def loss(self, out, target):
        d = out - target
        return float(np.mean(d * d))

    def backward(self, c, target):
        """Gradients of the loss with respect to every parameter (dict), in the device's rounding."""
        p, r, dt = self.p, self.r, self.dt
        n, src, dst, e = c["n"], c["src"], c["dst"], c["e"]
        g = {}
        dout = (2.0 / (n * FO)) * (c["out"] - target)
        dout = dout.astype(dt)
        hL = c[f"h{L}"]
        g["W_out"] = dout.T @ hL
        g["b_out"] = dout.sum(axis=0)
        dh = dout @ p["W_out"]
        for l in reversed(range(L)):
            h = c[f"h{l}"]
            u, m, pq = c[f"u{l}"], c[f"m{l}"], c[f"pq{l}"]
            du = r(dh * (u > 0))
            X = np.concatenate([h, m], axis=1)
            g[f"W_U{l}"] = du.T @ X
            g[f"b_U{l}"] = du.sum(axis=0)
            dX = du @ p[f"W_U{l}"]
            dh = dh + dX[:, :H]
            dm = dX[:, H:] * dt(MSG_SCALE)
            z = pq[src, :H] + pq[dst, H:] + (e @ p[f"W_E{l}"].T).astype(dt) + p[f"b_E{l}"]
            gz = dm[src] * (z > 0)
            dp = np.zeros((n, H), dt)
            dq = np.zeros((n, H), dt)
            np.add.at(dp, src, gz)
            np.add.at(dq, dst, gz)
            g[f"W_E{l}"] = gz.T @ e
            g[f"b_E{l}"] = gz.sum(axis=0)
            dpq = r(np.concatenate([dp, dq], axis=1))
            g[f"W_AB{l}"] = dpq.T @ h
            dh = dh + dpq @ p[f"W_AB{l}"]
        dpre = dh * (c["h0"] > 0)
        g["W_in"] = dpre.T @ c["x"]
        g["b_in"] = dpre.sum(axis=0)
        return g


def init_params(seed):
    """He-normal weights (std sqrt(2 / fan_in)), zero biases; the output layer scaled by 0.1."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_layout():
        if name.startswith("b_"):
            p[name] = np.zeros(shape)
        else:
            std = np.sqrt(2.0 / shape[1])
            if name == "W_out":
                std *= 0.1
            p[name] = rng.normal(0.0, std, shape)
    return flatten(p)
