"""CPU oracle of the path selector's network (test infrastructure only).

The reference has no learned component (its recursion follows the fixed cost
model of strassen.hpp:510-679), so this restates the network the library
defines: y = W3 tanh(W2 tanh(W1 x' + b1) + b2) + b3 with x' = (x - mu) / sd,
F -> 32 -> 32 -> O, parameters flattened as [W1 (32xF, row-major) | b1 | W2 |
b2 | W3 (Ox32) | b3], loss L = 1/(2N) sum_n sum_o m_no (y_no - t_no)^2 and
Adam (Kingma & Ba 2015, beta1 0.9, beta2 0.999, eps 1e-8).  The analytic
gradient here is pinned by central finite differences of the loss in the
tests; the device kernels are checked against it.
"""
import numpy as np

H = 32


def sizes(F, O):
    return [("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (O, H)), ("b3", (O,))]


def n_params(F, O):
    return sum(int(np.prod(s)) for _, s in sizes(F, O))


def unpack(prm, F, O):
    out, off = {}, 0
    for name, shp in sizes(F, O):
        n = int(np.prod(shp))
        out[name] = prm[off:off + n].reshape(shp)
        off += n
    return out


def init(F, O, seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in sizes(F, O):
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            parts.append(rng.uniform(-lim, lim, shp).ravel())
        else:
            parts.append(np.zeros(shp).ravel())
    return np.concatenate(parts)


def forward(prm, X, mu, sd, O, keep=False):
    F = X.shape[1]
    p = unpack(prm, F, O)
    xn = (X - mu) / sd
    h1 = np.tanh(xn @ p["W1"].T + p["b1"])
    h2 = np.tanh(h1 @ p["W2"].T + p["b2"])
    y = h2 @ p["W3"].T + p["b3"]
    return (y, xn, h1, h2) if keep else y


def loss(prm, X, T, M, mu, sd):
    y = forward(prm, X, mu, sd, T.shape[1])
    return 0.5 * np.sum(M * (y - T) ** 2) / X.shape[0]


def grad(prm, X, T, M, mu, sd):
    n, F = X.shape
    O = T.shape[1]
    p = unpack(prm, F, O)
    y, xn, h1, h2 = forward(prm, X, mu, sd, O, keep=True)
    dy = M * (y - T) / n
    d2 = (dy @ p["W3"]) * (1.0 - h2 ** 2)
    d1 = (d2 @ p["W2"]) * (1.0 - h1 ** 2)
    g = {"W3": dy.T @ h2, "b3": dy.sum(0), "W2": d2.T @ h1, "b2": d2.sum(0), "W1": d1.T @ xn, "b1": d1.sum(0)}
    return np.concatenate([g[name].ravel() for name, _ in sizes(F, O)])


def fd_grad(prm, X, T, M, mu, sd, idx, h=1e-6):
    """Central finite differences of the loss at the given parameter indices."""
    out = []
    for i in idx:
        e = np.zeros_like(prm)
        e[i] = h
        out.append((loss(prm + e, X, T, M, mu, sd) - loss(prm - e, X, T, M, mu, sd)) / (2 * h))
    return np.array(out)


def adam(prm, g, m1, m2, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return prm - lr * mh / (np.sqrt(vh) + eps), m1, m2
