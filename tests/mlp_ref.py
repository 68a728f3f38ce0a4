"""Float64 CPU reference of the sweep's warm-start network, the check for the device kernels (amp_mlp_*): forward and
backward of nIn -> H (tanh) -> H (tanh) -> nOut under the mean squared error (1 / (B nOut)) sum (y - t)^2, and one Adam
step, on the flat parameter layout W1 (H x nIn), b1, W2 (H x H), b2, W3 (nOut x H), b3."""
import numpy as np


def sizes(nIn, H, nOut):
    return [("W1", (H, nIn)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (nOut, H)), ("b3", (nOut,))]


def unpack(theta, nIn, H, nOut):
    out, o = {}, 0
    for name, shp in sizes(nIn, H, nOut):
        k = int(np.prod(shp))
        out[name] = theta[o:o + k].reshape(shp)
        o += k
    assert o == len(theta)
    return out


def forward(theta, X, nIn, H, nOut):
    p = unpack(np.asarray(theta, np.float64), nIn, H, nOut)
    X = np.asarray(X, np.float64)
    h1 = np.tanh(X @ p["W1"].T + p["b1"])
    h2 = np.tanh(h1 @ p["W2"].T + p["b2"])
    return h1, h2, h2 @ p["W3"].T + p["b3"]


def loss_and_grad(theta, X, T, nIn, H, nOut):
    p = unpack(np.asarray(theta, np.float64), nIn, H, nOut)
    X = np.asarray(X, np.float64)
    h1, h2, y = forward(theta, X, nIn, H, nOut)
    d = y - np.asarray(T, np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dz2 = (dy @ p["W3"]) * (1 - h2 * h2)
    g["W2"], g["b2"] = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ p["W2"]) * (1 - h1 * h1)
    g["W1"], g["b1"] = dz1.T @ X, dz1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in sizes(nIn, H, nOut)])


def adam(theta, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (step counts from 1); returns (theta, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    theta = theta - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return theta, m, v
