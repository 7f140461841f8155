"""CPU reference of the scheduler MLP (8 -> 32 -> 32 -> 4, ReLU), float64.

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward, backward and
training kernels.  The reference repository has no scheduler; this oracle is
pinned by central finite differences of its own loss (tests/test_mlp.py) and
by closed forms (zero hidden weights give the output bias; a one-hot input
through identity-like weights reproduces the input).

Parameter layout (1476 floats): W1[32][8], b1[32], W2[32][32], b2[32],
W3[4][32], b3[4], row-major.  Loss = sum_ij M_ij (y_ij - t_ij)^2 / sum M.
"""
import numpy as np

IN, H1, H2, OUT = 8, 32, 32, 4
SHAPES = [("W1", (H1, IN)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)),
          ("W3", (OUT, H2)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    w = np.asarray(w, np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in SHAPES])


def forward(w, X):
    p = unpack(w)
    X = np.asarray(X, np.float64)
    z1 = X @ p["W1"].T + p["b1"]
    a1 = np.maximum(z1, 0)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = np.maximum(z2, 0)
    return a2 @ p["W3"].T + p["b3"]


def loss_grad(w, X, T, M):
    p = unpack(w)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    M = np.asarray(M, np.float64)
    inv = 1.0 / M.sum()
    z1 = X @ p["W1"].T + p["b1"]
    a1 = np.maximum(z1, 0)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = np.maximum(z2, 0)
    y = a2 @ p["W3"].T + p["b3"]
    e = y - T
    loss = float((M * e * e).sum() * inv)
    d3 = 2.0 * M * e * inv
    d2 = (d3 @ p["W3"]) * (z2 > 0)
    d1 = (d2 @ p["W2"]) * (z1 > 0)
    g = {"W3": d3.T @ a2, "b3": d3.sum(0), "W2": d2.T @ a1, "b2": d2.sum(0),
         "W1": d1.T @ X, "b1": d1.sum(0)}
    return loss, pack(g)


def loss_only(w, X, T, M):
    y = forward(w, X)
    M = np.asarray(M, np.float64)
    return float((M * (y - T) ** 2).sum() / M.sum())


def init_weights(seed):
    """Kaiming-uniform weights, zero biases (seeded)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            bound = np.sqrt(6.0 / shape[1])
            p[name] = rng.uniform(-bound, bound, shape)
        else:
            p[name] = np.zeros(shape)
    return pack(p).astype(np.float32)


def features(F, EF, UV, UE, n, nnz, after_pull):
    """The scheduler inputs as the engine computes them on the device.

    F/EF: frontier vertices/edges, UV/UE: unvisited vertices/edges, n/nnz:
    graph vertices/edges, after_pull: the last step was bottom-up.
    """
    lg = lambda v: np.log2(1.0 + max(float(v), 0.0))
    return np.array([(lg(F) - lg(n)) / 32, (lg(EF) - lg(nnz)) / 32, (lg(UV) - lg(n)) / 32,
                     (lg(UE) - lg(nnz)) / 32, lg(EF / F if F > 0 else 0.0) / 16,
                     lg(nnz / n if n > 0 else 0.0) / 16, lg(n) / 32,
                     1.0 if after_pull else 0.0], np.float64)
