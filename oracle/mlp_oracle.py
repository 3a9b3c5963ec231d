"""NumPy FP64 oracle of the rank-prediction MLP (34 -> 64 -> 64 -> 1, ReLU).

TEST INFRASTRUCTURE ONLY (same import rule as oracle/tn_oracle.py).  The
reference has no learned component; this oracle defines the network the GPU
kernels (tncontract_amd/csrc/tn_mlp.hip) must reproduce, and is itself pinned
by central finite differences in tests/test_oracle_mlp.py.

Parameter vector layout (6465 values): W1[34][64], b1[64], W2[64][64], b2[64],
w3[64], b3.
"""
import numpy as np

NI, NH = 34, 64
SIZES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)),
         ("b3", (1,))]
NP = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[off:off + n].reshape(shape)
        off += n
    return out


def forward(p, X):
    q = unpack(p)
    z1 = X @ q["W1"] + q["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ q["W2"] + q["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ q["w3"] + q["b3"][0]
    return y, (h1, h2)


def loss(p, X, T):
    y, _ = forward(p, X)
    return float(np.mean((y - T) ** 2))


def grad(p, X, T):
    """Analytic gradient of mean squared error."""
    q = unpack(p)
    y, (h1, h2) = forward(p, X)
    B = X.shape[0]
    dy = 2.0 * (y - T) / B
    g = {}
    g["w3"] = h2.T @ dy
    g["b3"] = np.array([dy.sum()])
    dh2 = np.outer(dy, q["w3"]) * (h2 > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ q["W2"].T) * (h1 > 0)
    g["W1"] = X.T @ dh1
    g["b1"] = dh1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in SIZES])


def init_params(seed):
    """He-normal weights, zero biases (the same host init the GPU model uploads)."""
    rng = np.random.default_rng(seed)
    q = {"W1": rng.standard_normal((NI, NH)) * np.sqrt(2.0 / NI),
         "b1": np.zeros(NH),
         "W2": rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH),
         "b2": np.zeros(NH),
         "w3": rng.standard_normal(NH) * np.sqrt(1.0 / NH),
         "b3": np.zeros(1)}
    return np.concatenate([q[name].ravel() for name, _ in SIZES])
