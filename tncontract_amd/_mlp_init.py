"""Host-side initial weights of the rank MLP (He-normal weights, zero
biases), in the parameter layout of tn_mlp_* (W1[34][64], b1, W2[64][64],
b2, w3[64], b3)."""
import numpy as np


def he_init(seed):
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((34, 64)) * np.sqrt(2.0 / 34), np.zeros(64),
             rng.standard_normal((64, 64)) * np.sqrt(2.0 / 64), np.zeros(64),
             rng.standard_normal(64) * np.sqrt(1.0 / 64), np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)
