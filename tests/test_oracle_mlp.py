"""Pins for the MLP oracle: analytic gradient == central finite differences
(FP64), and the host initialisation matches the GPU model's."""
import numpy as np

from oracle import mlp_oracle as mo


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mo.init_params(1)
    X = rng.standard_normal((7, mo.NI))
    T = rng.standard_normal(7)
    g = mo.grad(p, X, T)
    idx = rng.choice(mo.NP, size=60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.loss(p + e, X, T) - mo.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_layout_and_init_agree_with_package():
    from tncontract_amd._mlp_init import he_init
    assert mo.NP == 6465
    np.testing.assert_allclose(he_init(3), mo.init_params(3).astype(np.float32))
