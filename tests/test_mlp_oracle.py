"""Pins of the FP64 oracle of the learned Krylov controller's network
(7 -> 64 -> 64 -> 2, exact-erf GELU): the analytic back-propagated gradient
against central finite differences, and the forward pass against
x Phi(x) from scipy for a network wired to one path."""
import numpy as np
import scipy.stats as st

import pyoracle as po

NI, NH, NO = 7, 64, 2
OW1, OB1 = 0, NH * NI
OW2, OB2 = OB1 + NH, OB1 + NH + NH * NH
OW3, OB3 = OB2 + NH, OB2 + NH + NO * NH


def _random_params(rng):
    P = np.zeros(po.MLP_NPARAM)
    P[OW1:OB1] = rng.uniform(-1, 1, NH * NI) * np.sqrt(6.0 / (NI + NH))
    P[OB1:OW2] = 0.1 * rng.standard_normal(NH)
    P[OW2:OB2] = rng.uniform(-1, 1, NH * NH) * np.sqrt(6.0 / (2 * NH))
    P[OB2:OW3] = 0.1 * rng.standard_normal(NH)
    P[OW3:OB3] = rng.uniform(-1, 1, NO * NH) * np.sqrt(6.0 / (NO + NH))
    P[OB3:] = 0.1 * rng.standard_normal(NO)
    return P


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = _random_params(rng)
    X = rng.standard_normal((6, NI))
    T = rng.standard_normal((6, NO))
    loss, G = po.mlp_loss_grad(P, X, T)
    Y = po.mlp_forward(P, X)
    assert abs(loss - 0.5 * np.mean(np.sum((Y - T) ** 2, axis=1))) < 1e-15 * max(1.0, loss)
    h = 1e-5
    fd = np.zeros_like(G)
    for p in range(P.size):
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        fd[p] = (po.mlp_loss_grad(Pp, X, T)[0] - po.mlp_loss_grad(Pm, X, T)[0]) / (2 * h)
    assert np.abs(G - fd).max() < 1e-8 * max(1.0, np.abs(G).max())
    # every block of parameters receives gradient
    for a, b in ((OW1, OB1), (OB1, OW2), (OW2, OB2), (OB2, OW3), (OW3, OB3), (OB3, po.MLP_NPARAM)):
        assert np.abs(G[a:b]).max() > 0


def test_forward_single_path_is_gelu_of_gelu():
    """W1[0][2] = 1, W2[0][0] = 1, W3[1][0] = 1, b3 = (0.25, -0.5): y1 = gelu(gelu(x_2)) - 0.5
    with gelu(x) = x Phi(x), Phi the standard normal CDF; y0 = 0.25."""
    P = np.zeros(po.MLP_NPARAM)
    P[OW1 + 0 * NI + 2] = 1.0
    P[OW2 + 0 * NH + 0] = 1.0
    P[OW3 + 1 * NH + 0] = 1.0
    P[OB3:] = (0.25, -0.5)
    X = np.zeros((5, NI))
    X[:, 2] = (-3.0, -0.7, 0.0, 0.4, 2.5)
    g = lambda x: x * st.norm.cdf(x)
    Y = po.mlp_forward(P, X)
    assert np.allclose(Y[:, 0], 0.25, rtol=0, atol=1e-15)
    assert np.allclose(Y[:, 1], g(g(X[:, 2])) - 0.5, rtol=0, atol=1e-15)
