"""numpy oracle of the learned step-size model (TEST INFRASTRUCTURE ONLY).

The model of finite_root_generation_amd/csrc/frg_mlp.h: 8 standardised features -> 64 -> 64
(tanh) -> 1, loss = mean (y - t)^2 / 2 with t the standardised target.  Parameter block layout:
W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3, xmu[8], xsd[8], ymu, ysd.
Everything is computed in float64 here, so it checks the GPU's FP32 kernels.
"""
import numpy as np

NIN, NH = 8, 64
W1, B1 = 0, NH * NIN
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH
NW = B3 + 1
XMU, XSD = NW, NW + NIN
YMU, YSD = XSD + NIN, XSD + NIN + 1
SIZE = YSD + 1


def unpack(m):
    m = np.asarray(m, np.float64)
    return (m[W1:B1].reshape(NH, NIN), m[B1:W2], m[W2:B2].reshape(NH, NH), m[B2:W3], m[W3:B3], m[B3],
            m[XMU:XSD], m[XSD:YMU], m[YMU], m[YSD])


def forward(m, feats):
    w1, b1, w2, b2, w3, b3, xmu, xsd, ymu, ysd = unpack(m)
    x = (np.asarray(feats, np.float64) - xmu) / xsd
    h1 = np.tanh(x @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    y = h2 @ w3 + b3
    return y * ysd + ymu


def loss_and_grad(m, records):
    """Mean loss and its gradient with respect to the 4801 weights."""
    w1, b1, w2, b2, w3, b3, xmu, xsd, ymu, ysd = unpack(m)
    rec = np.asarray(records, np.float64)
    x = (rec[:, :NIN] - xmu) / xsd
    t = (rec[:, NIN] - ymu) / ysd
    n = rec.shape[0]
    h1 = np.tanh(x @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    y = h2 @ w3 + b3
    r = y - t
    loss = 0.5 * np.mean(r * r)
    dy = r / n
    g = np.zeros(NW)
    g[W3:B3] = h2.T @ dy
    g[B3] = dy.sum()
    d2 = np.outer(dy, w3) * (1 - h2 * h2)
    g[W2:B2] = (d2.T @ h1).ravel()
    g[B2:W3] = d2.sum(0)
    d1 = (d2 @ w2) * (1 - h1 * h1)
    g[W1:B1] = (d1.T @ x).ravel()
    g[B1:W2] = d1.sum(0)
    return loss, g


def random_model(rng, records=None):
    m = np.zeros(SIZE)
    m[W1:B1] = rng.uniform(-0.5, 0.5, NH * NIN)
    m[B1:W2] = rng.uniform(-0.1, 0.1, NH)
    m[W2:B2] = rng.uniform(-0.2, 0.2, NH * NH)
    m[B2:W3] = rng.uniform(-0.1, 0.1, NH)
    m[W3:B3] = rng.uniform(-0.2, 0.2, NH)
    m[B3] = 0.05
    if records is None:
        m[XMU:XSD], m[XSD:YMU], m[YMU], m[YSD] = 0.0, 1.0, 0.0, 1.0
    else:
        rec = np.asarray(records, np.float64)
        m[XMU:XSD] = rec[:, :NIN].mean(0)
        m[XSD:YMU] = np.maximum(rec[:, :NIN].std(0), 1e-3)
        m[YMU] = rec[:, NIN].mean()
        m[YSD] = max(rec[:, NIN].std(), 1e-3)
    return m
