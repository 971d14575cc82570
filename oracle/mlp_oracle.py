"""FP64 numpy restatement of the smoother-weight MLP -- TEST INFRASTRUCTURE
ONLY (the checker for the HIP forward, backward and Adam kernels).

Network: in -> H (tanh) -> H (tanh) -> out (linear); loss
0.5 / B * sum (y - t)^2; flat parameter layout W1 [H x in], b1, W2 [H x H],
b2, W3 [out x H], b3. Adam as Kingma & Ba (2015), Algorithm 1.
"""
import numpy as np


def unpack(p, nin, h, nout):
    o = 0
    W1 = p[o:o + h * nin].reshape(h, nin); o += h * nin
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    W3 = p[o:o + nout * h].reshape(nout, h); o += nout * h
    b3 = p[o:o + nout]; o += nout
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, X, nin, h, nout):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, h, nout)
    H1 = np.tanh(X @ W1.T + b1)
    H2 = np.tanh(H1 @ W2.T + b2)
    return H2 @ W3.T + b3, H1, H2


def loss(p, X, T, nin, h, nout):
    Y, _, _ = forward(p, X, nin, h, nout)
    return 0.5 * np.sum((Y - T) ** 2) / X.shape[0]


def grad(p, X, T, nin, h, nout):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, h, nout)
    Y, H1, H2 = forward(p, X, nin, h, nout)
    B = X.shape[0]
    dY = (Y - T) / B
    gW3 = dY.T @ H2
    gb3 = dY.sum(0)
    d2 = (dY @ W3) * (1 - H2 ** 2)
    gW2 = d2.T @ H1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (1 - H1 ** 2)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def fd_grad(p, X, T, nin, h, nout, idx, eps=1e-6):
    """Central finite differences of the loss at the given parameter indices."""
    out = []
    for i in idx:
        e = np.zeros_like(p)
        e[i] = eps
        out.append((loss(p + e, X, T, nin, h, nout) - loss(p - e, X, T, nin, h, nout)) / (2 * eps))
    return np.array(out)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
