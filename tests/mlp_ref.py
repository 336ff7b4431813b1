"""CPU reference of the library's network (float64 numpy): tanh hidden layers, linear output,
normalised inputs/outputs, mean squared error, Adam.  Used to check the device kernels."""
import numpy as np


def unpack(p, layers):
    out, o = [], 0
    for (n_o, n_i) in layers:
        W = p[o:o + n_o * n_i].reshape(n_o, n_i)
        o += n_o * n_i
        b = p[o:o + n_o]
        o += n_o
        out.append((W, b))
    return out


def forward(p, layers, X, sc):
    in_shift, in_scale, out_shift, out_scale = sc
    H = [(X - in_shift) * in_scale]
    Ws = unpack(p, layers)
    for l, (W, b) in enumerate(Ws):
        Z = H[-1] @ W.T + b
        H.append(np.tanh(Z) if l + 1 < len(Ws) else Z)
    return H, H[-1] * out_scale + out_shift


def loss_and_grad(p, layers, X, T, sc):
    H, Y = forward(p, layers, X, sc)
    n = T.size
    d = Y - T
    loss = (d * d).sum() / n
    dZ = 2 * d * sc[3] / n
    Ws = unpack(p, layers)
    grads = []
    for l in range(len(Ws) - 1, -1, -1):
        W, b = Ws[l]
        grads.append((dZ.T @ H[l], dZ.sum(0)))
        if l > 0:
            dZ = (dZ @ W) * (1 - H[l] ** 2)
    g = []
    for gW, gb in reversed(grads):
        g += [gW.reshape(-1), gb]
    return loss, np.concatenate(g)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
