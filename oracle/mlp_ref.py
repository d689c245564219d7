"""FP64 restatement of the chemistry surrogate's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the device network's forward pass
and gradients (psphray_amd/csrc/psr_mlp.h, psr_mlp.hip).  The surrogate has
no counterpart in the reference; this file pins it by its definition:
22 -> 64 -> 64 -> 64 -> 4, SiLU, weights [in][out] row-major followed by the
bias, inputs normalised as (x - mean) / sigma; outputs are the changes of
logit x_HI, logit x_HeI, logit x_HeII and log10 T; loss = mean over samples
of the summed weighted squared errors (weights max(x (1 - x), 0.01) of the
result for the fractions, 1 for log10 T).  The backward pass is pinned by
central differences (tests/test_mlp.py), the features by their formula.
"""
import numpy as np

IN, HID, OUT = 22, 64, 4
SHAPES = [(IN, HID), (HID, HID), (HID, HID), (HID, OUT)]


def param_count():
    return sum(a * b + b for a, b in SHAPES)


def unpack(p):
    """[(W, b)] per layer from the flat parameter vector"""
    out, k = [], 0
    for a, b in SHAPES:
        W = p[k:k + a * b].reshape(a, b)
        k += a * b
        out.append((W, p[k:k + b]))
        k += b
    return out


def bf16(x):
    """round to bfloat16 (nearest even), returned as float64"""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, xn, emulate_bf16=False):
    """network outputs (the predicted changes) for normalised inputs xn[n][22]"""
    q = (lambda v: bf16(v)) if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    a = q(xn)
    layers = unpack(np.asarray(p, np.float64))
    for k, (W, b) in enumerate(layers):
        z = a @ q(W) + b
        a = q(silu(z)) if k < 3 else z
    return a


def logit(x):
    x = np.clip(np.asarray(x, np.float64), 1e-9, 1 - 1e-9)
    return np.log(x / (1 - x))


def target(f, y):
    """network targets t[n][4] and loss weights w[n][4] of raw feature rows
    f[n][22] and step results y[n][4]"""
    f = np.asarray(f, np.float64)
    y = np.asarray(y, np.float64)
    t = np.empty((len(f), 4))
    w = np.ones((len(f), 4))
    for j in range(3):
        t[:, j] = logit(y[:, j]) - logit(f[:, 1 + j])
        w[:, j] = np.maximum(y[:, j] * (1 - y[:, j]), 0.01)
    t[:, 3] = y[:, 3] - f[:, 4]
    return t, w


def predict(f, out):
    """step results from raw features and network outputs"""
    f = np.asarray(f, np.float64)
    y = np.empty((len(f), 4))
    for j in range(3):
        y[:, j] = 1.0 / (1.0 + np.exp(-(logit(f[:, 1 + j]) + out[:, j])))
    y[:, 3] = f[:, 4] + out[:, 3]
    return y


def loss_grad(p, xn, t, w=None):
    """loss and its gradient (FP64) for normalised inputs xn, targets t[n][4]
    and weights w[n][4] (default 1)"""
    layers = unpack(np.asarray(p, np.float64))
    n = len(xn)
    acts, zs = [np.asarray(xn, np.float64)], []
    for k, (W, b) in enumerate(layers):
        z = acts[-1] @ W + b
        zs.append(z)
        acts.append(silu(z) if k < 3 else z)
    w2 = np.ones_like(t) if w is None else np.asarray(w, np.float64) ** 2
    d = acts[-1] - t
    loss = float((w2 * d * d).sum() / n)
    delta = 2.0 * w2 * d / n
    grads = []
    for k in range(3, -1, -1):
        W, _ = layers[k]
        grads.append((acts[k].T @ delta, delta.sum(0)))
        if k:
            delta = (delta @ W.T) * silu_grad(zs[k - 1])
    grads.reverse()
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def features(lH, lHeI, lHeII, T, nH, gHI, gHeI, gHeII, heat, yeMET, time):
    """the 22 inputs of steps (arrays), by their definition"""
    n = np.broadcast(lH, T).size
    f = np.zeros((n, IN))
    lg = lambda v: np.log10(np.maximum(v, 1e-30))
    f[:, 0] = lg(nH)
    f[:, 1] = lH
    f[:, 2] = lHeI
    f[:, 3] = lHeII
    f[:, 4] = lg(T)
    f[:, 5] = lg(np.asarray(gHI) + 1e-12)
    f[:, 6] = lg(np.asarray(gHeI) + 1e-12)
    f[:, 7] = lg(np.asarray(gHeII) + 1e-12)
    f[:, 8] = lg(np.asarray(heat) + 1e-30)
    f[:, 9] = yeMET
    lf = lambda v: np.log(np.clip(v, 1e-6, 1 - 1e-6) / (1 - np.clip(v, 1e-6, 1 - 1e-6)))
    f[:, 10] = lf(np.asarray(lH, np.float64))
    f[:, 11] = lf(np.asarray(lHeI, np.float64))
    f[:, 12] = lf(np.asarray(lHeII, np.float64))
    f[:, 21] = lg(time)
    return f
