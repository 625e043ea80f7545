"""FP64 numpy restatement of the alignment MLP -- TEST INFRASTRUCTURE ONLY.

Forward pass, loss and analytic gradients of the network the library trains
(pysofft_amd/csrc/mlp.hip): hidden layers z = x W^T + b, a = GELU(z) = z Phi(z); the last
layer linear; the 6 outputs (a, b) -> rotation R = [x y z] by Gram-Schmidt
(x = a/|a|, y = (b - (x.b) x)/|.|, z = x cross y); loss = mean ||R - T||_F^2.
Only tests load this module.  Its gradients are pinned by finite differences in
tests/test_oracle_pins.py.

``forward_bf16`` / ``gradients_bf16`` round to bfloat16 (nearest even) where the device
network stores bfloat16 operands (inputs, weights, hidden activations after GELU, and the
backpropagated dZ), everything else FP64, so that they differ from the device only by
its FP32 accumulation.
"""
import numpy as np
from scipy.special import erf


def unpack(params, dims):
    out, o = [], 0
    for i in range(len(dims) - 1):
        nw = dims[i] * dims[i + 1]
        W = params[o:o + nw].reshape(dims[i + 1], dims[i]).astype(np.float64)
        o += nw
        b = params[o:o + dims[i + 1]].astype(np.float64)
        o += dims[i + 1]
        out.append((W, b))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def gram_schmidt(o):
    a, b = o[:, :3], o[:, 3:]
    x = a / np.linalg.norm(a, axis=1, keepdims=True)
    bp = b - np.sum(x * b, axis=1, keepdims=True) * x
    y = bp / np.linalg.norm(bp, axis=1, keepdims=True)
    z = np.cross(x, y)
    return np.stack([x, y, z], axis=2)  # columns


def forward(params, dims, X):
    layers = unpack(params, dims)
    a, zs, acts = X.astype(np.float64), [], [X.astype(np.float64)]
    for i, (W, b) in enumerate(layers):
        z = a @ W.T + b
        zs.append(z)
        a = gelu(z) if i + 1 < len(layers) else z
        acts.append(a)
    return a, zs, acts


def loss(params, dims, X, T):
    out, _, _ = forward(params, dims, X)
    R = gram_schmidt(out)
    return np.mean(np.sum((R - T.reshape(-1, 3, 3)) ** 2, axis=(1, 2)))


def grad_out(out, T):
    """d loss / d outputs by the chain rule through Gram-Schmidt (FP64)."""
    n = out.shape[0]
    a, b = out[:, :3], out[:, 3:]
    na = np.linalg.norm(a, axis=1, keepdims=True)
    x = a / na
    xb = np.sum(x * b, axis=1, keepdims=True)
    bp = b - xb * x
    nb = np.linalg.norm(bp, axis=1, keepdims=True)
    y = bp / nb
    z = np.cross(x, y)
    R = np.stack([x, y, z], axis=2)
    gR = 2.0 * (R - T.reshape(-1, 3, 3)) / n
    gx, gy, gz = gR[:, :, 0].copy(), gR[:, :, 1].copy(), gR[:, :, 2]
    gx += np.cross(y, gz)
    gy += np.cross(gz, x)
    gbp = (gy - np.sum(gy * y, axis=1, keepdims=True) * y) / nb
    gbx = np.sum(gbp * x, axis=1, keepdims=True)
    gb = gbp - gbx * x
    gx += -gbx * b - xb * gbp
    ga = (gx - np.sum(gx * x, axis=1, keepdims=True) * x) / na
    return np.concatenate([ga, gb], axis=1)


def gradients(params, dims, X, T):
    """Flat gradient vector in the parameter layout [W, b] per layer."""
    layers = unpack(params, dims)
    out, zs, acts = forward(params, dims, X)
    g = grad_out(out, T)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        grads[i] = (g.T @ acts[i], g.sum(0))
        if i == 0:
            break
        g = (g @ W) * gelu_grad(zs[i - 1])
    return np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])


def bf16_round(a):
    """float values rounded to bfloat16 (round to nearest even), returned as float64"""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(params, dims, X):
    layers = unpack(params, dims)
    a = bf16_round(X)
    zs, acts = [], [a]
    for i, (W, b) in enumerate(layers):
        z = a @ bf16_round(W).T + b
        zs.append(z)
        a = bf16_round(gelu(z)) if i + 1 < len(layers) else z
        acts.append(a)
    return a, zs, acts


def gradients_bf16(params, dims, X, T):
    """gradients with the device's bfloat16 rounding points: dW = bf16(dZ)^T bf16(A),
    db = sum of the unrounded dZ, dA = bf16(dZ) bf16(W)"""
    layers = unpack(params, dims)
    out, zs, acts = forward_bf16(params, dims, X)
    g = grad_out(out, T)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        gb = bf16_round(g)
        grads[i] = (gb.T @ acts[i], g.sum(0))
        if i == 0:
            break
        g = (gb @ bf16_round(W)) * gelu_grad(zs[i - 1])
    return np.concatenate([np.concatenate([gw.ravel(), gbias]) for gw, gbias in grads])


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam update (bias-corrected), FP64: returns (w, m, v)"""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return w - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
