"""TEST INFRASTRUCTURE ONLY: numpy restatement of the density-guess network's forward
pass, masked-MSE loss and backward pass, used to check the GPU kernels' gradients.

The network (no reference counterpart) is dims[0] -> SiLU hidden layers -> linear
output; parameters per layer: W [in x out] row-major then b [out].  With bf16=True the
rounding points of the GPU path are reproduced: GEMM operands (inputs, activations,
weights, back-propagated errors) are rounded to bfloat16 (round to nearest even) and
every product accumulates in float32; with bf16=False everything is float64, which is
what the finite-difference check of the backward formulas uses.
"""
import numpy as np


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), kept as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params, dims):
    out, o = [], 0
    for l in range(len(dims) - 1):
        nw = dims[l] * dims[l + 1]
        W = params[o:o + nw].reshape(dims[l], dims[l + 1])
        o += nw
        b = params[o:o + dims[l + 1]]
        o += dims[l + 1]
        out.append((W, b))
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, dims, X, bf16=True):
    dt = np.float32 if bf16 else np.float64
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=dt), dims)
    h = rnd(X)
    hs, zs = [h], []
    for l, (W, b) in enumerate(layers):
        z = (h.astype(dt) @ rnd(W).astype(dt)).astype(dt) + b.astype(dt)
        if l < len(layers) - 1:
            zs.append(z)
            h = rnd(_silu(z))
            hs.append(h)
        else:
            return z, hs, zs


def loss_grad(params, dims, X, T, M, bf16=True):
    """Masked mean squared error sum M (Y - T)^2 / sum M and its parameter gradient."""
    dt = np.float32 if bf16 else np.float64
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, hs, zs = forward(params, dims, X, bf16)
    T = np.asarray(T, dtype=dt)
    M = np.asarray(M, dtype=dt)
    norm = M.sum(dtype=np.float64)
    d = M * (y - T)
    loss = float((d * (y - T)).sum(dtype=np.float64) / norm)
    dz = (2.0 * d / norm).astype(dt)
    layers = unpack(np.asarray(params, dtype=dt), dims)
    grads = [None] * len(layers)
    dzb = rnd(dz)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        gW = (hs[l].astype(dt).T @ dzb.astype(dt)).astype(dt)
        gb = dz.sum(axis=0, dtype=dt)
        grads[l] = (gW, gb)
        if l > 0:
            dh = (dzb.astype(dt) @ rnd(W).astype(dt).T).astype(dt)
            dz = (dh * _silu_grad(zs[l - 1])).astype(dt)
            dzb = rnd(dz)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
