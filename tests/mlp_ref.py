"""float64 numpy restatement of the selector network (test reference).

Same parameter layout as the HIP implementation: per layer W (out x in,
row-major) then b; ReLU after hidden layers; masked mean squared error."""
import numpy as np


def unpack(params, dims):
    out, o = [], 0
    for fi, fo in zip(dims[:-1], dims[1:]):
        W = params[o:o + fi * fo].reshape(fo, fi)
        o += fi * fo
        b = params[o:o + fo]
        o += fo
        out.append((W, b))
    assert o == len(params)
    return out


def forward(params, dims, x):
    acts = [x]
    layers = unpack(np.asarray(params, np.float64), dims)
    for i, (W, b) in enumerate(layers):
        z = acts[-1] @ W.T + b
        acts.append(np.maximum(z, 0.0) if i < len(layers) - 1 else z)
    return acts


def near_kinks(params, dims, x, rel=1e-4):
    """Rows with a hidden pre-activation within rel of its condition scale
    (|A| |W|^T) of zero: there fp32 and float64 may disagree on the ReLU."""
    layers = unpack(np.asarray(params, np.float64), dims)
    a = np.asarray(x, np.float64)
    bad = np.zeros(len(a), bool)
    for i, (W, b) in enumerate(layers[:-1]):
        z = a @ W.T + b
        zs = np.abs(a) @ np.abs(W).T + np.abs(b)
        bad |= (np.abs(z) < rel * zs).any(axis=1)
        a = np.maximum(z, 0.0)
    return bad


def loss_and_grads(params, dims, x, t, mask, scale=False):
    """Loss and flat gradient; with scale=True also the gradient computed with
    every term replaced by its absolute value (the condition scale against
    which floating-point accumulation error is measured)."""
    x = np.asarray(x, np.float64)
    acts = forward(params, dims, x)
    y = acts[-1]
    m = mask > 0
    cnt = max(m.sum(), 1)
    d = np.where(m, y - t, 0.0)
    loss = (d * d).sum() / cnt
    dz = 2.0 * d / cnt
    layers = unpack(np.asarray(params, np.float64), dims)
    grads, scales = [], []
    adz = np.abs(dz)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        grads.append((dz.T @ acts[i], dz.sum(0)))
        scales.append((adz.T @ np.abs(acts[i]), adz.sum(0)))
        if i:
            dz = (dz @ W) * (acts[i] > 0)
            adz = (adz @ np.abs(W)) * (acts[i] > 0)
    flat, sflat = [], []
    for (gW, gb), (sW, sb) in zip(reversed(grads), reversed(scales)):
        flat += [gW.ravel(), gb]
        sflat += [sW.ravel(), sb]
    if scale:
        return loss, np.concatenate(flat), np.concatenate(sflat)
    return loss, np.concatenate(flat)
