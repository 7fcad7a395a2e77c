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


def loss_and_grads(params, dims, x, t, mask):
    x = np.asarray(x, np.float64)
    acts = forward(params, dims, x)
    y = acts[-1]
    m = mask > 0
    cnt = max(m.sum(), 1)
    d = np.where(m, y - t, 0.0)
    loss = (d * d).sum() / cnt
    dz = 2.0 * d / cnt
    layers = unpack(np.asarray(params, np.float64), dims)
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        grads.append((dz.T @ acts[i], dz.sum(0)))
        if i:
            dz = (dz @ W) * (acts[i] > 0)
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.ravel(), gb]
    return loss, np.concatenate(flat)
