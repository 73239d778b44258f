"""CPU restatement of the learned warm start's network (test infrastructure, not the product).

The product's network runs in endobeams.jl_amd/csrc/eb_mlp.hip (FP32 HIP kernels). This module
states the same function in numpy FP64 so that the device forward pass and back-propagation can be
checked: a tanh MLP whose parameters are, layer after layer, W (out×in, row-major) then b (out);
loss = mean over rows×outputs of (f(X) - T)². The gradient here is the textbook reverse-mode
chain rule; tests/test_oracle_mlp.py pins it against central finite differences.
The reference (EndoBeams.jl) has no learned component, so there is no reference file to cite.
"""
import numpy as np


def split(params, widths):
    """[(W, b)] views of a flat parameter vector."""
    out, o = [], 0
    for i, j in zip(widths[:-1], widths[1:]):
        W = params[o:o + i * j].reshape(j, i)
        o += i * j
        b = params[o:o + j]
        o += j
        out.append((W, b))
    assert o == params.size
    return out


def n_params(widths):
    return sum(i * j + j for i, j in zip(widths[:-1], widths[1:]))


def forward(params, widths, X):
    """Activations of every layer (the last is the linear output)."""
    acts = []
    h = np.asarray(X, dtype=np.float64)
    layers = split(np.asarray(params, dtype=np.float64), widths)
    for l, (W, b) in enumerate(layers):
        z = h @ W.T + b
        h = np.tanh(z) if l + 1 < len(layers) else z
        acts.append(h)
    return acts


def loss_and_gradient(params, widths, X, T):
    params = np.asarray(params, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    layers = split(params, widths)
    acts = forward(params, widths, X)
    Y = acts[-1]
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    grads = []
    dz = 2.0 * (Y - T) / n
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        hin = acts[l - 1] if l > 0 else X
        grads.append((dz.T @ hin, dz.sum(axis=0)))
        if l > 0:
            dz = (dz @ W) * (1.0 - hin ** 2)
    flat = []
    for gW, gb in reversed(grads):
        flat.append(gW.reshape(-1))
        flat.append(gb)
    return loss, np.concatenate(flat)
