"""CPU oracle of the scheduler's cost-model network -- TEST INFRASTRUCTURE ONLY.

Restates, in float64 numpy, the network the device trains
(gauxc_amd/csrc/device/mlp.hip): x -> [Linear -> ReLU] x L -> Linear, MSE
loss, parameters packed W1, b1, W2, b2, ... with W_l [out][in] row-major.
It has no counterpart in GauXC (GauXC orders tasks by the flop heuristic
XCTask::cost_exc_vxc, include/gauxc/xc_task.hpp:112-114); it is pinned by
central finite differences of its own loss (tests/test_oracle.py).
"""
import numpy as np


def layout(nin, nhidden, nlayers, nout):
    widths = [nin] + [nhidden] * nlayers + [nout]
    off, shapes = 0, []
    for l in range(1, len(widths)):
        shapes.append((off, widths[l], widths[l - 1]))
        off += widths[l] * widths[l - 1] + widths[l]
    return widths, shapes, off


def unpack(p, shapes):
    out = []
    for off, o, i in shapes:
        W = p[off:off + o * i].reshape(o, i)
        b = p[off + o * i:off + o * i + o]
        out.append((W, b))
    return out


def forward(p, shapes, x):
    a = np.asarray(x, dtype=np.float64)
    acts, pre = [a], []
    layers = unpack(np.asarray(p, dtype=np.float64), shapes)
    for k, (W, b) in enumerate(layers):
        z = a @ W.T + b
        pre.append(z)
        a = np.maximum(z, 0.0) if k < len(layers) - 1 else z
        acts.append(a)
    return acts, pre


def loss_and_grad(p, shapes, x, t):
    acts, pre = forward(p, shapes, x)
    y = acts[-1]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    g = np.zeros(len(p))
    dz = 2.0 * d / d.size
    layers = unpack(np.asarray(p, dtype=np.float64), shapes)
    for k in range(len(layers) - 1, -1, -1):
        off, o, i = shapes[k]
        if k < len(layers) - 1:
            dz = dz * (pre[k] > 0)
        g[off:off + o * i] = (dz.T @ acts[k]).ravel()
        g[off + o * i:off + o * i + o] = dz.sum(axis=0)
        if k > 0:
            dz = dz @ layers[k][0]
    return loss, g
