"""Learned task-cost model of the scheduler.

Per-task features (npts, screened nbe, number of shells, highest angular
momentum, bounding radius of the task's points, number of overlapping
shell pairs) feed a small FP32 MLP (3 hidden layers of 128, ReLU) that
predicts the task's density-stage and B^T Z-stage times.  The network is
trained on the device by the library's own kernels (xcq_mlp_*), on the
per-task wall-clock times the integrator records (xcq_enable_task_timing /
xcq_task_times).  Its predictions drive xcq_set_schedule: longest-first
bin-packing of work items into the persistent queues.

The fallback is GauXC's flop heuristic XCTask::cost_exc_vxc
(include/gauxc/xc_task.hpp:112-114), nbe (1 + nbe + n_deriv) npts, scaled by
a least-squares factor per stage so its units match the timings.
"""
import ctypes

import numpy as np

from . import _native as N

NFEAT = 6


def task_features(tasks, shells):
    """(ntasks, 6) raw features of a LoadBalancer.get_tasks() dict and the
    basis' xcq_shell array."""
    nsh = len(shells)
    org = np.array([[s.origin[k] for k in range(3)] for s in shells])
    rad = np.array([s.cutoff_radius for s in shells])
    lsh = np.array([s.l for s in shells])
    dist = np.linalg.norm(org[:, None, :] - org[None, :, :], axis=-1)
    overlap = dist < (rad[:, None] + rad[None, :])
    pts = np.asarray(tasks["points"]).reshape(-1, 3)
    nt = len(tasks["npts"])
    f = np.zeros((nt, NFEAT))
    for i in range(nt):
        s0, ns = int(tasks["shell_offset"][i]), int(tasks["nshells"][i])
        p0, npt = int(tasks["point_offset"][i]), int(tasks["npts"][i])
        sl = np.asarray(tasks["shell_list"][s0:s0 + ns], dtype=np.int64)
        P = pts[p0:p0 + npt]
        c = P.mean(axis=0)
        f[i, 0] = npt
        f[i, 1] = tasks["nbe"][i]
        f[i, 2] = ns
        f[i, 3] = lsh[sl].max() if ns else 0
        f[i, 4] = np.sqrt(((P - c) ** 2).sum(axis=1).max()) if npt else 0.0
        sub = overlap[np.ix_(sl, sl)]
        f[i, 5] = (sub.sum() + ns) / 2  # pairs i <= j
    assert nsh > 0
    return f


def heuristic_cost(tasks, n_deriv=3):
    """GauXC XCTask::cost_exc_vxc (xc_task.hpp:112-114)."""
    nbe = np.asarray(tasks["nbe"], dtype=np.float64)
    npts = np.asarray(tasks["npts"], dtype=np.float64)
    return nbe * (1 + nbe + n_deriv) * npts


class HeuristicModel:
    """cost_exc_vxc times one least-squares factor per stage."""

    def fit(self, tasks, times):
        h = heuristic_cost(tasks)
        self.scale = np.array([float(h @ times[:, k] / (h @ h)) for k in range(times.shape[1])])
        return self

    def predict(self, tasks):
        return heuristic_cost(tasks)[:, None] * self.scale[None, :]


class MLPCostModel:
    """Device-trained MLP on log-features -> log-times (two outputs)."""

    def __init__(self, device=0, hidden=128, layers=3, nout=2, seed=0):
        self.lib = N.device_lib()
        h = ctypes.c_void_p()
        N.xcq_check(self.lib.xcq_mlp_create(device, NFEAT, hidden, layers, nout, seed,
                                            ctypes.byref(h)), "xcq_mlp_create")
        self.h = h
        self.nout = nout
        self.shape = (NFEAT, hidden, layers, nout)

    def close(self):
        if self.h:
            self.lib.xcq_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # raw network access (tests, checkpoints)
    def num_params(self):
        n = ctypes.c_int64()
        N.xcq_check(self.lib.xcq_mlp_num_params(self.h, ctypes.byref(n)), "xcq_mlp_num_params")
        return n.value

    def get_params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        N.xcq_check(self.lib.xcq_mlp_get_params(self.h, p.ctypes.data), "xcq_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.xcq_check(self.lib.xcq_mlp_set_params(self.h, p.ctypes.data), "xcq_mlp_set_params")

    def net_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), self.nout), dtype=np.float32)
        N.xcq_check(self.lib.xcq_mlp_forward(self.h, len(x), x.ctypes.data, y.ctypes.data),
                    "xcq_mlp_forward")
        return y

    def net_gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = ctypes.c_float()
        N.xcq_check(self.lib.xcq_mlp_gradients(self.h, len(x), x.ctypes.data, t.ctypes.data,
                                               ctypes.byref(loss), g.ctypes.data),
                    "xcq_mlp_gradients")
        return loss.value, g

    def net_train(self, x, t, epochs=200, batch=512, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        N.xcq_check(self.lib.xcq_mlp_train(self.h, len(x), x.ctypes.data, t.ctypes.data, epochs,
                                           batch, lr, seed, ctypes.byref(loss)), "xcq_mlp_train")
        return loss.value

    # cost model on task features / times
    def _x(self, feats):
        z = np.log1p(np.asarray(feats, dtype=np.float64))
        return (z - self.mu) / self.sd

    def fit(self, feats, times, epochs=300, batch=512, lr=1e-3, seed=0):
        z = np.log1p(np.asarray(feats, dtype=np.float64))
        self.mu = z.mean(axis=0)
        self.sd = z.std(axis=0) + 1e-6
        y = np.log(np.maximum(np.asarray(times, dtype=np.float64), 1e-3))
        self.ymu = y.mean(axis=0)
        self.ysd = y.std(axis=0) + 1e-6
        return self.net_train(self._x(feats), (y - self.ymu) / self.ysd, epochs=epochs,
                              batch=min(batch, len(feats)), lr=lr, seed=seed)

    def predict(self, feats):
        y = self.net_forward(self._x(feats)).astype(np.float64)
        return np.exp(y * self.ysd + self.ymu)
