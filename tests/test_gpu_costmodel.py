"""The scheduler's cost model: the device network's forward and backward
kernels against the CPU oracle (oracle/mlp.py), training on the device,
and what the trained network gains over GauXC's flop heuristic on measured
per-task times."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def net(gx, gpu):
    from gauxc_amd.costmodel import MLPCostModel
    return MLPCostModel(0, hidden=128, layers=3, nout=2, seed=11)


def test_mlp_forward_and_gradients_match_oracle(net):
    import oracle.mlp as M
    widths, shapes, n = M.layout(6, 128, 3, 2)
    assert net.num_params() == n
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(300, 6)).astype(np.float32)
    t = rng.normal(size=(300, 2)).astype(np.float32)
    y = net.net_forward(x)
    acts, _ = M.forward(p, shapes, x.astype(np.float64))
    assert np.abs(y - acts[-1]).max() <= 1e-4 * max(1.0, np.abs(acts[-1]).max())
    loss, g = net.net_gradients(x, t)
    loss_ref, g_ref = M.loss_and_grad(p, shapes, x.astype(np.float64), t.astype(np.float64))
    _flush(f"loss {loss} ref {loss_ref}; max|dg| {np.abs(g - g_ref).max():.3e}")
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max() + 1e-7


def test_mlp_training_reduces_loss(gx, gpu):
    from gauxc_amd.costmodel import MLPCostModel
    m = MLPCostModel(0, hidden=64, layers=3, nout=2, seed=5)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, size=(4096, 6)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], x[:, 3] ** 2 - x[:, 4]], axis=1)
    t = t.astype(np.float32)
    l0, _ = m.net_gradients(x, t)
    m.net_train(x, t, epochs=60, batch=256, lr=2e-3, seed=2)
    l1, _ = m.net_gradients(x, t)
    _flush(f"loss {l0:.4f} -> {l1:.4f}")
    assert l1 < 0.05 * l0


def test_cost_model_gain_over_heuristic(gx, gpu):
    """Measured per-task times of an 8-water cluster: the MLP, trained on half
    the tasks, predicts the other half with a lower mean absolute error than
    GauXC's cost_exc_vxc heuristic fitted on the same half; driving the
    queued schedule with its predictions leaves the results unchanged."""
    from gauxc_amd.costmodel import HeuristicModel, MLPCostModel, task_features
    mol, bs, P, cfg = gx.workloads.make_config("c1", seed=2, n_water=8)
    lb = gx.workloads.build_tasks(mol, bs, grid="UltraFineGrid")
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    t = lb.get_tasks()
    ctx = gx.DeviceContext(0)
    ctx.upload_basis(bs.xcq_shells())
    ctx.upload_tasks(t)
    ctx.upload_density(P)
    ctx.enable_task_timing(True)
    reps = []
    for _ in range(4):
        ctx.exc_vxc("PBE", "mixed", True)
        reps.append(np.stack(ctx.task_times(), axis=1))
    times = np.median(np.stack(reps[1:]), axis=0)  # first run warms up
    e0, _, v0 = ctx.results()
    feats = task_features(t, bs.xcq_shells())
    nt = len(times)
    rng = np.random.default_rng(0)
    perm = rng.permutation(nt)
    tr, te = perm[: nt // 2], perm[nt // 2:]
    sub = lambda d, idx: {k: np.asarray(v)[idx] for k, v in d.items() if k in ("npts", "nbe")}
    heur = HeuristicModel().fit(sub(t, tr), times[tr])
    mlp = MLPCostModel(0, seed=4)
    mlp.fit(feats[tr], times[tr], epochs=300, batch=256, lr=1e-3)
    mae_h = np.abs(heur.predict(sub(t, te)) - times[te]).mean(axis=0)
    mae_m = np.abs(mlp.predict(feats[te]) - times[te]).mean(axis=0)
    _flush(f"{nt} tasks; MAE us (density, syr2k): heuristic {mae_h}, mlp {mae_m}; "
           f"mean time {times.mean(axis=0)}")
    assert mae_m.sum() < mae_h.sum()
    pred = mlp.predict(feats)
    ctx.enable_task_timing(False)
    ctx.set_schedule(1024, pred[:, 0], pred[:, 1])
    ctx.exc_vxc("PBE", "mixed", True)
    e1, _, v1 = ctx.results()
    assert abs(e1 - e0) <= 1e-9
    assert np.abs(v1 - v0).max() <= 1e-10
