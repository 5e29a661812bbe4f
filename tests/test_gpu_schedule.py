"""Scheduling machinery of the device path: tile-width variants, persistent
queued launches, graph capture/replay, batched molecules and per-task
timing.  Each must leave the results unchanged (to FP64 summation-order
noise) against the default one-workgroup-per-item launches, which the
parity tests pin to the oracle and GauXC's fixtures."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _tasks(gx, n_water, grid, seed=0, cfgname="c1"):
    mol, bs, P, cfg = gx.workloads.make_config(cfgname, seed=seed, n_water=n_water)
    lb = gx.workloads.build_tasks(mol, bs, grid=grid)
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    return bs, P, lb.get_tasks()


def _ctx(gx, bs, t, P):
    ctx = gx.DeviceContext(0)
    ctx.upload_basis(bs.xcq_shells())
    ctx.upload_tasks(t)
    ctx.upload_density(P)
    return ctx


def _run(ctx, prec, func="PBE"):
    ctx.exc_vxc(func, prec, True)
    return ctx.results()


@pytest.fixture(scope="module")
def small(gx, gpu):
    # 4 waters, UltraFineGrid: nbe spans ~20..250, so both tile widths occur.
    return _tasks(gx, 4, "UltraFineGrid")


@pytest.mark.parametrize("prec,tol_e,tol_v", [("fp64", 1e-11, 1e-12), ("mixed", 1e-9, 1e-10)])
def test_tile_policy_variants(gx, small, prec, tol_e, tol_v):
    """All-64, all-32 and split tile policies give the same EXC/VXC."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    ctx.set_tile_policy(0, 0)  # every task on 64-wide tiles
    e0, n0, v0 = _run(ctx, prec)
    for pol in [(100000, 100000), (48, 96), (96, 48)]:
        _flush(f"tile policy {pol} {prec}")
        ctx.set_tile_policy(*pol)
        e1, n1, v1 = _run(ctx, prec)
        assert abs(e1 - e0) <= tol_e
        assert abs(n1 - n0) <= tol_e
        assert np.abs(v1 - v0).max() <= tol_v


def test_density_occupancy_variants(gx, small):
    """The mixed density kernel's three-workgroups-per-CU configuration
    (two-block operand ring) computes the same arithmetic as the default
    one: identical EXC and NEL (fixed-order reductions) for every threshold,
    also with queues; VXC to the run-to-run noise of its FP64 atomics."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    ctx.set_tile_policy(48, 96)
    ctx.set_density_occupancy(0)
    e0, n0, v0 = _run(ctx, "mixed")
    for h, nq in [(100000, 0), (128, 0), (100000, 300)]:
        _flush(f"density occupancy threshold {h} queues {nq}")
        ctx.set_density_occupancy(h)
        ctx.set_schedule(nq)
        e1, n1, v1 = _run(ctx, "mixed")
        assert e1 == e0 and n1 == n0
        assert np.abs(v1 - v0).max() <= 1e-12
    ctx.set_schedule(0)
    ctx.set_density_occupancy(0)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_vxc_drop_tolerance(gx, small, prec):
    """Skipping scatter-adds below the drop tolerance (default 1e-18) leaves
    EXC/NEL unchanged and VXC within 1e-13 of adding every contribution."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    e0, n0, v0 = _run(ctx, prec)
    ctx.set_vxc_drop_tolerance(0.0)
    e1, n1, v1 = _run(ctx, prec)
    assert e1 == e0 and n1 == n0
    assert np.abs(v1 - v0).max() <= 1e-13
    ctx.set_vxc_drop_tolerance(1e-18)
    with pytest.raises(gx.GauXCError):
        ctx.set_vxc_drop_tolerance(-1.0)


def test_density_wide_tiles(gx, small):
    """128-function density tiles (mixed) for the larger tasks: the same X
    accumulation per element, the rho/grad-rho fold summed in another order,
    so EXC/NEL agree to summation noise."""
    bs, P, t = small
    assert t["nbe"].max() > 128  # some tasks span both 64-column halves of a tile
    ctx = _ctx(gx, bs, t, P)
    ctx.set_tile_policy(48, 96)
    f0, fn0, fv0 = _run(ctx, "fp64")
    e0, n0, v0 = _run(ctx, "mixed")
    for w, h in [(100, 0), (129, 100000), (0, 0)]:
        _flush(f"density 128-wide from nbe {w}, high occupancy to {h}")
        ctx.set_density_wide(w)
        ctx.set_density_occupancy(h)
        assert ctx.tile_policy()[4] == w
        e1, n1, v1 = _run(ctx, "mixed")
        assert abs(e1 - e0) <= 1e-11 * abs(e0)
        assert abs(n1 - n0) <= 1e-11 * abs(n0)
        assert np.abs(v1 - v0).max() <= 1e-12
        # FP64 runs 128-wide tasks on 64-wide tiles: the same result as before.
        f1, fn1, fv1 = _run(ctx, "fp64")
        assert abs(f1 - f0) <= 1e-12 * abs(f0) and abs(fn1 - fn0) <= 1e-12 * abs(fn0)
        assert np.abs(fv1 - fv0).max() <= 1e-13
    ctx.set_density_wide(1 << 30)
    ctx.set_density_occupancy(0)


@pytest.mark.parametrize("func", ["PBE", "SVWN5"])
def test_density_basis_eval_in_stage(gx, small, func):
    """The mixed density stage's own FP64 evaluation of B and grad B (for
    the fold and Z: mode 1; for the fold, with Z from a separate launch: 3)
    against reading the planes the collocation stage stored (mode 0: FP32,
    the value with its bfloat16 residual): all within the mixed bars of the
    FP64 path and close to each other, for 32-, 64- and 128-wide tiles and
    with queues."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    e0, n0, v0 = _run(ctx, "fp64", func)
    for dens32, wide, nq, hiocc in [(48, 1 << 30, 0, 0), (0, 0, 0, 0), (100000, 1 << 30, 0, 0),
                                    (48, 160, 300, 0), (0, 1 << 30, 0, 100000)]:
        ctx.set_tile_policy(dens32, 96)
        ctx.set_density_wide(wide)
        ctx.set_density_occupancy(hiocc)
        ctx.set_schedule(nq)
        res = {}
        for mode in (1, 3, 0):
            _flush(f"basis eval mode {mode} {func} dens32={dens32} wide={wide} nq={nq} "
                   f"hiocc={hiocc}")
            ctx.set_density_basis_eval(mode)
            assert ctx.density_basis_eval() == mode
            res[mode] = _run(ctx, "mixed", func)
            e, n, v = res[mode]
            assert abs(e - e0) <= 1e-6 and abs(n - n0) <= 1e-6
            assert np.abs(v - v0).max() <= 1e-7
        for mode in (1, 3):
            e, n, v = res[mode]
            assert abs(e - res[0][0]) <= 1e-7 and abs(n - res[0][1]) <= 1e-7
            assert np.abs(v - res[0][2]).max() <= 1e-8
        # Modes 1 and 3 evaluate the same fold and Z values in FP64.
        assert abs(res[1][0] - res[3][0]) <= 1e-11
        assert np.abs(res[1][2] - res[3][2]).max() <= 1e-12
    ctx.set_schedule(0)
    ctx.set_density_occupancy(0)
    ctx.set_density_basis_eval(3)
    for bad in (2, 4):
        with pytest.raises(gx.GauXCError):
            ctx.set_density_basis_eval(bad)


@pytest.mark.parametrize("func", ["PBE", "SVWN5"])
def test_density_basis_eval_fp64(gx, small, func):
    """FP64 path: the in-stage fold with the separate Z launch (modes 1 and 3
    both run it) against the stored planes: the same FP64 values up to the
    order of the spherical transform's sums, for 32- and 64-wide tiles, the
    three-workgroups-per-CU configuration and queues."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    for dens32, nq, hiocc in [(48, 0, 0), (0, 0, 100000), (100000, 300, 0), (0, 300, 100000)]:
        ctx.set_tile_policy(dens32, 96)
        ctx.set_density_occupancy(hiocc)
        ctx.set_schedule(nq)
        res = {}
        for mode in (3, 1, 0):
            _flush(f"fp64 basis eval mode {mode} {func} dens32={dens32} nq={nq} hiocc={hiocc}")
            ctx.set_density_basis_eval(mode)
            res[mode] = _run(ctx, "fp64", func)
        for mode in (3, 1):
            e, n, v = res[mode]
            assert abs(e - res[0][0]) <= 1e-11 and abs(n - res[0][1]) <= 1e-11
            assert np.abs(v - res[0][2]).max() <= 1e-12
    ctx.set_schedule(0)
    ctx.set_density_occupancy(0)
    ctx.set_density_basis_eval(3)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_queued_schedule_matches_direct(gx, small, prec):
    """Persistent queues (heuristic and explicit costs) reproduce the
    one-workgroup-per-item launch."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    e0, n0, v0 = _run(ctx, prec)
    nt = len(t["npts"])
    cost = t["npts"].astype(float) * t["nbe"].astype(float) ** 2
    for nq, dc, sc in [(256, None, None), (1000, cost, 0.5 * cost), (7, cost[::-1].copy(), cost)]:
        _flush(f"queues nq={nq} {prec}")
        ctx.set_schedule(nq, dc, sc)
        e1, n1, v1 = _run(ctx, prec)
        assert abs(e1 - e0) <= 1e-11
        assert np.abs(v1 - v0).max() <= 1e-12
    ctx.set_schedule(0)
    assert nt > 0


def test_graph_replay_matches_eager(gx, small):
    """Graph capture of the whole evaluation: several replays, with eager
    work and synchronisations between them, each equal to the eager run."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    e0, n0, v0 = _run(ctx, "fp64")
    em, nm, vm = _run(ctx, "mixed")
    ctx.use_graph(True)
    for rep in range(3):
        _flush(f"graph replay {rep}")
        e1, n1, v1 = _run(ctx, "fp64")
        assert abs(e1 - e0) <= 1e-11
        assert np.abs(v1 - v0).max() <= 1e-12
        assert ctx.stage_times()[6] > 0
        # eager-equivalent work in between: a mixed graph, then a host read
        e2, n2, v2 = _run(ctx, "mixed")
        assert abs(e2 - em) <= 1e-11
        assert np.abs(v2 - vm).max() <= 1e-12
        ctx.synchronize()
    ctx.use_graph(False)
    e3, n3, v3 = _run(ctx, "fp64")
    assert abs(e3 - e0) <= 1e-11


def test_batched_molecules_match_single(gx, gpu):
    """Three molecules in one context: each molecule's EXC/NEL/VXC equals its
    own single-molecule evaluation (FP64 and mixed)."""
    systems = [_tasks(gx, nw, "FineGrid", seed=s) for nw, s in [(2, 3), (3, 4), (2, 5)]]
    b = gx.BatchedDeviceContext(0)
    b.upload_molecules([bs.xcq_shells() for bs, _, _ in systems])
    b.upload_tasks([t for _, _, t in systems])
    for m, (_, P, _) in enumerate(systems):
        b.upload_density(m, P)
    for prec in ["fp64", "mixed"]:
        _flush(f"batched {prec}")
        b.exc_vxc("PBE", prec, True)
        for m, (bs, P, t) in enumerate(systems):
            e, n, v = b.results(m)
            ctx = _ctx(gx, bs, t, P)
            e1, n1, v1 = _run(ctx, prec)
            assert abs(e - e1) <= 1e-11, (m, e, e1)
            assert abs(n - n1) <= 1e-11
            assert np.abs(v - v1).max() <= 1e-12
    rc = b.lib.xcq_download_results(b.h, None, None, None, 0)
    assert rc != 0 and b"several molecules" in b.lib.xcq_last_error()


def test_task_timing_records_every_task(gx, small):
    """Per-item wall-clock stamps aggregate to a positive time per task."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    ctx.enable_task_timing(True)
    _run(ctx, "mixed")
    d, s = ctx.task_times()
    assert d.shape == (len(t["npts"]),)
    assert np.all(d > 0) and np.all(s > 0)
    # bigger tasks take longer on average
    w = t["npts"].astype(float) * t["nbe"].astype(float) ** 2
    big = w > np.quantile(w, 0.9)
    little = w < np.quantile(w, 0.1)
    assert d[big].mean() > d[little].mean()
    assert s[big].mean() > s[little].mean()


def test_autotune_keeps_results(gx, small):
    """The autotuner only picks tile thresholds and a queue count: the
    evaluation after tuning matches the one before."""
    bs, P, t = small
    ctx = _ctx(gx, bs, t, P)
    e0, n0, v0 = _run(ctx, "mixed")
    d, s, q, ms = ctx.autotune("PBE", "mixed")
    _flush(f"autotune -> {d} {s} {q} {ms:.3f} ms")
    cand = {0, 32, 48, 64, 96, 128, 192, 256, 1 << 30}
    assert d in cand and s in cand and q >= 0 and ms > 0
    assert ctx.density_basis_eval() in (0, 3)
    e1, n1, v1 = _run(ctx, "mixed")
    assert abs(e1 - e0) <= 1e-9
    assert np.abs(v1 - v0).max() <= 1e-10
