"""Graph mode (kiwi_hip_set_graphs): the sph-to-grid pipeline after the sizing pass captured into a HIP graph
and replayed.  Every replay is checked bit for bit against an eager call on the same inputs, with other
eager work and synchronisations between the replays; a replay reads the inputs current at its launch
(masses and velocities changed between replays), a call whose sizes change captures anew, batches replay
one graph per kernel path, and calls with host outputs run eagerly.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip, workloads  # noqa: E402

pytestmark = pytest.mark.gpu

COLS = workloads.COLS


def say(msg):
    print(msg, flush=True)


@pytest.fixture
def gctx(hip_ctx):
    yield hip_ctx
    hip_ctx.set_graphs(False)


def device(parts):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to("cuda:0") for k, v in parts.items()}


def outputs(res):
    import torch
    nc = int(np.prod(res))
    return tuple(torch.zeros(nc, dtype=torch.float64, device="cuda:0") for _ in range(4))


def host(out):
    return [t.cpu().numpy().copy() for t in out]


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("precision", [hip.MIXED, hip.FP64])
def test_replays_equal_eager(gctx, precision):
    import torch
    say(f"case: graph mode, 30,000 log-normal particles -> 48^3, precision {precision}: capture + 3 replays")
    parts, dom, res = workloads.lognormal_field(n_particles=30000, n=48, seed=9, h_max=12.0)
    dev = device(parts)
    out = outputs(res)
    gctx.set_graphs(False)
    *_, st_e = gctx.sph_to_grid(dev, dom, res, precision=precision, out=out)
    ref = host(out)
    other_parts, other_dom, other_res = workloads.lognormal_field(n_particles=5000, n=24, seed=10)
    other = device(other_parts)
    other_out = outputs(other_res)
    gctx.set_graphs(True)
    assert gctx.get_graphs()
    for rep in range(4):  # the capture, then three replays
        for t in out:
            t.fill_(-1.0)
        *_, st = gctx.sph_to_grid(dev, dom, res, precision=precision, out=out)
        got = host(out)
        assert same(got, ref), f"replay {rep} differs from the eager call"
        assert st.pairs_sphere == st_e.pairs_sphere and st.pairs_bbox == st_e.pairs_bbox
        assert st.lattice_points[:] == st_e.lattice_points[:] and st.fp64_fallback_blocks == st_e.fp64_fallback_blocks
        assert list(st.v_mean) == list(st_e.v_mean)
        # eager work between the replays: another problem in graph mode's eager path (host outputs), a
        # torch reduction and a synchronisation
        gctx.sph_to_grid(other_parts, other_dom, other_res, precision=precision)
        float(torch.sum(other["m"]))
        torch.cuda.synchronize()
    say(f"  4 graph-mode calls equal the eager call bit for bit ({st_e.pairs_sphere} sphere pairs)")
    assert other_out is not None


def test_replay_reads_current_inputs_and_recaptures(gctx):
    say("case: graph mode, masses and velocities changed between replays; positions changed -> new capture")
    parts, dom, res = workloads.lognormal_field(n_particles=20000, n=40, seed=11, h_max=8.0)
    dev = device(parts)
    out = outputs(res)
    gctx.set_graphs(True)
    gctx.sph_to_grid(dev, dom, res, precision=hip.MIXED, out=out)  # capture
    rng = np.random.default_rng(5)
    for rep in range(3):
        new = dict(parts)
        new["m"] = parts["m"] * (1.0 + 0.1 * rng.standard_normal(len(parts["m"])))
        new["vx"] = parts["vx"] + rng.standard_normal(len(parts["vx"]))
        for k in ("m", "vx"):
            dev[k].copy_(device({k: new[k]})[k])
        gctx.sph_to_grid(dev, dom, res, precision=hip.MIXED, out=out)
        got = host(out)
        gctx.set_graphs(False)
        ref_out = outputs(res)
        gctx.sph_to_grid(dev, dom, res, precision=hip.MIXED, out=ref_out)
        assert same(got, host(ref_out)), f"replay {rep} with new masses differs from the eager call"
        gctx.set_graphs(True)
        gctx.sph_to_grid(dev, dom, res, precision=hip.MIXED, out=out)  # capture again after the release
    # moved particles: other item counts, a new capture
    moved = dict(parts)
    moved["x"] = parts["x"] + 0.37
    devm = device(moved)
    gctx.sph_to_grid(devm, dom, res, precision=hip.MIXED, out=out)
    got = host(out)
    gctx.set_graphs(False)
    ref_out = outputs(res)
    gctx.sph_to_grid(devm, dom, res, precision=hip.MIXED, out=ref_out)
    assert same(got, host(ref_out))


def test_batch_graph_equals_eager(gctx):
    say("case: graph mode, a batch of 6 cutouts (M4 and WC2, mixed and FP64) replayed twice vs eager")
    batch = workloads.cutout_batch(n_problems=6, n_hi=20000, g_hi=64)
    probs = []
    for i, (p, d, r, kn) in enumerate(batch):
        probs.append({"parts": device(p), "domain": d, "res": r, "kernel": hip.kernel_spec(kn),
                      "precision": hip.MIXED if i % 3 else hip.FP64, "out": outputs(r)})
    gctx.set_graphs(False)
    gctx.sph_to_grid_batch(probs)
    ref = [host(q["out"]) for q in probs]
    gctx.set_graphs(True)
    for rep in range(3):
        for q in probs:
            for t in q["out"]:
                t.fill_(-1.0)
        gctx.sph_to_grid_batch(probs)
        for i, q in enumerate(probs):
            assert same(host(q["out"]), ref[i]), f"batch call {rep}, problem {i} differs"


def test_host_outputs_run_eagerly(gctx):
    say("case: graph mode with host outputs (eager path) equals graph mode off")
    parts, dom, res = workloads.lognormal_field(n_particles=8000, n=32, seed=12)
    gctx.set_graphs(False)
    a = gctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    gctx.set_graphs(True)
    b = gctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    assert same(a[:4], b[:4])
    assert b[4].device_ms_lines_kernel > 0  # timed per phase: the eager path
    assert gctx.L.kiwi_hip_set_graphs(gctx.h, 2) == -1  # KIWI_ERR_INVALID_ARG
