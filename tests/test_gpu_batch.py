"""GPU checks of kiwi_hip_sph_to_grid_batch (no reference counterpart: the reference maps one cutout per
process): every problem of a batch must return exactly what kiwi_hip_sph_to_grid returns for it alone
(bit-identical sums, the same counters), whatever the mix of kernels, precisions, grids and output
locations in the batch.  Parity of the single call itself is pinned in test_gpu_sph_to_grid.py.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip, workloads  # noqa: E402

pytestmark = pytest.mark.gpu

COLS = ("x", "y", "z", "vx", "vy", "vz", "m", "h")
COUNTERS = ("skipped_outside", "n_negative", "n_negative_warn", "pairs_bbox", "pairs_sphere", "plane_evals",
            "n_warn_records")


def say(msg):
    print(msg, flush=True)


def problems(seed=5):
    """A mixed batch: cubic FP64 and mixed, Wendland C2 and M6 (generic), tabulated cubic, an empty
    problem, one whose particles all miss the grid, aspect ratios up to 1:4."""
    rng = np.random.default_rng(seed)
    specs = [(800, (16, 24, 12), hip.kernel_spec("M4"), hip.FP64),
             (3000, (40, 10, 30), hip.kernel_spec("M4"), hip.MIXED),
             (1500, (20, 20, 20), hip.kernel_spec("WC2"), hip.FP64),
             (0, (8, 8, 8), hip.kernel_spec("M4"), hip.FP64),
             (2500, (32, 32, 8), hip.kernel_spec("M4"), hip.MIXED),
             (1200, (12, 48, 16), hip.kernel_spec("M6"), hip.FP64),
             (900, (24, 18, 30), hip.kernel_spec("M4", tabulated=True), hip.FP64),
             (400, (16, 16, 16), hip.kernel_spec("M4"), hip.FP64),
             (2000, (30, 30, 30), hip.kernel_spec("WC2"), hip.MIXED)]
    out = []
    for q, (n, dims, kern, prec) in enumerate(specs):
        if n > 0:
            parts, dom, res = workloads.lognormal_box(n, dims, seed=100 + q)
            parts = {k: np.ascontiguousarray(parts[k]) for k in COLS}
        else:
            parts = {k: np.zeros(0) for k in COLS}
            dom, res = np.array([0.0, dims[0], 0.0, dims[1], 0.0, dims[2]]), dims
        if q == 7:  # every particle far outside this problem's grid
            parts["x"] = parts["x"] + 1000.0
        out.append({"parts": parts, "domain": dom, "res": res, "kernel": kern, "precision": prec})
    rng.shuffle(out)
    return out


def test_batch_equals_single_calls(hip_ctx):
    say("case: 9 mixed problems (kernels, precisions, empty, all-outside) vs 9 single calls")
    probs = problems()
    got = hip_ctx.sph_to_grid_batch(probs)
    for q, (prob, res) in enumerate(zip(probs, got)):
        single = hip_ctx.sph_to_grid(prob["parts"], prob["domain"], prob["res"], kernel=prob["kernel"],
                                     precision=prob["precision"])
        for a, b in zip(res[:4], single[:4]):
            assert np.all(np.isfinite(a)) and np.all(np.isfinite(b)), q
            assert np.array_equal(a, b), q
        for f in COUNTERS:
            assert getattr(res[4], f) == getattr(single[4], f), (q, f)
        np.testing.assert_array_equal(np.array(res[4].v_mean), np.array(single[4].v_mean))
    say(f"  pairs per problem: {[r[4].pairs_sphere for r in got]}")


def test_batch_device_and_host_outputs(hip_ctx):
    import torch
    say("case: batch with device-resident particles and outputs for some problems")
    probs = problems(seed=6)
    ref = hip_ctx.sph_to_grid_batch(probs)
    dev = torch.device("cuda:0")
    for q, prob in enumerate(probs):
        if q % 2 == 0:
            prob["parts"] = {k: torch.from_numpy(v).to(dev) for k, v in prob["parts"].items()}
            nc = int(np.prod(prob["res"]))
            prob["out"] = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
    got = hip_ctx.sph_to_grid_batch(probs)
    torch.cuda.synchronize()
    for q, (a, b) in enumerate(zip(got, ref)):
        for x, y in zip(a[:4], b[:4]):
            xa = x.cpu().numpy() if hasattr(x, "cpu") else x
            assert np.array_equal(xa, y), q


def test_batch_errors_name_the_problem(hip_ctx):
    say("case: a malformed problem in a batch is reported with its index; the context stays usable")
    probs = problems()
    probs[4]["res"] = (0, 8, 8)
    with pytest.raises(hip.KiwiError, match="problem 4"):
        hip_ctx.sph_to_grid_batch(probs)
    assert len(hip_ctx.sph_to_grid_batch(problems()[:2])) == 2
    assert hip_ctx.sph_to_grid_batch([]) == []


def test_launch_shapes_do_not_change_results_and_autotune(hip_ctx):
    say("case: the cell kernel's launch shape (item grab, cursor slices) and the autotuner")
    probs = problems(seed=8)
    assert hip_ctx.get_tuning() == (16, 8)
    ref = hip_ctx.sph_to_grid_batch(probs)
    for shape in ((1, 1), (8, 1), (64, 16), (256, 3)):
        hip_ctx.set_tuning(*shape)
        assert hip_ctx.get_tuning() == shape
        got = hip_ctx.sph_to_grid_batch(probs)
        for q, (a, b) in enumerate(zip(got, ref)):
            for x, y in zip(a[:4], b[:4]):
                assert np.array_equal(x, y), (shape, q)
            assert a[4].pairs_sphere == b[4].pairs_sphere
    for bad in ((0, 8), (257, 8), (16, 0), (16, 17)):
        with pytest.raises(hip.KiwiError):
            hip_ctx.set_tuning(*bad)
    hip_ctx.set_tuning()
    assert hip_ctx.get_tuning() == (16, 8)
    shape, ms = hip_ctx.autotune(probs)
    say(f"  autotuned shape {shape}, {ms:.2f} ms per call")
    assert shape[0] in (8, 16, 32, 64) and shape[1] in (1, 4, 8, 16) and ms > 0
    assert hip_ctx.get_tuning() == shape
    got = hip_ctx.sph_to_grid_batch(probs)
    for a, b in zip(got, ref):
        for x, y in zip(a[:4], b[:4]):
            assert np.array_equal(x, y)
    hip_ctx.set_tuning()
