"""GPU checks of kiwi_hip_sph_to_grid_noalloc, the allocation-free variant of kiwi_hip_sph_to_grid (no
reference counterpart; for callers that cannot allocate device memory): called through the C ABI, its
sums and counters must equal kiwi_hip_sph_to_grid's bit for bit on the same device inputs, on calls in a
row whose second input is smaller (the shared scratch keeps the first call's contents), for the cubic
spline in both precisions and a generic kernel; oversize problems, host inputs and the tabulated
integral are rejected with their errors, and a later call still succeeds.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip, workloads  # noqa: E402

pytestmark = pytest.mark.gpu

COUNTERS = ("skipped_outside", "n_negative", "n_negative_warn", "pairs_bbox", "pairs_sphere", "plane_evals",
            "n_warn_records", "particle_blocks")


def say(msg):
    print(msg, flush=True)


def device(parts):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(parts[k], dtype=np.float64)).to("cuda:0") for k in workloads.COLS}


def outs(res):
    import torch
    nc = int(np.prod(res))
    return tuple(torch.zeros(nc, dtype=torch.float64, device="cuda:0") for _ in range(4))


def same(a, b):
    for u, v in zip(a[:4], b[:4]):
        assert np.array_equal(u.cpu().numpy(), v.cpu().numpy())
    for f in COUNTERS:
        assert getattr(a[4], f) == getattr(b[4], f), f
    assert list(a[4].v_mean) == list(b[4].v_mean)


def c1():
    s = np.load(os.path.join(ROOT, "tests", "golden", "c1_snapshot.npz"))
    return {k: s[k] for k in workloads.COLS}, s["domain"]


@pytest.mark.parametrize("precision", [hip.FP64, hip.MIXED])
def test_noalloc_equals_sph_to_grid_two_calls(hip_ctx, precision):
    parts, dom = c1()
    big, small = device(parts), device({k: v[::4] for k, v in parts.items()})
    for name, dp, res in (("C1 2000 particles -> 32^3", big, (32, 32, 32)),
                          ("every 4th C1 particle -> 20^3", small, (20, 20, 20))):
        say(f"case: noalloc vs sph_to_grid, {name}, precision {precision}")
        a = hip_ctx.sph_to_grid_noalloc(dp, dom, res, outs(res), precision=precision)
        b = hip_ctx.sph_to_grid(dp, dom, res, precision=precision, out=outs(res))
        same(a, b)


def test_noalloc_generic_kernel_and_lognormal(hip_ctx):
    parts, dom, res = workloads.lognormal_field(n_particles=6000, n=32, seed=9, h_max=4.0)
    dp = device(parts)
    for kern, prec in ((hip.kernel_spec("WC2"), hip.MIXED), (hip.kernel_spec("M6"), hip.FP64),
                       (hip.kernel_spec("M4"), hip.MIXED)):
        say(f"case: noalloc vs sph_to_grid, 6,000 log-normal particles -> 32^3, kernel {kern.kernel_id}, "
            f"precision {prec}")
        a = hip_ctx.sph_to_grid_noalloc(dp, dom, res, outs(res), kernel=kern, precision=prec)
        b = hip_ctx.sph_to_grid(dp, dom, res, kernel=kern, precision=prec, out=outs(res))
        same(a, b)


def test_noalloc_rejections_then_success(hip_ctx):
    parts, dom = c1()
    dp = device(parts)
    res = (32, 32, 32)
    say("case: noalloc rejects host inputs, the tabulated integral and an oversize problem")
    with pytest.raises(hip.KiwiError) as e:
        hip_ctx.sph_to_grid_noalloc({k: np.ascontiguousarray(parts[k]) for k in workloads.COLS}, dom, res, outs(res))
    assert e.value.code == -1
    with pytest.raises(hip.KiwiError) as e:
        hip_ctx.sph_to_grid_noalloc(dp, dom, res, outs(res), kernel=hip.kernel_spec("M4", tabulated=True))
    assert e.value.code == -4
    # about 1.3e8 bounding-box pairs in FP64 (~1 GB of per-pair sums): beyond the fixed scratch
    lp, ldom, lres = workloads.lognormal_field(n_particles=1_000_000, n=256, seed=2)
    sub = device(workloads.subsample(lp, 16))
    with pytest.raises(hip.KiwiError) as e:
        hip_ctx.sph_to_grid_noalloc(sub, ldom, lres, outs(lres), precision=hip.FP64)
    assert e.value.code == -5 and "scratch" in str(e.value)
    say("case: noalloc after the rejections")
    a = hip_ctx.sph_to_grid_noalloc(dp, dom, res, outs(res))
    b = hip_ctx.sph_to_grid(dp, dom, res, out=outs(res))
    same(a, b)
