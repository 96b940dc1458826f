"""GPU runs of the five workload configurations at their stated sizes (the parity suites elsewhere run
them small).  Each test prints its case before any device work.

  C2  1M log-normal clustered particles -> 256^3 (the benchmark workload, kiwi_amd.workloads seed 2):
      FP64 against the oracle on every 128th particle onto the same grid (7,813 particles, h up to 50
      cells), cells outside 1e-10 of the field maximum arbitrated in binary128 (tests/parity.py);
      the full set in FP64 and mixed precision (total mass within 1e-6, every field reported, repeat
      calls bit-identical); the padded-domain variant C2p conserving mass and momentum to 1e-9.
  C3  64 uneven cutouts (workloads.cutout_batch): the batch equals its 64 single calls bit for bit, in
      mixed precision and in FP64; the two smallest cubic-spline and Wendland C2 problems match the
      oracle in FP64; mixed within 1e-6 of FP64 in total mass per problem.
  C4  the adjoint on 2e5 log-normal particles -> 128^3, L = sum (rho - rho_target)^2 with rho_target
      from a copy with 5 % Gaussian mass perturbations: FP64 gradients against finite differences of the
      FP64 forward on random particles (1e-5), the invariants of L = total mass and L = total x
      momentum on a padded domain, mixed against FP64, and 200 Adam steps of mass recovery.
  C5  the Voronoi map of 2e5 log-normal particles plus the coverage-grid buffer (computed on the GPU,
      equal to the reference's add_coverage_grid_buffer) and one Lloyd step (the centroids computed on
      the GPU, equal to voro++'s within 1e-12 of the extent), on the domain padded by
      the largest support: mass and momentum within 1e-8 and cell volumes filling the domain within
      1e-4 (the reference's own thresholds, tests/real_data_test.py:250-257 of the reference), mixed
      within 1e-6 of FP64 in total mass, and the oracle's BFS on a 3,000-particle sample.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip, workloads  # noqa: E402
from parity import assert_cells_match, assert_voronoi_cells_match  # noqa: E402

pytestmark = pytest.mark.gpu

COLS = workloads.COLS


def say(msg):
    print(msg, flush=True)


def rel_max(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


def threads():
    return min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def c2():
    parts, dom, res = workloads.lognormal_field(n_particles=1_000_000, n=256, seed=2)
    return {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in COLS}, dom, res


# --------------------------------------------------------------------------------------------------- C2
def test_c2_fp64_sample_matches_oracle(hip_ctx, c2):
    from oracle import oracle as O
    parts, dom, res = c2
    sample = workloads.subsample(parts, 128)
    say(f"case: C2 sample, every 128th particle ({len(sample['x'])}, h {sample['h'].min():.2f}-"
        f"{sample['h'].max():.2f} cells) -> 256^3, FP64 vs oracle")
    mass, px, py, pz, st = hip_ctx.sph_to_grid(sample, dom, res, precision=hip.FP64)
    om, opx, opy, opz, ost = O.sph_to_grid(sample, dom, *res, threads=threads())
    nnz = int(np.count_nonzero(om))
    n = assert_cells_match((mass, px, py, pz), (om, opx, opy, opz), sample, dom, res,
                           max_disputed=max(2000, nnz // 500))
    say(f"  {nnz} nonzero cells, {n} arbitrated in binary128")
    assert st.pairs_bbox == ost.pairs_bbox and st.skipped_outside == ost.skipped_outside
    assert st.n_negative_warn == ost.n_negative_warn


def test_c2_full_mixed_against_fp64(hip_ctx, c2):
    parts, dom, res = c2
    say("case: C2 1,000,000 particles -> 256^3, FP64 and mixed, repeat calls")
    f = hip_ctx.sph_to_grid(parts, dom, res, precision=hip.FP64)
    m = hip_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    tot64 = f[0].sum()
    err = abs(m[0].sum() - tot64) / tot64
    l1 = float(np.abs(m[0] - f[0]).sum() / np.abs(f[0]).sum())
    fields = {name: rel_max(a, b) for name, a, b in zip(("mass", "px", "py", "pz"), m[:4], f[:4])}
    say(f"  total mass rel err {err:.3e}; mass-weighted L1 {l1:.2e}; per-field max err / max {fields}")
    assert err <= 1e-6
    assert max(fields.values()) <= 1e-5
    assert f[4].pairs_bbox == m[4].pairs_bbox == 2_137_844_450
    assert abs(m[4].pairs_sphere - f[4].pairs_sphere) <= 1e-6 * f[4].pairs_sphere
    again = hip_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    for a, b in zip(again[:4], m[:4]):
        assert np.array_equal(a, b)


def test_c2p_padded_conservation(hip_ctx, c2):
    parts, _, res = c2
    pad = workloads.padded_domain(parts)
    say(f"case: C2p 1,000,000 particles -> 256^3 over the padded domain {np.round(pad, 2).tolist()}")
    M = parts["m"].sum()
    for prec, tol in ((hip.FP64, 1e-9), (hip.MIXED, 1e-6)):
        mass, px, py, pz, st = hip_ctx.sph_to_grid(parts, pad, res, precision=prec)
        em = abs(mass.sum() - M) / M
        ep = [abs(p.sum() - (parts["m"] * parts[c]).sum()) / np.abs(parts["m"] * parts[c]).sum()
              for p, c in ((px, "vx"), (py, "vy"), (pz, "vz"))]
        say(f"  precision {prec}: mass rel err {em:.2e}, momentum rel err {max(ep):.2e}")
        assert st.skipped_outside == 0
        assert em <= tol and max(ep) <= tol


# --------------------------------------------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3():
    return workloads.cutout_batch(n_problems=64)


def c3_problems(batch, precision):
    return [{"parts": {k: np.ascontiguousarray(p[k], dtype=np.float64) for k in COLS}, "domain": d, "res": r,
             "kernel": hip.kernel_spec(kn), "precision": precision} for p, d, r, kn in batch]


COUNTERS = ("skipped_outside", "n_negative", "n_negative_warn", "pairs_bbox", "pairs_sphere", "plane_evals")


@pytest.mark.parametrize("precision", [hip.MIXED, hip.FP64])
def test_c3_batch_equals_single_calls(hip_ctx, c3, precision):
    probs = c3_problems(c3, precision)
    say(f"case: C3 64 cutouts ({sum(len(p['parts']['x']) for p in probs)} particles, "
        f"{sum(int(np.prod(p['res'])) for p in probs)} cells), precision {precision}: batch vs single calls")
    got = hip_ctx.sph_to_grid_batch(probs)
    for q, (prob, res) in enumerate(zip(probs, got)):
        single = hip_ctx.sph_to_grid(prob["parts"], prob["domain"], prob["res"], kernel=prob["kernel"],
                                     precision=precision)
        for a, b in zip(res[:4], single[:4]):
            assert np.array_equal(a, b), q
        for f in COUNTERS:
            assert getattr(res[4], f) == getattr(single[4], f), (q, f)


def test_c3_smallest_problems_match_oracle_and_mixed(hip_ctx, c3):
    from oracle import oracle as O
    p64 = c3_problems(c3, hip.FP64)
    p32 = c3_problems(c3, hip.MIXED)
    g64 = hip_ctx.sph_to_grid_batch(p64)
    g32 = hip_ctx.sph_to_grid_batch(p32)
    for q, (a, b) in enumerate(zip(g32, g64)):
        err = abs(a[0].sum() - b[0].sum()) / b[0].sum()
        assert err <= 1e-6, (q, err)
    sizes = [len(p["parts"]["x"]) for p in p64]
    for kname in ("M4", "WC2"):
        idx = sorted((q for q in range(len(p64)) if c3[q][3] == kname), key=lambda q: sizes[q])[:2]
        for q in idx:
            prob = p64[q]
            say(f"case: C3 problem {q} ({kname}, {sizes[q]} particles -> {tuple(prob['res'])}) FP64 vs oracle")
            ok = O.kernel(kname, generic=(kname != "M4"))
            ref = O.sph_to_grid(prob["parts"], prob["domain"], *prob["res"], k=ok, threads=threads())
            tol = 1e-10 if kname == "M4" else 2e-10
            n = assert_cells_match(g64[q][:4], ref[:4], prob["parts"], prob["domain"], prob["res"], tol=tol,
                                   quad_tol=tol if kname != "M4" else 1e-11, k=ok)
            say(f"  cells arbitrated in binary128: {n}")
            assert g64[q][4].pairs_bbox == ref[4].pairs_bbox


# --------------------------------------------------------------------------------------------------- C4
@pytest.fixture(scope="module")
def c4():
    import torch
    parts, dom, res = workloads.lognormal_field(n_particles=200_000, n=128, seed=4)
    parts = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in COLS}
    dev = torch.device("cuda:0")
    return parts, dom, res, {k: torch.from_numpy(v).to(dev) for k, v in parts.items()}


def test_c4_adjoint_finite_differences(hip_ctx, c4):
    import torch
    parts, dom, res, t = c4
    n = len(parts["x"])
    dev = t["x"].device
    nc = int(np.prod(res))
    V = float(np.prod((dom[1::2] - dom[0::2]) / np.array(res)))
    rng = np.random.default_rng(12)
    tp = dict(t)
    tp["m"] = torch.from_numpy(parts["m"] * (1 + 0.05 * rng.standard_normal(n))).to(dev)

    def forward(tt, precision=hip.FP64):
        out = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
        hip_ctx.sph_to_grid(tt, dom, res, precision=precision, out=out)
        return out

    say("case: C4 adjoint, 200,000 particles -> 128^3, L = sum (rho - rho_target)^2, FP64 and mixed")
    target = forward(tp)[0] / V
    mass = forward(t)[0]
    gm = (2 * (mass / V - target) / V).contiguous()
    zero = torch.zeros(nc, dtype=torch.float64, device=dev)
    g64, st = hip_ctx.sph_to_grid_adjoint(t, dom, res, (gm, zero, zero, zero))
    g32, _ = hip_ctx.sph_to_grid_adjoint(t, dom, res, (gm, zero, zero, zero), precision=hip.MIXED)
    mixed = {f: float(np.max(np.abs(g32[f] - g64[f])) / max(np.abs(g64[f]).max(), 1e-300)) for f in hip.GRAD_FIELDS}
    say(f"  {st.cells} in-support pairs; mixed vs FP64 (max err / max): {mixed}")
    assert mixed["m"] <= 2e-4 and max(mixed[f] for f in ("x", "y", "z", "h")) <= 3e-2

    one = {k: torch.empty(1, dtype=torch.float64, device=dev) for k in t}
    single = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))

    def contribution(i, f=None, value=None):
        for k in t:
            one[k].copy_(t[k][i:i + 1])
        if f is not None:
            one[f].fill_(value)
        hip_ctx.sph_to_grid(one, dom, res, out=single)
        return single[0].clone()

    idx = rng.choice(n, 60, replace=False)
    for f in ("m", "x", "y", "z", "h"):
        scale = float(np.sqrt(np.mean(g64[f] ** 2)))
        worst = 0.0
        for i in (int(v) for v in idx):
            base = float(parts[f][i])
            own = contribution(i)
            if f == "m":  # L is quadratic in m: the central difference is exact up to rounding
                eps = 1e-6 * abs(base)
                rp = (mass - own + contribution(i, f, base + eps)) / V - target
                rm = (mass - own + contribution(i, f, base - eps)) / V - target
                fd = float(torch.sum(rp * rp - rm * rm)) / (2 * eps)
            else:  # fourth-order central difference
                eps = 1e-3 * float(parts["h"][i])
                d = []
                for k in (1, 2):
                    rp = (mass - own + contribution(i, f, base + k * eps)) / V - target
                    rm = (mass - own + contribution(i, f, base - k * eps)) / V - target
                    d.append(float(torch.sum(rp * rp - rm * rm)))
                fd = (8 * d[0] - d[1]) / (12 * eps)
            worst = max(worst, abs(g64[f][i] - fd) / max(abs(fd), 1e-2 * scale))
        say(f"  d/d{f}: worst relative error {worst:.2e} over {len(idx)} particles")
        assert worst <= 1e-5, f


def test_c4_adjoint_invariants_padded(hip_ctx, c4):
    parts, _, res, _ = c4
    pad = workloads.padded_domain(parts)
    nc = int(np.prod(res))
    one, zero = np.ones(nc), np.zeros(nc)
    say("case: C4 adjoint invariants on the padded domain: L = total mass, L = total x momentum")
    g, _ = hip_ctx.sph_to_grid_adjoint(parts, pad, res, (one, zero, zero, zero))
    m_over_h = parts["m"] / parts["h"]
    errs = {"m": float(np.max(np.abs(g["m"] - 1.0)))}
    errs.update({f: float(np.max(np.abs(g[f]) / m_over_h)) for f in ("x", "y", "z", "h")})
    say(f"  L = total mass: max |dL/dm - 1| and max |dL/dq| / (m/h): {errs}")
    assert errs["m"] <= 1e-11
    for f in ("x", "y", "z", "h"):
        assert errs[f] < 1e-10, f
    g, _ = hip_ctx.sph_to_grid_adjoint(parts, pad, res, (zero, one, zero, zero))
    np.testing.assert_allclose(g["vx"], parts["m"], rtol=1e-11, atol=0)
    np.testing.assert_allclose(g["m"], parts["vx"], rtol=0, atol=1e-11 * np.abs(parts["vx"]).max())


def test_c4_mass_recovery(hip_ctx, c4):
    import torch
    from kiwi_amd.autograd import sph_to_grid
    parts, dom, res, t = c4
    n = len(parts["x"])
    nc = int(np.prod(res))
    V = float(np.prod((dom[1::2] - dom[0::2]) / np.array(res)))
    say("case: C4 mass recovery, 200 Adam steps from 5 % perturbed masses")
    out = tuple(torch.zeros(nc, dtype=torch.float64, device=t["x"].device) for _ in range(4))
    hip_ctx.sph_to_grid(t, dom, res, out=out)
    truth = out[0] / V
    rng = np.random.default_rng(12)
    m = torch.from_numpy(parts["m"] * (1 + 0.05 * rng.standard_normal(n))).to(t["x"].device).requires_grad_(True)
    opt = torch.optim.Adam([m], lr=0.01 * float(parts["m"].mean()))
    losses = []
    for step in range(200):
        opt.zero_grad()
        rho = sph_to_grid(m, t["x"], t["y"], t["z"], t["h"], t["vx"], t["vy"], t["vz"], hip_ctx, dom, res)[0] / V
        loss = torch.sum((rho - truth) ** 2)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    err = float(torch.sqrt(torch.mean(((m.detach() - t["m"]) / t["m"]) ** 2)))
    say(f"  loss {losses[0]:.3e} -> {losses[-1]:.3e}; rms relative mass error {err:.4f}")
    assert losses[-1] < 1e-2 * losses[0]
    assert err < 0.05


# --------------------------------------------------------------------------------------------------- C5
@pytest.fixture(scope="module")
def c5(hip_ctx):
    from oracle import oracle as O
    assert os.path.exists(O.VORO_SO), f"voro++ shim not built at {O.VORO_SO}: the tessellation is host input"
    parts, _, _ = workloads.lognormal_field(n_particles=200_000, n=256, seed=5)
    parts = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in COLS}
    dom = workloads.padded_domain(parts)
    say(f"case: C5 setup, 200,000 particles, padded domain {np.round(dom, 2).tolist()}: coverage buffer, "
        f"tessellation, one Lloyd step, tessellation")
    ids, xyz, _ = hip_ctx.coverage_buffer(parts, dom, 0.05, 2.0)
    rid, rxyz = O.ref_coverage_buffer(parts["x"], parts["y"], parts["z"], parts["h"], dom, 0.05, 2.0)
    assert np.array_equal(ids, rid) and np.array_equal(xyz, rxyz)
    pos = np.stack([parts["x"], parts["y"], parts["z"]], 1)
    sites = np.concatenate([pos, xyz])
    sid = np.concatenate([np.arange(1, len(pos) + 1, dtype=np.int32), ids.astype(np.int32)])
    g0 = O.tessellate(sid, sites, dom, np.zeros((0, 3)))
    c0 = {k: np.ascontiguousarray(g0[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr")}
    c0["vert_xyz"] = np.ascontiguousarray(g0["vert_xyz"].reshape(-1))
    c0["domain"] = dom
    cen, vol = hip_ctx.voronoi_centroids(c0, np.ascontiguousarray(g0["site"]))  # the Lloyd step on the GPU
    ext = float(np.max(dom[1::2] - dom[0::2]))
    say(f"  Lloyd step: centroids within {np.max(np.abs(cen - g0['centroid'])) / ext:.1e} of the extent of "
        f"voro++'s, volumes within {np.max(np.abs(vol - g0['volume']) / g0['volume']):.1e}")
    assert np.max(np.abs(cen - g0["centroid"])) <= 1e-12 * ext
    g = O.tessellate(g0["cell_id"], cen, dom, pos)  # the final cells, from the Lloyd step's sites
    cells = {k: np.ascontiguousarray(g[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr")}
    cells["vert_xyz"] = np.ascontiguousarray(g["vert_xyz"].reshape(-1))
    cells["domain"] = dom
    say(f"  {len(ids)} buffer sites, {len(g['volume'])} cells, {len(g['face_nbr'])} faces")
    return parts, g["home"].astype(np.int32), cells, g, dom, g0


def test_c5_map_conservation_and_mixed(hip_ctx, c5):
    parts, home, cells, g, dom, _ = c5
    say("case: C5 map, FP64 and mixed: mass, momentum, volume fill")
    f = hip_ctx.voronoi_map(parts, home, cells, precision=hip.FP64)
    m = hip_ctx.voronoi_map(parts, home, cells, precision=hip.MIXED)
    M = parts["m"].sum()
    mabsv = sum(np.abs(parts["m"] * parts[c]).sum() for c in ("vx", "vy", "vz"))
    em = abs(f[0].sum() - M) / M
    dmom = max(abs(f[k].sum() - (parts["m"] * parts[c]).sum()) for k, c in ((1, "vx"), (2, "vy"), (3, "vz"))) / mabsv
    vol = float(np.prod(dom[1::2] - dom[0::2]))
    ev = abs(g["volume"].sum() - vol) / vol
    e32 = abs(m[0].sum() - f[0].sum()) / f[0].sum()
    say(f"  mass rel err {em:.2e}, momentum err/sum(m|v|) {dmom:.2e}, volume fill {ev:.2e}, mixed total mass "
        f"{e32:.2e}; visited pairs {f[4].pairs_visited}, face integrals {f[4].face_evals}")
    assert f[4].unresolved == 0
    assert em < 1e-8 and dmom < 1e-8 and ev < 1e-4
    assert e32 <= 1e-6
    assert np.all(f[0] >= 0.0) and np.all(g["volume"] > 0.0)


def test_c5_sample_matches_oracle(hip_ctx, c5):
    from oracle import oracle as O
    parts, home, cells, _, _, _ = c5
    k = 3000
    sample = {c: np.ascontiguousarray(v[:k]) for c, v in parts.items()}
    sh = np.ascontiguousarray(home[:k])
    say(f"case: C5 map of the first {k} particles over the full tessellation, FP64 vs oracle BFS")
    got = hip_ctx.voronoi_map(sample, sh, cells, precision=hip.FP64)
    pid = np.arange(1, k + 1, dtype=np.int32)
    ref = O.voronoi_map(sample, pid, sh, cells["cell_face_ptr"], cells["face_vert_ptr"], cells["vert_xyz"],
                        cells["face_nbr"])
    n = assert_voronoi_cells_match(got[:4], ref[:4], sample, pid, sh, cells)
    say(f"  cells arbitrated in binary128: {n}")
    assert got[4].pairs_visited == ref[4][4] and got[4].face_evals == ref[4][5]
