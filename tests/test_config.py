"""The host configuration and provenance manifest (libkiwi_io: kiwi_config_load, kiwi_manifest_write,
kiwi_manifest_check) against the reference CLI built from its sources here (oracle/_ref/kiwi, by
oracle/ref.mk).  CPU only.

  * Every malformed config of the reference's tests/bad_input_test.py that fails in the config reader,
    plus further wrong types, structures and values: the message equals the reference's "[kiwi] error:"
    line exactly.  For JSON syntax errors the reference's message carries nlohmann::json's parser
    diagnostics; the text is compared up to and including the line number (the column and the
    diagnostic wording of the library's parser are not restated: parity unpinned for them).
  * Valid configs: Config::to_json equals the "config" section of the manifest the reference writes.
  * manifest.json after a reference run (a small snapshot, and a 70 MB one whose fingerprint takes the
    head-and-tail path): the bytes written here equal the reference's, the write time aside.
  * A changed config in the same run directory: the warning text equals the reference's.
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import io, workloads  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.exists(O.REF_CLI), reason="oracle/_ref not built (needs the "
                                                                      "reference sources)")

GOOD_ROWS = ["0.1 0.1 0.1 0 0 0 1 0.3 1", "0.2 -0.3 0.1 0 0 0 1 0.3 1", "-0.5 0.4 0.3 0 0 0 1 0.3 1"]
BASE = {"run_name": "bad", "domain": {"x_min": -1, "x_max": 1, "y_min": -1, "y_max": 1, "z_min": -1, "z_max": 1},
        "input_snapshot": "good.out", "input_snapshot_format": "plain_ascii", "work_folder": "work"}


def cfg(**changes):
    c = json.loads(json.dumps(BASE))
    for k, v in changes.items():
        if isinstance(v, dict) and isinstance(c.get(k), dict):
            c[k].update(v)
        else:
            c[k] = v
    return c


CASES = [
    ("unknown snapshot format", cfg(input_snapshot_format="fits")),
    ("removed format name", cfg(input_snapshot_format="dale_phantom")),
    ("graded_shells without its parameters", cfg(buffer={"enabled": True, "strategy": "graded_shells"})),
    ("unknown kernel", cfg(kernel={"name": "gaussian"})),
    ("zero resolution", cfg(cartesian_grid={"resolution": 0})),
    ("negative resolution", cfg(cartesian_grid={"resolution": -4})),
    ("negative resolution_y", cfg(cartesian_grid={"resolution_y": -1})),
    ("negative Lloyd iterations", cfg(lloyd={"iterations": -1})),
    ("inverted domain x", cfg(domain={"x_min": 1, "x_max": -1})),
    ("inverted domain z", cfg(domain={"z_min": 1, "z_max": 1})),
    ("domain value is a string", cfg(domain={"x_min": "a"})),
    ("domain value is a boolean", cfg(domain={"y_max": True})),
    ("domain is an array", cfg(domain=[1, 2])),
    ("missing domain key", {"run_name": "x", "domain": {"x_min": -1}}),
    ("missing run_name", {"domain": BASE["domain"]}),
    ("missing input_snapshot", {k: v for k, v in BASE.items() if k != "input_snapshot"}),
    ("missing work_folder", {k: v for k, v in BASE.items() if k != "work_folder"}),
    ("run_name is a number", cfg(run_name=3)),
    ("root is an array", [1, 2, 3]),
    ("block_scale_factor zero", cfg(grid_blocks={"block_scale_factor": 0})),
    ("grid_blocks is a number", cfg(grid_blocks=4)),
    ("coverage spacing zero", cfg(buffer={"enabled": True, "coverage_grid": {"spacing_fraction": 0}})),
    ("coverage factor negative", cfg(buffer={"enabled": True, "coverage_grid": {"coverage_radius_factor": -1.5}})),
    ("coverage_grid is a string", cfg(buffer={"enabled": True, "coverage_grid": "x"})),
    ("buffer enabled is a string", cfg(buffer={"enabled": "yes"})),
    ("graded_shells missing dz", cfg(buffer={"enabled": True, "strategy": "graded_shells",
                                             "graded_shells": {"dx": 1, "dy": 1}})),
    ("lloyd iterations float", cfg(lloyd={"iterations": "2"})),
    ("diagnostics empty python", cfg(diagnostics={"python_executable": ""})),
    ("cartesian_map resolution zero", cfg(cartesian_map={"resolution": 0})),
    ("negative support radius", cfg(kernel={"support_radius_in_h": -1})),
    ("negative hfact", cfg(kernel={"name": "wendland_c2", "hfact": -0.1})),
    ("tabulated with a generic kernel", cfg(kernel={"name": "quintic"}, cartesian_grid={"use_tabulated_integral": True})),
    ("tabulated with a wider support", cfg(kernel={"support_radius_in_h": 2.5}, mapping={"use_tabulated_integral": True})),
    ("phantom units zero mass", cfg(phantom_units={"mass_g": 0})),
    ("allow_missing_velocity is a number", cfg(allow_missing_velocity=1)),
    ("polaris sim_id is a string", cfg(polaris_writer={"sim_id": "a"})),
]

SYNTAX = [("invalid JSON", "{not json"), ("empty file", ""), ("trailing comma", '{"run_name": "x",}'),
          ("unterminated string", '{"run_name": "x'), ("second line", '{\n  "run_name": x\n}')]


def ref_error(cfg_text, tmp):
    """The reference CLI's error line (without the "[kiwi] error: " prefix) for a config."""
    path = os.path.join(tmp, "case.json")
    with open(path, "w") as f:
        f.write(cfg_text)
    r = subprocess.run([O.REF_CLI, "sph-to-grid", path], capture_output=True, text=True, cwd=tmp, timeout=60)
    lines = [ln for ln in r.stderr.splitlines() if ln.startswith("[kiwi] error: ")]
    assert r.returncode == 1 and len(lines) == 1, (r.returncode, r.stderr[-500:])
    return lines[0][len("[kiwi] error: "):], path


@pytest.fixture
def tmp():
    """A scratch directory that is also the working directory (the configs' paths are relative)."""
    d = tempfile.mkdtemp(prefix="kiwi_cfg_")
    with open(os.path.join(d, "good.out"), "w") as f:
        f.write("\n".join(GOOD_ROWS) + "\n")
    cwd = os.getcwd()
    os.chdir(d)
    yield d
    os.chdir(cwd)
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("name,c", CASES, ids=[c[0] for c in CASES])
def test_config_errors_equal_reference(tmp, name, c):
    ref, path = ref_error(json.dumps(c), tmp)
    with pytest.raises(io.ConfigError) as e:
        io.load_config(path)
    assert str(e.value) == ref


@pytest.mark.parametrize("name,text", SYNTAX, ids=[s[0] for s in SYNTAX])
def test_config_syntax_errors(tmp, name, text):
    ref, path = ref_error(text, tmp)
    with pytest.raises(io.ConfigError) as e:
        io.load_config(path)
    ours = str(e.value)
    head = re.match(r"(config file is not valid JSON \(.*\): \[json\.exception\.parse_error\.101\] parse error at "
                    r"line \d+), column", ref)
    assert head, ref
    assert ours.startswith(head.group(1) + ", column "), (ours, ref)


def test_missing_config_file(tmp):
    r = subprocess.run([O.REF_CLI, "sph-to-grid", os.path.join(tmp, "none.json")], capture_output=True, text=True)
    line = [ln for ln in r.stderr.splitlines() if ln.startswith("[kiwi] error: ")][0]
    with pytest.raises(io.ConfigError) as e:
        io.load_config(os.path.join(tmp, "none.json"))
    assert "[kiwi] error: " + str(e.value) == line


def run_ref(config, tmp, stage="sph-to-grid"):
    path = os.path.join(tmp, "run.json")
    with open(path, "w") as f:
        json.dump(config, f, indent=1)
    r = subprocess.run([O.REF_CLI, stage, path], capture_output=True, text=True, cwd=tmp, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert r.returncode == 0, r.stderr[-2000:]
    return path, r.stderr


def manifest_bytes(path):
    with open(path, "rb") as f:
        raw = f.read()
    return re.sub(rb'"written_at_unix": \d+', b'"written_at_unix": 0', raw)


VALID = [
    cfg(cartesian_grid={"resolution": 8}),
    cfg(cartesian_grid={"resolution": 6, "resolution_x": 4, "use_tabulated_integral": True},
        kernel={"name": "cubic-spline"}, grid_blocks={"block_scale_factor": 2.5}),
    cfg(cartesian_grid={"resolution": 5}, kernel={"name": "Wendland C4", "support_radius_in_h": 2.25, "hfact": 1.4},
        buffer={"enabled": True, "coverage_grid": {"spacing_fraction": 0.125}}, lloyd={"iterations": 2},
        phantom_units={"length_cm": 1e17}, allow_missing_velocity=True, diagnostics={"enabled": False},
        polaris_writer={"sim_id": 7, "gas_temperature_k": 10.5}, cartesian_map={"resolution": 33},
        mapping={"density_table": "t.bin"}),
    cfg(cartesian_grid={"resolution": 4}, buffer={"enabled": False, "strategy": "graded_shells",
                                                   "graded_shells": {"dx": 0.1, "dy": 0.2, "dz": 0.3}},
        domain={"x_min": -1e-5, "x_max": 123456789.125, "y_min": -0.0001, "y_max": 1e16, "z_min": 0.1,
                "z_max": 3}),
]


@pytest.mark.parametrize("k", range(len(VALID)))
def test_manifest_equals_reference(tmp, k):
    path, _ = run_ref(VALID[k], tmp)
    c = io.load_config(path)
    ref_manifest = os.path.join(tmp, c["run_dir"], "manifest.json")
    want = manifest_bytes(ref_manifest)
    assert json.loads(want)["config"] == json.loads(io.config_json(path))
    os.rename(ref_manifest, ref_manifest + ".ref")
    io.write_manifest(path)
    got = manifest_bytes(ref_manifest)
    assert got == want, (got.decode()[:1500], want.decode()[:1500])
    # the derived values the stages use
    kid = {"cubic_spline": 0, "cubic-spline": 0, "Wendland C4": 4}[VALID[k].get("kernel", {}).get("name", "cubic_spline")]
    assert c["kernel_id"] == kid
    res = VALID[k]["cartesian_grid"]
    assert (c["grid_nx"], c["grid_ny"], c["grid_nz"]) == tuple(res.get(f"resolution_{a}", 0) or res["resolution"]
                                                                for a in "xyz")


def test_manifest_large_input_and_changed_config_warning(tmp):
    # a plain_ascii snapshot above 64 MB: the fingerprint hashes 32 MB from each end and the size
    rng = np.random.default_rng(3)
    n = 480_000
    parts = {c: rng.uniform(-0.9, 0.9, n) for c in ("x", "y", "z")}
    parts.update({c: rng.normal(size=n) for c in ("vx", "vy", "vz")})
    parts["m"] = np.full(n, 1e-6)
    parts["h"] = np.full(n, 0.01)
    workloads.write_plain_ascii(parts, os.path.join(tmp, "big.out"))
    assert os.path.getsize(os.path.join(tmp, "big.out")) > (64 << 20)
    base = cfg(input_snapshot="big.out", cartesian_grid={"resolution": 4})
    path, _ = run_ref(base, tmp)
    c = io.load_config(path)
    ref_manifest = os.path.join(tmp, c["run_dir"], "manifest.json")
    want = manifest_bytes(ref_manifest)
    os.rename(ref_manifest, ref_manifest + ".ref")
    io.write_manifest(path)
    assert manifest_bytes(ref_manifest) == want
    # the same run directory with a changed config: the reference warns before loading its cached grid
    changed = cfg(input_snapshot="big.out", cartesian_grid={"resolution": 5}, lloyd={"iterations": 1})
    path2, err = run_ref(changed, tmp)
    ref_warn = [ln for ln in err.splitlines() if ln.startswith("[kiwi] WARNING") or ln.startswith("[kiwi]   ")
                or ln.startswith("[kiwi] Delete")]
    assert ref_warn
    # restore the manifest of the first run, then check the changed config against it here
    os.replace(ref_manifest + ".ref", ref_manifest)
    code = ("import sys; sys.path.insert(0, %r); from kiwi_amd import io; sys.exit(io.check_manifest(%r))"
            % (ROOT, path2))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=tmp)
    assert r.returncode == 1
    assert r.stderr.splitlines() == ref_warn
