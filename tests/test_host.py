"""The C++ host driver (parallel-critical-collapse_amd/cc_newton) without a GPU:
it reads the reference's simulation JSON keys (SimulationConfig.hpp:73-137)
and rejects bad input the way the reference does."""
import json
import os
import subprocess

import numpy as np
import pytest

from _helpers import ROOT, W

CLI = os.path.join(ROOT, "parallel-critical-collapse_amd", "cc_newton")


def config_dict(shape, Up, psic, fc, Delta, **over):
    d = {"Ntau": shape.Ntau, "Dim": shape.Dim, "XLeft": shape.XLeft, "XMid": shape.XMid,
         "XRight": shape.XRight, "EpsNewton": shape.EpsNewton, "PrecisionNewton": 1e-10,
         "MaxIterNewton": 10, "Verbose": True, "SlowError": shape.SlowError, "NLeft": shape.NLeft,
         "NRight": shape.NRight, "SchemeIRK": 2, "PrecisionIRK": shape.PrecisionIRK,
         "MaxIterIRK": shape.MaxIterIRK, "Converged": False, "Debug": False,
         "Initial_Conditions": {"Delta": Delta, "fc": list(fc), "psic": list(psic), "Up": list(Up)}}
    d.update(over)
    return d


def run(args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


@pytest.fixture
def cfg_file(tmp_path):
    shape = W.Shape(Ntau=128, X=1000)
    rng = np.random.default_rng(1)
    Up, psic, fc = (rng.standard_normal(128) * s for s in (1e-3, 1e-2, 1.0))
    d = config_dict(shape, Up, psic, fc, 3.4453)
    p = tmp_path / "sim.json"
    p.write_text(json.dumps(d, indent=4))
    return p, d


def test_dry_run_reads_every_key(cfg_file):
    p, d = cfg_file
    r = run(["--dry-run", "-i", str(p)])
    assert r.returncode == 0, r.stderr
    back = json.loads(r.stdout)
    assert back.keys() == d.keys()
    for k, v in d.items():
        if k != "Initial_Conditions":
            assert back[k] == v, k
    for k in ("fc", "psic", "Up"):
        assert back["Initial_Conditions"][k] == d["Initial_Conditions"][k]       # exact round trip
    assert back["Initial_Conditions"]["Delta"] == d["Initial_Conditions"]["Delta"]


@pytest.mark.parametrize("edit,msg", [
    (lambda d: d.pop("NLeft"), "lacks key 'NLeft'"),
    (lambda d: d.update(SchemeIRK=4), "Wrong IRK Scheme stage supplied!"),
])
def test_bad_input_is_rejected(tmp_path, cfg_file, edit, msg):
    _, d = cfg_file
    edit(d)
    p = tmp_path / "bad.json"
    p.write_text(json.dumps(d))
    r = run(["--dry-run", "-i", str(p)])
    assert r.returncode == 1 and msg in r.stderr


def test_missing_input_path():
    r = run(["-i", "/nonexistent/sim.json"])
    assert r.returncode == 1 and "Invalid simulation input path!" in r.stderr


# --- multiple run (main.cpp:236-378, SimulationSuite in SimulationConfig.hpp:213-315) ---

def suite_member(D, converged, fields=None, N=8):
    shape = W.Shape(Ntau=N, X=200, Dim=D)
    d = config_dict(shape, [], [], [], 0.0, Converged=converged)
    d["Initial_Conditions"] = fields if fields is not None else {}
    return d


def quadratic_fields(D, N=8):
    """Fields and Delta that are quadratic in D: three-point quadratic
    extrapolation reproduces them exactly."""
    j = np.arange(N)
    return {"fc": list(0.2 + 0.01 * D * D + 0.1 * np.cos(j) * D), "psic": list(np.sin(j) * (D - 3.0) ** 2),
            "Up": list(np.cos(2 * j) * (1.0 + 0.5 * D)), "Delta": 3.4 + 0.02 * (D - 4) - 0.003 * (D - 4) ** 2}


def dry_suite(tmp_path, suite, *flags):
    p, o = tmp_path / "suite.json", tmp_path / "planned.json"
    p.write_text(json.dumps(suite))
    r = run(["-m", "--dry-run", "-i", str(p), "-o", str(o), *flags])
    assert r.returncode == 0, r.stderr
    order = [l.split("D=")[1].rstrip(".") for l in r.stdout.splitlines() if l.startswith("Starting simulation for D=")]
    return order, json.loads(o.read_text()), r.stdout


def test_suite_order(tmp_path):
    """String-sorted keys, descending by default, ascending with -r; of the
    converged dimensions only the last three of that order run first, as
    seeds; --ignore-converged runs every key in order."""
    conv = ["3.900", "3.950", "4.000", "4.050"]
    todo = ["4.100", "3.800", "4.200"]
    suite = {k: suite_member(float(k), True, quadratic_fields(float(k))) for k in conv}
    suite.update({k: suite_member(float(k), False) for k in todo})
    order, _, _ = dry_suite(tmp_path, suite)
    # descending: converged sorted descending [4.050, 4.000, 3.950, 3.900], last three seeded
    assert order == ["4.000", "3.950", "3.900", "4.200", "4.100", "3.800"]
    order, _, _ = dry_suite(tmp_path, suite, "-r")
    assert order == ["3.950", "4.000", "4.050", "3.800", "4.100", "4.200"]
    order, _, _ = dry_suite(tmp_path, suite, "--ignore-converged")
    assert order == sorted(conv + todo, reverse=True)
    order, _, _ = dry_suite(tmp_path, suite, "--ignore-converged", "-r")
    assert order == sorted(conv + todo)


def test_suite_initial_data_copy_and_extrapolation(tmp_path):
    """Second and third dimensions copy the previous converged solution; from
    the fourth on, data quadratic in D is extrapolated exactly."""
    suite = {"4.000": suite_member(4.0, True, quadratic_fields(4.0)),
             "4.050": suite_member(4.05, True, quadratic_fields(4.05)),
             "4.100": suite_member(4.1, True, quadratic_fields(4.1)),
             "4.150": suite_member(4.15, False),
             "4.200": suite_member(4.2, False)}
    order, planned, out = dry_suite(tmp_path, suite, "-r")
    assert order == ["4.000", "4.050", "4.100", "4.150", "4.200"]
    ic = planned["4.150"]["Initial_Conditions"]
    exact = quadratic_fields(4.15)
    for k in ("fc", "psic", "Up"):
        assert np.max(np.abs(np.array(ic[k]) - np.array(exact[k]))) < 1e-13, k
    assert abs(ic["Delta"] - exact["Delta"]) < 1e-14
    assert "Previous dimensions not converged" in out          # 4.200 needs 4.150 solved first
    # copy: a single converged seed, then the next dimension borrows its data
    suite = {"4.000": suite_member(4.0, True, quadratic_fields(4.0)), "4.010": suite_member(4.01, False)}
    _, planned, _ = dry_suite(tmp_path, suite, "-r")
    assert planned["4.010"]["Initial_Conditions"] == {**quadratic_fields(4.0)}


def test_suite_first_dimension_needs_data(tmp_path):
    p = tmp_path / "suite.json"
    p.write_text(json.dumps({"4.000": suite_member(4.0, False)}))
    r = run(["-m", "-i", str(p)])
    assert r.returncode != 0
    assert "No converged solution as initial conditions available for 4.000!" in r.stderr


LAYERS = os.path.join(ROOT, "parallel-critical-collapse_amd", "cc_layers")


def run_layers(tmp_path, text):
    p, q = tmp_path / "req.json", tmp_path / "out.json"
    p.write_text(text)
    r = subprocess.run([LAYERS, str(p), str(q)], capture_output=True, text=True, timeout=60)
    return r, (json.loads(q.read_text())["results"] if r.returncode == 0 else None)


def test_layers_driver_rejects_bad_requests(tmp_path):
    """cc_layers (the SpectralTransformer / ODEStepper / ShootingSolver driver)
    checks its request before any device work: a malformed file, a missing
    call list or an IRK scheme other than 1, 2, 3 end the run; a transform
    length the library lacks and an unknown method fail their call only."""
    for text, msg in (("{", "json"), (json.dumps({"Ntau": 128}), "calls"),
                      (json.dumps({"Ntau": 128, "SchemeIRK": 4, "calls": []}), "Wrong IRK Scheme")):
        r, _ = run_layers(tmp_path, text)
        assert r.returncode != 0 and msg in r.stderr, (text, r.stderr)
    calls = [{"method": "differentiate", "in": [1.0, 0.0]}, {"method": "nope"}]
    r, res = run_layers(tmp_path, json.dumps({"Ntau": 100, "calls": calls}))
    assert r.returncode == 0
    assert res[0]["status"] == 4 and "Ntau must be 128, 256, 512, 1024 or 2048" in res[0]["error"]
    assert res[1]["status"] == 4 and "unknown method 'nope'" in res[1]["error"]
    r, _ = run_layers(tmp_path, "{}")
    assert r.returncode != 0
    r = subprocess.run([LAYERS], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "usage" in r.stderr
