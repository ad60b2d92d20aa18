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
