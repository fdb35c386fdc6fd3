"""Input layer of the C++ run driver (langevina_amd/bin/lgv_superpions) on
the CPU: input.json keys, defaults and printout of ModelAData
(ModelG/ModelA.h:28-248) and the error paths of SuperPions.cxx main.  With
-quit the driver stops after printing the parameters, before any device
call; an unknown evolverType is rejected before the device is touched."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "langevina_amd", "bin", "lgv_superpions")


def run(tmp_path, text, *args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} missing: build it with python -c "
                    "'from langevina_amd.build import build_driver; build_driver()'")
    f = tmp_path / "input.json"
    f.write_text(text)
    r = subprocess.run([EXE, "-input", str(f), *args], capture_output=True, text=True,
                       timeout=60)
    return r.returncode, r.stdout, r.stderr


def params(stdout):
    out = {}
    for line in stdout.splitlines():
        if " = " in line:
            k, v = line.split(" = ", 1)
            out[k.strip()] = v.strip()
    return out


def test_defaults(tmp_path):
    """Only NX given: every other value is the default of ModelA.h."""
    rc, out, err = run(tmp_path, json.dumps({"NX": 16}), "-quit")
    assert rc == 0, err
    p = params(out)
    expect = {
        "NX": "16", "NY": "16", "NZ": "16", "LX": "1.600000e+01",   # ModelA.h:218-227
        "initialtime": "0.000000e+00", "deltat": "2.400000e-01",     # ModelA.h:33-35
        # finaltime 10 trimmed to whole saveFrequency*deltat blocks (ModelA.h:48-52):
        # 0.72 * int(10 / 0.72) = 9.36
        "finaltime": "9.360000e+00",
        "mass0": "-4.700520e+00", "dmassdt": "0.000000e+00",          # ModelA.h:73-74
        "lambda": "4.000000e+00", "H": "3.000000e-03", "chi": "5.000000e+00",
        "gamma": "1.000000e+00", "diffusion": "3.333333e-01",         # gamma / 3, ModelA.h:98
        "evolverType": "PV2HBSplit23", "seed": "10", "restart": "false",
        "outputfiletag": "o4output", "saveFrequency": "3",
        "thermalization_time": "0.000000e+00", "quench_mode": "false",
        "quench_mode_mass0": "-4.700520e+00", "eventmode": "false", "nevents": "1",
    }
    for k, v in expect.items():
        assert p.get(k) == v, (k, p.get(k), v)


def test_values_comments_and_trailing_commas(tmp_path):
    text = """{
        // jsoncpp accepts comments
        "NX": 32, "LX": 64, "finaltime": 10.01, "initialtime": 1.0, "deltat": 0.04,
        "mass0": -4.813, "dmassdt": 0.5, "lambda": 1, "H": 0.01, "chi": 2.0,
        "gamma": 2.0, "seed": 122335456, "saveFrequency": 5,
        "outputfiletag": "run1", "evolverType": "PV2HBSplit23NoDiffuse",
        "thermalization_time": 2.5, "quench_mode": true, "quench_mode_mass0": -5.0,
        "eventmode": true, "nevents": 4, "restart": false,
    }"""
    rc, out, err = run(tmp_path, text, "-quit")
    assert rc == 0, err
    p = params(out)
    assert p["LX"] == "6.400000e+01" and p["NZ"] == "32"
    # 1 + 0.2 * int(9.01 / 0.2) = 1 + 0.2 * 45 = 10
    assert p["finaltime"] == "1.000000e+01"
    assert p["mass0"] == "-4.813000e+00" and p["dmassdt"] == "5.000000e-01"
    assert p["diffusion"] == "6.666667e-01"
    assert p["evolverType"] == "PV2HBSplit23NoDiffuse" and p["seed"] == "122335456"
    assert p["quench_mode"] == "true" and p["quench_mode_mass0"] == "-5.000000e+00"
    assert p["eventmode"] == "true" and p["nevents"] == "4"
    assert p["thermalization_time"] == "2.500000e+00" and p["outputfiletag"] == "run1"


def test_explicit_diffusion(tmp_path):
    rc, out, _ = run(tmp_path, json.dumps({"NX": 8, "gamma": 3.0, "diffusion": 0.25}), "-quit")
    assert rc == 0 and params(out)["diffusion"] == "2.500000e-01"


@pytest.mark.parametrize("text, msg", [
    ("{}", "NX is required"),
    ('{"NX": 16, "deltat": 0.05,', "JSON parse error"),
    ('{"NX": "sixteen"}', "not convertible to Int"),
])
def test_bad_input_exits_with_an_error(tmp_path, text, msg):
    rc, out, err = run(tmp_path, text, "-quit")
    assert rc == 1 and msg in err


def test_unknown_stepper_is_rejected_before_any_device_call(tmp_path):
    rc, out, _ = run(tmp_path, json.dumps({"NX": 8, "evolverType": "RK4"}))
    assert rc == 0 and "Unrecognized stepper type RK4. Aborting..." in out


def test_missing_input_file(tmp_path):
    r = subprocess.run([EXE, "-input", str(tmp_path / "none.json")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0 and "Unable to open input file" in r.stdout
