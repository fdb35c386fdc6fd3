"""The C++ run driver (langevina_amd/bin/lgv_superpions, the control flow of
ModelG/SuperPions.cxx) on the GPU, against the oracle replaying the same
run_event loop: thermalization, measure-before-step at every saveFrequency,
the HDF5 layout of measurer_output.cxx / ntuple.h, and restart continuity.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from h5tools import H5DUMP, h5_read

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "langevina_amd", "bin", "lgv_superpions")
H5COPY = shutil.which("h5copy") or "/opt/conda/bin/h5copy"

PAR = {"NX": 8, "finaltime": 0.4, "initialtime": 0, "deltat": 0.04, "mass0": -4.813,
       "dmassdt": 0.3, "lambda": 1.0, "H": 0.01, "chi": 2.0, "gamma": 1,
       "evolverType": "PV2HBSplit23", "seed": 42, "outputfiletag": "run",
       "saveFrequency": 2, "thermalization_time": 0.08}


def log(*a):
    print(*a, flush=True)


def run_driver(tmp, params, name="in.json"):
    path = os.path.join(tmp, name)
    with open(path, "w") as f:
        f.write(json.dumps(params, indent=1)[:-2] + ",\n}\n")  # trailing comma, as input.json
    r = subprocess.run([DRIVER, "-input", path], cwd=tmp, capture_output=True, text=True,
                       timeout=120)
    log(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    return r.stdout


def oracle_run(orc, par, nsteps_run, lat=None, thermalize=True):
    """run_event (SuperPions.cxx:78-161) on the oracle: the clock reset,
    thermalize_event (with the quench of SuperPions.cxx:31-75 when
    quench_mode is set: mass0 = quench_mode_mass0 and dmassdt = 0 while
    thermalizing, then restored), then measure-before-step."""
    N = par["NX"]
    if lat is None:
        lat = orc.Lattice(N, mass0=par["mass0"], dmassdt=par["dmassdt"], lam=par["lambda"],
                          H=par["H"], chi=par["chi"], gamma=par["gamma"], seed=par["seed"],
                          rng_mode=orc.RNG_PER_SITE)
    dt = par["deltat"]
    t = float(par["initialtime"])
    lat.t = t
    if thermalize:
        quench = par.get("quench_mode", False)
        if quench:
            lat.c.mass0, lat.c.dmassdt = par["quench_mode_mass0"], 0.0
        lat.thermalize(dt, int(par["thermalization_time"] / dt))
        if quench:
            lat.c.mass0, lat.c.dmassdt = par["mass0"], par["dmassdt"]
    rows = []
    for steps in range(nsteps_run):
        if steps % par["saveFrequency"] == 0:
            w, wp, oa = orc.measure_walls(lat.u)
            rows.append((t, lat.c.mass0 + lat.c.dmassdt * t, oa, w, wp))
        lat.t = t
        lat.step(dt)
        t += dt
    return lat, rows


def test_driver_layout_and_values(tmp_path, orc):
    log("driver")
    tmp = str(tmp_path)
    out = run_driver(tmp, PAR)
    assert "Statistics for Phi HB Steps" in out and "Statistics for Charge HB Steps" in out
    h5 = os.path.join(tmp, "run.h5")
    N, nk = 8, 5
    nsteps = 10  # finaltime 0.4 / deltat 0.04, snapped to saveFrequency multiples
    lat, rows = oracle_run(orc, PAR, nsteps)
    phi = h5_read(h5, "/phi")
    timeout = h5_read(h5, "/timeout")
    assert phi.shape == (len(rows), 11) and timeout.shape == (len(rows), 2)
    for d in "xyz":
        assert h5_read(h5, f"/wall{d}_k").shape == (len(rows), 11, nk, 2)
        assert h5_read(h5, f"/wall{d}_k_rotated").shape == (len(rows), 14, nk, 2)
        assert h5_read(h5, f"/wall{d}_phase_k").shape == (len(rows), 14, nk, 2)
        hdr = subprocess.run([H5DUMP, "-H", "-d", f"/wall{d}", h5], capture_output=True,
                             text=True).stdout
        assert "( 0, 11, 8 ) / ( H5S_UNLIMITED, 11, 8 )" in hdr  # created, never filled
    wk = h5_read(h5, "/wallz_k")
    wkr = h5_read(h5, "/wallz_k_rotated")
    wpk = h5_read(h5, "/wallz_phase_k")
    for i, (t, m, oa, w, wp) in enumerate(rows):
        assert timeout[i, 0] == t and abs(timeout[i, 1] - m) <= 1e-15
        np.testing.assert_allclose(phi[i], oa, rtol=1e-10, atol=1e-13)
        k_or = orc.dft_rows(w[2])
        np.testing.assert_allclose(wk[i], k_or, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(wpk[i], orc.dft_rows(wp[2]), rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(wkr[i], orc.rotate_to_vev(k_or), rtol=1e-9, atol=1e-12)
    # final fields written by ModelA::finalize
    fin = h5_read(os.path.join(tmp, "run_save.h5"), "/o4fields")
    assert fin.shape == (N, N, N, 10)
    np.testing.assert_allclose(fin, lat.u, rtol=1e-10, atol=1e-12)
    s0, s1 = open(os.path.join(tmp, "run_0.rng")).read().split()
    assert int(s0) >= 0 and int(s1) >= 0

    # the offline pipeline of ModelG/x2k.py on this output: x2k of the (empty)
    # real-space wall, then x2k_rotated of the measured transform
    from test_x2k import X2K, X2KR, rotated_numpy
    r = subprocess.run([X2K, h5, "wallz"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out_h5 = os.path.join(tmp, "run_out.h5")
    assert h5_read(out_h5, "/wallz_k").shape == (0, 11, nk, 2)
    subprocess.run([H5COPY, "-i", h5, "-o", out_h5, "-s", "/wallz_k", "-d", "/wallz_k_meas"],
                   check=True, capture_output=True)
    r = subprocess.run([X2KR, out_h5, "wallz_k_meas"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rot = h5_read(out_h5, "/wallz_k_meas_rotated")
    ref = rotated_numpy(wk)
    np.testing.assert_allclose(rot[..., 0] + 1j * rot[..., 1], ref, rtol=0, atol=1e-13)


def test_driver_restart_continues_the_stream(tmp_path, orc):
    log("restart")
    tmp = str(tmp_path)
    run_driver(tmp, PAR)
    par2 = dict(PAR, restart=True)
    run_driver(tmp, par2, "in2.json")
    h5 = os.path.join(tmp, "run.h5")
    phi = h5_read(h5, "/phi")
    assert phi.shape[0] == 10  # 5 rows appended to 5 rows
    # oracle: the same two segments; the restarted run resets the clock to
    # initialtime and skips thermalization (SuperPions.cxx:82-86)
    lat, rows1 = oracle_run(orc, PAR, 10)
    lat, rows2 = oracle_run(orc, PAR, 10, lat=lat, thermalize=False)
    fin = h5_read(os.path.join(tmp, "run_save.h5"), "/o4fields")
    np.testing.assert_allclose(fin, lat.u, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(phi[5:], np.array([r[2] for r in rows2]), rtol=1e-10, atol=1e-13)


def test_driver_event_and_quench_modes(tmp_path, orc):
    log("events")
    tmp = str(tmp_path)
    par = dict(PAR, eventmode=True, nevents=2, quench_mode=True, quench_mode_mass0=-4.0)
    out = run_driver(tmp, par)
    assert os.path.exists(os.path.join(tmp, "run_0000.h5"))
    assert os.path.exists(os.path.join(tmp, "run_0001.h5"))
    assert not os.path.exists(os.path.join(tmp, "run_save.h5"))  # no save in event mode
    assert out.count("Settinng up a quench initial condition") == 2
    assert h5_read(os.path.join(tmp, "run_0001.h5"), "/phi").shape == (5, 11)
    # the oracle replays both events: each thermalizes the previous event's
    # final fields at the quench mass (fresh charges), then runs at the
    # nominal m^2(t); the measured rows and k-space walls must agree
    lat = None
    for ev in range(2):
        lat, rows = oracle_run(orc, par, 10, lat=lat)
        h5 = os.path.join(tmp, f"run_{ev:04d}.h5")
        phi = h5_read(h5, "/phi")
        timeout = h5_read(h5, "/timeout")
        wk = h5_read(h5, "/wallz_k")
        assert phi.shape == (len(rows), 11)
        for i, (t, m, oa, w, wp) in enumerate(rows):
            assert timeout[i, 0] == t and abs(timeout[i, 1] - m) <= 1e-15, (ev, i)
            np.testing.assert_allclose(phi[i], oa, rtol=1e-10, atol=1e-13, err_msg=f"event {ev}")
            np.testing.assert_allclose(wk[i], orc.dft_rows(w[2]), rtol=1e-10, atol=1e-13)
    # the quench changed the thermalized state: the same run without it differs
    lat_nq, rows_nq = oracle_run(orc, dict(par, quench_mode=False), 10)
    phi0 = h5_read(os.path.join(tmp, "run_0000.h5"), "/phi")
    assert np.abs(phi0[0] - rows_nq[0][2]).max() > 1e-6
