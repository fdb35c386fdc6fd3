"""The C-ABI library (include/lgv.h) on a machine without a GPU: it loads,
exports every entry point the header declares, and rejects bad arguments
with the documented codes before touching a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lgv.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(lgv_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def L():
    from langevina_amd import lgv
    if not os.path.exists(lgv.LIB_PATH):
        pytest.fail(f"{lgv.LIB_PATH} missing: build it with python -m langevina_amd.build")
    return lgv


def test_header_declares_the_interface():
    names = declared()
    for must in ("lgv_create", "lgv_step", "lgv_run", "lgv_measure_full", "lgv_em_forward",
                 "lgv_em_backward", "lgv_rng_get_state"):
        assert must in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol(L):
    lib = L.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(L.EXPORTS) == declared()


def test_version_and_error_text(L):
    lib = L.lib()
    assert lib.lgv_version() >= 1
    assert isinstance(lib.lgv_last_error(), bytes)


def _create(L, params, **kw):
    from langevina_amd.lgv import Config, ModelParams
    arr = (ModelParams * len(params))(*params)
    cfg = Config()
    cfg.nreplicas = len(params)
    cfg.replicas = C.cast(arr, C.POINTER(ModelParams))
    cfg.precision = kw.get("precision", 0)
    cfg.scounts[0], cfg.scounts[1] = kw.get("scounts", (2, 3))
    cfg.device = kw.get("device", 0)
    h = C.c_void_p()
    rc = L.lib().lgv_create(C.byref(cfg), C.byref(h))
    return rc, L.lib().lgv_last_error().decode()


@pytest.mark.parametrize("bad, text", [
    (dict(NX=15), "even"),
    (dict(NX=2), "even"),
    (dict(chi=0.0), "chi"),
])
def test_create_rejects_bad_replicas(L, bad, text):
    p = L.model_params(16)
    for k, v in bad.items():
        setattr(p, k, v)
    rc, msg = _create(L, [p])
    assert rc == -1 and text in msg


def test_create_rejects_bad_config(L):
    rc, msg = _create(L, [L.model_params(16)], precision=7)
    assert rc == -1 and "precision" in msg
    rc, msg = _create(L, [L.model_params(16)], scounts=(2, 100))
    assert rc == -1 and "scounts" in msg
    lib = L.lib()
    assert lib.lgv_create(None, None) == -1
    assert lib.lgv_step(None, C.c_double(0.04)) == -1


def test_create_failure_keeps_the_first_error(L):
    """A HIP failure inside lgv_create (here: a device index that does not
    exist) returns LGV_ERR_HIP with the failing call in the message; the
    partly built context is released without overwriting that message."""
    rc, msg = _create(L, [L.model_params(16)], device=999)
    assert rc == -2 and "hipSetDevice" in msg, msg


def test_em_create_rejects_bad_arguments(L):
    from langevina_amd.em import EmParams, _bind
    lib = _bind()
    p = EmParams()
    p.N, p.dt = 2, 0.02
    h = C.c_void_p()
    assert lib.lgv_em_create(C.byref(p), 10, 5, 0, C.byref(h)) == -1
    assert b"N=2" in lib.lgv_em_last_error()
    p.N = 16
    assert lib.lgv_em_create(C.byref(p), 0, 5, 0, C.byref(h)) == -1
    p.noise_kind = 3
    assert lib.lgv_em_create(C.byref(p), 10, 5, 0, C.byref(h)) == -1
    assert b"noise_kind" in lib.lgv_em_last_error()


def test_em_step_entry_points_reject_bad_arguments(L):
    """The single-step entry points check their arguments before any device
    call: null handle or buffers, in == out, a negative step, and (the entry
    point without allocation) a lattice beyond its static scratch array."""
    from langevina_amd.em import EmParams, _bind
    lib = _bind()
    buf = (C.c_double * 8)()
    a, b = C.cast(buf, C.c_void_p), C.c_void_p(C.addressof(buf) + 32)
    assert lib.lgv_em_step(None, a, b, 0, None) == -1
    assert lib.lgv_em_series(None, a, 10, None, a) == -1
    p = EmParams()
    p.N, p.dt, p.gamma, p.noise_kind = 258, 0.01, 1.0, 0
    assert lib.lgv_em_step_noalloc(C.byref(p), a, b, 0, None, 0, None) == -4
    assert b"N=258" in lib.lgv_em_last_error()
    p.N = 16
    assert lib.lgv_em_step_noalloc(C.byref(p), a, a, 0, None, 0, None) == -1
    assert b"differ" in lib.lgv_em_last_error()
    assert lib.lgv_em_step_noalloc(C.byref(p), a, b, -1, None, 0, None) == -1
    p.noise_kind = 5
    assert lib.lgv_em_step_noalloc(C.byref(p), a, b, 0, None, 0, None) == -1
    assert lib.lgv_em_step_noalloc(None, a, b, 0, None, 0, None) == -1


def test_tile_and_graph_setters_reject_a_null_context(L):
    lib = L.lib()
    assert lib.lgv_set_sweep_tiles(None, 0, 0, 0, 0, 0) == -1
    assert lib.lgv_set_ideal_tile(None, 0, 32, 16) == -1
    assert lib.lgv_use_graph(None, 1) == -1
    assert lib.lgv_set_phi_kernel(None, 0) == -1
    assert lib.lgv_timing_enable(None, 1) == -1
    k = C.c_int(-5)
    assert lib.lgv_get_phi_kernel(None, C.byref(k)) == -1
