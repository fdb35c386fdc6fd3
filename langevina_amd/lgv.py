"""ctypes binding of liblgv_hip.so (include/lgv.h).

There is no CPU fallback: if the HIP library is missing or fails to load,
every entry point raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblgv_hip.so")

FP64 = 0
FP32 = 1
MON_PHI, MON_CHARGE, MON_IDEAL = 0, 1, 2


class LgvError(RuntimeError):
    pass


class ModelParams(C.Structure):
    """lgv_model_params: ModelAData + ModelACoefficients + seed
    (ModelG/ModelA.h:71-111, 185-248)."""
    _fields_ = [("NX", C.c_int), ("LX", C.c_int),
                ("mass0", C.c_double), ("dmassdt", C.c_double),
                ("lam", C.c_double), ("H", C.c_double), ("chi", C.c_double),
                ("gamma", C.c_double), ("diffusion", C.c_double),
                ("seed", C.c_int), ("initialtime", C.c_double)]


class Config(C.Structure):
    _fields_ = [("nreplicas", C.c_int), ("replicas", C.POINTER(ModelParams)),
                ("precision", C.c_int), ("scounts", C.c_uint * 2),
                ("nodiffuse", C.c_int), ("onlydiffuse", C.c_int),
                ("track_energy", C.c_int), ("device", C.c_int)]


class Monitor(C.Structure):
    _fields_ = [("down_counts", C.c_int64), ("up_counts", C.c_int64),
                ("up_yes", C.c_int64), ("up_no", C.c_int64),
                ("down_dS", C.c_double), ("up_dS", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RngState(C.Structure):
    _fields_ = [("s0", C.c_uint64), ("s1", C.c_uint64), ("saved_normal", C.c_double),
                ("has_saved_normal", C.c_int), ("sweep_ctr", C.c_uint64)]


EXPORTS = [
    "lgv_last_error", "lgv_version", "lgv_create", "lgv_destroy", "lgv_set_field",
    "lgv_get_field", "lgv_set_time", "lgv_get_time", "lgv_step", "lgv_run",
    "lgv_use_graph", "lgv_ideal_step", "lgv_hb_phi_step", "lgv_charge_step",
    "lgv_init_gaussian_charges", "lgv_thermalize", "lgv_monitor_get", "lgv_energy_get",
    "lgv_rng_get_state", "lgv_rng_set_state", "lgv_measure", "lgv_synchronize",
    "lgv_field_device_ptr", "lgv_timing_enable", "lgv_timing_get", "lgv_measure_full",
    "lgv_set_mass", "lgv_monitor_reset", "lgv_set_ideal_tile", "lgv_set_sweep_tiles",
    "lgv_em_last_error", "lgv_em_create", "lgv_em_set_stream", "lgv_em_forward",
    "lgv_em_backward", "lgv_em_destroy", "lgv_em_series", "lgv_em_step",
    "lgv_em_step_noalloc", "lgv_set_phi_kernel", "lgv_get_phi_kernel",
]

# lgv_set_phi_kernel values
PHI_AUTO, PHI_SITE, PHI_MARCH = 0, 1, 2


class Measurement(C.Structure):
    """lgv_measurement: output pointers of lgv_measure_full (NULL = skip)."""
    _fields_ = [("wall", C.POINTER(C.c_double)), ("wallphase", C.POINTER(C.c_double)),
                ("oavg", C.POINTER(C.c_double)), ("wall_k", C.POINTER(C.c_double)),
                ("wall_k_rotated", C.POINTER(C.c_double)),
                ("wallphase_k", C.POINTER(C.c_double))]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LgvError(f"{LIB_PATH} is not built: run python -m langevina_amd.build")
    L = C.CDLL(LIB_PATH)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ctx = C.c_void_p
    L.lgv_last_error.restype = C.c_char_p
    L.lgv_create.argtypes = [C.POINTER(Config), C.POINTER(ctx)]
    L.lgv_destroy.argtypes = [ctx]
    L.lgv_set_field.argtypes = [ctx, C.c_int, dp]
    L.lgv_get_field.argtypes = [ctx, C.c_int, dp]
    L.lgv_set_time.argtypes = [ctx, C.c_int, C.c_double]
    L.lgv_get_time.argtypes = [ctx, C.c_int, C.POINTER(C.c_double)]
    L.lgv_step.argtypes = [ctx, C.c_double]
    L.lgv_run.argtypes = [ctx, C.c_double, C.c_int]
    L.lgv_use_graph.argtypes = [ctx, C.c_int]
    L.lgv_ideal_step.argtypes = [ctx, C.c_double]
    L.lgv_hb_phi_step.argtypes = [ctx, C.c_double]
    L.lgv_charge_step.argtypes = [ctx, C.c_double]
    L.lgv_init_gaussian_charges.argtypes = [ctx, C.c_int]
    L.lgv_thermalize.argtypes = [ctx, C.c_double, C.c_int]
    L.lgv_monitor_get.argtypes = [ctx, C.c_int, C.POINTER(Monitor)]
    L.lgv_energy_get.argtypes = [ctx, C.c_int, dp]
    L.lgv_rng_get_state.argtypes = [ctx, C.c_int, C.POINTER(RngState)]
    L.lgv_rng_set_state.argtypes = [ctx, C.c_int, C.POINTER(RngState)]
    L.lgv_measure.argtypes = [ctx, C.c_int, dp, dp, dp]
    L.lgv_timing_enable.argtypes = [ctx, C.c_int]
    L.lgv_timing_get.argtypes = [ctx, dp, np.ctypeslib.ndpointer(dtype=np.int64,
                                                                 flags="C_CONTIGUOUS")]
    L.lgv_measure_full.argtypes = [ctx, C.c_int, C.POINTER(Measurement)]
    L.lgv_set_mass.argtypes = [ctx, C.c_int, C.c_double, C.c_double]
    L.lgv_monitor_reset.argtypes = [ctx]
    L.lgv_set_ideal_tile.argtypes = [ctx, C.c_int, C.c_int, C.c_int]
    L.lgv_set_sweep_tiles.argtypes = [ctx, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.lgv_synchronize.argtypes = [ctx]
    L.lgv_field_device_ptr.argtypes = [ctx, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise LgvError(f"lgv error {rc}: {lib().lgv_last_error().decode()}")


def model_params(NX, LX=0, mass0=-4.70052, dmassdt=0.0, lam=4.0, H=0.003, chi=5.0,
                 gamma=1.0, diffusion=None, seed=10, initialtime=0.0):
    """Defaults are ModelACoefficients' and ModelAHandlerData's
    (ModelG/ModelA.h:73-81, 119); diffusion defaults to gamma/3 (ModelA.h:98)."""
    p = ModelParams()
    p.NX, p.LX = NX, LX
    p.mass0, p.dmassdt, p.lam, p.H, p.chi, p.gamma = mass0, dmassdt, lam, H, chi, gamma
    p.diffusion = 1.0 / 3.0 * gamma if diffusion is None else diffusion
    p.seed, p.initialtime = seed, initialtime
    return p


def _xoro_uniforms(seed, n):
    """n uniform() draws (NoiseGenerator::uniform, NoiseGenerator.h:68-69) of
    a xoroshiro128+ engine seeded by xoroshiro128plus::seed(seed)
    (xoroshiro128plus.h:86-107): generate_canonical of one 64-bit output."""
    m = 2 ** 64
    s0 = (0x01d353e5f3993bb0 + seed) % m
    s1 = (0x7b9c0df6cb193b20 * (seed + 1)) % m
    out = []
    for _ in range(n):
        res = (s0 + s1) % m
        t = s1 ^ s0
        s0 = (((s0 << 55) | (s0 >> 9)) % m) ^ t ^ ((t << 14) % m)
        s1 = ((t << 36) | (t >> 28)) % m
        u = float(res) / m
        out.append(u if u < 1.0 else float.fromhex("0x1.fffffffffffffp-1"))
    return out


def ensemble_params(sizes, seed0=1000, H_range=(0.003, 0.03), H_seed=12345, **par):
    """Parameters of a batch of independent replicas: replica r has side
    sizes[r], seed seed0 + r and H uniform in H_range, drawn from one
    xoroshiro128+ stream seeded with H_seed; par sets the other couplings."""
    lo, hi = H_range
    us = _xoro_uniforms(H_seed, len(sizes))
    return [model_params(N, seed=seed0 + r, **dict(par, H=lo + (hi - lo) * us[r]))
            for r, N in enumerate(sizes)]


class Context:
    """One GPU context holding one or more replicas stepped together."""

    def __init__(self, replicas, precision=FP64, scounts=(2, 3), nodiffuse=False,
                 onlydiffuse=False, track_energy=False, device=0):
        if isinstance(replicas, ModelParams):
            replicas = [replicas]
        self._params = (ModelParams * len(replicas))(*replicas)
        cfg = Config()
        cfg.nreplicas = len(replicas)
        cfg.replicas = C.cast(self._params, C.POINTER(ModelParams))
        cfg.precision = precision
        cfg.scounts[0], cfg.scounts[1] = scounts
        cfg.nodiffuse, cfg.onlydiffuse = int(nodiffuse), int(onlydiffuse)
        cfg.track_energy = int(track_energy)
        cfg.device = device
        self.h = C.c_void_p()
        _check(lib().lgv_create(C.byref(cfg), C.byref(self.h)))
        self.N = [p.NX for p in replicas]
        self.nrep = len(replicas)
        self.precision = precision

    def close(self):
        if self.h:
            lib().lgv_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # fields in the o4fields layout [z][y][x][10]
    def set_field(self, u, replica=0):
        N = self.N[replica]
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (N, N, N, 10):
            raise ValueError(f"field shape {u.shape} != {(N, N, N, 10)}")
        _check(lib().lgv_set_field(self.h, replica, u.reshape(-1)))

    def get_field(self, replica=0):
        N = self.N[replica]
        u = np.zeros((N, N, N, 10))
        _check(lib().lgv_get_field(self.h, replica, u.reshape(-1)))
        return u

    def set_time(self, t, replica=-1):
        _check(lib().lgv_set_time(self.h, replica, t))

    def time(self, replica=0):
        t = C.c_double()
        _check(lib().lgv_get_time(self.h, replica, C.byref(t)))
        return t.value

    def step(self, dt):
        _check(lib().lgv_step(self.h, dt))

    def run(self, dt, nsteps):
        _check(lib().lgv_run(self.h, dt, nsteps))

    def use_graph(self, enable=True):
        """False: stream launches; True or K: replay a graph of K steps."""
        _check(lib().lgv_use_graph(self.h, int(enable)))

    def ideal_step(self, dt):
        _check(lib().lgv_ideal_step(self.h, dt))

    def hb_phi_step(self, dt):
        _check(lib().lgv_hb_phi_step(self.h, dt))

    def charge_step(self, dt):
        _check(lib().lgv_charge_step(self.h, dt))

    def init_gaussian_charges(self, replica=0):
        _check(lib().lgv_init_gaussian_charges(self.h, replica))

    def thermalize(self, dt, nsteps):
        _check(lib().lgv_thermalize(self.h, dt, nsteps))

    def monitors(self, replica=0):
        m = (Monitor * 3)()
        _check(lib().lgv_monitor_get(self.h, replica, m))
        return {"phi": m[0].as_dict(), "charge": m[1].as_dict(), "ideal": m[2].as_dict()}

    def energy(self, replica=0):
        e = np.zeros(2)
        _check(lib().lgv_energy_get(self.h, replica, e))
        return e

    def rng_state(self, replica=0):
        s = RngState()
        _check(lib().lgv_rng_get_state(self.h, replica, C.byref(s)))
        return s

    def set_rng_state(self, state, replica=0):
        _check(lib().lgv_rng_set_state(self.h, replica, C.byref(state)))

    def measure(self, replica=0):
        N = self.N[replica]
        wall = np.zeros((3, 11, N))
        wallphase = np.zeros((3, 14, N))
        oavg = np.zeros(11)
        _check(lib().lgv_measure(self.h, replica, wall.reshape(-1), wallphase.reshape(-1), oavg))
        return wall, wallphase, oavg

    def measure_full(self, replica=0):
        """Measurer::measure: walls, OAverage and the k-space walls
        (wall_k, wall_k_rotated, wallphase_k as complex arrays, k = 0..N/2)."""
        N = self.N[replica]
        K = N // 2 + 1
        out = {"wall": np.zeros((3, 11, N)), "wallphase": np.zeros((3, 14, N)),
               "oavg": np.zeros(11), "wall_k": np.zeros((3, 11, K, 2)),
               "wall_k_rotated": np.zeros((3, 14, K, 2)), "wallphase_k": np.zeros((3, 14, K, 2))}
        m = Measurement(*[v.ctypes.data_as(C.POINTER(C.c_double)) for v in out.values()])
        _check(lib().lgv_measure_full(self.h, replica, C.byref(m)))
        for k in ("wall_k", "wall_k_rotated", "wallphase_k"):
            out[k] = out[k][..., 0] + 1j * out[k][..., 1]
        return out

    def set_mass(self, mass0, dmassdt=0.0, replica=-1):
        _check(lib().lgv_set_mass(self.h, replica, mass0, dmassdt))

    def monitor_reset(self):
        _check(lib().lgv_monitor_reset(self.h))

    def synchronize(self):
        _check(lib().lgv_synchronize(self.h))

    def set_phi_kernel(self, kernel):
        """PHI_AUTO (size rule), PHI_SITE or PHI_MARCH for the phi half-sweeps."""
        _check(lib().lgv_set_phi_kernel(self.h, int(kernel)))

    def phi_kernel(self):
        """The phi half-sweep kernel in use: PHI_SITE or PHI_MARCH."""
        k = C.c_int()
        _check(lib().lgv_get_phi_kernel(self.h, C.byref(k)))
        return k.value

    def timing_enable(self, enable=True):
        """False: off; True: time every step; k > 1: every k-th step."""
        _check(lib().lgv_timing_enable(self.h, int(enable)))

    def timing(self):
        """Per kernel class (ideal, phi, charge, energy): total ms and launches."""
        ms = np.zeros(4)
        n = np.zeros(4, dtype=np.int64)
        _check(lib().lgv_timing_get(self.h, ms, n))
        names = ("ideal", "phi", "charge", "energy")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(names)}

    def field_device_ptr(self, replica=0):
        p, e = C.c_void_p(), C.c_int()
        _check(lib().lgv_field_device_ptr(self.h, replica, C.byref(p), C.byref(e)))
        return p.value, e.value
