"""ctypes front end of the CPU oracle (oracle/lgv_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the langevina_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblgv_oracle.so")
REF_RNG = os.path.join(HERE, "_ref", "ref_rng")

RNG_REFERENCE = 0
RNG_PER_SITE = 1
FP64 = 0
FP32 = 1


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Xoro(C.Structure):
    _fields_ = [("s0", C.c_uint64), ("s1", C.c_uint64)]


class Noise(C.Structure):
    _fields_ = [("g", Xoro), ("saved", C.c_double), ("has_saved", C.c_int)]


class Monitor(C.Structure):
    _fields_ = [("down_counts", C.c_int64), ("up_counts", C.c_int64),
                ("up_yes", C.c_int64), ("up_no", C.c_int64),
                ("down_dS", C.c_double), ("up_dS", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Ctx(C.Structure):
    _fields_ = [("N", C.c_int), ("hx", C.c_double),
                ("mass0", C.c_double), ("dmassdt", C.c_double),
                ("lam", C.c_double), ("H", C.c_double), ("chi", C.c_double),
                ("gamma", C.c_double), ("diffusion", C.c_double),
                ("seed", C.c_int), ("rng_mode", C.c_int), ("precision", C.c_int),
                ("scounts", C.c_uint * 2), ("nodiffuse", C.c_int),
                ("onlydiffuse", C.c_int), ("t", C.c_double),
                ("sweep_ctr", C.c_uint64), ("base_key", C.c_uint64),
                ("noise", Noise), ("mon_phi", Monitor), ("mon_charge", Monitor),
                ("mon_ideal", Monitor), ("last_energy", C.c_double * 2),
                ("last_order", C.c_int * 5)]


class EmParams(C.Structure):
    _fields_ = [("N", C.c_int), ("m2", C.c_double), ("lam", C.c_double),
                ("H", C.c_double), ("gamma", C.c_double), ("dt", C.c_double),
                ("base_key", C.c_uint64), ("noise_kind", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.orc_xoro_seed.argtypes = [C.POINTER(Xoro), C.c_uint64]
        L.orc_xoro_next.argtypes = [C.POINTER(Xoro)]
        L.orc_xoro_next.restype = C.c_uint64
        L.orc_canonical.argtypes = [C.c_uint64]
        L.orc_canonical.restype = C.c_double
        L.orc_uniform.argtypes = [C.POINTER(Xoro)]
        L.orc_uniform.restype = C.c_double
        L.orc_variance1.argtypes = [C.POINTER(Xoro)]
        L.orc_variance1.restype = C.c_double
        L.orc_seed_seq_generate.argtypes = [C.POINTER(C.c_uint32), C.c_size_t,
                                            C.POINTER(C.c_uint32), C.c_size_t]
        L.orc_uniform_int.argtypes = [C.POINTER(Xoro), C.c_uint64]
        L.orc_uniform_int.restype = C.c_uint64
        L.orc_shuffle.argtypes = [C.POINTER(Xoro), C.POINTER(C.c_int), C.c_int]
        L.orc_noise_init.argtypes = [C.POINTER(Noise), C.c_int]
        L.orc_normal.argtypes = [C.POINTER(Noise)]
        L.orc_normal.restype = C.c_double
        for f in ("orc_mix64",):
            getattr(L, f).argtypes = [C.c_uint64]
            getattr(L, f).restype = C.c_uint64
        L.orc_one_site_chain.argtypes = [C.c_int, C.c_long, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_uint32, dp]
        L.orc_div_by_const_mismatches.argtypes = [C.c_double, C.c_long, C.c_uint64]
        L.orc_div_by_const_mismatches.restype = C.c_long
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_base_key.argtypes = [C.c_uint32]
        L.orc_base_key.restype = C.c_uint64
        L.orc_stream_key.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_stream_key.restype = C.c_uint64
        L.orc_site_key.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_site_key.restype = C.c_uint64
        L.orc_stream_moments.argtypes = [C.c_uint64, C.c_uint64, C.c_long, C.c_int, dp, dp]
        L.orc_expmul.argtypes = [dp, dp]
        L.orc_quatmul.argtypes = [dp, dp, dp]
        L.orc_o4_rotation.argtypes = [dp, dp, dp]
        cp = C.POINTER(Ctx)
        L.orc_ctx_init.argtypes = [cp]
        L.orc_mass.argtypes = [cp]
        L.orc_mass.restype = C.c_double
        L.orc_energy.argtypes = [cp, dp]
        L.orc_energy.restype = C.c_double
        L.orc_rotate_phi.argtypes = [cp, dp, C.c_double]
        L.orc_ideal_kick.argtypes = [cp, dp, C.c_double]
        L.orc_ideal_step.argtypes = [cp, dp, C.c_double]
        L.orc_hb_phi_halfsweep.argtypes = [cp, dp, C.c_double, C.c_int]
        L.orc_hb_phi_step.argtypes = [cp, dp, C.c_double]
        L.orc_charge_subsweep.argtypes = [cp, dp, C.c_double, C.c_int, C.c_int, C.c_uint64]
        L.orc_charge_step.argtypes = [cp, dp, C.c_double]
        L.orc_update_charge_pair.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.POINTER(Xoro), C.POINTER(Monitor)]
        L.orc_update_charge_pair.restype = C.c_double
        L.orc_pv2hbsplit_step.argtypes = [cp, dp, C.c_double]
        L.orc_init_gaussian_charges.argtypes = [cp, dp]
        L.orc_thermalize.argtypes = [cp, dp, C.c_double, C.c_int]
        L.orc_round_fp32.argtypes = [dp, C.c_size_t]
        L.orc_measure_walls.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_dft_rows.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_rotate_to_vev.argtypes = [C.c_int, dp, dp]
        ep = C.POINTER(EmParams)
        L.orc_em_noise.argtypes = [ep, C.c_uint64, dp]
        L.orc_em_step.argtypes = [ep, C.c_uint64, dp, dp]
        L.orc_em_run.argtypes = [ep, dp, C.c_int, C.c_int, dp, C.c_void_p]
        L.orc_em_run.restype = C.c_double
        L.orc_em_adjoint.argtypes = [ep, dp, C.c_int, C.c_int, dp, dp]
        L.orc_em_series.argtypes = [ep, dp, C.c_int, C.c_int, dp, C.c_void_p]
        _lib = L
    return _lib


# ---------------------------------------------------------------- RNG


class Engine:
    """xoroshiro128plus with the reference's seed(uint64) formula."""

    def __init__(self, seed=None, state=None):
        self.g = Xoro()
        if state is not None:
            self.g.s0, self.g.s1 = state
        else:
            lib().orc_xoro_seed(C.byref(self.g), C.c_uint64(seed))

    def next(self):
        return lib().orc_xoro_next(C.byref(self.g))

    def uniform(self):
        return lib().orc_uniform(C.byref(self.g))

    def variance1(self):
        return lib().orc_variance1(C.byref(self.g))

    def uniform_int(self, b):
        return lib().orc_uniform_int(C.byref(self.g), b)

    def charge_pair(self, chi, rms, nA, nB, monitor):
        return lib().orc_update_charge_pair(chi, rms, nA, nB, C.byref(self.g), C.byref(monitor))

    def shuffle(self, a):
        arr = (C.c_int * len(a))(*a)
        lib().orc_shuffle(C.byref(self.g), arr, len(a))
        return list(arr)

    @property
    def state(self):
        return (self.g.s0, self.g.s1)


def seed_seq_generate(values, n):
    v = (C.c_uint32 * max(1, len(values)))(*[x & 0xFFFFFFFF for x in values])
    out = (C.c_uint32 * n)()
    lib().orc_seed_seq_generate(v, len(values), out, n)
    return list(out)


class NoiseGenerator:
    """NoiseGenerator(baseSeed) on one rank (ModelG/NoiseGenerator.h:68-82)."""

    def __init__(self, base_seed):
        self.ng = Noise()
        lib().orc_noise_init(C.byref(self.ng), base_seed)

    def uniform(self):
        return lib().orc_uniform(C.byref(self.ng.g))

    def variance1(self):
        return lib().orc_variance1(C.byref(self.ng.g))

    def normal(self):
        return lib().orc_normal(C.byref(self.ng))


def canonical(x):
    return lib().orc_canonical(C.c_uint64(x))


def mix64(z):
    return lib().orc_mix64(C.c_uint64(z))


def base_key(seed32):
    return lib().orc_base_key(seed32)


def stream_key(base, ctr):
    return lib().orc_stream_key(C.c_uint64(base), C.c_uint64(ctr))


def site_key(skey, site):
    return lib().orc_site_key(C.c_uint64(skey), C.c_uint64(site))


def ref_rng(*args):
    """Runs oracle/_ref/ref_rng (built from the reference's xoroshiro128plus.h)."""
    out = subprocess.run([REF_RNG] + [str(a) for a in args], check=True,
                         capture_output=True, text=True).stdout
    return out.split("\n")[:-1]


# ---------------------------------------------------------------- O(4)


def o4_rotation(V, A, phi):
    phi = np.array(phi, dtype=np.float64)
    lib().orc_o4_rotation(np.ascontiguousarray(V, dtype=np.float64),
                          np.ascontiguousarray(A, dtype=np.float64), phi)
    return phi


def expmul(v):
    w = np.zeros(4)
    lib().orc_expmul(np.ascontiguousarray(v, dtype=np.float64), w)
    return w


def quatmul(v1, v2):
    w = np.zeros(4)
    lib().orc_quatmul(np.ascontiguousarray(v1, dtype=np.float64),
                      np.ascontiguousarray(v2, dtype=np.float64), w)
    return w


# ---------------------------------------------------------------- lattice


class Lattice:
    """A single-rank ModelA state and its steppers, on the CPU."""

    def __init__(self, N, mass0=-4.813, dmassdt=0.0, lam=1.0, H=0.01, chi=2.0,
                 gamma=1.0, diffusion=None, seed=42, rng_mode=RNG_PER_SITE,
                 precision=FP64, scounts=(2, 3), nodiffuse=False,
                 onlydiffuse=False, LX=None, t=0.0):
        c = Ctx()
        c.N = N
        c.hx = (N if LX is None else LX) / N
        c.mass0, c.dmassdt, c.lam, c.H, c.chi, c.gamma = mass0, dmassdt, lam, H, chi, gamma
        c.diffusion = 1.0 / 3.0 * gamma if diffusion is None else diffusion
        c.seed = seed
        c.rng_mode = rng_mode
        c.precision = precision
        c.scounts[0], c.scounts[1] = scounts
        c.nodiffuse, c.onlydiffuse = int(nodiffuse), int(onlydiffuse)
        c.t = t
        lib().orc_ctx_init(C.byref(c))
        self.c = c
        self.N = N
        self.u = np.zeros((N, N, N, 10))

    def _u(self):
        assert self.u.flags.c_contiguous and self.u.dtype == np.float64
        return self.u.reshape(-1)

    def energy(self):
        return lib().orc_energy(C.byref(self.c), self._u())

    def ideal_step(self, dt):
        lib().orc_ideal_step(C.byref(self.c), self._u(), dt)

    def rotate_phi(self, tau):
        lib().orc_rotate_phi(C.byref(self.c), self._u(), tau)

    def ideal_kick(self, dt):
        lib().orc_ideal_kick(C.byref(self.c), self._u(), dt)

    def hb_phi_halfsweep(self, dt, ieo):
        lib().orc_hb_phi_halfsweep(C.byref(self.c), self._u(), dt, ieo)

    def hb_phi_step(self, dt):
        lib().orc_hb_phi_step(C.byref(self.c), self._u(), dt)

    def charge_subsweep(self, rms, direction, eoA, ctr):
        lib().orc_charge_subsweep(C.byref(self.c), self._u(), rms, direction, eoA, ctr)

    def charge_step(self, dt):
        lib().orc_charge_step(C.byref(self.c), self._u(), dt)

    def step(self, dt):
        lib().orc_pv2hbsplit_step(C.byref(self.c), self._u(), dt)

    def init_gaussian_charges(self):
        lib().orc_init_gaussian_charges(C.byref(self.c), self._u())

    def thermalize(self, dt, nsteps):
        lib().orc_thermalize(C.byref(self.c), self._u(), dt, nsteps)

    def round_fp32(self):
        lib().orc_round_fp32(self._u(), self.u.size)

    @property
    def t(self):
        return self.c.t

    @t.setter
    def t(self, v):
        self.c.t = v

    def monitors(self):
        return {"phi": self.c.mon_phi.as_dict(), "charge": self.c.mon_charge.as_dict(),
                "ideal": self.c.mon_ideal.as_dict()}

    def measure(self):
        return measure_walls(self.u)


def measure_walls(u):
    N = u.shape[0]
    wall = np.zeros((3, 11, N))
    wallphase = np.zeros((3, 14, N))
    oavg = np.zeros(11)
    lib().orc_measure_walls(N, np.ascontiguousarray(u).reshape(-1), wall.reshape(-1),
                            wallphase.reshape(-1), oavg)
    return wall, wallphase, oavg


def dft_rows(rows):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    R, N = rows.shape
    out = np.zeros((R, N // 2 + 1, 2))
    lib().orc_dft_rows(N, R, rows.reshape(-1), out.reshape(-1))
    return out


def rotate_to_vev(wallk):
    wallk = np.ascontiguousarray(wallk, dtype=np.float64)
    nk = wallk.shape[1]
    out = np.zeros((14, nk, 2))
    lib().orc_rotate_to_vev(nk, wallk.reshape(-1), out.reshape(-1))
    return out


# ---------------------------------------------------------------- EM


def em_params(N, m2, lam, H, gamma, dt, seed, noise_kind=0):
    p = EmParams()
    p.N, p.m2, p.lam, p.H, p.gamma, p.dt = N, m2, lam, H, gamma, dt
    p.base_key = base_key(seed_seq_generate([seed], 1)[0])
    p.noise_kind = noise_kind
    return p


def em_noise(p, step):
    xi = np.zeros((p.N, p.N, p.N, 4))
    lib().orc_em_noise(C.byref(p), step, xi.reshape(-1))
    return xi


def em_step(p, step, phi):
    out = np.zeros_like(phi)
    lib().orc_em_step(C.byref(p), step, np.ascontiguousarray(phi).reshape(-1), out.reshape(-1))
    return out


def em_run(p, phi0, nsteps, window):
    phibar = np.zeros(nsteps)
    fin = np.zeros_like(phi0)
    loss = lib().orc_em_run(C.byref(p), np.ascontiguousarray(phi0).reshape(-1), nsteps,
                            window, phibar, fin.ctypes.data_as(C.c_void_p))
    return loss, phibar, fin


def em_series(p, phi0, nsteps, every):
    """rows[r] = (phibar_0..3, <|phi|^2>, V |phibar|^2) of the state after
    r*every steps (r*every < nsteps), and the final state."""
    rows = np.zeros(((nsteps + every - 1) // every, 6))
    fin = np.zeros_like(phi0)
    lib().orc_em_series(C.byref(p), np.ascontiguousarray(phi0).reshape(-1), nsteps, every,
                        rows.reshape(-1), fin.ctypes.data_as(C.c_void_p))
    return rows, fin


def em_adjoint(p, phi0, nsteps, window):
    g = np.zeros_like(phi0)
    gp = np.zeros(3)
    lib().orc_em_adjoint(C.byref(p), np.ascontiguousarray(phi0).reshape(-1), nsteps,
                         window, g.reshape(-1), gp)
    return g, gp


def set_threads(n):
    """OpenMP threads for the per-site-stream mode's loops (1 = serial)."""
    lib().orc_set_threads(int(n))
