"""ctypes binding of the Model A Euler--Maruyama stepper and its adjoint
(lgv_em_* in include/lgv.h).  Device buffers are passed as raw addresses
(e.g. torch.Tensor.data_ptr()); no CPU fallback exists."""
import ctypes as C

import numpy as np

from .lgv import LgvError, lib

NOISE_UNIFORM = 0   # variance1: U(-sqrt3, sqrt3), NoiseGenerator.h:105
NOISE_GAUSSIAN = 1  # Box-Muller from two canonical draws
NOBS = 6  # phibar_0..3, <|phi|^2>, V |phibar|^2 (LGV_EM_NOBS)


class EmParams(C.Structure):
    _fields_ = [("N", C.c_int), ("m2", C.c_double), ("lam", C.c_double), ("H", C.c_double),
                ("gamma", C.c_double), ("dt", C.c_double), ("seed", C.c_int),
                ("noise_kind", C.c_int)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    L.lgv_em_last_error.restype = C.c_char_p
    L.lgv_em_create.argtypes = [C.POINTER(EmParams), C.c_int, C.c_int, C.c_int,
                                C.POINTER(C.c_void_p)]
    L.lgv_em_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.lgv_em_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_double)]
    L.lgv_em_backward.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
    L.lgv_em_destroy.argtypes = [C.c_void_p]
    L.lgv_em_series.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.lgv_em_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    L.lgv_em_step_noalloc.argtypes = [C.POINTER(EmParams), C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_void_p, C.c_int, C.c_void_p]
    _bound = True
    return L


def _check(rc):
    if rc != 0:
        raise LgvError(f"lgv_em error {rc}: {_bind().lgv_em_last_error().decode()}")


class EulerMaruyama:
    """Model A EM stepper on an N^3 lattice with checkpointed reverse mode."""

    def __init__(self, N, m2, lam, H, gamma, dt, seed, nsteps, ckpt_every,
                 noise_kind=NOISE_UNIFORM, device=0):
        L = _bind()
        p = EmParams(N, m2, lam, H, gamma, dt, seed, noise_kind)
        self.h = C.c_void_p()
        _check(L.lgv_em_create(C.byref(p), nsteps, ckpt_every, device, C.byref(self.h)))
        self.N, self.nsteps = N, nsteps

    def set_stream(self, stream):
        _check(_bind().lgv_em_set_stream(self.h, C.c_void_p(stream or None)))

    def forward(self, phi0_ptr, window, phi_final_ptr=None):
        phibar = np.zeros(self.nsteps)
        loss = C.c_double()
        _check(_bind().lgv_em_forward(self.h, C.c_void_p(phi0_ptr), window,
                                      C.c_void_p(phi_final_ptr) if phi_final_ptr else None,
                                      phibar.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return loss.value, phibar

    def series(self, phi0_ptr, every, phi_final_ptr=None):
        """Observables (phibar_0..3, <|phi|^2>, V |phibar|^2) of the states
        0, every, 2 every, ... below nsteps: an array (ceil(nsteps/every), 6)."""
        rows = np.zeros(((self.nsteps + every - 1) // every, NOBS))
        _check(_bind().lgv_em_series(self.h, C.c_void_p(phi0_ptr), every,
                                     C.c_void_p(phi_final_ptr) if phi_final_ptr else None,
                                     rows.ctypes.data_as(C.c_void_p)))
        return rows

    def step(self, in_ptr, out_ptr, n, obs_ptr=None):
        """Step n from in to out (device addresses), stream-ordered; obs_ptr
        (device, 6 doubles) receives the observables of out."""
        _check(_bind().lgv_em_step(self.h, C.c_void_p(in_ptr), C.c_void_p(out_ptr), n,
                                   C.c_void_p(obs_ptr) if obs_ptr else None))

    def backward(self, grad_phi0_ptr):
        gp = (C.c_double * 3)()
        _check(_bind().lgv_em_backward(self.h, C.c_void_p(grad_phi0_ptr), gp))
        return np.array(gp[:])  # d/d m2, d/d H, d/d lambda

    def close(self):
        if self.h:
            _bind().lgv_em_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_noalloc(N, m2, lam, H, gamma, dt, seed, in_ptr, out_ptr, n, obs_ptr=None,
                 noise_kind=NOISE_UNIFORM, device=0, stream=None):
    """lgv_em_step_noalloc: one step without a handle and without any
    device allocation (N <= 256; calls overlapping on one GPU race)."""
    p = EmParams(N, m2, lam, H, gamma, dt, seed, noise_kind)
    _check(_bind().lgv_em_step_noalloc(C.byref(p), C.c_void_p(in_ptr), C.c_void_p(out_ptr), n,
                                       C.c_void_p(obs_ptr) if obs_ptr else None, device,
                                       C.c_void_p(stream) if stream else None))


def series_susceptibility(rows, V):
    """V (<|phibar|^2> - <|phibar|>^2) over a measurement series: the O(4)
    order-parameter susceptibility from the rows of EulerMaruyama.series."""
    mag = np.sqrt((rows[:, :4] ** 2).sum(axis=1))
    return V * ((mag ** 2).mean() - mag.mean() ** 2)
