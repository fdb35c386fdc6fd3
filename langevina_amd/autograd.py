"""torch.autograd.Function over the HIP Euler--Maruyama stepper.

    loss = modela_em_loss(m2, H, lam, phi0, N=..., gamma=1.0, dt=0.02,
                          nsteps=500, window=250, ckpt_every=50, seed=42)
    loss.backward()   # m2.grad, H.grad, lam.grad, phi0.grad

m2, H and lam are 0-d float64 tensors, phi0 a float64 CUDA tensor of shape
(4, N, N, N).  The forward and adjoint kernels run on torch's current
stream; the noise of every step is replayed from the per-site counter
streams, so the backward pass differentiates the forward trajectory.
"""
import torch

from .em import NOISE_UNIFORM, EulerMaruyama


class ModelAEMLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m2, H, lam, phi0, cfg):
        if phi0.dtype != torch.float64 or not phi0.is_cuda:
            raise TypeError("phi0 must be a float64 CUDA tensor of shape (4, N, N, N)")
        N = cfg["N"]
        if tuple(phi0.shape) != (4, N, N, N):
            raise ValueError(f"phi0 shape {tuple(phi0.shape)} != {(4, N, N, N)}")
        h = EulerMaruyama(N, float(m2), float(lam), float(H), cfg.get("gamma", 1.0),
                          cfg["dt"], cfg.get("seed", 42), cfg["nsteps"],
                          cfg.get("ckpt_every", 50), cfg.get("noise_kind", NOISE_UNIFORM),
                          device=phi0.device.index or 0)
        h.set_stream(torch.cuda.current_stream(phi0.device).cuda_stream)
        phi0c = phi0.detach().contiguous()
        loss, phibar = h.forward(phi0c.data_ptr(), cfg["window"])
        ctx.h = h
        ctx.keep = phi0c
        ctx.phibar = phibar
        return torch.tensor(loss, dtype=torch.float64, device=phi0.device)

    @staticmethod
    def backward(ctx, gout):
        h = ctx.h
        g = torch.empty_like(ctx.keep)
        gp = h.backward(g.data_ptr())
        h.close()
        dev = ctx.keep.device
        gm2, gH, glam = (gout * torch.tensor(v, dtype=torch.float64, device=dev) for v in gp)
        return gm2, gH, glam, gout * g, None


def modela_em_loss(m2, H, lam, phi0, **cfg):
    cfg.setdefault("N", phi0.shape[-1])
    return ModelAEMLoss.apply(m2, H, lam, phi0, cfg)
