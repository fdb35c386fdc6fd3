"""lgv_em_step_noalloc against lgv_em_step: the same bits (fields and
observables) on two calls in a row, the second on a smaller lattice, and
the range check beyond the static scratch array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _pair(torch, N, seed, n, kind, with_obs):
    from langevina_amd.em import EulerMaruyama, step_noalloc
    par = dict(m2=-4.813, lam=1.0, H=0.01, gamma=1.0, dt=0.04)
    rng = np.random.default_rng(seed)
    src = torch.tensor(rng.normal(size=(4, N, N, N)) * 0.5, device="cuda")
    h = EulerMaruyama(N, par["m2"], par["lam"], par["H"], par["gamma"], par["dt"], seed, 4, 4,
                      noise_kind=kind)
    stream = torch.cuda.current_stream().cuda_stream
    h.set_stream(stream)
    want, got = torch.empty_like(src), torch.full_like(src, float("nan"))
    ow = torch.zeros(6, dtype=torch.float64, device="cuda")
    og = torch.full((6,), float("nan"), dtype=torch.float64, device="cuda")
    h.step(src.data_ptr(), want.data_ptr(), n, ow.data_ptr() if with_obs else None)
    step_noalloc(N, par["m2"], par["lam"], par["H"], par["gamma"], par["dt"], seed,
                 src.data_ptr(), got.data_ptr(), n, og.data_ptr() if with_obs else None,
                 noise_kind=kind, device=0, stream=stream)
    torch.cuda.synchronize()
    h.close()
    assert torch.equal(got, want), (N, n)
    if with_obs:
        assert torch.equal(og, ow), (og, ow)
        assert ow[4].item() > 0


def test_noalloc_matches_handle_step_twice(torch):
    print("noalloc N=32 then N=12", flush=True)
    _pair(torch, 32, 7, 3, 1, True)
    _pair(torch, 12, 9, 0, 0, True)
    _pair(torch, 12, 9, 5, 0, False)


def test_noalloc_chain_matches_series(torch):
    print("noalloc chain vs series", flush=True)
    from langevina_amd.em import EulerMaruyama, step_noalloc
    N, nsteps = 16, 6
    phi0 = torch.tensor(np.random.default_rng(1).normal(size=(4, N, N, N)) * 0.2, device="cuda")
    h = EulerMaruyama(N, -0.5, 1.0, 0.02, 1.0, 0.03, 11, nsteps, 3, noise_kind=1)
    fin = torch.empty_like(phi0)
    rows = h.series(phi0.data_ptr(), 1, fin.data_ptr())
    h.close()
    a, b = phi0.clone(), torch.empty_like(phi0)
    obs = torch.zeros(nsteps, 6, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for n in range(nsteps):
        step_noalloc(N, -0.5, 1.0, 0.02, 1.0, 0.03, 11, a.data_ptr(), b.data_ptr(), n,
                     obs[n].data_ptr(), noise_kind=1, stream=st)
        a, b = b, a
    torch.cuda.synchronize()
    assert torch.equal(a, fin)
    np.testing.assert_array_equal(obs[:-1].cpu().numpy(), rows[1:])


def test_noalloc_rejects_lattices_beyond_its_scratch(torch):
    from langevina_amd.em import step_noalloc
    from langevina_amd.lgv import LgvError
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(LgvError, match="outside"):
        step_noalloc(258, -1.0, 1.0, 0.0, 1.0, 0.01, 1, x.data_ptr(), x.data_ptr() + 8, 0)
