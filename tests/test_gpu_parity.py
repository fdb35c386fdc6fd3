"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bars: the phi and charge Metropolis sweeps (random streams, proposals,
accept/reject decisions, stored fields, monitor counts) are bit-exact; one
IdealPV2 substep and one full PV2HBSplit23 step agree to 1e-12 relative
(device sin/cos differ from glibc by an ulp); FP32 storage matches the
oracle's FP32 mode to one float ulp.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PAR = dict(mass0=-4.813, lam=1.0, H=0.01, chi=2.0, gamma=1.0)


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def L():
    import langevina_amd
    langevina_amd.lib()
    return langevina_amd


def make_pair(L, orc, N, seed=42, precision=0, scale=1.0, state_seed=0, **kw):
    par = dict(PAR)
    par.update(kw)
    ctx = L.Context(L.model_params(N, seed=seed, **par), precision=precision)
    lat = orc.Lattice(N, mass0=par["mass0"], lam=par["lam"], H=par["H"], chi=par["chi"],
                      gamma=par["gamma"], seed=seed, rng_mode=orc.RNG_PER_SITE,
                      precision=precision)
    rng = np.random.default_rng(state_seed)
    u = rng.normal(size=(N, N, N, 10)) * scale
    if precision == 1:
        u = u.astype(np.float32).astype(np.float64)
    lat.u[...] = u
    ctx.set_field(u)
    return ctx, lat


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("N", [8, 16, 32])
def test_field_roundtrip(L, orc, N):
    log("roundtrip", N)
    ctx = L.Context(L.model_params(N))
    u = np.random.default_rng(N).normal(size=(N, N, N, 10))
    ctx.set_field(u)
    np.testing.assert_array_equal(ctx.get_field(), u)
    ctx.close()


@pytest.mark.parametrize("N,prec", [(8, 0), (16, 0), (16, 1), (32, 0)])
def test_hb_phi_sweep_bit_exact(L, orc, N, prec):
    log("hb", N, prec)
    ctx, lat = make_pair(L, orc, N, precision=prec, scale=0.7)
    for dt in (0.04 / 6, 0.3):
        ctx.hb_phi_step(dt)
        lat.hb_phi_step(dt)
        np.testing.assert_array_equal(ctx.get_field(), lat.u)
    m_dev, m_orc = ctx.monitors()["phi"], lat.monitors()["phi"]
    for k in ("down_counts", "up_counts", "up_yes", "up_no"):
        assert m_dev[k] == m_orc[k], k
    for k in ("down_dS", "up_dS"):
        assert abs(m_dev[k] - m_orc[k]) <= 1e-9 * abs(m_orc[k]) + 1e-12
    ctx.close()


# N = 4 ... 64: the x sweep in row segments of NH lanes (NH = 2 ... 32); 6
# and 12 (NH = 3, 6): the x sweep in charge_line
@pytest.mark.parametrize("N,prec", [(4, 0), (6, 0), (8, 0), (12, 1), (16, 0), (16, 1), (32, 0),
                                    (64, 1)])
def test_charge_sweep_bit_exact(L, orc, N, prec):
    log("charge", N, prec)
    ctx, lat = make_pair(L, orc, N, precision=prec, seed=7)
    for dt in (0.04, 0.5):
        ctx.charge_step(dt)
        lat.charge_step(dt)
        np.testing.assert_array_equal(ctx.get_field(), lat.u)
    m_dev, m_orc = ctx.monitors()["charge"], lat.monitors()["charge"]
    for k in ("down_counts", "up_counts", "up_yes", "up_no"):
        assert m_dev[k] == m_orc[k], k
    ctx.close()


@pytest.mark.parametrize("N,prec", [(8, 0), (16, 0), (16, 1)])
def test_ideal_step(L, orc, N, prec):
    log("ideal", N, prec)
    ctx, lat = make_pair(L, orc, N, precision=prec, seed=3, scale=0.8)
    ctx.ideal_step(0.04 / 3)
    lat.ideal_step(0.04 / 3)
    got = ctx.get_field()
    if prec == 0:
        assert rel_err(got, lat.u) <= 1e-12
    else:
        # FP32 storage: at most one float ulp apart
        ulp = np.spacing(np.abs(lat.u).astype(np.float32)).astype(np.float64)
        assert (np.abs(got - lat.u) <= ulp).all()
    ctx.close()


@pytest.mark.parametrize("N", [8, 16, 32])
def test_full_step_fp64(L, orc, N):
    log("full step", N)
    ctx, lat = make_pair(L, orc, N, seed=11, scale=0.6)
    for _ in range(2):
        ctx.step(0.04)
        lat.step(0.04)
    assert rel_err(ctx.get_field(), lat.u) <= 1e-12
    assert ctx.rng_state().sweep_ctr == lat.c.sweep_ctr
    m_dev, m_orc = ctx.monitors(), lat.monitors()
    for kind in ("phi", "charge"):
        assert m_dev[kind]["up_counts"] + m_dev[kind]["down_counts"] == \
            m_orc[kind]["up_counts"] + m_orc[kind]["down_counts"]
    ctx.close()


def test_thermalize_and_run_with_clock(L, orc):
    log("thermalize")
    N = 16
    ctx = L.Context(L.model_params(N, seed=42, dmassdt=0.5, **PAR))
    lat = orc.Lattice(N, seed=42, dmassdt=0.5, rng_mode=orc.RNG_PER_SITE, **{
        "mass0": PAR["mass0"], "lam": PAR["lam"], "H": PAR["H"], "chi": PAR["chi"]})
    ctx.thermalize(0.04, 3)
    lat.thermalize(0.04, 3)
    np.testing.assert_array_equal(ctx.get_field(), lat.u)
    ctx.run(0.04, 3)
    for _ in range(3):
        lat.step(0.04)
        lat.t = lat.t + 0.04
    assert ctx.time() == lat.t
    assert rel_err(ctx.get_field(), lat.u) <= 1e-12
    ctx.close()


def test_graph_replay_matches_stream(L, orc):
    """Every replay of the captured step, with eager work and host
    synchronisation between replays (a field copy, a measurement), gives the
    bits of the same step launched on the stream; then a multi-step replay."""
    log("graph")
    N = 16
    a, _ = make_pair(L, orc, N, seed=5)
    b, _ = make_pair(L, orc, N, seed=5)
    b.use_graph(True)
    for i in range(4):
        a.run(0.04, 1)
        b.run(0.04, 1)
        np.testing.assert_array_equal(a.get_field(), b.get_field(), err_msg=f"replay {i}")
        assert a.time() == b.time()
        for x, y in zip(a.measure(), b.measure()):
            np.testing.assert_array_equal(x, y)
    a.run(0.04, 3)
    b.run(0.04, 3)
    np.testing.assert_array_equal(a.get_field(), b.get_field())
    assert a.monitors() == b.monitors()
    a.close()
    b.close()


def test_replicas_in_one_context(L, orc):
    log("replicas")
    sizes = [(8, 1), (16, 2), (12, 3)]
    ctx = L.Context([L.model_params(N, seed=s, **dict(PAR, H=0.01 * s)) for N, s in sizes])
    lats = []
    for r, (N, s) in enumerate(sizes):
        lat = orc.Lattice(N, seed=s, rng_mode=orc.RNG_PER_SITE, mass0=PAR["mass0"],
                          lam=PAR["lam"], H=0.01 * s, chi=PAR["chi"])
        u = np.random.default_rng(r).normal(size=(N, N, N, 10))
        lat.u[...] = u
        ctx.set_field(u, replica=r)
        lats.append(lat)
    ctx.step(0.04)
    for r, lat in enumerate(lats):
        lat.step(0.04)
        assert rel_err(ctx.get_field(r), lat.u) <= 1e-12, r
    ctx.close()


def test_energy_tracking(L, orc):
    log("energy")
    N = 16
    par = dict(PAR)
    ctx = L.Context(L.model_params(N, seed=9, **par), track_energy=True)
    lat = orc.Lattice(N, seed=9, rng_mode=orc.RNG_PER_SITE, mass0=par["mass0"],
                      lam=par["lam"], H=par["H"], chi=par["chi"])
    u = np.random.default_rng(1).normal(size=(N, N, N, 10)) * 0.5
    lat.u[...] = u
    ctx.set_field(u)
    ctx.ideal_step(0.01)
    lat.ideal_step(0.01)
    e = ctx.energy()
    np.testing.assert_allclose(e, list(lat.c.last_energy), rtol=1e-12)
    ctx.close()


def test_measure_walls(L, orc):
    log("measure")
    N = 16
    ctx, lat = make_pair(L, orc, N, seed=2)
    wall, wallphase, oavg = ctx.measure()
    w, wp, oa = orc.measure_walls(lat.u)
    np.testing.assert_allclose(wall, w, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(wallphase, wp, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(oavg, oa, rtol=1e-12, atol=1e-14)
    ctx.close()


def test_64_cubed_step(L, orc):
    log("64^3")
    N = 64
    ctx, lat = make_pair(L, orc, N, seed=42, scale=0.5)
    ctx.step(0.04)
    lat.step(0.04)
    assert rel_err(ctx.get_field(), lat.u) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("N", [16, 48])
def test_ideal_tile_does_not_change_results(L, orc, N):
    log("tiles", N)
    from langevina_amd.autotune import set_ideal_tile
    u = np.random.default_rng(9).normal(size=(N, N, N, 10)) * 0.5
    ref = None
    for cap, tz in ((32, 16), (8, 4), (16, 32), (64, 8), (16, 2), (8, 1)):
        ctx = L.Context(L.model_params(N, seed=3, **PAR))
        set_ideal_tile(ctx, 0, cap, tz)
        ctx.set_field(u)
        ctx.step(0.04)
        got = ctx.get_field()
        ctx.close()
        if ref is None:
            ref = got
        else:
            np.testing.assert_array_equal(got, ref)


def test_ideal_tile_two_pass_ring(L):
    """At N = 128 a 64-half-column tile is 4 rows high: its halo ring of
    2*4 + 4*64 = 264 sites is more than the 256 threads of a block, and the
    remainder takes a second pass.  It gives the bits of the default tile,
    also with a chunk length that leaves a short last chunk."""
    log("tile two-pass ring")
    from langevina_amd.autotune import set_ideal_tile
    N = 128
    u = np.random.default_rng(4).normal(size=(N, N, N, 10)) * 0.5
    ref = None
    for cap, tz in ((32, 16), (64, 16), (64, 5)):
        ctx = L.Context(L.model_params(N, seed=3, **PAR))
        set_ideal_tile(ctx, 0, cap, tz)
        ctx.set_field(u)
        ctx.step(0.04)
        got = ctx.get_field()
        ctx.close()
        if ref is None:
            ref = got
        else:
            np.testing.assert_array_equal(got, ref, err_msg=f"cap {cap} tz {tz}")


@pytest.mark.parametrize("variant", ["nodiffuse", "onlydiffuse"])
def test_step_variants(L, orc, variant):
    """PV2HBSplit23NoDiffuse / OnlyDiffuse (SuperPions.cxx:209-225)."""
    log("variant", variant)
    N = 16
    kw = {variant: True}
    ctx = L.Context(L.model_params(N, seed=5, **PAR), **kw)
    lat = orc.Lattice(N, seed=5, rng_mode=orc.RNG_PER_SITE, **kw, **{
        "mass0": PAR["mass0"], "lam": PAR["lam"], "H": PAR["H"], "chi": PAR["chi"]})
    u = np.random.default_rng(4).normal(size=(N, N, N, 10)) * 0.6
    lat.u[...] = u
    ctx.set_field(u)
    for _ in range(2):
        ctx.step(0.04)
        lat.step(0.04)
    assert rel_err(ctx.get_field(), lat.u) <= 1e-12
    assert ctx.rng_state().sweep_ctr == lat.c.sweep_ctr
    ctx.close()


@pytest.mark.parametrize("N, precision", [(128, 0), (128, 1), (256, 0)])
def test_charge_rows_in_one_wavefront(L, orc, N, precision):
    """At N = 128 and 256 a wavefront owns whole x-rows of the charge sweep
    (first-parity results passed between lanes); the charge-only step gives
    the oracle's bits."""
    log("charge rows", N, precision)
    ctx = L.Context(L.model_params(N, seed=6, **PAR), precision=precision, onlydiffuse=True)
    lat = orc.Lattice(N, seed=6, rng_mode=orc.RNG_PER_SITE, precision=precision,
                      onlydiffuse=True, mass0=PAR["mass0"], lam=PAR["lam"], H=PAR["H"],
                      chi=PAR["chi"])
    u = np.random.default_rng(6).normal(size=(N, N, N, 10)) * 0.8
    if precision == 1:
        u = u.astype(np.float32).astype(np.float64)
    lat.u[...] = u
    ctx.set_field(u)
    del u
    for _ in range(2):
        ctx.step(0.04)
        lat.step(0.04)
    np.testing.assert_array_equal(ctx.get_field(), lat.u)
    ctx.close()


def test_lattice_spacing_and_moving_mass(L, orc):
    """LX = 2 NX (h = 2, the kick's 1/h^2 factor, ModelA.h:229) and
    m^2(t) = mass0 + dmassdt t."""
    log("spacing")
    N = 16
    par = dict(PAR, dmassdt=-0.7)
    ctx = L.Context(L.model_params(N, LX=2 * N, seed=8, **par))
    lat = orc.Lattice(N, LX=2 * N, seed=8, rng_mode=orc.RNG_PER_SITE, dmassdt=-0.7, **{
        "mass0": PAR["mass0"], "lam": PAR["lam"], "H": PAR["H"], "chi": PAR["chi"]})
    u = np.random.default_rng(6).normal(size=(N, N, N, 10)) * 0.6
    lat.u[...] = u
    ctx.set_field(u)
    ctx.run(0.04, 3)
    for _ in range(3):
        lat.step(0.04)
        lat.t = lat.t + 0.04
    assert rel_err(ctx.get_field(), lat.u) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("N", [4, 6, 10])
def test_small_lattices(L, orc, N):
    """Sizes where every tile wraps onto itself (N = 4, 6) or is ragged (10)."""
    log("small", N)
    ctx, lat = make_pair(L, orc, N, seed=13, scale=0.6)
    for _ in range(2):
        ctx.step(0.04)
        lat.step(0.04)
    assert rel_err(ctx.get_field(), lat.u) <= 1e-12
    ctx.close()


def test_full_step_fp32(L, orc):
    """FP32 fields: one full step against the oracle's FP32 mode."""
    log("full step fp32")
    N = 16
    ctx, lat = make_pair(L, orc, N, seed=21, precision=1, scale=0.6)
    ctx.step(0.04)
    lat.step(0.04)
    got = ctx.get_field()
    assert np.array_equal(got, got.astype(np.float32).astype(np.float64))
    # the ideal substep's FP64 arithmetic differs from glibc's by an ulp, which
    # can move a stored float by one ulp
    np.testing.assert_allclose(got, lat.u, rtol=2.5e-7, atol=1e-30)
    ctx.close()


def test_nontemporal_store_kernels_match(L, monkeypatch):
    """The large-field mode (chosen for fields above 1 GiB: the FP64 phi march
    and non-temporal stores in it and in the ideal kernel) is forced on a small
    lattice and gives the bits of the default mode (per-site phi kernel, plain
    stores), in both storage precisions."""
    log("nontemporal stores")
    N = 32
    u = np.random.default_rng(23).normal(size=(N, N, N, 10)) * 0.6
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("LGV_NT_STORES", mode)
        for prec in (0, 1):
            ctx = L.Context(L.model_params(N, seed=4, **PAR), precision=prec)
            ctx.set_field(u.astype(np.float32).astype(np.float64) if prec else u)
            ctx.run(0.04, 2)
            out[mode, prec] = ctx.get_field()
            ctx.close()
    for prec in (0, 1):
        np.testing.assert_array_equal(out["0", prec], out["1", prec])


@pytest.mark.parametrize("N", [8, 12, 20, 64, 132])
def test_phi_half_sweep_kernels_match(L, monkeypatch, N):
    """The two phi half-sweep kernels (one thread per site, the z-march with
    the other parity's plane in LDS) give the same bits and the same
    acceptance counts, in both storage precisions.  N = 132 puts several
    64 x 4 x 16 march tiles on the lattice, the last of them short in x and
    z; N = 64 has 8 rows of 32 x 8 tiles, so the march numbers its blocks by
    y bands (one band per XCD); for N <= 20 one tile spans the whole
    cross-section, so its ring wraps onto its own half-columns."""
    log(f"phi sweep kernels N={N}")
    u = np.random.default_rng(N).normal(size=(N, N, N, 10)) * 0.6
    out, mons = {}, {}
    for kern in ("site", "march"):
        monkeypatch.setenv("LGV_PHI_KERNEL", kern)
        for prec in (0, 1):
            ctx = L.Context(L.model_params(N, seed=6, **PAR), precision=prec)
            ctx.set_field(u.astype(np.float32).astype(np.float64) if prec else u)
            ctx.run(0.04, 2)
            out[kern, prec] = ctx.get_field()
            mons[kern, prec] = ctx.monitors()["phi"]
            ctx.close()
    for prec in (0, 1):
        np.testing.assert_array_equal(out["march", prec], out["site", prec])
        a, b = mons["march", prec], mons["site", prec]
        for k in a:
            if isinstance(a[k], int):
                assert a[k] == b[k], (prec, k)
            else:
                assert a[k] == pytest.approx(b[k], rel=1e-9, abs=1e-12), (prec, k)


def test_phi_kernel_choice_is_checked(L, monkeypatch):
    monkeypatch.setenv("LGV_PHI_KERNEL", "fused")
    with pytest.raises(L.LgvError, match="expected site or march"):
        L.Context(L.model_params(8, seed=1, **PAR))


def test_field_transfer_in_slabs(L, orc, monkeypatch):
    """set_field / get_field move the AoS field through a staging buffer in
    slabs of z-planes (one slab up to 2 GiB); forced to 3-plane slabs (the last
    one short) the round trip is exact and a step still matches the oracle."""
    log("staging slabs")
    monkeypatch.setenv("LGV_STAGE_PLANES", "3")
    N = 16
    ctx, lat = make_pair(L, orc, N, seed=8)
    np.testing.assert_array_equal(ctx.get_field(), lat.u)
    ctx.step(0.04)
    lat.step(0.04)
    assert rel_err(ctx.get_field(), lat.u) <= 1e-12
    ctx.close()


def test_64bit_offset_kernels_match(L, orc, monkeypatch):
    """The 64-bit-offset instantiation (planes of 4 GiB or more) is forced on a
    small lattice and must give the same bits as the 32-bit one."""
    log("wide offsets")
    N = 16
    u = np.random.default_rng(17).normal(size=(N, N, N, 10)) * 0.6
    out = {}
    for nt in ("0", "1"):   # both phi kernels and store modes
        monkeypatch.setenv("LGV_NT_STORES", nt)
        for wide in ("0", "1"):
            monkeypatch.setenv("LGV_FORCE_WIDE", wide)
            for prec in (0, 1):
                ctx = L.Context(L.model_params(N, seed=3, **PAR), precision=prec)
                ctx.set_field(u.astype(np.float32).astype(np.float64) if prec else u)
                ctx.run(0.04, 2)
                out[nt, wide, prec] = ctx.get_field()
                ctx.close()
    for key, val in out.items():
        np.testing.assert_array_equal(val, out["0", "0", key[2]], err_msg=str(key))


@pytest.mark.parametrize("N,prec", [(48, 0), (128, 0), (48, 1), (128, 1)])
def test_sweep_tiles_do_not_change_results(L, monkeypatch, N, prec):
    """The tiles of the phi z-march and of the charge march (including tiles
    that turn the whole-row x sweep of N = 128 off, ragged tiles at N = 48,
    short chunks, and odd z-chunks, which end the FP32 march's two-plane
    prefetch on one plane) give the bits and acceptance counts of the
    defaults; a retile in the middle of a run keeps the monitor counts."""
    log("sweep tiles", N, prec)
    from langevina_amd.autotune import set_sweep_tiles
    monkeypatch.setenv("LGV_PHI_KERNEL", "march")
    u = np.random.default_rng(N + 1).normal(size=(N, N, N, 10)) * 0.6
    if prec:
        u = u.astype(np.float32).astype(np.float64)
    ref = None
    for tiles in ((0, 0, 0, 0), (16, 8, 32, 4), (64, 32, 16, 32), (32, 16, 64, 8),
                  (16, 4, 16, 2), (32, 5, 0, 0), (64, 1, 0, 0)):
        ctx = L.Context(L.model_params(N, seed=9, **PAR), precision=prec)
        ctx.set_field(u)
        ctx.run(0.04, 1)
        set_sweep_tiles(ctx, 0, *tiles)
        ctx.run(0.04, 1)
        got, mon = ctx.get_field(), ctx.monitors()
        ctx.close()
        if ref is None:
            ref = (got, mon)
        else:
            np.testing.assert_array_equal(got, ref[0], err_msg=str(tiles))
            for kind in ("phi", "charge"):
                for k in ("down_counts", "up_yes", "up_no"):
                    assert mon[kind][k] == ref[1][kind][k], (tiles, kind, k)


@pytest.mark.parametrize("prec", [0, 1])
def test_phi_kernel_switch_keeps_results(L, prec):
    """lgv_set_phi_kernel: switching between the per-site kernel and the
    z-march in the middle of a run gives the bits and acceptance counts of a
    run on either kernel alone; LGV_PHI_AUTO restores the size rule's
    choice (the per-site kernel at 24^3); other values are rejected."""
    log("phi kernel switch", prec)
    N = 24
    u = np.random.default_rng(7).normal(size=(N, N, N, 10)) * 0.6
    runs = {}
    for name, kernels in (("site", (L.PHI_SITE, L.PHI_SITE)),
                          ("march", (L.PHI_MARCH, L.PHI_MARCH)),
                          ("switch", (L.PHI_MARCH, L.PHI_SITE))):
        ctx = L.Context(L.model_params(N, seed=5, **PAR), precision=prec)
        ctx.set_field(u)
        for k in kernels:
            ctx.set_phi_kernel(k)
            assert ctx.phi_kernel() == k
            ctx.run(0.04, 2)
        runs[name] = (ctx.get_field(), ctx.monitors())
        if name == "switch":
            ctx.set_phi_kernel(L.PHI_AUTO)
            assert ctx.phi_kernel() == L.PHI_SITE
            with pytest.raises(L.LgvError, match="phi kernel"):
                ctx.set_phi_kernel(3)
        ctx.close()
    for name in ("march", "switch"):
        np.testing.assert_array_equal(runs[name][0], runs["site"][0], err_msg=name)
        for kind in ("phi", "charge"):
            for k in ("down_counts", "up_yes", "up_no"):
                assert runs[name][1][kind][k] == runs["site"][1][kind][k], (name, kind, k)


def test_sampled_timing_counts_every_kth_step(L):
    """lgv_timing_enable(ctx, k): the kernels of every k-th step (the first
    after enabling included) are timed, the others run untimed; results are
    those of an untimed run."""
    log("sampled timing")
    N = 16
    u = np.random.default_rng(3).normal(size=(N, N, N, 10)) * 0.6
    out = []
    for k in (0, 3):
        ctx = L.Context(L.model_params(N, seed=4, **PAR))
        ctx.set_field(u)
        ctx.timing_enable(k)
        ctx.run(0.04, 7)
        tm = ctx.timing()
        if k:
            # steps 0, 3 and 6: 3 ideal substeps, 12 phi half-sweeps and 3
            # charge directions each
            assert (tm["ideal"][1], tm["phi"][1], tm["charge"][1]) == (9, 36, 9)
            assert all(tm[c][0] > 0 for c in ("ideal", "phi", "charge"))
        else:
            assert tm["ideal"][1] == tm["phi"][1] == tm["charge"][1] == 0
        with pytest.raises(L.LgvError, match="timing_enable"):
            ctx.timing_enable(-1)
        out.append(ctx.get_field())
        ctx.close()
    np.testing.assert_array_equal(out[0], out[1])


def test_sweep_tiles_are_checked(L):
    from langevina_amd.autotune import set_sweep_tiles
    ctx = L.Context(L.model_params(16, seed=1, **PAR))
    with pytest.raises(L.LgvError, match="out of range"):
        set_sweep_tiles(ctx, 0, 0, 0, 16, 5)
    with pytest.raises(L.LgvError, match="out of range"):
        set_sweep_tiles(ctx, 0, 2, 0, 0, 0)
    ctx.close()


def test_multi_step_graph_matches_stream(L, orc):
    """A graph of 4 steps replayed for a run of 11 (two replays and three
    one-step graphs) gives the bits, clock and counters of stream launches."""
    log("multi-step graph")
    N = 16
    a, _ = make_pair(L, orc, N, seed=15)
    b, _ = make_pair(L, orc, N, seed=15)
    b.use_graph(4)
    for n in (11, 4, 1):
        a.run(0.04, n)
        b.run(0.04, n)
        np.testing.assert_array_equal(a.get_field(), b.get_field(), err_msg=f"run of {n}")
        assert a.time() == b.time()
        assert a.rng_state().sweep_ctr == b.rng_state().sweep_ctr
    assert a.monitors() == b.monitors()
    a.close()
    b.close()


@pytest.mark.parametrize("N,prec", [(16, 0), (32, 1)])
def test_concurrent_charge_sweep_matches_sequential(L, orc, monkeypatch, N, prec):
    """The charge sweep on the second stream, concurrent with the last phi
    sweeps of the step, gives the bits, counters and monitors of the
    sequential order, on the stream and through a graph; and the oracle's."""
    log("overlap", N, prec)
    out = {}
    for ov in ("0", "1"):
        monkeypatch.setenv("LGV_OVERLAP", ov)
        for graph in (0, 3):
            ctx, lat = make_pair(L, orc, N, seed=19, precision=prec, scale=0.6)
            ctx.use_graph(graph)
            ctx.run(0.04, 4)
            out[ov, graph] = (ctx.get_field(), ctx.monitors(), ctx.rng_state().sweep_ctr)
            ctx.close()
    ref = out["0", 0]
    for key, (f, m, ctr) in out.items():
        np.testing.assert_array_equal(f, ref[0], err_msg=str(key))
        assert m == ref[1] and ctr == ref[2], key
    for _ in range(4):
        lat.step(0.04)
        lat.t = lat.t + 0.04
    if prec == 0:
        assert rel_err(ref[0], lat.u) <= 1e-12
    else:
        ulp = np.spacing(np.abs(lat.u).astype(np.float32)).astype(np.float64)
        assert (np.abs(ref[0] - lat.u) <= 4 * ulp).all()
