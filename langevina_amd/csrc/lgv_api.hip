// lgv_api.hip -- the C ABI of include/lgv.h: context, replica storage,
// composition of PV2HBSplit::step from the kernels, host-side global engine
// (initial charges), monitors and measurement.
#include <cfloat>
#include <array>
#include <cmath>
#include <complex>
#include <limits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "lgv.h"
#include "lgv_device.h"
#include "lgv_internal.h"

using namespace lgv;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHECK(expr)                                                          \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return fail(LGV_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                          \
  } while (0)

// std::seed_seq{base}.generate(1 word): the seed NoiseGenerator gives rank 0
// (ModelG/NoiseGenerator.h:68-82).
uint32_t rank0_seed(int base) {
  std::seed_seq seq = {base};
  uint32_t out = 0;
  seq.generate(&out, &out + 1);
  return out;
}

// NoiseGenerator's engine and its normal() (libstdc++ 11 Marsaglia polar
// method with one cached value, bits/random.tcc:1802-1835).
struct HostNoise {
  Xoro g;
  double saved = 0.;
  int has_saved = 0;
  double normal() {
    if (has_saved) {
      has_saved = 0;
      return saved * 1.0 + 0.0;
    }
    double x, y, r2;
    do {
      x = 2.0 * canonical(xoro_next(g)) - 1.0;
      y = 2.0 * canonical(xoro_next(g)) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = std::sqrt(-2 * std::log(r2) / r2);
    saved = x * mult;
    has_saved = 1;
    return (y * mult) * 1.0 + 0.0;
  }
};

}  // namespace

struct lgv_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // the charge sweep of a step runs on stream2, concurrently with the phi
  // sweeps that follow the last ideal substep (see enqueue_step)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool overlap = true;
  bool fp32 = false;
  unsigned s0 = 2, s1 = 3;
  bool nodiffuse = false, onlydiffuse = false, track_energy = false;
  int nrep = 0;
  std::vector<lgv_model_params> mp;
  std::vector<ReplicaParams> hp;
  std::vector<HostNoise> noise;
  ReplicaParams *dP = nullptr;
  ReplicaState *dS = nullptr;
  std::vector<void *> fields;
  Plan plan_half, plan_full, plan_ideal, plan_charge, plan_hb;
  bool wide = false;    // some plane is >= 4 GiB: kernels use 64-bit offsets
  bool nt_stores = false;  // one buffer's fields exceed kStreamBytes
  bool phi_march = false;  // the phi half-sweep runs the z-march kernel
  bool phi_march_auto = false;  // the size rule's choice (LGV_PHI_AUTO)
  MonSlot *mon_phi = nullptr, *mon_charge = nullptr, *mon_ideal = nullptr;
  // monitor counts folded out of the per-block slots when a retile
  // reallocated them: [replica][phi, charge]
  std::vector<std::array<lgv_monitor, 2>> mon_carry;
  EnergySlot *eslots = nullptr;
  double coef_hb = NAN, coef_charge = NAN, coef_ideal = NAN;
  double *d_aos = nullptr;
  int maxN = 0;
  double *d_partXZ = nullptr, *d_partXY = nullptr;
  // walls, phase walls and their transforms of one measurement, packed for
  // the current N (see MeasBuf) so that one copy brings them to the host
  double *d_meas = nullptr;
  bool use_graph = false;
  int graph_steps = 1;                 // steps per captured graph
  hipGraphExec_t graph_exec = nullptr;  // graph_steps steps
  hipGraphExec_t graph_one = nullptr;   // one step (the remainder of a run)
  double graph_dt = NAN;
  // HIP-event timing of kernel classes on the context stream
  bool timing = false;      // HIP-event timing configured (lgv_timing_enable)
  bool timing_now = false;  // the launches being enqueued are timed
  int timing_every = 1;     // time every k-th step of lgv_step / lgv_run
  uint64_t timing_ctr = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pairs;
  double t_ms[LGV_NCLASS] = {0};
  long long t_launch[LGV_NCLASS] = {0};
};

namespace {

// z-planes per staging slab; LGV_STAGE_PLANES (a test hook) lowers it.
int slab_planes(int N) {
  int n = stage_planes(N);
  if (const char *e = getenv("LGV_STAGE_PLANES")) {
    const int v = atoi(e);
    if (v >= 1 && v < n) n = v;
  }
  return n;
}

// Layout of one measurement in a packed buffer for side N: walls [3][11][N],
// phase walls [3][14][N], their transforms [3][11][nk][2] and [3][14][nk][2].
struct MeasBuf {
  double *wall, *wallphase, *wallk, *wallphasek;
  int64_t total;  // doubles
  static int64_t size(int N) { return 3 * 25 * (int64_t)N + 3 * 25 * (int64_t)(N / 2 + 1) * 2; }
  MeasBuf(double *base, int N) {
    const int64_t nk = N / 2 + 1;
    wall = base;
    wallphase = wall + 3 * 11 * (int64_t)N;
    wallk = wallphase + 3 * 14 * (int64_t)N;
    wallphasek = wallk + 3 * 11 * nk * 2;
    total = size(N);
  }
};

int build_plan(Plan &plan, const std::vector<int64_t> &work, bool wide) {
  plan.wide = wide;
  plan.first.clear();
  std::vector<int> rep, first;
  for (size_t r = 0; r < work.size(); r++) {
    first.push_back((int)rep.size());
    const int64_t nb = (work[r] + kBlock - 1) / kBlock;
    for (int64_t b = 0; b < nb; b++) rep.push_back((int)r);
  }
  plan.nblocks = (int)rep.size();
  plan.first = first;
  HIPCHECK(hipMalloc(&plan.d_blk_rep, rep.size() * sizeof(int)));
  HIPCHECK(hipMalloc(&plan.d_blk_first, first.size() * sizeof(int)));
  HIPCHECK(hipMemcpy(plan.d_blk_rep, rep.data(), rep.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(plan.d_blk_first, first.data(), first.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  return LGV_OK;
}

// Releases a plan's device arrays (the stream must be idle).
int free_plan(Plan &plan) {
  if (plan.d_blk_rep) HIPCHECK(hipFree(plan.d_blk_rep));
  if (plan.d_blk_first) HIPCHECK(hipFree(plan.d_blk_first));
  plan = Plan{};
  return LGV_OK;
}

// Blocks of the fused ideal kernel: one per (x-tile, y-tile, z-chunk) of
// every replica, with that replica's tile shape.
int build_ideal_plan(lgv_ctx *c) {
  std::vector<int64_t> work;
  for (int r = 0; r < c->nrep; r++) {
    const ReplicaParams &p = c->hp[r];
    int TXH, TY, TZ;
    ideal_tile(p.N, p.tile_cap, p.tile_tz, TXH, TY, TZ);
    if ((TY + 2) * (2 * TXH + 2) > ideal_lds(c->fp32))
      return fail(LGV_ERR_ARG, "replica %d: ideal tile %dx%d exceeds the LDS plane", r, TXH, TY);
    work.push_back((int64_t)((p.NH + TXH - 1) / TXH) * ((p.N + TY - 1) / TY) *
                   ((p.N + TZ - 1) / TZ) * kBlock);
  }
  if (c->plan_ideal.d_blk_rep) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipFree(c->plan_ideal.d_blk_rep));
    HIPCHECK(hipFree(c->plan_ideal.d_blk_first));
    c->plan_ideal = Plan{};
  }
  const int rc = build_plan(c->plan_ideal, work, c->wide);
  c->plan_ideal.stream = c->nt_stores;
  return rc;
}

// Blocks of the phi march kernel: (x-tile, y-tile, z-chunk) per replica.
int build_hb_plan(lgv_ctx *c) {
  std::vector<int64_t> work;
  for (int r = 0; r < c->nrep; r++) {
    const ReplicaParams &p = c->hp[r];
    int TXH, TY, TZ;
    hb_tile(p.N, p.hb_cap, p.hb_tz, TXH, TY, TZ);
    if ((TY + 2) * (TXH + 2) > kHbMarchGrid || 2 * TY + 2 * TXH > kBlock)
      return fail(LGV_ERR_ARG, "replica %d: phi tile %dx%d does not fit", r, TXH, TY);
    work.push_back((int64_t)((p.NH + TXH - 1) / TXH) * ((p.N + TY - 1) / TY) *
                   ((p.N + TZ - 1) / TZ) * kBlock);
  }
  int rc = free_plan(c->plan_hb);
  if (rc) return rc;
  rc = build_plan(c->plan_hb, work, c->wide);
  c->plan_hb.stream = c->nt_stores;
  return rc;
}

// Blocks of the charge march kernel: (x-tile, line tile, chunk) per replica.
int build_charge_plan(lgv_ctx *c) {
  std::vector<int64_t> work;
  for (int r = 0; r < c->nrep; r++) {
    const ReplicaParams &p = c->hp[r];
    int TXH, TQ, C;
    charge_tile(p.N, p.chg_cap, p.chg_chunk, TXH, TQ, C);
    if (C % 2 != 0) return fail(LGV_ERR_ARG, "replica %d: charge chunk %d is odd", r, C);
    work.push_back((int64_t)((p.NH + TXH - 1) / TXH) * ((p.N + TQ - 1) / TQ) *
                   ((p.N + C - 1) / C) * kBlock);
  }
  int rc = free_plan(c->plan_charge);
  if (rc) return rc;
  rc = build_plan(c->plan_charge, work, c->wide);
  // the x sub-sweeps with whole rows per wavefront need NH = 64 or 128 and
  // a tile of 64 half-columns x 4 lines
  bool xw = true;
  for (int r = 0; r < c->nrep; r++) {
    const ReplicaParams &p = c->hp[r];
    int TXH, TQ, C;
    charge_tile(p.N, p.chg_cap, p.chg_chunk, TXH, TQ, C);
    xw = xw && (p.NH == 64 || p.NH == 128) && TXH == 64 && TQ == 4;
  }
  c->plan_charge.xwave = xw;
  return rc;
}

// Step coefficients with the reference's expressions:
//   EulerLangevinHB  rdtg = sqrt(2.*dt*gamma)          Stepper.cxx:301
//   ModelGChargeHB   rdtsigma = sqrt(2.*dt*sigma())    Stepper.cxx:456
//   IdealPV2         rotatePhi(dt/2.0)                 Stepper.cxx:73,169
bool same(double a, double b) { return a == b || (std::isnan(a) && std::isnan(b)); }

int ensure_coeffs(lgv_ctx *c, double dt_hb, double dt_charge, double dt_ideal) {
  if (same(dt_hb, c->coef_hb) && same(dt_charge, c->coef_charge) &&
      same(dt_ideal, c->coef_ideal))
    return LGV_OK;
  for (int r = 0; r < c->nrep; r++) {
    const lgv_model_params &m = c->mp[r];
    ReplicaParams &p = c->hp[r];
    p.rdtg = std::sqrt(2. * dt_hb * m.gamma);
    p.rms = std::sqrt(2. * dt_charge * (m.diffusion * m.chi));
    p.dt_ideal = dt_ideal;
    p.tau_ideal = dt_ideal / 2.0;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(c->dP, c->hp.data(), c->nrep * sizeof(ReplicaParams),
                     hipMemcpyHostToDevice));
  c->coef_hb = dt_hb;
  c->coef_charge = dt_charge;
  c->coef_ideal = dt_ideal;
  return LGV_OK;
}

// Brackets one kernel (class cls) with events when timing is on.
struct Timed {
  lgv_ctx *c;
  int cls;
  hipEvent_t e1 = nullptr;
  Timed(lgv_ctx *cc, int k) : c(cc), cls(k) {
    if (!c->timing_now) return;
    if (c->ev_used + 2 > c->ev_pool.size()) {
      for (int i = 0; i < 64; i++) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { c->timing = c->timing_now = false; return; }
        c->ev_pool.push_back(e);
      }
    }
    hipEvent_t e0 = c->ev_pool[c->ev_used++];
    if (hipEventRecord(e0, c->stream) != hipSuccess) {  // timing is best effort
      c->timing = c->timing_now = false;
      return;
    }
    e1 = c->ev_pool[c->ev_used++];
    c->ev_pairs.push_back({cls, {e0, e1}});
  }
  ~Timed() {
    if (!e1) return;
    if (hipEventRecord(e1, c->stream) != hipSuccess) {
      c->timing = c->timing_now = false;
      c->ev_pairs.pop_back();  // the pair this object pushed: never resolved
    }
  }
};

// Resolves recorded event pairs into the per-class totals.
int drain_timing(lgv_ctx *c) {
  if (c->ev_pairs.empty()) return LGV_OK;
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (auto &pr : c->ev_pairs) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, pr.second.first, pr.second.second));
    c->t_ms[pr.first] += ms;
    c->t_launch[pr.first] += 1;
  }
  c->ev_pairs.clear();
  c->ev_used = 0;
  return LGV_OK;
}

// The launches of a step read the live buffers of the replica state XOR
// `flip` (bit 0 phi, bit 1 charges), the flips of the step's earlier
// launches; k_advance applies the step's net flip at its end, so no launch of
// a step writes the state's buffer bits.  Each enqueue_* returns the flips
// of its own launches.
int enqueue_ideal(lgv_ctx *c, int off, int flip) {
  if (c->timing_now && c->ev_used + 16 > 4096) drain_timing(c);
  if (c->track_energy) {
    Timed tm(c, LGV_CLASS_ENERGY);
    launch_energy(c->fp32, c->plan_full, c->dP, c->dS, flip, c->eslots, c->stream);
    launch_energy_finish(c->nrep, c->plan_full, c->dP, c->dS, c->eslots, 0, off,
                         c->mon_ideal, c->stream);
  }
  {
    Timed tm(c, LGV_CLASS_IDEAL);
    launch_ideal(c->fp32, c->plan_ideal, c->dP, c->dS, flip, c->stream);
  }
  if (c->track_energy) {
    Timed tm(c, LGV_CLASS_ENERGY);
    launch_energy(c->fp32, c->plan_full, c->dP, c->dS, flip ^ 3, c->eslots, c->stream);
    launch_energy_finish(c->nrep, c->plan_full, c->dP, c->dS, c->eslots, 1, off,
                         c->mon_ideal, c->stream);
  }
  return 3;  // out of place: phi and charges
}

// EulerLangevinHB::step (counters off, off+1 for the two parities), one
// in-place launch per parity: the z-march with the other parity's plane in
// LDS or the one-thread-per-site kernel, by the size rule (the march for
// FP64 fields above 1 GiB and FP32 fields above 600 MiB) unless
// lgv_set_phi_kernel chose.  Measured per half-sweep before the march's
// prefetch, FP64: 256^3 march 157 us against 204 us per site; 192^3 85
// against 75 (70 in plane bands), 128^3 35 against 29 (26), 64^3 30 against
// 9 (below 1 GiB the fields stay largely cache-resident and the per-site
// kernel's 8 waves per SIMD win); FP32 256^3 110 against 109, but the
// march's lower traffic speeds up the kernels that follow (kMarchBytesFp32).
bool use_march(const lgv_ctx *c) { return c->phi_march; }
const Plan &hb_plan(const lgv_ctx *c) { return use_march(c) ? c->plan_hb : c->plan_half; }
int enqueue_hb(lgv_ctx *c, int off, int flip) {
  for (int ieo = 0; ieo < 2; ieo++) {
    Timed tm(c, LGV_CLASS_PHI);
    if (use_march(c))
      launch_hb_march(c->fp32, c->plan_hb, c->dP, c->dS, ieo, off + ieo, flip, c->mon_phi,
                      c->stream);
    else
      launch_hb_phi(c->fp32, c->plan_half, c->dP, c->dS, ieo, off + ieo, flip, c->mon_phi,
                    c->stream);
  }
  return 0;  // in place
}

// ModelGChargeHB::step: the shuffle (counter off, drawn by every charge
// launch), then per direction both parities in one pass (counters
// off+1+2*ixyz and off+2+2*ixyz, as the unfused order), each launch writing
// the other charge buffer.
int enqueue_charge(lgv_ctx *c, int off, int flip, hipStream_t st) {
  for (int ixyz = 0; ixyz < 3; ixyz++) {
    Timed tm(c, LGV_CLASS_CHARGE);
    launch_charge_march(c->fp32, c->plan_charge, c->dP, c->dS, ixyz, off + 1 + 2 * ixyz,
                        flip ^ (ixyz & 1 ? 2 : 0), c->mon_charge, st);
  }
  return 2;  // three flips of the charge buffers
}

// PV2HBSplit::step (Stepper.cxx:542-580); returns the number of stream
// counters consumed.  The charge sweep reads and writes only the charges
// and the phi sweeps only phi, so after the last ideal substep the charge
// sweep runs on the second stream while the phi sweeps run on the first
// (fork and join by events, which a graph capture records as well).  Every
// sweep keeps its counter (random streams) and its monitor slots, so the
// result is the sequential one bit for bit.  Contexts of one lattice whose
// fields exceed 600 MiB keep the sequential order (their launches fill the
// GPU on their own), and so do timed steps (events on one stream).  flip
// returns the step's net live-buffer flip.
int enqueue_step(lgv_ctx *c, int &flip) {
  int off = 0;
  flip = 0;
  bool charge_queued = false, join_event = false;
  c->timing_now = c->timing && c->timing_ctr++ % c->timing_every == 0;
  // a timed step keeps one stream (events on it bracket every launch)
  const bool fork = c->overlap && !c->timing_now && !c->onlydiffuse && !c->nodiffuse;
  if (!c->onlydiffuse) {
    for (unsigned i1 = 0; i1 < c->s1; i1++) {
      flip ^= enqueue_ideal(c, off, flip);
      off += 1;
      if (fork && i1 + 1 == c->s1 && hipEventRecord(c->ev_fork, c->stream) == hipSuccess &&
          hipStreamWaitEvent(c->stream2, c->ev_fork, 0) == hipSuccess) {
        // the phi sweeps that follow do not touch the charge bit of flip
        flip ^= enqueue_charge(c, off + 2 * (int)c->s0, flip, c->stream2);
        charge_queued = true;
        join_event = hipEventRecord(c->ev_join, c->stream2) == hipSuccess;
      }
      for (unsigned i0 = 0; i0 < c->s0; i0++) {
        flip ^= enqueue_hb(c, off, flip);
        off += 2;
      }
    }
  }
  if (!c->nodiffuse) {
    if (!charge_queued) {
      flip ^= enqueue_charge(c, off, flip, c->stream);
    } else if (!join_event || hipStreamWaitEvent(c->stream, c->ev_join, 0) != hipSuccess) {
      (void)hipStreamSynchronize(c->stream2);  // join on the host instead
    }
    off += 7;
  }
  return off;
}

int check_launch() {
  HIPCHECK(hipGetLastError());
  return LGV_OK;
}

int check_replica(lgv_ctx *c, int r) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (r < 0 || r >= c->nrep) return fail(LGV_ERR_RANGE, "replica %d out of range [0,%d)", r, c->nrep);
  return LGV_OK;
}

}  // namespace

extern "C" {

const char *lgv_last_error(void) { return g_err.c_str(); }
int lgv_version(void) { return 1; }

int lgv_create(const lgv_config *cfg, lgv_ctx **out) {
  if (!cfg || !out) return fail(LGV_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->nreplicas <= 0 || !cfg->replicas) return fail(LGV_ERR_ARG, "no replicas");
  if (cfg->precision != LGV_FP64 && cfg->precision != LGV_FP32)
    return fail(LGV_ERR_ARG, "precision must be LGV_FP64 or LGV_FP32");
  const unsigned s0 = cfg->scounts[0] ? cfg->scounts[0] : 2;
  const unsigned s1 = cfg->scounts[1] ? cfg->scounts[1] : 3;
  if (s0 > 64 || s1 > 64) return fail(LGV_ERR_ARG, "scounts too large");
  for (int r = 0; r < cfg->nreplicas; r++) {
    const lgv_model_params &m = cfg->replicas[r];
    if (m.NX < 4 || (m.NX & 1) || m.NX > kMaxN)
      return fail(LGV_ERR_ARG, "replica %d: NX=%d must be even and in [4, %d]", r, m.NX, kMaxN);
    if (!(m.chi > 0)) return fail(LGV_ERR_ARG, "replica %d: chi must be positive", r);
  }
  // on any failure below, everything allocated so far is released (the
  // error message of the failure is kept)
  auto cleanup = [](lgv_ctx *p) {
    const std::string msg = g_err;
    lgv_destroy(p);
    g_err = msg;
  };
  std::unique_ptr<lgv_ctx, decltype(cleanup)> c(new lgv_ctx, cleanup);
  c->device = cfg->device;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  HIPCHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIPCHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  c->fp32 = cfg->precision == LGV_FP32;
  c->s0 = s0;
  c->s1 = s1;
  c->nodiffuse = cfg->nodiffuse != 0;
  c->onlydiffuse = cfg->onlydiffuse != 0;
  c->track_energy = cfg->track_energy != 0;
  c->nrep = cfg->nreplicas;
  c->mp.assign(cfg->replicas, cfg->replicas + cfg->nreplicas);
  c->hp.resize(c->nrep);
  c->noise.resize(c->nrep);
  const size_t esz = c->fp32 ? sizeof(float) : sizeof(double);
  std::vector<int64_t> whalf, wfull;
  uint64_t buffer_bytes = 0;  // one buffer: 10 planes x 2 parities
  std::vector<ReplicaState> hs(c->nrep);
  for (int r = 0; r < c->nrep; r++) {
    const lgv_model_params &m = c->mp[r];
    const int N = m.NX;
    const int64_t half = (int64_t)N * N * N / 2;
    ReplicaParams &p = c->hp[r];
    memset(&p, 0, sizeof p);
    for (int b = 0; b < 2; b++) {
      void *field = nullptr;
      HIPCHECK(hipMalloc(&field, 20 * half * esz));
      HIPCHECK(hipMemset(field, 0, 20 * half * esz));
      c->fields.push_back(field);
      p.fields[b] = field;
    }
    p.tile_cap = 32;
    p.tile_tz = 16;
    p.hb_cap = p.hb_tz = p.chg_cap = p.chg_chunk = 0;
    p.N = N;
    p.NH = N / 2;
    p.half = half;
    p.mass0 = m.mass0;
    p.dmassdt = m.dmassdt;
    p.lambda = m.lambda;
    p.H = m.H;
    p.chi = m.chi;
    const double hx = (double)(m.LX > 0 ? m.LX : N) / N;
    p.axx = std::pow(1. / hx, 2);
    const uint32_t s32 = rank0_seed(m.seed);
    p.base_key = base_key(s32);
    c->noise[r].g = xoro_seeded((uint64_t)s32);
    hs[r] = ReplicaState{};
    hs[r].t = m.initialtime;
    whalf.push_back(half);
    wfull.push_back(2 * half);
    if ((uint64_t)half * esz > 0xFFFFFFFFull) c->wide = true;
    buffer_bytes += (uint64_t)20 * half * esz;
    // test hook: LGV_FORCE_WIDE=1 runs the 64-bit-offset kernels on any size
    if (const char *fw = getenv("LGV_FORCE_WIDE")) c->wide = c->wide || atoi(fw) != 0;
    c->maxN = std::max(c->maxN, N);
  }
  c->nt_stores = buffer_bytes > kStreamBytes;
  // test hook: LGV_NT_STORES=0/1 forces the store mode on any size
  if (const char *nt = getenv("LGV_NT_STORES")) c->nt_stores = atoi(nt) != 0;
  c->phi_march = c->fp32 ? buffer_bytes > kMarchBytesFp32 : c->nt_stores;
  c->phi_march_auto = c->phi_march;
  // one lattice whose launches each fill the GPU many times over gains
  // nothing from running the charge sweep beside the phi sweeps (256^3:
  // 1 % slower); smaller and batched lattices do (128^3 FP64 2.53e9 ->
  // 2.67e9 site-updates/s, the 96-replica FP32 batch +2 %)
  c->overlap = !(cfg->nreplicas == 1 && buffer_bytes > kMarchBytesFp32);
  // test hook: LGV_OVERLAP=0/1 forces the sequential / concurrent order
  if (const char *ov = getenv("LGV_OVERLAP")) c->overlap = atoi(ov) != 0;
  // test hook: LGV_PHI_KERNEL=site/march forces the phi half-sweep kernel
  if (const char *pk = getenv("LGV_PHI_KERNEL")) {
    const std::string k(pk);
    if (k != "site" && k != "march")
      return fail(LGV_ERR_ARG, "LGV_PHI_KERNEL=%s: expected site or march", pk);
    c->phi_march = k == "march";
  }
  HIPCHECK(hipMalloc(&c->dP, c->nrep * sizeof(ReplicaParams)));
  HIPCHECK(hipMemcpy(c->dP, c->hp.data(), c->nrep * sizeof(ReplicaParams),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMalloc(&c->dS, c->nrep * sizeof(ReplicaState)));
  HIPCHECK(hipMemcpy(c->dS, hs.data(), c->nrep * sizeof(ReplicaState), hipMemcpyHostToDevice));
  int rc;
  if ((rc = build_plan(c->plan_half, whalf, c->wide)) != LGV_OK) return rc;
  if ((rc = build_plan(c->plan_full, wfull, c->wide)) != LGV_OK) return rc;
  if ((rc = build_ideal_plan(c.get())) != LGV_OK) return rc;
  if ((rc = build_charge_plan(c.get())) != LGV_OK) return rc;
  if ((rc = build_hb_plan(c.get())) != LGV_OK) return rc;
  HIPCHECK(hipMalloc(&c->mon_phi, hb_plan(c.get()).nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMemset(c->mon_phi, 0, hb_plan(c.get()).nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMalloc(&c->mon_charge, c->plan_charge.nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMemset(c->mon_charge, 0, c->plan_charge.nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMalloc(&c->mon_ideal, c->nrep * sizeof(MonSlot)));
  HIPCHECK(hipMemset(c->mon_ideal, 0, c->nrep * sizeof(MonSlot)));
  c->mon_carry.assign(c->nrep, {lgv_monitor{}, lgv_monitor{}});
  HIPCHECK(hipMalloc(&c->eslots, c->plan_full.nblocks * sizeof(EnergySlot)));
  HIPCHECK(hipMalloc(&c->d_aos, (int64_t)stage_planes(c->maxN) * c->maxN * c->maxN * 10 *
                                     sizeof(double)));
  const int64_t part = 25 * std::max(kMeasureColumns, (int64_t)c->maxN * c->maxN);
  HIPCHECK(hipMalloc(&c->d_partXZ, part * sizeof(double)));
  HIPCHECK(hipMalloc(&c->d_partXY, part * sizeof(double)));
  HIPCHECK(hipMalloc(&c->d_meas, MeasBuf::size(c->maxN) * sizeof(double)));
  HIPCHECK(hipDeviceSynchronize());
  *out = c.release();
  return LGV_OK;
}

int lgv_destroy(lgv_ctx *c) {
  if (!c) return LGV_OK;
  // releases everything even after a failure and reports the first one
  hipError_t first = hipSuccess;
  auto note = [&](hipError_t e) {
    if (first == hipSuccess && e != hipSuccess) first = e;
  };
  note(hipSetDevice(c->device));
  if (c->stream) note(hipStreamSynchronize(c->stream));
  if (c->stream2) note(hipStreamSynchronize(c->stream2));
  if (c->graph_exec) note(hipGraphExecDestroy(c->graph_exec));
  if (c->graph_one) note(hipGraphExecDestroy(c->graph_one));
  for (hipEvent_t e : c->ev_pool) note(hipEventDestroy(e));
  for (void *f : c->fields) note(hipFree(f));
  note(hipFree(c->dP));
  note(hipFree(c->dS));
  for (Plan *pl : {&c->plan_half, &c->plan_full, &c->plan_ideal, &c->plan_charge, &c->plan_hb}) {
    note(hipFree(pl->d_blk_rep));
    note(hipFree(pl->d_blk_first));
  }
  for (void *q : {(void *)c->mon_phi, (void *)c->mon_charge, (void *)c->mon_ideal,
                  (void *)c->eslots, (void *)c->d_aos, (void *)c->d_partXZ, (void *)c->d_partXY,
                  (void *)c->d_meas})
    note(hipFree(q));
  if (c->ev_fork) note(hipEventDestroy(c->ev_fork));
  if (c->ev_join) note(hipEventDestroy(c->ev_join));
  if (c->stream2) note(hipStreamDestroy(c->stream2));
  if (c->stream) note(hipStreamDestroy(c->stream));
  delete c;
  if (first != hipSuccess)
    return fail(LGV_ERR_HIP, "lgv_destroy: %s", hipGetErrorString(first));
  return LGV_OK;
}

int lgv_set_field(lgv_ctx *c, int r, const double *aos) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!aos) return fail(LGV_ERR_ARG, "null field");
  HIPCHECK(hipSetDevice(c->device));
  const int N = c->hp[r].N;
  const int64_t P = (int64_t)N * N * 10;  // doubles per z-plane
  const int slab = slab_planes(N);
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int z0 = 0; z0 < N; z0 += slab) {
    const int nz = std::min(slab, N - z0);
    HIPCHECK(hipMemcpy(c->d_aos, aos + z0 * P, nz * P * sizeof(double), hipMemcpyHostToDevice));
    launch_aos_to_rb(c->fp32, c->dP, c->dS, r, N, z0, nz, c->d_aos, c->stream);
    if ((rc = check_launch())) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  return LGV_OK;
}

int lgv_get_field(lgv_ctx *c, int r, double *aos) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!aos) return fail(LGV_ERR_ARG, "null field");
  HIPCHECK(hipSetDevice(c->device));
  const int N = c->hp[r].N;
  const int64_t P = (int64_t)N * N * 10;  // doubles per z-plane
  const int slab = slab_planes(N);
  for (int z0 = 0; z0 < N; z0 += slab) {
    const int nz = std::min(slab, N - z0);
    launch_rb_to_aos(c->fp32, c->dP, c->dS, r, N, z0, nz, c->d_aos, c->stream);
    if ((rc = check_launch())) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(aos + z0 * P, c->d_aos, nz * P * sizeof(double), hipMemcpyDeviceToHost));
  }
  return LGV_OK;
}

int lgv_set_time(lgv_ctx *c, int r, double t) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (r >= c->nrep) return fail(LGV_ERR_RANGE, "replica out of range");
  HIPCHECK(hipSetDevice(c->device));
  launch_set_time(c->nrep, r, c->dS, t, c->stream);
  return check_launch();
}

int lgv_get_time(lgv_ctx *c, int r, double *t) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHECK(hipSetDevice(c->device));
  ReplicaState s;
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(&s, c->dS + r, sizeof s, hipMemcpyDeviceToHost));
  *t = s.t;
  return LGV_OK;
}

int lgv_step(lgv_ctx *c, double dt) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  HIPCHECK(hipSetDevice(c->device));
  int rc = ensure_coeffs(c, dt / (c->s0 * c->s1), dt, dt / (c->s1));
  if (rc) return rc;
  int flip;
  const int n = enqueue_step(c, flip);
  launch_advance(c->nrep, c->dS, dt, 0, (uint64_t)n, flip, c->stream);
  return check_launch();
}

// Releases the captured step graphs (they capture plans, tiles and dt).
static int drop_graphs(lgv_ctx *c) {
  for (hipGraphExec_t *g : {&c->graph_exec, &c->graph_one}) {
    if (*g) HIPCHECK(hipGraphExecDestroy(*g));
    *g = nullptr;
  }
  return LGV_OK;
}

// The per-block monitor slots follow the plans: before a replanning,
// fold_monitors adds them into the per-replica carry; realloc_monitors
// then allocates zeroed slots for the plans in use.
static int fold_monitors(lgv_ctx *c) {
  lgv_monitor m[3];
  for (int i = 0; i < c->nrep; i++) {
    int rc = lgv_monitor_get(c, i, m);
    if (rc) return rc;
    c->mon_carry[i] = {m[0], m[1]};
  }
  return LGV_OK;
}
static int realloc_monitors(lgv_ctx *c) {
  HIPCHECK(hipFree(c->mon_phi));
  HIPCHECK(hipFree(c->mon_charge));
  c->mon_phi = c->mon_charge = nullptr;
  HIPCHECK(hipMalloc(&c->mon_phi, hb_plan(c).nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMemset(c->mon_phi, 0, hb_plan(c).nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMalloc(&c->mon_charge, c->plan_charge.nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMemset(c->mon_charge, 0, c->plan_charge.nblocks * sizeof(MonSlot)));
  return LGV_OK;
}

int lgv_use_graph(lgv_ctx *c, int enable) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (enable < 0 || enable > 1024) return fail(LGV_ERR_ARG, "graph steps %d out of range", enable);
  HIPCHECK(hipSetDevice(c->device));
  c->use_graph = enable != 0;
  if (enable > 0 && enable != c->graph_steps) {
    int rc = drop_graphs(c);
    if (rc) return rc;
    c->graph_steps = enable;
  }
  return LGV_OK;
}

int lgv_run(lgv_ctx *c, double dt, int nsteps) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (nsteps < 0) return fail(LGV_ERR_ARG, "nsteps < 0");
  HIPCHECK(hipSetDevice(c->device));
  int rc = ensure_coeffs(c, dt / (c->s0 * c->s1), dt, dt / (c->s1));
  if (rc) return rc;
  if (!c->use_graph || c->timing) {
    for (int i = 0; i < nsteps; i++) {
      int flip;
      const int n = enqueue_step(c, flip);
      launch_advance(c->nrep, c->dS, dt, 1, (uint64_t)n, flip, c->stream);
    }
    return check_launch();
  }
  if (!c->graph_exec || c->graph_dt != dt) {
    if ((rc = drop_graphs(c))) return rc;
    // graph_steps steps in one graph, and (when that is more than one) one
    // step for the remainder of a run
    for (int g = 0; g < (c->graph_steps > 1 ? 2 : 1); g++) {
      const int k = g == 0 ? c->graph_steps : 1;
      hipGraph_t graph;
      HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < k; i++) {
        int flip;
        const int n = enqueue_step(c, flip);
        launch_advance(c->nrep, c->dS, dt, 1, (uint64_t)n, flip, c->stream);
      }
      HIPCHECK(hipStreamEndCapture(c->stream, &graph));
      // the captured graph is released whether or not it instantiates
      const hipError_t ei = hipGraphInstantiate(g == 0 ? &c->graph_exec : &c->graph_one,
                                                graph, nullptr, nullptr, 0);
      HIPCHECK(hipGraphDestroy(graph));
      if (ei != hipSuccess) {
        (g == 0 ? c->graph_exec : c->graph_one) = nullptr;
        drop_graphs(c);
        return fail(LGV_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
      }
    }
    c->graph_dt = dt;
  }
  const int K = c->graph_steps;
  for (int i = 0; i < nsteps / K; i++) HIPCHECK(hipGraphLaunch(c->graph_exec, c->stream));
  for (int i = 0; i < nsteps % K; i++) HIPCHECK(hipGraphLaunch(c->graph_one, c->stream));
  return check_launch();
}

int lgv_ideal_step(lgv_ctx *c, double dt) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  HIPCHECK(hipSetDevice(c->device));
  int rc = ensure_coeffs(c, c->coef_hb, c->coef_charge, dt);
  if (rc) return rc;
  c->timing_now = c->timing;  // single sweeps are always timed
  const int flip = enqueue_ideal(c, 0, 0);
  launch_advance(c->nrep, c->dS, 0., 0, 1, flip, c->stream);
  return check_launch();
}

int lgv_hb_phi_step(lgv_ctx *c, double dt) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  HIPCHECK(hipSetDevice(c->device));
  int rc = ensure_coeffs(c, dt, c->coef_charge, c->coef_ideal);
  if (rc) return rc;
  c->timing_now = c->timing;  // single sweeps are always timed
  const int flip = enqueue_hb(c, 0, 0);
  launch_advance(c->nrep, c->dS, 0., 0, 2, flip, c->stream);
  return check_launch();
}

int lgv_charge_step(lgv_ctx *c, double dt) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  HIPCHECK(hipSetDevice(c->device));
  int rc = ensure_coeffs(c, c->coef_hb, dt, c->coef_ideal);
  if (rc) return rc;
  c->timing_now = c->timing;  // single sweeps are always timed
  const int flip = enqueue_charge(c, 0, 0, c->stream);
  launch_advance(c->nrep, c->dS, 0., 0, 7, flip, c->stream);
  return check_launch();
}

int lgv_init_gaussian_charges(lgv_ctx *c, int r) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  const int N = c->hp[r].N;
  const size_t V = (size_t)N * N * N;
  std::vector<double> u(V * 10);
  if ((rc = lgv_get_field(c, r, u.data()))) return rc;
  double sum[10] = {0};
  const double chi = c->mp[r].chi;
  HostNoise &ng = c->noise[r];
  for (size_t s = 0; s < V; s++)
    for (int L = 4; L < 10; L++) {
      u[s * 10 + L] = std::sqrt(chi) * ng.normal();
      sum[L] += u[s * 10 + L];
    }
  const double vol = (double)N * N * N;
  for (size_t s = 0; s < V; s++)
    for (int L = 4; L < 10; L++) u[s * 10 + L] -= sum[L] / vol;
  return lgv_set_field(c, r, u.data());
}

int lgv_thermalize(lgv_ctx *c, double dt, int nsteps) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  int rc;
  for (int r = 0; r < c->nrep; r++)
    if ((rc = lgv_init_gaussian_charges(c, r))) return rc;
  for (int i = 0; i < nsteps; i++) {
    const int substeps = 6;
    for (int j = 0; j < substeps; j++)
      if ((rc = lgv_hb_phi_step(c, dt / substeps))) return rc;
  }
  return LGV_OK;
}

int lgv_monitor_get(lgv_ctx *c, int r, lgv_monitor out[3]) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  MonSlot *src[2] = {c->mon_phi, c->mon_charge};
  const Plan *plan[2] = {&hb_plan(c), &c->plan_charge};
  for (int k = 0; k < 2; k++) {
    const int b0 = plan[k]->first[r];
    const int b1 = r + 1 < c->nrep ? plan[k]->first[r + 1] : plan[k]->nblocks;
    std::vector<MonSlot> slots(b1 - b0);
    HIPCHECK(hipMemcpy(slots.data(), src[k] + b0, slots.size() * sizeof(MonSlot),
                       hipMemcpyDeviceToHost));
    lgv_monitor m = c->mon_carry[r][k];
    for (const MonSlot &s : slots) {
      m.down_counts += s.down_counts;
      m.up_counts += s.up_counts;
      m.up_yes += s.up_yes;
      m.up_no += s.up_no;
      m.down_dS += s.down_dS;
      m.up_dS += s.up_dS;
    }
    out[k] = m;
  }
  MonSlot mi;
  HIPCHECK(hipMemcpy(&mi, c->mon_ideal + r, sizeof mi, hipMemcpyDeviceToHost));
  out[2] = lgv_monitor{mi.down_counts, mi.up_counts, mi.up_yes, mi.up_no, mi.down_dS, mi.up_dS};
  return LGV_OK;
}

int lgv_energy_get(lgv_ctx *c, int r, double out[2]) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  ReplicaState s;
  HIPCHECK(hipMemcpy(&s, c->dS + r, sizeof s, hipMemcpyDeviceToHost));
  out[0] = s.E_old;
  out[1] = s.E_new;
  return LGV_OK;
}

int lgv_rng_get_state(lgv_ctx *c, int r, lgv_rng_state *st) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  ReplicaState s;
  HIPCHECK(hipMemcpy(&s, c->dS + r, sizeof s, hipMemcpyDeviceToHost));
  st->s0 = c->noise[r].g.s0;
  st->s1 = c->noise[r].g.s1;
  st->saved_normal = c->noise[r].saved;
  st->has_saved_normal = c->noise[r].has_saved;
  st->sweep_ctr = s.sweep_ctr;
  return LGV_OK;
}

int lgv_rng_set_state(lgv_ctx *c, int r, const lgv_rng_state *st) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->noise[r].g.s0 = st->s0;
  c->noise[r].g.s1 = st->s1;
  c->noise[r].saved = st->saved_normal;
  c->noise[r].has_saved = st->has_saved_normal;
  ReplicaState s;
  HIPCHECK(hipMemcpy(&s, c->dS + r, sizeof s, hipMemcpyDeviceToHost));
  s.sweep_ctr = st->sweep_ctr;
  HIPCHECK(hipMemcpy(c->dS + r, &s, sizeof s, hipMemcpyHostToDevice));
  return LGV_OK;
}

int lgv_measure(lgv_ctx *c, int r, double *wall, double *wallphase, double *oavg) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!wall || !wallphase || !oavg) return fail(LGV_ERR_ARG, "null output");
  HIPCHECK(hipSetDevice(c->device));
  const int N = c->hp[r].N;
  const MeasBuf mb(c->d_meas, N);
  launch_measure(c->fp32, c->dP, c->dS, r, N, c->d_partXZ, c->d_partXY, mb.wall, mb.wallphase,
                 c->stream);
  if ((rc = check_launch())) return rc;
  std::vector<double> h(3 * 25 * (size_t)N);  // walls and phase walls, one copy
  HIPCHECK(hipMemcpyAsync(h.data(), mb.wall, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  memcpy(wall, h.data(), 3 * 11 * N * sizeof(double));
  memcpy(wallphase, h.data() + 3 * 11 * N, 3 * 14 * N * sizeof(double));
  // computeDerivedObs's OAverage (measurer.cxx:288-300): mean of wallX rows
  for (int l = 0; l < 11; l++) {
    double s = 0.;
    for (int i = 0; i < N; i++) s += wall[l * N + i];
    oavg[l] = s / (double)N;
  }
  return LGV_OK;
}

int lgv_timing_enable(lgv_ctx *c, int enable) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (enable < 0) return fail(LGV_ERR_ARG, "timing_enable %d < 0", enable);
  HIPCHECK(hipSetDevice(c->device));
  int rc = drain_timing(c);
  if (rc) return rc;
  c->timing = enable != 0;
  c->timing_now = c->timing;
  c->timing_every = enable > 1 ? enable : 1;
  c->timing_ctr = 0;
  for (int k = 0; k < LGV_NCLASS; k++) {
    c->t_ms[k] = 0.;
    c->t_launch[k] = 0;
  }
  return LGV_OK;
}

int lgv_timing_get(lgv_ctx *c, double ms[LGV_NCLASS], long long launches[LGV_NCLASS]) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  HIPCHECK(hipSetDevice(c->device));
  int rc = drain_timing(c);
  if (rc) return rc;
  for (int k = 0; k < LGV_NCLASS; k++) {
    ms[k] = c->t_ms[k];
    launches[k] = c->t_launch[k];
  }
  return LGV_OK;
}

// rotate_to_vev (measurer.cxx:231-286): rotates the k-space walls into the
// frame of the k = 0 phi direction.  wk[11][nk][2] -> rot[14][nk][2].
static void rotate_to_vev(int nk, const double *wk, double *rot) {
  using cd = std::complex<double>;
  auto W = [&](int a, int k) { return cd(wk[(a * nk + k) * 2], wk[(a * nk + k) * 2 + 1]); };
  auto contract = [](const cd *n, const cd *rho, cd *out) {
    out[0] = -(n[1] * rho[0]) - n[2] * rho[1] - n[3] * rho[2];
    out[1] = n[0] * rho[0] + n[3] * rho[4] - n[2] * rho[5];
    out[2] = n[0] * rho[1] - n[3] * rho[3] + n[1] * rho[5];
    out[3] = n[0] * rho[2] + n[2] * rho[3] - n[1] * rho[4];
  };
  cd n[4];
  double norm = 0.;
  for (int a = 0; a < 4; a++) {
    n[a] = W(a, 0);
    norm += std::norm(n[a]);
  }
  norm = std::sqrt(norm);
  if (norm < std::numeric_limits<double>::min()) {
    n[0] = 1; n[1] = n[2] = n[3] = 0;
  } else {
    for (int a = 0; a < 4; a++) n[a] /= norm;
  }
  for (int k = 0; k < nk; k++) {
    cd phi[4], rho[6], drho[6], phir[4], A[4], V[4];
    for (int a = 0; a < 4; a++) phi[a] = W(a, k);
    for (int ab = 0; ab < 6; ab++) rho[ab] = W(4 + ab, k);
    for (int l = 0; l < 3; l++) { drho[l] = rho[l + 3]; drho[l + 3] = rho[l]; }
    const cd phi2k = W(10, k);
    cd phis(0.);
    for (int a = 0; a < 4; a++) phis += n[a] * phi[a];
    for (int a = 0; a < 4; a++) phir[a] = phi[a] - n[a] * phis;
    contract(n, rho, A);
    contract(n, drho, V);
    auto put = [&](int a, cd v) {
      rot[(a * nk + k) * 2] = v.real();
      rot[(a * nk + k) * 2 + 1] = v.imag();
    };
    put(0, phis);
    for (int a = 0; a < 4; a++) {
      put(1 + a, phir[a]);
      put(5 + a, A[a]);
      put(9 + a, V[a]);
    }
    put(13, phi2k);
  }
}

int lgv_measure_full(lgv_ctx *c, int r, lgv_measurement *m) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!m) return fail(LGV_ERR_ARG, "null measurement");
  HIPCHECK(hipSetDevice(c->device));
  const int N = c->hp[r].N, nk = N / 2 + 1;
  std::vector<double> wall(3 * 11 * N), wallphase(3 * 14 * N), oavg(11);
  const MeasBuf mb(c->d_meas, N);
  launch_measure(c->fp32, c->dP, c->dS, r, N, c->d_partXZ, c->d_partXY, mb.wall, mb.wallphase,
                 c->stream);
  launch_wall_dft(N, mb.wall, mb.wallphase, mb.wallk, mb.wallphasek, c->stream);
  if ((rc = check_launch())) return rc;
  std::vector<double> h(mb.total);  // everything in one copy
  HIPCHECK(hipMemcpyAsync(h.data(), mb.wall, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  std::vector<double> wk(3 * 11 * nk * 2), wpk(3 * 14 * nk * 2);
  memcpy(wall.data(), h.data(), wall.size() * sizeof(double));
  memcpy(wallphase.data(), h.data() + (mb.wallphase - mb.wall), wallphase.size() * sizeof(double));
  memcpy(wk.data(), h.data() + (mb.wallk - mb.wall), wk.size() * sizeof(double));
  memcpy(wpk.data(), h.data() + (mb.wallphasek - mb.wall), wpk.size() * sizeof(double));
  for (int l = 0; l < 11; l++) {
    double s = 0.;
    for (int i = 0; i < N; i++) s += wall[l * N + i];
    oavg[l] = s / (double)N;
  }
  if (m->wall) memcpy(m->wall, wall.data(), wall.size() * sizeof(double));
  if (m->wallphase) memcpy(m->wallphase, wallphase.data(), wallphase.size() * sizeof(double));
  if (m->oavg) memcpy(m->oavg, oavg.data(), 11 * sizeof(double));
  if (m->wall_k) memcpy(m->wall_k, wk.data(), wk.size() * sizeof(double));
  if (m->wallphase_k) memcpy(m->wallphase_k, wpk.data(), wpk.size() * sizeof(double));
  if (m->wall_k_rotated)
    for (int d = 0; d < 3; d++)
      rotate_to_vev(nk, wk.data() + (size_t)d * 11 * nk * 2,
                    m->wall_k_rotated + (size_t)d * 14 * nk * 2);
  return LGV_OK;
}

int lgv_set_mass(lgv_ctx *c, int r, double mass0, double dmassdt) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (r >= c->nrep) return fail(LGV_ERR_RANGE, "replica out of range");
  HIPCHECK(hipSetDevice(c->device));
  for (int i = 0; i < c->nrep; i++) {
    if (r >= 0 && i != r) continue;
    c->mp[i].mass0 = c->hp[i].mass0 = mass0;
    c->mp[i].dmassdt = c->hp[i].dmassdt = dmassdt;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(c->dP, c->hp.data(), c->nrep * sizeof(ReplicaParams),
                     hipMemcpyHostToDevice));
  if (int rg = drop_graphs(c)) return rg;
  return LGV_OK;
}

int lgv_set_ideal_tile(lgv_ctx *c, int r, int txh_cap, int tz) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (r >= c->nrep) return fail(LGV_ERR_RANGE, "replica out of range");
  if (txh_cap < 2 || txh_cap > 64 || tz < 1 || tz > 1024)
    return fail(LGV_ERR_ARG, "tile cap %d / tz %d out of range", txh_cap, tz);
  HIPCHECK(hipSetDevice(c->device));
  std::vector<ReplicaParams> old = c->hp;
  for (int i = 0; i < c->nrep; i++) {
    if (r >= 0 && i != r) continue;
    c->hp[i].tile_cap = txh_cap;
    c->hp[i].tile_tz = tz;
  }
  int rc = build_ideal_plan(c);
  if (rc) {
    c->hp = old;
    build_ideal_plan(c);
    return rc;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(c->dP, c->hp.data(), c->nrep * sizeof(ReplicaParams),
                     hipMemcpyHostToDevice));
  if (int rg = drop_graphs(c)) return rg;
  return LGV_OK;
}

int lgv_set_sweep_tiles(lgv_ctx *c, int r, int hb_cap, int hb_tz, int chg_cap,
                        int chg_chunk) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (r >= c->nrep) return fail(LGV_ERR_RANGE, "replica out of range");
  auto bad = [](int v, int lo, int hi) { return v != 0 && (v < lo || v > hi); };
  if (bad(hb_cap, 4, 64) || bad(hb_tz, 1, 1024) || bad(chg_cap, 4, 64) ||
      bad(chg_chunk, 2, 1024) || chg_chunk % 2 != 0)
    return fail(LGV_ERR_ARG, "sweep tile %d/%d, %d/%d out of range", hb_cap, hb_tz, chg_cap,
                chg_chunk);
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (int rm = fold_monitors(c)) return rm;
  std::vector<ReplicaParams> old = c->hp;
  for (int i = 0; i < c->nrep; i++) {
    if (r >= 0 && i != r) continue;
    c->hp[i].hb_cap = hb_cap;
    c->hp[i].hb_tz = hb_tz;
    c->hp[i].chg_cap = chg_cap;
    c->hp[i].chg_chunk = chg_chunk;
  }
  int rc = build_hb_plan(c);
  if (!rc) rc = build_charge_plan(c);
  if (rc) {
    c->hp = old;
    build_hb_plan(c);
    build_charge_plan(c);
  }
  if (int rm = realloc_monitors(c)) return rm;
  HIPCHECK(hipMemcpy(c->dP, c->hp.data(), c->nrep * sizeof(ReplicaParams),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipDeviceSynchronize());
  if (int rg = drop_graphs(c)) return rg;
  return rc;
}

int lgv_set_phi_kernel(lgv_ctx *c, int kernel) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  if (kernel != LGV_PHI_AUTO && kernel != LGV_PHI_SITE && kernel != LGV_PHI_MARCH)
    return fail(LGV_ERR_ARG, "phi kernel %d: expected LGV_PHI_AUTO, _SITE or _MARCH", kernel);
  const bool march = kernel == LGV_PHI_AUTO ? c->phi_march_auto : kernel == LGV_PHI_MARCH;
  if (march == c->phi_march) return LGV_OK;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (int rm = fold_monitors(c)) return rm;
  c->phi_march = march;
  if (int rm = realloc_monitors(c)) return rm;
  HIPCHECK(hipDeviceSynchronize());
  return drop_graphs(c);
}

int lgv_get_phi_kernel(lgv_ctx *c, int *kernel) {
  if (!c || !kernel) return fail(LGV_ERR_ARG, "null argument");
  *kernel = c->phi_march ? LGV_PHI_MARCH : LGV_PHI_SITE;
  return LGV_OK;
}

int lgv_monitor_reset(lgv_ctx *c) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemset(c->mon_phi, 0, hb_plan(c).nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMemset(c->mon_charge, 0, c->plan_charge.nblocks * sizeof(MonSlot)));
  HIPCHECK(hipMemset(c->mon_ideal, 0, c->nrep * sizeof(MonSlot)));
  HIPCHECK(hipDeviceSynchronize());
  for (auto &cr : c->mon_carry) cr = {lgv_monitor{}, lgv_monitor{}};
  return LGV_OK;
}

int lgv_synchronize(lgv_ctx *c) {
  if (!c) return fail(LGV_ERR_ARG, "null context");
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return LGV_OK;
}

int lgv_field_device_ptr(lgv_ctx *c, int r, void **ptr, int *elem_bytes) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  *ptr = c->fields[r];
  *elem_bytes = c->fp32 ? 4 : 8;
  return LGV_OK;
}

}  // extern "C"
