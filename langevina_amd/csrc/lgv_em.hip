// lgv_em.hip -- Model A Euler--Maruyama stepper and its adjoint (new
// functionality: the reference has no EM stepper; its Metropolis "Model A"
// is EulerLangevinHB, ModelG/Stepper.cxx:287-389, and the forward-Euler
// recursion is the one of ModelG/freeCase.ipynb cells 5, 7).
//
//   phi_{n+1} = phi_n - G dt (-Lap phi_n + m2 phi_n + lambda |phi_n|^2 phi_n - H e0)
//               + sqrt(2 G dt) xi_n
//
// with the reference's lattice Laplacian (six neighbours minus 6 phi) and
// xi_n from per-site counter streams (replayed exactly in the backward
// pass).  Fields are caller-owned device arrays phi[a][z][y][x] (FP64).
// The loss is  mean_w(phibar) + V var_w(phibar)  over the last `window`
// steps, phibar_n = volume mean of phi_0 after step n.  The adjoint uses
// the symmetric one-step Jacobian J = I - G dt (-Lap + m2 + lambda (|phi|^2
// + 2 phi phi^T)), checkpoints every K steps and recomputation inside a
// segment.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "lgv.h"
#include "lgv_device.h"

using namespace lgv;

namespace {

constexpr int kEmBlock = 256;
constexpr uint64_t kEmTag = 0x4000000000000000ULL;

// Tile of the two stencil marches: TX sites in x by TY rows in y, chunks of
// TZ planes in z.  A block's plane (tile plus a one-site ring, no corners)
// sits in LDS; each thread keeps its own column's planes z-1, z, z+1 in
// registers.
constexpr int kEmGrid = 400;  // (TY+2)(TX+2) <= 6 * 66
__host__ __device__ inline void em_tile(int N, int &TX, int &TY, int &TZ) {
  TX = N < 64 ? N : 64;
  TY = kEmBlock / TX;
  if (TY > N) TY = N;
  TZ = N < 8 ? N : 8;
}
__host__ __device__ inline int em_blocks(int N) {
  int TX, TY, TZ;
  em_tile(N, TX, TY, TZ);
  return ((N + TX - 1) / TX) * ((N + TY - 1) / TY) * ((N + TZ - 1) / TZ);
}

struct GradSlot {
  double m2_s, m2_c, h_s, h_c, l_s, l_c;
};

__device__ __forceinline__ void em_noise(uint64_t skey, int64_t s, int kind, double xi[4]) {
  Xoro g = xoro_seeded(site_key(skey, (uint64_t)s));
  if (kind == 0) {
    for (int a = 0; a < 4; a++) xi[a] = variance1(g);
  } else {
    for (int a = 0; a < 4; a += 2) {
      const double u1 = canonical(xoro_next(g));
      const double u2 = canonical(xoro_next(g));
      const double r = sqrt(-2.0 * log(1.0 - u1));
      const double th = 6.283185307179586 * u2;
      xi[a] = r * cos(th);
      xi[a + 1] = r * sin(th);
    }
  }
}

// Volume sums of one launch: NS = 0 (none), 1 (phi_0: phibar) or 5 (phi_0..3
// and |phi|^2: the observables of a measurement).  Every block stores its
// compensated partials in its own slot.  The slots of step n are merged,
// in block order, by block 0 of the launch of step n+1 (the kernel
// boundary orders the two, so no fence or atomic is needed), or by
// k_em_merge after the last step of a run.
constexpr int kObsSums = 5;
constexpr int kPartStride = 2 * kObsSums;  // doubles per block slot
constexpr int kNobs = 6;  // phibar_0..3, <|phi|^2>, V |phibar|^2
static_assert(kNobs == LGV_EM_NOBS, "observable count of the C ABI");

// Slots an earlier launch left, and where their merged values go.
struct EmMerge {
  const double *slots;  // nblocks * kPartStride doubles, or nullptr: nothing
  int ns;               // 1 or kObsSums sums per slot
  int nblocks;
  double *phibar;       // volume mean of phi_0, or nullptr
  double *obs;          // kNobs observables (ns == kObsSums), or nullptr
};

template <int NS>
__device__ __forceinline__ void wave_merge(KSum *k) {
#pragma unroll
  for (int i = 0; i < NS; i++)
    for (int o = 32; o > 0; o >>= 1) {
      KSum t;
      t.s = __shfl_xor(k[i].s, o, 64);
      t.c = __shfl_xor(k[i].c, o, 64);
      ksum_merge(k[i], t);
    }
}

// This block's NS sums into its slot.
template <int NS>
__device__ void em_store_slot(KSum *acc, double *slots) {
  __shared__ KSum sh[NS][kEmBlock / 64];
  wave_merge<NS>(acc);
  if ((threadIdx.x & 63) == 0)
    for (int i = 0; i < NS; i++) sh[i][threadIdx.x >> 6] = acc[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double *slot = slots + (size_t)blockIdx.x * kPartStride;
    for (int i = 0; i < NS; i++) {
      KSum a = sh[i][0];
      for (int q = 1; q < kEmBlock / 64; q++) ksum_merge(a, sh[i][q]);
      slot[2 * i] = a.s;
      slot[2 * i + 1] = a.c;
    }
  }
}

// Merges the slots of m in block order with the whole (kEmBlock-thread)
// block and writes phibar / obs; every thread of the block calls it.
template <int NS>
__device__ void em_merge_ns(const EmMerge &m, int64_t V) {
  __shared__ KSum sh[NS][kEmBlock / 64];
  KSum b[NS];
  for (int i = 0; i < NS; i++) b[i] = KSum{0., 0.};
  for (int q = threadIdx.x; q < m.nblocks; q += kEmBlock) {
    const double *slot = m.slots + (size_t)q * kPartStride;
    for (int i = 0; i < NS; i++) ksum_merge(b[i], KSum{slot[2 * i], slot[2 * i + 1]});
  }
  wave_merge<NS>(b);
  if ((threadIdx.x & 63) == 0)
    for (int i = 0; i < NS; i++) sh[i][threadIdx.x >> 6] = b[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double v[NS];
    for (int i = 0; i < NS; i++) {
      KSum a = sh[i][0];
      for (int q = 1; q < kEmBlock / 64; q++) ksum_merge(a, sh[i][q]);
      v[i] = ksum_value(a) / (double)V;
    }
    if (m.phibar) *m.phibar = v[0];
    if constexpr (NS == kObsSums) {
      if (m.obs) {
        for (int i = 0; i < NS; i++) m.obs[i] = v[i];
        m.obs[5] = (double)V * (((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3]);
      }
    }
  }
  __syncthreads();
}
__device__ __forceinline__ void em_merge(const EmMerge &m, int64_t V) {
  if (m.ns == kObsSums) em_merge_ns<kObsSums>(m, V);
  else em_merge_ns<1>(m, V);
}

// Adds one site's phi to the NS sums.
template <int NS>
__device__ __forceinline__ void em_accumulate(KSum *acc, const double o[4]) {
  if constexpr (NS >= 1) ksum_add(acc[0], o[0]);
  if constexpr (NS == kObsSums) {
    ksum_add(acc[1], o[1]);
    ksum_add(acc[2], o[2]);
    ksum_add(acc[3], o[3]);
    ksum_add(acc[4], ((o[0] * o[0] + o[1] * o[1]) + o[2] * o[2]) + o[3] * o[3]);
  }
}

// Zeroes n doubles.  A kernel rather than hipMemsetAsync, so that a caller
// capturing the backward pass into a graph replays a plain kernel node.
__global__ void k_em_zero(double *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.;
}

// Geometry of one march block: tile origin and extent, this thread's column
// and its ring site.
struct EmMarch {
  int N, x0, y0, z0, tw, th, tz, GW;
  int tx, ty, x, y;
  bool own, halo;
  int hx, hy, hpos, pos;
  __device__ explicit EmMarch(int n) : N(n) {
    int TX, TY, TZ;
    em_tile(N, TX, TY, TZ);
    const int ntx = (N + TX - 1) / TX, nty = (N + TY - 1) / TY;
    int bx, by, bz;
    yband_tile(blockIdx.x, ntx, nty, bx, by, bz);
    x0 = bx * TX;
    y0 = by * TY;
    z0 = bz * TZ;
    tw = min(TX, N - x0);
    th = min(TY, N - y0);
    tz = min(TZ, N - z0);
    GW = tw + 2;
    const int t = threadIdx.x;
    tx = t % TX;
    ty = t / TX;
    own = t < TX * TY && tx < tw && ty < th;
    x = x0 + tx;
    y = y0 + ty;
    pos = (ty + 1) * GW + (tx + 1);
    halo = t < 2 * th + 2 * tw;
    int hlx, hly;
    if (t < th) { hlx = -1; hly = t; }
    else if (t < 2 * th) { hlx = tw; hly = t - th; }
    else if (t < 2 * th + tw) { hlx = t - 2 * th; hly = -1; }
    else { hlx = t - 2 * th - tw; hly = th; }
    hx = x0 + hlx;
    hx = hx < 0 ? hx + N : (hx >= N ? hx - N : hx);
    hy = y0 + hly;
    hy = hy < 0 ? hy + N : (hy >= N ? hy - N : hy);
    hpos = (hly + 1) * GW + (hlx + 1);
  }
  __device__ uint32_t site(int xx, int yy, int z) const {
    return ((uint32_t)z * N + yy) * N + xx;
  }
};

// One forward step in -> out, marching z with the plane in LDS; the NS
// volume sums of `out` go to `red`.
template <int NS>
__global__ __launch_bounds__(kEmBlock) void k_em_step(const double *__restrict__ in,
                                                      double *__restrict__ out, int N,
                                                      double m2, double lam, double H,
                                                      double gdt, double amp, uint64_t skey,
                                                      int kind, double *slots, EmMerge prev) {
  __shared__ double Q[2][4][kEmGrid];
  const int64_t V = (int64_t)N * N * N;
  if (prev.slots && blockIdx.x == 0) em_merge(prev, V);
  const EmMarch m(N);
  KSum acc[NS > 0 ? NS : 1];
  for (int i = 0; i < NS; i++) acc[i] = KSum{0., 0.};
  double cm[4], cc[4], cp[4];  // own column at z-1, z, z+1
  if (m.own) {
    const int zm = m.z0 == 0 ? N - 1 : m.z0 - 1;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      cm[a] = in[a * V + m.site(m.x, m.y, zm)];
      cc[a] = in[a * V + m.site(m.x, m.y, m.z0)];
    }
  }
  for (int j = 0; j < m.tz; j++) {
    const int z = m.z0 + j;
    const int zp = z + 1 == N ? 0 : z + 1;
    double(*g)[kEmGrid] = Q[j & 1];
    if (m.own) {
#pragma unroll
      for (int a = 0; a < 4; a++) {
        cp[a] = in[a * V + m.site(m.x, m.y, zp)];
        g[a][m.pos] = cc[a];
      }
    }
    if (m.halo) {
#pragma unroll
      for (int a = 0; a < 4; a++) g[a][m.hpos] = in[a * V + m.site(m.hx, m.hy, z)];
    }
    __syncthreads();
    if (m.own) {
      const uint32_t s = m.site(m.x, m.y, z);
      const double s2 = cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2] + cc[3] * cc[3];
      double xi[4], o[4];
      em_noise(skey, s, kind, xi);
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const double nb = (g[a][m.pos + 1] + g[a][m.pos - 1]) +
                          (g[a][m.pos + m.GW] + g[a][m.pos - m.GW]) + (cp[a] + cm[a]);
        const double lap = nb - 6.0 * cc[a];
        double drift = (-lap + m2 * cc[a]) + (lam * s2) * cc[a];
        if (a == 0) drift = drift - H;
        o[a] = (cc[a] - gdt * drift) + amp * xi[a];
        out[a * V + s] = o[a];
      }
      em_accumulate<NS>(acc, o);
#pragma unroll
      for (int a = 0; a < 4; a++) {
        cm[a] = cc[a];
        cc[a] = cp[a];
      }
    }
  }
  if constexpr (NS > 0) em_store_slot<NS>(acc, slots);
}

// The observable sums of a state without stepping it (the measurement
// before the first step); the same blocks and slots as k_em_step.
__global__ __launch_bounds__(kEmBlock) void k_em_observe(const double *__restrict__ in, int N,
                                                         double *slots) {
  const int64_t V = (int64_t)N * N * N;
  const EmMarch m(N);
  KSum acc[kObsSums];
  for (int i = 0; i < kObsSums; i++) acc[i] = KSum{0., 0.};
  if (m.own) {
    for (int j = 0; j < m.tz; j++) {
      const uint32_t s = m.site(m.x, m.y, m.z0 + j);
      double o[4];
#pragma unroll
      for (int a = 0; a < 4; a++) o[a] = in[a * V + s];
      em_accumulate<kObsSums>(acc, o);
    }
  }
  em_store_slot<kObsSums>(acc, slots);
}

// The merge of the last launch of a run.
__global__ __launch_bounds__(kEmBlock) void k_em_merge(EmMerge m, int64_t V) { em_merge(m, V); }

// One adjoint step: lam_out = J(phi)^T (lam_in + dl e0), marching z with
// lam_in's plane in LDS; accumulates the parameter gradients into the
// block's slot.
__global__ __launch_bounds__(kEmBlock) void k_em_adjoint(const double *__restrict__ phi,
                                                         const double *__restrict__ lin,
                                                         double *__restrict__ lout, int N,
                                                         double m2, double lam, double gdt,
                                                         double dl,
                                                         GradSlot *__restrict__ slots) {
  __shared__ double Q[2][4][kEmGrid];
  const int64_t V = (int64_t)N * N * N;
  const EmMarch m(N);
  KSum ga{0., 0.}, gb{0., 0.}, gc{0., 0.};  // d/d m2, d/d H, d/d lambda
  double lm[4], lc[4], lp[4];  // lam_in of the own column at z-1, z, z+1
  if (m.own) {
    const int zm = m.z0 == 0 ? N - 1 : m.z0 - 1;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      lm[a] = lin[a * V + m.site(m.x, m.y, zm)];
      lc[a] = lin[a * V + m.site(m.x, m.y, m.z0)];
    }
  }
  for (int j = 0; j < m.tz; j++) {
    const int z = m.z0 + j;
    const int zp = z + 1 == N ? 0 : z + 1;
    double(*g)[kEmGrid] = Q[j & 1];
    double c[4];
    if (m.own) {
      const uint32_t s = m.site(m.x, m.y, z);
#pragma unroll
      for (int a = 0; a < 4; a++) {
        lp[a] = lin[a * V + m.site(m.x, m.y, zp)];
        c[a] = phi[a * V + s];
        g[a][m.pos] = lc[a];
      }
    }
    if (m.halo) {
#pragma unroll
      for (int a = 0; a < 4; a++) g[a][m.hpos] = lin[a * V + m.site(m.hx, m.hy, z)];
    }
    __syncthreads();
    if (m.own) {
      const uint32_t s = m.site(m.x, m.y, z);
      double l[4];
#pragma unroll
      for (int a = 0; a < 4; a++) l[a] = lc[a] + (a == 0 ? dl : 0.0);
      const double s2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
      const double pl = c[0] * l[0] + c[1] * l[1] + c[2] * l[2] + c[3] * l[3];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const double e = a == 0 ? dl : 0.0;
        const double nb = ((g[a][m.pos + 1] + e) + (g[a][m.pos - 1] + e)) +
                          ((g[a][m.pos + m.GW] + e) + (g[a][m.pos - m.GW] + e)) +
                          ((lp[a] + e) + (lm[a] + e));
        const double lapl = nb - 6.0 * l[a];
        const double jt = (-lapl + m2 * l[a]) + lam * (s2 * l[a] + 2.0 * c[a] * pl);
        lout[a * V + s] = l[a] - gdt * jt;
      }
      ksum_add(ga, -gdt * pl);
      ksum_add(gb, gdt * l[0]);
      ksum_add(gc, -gdt * s2 * pl);
#pragma unroll
      for (int a = 0; a < 4; a++) {
        lm[a] = lc[a];
        lc[a] = lp[a];
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    KSum t;
    t.s = __shfl_xor(ga.s, o, 64); t.c = __shfl_xor(ga.c, o, 64); ksum_merge(ga, t);
    t.s = __shfl_xor(gb.s, o, 64); t.c = __shfl_xor(gb.c, o, 64); ksum_merge(gb, t);
    t.s = __shfl_xor(gc.s, o, 64); t.c = __shfl_xor(gc.c, o, 64); ksum_merge(gc, t);
  }
  __shared__ KSum sh[3][kEmBlock / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[0][w] = ga; sh[1][w] = gb; sh[2][w] = gc; }
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum A = sh[0][0], B = sh[1][0], Cc = sh[2][0];
    for (int i = 1; i < kEmBlock / 64; i++) {
      ksum_merge(A, sh[0][i]); ksum_merge(B, sh[1][i]); ksum_merge(Cc, sh[2][i]);
    }
    GradSlot &gs = slots[blockIdx.x];
    KSum mm{gs.m2_s, gs.m2_c}, hh{gs.h_s, gs.h_c}, ll{gs.l_s, gs.l_c};
    ksum_merge(mm, A); ksum_merge(hh, B); ksum_merge(ll, Cc);
    gs = GradSlot{mm.s, mm.c, hh.s, hh.c, ll.s, ll.c};
  }
}

thread_local std::string g_em_err;
int em_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_em_err = buf;
  return code;
}
#define EMCHECK(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return em_fail(LGV_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                          \
  } while (0)

}  // namespace

struct lgv_em {
  lgv_em_params p;
  int nsteps = 0, K = 0, nseg = 0, window = 0, device = 0;
  int64_t V = 0;
  int nblocks = 0;
  uint64_t base = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  double *ckpt = nullptr;   // nseg states
  double *seg = nullptr;    // K states of the segment being reversed
  double *work[2] = {nullptr, nullptr};
  double *part = nullptr;   // per-block sums of two consecutive launches
  double *phibar = nullptr; // nsteps
  double *rows = nullptr;   // nsteps x kNobs: the measurement series
  GradSlot *gslots = nullptr;
  std::vector<double> phibar_host;
  bool have_forward = false;
};

namespace {

// Step n: in -> out, its NS volume sums into slots, and block 0 merges
// the slots of an earlier launch (prev).
void launch_step(int NS, const lgv_em_params &p, uint64_t base, int nblocks, const double *in,
                 double *out, int n, double *slots, const EmMerge &prev, hipStream_t st) {
  const double gdt = p.gamma * p.dt;
  const double amp = std::sqrt(2.0 * gdt);
  const uint64_t skey = stream_key(base, kEmTag | (uint64_t)n);
  if (NS == 0)
    k_em_step<0><<<nblocks, kEmBlock, 0, st>>>(in, out, p.N, p.m2, p.lambda, p.H, gdt, amp,
                                               skey, p.noise_kind, slots, prev);
  else if (NS == 1)
    k_em_step<1><<<nblocks, kEmBlock, 0, st>>>(in, out, p.N, p.m2, p.lambda, p.H, gdt, amp,
                                               skey, p.noise_kind, slots, prev);
  else
    k_em_step<kObsSums><<<nblocks, kEmBlock, 0, st>>>(in, out, p.N, p.m2, p.lambda, p.H, gdt,
                                                      amp, skey, p.noise_kind, slots, prev);
}

// Steps n0..n1-1 starting from `start`.  Forward mode (states == nullptr):
// records phibar of every step, ping-pongs through the work buffers and
// leaves phi_{n1} in `last`; with every > 0 it also writes the observables
// of the states every, 2 every, ... below nsteps into rows (row = state /
// every).  `pending` (may be nullptr) is a merge an earlier launch left for
// the first step.  Recompute mode: states[i] = phi_{n0+i} for i = 0..n1-n0
// (start is copied to states[0]), no sums; the work buffers, which hold
// the adjoint state during the backward pass, are not touched.
int em_run_segment(lgv_em *h, const double *start, int n0, int n1, double *states,
                   double *last, int every = 0, double *rows = nullptr,
                   const EmMerge *pending = nullptr) {
  const size_t st = 4 * h->V * sizeof(double);
  const double *in = start;
  if (states) {
    EMCHECK(hipMemcpyAsync(states, start, st, hipMemcpyDeviceToDevice, h->stream));
    in = states;
  }
  EmMerge prev{nullptr, 1, h->nblocks, nullptr, nullptr};
  if (pending) prev = *pending;
  for (int n = n0; n < n1; n++) {
    double *out = states ? states + (int64_t)(n - n0 + 1) * 4 * h->V
                         : ((n + 1 == n1) ? last : h->work[(n - n0) & 1]);
    double *slots = h->part + (size_t)(n & 1) * h->nblocks * kPartStride;
    EmMerge mine{slots, 1, h->nblocks, nullptr, nullptr};
    int NS = 0;
    if (!states) {  // the reverse pass uses the forward's phibar
      NS = 1;
      mine.phibar = h->phibar + n;
      if (every > 0 && (n + 1) % every == 0 && n + 1 < h->nsteps) {
        NS = mine.ns = kObsSums;
        mine.obs = rows + (size_t)((n + 1) / every) * kNobs;
      }
    }
    launch_step(NS, h->p, h->base, h->nblocks, in, out, n, slots, prev, h->stream);
    prev = NS ? mine : EmMerge{nullptr, 1, h->nblocks, nullptr, nullptr};
    in = out;
  }
  if (prev.slots) k_em_merge<<<1, kEmBlock, 0, h->stream>>>(prev, h->V);
  if (!states && n1 == n0 && last != start)
    EMCHECK(hipMemcpyAsync(last, start, st, hipMemcpyDeviceToDevice, h->stream));
  EMCHECK(hipGetLastError());
  return LGV_OK;
}

// std::seed_seq{v}.generate() of one word ([rand.util.seedseq], libstdc++ 11
// bits/random.tcc:3250-3330) without the heap: with one input and one
// output word every index of the algorithm is 0, the first loop leaves
// b = 1664525 T(b0) + 1 + ... , and the result is
// 1566083941 T(3 (1664525 T(r) + v)), T(x) = x ^ (x >> 27), r = 1664525
// T(0x8b8b8b8b) + 1, all in 32-bit arithmetic.
uint32_t seed_word(uint32_t v) {
  auto T = [](uint32_t x) { return x ^ (x >> 27); };
  const uint32_t r0 = 1664525u * T(0x8b8b8b8bu) + 1u;
  const uint32_t b1 = 1664525u * T(r0) + v;
  return 1566083941u * T(3u * b1);
}

// ---- the entry point that allocates nothing (lgv_em_step_noalloc)
//
// Its reduction slots live in one file-scope device array per GPU, sized
// for em_blocks(N) <= kScratchBlocks (every N <= 256).  Host code
// never passes `scratch` itself to a kernel (on the host it names a
// stand-in, not device memory): the device address is looked up once per
// device with hipGetSymbolAddress and kept in g_scratch.
constexpr int kScratchBlocks = 8192;  // em_blocks(256) = 4 x 64 x 32
constexpr size_t kScratchDoubles = (size_t)kScratchBlocks * kPartStride;  // 640 KiB
}  // namespace
static __device__ __align__(256) double scratch[kScratchDoubles];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices];

// The device address of `scratch` on `device` (current), looked up on the
// first call for that device.  Nothing in it needs clearing: every launch
// writes the slots it merges.
int scratch_on(int device, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    EMCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double *>(ptr);
  }
  *out = g_scratch[device];
  return LGV_OK;
}

}  // namespace

extern "C" {

const char *lgv_em_last_error(void) { return g_em_err.c_str(); }

int lgv_em_create(const lgv_em_params *p, int nsteps, int ckpt_every, int device,
                  lgv_em **out) {
  if (!p || !out) return em_fail(LGV_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->N < 4 || p->N > 1024) return em_fail(LGV_ERR_ARG, "N=%d outside [4,1024]", p->N);
  if (nsteps <= 0 || ckpt_every <= 0) return em_fail(LGV_ERR_ARG, "nsteps and ckpt_every must be positive");
  if (p->noise_kind != 0 && p->noise_kind != 1) return em_fail(LGV_ERR_ARG, "noise_kind must be 0 or 1");
  lgv_em *h = new lgv_em;
  h->p = *p;
  h->nsteps = nsteps;
  h->K = ckpt_every;
  h->nseg = (nsteps + ckpt_every - 1) / ckpt_every;
  h->device = device;
  h->V = (int64_t)p->N * p->N * p->N;
  h->nblocks = em_blocks(p->N);
  std::seed_seq seq = {p->seed};
  uint32_t s32 = 0;
  seq.generate(&s32, &s32 + 1);
  h->base = base_key(s32);
  const size_t st = 4 * h->V * sizeof(double);
  auto cleanup_fail = [&](hipError_t e) {
    lgv_em_destroy(h);
    return em_fail(LGV_ERR_HIP, "lgv_em_create: %s", hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return cleanup_fail(e);
  if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) return cleanup_fail(e);
  h->own_stream = true;
  if ((e = hipMalloc(&h->ckpt, st * h->nseg)) != hipSuccess) return cleanup_fail(e);
  if ((e = hipMalloc(&h->seg, st * h->K)) != hipSuccess) return cleanup_fail(e);
  for (int i = 0; i < 2; i++)
    if ((e = hipMalloc(&h->work[i], st)) != hipSuccess) return cleanup_fail(e);
  if ((e = hipMalloc(&h->part, 2 * (size_t)h->nblocks * kPartStride * sizeof(double))) !=
      hipSuccess)
    return cleanup_fail(e);
  if ((e = hipMalloc(&h->phibar, nsteps * sizeof(double))) != hipSuccess) return cleanup_fail(e);
  if ((e = hipMalloc(&h->rows, (size_t)nsteps * kNobs * sizeof(double))) != hipSuccess)
    return cleanup_fail(e);
  if ((e = hipMalloc(&h->gslots, h->nblocks * sizeof(GradSlot))) != hipSuccess) return cleanup_fail(e);
  *out = h;
  return LGV_OK;
}

int lgv_em_set_stream(lgv_em *h, void *stream) {
  if (!h) return em_fail(LGV_ERR_ARG, "null handle");
  if (h->own_stream && h->stream) {
    EMCHECK(hipStreamSynchronize(h->stream));
    EMCHECK(hipStreamDestroy(h->stream));
    h->stream = nullptr;
  }
  h->own_stream = stream == nullptr;
  if (stream) {
    h->stream = (hipStream_t)stream;
  } else {
    EMCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  }
  return LGV_OK;
}

int lgv_em_forward(lgv_em *h, const double *phi0, int window, double *phi_final,
                   double *phibar_host, double *loss) {
  if (!h || !phi0 || !loss) return em_fail(LGV_ERR_ARG, "null argument");
  if (window <= 0 || window > h->nsteps) return em_fail(LGV_ERR_ARG, "window outside [1,nsteps]");
  EMCHECK(hipSetDevice(h->device));
  h->window = window;
  const size_t st = 4 * h->V * sizeof(double);
  EMCHECK(hipMemcpyAsync(h->ckpt, phi0, st, hipMemcpyDeviceToDevice, h->stream));
  int rc;
  for (int j = 0; j < h->nseg; j++) {
    const int n0 = j * h->K, n1 = std::min(h->nsteps, n0 + h->K);
    double *last = (j + 1 < h->nseg) ? h->ckpt + (int64_t)(j + 1) * 4 * h->V
                                     : (phi_final ? phi_final : h->seg);
    if ((rc = em_run_segment(h, h->ckpt + (int64_t)j * 4 * h->V, n0, n1, nullptr, last)))
      return rc;
  }
  h->phibar_host.resize(h->nsteps);
  EMCHECK(hipMemcpyAsync(h->phibar_host.data(), h->phibar, h->nsteps * sizeof(double),
                         hipMemcpyDeviceToHost, h->stream));
  EMCHECK(hipStreamSynchronize(h->stream));
  long double m = 0, m2 = 0;
  for (int n = h->nsteps - window; n < h->nsteps; n++) {
    m += h->phibar_host[n];
    m2 += (long double)h->phibar_host[n] * h->phibar_host[n];
  }
  m /= window;
  m2 /= window;
  *loss = (double)(m + (long double)h->V * (m2 - m * m));
  if (phibar_host) memcpy(phibar_host, h->phibar_host.data(), h->nsteps * sizeof(double));
  h->have_forward = true;
  return LGV_OK;
}

int lgv_em_backward(lgv_em *h, double *grad_phi0, double grad_params[3]) {
  if (!h || !grad_phi0 || !grad_params) return em_fail(LGV_ERR_ARG, "null argument");
  if (!h->have_forward) return em_fail(LGV_ERR_STATE, "lgv_em_backward before lgv_em_forward");
  EMCHECK(hipSetDevice(h->device));
  const size_t st = 4 * h->V * sizeof(double);
  const int W = h->window;
  long double mean = 0;
  for (int n = h->nsteps - W; n < h->nsteps; n++) mean += h->phibar_host[n];
  mean /= W;
  const double gdt = h->p.gamma * h->p.dt;
  double *lam = h->work[0], *nxt = h->work[1];
  static_assert(sizeof(GradSlot) == 6 * sizeof(double), "GradSlot is six doubles");
  k_em_zero<<<1024, 256, 0, h->stream>>>(lam, 4 * h->V);
  k_em_zero<<<64, 256, 0, h->stream>>>(reinterpret_cast<double *>(h->gslots),
                                       (int64_t)6 * h->nblocks);
  EMCHECK(hipGetLastError());
  int rc;
  for (int j = h->nseg - 1; j >= 0; j--) {
    const int n0 = j * h->K, n1 = std::min(h->nsteps, n0 + h->K);
    // recompute phi_{n0} .. phi_{n1-1} into seg (the forward noise is
    // replayed from the same counters)
    if ((rc = em_run_segment(h, h->ckpt + (int64_t)j * 4 * h->V, n0, n1 - 1, h->seg,
                             nullptr)))
      return rc;
    for (int n = n1 - 1; n >= n0; n--) {
      double dl = 0.;
      if (n >= h->nsteps - W)
        dl = (1.0 / W + (double)h->V * 2.0 * (h->phibar_host[n] - (double)mean) / W) /
             (double)h->V;
      k_em_adjoint<<<h->nblocks, kEmBlock, 0, h->stream>>>(
          h->seg + (int64_t)(n - n0) * 4 * h->V, lam, nxt, h->p.N, h->p.m2, h->p.lambda, gdt,
          dl, h->gslots);
      std::swap(lam, nxt);
    }
  }
  EMCHECK(hipGetLastError());
  EMCHECK(hipMemcpyAsync(grad_phi0, lam, st, hipMemcpyDeviceToDevice, h->stream));
  std::vector<GradSlot> gs(h->nblocks);
  EMCHECK(hipMemcpyAsync(gs.data(), h->gslots, h->nblocks * sizeof(GradSlot),
                         hipMemcpyDeviceToHost, h->stream));
  EMCHECK(hipStreamSynchronize(h->stream));
  KSum m{0, 0}, hh{0, 0}, l{0, 0};
  for (const GradSlot &g : gs) {
    ksum_merge(m, KSum{g.m2_s, g.m2_c});
    ksum_merge(hh, KSum{g.h_s, g.h_c});
    ksum_merge(l, KSum{g.l_s, g.l_c});
  }
  grad_params[0] = ksum_value(m);
  grad_params[1] = ksum_value(hh);
  grad_params[2] = ksum_value(l);
  // work[0/1] were used as lam buffers; restore the pointer roles
  h->work[0] = lam;
  h->work[1] = nxt;
  return LGV_OK;
}

int lgv_em_series(lgv_em *h, const double *phi0, int every, double *phi_final,
                  double *rows_host) {
  if (!h || !phi0 || !rows_host) return em_fail(LGV_ERR_ARG, "null argument");
  if (every <= 0) return em_fail(LGV_ERR_ARG, "every must be positive");
  EMCHECK(hipSetDevice(h->device));
  const int nrows = (h->nsteps + every - 1) / every;
  // the measurement before the first step, then one after every `every`
  // steps, inside the step launch that produces the state
  // (slots of parity 1, merged by step 0's launch)
  double *slots = h->part + (size_t)h->nblocks * kPartStride;
  k_em_observe<<<h->nblocks, kEmBlock, 0, h->stream>>>(phi0, h->p.N, slots);
  EMCHECK(hipGetLastError());
  const EmMerge row0{slots, kObsSums, h->nblocks, nullptr, h->rows};
  int rc;
  if ((rc = em_run_segment(h, phi0, 0, h->nsteps, nullptr, phi_final ? phi_final : h->seg,
                           every, h->rows, &row0)))
    return rc;
  EMCHECK(hipMemcpyAsync(rows_host, h->rows, (size_t)nrows * kNobs * sizeof(double),
                         hipMemcpyDeviceToHost, h->stream));
  EMCHECK(hipStreamSynchronize(h->stream));
  return LGV_OK;
}

int lgv_em_step(lgv_em *h, const double *in, double *out, int n, double *obs) {
  if (!h || !in || !out) return em_fail(LGV_ERR_ARG, "null argument");
  if (n < 0) return em_fail(LGV_ERR_ARG, "negative step index");
  if (in == out) return em_fail(LGV_ERR_ARG, "in and out must differ");
  EMCHECK(hipSetDevice(h->device));
  const EmMerge none{nullptr, 1, h->nblocks, nullptr, nullptr};
  launch_step(obs ? kObsSums : 0, h->p, h->base, h->nblocks, in, out, n, h->part, none,
              h->stream);
  if (obs)
    k_em_merge<<<1, kEmBlock, 0, h->stream>>>(EmMerge{h->part, kObsSums, h->nblocks, nullptr, obs},
                                              h->V);
  EMCHECK(hipGetLastError());
  return LGV_OK;
}

int lgv_em_step_noalloc(const lgv_em_params *p, const double *in, double *out, int n,
                        double *obs, int device, void *stream) {
  if (!p || !in || !out) return em_fail(LGV_ERR_ARG, "null argument");
  if (in == out) return em_fail(LGV_ERR_ARG, "in and out must differ");
  if (p->N < 4 || p->N > 256 || em_blocks(p->N) > kScratchBlocks)
    return em_fail(LGV_ERR_RANGE, "lgv_em_step_noalloc: N=%d outside [4,256]", p->N);
  if (p->noise_kind != 0 && p->noise_kind != 1)
    return em_fail(LGV_ERR_ARG, "noise_kind must be 0 or 1");
  if (n < 0) return em_fail(LGV_ERR_ARG, "negative step index");
  if (device < 0 || device >= kMaxDevices)
    return em_fail(LGV_ERR_ARG, "device %d outside [0,%d)", device, kMaxDevices);
  EMCHECK(hipSetDevice(device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  double *d = nullptr;
  int rc;
  if ((rc = scratch_on(device, &d))) return rc;
  const int nblocks = em_blocks(p->N);
  const EmMerge none{nullptr, 1, nblocks, nullptr, nullptr};
  launch_step(obs ? kObsSums : 0, *p, base_key(seed_word((uint32_t)p->seed)), nblocks, in, out,
              n, d, none, st);
  if (obs) {
    const int64_t V = (int64_t)p->N * p->N * p->N;
    k_em_merge<<<1, kEmBlock, 0, st>>>(EmMerge{d, kObsSums, nblocks, nullptr, obs}, V);
  }
  EMCHECK(hipGetLastError());
  return LGV_OK;
}

int lgv_em_destroy(lgv_em *h) {
  if (!h) return LGV_OK;
  // releases everything even after a failure and reports the first one
  hipError_t first = hipSuccess;
  auto note = [&](hipError_t e) {
    if (first == hipSuccess && e != hipSuccess) first = e;
  };
  note(hipSetDevice(h->device));
  if (h->stream) note(hipStreamSynchronize(h->stream));
  for (void *q : {(void *)h->ckpt, (void *)h->seg, (void *)h->work[0], (void *)h->work[1],
                  (void *)h->part, (void *)h->phibar, (void *)h->rows, (void *)h->gslots})
    note(hipFree(q));
  if (h->own_stream && h->stream) note(hipStreamDestroy(h->stream));
  delete h;
  if (first != hipSuccess)
    return em_fail(LGV_ERR_HIP, "lgv_em_destroy: %s", hipGetErrorString(first));
  return LGV_OK;
}

}  // extern "C"
