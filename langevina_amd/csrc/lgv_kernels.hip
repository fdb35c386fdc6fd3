// lgv_kernels.hip -- HIP kernels of the ModelG time step for gfx950.
//
// Blocks are mapped to replicas through a launch plan, so one launch steps
// many independent lattices.  Fields live in the red-black
// structure-of-arrays layout of lgv_device.h, double-buffered: the
// Metropolis sweeps update the live buffer in place (a checkerboard never
// reads what it writes), the fused ideal substep writes the other buffer.
// All FP arithmetic is FP64 (FP32 storage is widened on load and rounded
// on store).  The Metropolis sweeps follow the reference's expression order
// bit for bit; the library is built with -ffp-contract=off.
#include <float.h>

#include "lgv_device.h"
#include "lgv_internal.h"

namespace lgv {

// `cur` packs the live buffer of the phi planes (bit 0) and of the charge
// planes (bit 1).  A select instead of p.fields[b]: a runtime index into the
// register copy of ReplicaParams would push the whole struct to scratch.
template <typename T>
__device__ __forceinline__ T *plane(const ReplicaParams &p, int cur, int c, int par) {
  const int b = c < 4 ? (cur & 1) : (cur >> 1);
  T *base = reinterpret_cast<T *>(b ? p.fields[1] : p.fields[0]);
  return base + ((int64_t)(c * 2 + par)) * p.half;
}
__device__ __forceinline__ int live(const ReplicaState &s) {
  return s.cur_phi | (s.cur_chg << 1);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Renumber blocks so that XCD k runs the k-th contiguous eighth of the
// plan: blocks that share neighbour planes then share an L2.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, n = gridDim.x, q = n >> 3;
  return b < (q << 3) ? (b & 7) * q + (b >> 3) : b;
}

__device__ __forceinline__ int wrapp(int a, int N) { return a == N ? 0 : a; }
__device__ __forceinline__ int wrapm(int a, int N) { return a < 0 ? N - 1 : a; }

// Deterministic block reduction of the monitor increments; thread 0 adds the
// block total into the block's own slot (stream order makes this race-free).
__device__ void block_monitor_add(MonSlot *slot, long long down, long long upyes,
                                  long long upno, double dSd, double dSu) {
  __shared__ long long sc[3][kBlock / 64];
  __shared__ double sd[2][kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    down += __shfl_xor(down, o, 64);
    upyes += __shfl_xor(upyes, o, 64);
    upno += __shfl_xor(upno, o, 64);
    dSd += __shfl_xor(dSd, o, 64);
    dSu += __shfl_xor(dSu, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sc[0][w] = down; sc[1][w] = upyes; sc[2][w] = upno;
    sd[0][w] = dSd; sd[1][w] = dSu;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long a = 0, b = 0, c = 0;
    double d = 0, e = 0;
    for (int i = 0; i < kBlock / 64; i++) {
      a += sc[0][i]; b += sc[1][i]; c += sc[2][i];
      d += sd[0][i]; e += sd[1][i];
    }
    slot->down_counts += a;
    slot->up_counts += b + c;
    slot->up_yes += b;
    slot->up_no += c;
    slot->down_dS += d;
    slot->up_dS += e;
  }
}

// ------------------------------------------------------------------------
// One parity of EulerLangevinHB::step (ModelG/Stepper.cxx:309-386).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_hb_phi(const ReplicaParams *__restrict__ P,
                                                   const ReplicaState *__restrict__ S,
                                                   const int *__restrict__ blk_rep,
                                                   const int *__restrict__ blk_first, int ieo,
                                                   int ctr_off, MonSlot *__restrict__ mon) {
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  const ReplicaParams p = P[r];
  const int64_t idx = (int64_t)(bid - blk_first[r]) * kBlock + threadIdx.x;
  long long down = 0, upyes = 0, upno = 0;
  double dSd = 0., dSu = 0.;
  if (idx < p.half) {
    const int N = p.N, NH = p.NH;
    const int xh = (int)(idx % NH);
    const int64_t rowi = idx / NH;
    const int y = (int)(rowi % N), z = (int)(rowi / N);
    const int x = site_x(xh, y, z, ieo);
    const int q = ieo ^ 1;
    const int64_t row = rowi * NH;
    const int64_t nb[6] = {row + (wrapp(x + 1, N) >> 1), row + (wrapm(x - 1, N) >> 1),
                           ((int64_t)z * N + wrapp(y + 1, N)) * NH + xh,
                           ((int64_t)z * N + wrapm(y - 1, N)) * NH + xh,
                           ((int64_t)wrapp(z + 1, N) * N + y) * NH + xh,
                           ((int64_t)wrapm(z - 1, N) * N + y) * NH + xh};
    const ReplicaState s = S[r];
    const double mass = p.mass0 + p.dmassdt * s.t;
    const double Lambda2 = 0.5 * (2. * 3. + mass);
    const double lambda = p.lambda;
    const double H = p.H;
    const double rdtg = p.rdtg;
    Xoro g = xoro_seeded(site_key(stream_key(p.base_key, s.sweep_ctr + ctr_off),
                                  ((uint64_t)z * N + y) * N + x));
    double hphi_o = 0., hphi_n = 0., s_o = 0., s_n = 0.;
    double phi_o[4], phi_n[4];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      const T *fq = plane<T>(p, live(s), l, q);
      const double heff = ((double)fq[nb[0]] + (double)fq[nb[1]]) +
                          ((double)fq[nb[2]] + (double)fq[nb[3]]) +
                          ((double)fq[nb[4]] + (double)fq[nb[5]]);
      phi_o[l] = (double)plane<T>(p, live(s), l, ieo)[idx];
      phi_n[l] = phi_o[l] + rdtg * variance1(g);
      s_o += phi_o[l] * phi_o[l];
      s_n += phi_n[l] * phi_n[l];
      hphi_o += heff * phi_o[l];
      hphi_n += heff * phi_n[l];
    }
    hphi_o += (H * phi_o[0]);
    hphi_n += (H * phi_n[0]);
    const double dS = -(hphi_n - hphi_o) + Lambda2 * (s_n - s_o) +
                      lambda / 4. * (s_n * s_n - s_o * s_o);
    bool accept;
    if (dS < 0) {
      accept = true;
      down = 1;
      dSd = dS;
    } else {
      const double rr = uniform01(g);
      accept = rr < exp(-dS);
      if (accept) { upyes = 1; dSu = dS; } else { upno = 1; }
    }
    if (accept) {
#pragma unroll
      for (int l = 0; l < 4; l++) plane<T>(p, live(s), l, ieo)[idx] = (T)phi_n[l];
    }
  }
  block_monitor_add(mon + bid, down, upyes, upno, dSd, dSu);
}

// ------------------------------------------------------------------------
// The shuffled sub-sweep order of ModelGChargeHB::step (Stepper.cxx:461-468)
// from the replica's per-step shuffle stream.
__global__ void k_shuffle(ReplicaState *S, const ReplicaParams *P, int nrep, int ctr_off) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  Xoro g = xoro_seeded(site_key(stream_key(P[r].base_key, S[r].sweep_ctr + ctr_off), 0));
  int order[5];
  shuffle_orders(g, order);
  for (int i = 0; i < 5; i++) S[r].order[i] = order[i];
}

// ------------------------------------------------------------------------
// The two sub-sweeps of one direction of ModelGChargeHB::step (the shuffled
// order always runs both parities of a direction back to back,
// Stepper.cxx:471-473) in one pass.  Thread = site A' of the second parity
// e2; along the direction its line reads L = A'-e, A', B' = A'+e, R = A'+2e.
// First sub-sweep (parity e1): pairs (L, A') and (B', R), keyed by L and B';
// second: pair (A', B') keyed by A' on the first sub-sweep's results.  Each
// first-parity pair is computed by the two threads that need it (identical
// counter streams, identical bits); the monitor counts it once, from the
// thread on its right.  Reads the live charge buffer, writes the other one.
__device__ __forceinline__ void charge_pair(double nA, double nB, double rms, double chi,
                                            Xoro &g, double &outA, double &outB,
                                            long long &down, long long &upyes,
                                            long long &upno, double &dSd, double &dSu,
                                            bool count) {
  double q = rms * variance1(g);
  const double dS = ((nA - q) * (nA - q) + (nB + q) * (nB + q) - nA * nA - nB * nB) /
                    (2. * chi);
  if (dS < 0) {
    if (count) { down++; dSd += dS; }
  } else {
    const double rr = uniform01(g);
    if (rr < exp(-dS)) {
      if (count) { upyes++; dSu += dS; }
    } else {
      if (count) upno++;
      q = 0.;
    }
  }
  outA = nA + (-q);
  outB = nB + q;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_charge_dir(const ReplicaParams *__restrict__ P,
                                                       const ReplicaState *__restrict__ S,
                                                       const int *__restrict__ blk_rep,
                                                       const int *__restrict__ blk_first,
                                                       int ixyz, int ctr_off,
                                                       MonSlot *__restrict__ mon) {
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  const ReplicaParams p = P[r];
  const int64_t idx = (int64_t)(bid - blk_first[r]) * kBlock + threadIdx.x;
  long long down = 0, upyes = 0, upno = 0;
  double dSd = 0., dSu = 0.;
  if (idx < p.half) {
    const ReplicaState s = S[r];
    const int cur = live(s), nxt = cur ^ 2;
    const int dir = S[r].order[ixyz];
    const int e1 = S[r].order[3], e2 = S[r].order[4];
    const int N = p.N, NH = p.NH;
    const int xh = (int)(idx % NH);
    const int64_t rowi = idx / NH;
    const int y = (int)(rowi % N), z = (int)(rowi / N);
    const int x = site_x(xh, y, z, e2);
    // the four sites along the direction: L, A', B', R (parities e1 e2 e1 e2)
    const int a0 = dir == 0 ? x : (dir == 1 ? y : z);
    const int am = a0 == 0 ? N - 1 : a0 - 1;
    const int ap = a0 + 1 == N ? 0 : a0 + 1;
    const int app = ap + 1 == N ? 0 : ap + 1;
    int64_t ii[4];
    uint64_t lin[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int a = k == 0 ? am : (k == 1 ? a0 : (k == 2 ? ap : app));
      const int xx = dir == 0 ? a : x, yy = dir == 1 ? a : y, zz = dir == 2 ? a : z;
      ii[k] = ((int64_t)zz * N + yy) * NH + (xx >> 1);
      lin[k] = ((uint64_t)zz * N + yy) * N + xx;
    }
    const double chi = p.chi, rms = p.rms;
    const uint64_t k1 = stream_key(p.base_key, s.sweep_ctr + ctr_off);
    const uint64_t k2 = stream_key(p.base_key, s.sweep_ctr + ctr_off + 1);
    Xoro gL = xoro_seeded(site_key(k1, lin[0]));
    Xoro gB = xoro_seeded(site_key(k1, lin[2]));
    Xoro gA = xoro_seeded(site_key(k2, lin[1]));
    long long d0 = 0, y0 = 0, n0 = 0;
    double sd0 = 0., su0 = 0.;
#pragma unroll
    for (int L = 4; L < 10; L++) {
      const double vL = (double)plane<T>(p, cur, L, e1)[ii[0]];
      const double vA = (double)plane<T>(p, cur, L, e2)[ii[1]];
      const double vB = (double)plane<T>(p, cur, L, e1)[ii[2]];
      const double vR = (double)plane<T>(p, cur, L, e2)[ii[3]];
      double tL, a1, b1, tR;
      charge_pair(vL, vA, rms, chi, gL, tL, a1, down, upyes, upno, dSd, dSu, true);
      charge_pair(vB, vR, rms, chi, gB, b1, tR, d0, y0, n0, sd0, su0, false);
      // the first sub-sweep's results as stored (rounded in FP32 storage)
      a1 = (double)(T)a1;
      b1 = (double)(T)b1;
      double a2, b2;
      charge_pair(a1, b1, rms, chi, gA, a2, b2, down, upyes, upno, dSd, dSu, true);
      plane<T>(p, nxt, L, e2)[ii[1]] = (T)a2;
      plane<T>(p, nxt, L, e1)[ii[2]] = (T)b2;
    }
  }
  block_monitor_add(mon + bid, down, upyes, upno, dSd, dSu);
}

__global__ void k_flip_chg(ReplicaState *S, int nrep) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrep) S[r].cur_chg ^= 1;
}

// ------------------------------------------------------------------------
// IdealPV2::step_no_reject (Stepper.cxx:68-176) fused into one pass:
// rotate by dt/2, kick the charges with the rotated neighbours, rotate by
// dt/2 with the kicked charges.  A block owns a TXH x TY x-y tile (both
// parities of each half-column) and marches over TZ planes; the rotated phi
// of the current plane and a one-site x-y halo (recomputed from the raw
// fields) sit in LDS, the rotated phi of the planes below and above in
// registers.  Reads the live buffer, writes the other one.
template <typename T>
__device__ __forceinline__ void load_site(const ReplicaParams &p, int cur, int x, int y, int z,
                                          double f[10]) {
  const int par = (x + y + z) & 1;
  const int64_t i = ((int64_t)z * p.N + y) * p.NH + (x >> 1);
#pragma unroll
  for (int c = 0; c < 10; c++) f[c] = (double)plane<T>(p, cur, c, par)[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ideal_fused(const ReplicaParams *__restrict__ P,
                                                        const ReplicaState *__restrict__ S,
                                                        const int *__restrict__ blk_rep,
                                                        const int *__restrict__ blk_first) {
  __shared__ double lds[4][kIdealLds];
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  // a reference, not a copy: a private copy would be staged in scratch memory
  const ReplicaParams &p = P[r];
  const int cur = live(S[r]), nxt = cur ^ 3;
  const int N = p.N, NH = p.NH;
  int TXH, TY, TZ;
  ideal_tile(N, p.tile_cap, p.tile_tz, TXH, TY, TZ);
  const int ntx = (NH + TXH - 1) / TXH, nty = (N + TY - 1) / TY;
  int b = bid - blk_first[r];
  const int bx = b % ntx;
  b /= ntx;
  const int by = b % nty, bz = b / nty;
  const int xh0 = bx * TXH, y0 = by * TY, z0 = bz * TZ;
  const int tw = min(TXH, NH - xh0), th = min(TY, N - y0), tz = min(TZ, N - z0);
  const int W = 2 * tw + 2;
  const int t = threadIdx.x;
  const int tx = t % TXH, ty = t / TXH;
  const bool own = ty < th && tx < tw;
  const int xh = xh0 + tx, y = y0 + ty;
  // halo site of this thread (one-site ring around the tile, no corners)
  const int nh = 2 * th + 4 * tw;
  const bool halo = t < nh;
  int hlx = 0, hly = 0;
  if (t < th) { hlx = -1; hly = t; }
  else if (t < 2 * th) { hlx = 2 * tw; hly = t - th; }
  else if (t < 2 * th + 2 * tw) { hlx = t - 2 * th; hly = -1; }
  else { hlx = t - 2 * th - 2 * tw; hly = th; }
  const int hgx = (2 * xh0 + hlx + N) % N, hgy = (y0 + hly + N) % N;
  const double h = 0.5 * (p.tau_ideal / p.chi);
  const double axx = p.axx, H0 = p.H, dt = p.dt_ideal;

  double prv[2][4], cen[2][4], nex[2][4];  // rotated phi, own sites, z-1, z, z+1
  double chg[2][6], chn[2][6];             // raw charges of planes z and z+1
  auto rotate_own = [&](int z, double out[2][4], double ch[2][6]) __attribute__((always_inline)) {
    if (!own) return;
    const int64_t i = ((int64_t)z * N + y) * NH + xh;
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int par = (s + y + z) & 1;
      double f[10];
#pragma unroll
      for (int c = 0; c < 10; c++) f[c] = (double)plane<T>(p, cur, c, par)[i];
      rotate_site_fast(f, f + 4, f + 7, h);
#pragma unroll
      for (int c = 0; c < 4; c++) out[s][c] = f[c];
#pragma unroll
      for (int c = 0; c < 6; c++) ch[s][c] = f[4 + c];
    }
  };
  auto fill_lds = [&](int z, const double o[2][4]) __attribute__((always_inline)) {
    if (own) {
#pragma unroll
      for (int s = 0; s < 2; s++)
#pragma unroll
        for (int c = 0; c < 4; c++) lds[c][(ty + 1) * W + (2 * tx + s + 1)] = o[s][c];
    }
    if (halo) {
      double f[10];
      load_site<T>(p, cur, hgx, hgy, z, f);
      rotate_site_fast(f, f + 4, f + 7, h);
#pragma unroll
      for (int c = 0; c < 4; c++) lds[c][(hly + 1) * W + (hlx + 1)] = f[c];
    }
  };

  double dummy[2][6];
  rotate_own((z0 - 1 + N) % N, prv, dummy);
  rotate_own(z0, cen, chg);
  fill_lds(z0, cen);
  for (int k = 0; k < tz; k++) {
    const int z = z0 + k;
    const int zn = (z + 1) % N;
    rotate_own(zn, nex, chn);
    __syncthreads();
    if (own) {
      const int64_t i = ((int64_t)z * N + y) * NH + xh;
#pragma unroll
      for (int s = 0; s < 2; s++) {
        const int lx = 2 * tx + s + 1, ly = ty + 1;
        double cp[4], xp[4], xm[4], yp[4], ym[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          cp[c] = cen[s][c];
          xp[c] = lds[c][ly * W + lx + 1];
          xm[c] = lds[c][ly * W + lx - 1];
          yp[c] = lds[c][(ly + 1) * W + lx];
          ym[c] = lds[c][(ly - 1) * W + lx];
        }
        const double *zp = nex[s], *zm = prv[s];
        double A[3], V[3];
#pragma unroll
        for (int l = 0; l < 3; l++) {
          A[l] = chg[s][l];
          V[l] = chg[s][3 + l];
        }
#pragma unroll
        for (int l = 0; l < 3; l++) {
          const double advxx = (-xp[0] * cp[l + 1] + cp[0] * xp[l + 1] + cp[0] * xm[l + 1] -
                                xm[0] * cp[l + 1]) * axx;
          const double advyy = (-yp[0] * cp[l + 1] + cp[0] * yp[l + 1] + cp[0] * ym[l + 1] -
                                ym[0] * cp[l + 1]) * axx;
          const double advzz = (-zp[0] * cp[l + 1] + cp[0] * zp[l + 1] + cp[0] * zm[l + 1] -
                                zm[0] * cp[l + 1]) * axx;
          A[l] += dt * (advxx + advyy + advzz - H0 * cp[l + 1]);
        }
#pragma unroll
        for (int s3 = 0; s3 < 3; s3++) {
          const int s1 = (s3 + 1) % 3, s2 = (s3 + 2) % 3;
          const double advxx = (xp[1 + s2] * cp[1 + s1] - cp[1 + s2] * xm[1 + s1] -
                                xp[1 + s1] * cp[1 + s2] + cp[1 + s1] * xm[1 + s2]) * axx;
          const double advyy = (yp[1 + s2] * cp[1 + s1] - cp[1 + s2] * ym[1 + s1] -
                                yp[1 + s1] * cp[1 + s2] + cp[1 + s1] * ym[1 + s2]) * axx;
          const double advzz = (zp[1 + s2] * cp[1 + s1] - cp[1 + s2] * zm[1 + s1] -
                                zp[1 + s1] * cp[1 + s2] + cp[1 + s1] * zm[1 + s2]) * axx;
          V[s3] += dt * (advxx + advyy + advzz);
        }
        rotate_site_fast(cp, A, V, h);
        const int par = (s + y + z) & 1;
#pragma unroll
        for (int c = 0; c < 4; c++) plane<T>(p, nxt, c, par)[i] = (T)cp[c];
#pragma unroll
        for (int c = 0; c < 3; c++) plane<T>(p, nxt, 4 + c, par)[i] = (T)A[c];
#pragma unroll
        for (int c = 0; c < 3; c++) plane<T>(p, nxt, 7 + c, par)[i] = (T)V[c];
      }
    }
    __syncthreads();
    if (k + 1 < tz) fill_lds(zn, nex);
#pragma unroll
    for (int s = 0; s < 2; s++) {
#pragma unroll
      for (int c = 0; c < 4; c++) {
        prv[s][c] = cen[s][c];
        cen[s][c] = nex[s][c];
      }
#pragma unroll
      for (int c = 0; c < 6; c++) chg[s][c] = chn[s][c];
    }
  }
}

__global__ void k_flip(ReplicaState *S, int nrep) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrep) {
    S[r].cur_phi ^= 1;
    S[r].cur_chg ^= 1;
  }
}

// ------------------------------------------------------------------------
// IdealPV2::computeEnergy (Stepper.cxx:204-268): per-block compensated
// partial sums of grad2, hEn and nab2.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy(const ReplicaParams *__restrict__ P,
                                                   const ReplicaState *__restrict__ S,
                                                   const int *__restrict__ blk_rep,
                                                   const int *__restrict__ blk_first,
                                                   EnergySlot *__restrict__ slots) {
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  const ReplicaParams p = P[r];
  const int cur = live(S[r]);
  int64_t idx = (int64_t)(bid - blk_first[r]) * kBlock + threadIdx.x;
  double g2 = 0., hE = 0., n2 = 0.;
  if (idx < 2 * p.half) {
    const int par = idx >= p.half;
    idx -= par * p.half;
    const int N = p.N, NH = p.NH;
    const int xh = (int)(idx % NH);
    const int64_t rowi = idx / NH;
    const int y = (int)(rowi % N), z = (int)(rowi / N);
    const int x = site_x(xh, y, z, par);
    const int q = par ^ 1;
    const int64_t ixp = rowi * NH + (wrapp(x + 1, N) >> 1);
    const int64_t iyp = ((int64_t)z * N + wrapp(y + 1, N)) * NH + xh;
    const int64_t izp = ((int64_t)wrapp(z + 1, N) * N + y) * NH + xh;
    for (int s = 0; s < 4; s++) {
      const double phimid = (double)plane<T>(p, cur, s, par)[idx];
      const T *o = plane<T>(p, cur, s, q);
      double d;
      d = (double)o[izp] - phimid; g2 += d * d;
      d = (double)o[iyp] - phimid; g2 += d * d;
      d = (double)o[ixp] - phimid; g2 += d * d;
      if (s == 0) hE = phimid * p.H;
    }
    for (int s = 0; s < 3; s++) { const double v = (double)plane<T>(p, cur, 7 + s, par)[idx]; n2 += v * v; }
    for (int s = 0; s < 3; s++) { const double a = (double)plane<T>(p, cur, 4 + s, par)[idx]; n2 += a * a; }
  }
  KSum kg{g2, 0.}, kh{hE, 0.}, kn{n2, 0.};
  for (int o = 32; o > 0; o >>= 1) {
    KSum tt;
    tt.s = __shfl_xor(kg.s, o, 64); tt.c = __shfl_xor(kg.c, o, 64); ksum_merge(kg, tt);
    tt.s = __shfl_xor(kh.s, o, 64); tt.c = __shfl_xor(kh.c, o, 64); ksum_merge(kh, tt);
    tt.s = __shfl_xor(kn.s, o, 64); tt.c = __shfl_xor(kn.c, o, 64); ksum_merge(kn, tt);
  }
  __shared__ KSum sh[3][kBlock / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[0][w] = kg; sh[1][w] = kh; sh[2][w] = kn; }
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum a = sh[0][0], b = sh[1][0], c = sh[2][0];
    for (int i = 1; i < kBlock / 64; i++) {
      ksum_merge(a, sh[0][i]); ksum_merge(b, sh[1][i]); ksum_merge(c, sh[2][i]);
    }
    slots[bid] = EnergySlot{a.s, a.c, b.s, b.c, c.s, c.c};
  }
}

// Combines a replica's energy partials; which = 0 stores E_old, 1 stores
// E_new and applies IdealPV2::step's statistics (Stepper.cxx:37-55): if
// deltaE > 0 one uniform is drawn and compared with exp(-deltaE).
__global__ void k_energy_finish(const ReplicaParams *P, ReplicaState *S,
                                const int *blk_first, int nrep, int nblocks,
                                const EnergySlot *slots, int which, int ctr_off,
                                MonSlot *mon_ideal) {
  const int r = blockIdx.x;
  if (r >= nrep) return;
  const int b0 = blk_first[r];
  const int b1 = (r + 1 < nrep) ? blk_first[r + 1] : nblocks;
  KSum g{0, 0}, h{0, 0}, n{0, 0};
  for (int b = b0 + (int)threadIdx.x; b < b1; b += 64) {
    ksum_merge(g, KSum{slots[b].g_s, slots[b].g_c});
    ksum_merge(h, KSum{slots[b].h_s, slots[b].h_c});
    ksum_merge(n, KSum{slots[b].n_s, slots[b].n_c});
  }
  for (int o = 32; o > 0; o >>= 1) {
    KSum t;
    t.s = __shfl_xor(g.s, o, 64); t.c = __shfl_xor(g.c, o, 64); ksum_merge(g, t);
    t.s = __shfl_xor(h.s, o, 64); t.c = __shfl_xor(h.c, o, 64); ksum_merge(h, t);
    t.s = __shfl_xor(n.s, o, 64); t.c = __shfl_xor(n.c, o, 64); ksum_merge(n, t);
  }
  if (threadIdx.x != 0) return;
  const ReplicaParams &p = P[r];
  double e = 0;
  e += 0.5 / p.chi * ksum_value(n);
  e += 0.5 * ksum_value(g);
  e -= ksum_value(h);
  if (which == 0) {
    S[r].E_old = e;
    return;
  }
  S[r].E_new = e;
  const double deltaE = e - S[r].E_old;
  MonSlot &m = mon_ideal[r];
  if (deltaE > 0) {
    Xoro gg = xoro_seeded(site_key(stream_key(p.base_key, S[r].sweep_ctr + ctr_off), 0));
    const double prob = uniform01(gg);
    if (prob > exp(-deltaE)) { m.up_counts++; m.up_no++; }
    else { m.up_counts++; m.up_yes++; m.up_dS += deltaE; }
  } else {
    m.down_counts++;
    m.down_dS += deltaE;
  }
}

// ------------------------------------------------------------------------
__global__ void k_advance(ReplicaState *S, int nrep, double dt, int advance_time,
                          uint64_t nctr) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  if (advance_time) S[r].t += dt;
  S[r].sweep_ctr += nctr;
}

__global__ void k_set_time(ReplicaState *S, int nrep, int replica, double t) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  if (replica < 0 || r == replica) S[r].t = t;
}

// ------------------------------------------------------------------------
// o4fields [z][y][x][10] <-> red-black planes of the live buffer.
template <typename T>
__global__ void k_aos_to_rb(const ReplicaParams *P, const ReplicaState *S, int replica,
                            const double *aos) {
  const ReplicaParams p = P[replica];
  const int cur = live(S[replica]);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t V = 2 * p.half;
  if (s >= V) return;
  const int N = p.N;
  const int x = (int)(s % N);
  const int64_t yz = s / N;
  const int y = (int)(yz % N), z = (int)(yz / N);
  const int par = (x + y + z) & 1;
  const int64_t i = yz * p.NH + (x >> 1);
  for (int c = 0; c < 10; c++) plane<T>(p, cur, c, par)[i] = (T)aos[s * 10 + c];
}

template <typename T>
__global__ void k_rb_to_aos(const ReplicaParams *P, const ReplicaState *S, int replica,
                            double *aos) {
  const ReplicaParams p = P[replica];
  const int cur = live(S[replica]);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t V = 2 * p.half;
  if (s >= V) return;
  const int N = p.N;
  const int x = (int)(s % N);
  const int64_t yz = s / N;
  const int y = (int)(yz % N), z = (int)(yz / N);
  const int par = (x + y + z) & 1;
  const int64_t i = yz * p.NH + (x >> 1);
  for (int c = 0; c < 10; c++) aos[s * 10 + c] = (double)plane<T>(p, cur, c, par)[i];
}

// ------------------------------------------------------------------------
// Measurer observables of one site (measurer.cxx:245-269 and :351-392):
// o[0..10] = phi, A, V, phi^2; o[11..24] = |phi|, n, A', V', |phi|^2.
__device__ __forceinline__ void contract_rho(const double n[4], const double rho[6],
                                             double out[4]) {
  out[0] = -(n[1] * rho[0]) - n[2] * rho[1] - n[3] * rho[2];
  out[1] = n[0] * rho[0] + n[3] * rho[4] - n[2] * rho[5];
  out[2] = n[0] * rho[1] - n[3] * rho[3] + n[1] * rho[5];
  out[3] = n[0] * rho[2] + n[2] * rho[3] - n[1] * rho[4];
}

template <typename T>
__device__ __forceinline__ void site_obs(const ReplicaParams &p, int cur, int par, int64_t i,
                                         double o[25]) {
  double f[10];
  for (int c = 0; c < 10; c++) f[c] = (double)plane<T>(p, cur, c, par)[i];
  double phi2 = 0.0;
  for (int l = 0; l < 4; l++) phi2 += f[l] * f[l];
  for (int c = 0; c < 10; c++) o[c] = f[c];
  o[10] = phi2;
  const double sigma = sqrt(phi2);
  double n[4];
  if (phi2 > 100. * DBL_MIN) {
    for (int l = 0; l < 4; l++) n[l] = f[l] / sigma;
  } else {
    n[0] = 1; n[1] = 0; n[2] = 0; n[3] = 0;
  }
  const double rho[6] = {f[4], f[5], f[6], f[7], f[8], f[9]};
  const double drho[6] = {f[7], f[8], f[9], f[4], f[5], f[6]};
  double A[4], V[4];
  contract_rho(n, rho, A);
  contract_rho(n, drho, V);
  o[11] = sigma;
  for (int l = 0; l < 4; l++) {
    o[12 + l] = n[l];
    o[16 + l] = A[l];
    o[20 + l] = V[l];
  }
  o[24] = sigma * sigma;
}

// One thread per (x, z): compensated sum over y -> partXZ[o][z][x];
// one thread per (x, y): compensated sum over z -> partXY[o][y][x].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_measure_cols(const ReplicaParams *P,
                                                         const ReplicaState *S, int replica,
                                                         int along_z, double *part) {
  const ReplicaParams p = P[replica];
  const int cur = live(S[replica]);
  const int N = p.N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)N * N) return;
  const int x = (int)(t % N), w = (int)(t / N);
  KSum acc[25];
  for (int o = 0; o < 25; o++) acc[o] = KSum{0., 0.};
  for (int v = 0; v < N; v++) {
    const int y = along_z ? w : v;
    const int z = along_z ? v : w;
    const int par = (x + y + z) & 1;
    const int64_t i = ((int64_t)z * N + y) * p.NH + (x >> 1);
    double o[25];
    site_obs<T>(p, cur, par, i, o);
    for (int k = 0; k < 25; k++) ksum_add(acc[k], o[k]);
  }
  for (int k = 0; k < 25; k++) part[((int64_t)k * N + w) * N + x] = ksum_value(acc[k]);
}

// wall[d][obs][n]: d = 0 (x) sums partXZ over z, d = 2 (z) sums partXZ over
// x, d = 1 (y) sums partXY over x; then / N^2 as measurer.cxx:274-280.
__global__ void k_measure_reduce(int N, const double *partXZ, const double *partXY,
                                 double *wall, double *wallphase) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * 25 * N) return;
  const int n = t % N;
  const int k = (t / N) % 25;
  const int d = t / (25 * N);
  KSum acc{0., 0.};
  for (int m = 0; m < N; m++) {
    double v;
    if (d == 0) v = partXZ[((int64_t)k * N + m) * N + n];
    else if (d == 2) v = partXZ[((int64_t)k * N + n) * N + m];
    else v = partXY[((int64_t)k * N + n) * N + m];
    ksum_add(acc, v);
  }
  const double val = ksum_value(acc) / (double)(N * N);
  if (k < 11) wall[((int64_t)d * 11 + k) * N + n] = val;
  else wallphase[((int64_t)d * 14 + (k - 11)) * N + n] = val;
}

// measurer_fft::execute_row (measurer.h:33-41): FFTW's r2c forward
// transform divided by N, out[k] = (1/N) sum_x in[x] e^{-2 pi i k x / N},
// k = 0..N/2, as a direct DFT with exactly reduced twiddles and compensated
// sums.  One thread per (row, k); rows = 3 x 25 walls.
__global__ void k_wall_dft(int N, const double *wall, const double *wallphase, double *wall_k,
                           double *wallphase_k) {
  const int nk = N / 2 + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * 25 * nk) return;
  const int k = t % nk;
  const int row = t / nk;  // d*25 + obs
  const int d = row / 25, o = row % 25;
  const double *in = o < 11 ? wall + ((int64_t)d * 11 + o) * N
                            : wallphase + ((int64_t)d * 14 + (o - 11)) * N;
  KSum re{0., 0.}, im{0., 0.};
  for (int x = 0; x < N; x++) {
    const int m = (int)(((int64_t)k * x) % N);
    double s, c;
    sincospi(2.0 * m / N, &s, &c);
    ksum_add(re, in[x] * c);
    ksum_add(im, -(in[x] * s));
  }
  double *out = o < 11 ? wall_k + (((int64_t)d * 11 + o) * nk + k) * 2
                       : wallphase_k + (((int64_t)d * 14 + (o - 11)) * nk + k) * 2;
  out[0] = ksum_value(re) / N;
  out[1] = ksum_value(im) / N;
}

// ======================================================================== launchers

void launch_wall_dft(int N, const double *wall, const double *wallphase, double *wall_k,
                     double *wallphase_k, hipStream_t st) {
  const int n = 3 * 25 * (N / 2 + 1);
  k_wall_dft<<<(n + 127) / 128, 128, 0, st>>>(N, wall, wallphase, wall_k, wallphase_k);
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_hb_phi(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                   int ieo, int ctr_off, MonSlot *mon, hipStream_t st) {
  if (fp32)
    k_hb_phi<float><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep, plan.d_blk_first,
                                                     ieo, ctr_off, mon);
  else
    k_hb_phi<double><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep, plan.d_blk_first,
                                                      ieo, ctr_off, mon);
}

void launch_shuffle(int nrep, ReplicaState *S, const ReplicaParams *P, int ctr_off,
                    hipStream_t st) {
  k_shuffle<<<nblk(nrep, 64), 64, 0, st>>>(S, P, nrep, ctr_off);
}

void launch_charge_dir(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                       int nrep, int ixyz, int ctr_off, MonSlot *mon, hipStream_t st) {
  if (fp32)
    k_charge_dir<float><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep,
                                                         plan.d_blk_first, ixyz, ctr_off, mon);
  else
    k_charge_dir<double><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep,
                                                          plan.d_blk_first, ixyz, ctr_off, mon);
  k_flip_chg<<<nblk(nrep, 64), 64, 0, st>>>(S, nrep);
}

void launch_ideal(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                  int nrep, hipStream_t st) {
  if (fp32)
    k_ideal_fused<float><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep,
                                                          plan.d_blk_first);
  else
    k_ideal_fused<double><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep,
                                                           plan.d_blk_first);
  k_flip<<<nblk(nrep, 64), 64, 0, st>>>(S, nrep);
}

void launch_energy(bool fp32, const Plan &plan, const ReplicaParams *P,
                   const ReplicaState *S, EnergySlot *slots, hipStream_t st) {
  if (fp32)
    k_energy<float><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep, plan.d_blk_first,
                                                     slots);
  else
    k_energy<double><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep, plan.d_blk_first,
                                                      slots);
}

void launch_energy_finish(int nrep, const Plan &plan, const ReplicaParams *P,
                          ReplicaState *S, const EnergySlot *slots, int which,
                          int ctr_off, MonSlot *mon_ideal, hipStream_t st) {
  k_energy_finish<<<nrep, 64, 0, st>>>(P, S, plan.d_blk_first, nrep, plan.nblocks, slots,
                                       which, ctr_off, mon_ideal);
}

void launch_advance(int nrep, ReplicaState *S, double dt, int advance_time, uint64_t nctr,
                    hipStream_t st) {
  k_advance<<<nblk(nrep, 64), 64, 0, st>>>(S, nrep, dt, advance_time, nctr);
}

void launch_set_time(int nrep, int replica, ReplicaState *S, double t, hipStream_t st) {
  k_set_time<<<nblk(nrep, 64), 64, 0, st>>>(S, nrep, replica, t);
}

void launch_aos_to_rb(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                      int N, const double *aos, hipStream_t st) {
  const int64_t V = (int64_t)N * N * N;
  if (fp32) k_aos_to_rb<float><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, aos);
  else k_aos_to_rb<double><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, aos);
}

void launch_rb_to_aos(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                      int N, double *aos, hipStream_t st) {
  const int64_t V = (int64_t)N * N * N;
  if (fp32) k_rb_to_aos<float><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, aos);
  else k_rb_to_aos<double><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, aos);
}

void launch_measure(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                    int N, double *partXZ, double *partXY, double *wall, double *wallphase,
                    hipStream_t st) {
  const int64_t cols = (int64_t)N * N;
  if (fp32) {
    k_measure_cols<float><<<nblk(cols, kBlock), kBlock, 0, st>>>(P, S, replica, 0, partXZ);
    k_measure_cols<float><<<nblk(cols, kBlock), kBlock, 0, st>>>(P, S, replica, 1, partXY);
  } else {
    k_measure_cols<double><<<nblk(cols, kBlock), kBlock, 0, st>>>(P, S, replica, 0, partXZ);
    k_measure_cols<double><<<nblk(cols, kBlock), kBlock, 0, st>>>(P, S, replica, 1, partXY);
  }
  k_measure_reduce<<<nblk(3 * 25 * N, 256), 256, 0, st>>>(N, partXZ, partXY, wall, wallphase);
}

}  // namespace lgv
