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

// Global-memory view of one plane for the sweep kernels.  The base is
// uniform over a block (it sits in SGPRs) and, with W == false, the element
// offset is a 32-bit byte offset, so every access is one global load/store
// in the saddr + 32-bit voffset form with no 64-bit address arithmetic.
// W == true (a plane of 4 GiB or more) widens the offset to 64 bits.
typedef __attribute__((address_space(1))) char gbyte;
template <typename T, bool W>
struct GPlane {
  gbyte *b;
  __device__ __forceinline__ static auto off(uint32_t i) {
    if constexpr (W) return (uint64_t)i * sizeof(T);
    else return (uint32_t)(i * (uint32_t)sizeof(T));
  }
  __device__ __forceinline__ T raw(uint32_t i) const {
    return *reinterpret_cast<const __attribute__((address_space(1))) T *>(b + off(i));
  }
  __device__ __forceinline__ double ld(uint32_t i) const { return (double)raw(i); }
  __device__ __forceinline__ void st(uint32_t i, double v) const {
    *reinterpret_cast<__attribute__((address_space(1))) T *>(b + off(i)) = (T)v;
  }
  // NTS: a non-temporal load, for a value no other thread reads
  template <bool NTS>
  __device__ __forceinline__ T get_raw(uint32_t i) const {
    if constexpr (NTS)
      return __builtin_nontemporal_load(
          reinterpret_cast<const __attribute__((address_space(1))) T *>(b + off(i)));
    else
      return raw(i);
  }
  template <bool NTS>
  __device__ __forceinline__ double get(uint32_t i) const { return (double)get_raw<NTS>(i); }
  // NTS: a non-temporal store (not kept in the caches)
  template <bool NTS>
  __device__ __forceinline__ void put_raw(uint32_t i, T v) const {
    if constexpr (NTS)
      __builtin_nontemporal_store(
          v, reinterpret_cast<__attribute__((address_space(1))) T *>(b + off(i)));
    else
      *reinterpret_cast<__attribute__((address_space(1))) T *>(b + off(i)) = v;
  }
  template <bool NTS>
  __device__ __forceinline__ void put(uint32_t i, double v) const {
    put_raw<NTS>(i, (T)v);
  }
};
template <typename T, bool W>
__device__ __forceinline__ GPlane<T, W> gplane(const ReplicaParams &p, int cur, int c, int par) {
  const int buf = c < 4 ? (cur & 1) : (cur >> 1);
  char *base = reinterpret_cast<char *>(buf ? p.fields[1] : p.fields[0]);
  return {(gbyte *)(base + (int64_t)(c * 2 + par) * p.half * (int64_t)sizeof(T))};
}
// The two field buffers of a replica read once into registers: a kernel
// that stores between its loads then forms every plane base from these
// values, where gplane(p, ...) would reload p.fields after each store (the
// compiler cannot prove that a field store misses the parameter block).
struct FieldBufs {
  char *f0, *f1;
  int64_t half_bytes;  // one half-plane
};
template <typename T>
__device__ __forceinline__ FieldBufs field_bufs(const ReplicaParams &p) {
  return {reinterpret_cast<char *>(p.fields[0]), reinterpret_cast<char *>(p.fields[1]),
          p.half * (int64_t)sizeof(T)};
}
template <typename T, bool W>
__device__ __forceinline__ GPlane<T, W> gplane(const FieldBufs &fb, int cur, int c, int par) {
  const int buf = c < 4 ? (cur & 1) : (cur >> 1);
  return {(gbyte *)((buf ? fb.f1 : fb.f0) + (int64_t)(c * 2 + par) * fb.half_bytes)};
}

// Block index of this workgroup within the launch plan.  (A renumbering that
// gives each XCD a contiguous eighth of the plan was measured on MI355X and
// made the ideal and charge sweeps slower; the ideal kernel and the phi march
// map one-replica launches with yband_tile instead, the charge sweep keeps
// dispatch order.)
__device__ __forceinline__ int xcd_block() { return blockIdx.x; }

__device__ __forceinline__ int wrapp(int a, int N) { return a == N ? 0 : a; }
__device__ __forceinline__ int wrapm(int a, int N) { return a < 0 ? N - 1 : a; }

// Deterministic block reduction of the monitor increments; thread 0 adds the
// block total into the block's own slot (stream order makes this race-free).
// The three counts travel packed in one 64-bit word, 21 bits each (a block
// counts at most 12 decisions per thread).
__device__ __forceinline__ uint64_t mon_pack(uint64_t down, uint64_t upyes, uint64_t upno) {
  return down | (upyes << 21) | (upno << 42);
}
__device__ void block_monitor_add(MonSlot *slot, uint64_t cnt, double dSd, double dSu) {
  __shared__ uint64_t sc[kBlock / 64];
  __shared__ double sd[2][kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    cnt += __shfl_xor(cnt, o, 64);
    dSd += __shfl_xor(dSd, o, 64);
    dSu += __shfl_xor(dSu, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sc[w] = cnt;
    sd[0][w] = dSd; sd[1][w] = dSu;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t c = 0;
    double d = 0, e = 0;
    for (int i = 0; i < kBlock / 64; i++) {
      c += sc[i];
      d += sd[0][i]; e += sd[1][i];
    }
    const long long a = (long long)(c & 0x1fffff), b = (long long)((c >> 21) & 0x1fffff),
                    n = (long long)(c >> 42);
    slot->down_counts += a;
    slot->up_counts += b + n;
    slot->up_yes += b;
    slot->up_no += n;
    slot->down_dS += d;
    slot->up_dS += e;
  }
}

// ------------------------------------------------------------------------
// The Metropolis test canonical(u) < exp(x) for the draw u and x = -dS <= 0,
// decided as the double test decides it.  Fast path (x > -11, exp(x) >
// 1.6e-5): the draw's top 32 bits in single precision against single-
// precision bounds e (1 -+ 1e-3) 2^32, e = __expf(x).  With r = 2^-32
// float(u >> 32), |canonical(u) - r| <= 2^-32 + 2^-24 r, and e is within
// 2e-6 of exp(x) relative (x and x log2(e) rounded to float, one-ulp
// v_exp_f32), so a draw below the lower bound lies below exp(x) and one at
// or above the upper bound does not: the margin, about 1e-3 exp(x), exceeds
// the 2.4e-10 of the draw's rounding whenever exp(x) > 2.4e-7.  The double
// draw and the double exp decide the rest (about 2e-3 of the tests).
// The double test, out of line: it runs for about 2e-3 of the tests, and
// inlined at every decision of a sweep kernel its exp makes up a sizeable
// part of the kernel's code.
__device__ __attribute__((noinline)) bool accept_exact(uint64_t u, double x) {
  return canonical(u) < exp(x);
}
__device__ __forceinline__ bool accept_draw(uint64_t u, double x) {
  // the bounds are formed for every x (a very negative x gives e = 0, no
  // NaN) and the test falls back to the double path outside x > -11; the
  // empty asm keeps the draw's top half a 32-bit value (otherwise it is
  // converted as a 64-bit one, six instructions instead of one)
  uint32_t h = (uint32_t)(u >> 32);
  __asm__("" : "+v"(h));
  const float e = __expf((float)x);
  const float r = (float)h;
  const bool below = r < e * (0.999f * 0x1p32f);
  const bool above = r >= e * (1.001f * 0x1p32f);
  if (x > -11. && (below || above)) return below;
  // the draw rebuilt from h (the same bits), so that u's top half need not
  // be kept beside h
  return accept_exact(((uint64_t)h << 32) | (uint32_t)u, x);
}

// The Metropolis decision of one phi site (Stepper.cxx:330-370) in the
// reference's operand order: heff[l] is the neighbour sum of component l,
// ((x+1 + x-1) + (y+1 + y-1)) + (z+1 + z-1).  Returns the new value (the
// old one on reject) and adds the decision to the monitor when count.
__device__ __forceinline__ void hb_decide(const double phi_o[4], const double heff[4],
                                          double Lambda2, double lambda, double H, double rdtg,
                                          Xoro &g, double phi_new[4], uint64_t &cnt,
                                          double &dSd, double &dSu, bool count) {
  double hphi_o = 0., hphi_n = 0., s_o = 0., s_n = 0.;
  double phi_n[4];
#pragma unroll
  for (int l = 0; l < 4; l++) {
    phi_n[l] = phi_o[l] + rdtg * variance1(g);
    s_o += phi_o[l] * phi_o[l];
    s_n += phi_n[l] * phi_n[l];
    hphi_o += heff[l] * phi_o[l];
    hphi_n += heff[l] * phi_n[l];
  }
  hphi_o += (H * phi_o[0]);
  hphi_n += (H * phi_n[0]);
  const double dS = -(hphi_n - hphi_o) + Lambda2 * (s_n - s_o) +
                    lambda / 4. * (s_n * s_n - s_o * s_o);
  bool accept;
  if (dS < 0) {
    accept = true;
    if (count) { cnt += mon_pack(1, 0, 0); dSd += dS; }
  } else {
    accept = accept_draw(xoro_next(g), -dS);
    if (count) {
      if (accept) { cnt += mon_pack(0, 1, 0); dSu += dS; } else { cnt += mon_pack(0, 0, 1); }
    }
  }
#pragma unroll
  for (int l = 0; l < 4; l++) phi_new[l] = accept ? phi_n[l] : phi_o[l];
}

// One parity of EulerLangevinHB::step (ModelG/Stepper.cxx:309-386), one
// thread per site, in place.
template <typename T, bool W>
__global__ __launch_bounds__(kBlock) void k_hb_phi(const ReplicaParams *__restrict__ P,
                                                   const ReplicaState *__restrict__ S,
                                                   const int *__restrict__ blk_rep,
                                                   const int *__restrict__ blk_first, int ieo,
                                                   int ctr_off, int flip,
                                                   MonSlot *__restrict__ mon) {
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  const ReplicaParams &p = P[r];
  // FP64 one-replica launches sweep plane bands per XCD (measured: 128^3
  // 29.0 -> 26.3 us, 192^3 74.4 -> 70.1 us; FP32 256^3 1 % slower)
  int pb = bid - blk_first[r];
  const int bpp = p.N * p.NH / kBlock;  // blocks per plane
  if (sizeof(T) == 8 && (p.N * p.NH) % kBlock == 0 && (bpp & 7) == 0 &&
      (int)gridDim.x == p.N * bpp)
    pb = band_block(pb, bpp);
  const uint32_t idx = (uint32_t)pb * kBlock + threadIdx.x;
  uint64_t cnt = 0;
  double dSd = 0., dSu = 0.;
  if (idx < (uint64_t)p.half) {
    const uint32_t N = p.N, NH = p.NH;
    const uint32_t rowi = idx / NH, xh = idx - rowi * NH;
    const uint32_t z = rowi / N, y = rowi - z * N;
    const uint32_t x = site_x(xh, y, z, ieo);
    const int q = ieo ^ 1;
    const uint32_t row = rowi * NH;
    const uint32_t xp = x + 1 == N ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
    const uint32_t yp = y + 1 == N ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
    const uint32_t zp = z + 1 == N ? 0 : z + 1, zm = z == 0 ? N - 1 : z - 1;
    const uint32_t nb[6] = {row + (xp >> 1), row + (xm >> 1), (z * N + yp) * NH + xh,
                            (z * N + ym) * NH + xh, (zp * N + y) * NH + xh,
                            (zm * N + y) * NH + xh};
    const ReplicaState &s = S[r];
    const int cur = s.cur_phi ^ (flip & 1);  // phi planes only
    const double mass = p.mass0 + p.dmassdt * s.t;
    const double Lambda2 = 0.5 * (2. * 3. + mass);
    Xoro g = xoro_seeded(site_key(stream_key(p.base_key, s.sweep_ctr + ctr_off),
                                  (uint64_t)rowi * N + x));
    double phi_o[4], heff[4], phi_n[4];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      const GPlane<T, W> fq = gplane<T, W>(p, cur, l, q);
      heff[l] = (fq.ld(nb[0]) + fq.ld(nb[1])) + (fq.ld(nb[2]) + fq.ld(nb[3])) +
                (fq.ld(nb[4]) + fq.ld(nb[5]));
      phi_o[l] = gplane<T, W>(p, cur, l, ieo).ld(idx);
    }
    hb_decide(phi_o, heff, Lambda2, p.lambda, p.H, p.rdtg, g, phi_n, cnt, dSd, dSu, true);
#pragma unroll
    for (int l = 0; l < 4; l++) gplane<T, W>(p, cur, l, ieo).st(idx, phi_n[l]);
  }
  block_monitor_add(mon + bid, cnt, dSd, dSu);
}

__device__ __forceinline__ int wrapz(int z, int N) { return z < 0 ? z + N : (z >= N ? z - N : z); }

// One parity of EulerLangevinHB::step (Stepper.cxx:309-386), in place, by
// marching z through a TXH x TY tile of half-columns (the tile of hb_tile).
// The update parity ieo reads only the other parity q, which no thread of
// the launch writes: per plane, the q values of the tile and of a one-site
// ring sit in a packed LDS grid (one entry per half-column and row; two
// grids alternate, one barrier per plane), the q values of the thread's own
// half-column at z-1 and z+1 in registers.  Each q value is loaded once per
// plane (plus the ring), each decision is made once.
template <typename T, bool W, bool NTS>
__global__ __launch_bounds__(kBlock) void k_hb_march(const ReplicaParams *__restrict__ P,
                                                     const ReplicaState *__restrict__ S,
                                                     const int *__restrict__ blk_rep,
                                                     const int *__restrict__ blk_first, int ieo,
                                                     int ctr_off, int flip,
                                                     MonSlot *__restrict__ mon) {
  constexpr int kGrid = kHbMarchGrid;
  // values held widened to FP64 (exact for FP32 fields): each is read by up
  // to four sites, so one conversion on the way in replaces four on the way out
  __shared__ double Q[2][4][kGrid];
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  const ReplicaParams &p = P[r];
  const ReplicaState &s = S[r];
  const int cur = s.cur_phi ^ (flip & 1);  // phi planes only
  const int q = ieo ^ 1;
  const int N = p.N, NH = p.NH;
  int TXH, TY, TZ;
  hb_tile(N, p.hb_cap, p.hb_tz, TXH, TY, TZ);
  const int ntx = (NH + TXH - 1) / TXH, nty = (N + TY - 1) / TY;
  int b = bid - blk_first[r];
  int bx, by, bz;
  if ((int)gridDim.x == ntx * nty * ((N + TZ - 1) / TZ)) {
    yband_tile(b, ntx, nty, bx, by, bz);  // one replica
  } else {
    bx = b % ntx;
    b /= ntx;
    by = b % nty;
    bz = b / nty;
  }
  const int xh0 = bx * TXH, y0 = by * TY, z0 = bz * TZ;
  const int tw = min(TXH, NH - xh0), th = min(TY, N - y0), tz = min(TZ, N - z0);
  const int GW = tw + 2;  // grid row width: half-columns xh0-1 .. xh0+tw
  const int t = threadIdx.x;
  const int tx = t % TXH, ty = t / TXH;
  const bool own = t < TXH * TY && ty < th && tx < tw;
  const int xh = xh0 + tx, y = y0 + ty;
  // ring entry of this thread: left/right half-column, rows above/below
  const bool halo = t < 2 * th + 2 * tw;
  int hlx, hly;
  if (t < th) { hlx = -1; hly = t; }
  else if (t < 2 * th) { hlx = tw; hly = t - th; }
  else if (t < 2 * th + tw) { hlx = t - 2 * th; hly = -1; }
  else { hlx = t - 2 * th - tw; hly = th; }
  const int hgxh = wrapz(xh0 + hlx, NH), hgy = wrapz(y0 + hly, N);
  const int hpos = (hly + 1) * GW + (hlx + 1);

  const double mass = p.mass0 + p.dmassdt * s.t;
  const double Lambda2 = 0.5 * (2. * 3. + mass);
  const double lambda = p.lambda, H = p.H, rdtg = p.rdtg;
  const uint64_t key = stream_key(p.base_key, s.sweep_ctr + ctr_off);
  uint64_t cnt = 0;
  double dSd = 0., dSu = 0.;
  auto idx = [&](int xhh, int yy, int z) __attribute__((always_inline)) {
    return ((uint32_t)z * N + yy) * NH + xhh;
  };
  // the eight plane bases, formed once (a store would otherwise force
  // p.fields to be reloaded)
  GPlane<T, W> gq[4], go[4];
#pragma unroll
  for (int l = 0; l < 4; l++) {
    gq[l] = gplane<T, W>(p, cur, l, q);
    go[l] = gplane<T, W>(p, cur, l, ieo);
  }
  auto ld4 = [&](uint32_t i, double v[4]) __attribute__((always_inline)) {
#pragma unroll
    for (int l = 0; l < 4; l++) v[l] = gq[l].ld(i);
  };
  // The loads of plane z+1 (the own half-column's q values at z+2, its own
  // update-parity value at z+1, the ring entry at z+1) are issued before
  // plane z's barrier, into the registers nq, nphi, nh as stored, so that
  // they are in flight during plane z's decisions.
  auto raw4 = [&](uint32_t i, T v[4]) __attribute__((always_inline)) {
#pragma unroll
    for (int l = 0; l < 4; l++) v[l] = gq[l].raw(i);
  };
  // the update parity's own value: read by this thread only
  auto own4 = [&](uint32_t i, T v[4]) __attribute__((always_inline)) {
#pragma unroll
    for (int l = 0; l < 4; l++) v[l] = go[l].template get_raw<NTS>(i);
  };
  // Plane z's new values are stored in iteration z+1, after the loads of
  // plane z+2 are issued and before the barrier: one in-order counter covers
  // loads and stores, so the next iteration's wait for those loads also
  // waits for these stores, which have had the decisions of plane z+1 to
  // complete (a store at the end of an iteration would be waited for at
  // once).  Only this thread reads its own sites.
  auto store4 = [&](uint32_t i, const T v[4]) __attribute__((always_inline)) {
#pragma unroll
    for (int l = 0; l < 4; l++) go[l].template put_raw<NTS>(i, v[l]);
  };
  // FP32 fields: the loads run two planes ahead in two register sets (a
  // set holds half the bytes of an FP64 one, so one set in flight per wave
  // leaves the memory system short of requests; 256^3 half-sweep 94 -> 92
  // us, 119 registers, still four waves per SIMD); FP64: one plane ahead.
  constexpr int kDepth = sizeof(T) == 4 ? 2 : 1;
  double qm[4], qc[4], qp[4];  // q values of the own half-column at z-1, z, z+1
  // set s: the own q values at z+1, the own update-parity value at z and the
  // ring entry at z, for the planes z = z0 + s, z0 + s + kDepth, ...
  T nq[kDepth][4], nphi[kDepth][4], nh[kDepth][4];
  T pend[4];         // plane z-1's new values, stored at the top of iteration z
  uint32_t pend_i = 0;
  if (own) {
    ld4(idx(xh, y, wrapz(z0 - 1, N)), qm);
    ld4(idx(xh, y, z0), qc);
  }
#pragma unroll
  for (int sidx = 0; sidx < kDepth; sidx++) {
    if (sidx < tz) {  // plane z0 + sidx < z0 + tz <= N
      const int zs = z0 + sidx;
      if (own) {
        raw4(idx(xh, y, zs + 1 == N ? 0 : zs + 1), nq[sidx]);
        own4(idx(xh, y, zs), nphi[sidx]);
      }
      if (halo) raw4(idx(hgxh, hgy, zs), nh[sidx]);
    }
  }
  // plane z0 + j from register set sq / sphi / sh, which then takes the
  // loads of plane z0 + j + kDepth
  auto plane = [&](int j, T (&sq)[4], T (&sphi)[4], T (&sh)[4]) __attribute__((always_inline)) {
    const int z = z0 + j;
    double(*g)[kGrid] = Q[j & 1];
    double phi_o[4];
    if (own) {
      const int pos = (ty + 1) * GW + (tx + 1);
#pragma unroll
      for (int l = 0; l < 4; l++) {
        qp[l] = sq[l];
        phi_o[l] = sphi[l];
        g[l][pos] = qc[l];
      }
    }
    if (halo) {
#pragma unroll
      for (int l = 0; l < 4; l++) g[l][hpos] = sh[l];
    }
    if (j + kDepth < tz) {  // plane zn = z + kDepth < z0 + tz <= N
      const int zn = z + kDepth;
      if (own) {
        raw4(idx(xh, y, zn + 1 == N ? 0 : zn + 1), sq);
        own4(idx(xh, y, zn), sphi);
      }
      if (halo) raw4(idx(hgxh, hgy, zn), sh);
    }
    if (own && j > 0) store4(pend_i, pend);
    __syncthreads();
    if (own) {
      const int o = (ieo + y + z) & 1;  // x = 2 xh + o
      const int pos = (ty + 1) * GW + (tx + 1);
      double heff[4], phi_n[4];
#pragma unroll
      for (int l = 0; l < 4; l++)
        heff[l] = (g[l][pos + o] + g[l][pos - 1 + o]) +
                  (g[l][pos + GW] + g[l][pos - GW]) + (qp[l] + qm[l]);
      Xoro gg = xoro_seeded(site_key(key, ((uint64_t)z * N + y) * N + 2 * xh + o));
      hb_decide(phi_o, heff, Lambda2, lambda, H, rdtg, gg, phi_n, cnt, dSd, dSu, true);
      pend_i = idx(xh, y, z);
#pragma unroll
      for (int l = 0; l < 4; l++) {
        pend[l] = (T)phi_n[l];
        qm[l] = qc[l];
        qc[l] = qp[l];
      }
    }
  };
  // tz is the same for the whole block, so every thread makes the same calls
  // (and reaches the same barriers)
  for (int j = 0; j < tz; j += kDepth) {
    plane(j, nq[0], nphi[0], nh[0]);
    if constexpr (kDepth == 2) {
      if (j + 1 < tz) plane(j + 1, nq[kDepth - 1], nphi[kDepth - 1], nh[kDepth - 1]);
    }
  }
  if (own && tz > 0) store4(pend_i, pend);
  block_monitor_add(mon + bid, cnt, dSd, dSu);
}


// ------------------------------------------------------------------------
// modelg_update_charge_pair (Stepper.cxx:419-441) for one component of one
// pair, applied as nA + (-q), nB + q; counts the decision when count.
// The constants of a charge sweep: the proposal width and the divisor 2 chi
// of dS with its rounded reciprocal.
struct ChargeK {
  double rms, d, inv;
};
__device__ __forceinline__ ChargeK charge_k(double rms, double chi) {
  const double d = 2. * chi;
  return {rms, d, 1. / d};
}
// x / d from the rounded reciprocal with one Markstein correction: q0 = x *
// inv, the residual x - q0 d exact with one FMA, q0 + residual * inv rounded
// once.  It equals the IEEE quotient for these divisors (no difference in
// 1.6e8 random dividends over eight values of chi, including the defaults;
// the CPU suite checks it); a one-ulp difference could only change a
// decision if the uniform draw fell within that ulp of exp(-dS).  Three
// instructions instead of the division's ten.
__device__ __forceinline__ double div_by(double x, const ChargeK &k) {
  const double q0 = x * k.inv;
  return fma(fma(-q0, k.d, x), k.inv, q0);
}

__device__ __forceinline__ void charge_pair(double nA, double nB, const ChargeK &k,
                                            Xoro &g, double &outA, double &outB, uint64_t &cnt,
                                            double &dSd, double &dSu, bool count) {
  double q = k.rms * variance1(g);
  const double dS = div_by((nA - q) * (nA - q) + (nB + q) * (nB + q) - nA * nA - nB * nB, k);
  if (dS < 0) {
    if (count) { cnt += mon_pack(1, 0, 0); dSd += dS; }
  } else {
    if (accept_draw(xoro_next(g), -dS)) {
      if (count) { cnt += mon_pack(0, 1, 0); dSu += dS; }
    } else {
      if (count) cnt += mon_pack(0, 0, 1);
      q = 0.;
    }
  }
  outA = nA + (-q);
  outB = nB + q;
}

// The two sub-sweeps of one direction (the shuffled order always runs both
// parities of a direction back to back, Stepper.cxx:471-473) for one site
// A' of the second parity e2: along the direction its line reads
// L = A'-e, A', B' = A'+e, R = A'+2e.  First sub-sweep (parity e1): pairs
// (L, A') and (B', R), keyed by L and B'; second: pair (A', B') keyed by A'
// on the first sub-sweep's results.  Each first-parity pair is computed by
// the two threads that need it (identical counter streams, identical bits);
// the monitor counts it once, from the thread on its right.  Reads the live
// charge buffer, writes A' and B' into the other one.
template <typename T, bool W>
__device__ __forceinline__ void charge_line(const ReplicaParams &p, int cur, int nxt, int dir,
                                            int e1, int e2, uint32_t xh, uint32_t y, uint32_t z,
                                            uint64_t k1, uint64_t k2, uint64_t &cnt,
                                            double &dSd, double &dSu) {
  const uint32_t N = p.N, NH = p.NH;
  const uint32_t x = site_x(xh, y, z, e2);
  // the four sites along the direction: L, A', B', R (parities e1 e2 e1 e2)
  const uint32_t a0 = dir == 0 ? x : (dir == 1 ? y : z);
  const uint32_t am = a0 == 0 ? N - 1 : a0 - 1;
  const uint32_t ap = a0 + 1 == N ? 0 : a0 + 1;
  const uint32_t app = ap + 1 == N ? 0 : ap + 1;
  uint32_t ii[4], lin[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t a = k == 0 ? am : (k == 1 ? a0 : (k == 2 ? ap : app));
    const uint32_t xx = dir == 0 ? a : x, yy = dir == 1 ? a : y, zz = dir == 2 ? a : z;
    ii[k] = (zz * N + yy) * NH + (xx >> 1);
    lin[k] = (zz * N + yy) * N + xx;
  }
  const ChargeK ck = charge_k(p.rms, p.chi);
  const FieldBufs fb = field_bufs<T>(p);
  Xoro gL = xoro_seeded(site_key(k1, lin[0]));
  Xoro gB = xoro_seeded(site_key(k1, lin[2]));
  Xoro gA = xoro_seeded(site_key(k2, lin[1]));
  uint64_t c0 = 0;
  double sd0 = 0., su0 = 0.;
  // every load of the line before the first store (see charge_column)
  double w[6][4];
#pragma unroll
  for (int L = 4; L < 10; L++) {
    const GPlane<T, W> f1 = gplane<T, W>(fb, cur, L, e1), f2 = gplane<T, W>(fb, cur, L, e2);
    w[L - 4][0] = f1.ld(ii[0]);
    w[L - 4][1] = f2.ld(ii[1]);
    w[L - 4][2] = f1.ld(ii[2]);
    w[L - 4][3] = f2.ld(ii[3]);
  }
#pragma unroll
  for (int L = 4; L < 10; L++) {
    const double vL = w[L - 4][0];
    const double vA = w[L - 4][1];
    const double vB = w[L - 4][2];
    const double vR = w[L - 4][3];
    double tL, a1, b1, tR;
    charge_pair(vL, vA, ck, gL, tL, a1, cnt, dSd, dSu, true);
    charge_pair(vB, vR, ck, gB, b1, tR, c0, sd0, su0, false);
    // the first sub-sweep's results as stored (rounded in FP32 storage)
    a1 = (double)(T)a1;
    b1 = (double)(T)b1;
    double a2, b2;
    charge_pair(a1, b1, ck, gA, a2, b2, cnt, dSd, dSu, true);
    gplane<T, W>(fb, nxt, L, e2).st(ii[1], a2);
    gplane<T, W>(fb, nxt, L, e1).st(ii[2], b2);
  }
}

// Lane 0's value of v, uniform over the wave (v_readfirstlane per word).
__device__ __forceinline__ float first_lane(float v) {
  return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}
__device__ __forceinline__ double first_lane(double v) {
  const long long bits = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)bits);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(bits >> 32));
  return __longlong_as_double(((long long)hi << 32) | lo);
}

// The x direction with a whole row in one wavefront or in one aligned
// segment of seg lanes of it (NH = seg H; seg = NH <= 64 a power of two for
// H = 1, seg = 64 for H = 2): lane l of the segment owns the second-parity
// sites A'_h of half-columns h = l + 64 j, j < H.  First sub-sweep: the lane
// updates the pair (L_h, A'_h), L_h = A'_h - e_x.  The updated L_{h+1} is
// B'_h = A'_h + e_x; it comes from the lane that owns h + 1 (lane l + 1, or
// for the segment's last lane its lane 0, at j + 1 for H = 2, the row
// wrapping periodically).  Second sub-sweep: the lane updates (A'_h, B'_h)
// and writes both.  Every pair is computed once and counted once, from the
// same counter streams as charge_line.  Index logic checked on the CPU
// against the plain two sub-sweeps (exact integer stand-in, N = 4 ... 256).
template <typename T, bool W, int H>
__device__ __forceinline__ void charge_row_wave(const ReplicaParams &p, int cur, int nxt, int e1,
                                                int e2, uint32_t y, uint32_t z, int lane, int seg,
                                                uint64_t k1, uint64_t k2, uint64_t &cnt,
                                                double &dSd, double &dSu) {
  static_assert(H == 1 || H == 2, "rows of up to 128 half-columns");
  const uint32_t N = p.N, NH = p.NH;
  const uint32_t row = z * N + y;
  const ChargeK ck = charge_k(p.rms, p.chi);
  const FieldBufs fb = field_bufs<T>(p);
  // a1: updated A'_h; r: updated L_{h+1} from lane + 1; l0: lane 0's updated
  // L (H = 2 only); all as stored (T), which is how the second sub-sweep reads
  // them, so that FP32 fields hold and shuffle half the registers
  T a1[H][6], r[H][6], l0[H][6];
#pragma unroll
  for (int j = 0; j < H; j++) {
    const uint32_t h = lane + 64 * j;
    const uint32_t x = site_x(h, y, z, e2);
    const uint32_t xl = x == 0 ? N - 1 : x - 1;
    const uint32_t iA = row * NH + h, iL = row * NH + (xl >> 1);
    T wl[6], wa[6];  // as stored: converted where used
#pragma unroll
    for (int L = 4; L < 10; L++) {
      wl[L - 4] = gplane<T, W>(fb, cur, L, e1).raw(iL);
      wa[L - 4] = gplane<T, W>(fb, cur, L, e2).raw(iA);
    }
    Xoro gL = xoro_seeded(site_key(k1, row * N + xl));
#pragma unroll
    for (int c = 0; c < 6; c++) {
      double tl, ta;
      charge_pair((double)wl[c], (double)wa[c], ck, gL, tl, ta, cnt, dSd, dSu, true);
      // the first sub-sweep's results as stored (rounded in FP32 storage)
      const T tls = (T)tl;
      a1[j][c] = (T)ta;
      if constexpr (H == 1) {
        r[j][c] = __shfl(tls, (lane + 1) & (seg - 1), seg);
      } else {
        r[j][c] = __shfl(tls, (lane + 1) & 63, 64);
        l0[j][c] = first_lane(tls);  // lane 0's value, through scalar registers
      }
    }
  }
#pragma unroll
  for (int j = 0; j < H; j++) {
    const uint32_t h = lane + 64 * j;
    const uint32_t x = site_x(h, y, z, e2);
    const uint32_t xb = x + 1 == N ? 0 : x + 1;
    Xoro gA = xoro_seeded(site_key(k2, row * N + x));
#pragma unroll
    for (int c = 0; c < 6; c++) {
      // H = 2: lane 63's B'_h is lane 0's L of the next j (wrapping to j = 0)
      const double nb = (double)((H > 1 && lane == 63) ? l0[(j + 1) % H][c] : r[j][c]);
      double a2, b2;
      charge_pair((double)a1[j][c], nb, ck, gA, a2, b2, cnt, dSd, dSu, true);
      gplane<T, W>(fb, nxt, 4 + c, e2).st(row * NH + h, a2);
      gplane<T, W>(fb, nxt, 4 + c, e1).st(row * NH + (xb >> 1), b2);
    }
  }
}

// The same two sub-sweeps with every site loaded once and every pair
// computed once (except one first-parity pair per chunk end).  Blocks tile
// the lattice as TXH half-columns x TQ lines x a chunk of C sites (charge_tile):
//  * d = y or z: a thread owns the two columns x = 2xh, 2xh+1 along d and
//    marches d through its chunk.  Along a column the parity alternates; at
//    each step it updates the next first-parity pair (m, m+1) and then the
//    second-parity pair (m-1, m) from the previous pair's upper site, kept in
//    registers, and this pair's lower site.  The chunk owns the
//    second-parity pairs whose lower site lies in it and the first-parity
//    pairs whose lower site lies in it (the monitor counts those).
//  * d = x: pairs lie in rows; the thread runs charge_line for the row
//    (y, z) of each z of its chunk.
template <typename T, bool W, bool PIPE>
__device__ __forceinline__ void charge_column(const ReplicaParams &p, int cur, int nxt, int dir,
                                              int e1, int e2, uint32_t x, uint32_t q, int c0,
                                              int c1, uint64_t k1, uint64_t k2, uint64_t &cnt,
                                              double &dSd, double &dSu) {
  const int N = p.N, NH = p.NH;
  const ChargeK ck = charge_k(p.rms, p.chi);
  const FieldBufs fb = field_bufs<T>(p);
  // index / linear site of position m along the column
  auto at = [&](int m, uint32_t &ii, uint32_t &lin) __attribute__((always_inline)) {
    const uint32_t row = dir == 2 ? (uint32_t)m * N + q : q * N + (uint32_t)m;
    ii = row * NH + (x >> 1);
    lin = row * N + x;
  };
  // lower site of the first first-parity pair: c0-1 or c0
  const int mm0 = (int)((x + q + (uint32_t)(c0 + N - 1)) & 1) == e1 ? c0 - 1 : c0;
  // pair k: lower site m1 (parity e1), upper site m2 = m1 + 1 (parity e2)
  auto pair_at = [&](int k, int &m1, uint32_t &i1, uint32_t &l1, uint32_t &i2, uint32_t &l2)
      __attribute__((always_inline)) {
    const int mm = mm0 + 2 * k;
    m1 = mm < 0 ? mm + N : (mm >= N ? mm - N : mm);
    const int m2 = m1 + 1 == N ? 0 : m1 + 1;
    at(m1, i1, l1);
    at(m2, i2, l2);
  };
  // PIPE: pair k+1's loads are issued before pair k's arithmetic, into n1, n2
  // as stored, and pair k's results are stored in iteration k+1 after those
  // loads (one in-order counter covers loads and stores: a store issued
  // between a load and the wait for it would be waited for as well)
  T n1[6], n2[6], s0[6], s1[6];
  uint32_t si0 = 0, si1 = 0;
  auto load_pair = [&](uint32_t i1, uint32_t i2) __attribute__((always_inline)) {
#pragma unroll
    for (int L = 4; L < 10; L++) {
      n1[L - 4] = gplane<T, W>(fb, cur, L, e1).raw(i1);
      n2[L - 4] = gplane<T, W>(fb, cur, L, e2).raw(i2);
    }
  };
  auto store_pending = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int L = 4; L < 10; L++) {
      gplane<T, W>(fb, nxt, L, e2).template put_raw<false>(si0, s0[L - 4]);
      gplane<T, W>(fb, nxt, L, e1).template put_raw<false>(si1, s1[L - 4]);
    }
  };
  const int npair = (c1 - c0) / 2;
  if (PIPE) {
    int m1;
    uint32_t i1, l1, i2, l2;
    pair_at(0, m1, i1, l1, i2, l2);
    load_pair(i1, i2);
  }
  double up[6];  // upper site of the previous first-parity pair
  for (int k = 0; k <= npair; k++) {
    const int mm = mm0 + 2 * k;
    int m1;
    uint32_t i1, l1, i2, l2;
    pair_at(k, m1, i1, l1, i2, l2);
    const bool own1 = mm >= c0 && mm < c1;
    Xoro g1 = xoro_seeded(site_key(k1, l1));
    Xoro g2;
    uint32_t i0 = 0, l0;
    if (k > 0) {
      const int m0 = m1 == 0 ? N - 1 : m1 - 1;
      at(m0, i0, l0);
      g2 = xoro_seeded(site_key(k2, l0));
    }
    double w1[6], w2[6];
    if (PIPE) {
#pragma unroll
      for (int L = 4; L < 10; L++) {
        w1[L - 4] = (double)n1[L - 4];
        w2[L - 4] = (double)n2[L - 4];
      }
      if (k < npair) {
        int mn;
        uint32_t j1, ln1, j2, ln2;
        pair_at(k + 1, mn, j1, ln1, j2, ln2);
        load_pair(j1, j2);
      }
      if (k > 1) store_pending();
    } else {
      // every load of the pair is issued before the first store: the
      // compiler cannot prove that the stores (other buffer) miss the loaded
      // planes and would otherwise wait for each component's load after the
      // previous store
#pragma unroll
      for (int L = 4; L < 10; L++) {
        w1[L - 4] = gplane<T, W>(fb, cur, L, e1).ld(i1);
        w2[L - 4] = gplane<T, W>(fb, cur, L, e2).ld(i2);
      }
    }
#pragma unroll
    for (int L = 4; L < 10; L++) {
      const double v1 = w1[L - 4];
      const double v2 = w2[L - 4];
      double a, b;
      charge_pair(v1, v2, ck, g1, a, b, cnt, dSd, dSu, own1);
      a = (double)(T)a;
      b = (double)(T)b;
      if (k > 0) {
        double o0, o1;
        charge_pair(up[L - 4], a, ck, g2, o0, o1, cnt, dSd, dSu, true);
        if (PIPE) {
          s0[L - 4] = (T)o0;
          s1[L - 4] = (T)o1;
        } else {
          gplane<T, W>(fb, nxt, L, e2).st(i0, o0);
          gplane<T, W>(fb, nxt, L, e1).st(i1, o1);
        }
      }
      up[L - 4] = b;
    }
    if (PIPE && k > 0) {
      si0 = i0;
      si1 = i1;
    }
  }
  if (PIPE && npair > 0) store_pending();
}

template <typename T, bool W, bool XW>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && XW ? 4 : 1))) void k_charge_march(const ReplicaParams *__restrict__ P,
                                                         const ReplicaState *__restrict__ S,
                                                         const int *__restrict__ blk_rep,
                                                         const int *__restrict__ blk_first,
                                                         int ixyz, int ctr_off, int flip,
                                                         MonSlot *__restrict__ mon) {
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  const ReplicaParams &p = P[r];
  const ReplicaState &s = S[r];
  const int N = p.N, NH = p.NH;
  int TXH, TQ, C;
  charge_tile(N, p.chg_cap, p.chg_chunk, TXH, TQ, C);
  const int nbx = (NH + TXH - 1) / TXH, nbq = (N + TQ - 1) / TQ;
  int b = bid - blk_first[r];
  const int bx = b % nbx;
  b /= nbx;
  const int bq = b % nbq, bc = b / nbq;
  const int t = threadIdx.x;
  const uint32_t xh = bx * TXH + t % TXH, q = bq * TQ + t / TXH;
  const int c0 = bc * C, c1 = min(c0 + C, N);
  uint64_t cnt = 0;
  double dSd = 0., dSu = 0.;
  if (xh < (uint32_t)NH && q < (uint32_t)N && t < TXH * TQ) {
    const int cur = live(s) ^ flip, nxt = cur ^ 2;
    // the shuffled sub-sweep order of ModelGChargeHB::step
    // (Stepper.cxx:461-468) from the sweep's shuffle stream, the counter
    // before the first direction's two; every thread draws the same order
    int order[5];
    {
      Xoro g = xoro_seeded(site_key(stream_key(p.base_key, s.sweep_ctr + ctr_off - 1 - 2 * ixyz), 0));
      shuffle_orders(g, order);
    }
    const int dir = order[ixyz];
    const int e1 = order[3], e2 = order[4];
    const uint64_t k1 = stream_key(p.base_key, s.sweep_ctr + ctr_off);
    const uint64_t k2 = stream_key(p.base_key, s.sweep_ctr + ctr_off + 1);
    if (XW && dir == 0) {  // every replica of the launch has NH = 64 or 128
      // whole rows in one wavefront (TXH = 64, TQ = 4: wave w takes row q);
      // the chunk's planes are split over the nbx x-blocks
      const int lane = t & 63;
      const int len = c1 - c0;
      const int zs = c0 + bx * len / nbx, ze = c0 + (bx + 1) * len / nbx;
      for (int z = zs; z < ze; z++) {
        if (NH == 64)
          charge_row_wave<T, W, 1>(p, cur, nxt, e1, e2, q, (uint32_t)z, lane, 64, k1, k2, cnt,
                                   dSd, dSu);
        else
          charge_row_wave<T, W, 2>(p, cur, nxt, e1, e2, q, (uint32_t)z, lane, 64, k1, k2, cnt,
                                   dSd, dSu);
      }
    } else if (dir == 0 && TXH == NH && NH <= 64 && (NH & (NH - 1)) == 0) {
      // whole rows in aligned segments of NH lanes: with TXH = NH thread t
      // takes half-column t % NH of the tile's row t / NH, and NH divides
      // the wavefront, so a row never straddles two waves
      for (int z = c0; z < c1; z++)
        charge_row_wave<T, W, 1>(p, cur, nxt, e1, e2, q, (uint32_t)z, (int)xh, NH, k1, k2, cnt,
                                 dSd, dSu);
    } else if (dir == 0) {
      for (int z = c0; z < c1; z++)
        charge_line<T, W>(p, cur, nxt, 0, e1, e2, xh, q, (uint32_t)z, k1, k2, cnt, dSd, dSu);
    } else {
#pragma unroll 1
      for (int sx = 0; sx < 2; sx++)
        charge_column<T, W, XW>(p, cur, nxt, dir, e1, e2, 2 * xh + sx, q, c0, c1, k1, k2, cnt,
                                dSd, dSu);
    }
  }
  block_monitor_add(mon + bid, cnt, dSd, dSu);
}

// ------------------------------------------------------------------------
// IdealPV2::step_no_reject (Stepper.cxx:68-176) fused into one pass:
// rotate by dt/2, kick the charges with the rotated neighbours, rotate by
// dt/2 with the kicked charges.  A block owns a TXH x TY x-y tile (both
// parities of each half-column) and marches over TZ planes; the rotated phi
// of the current plane and a one-site x-y halo (recomputed from the raw
// fields) sit in LDS, the rotated phi of the plane above in registers and of
// the plane below in thread-private LDS slots.  Reads the live buffer,
// writes the other one.
template <typename T, bool W>
__device__ __forceinline__ void load_site(const ReplicaParams &p, int cur, int x, int y, int z,
                                          double f[10]) {
  const int par = (x + y + z) & 1;
  const uint32_t i = ((uint32_t)z * p.N + y) * p.NH + (x >> 1);
#pragma unroll
  for (int c = 0; c < 10; c++) f[c] = gplane<T, W>(p, cur, c, par).ld(i);
}

// FP32 fields: at least 3 waves per SIMD (168 registers)
template <typename T, bool W, bool NTS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1)))
void k_ideal_fused(const ReplicaParams *__restrict__ P, const ReplicaState *__restrict__ S,
                   const int *__restrict__ blk_rep, const int *__restrict__ blk_first,
                   int flip) {
  __shared__ double lds[4][ideal_lds(sizeof(T) == 4)];
  // rotated phi of plane z-1 at the thread's own sites: thread-private slots
  __shared__ double lprv[2][4][kBlock];
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  // a reference, not a copy: a private copy would be staged in scratch memory
  const ReplicaParams &p = P[r];
  const int cur = live(S[r]) ^ flip, nxt = cur ^ 3;
  const int N = p.N, NH = p.NH;
  int TXH, TY, TZ;
  ideal_tile(N, p.tile_cap, p.tile_tz, TXH, TY, TZ);
  const int ntx = (NH + TXH - 1) / TXH, nty = (N + TY - 1) / TY;
  int b = bid - blk_first[r];
  int bx, by, bz;
  if ((int)gridDim.x == ntx * nty * ((N + TZ - 1) / TZ)) {
    yband_tile(b, ntx, nty, bx, by, bz);  // one replica
  } else {
    bx = b % ntx;
    b /= ntx;
    by = b % nty;
    bz = b / nty;
  }
  const int xh0 = bx * TXH, y0 = by * TY, z0 = bz * TZ;
  const int tw = min(TXH, NH - xh0), th = min(TY, N - y0), tz = min(TZ, N - z0);
  const int LW = 2 * tw + 2;  // LDS row width
  const int t = threadIdx.x;
  const int tx = t % TXH, ty = t / TXH;
  const bool own = ty < th && tx < tw;
  const int xh = xh0 + tx, y = y0 + ty;
  // the one-site ring around the tile (no corners): ring site j = t, t + kBlock, ..
  // (a second pass only when the ring has more sites than the block threads)
  const int nh = 2 * th + 4 * tw;
  // ring site j: global (x, y) and LDS slot
  auto ring_site = [&](int j, int &gx, int &gy, int &pos) __attribute__((always_inline)) {
    int hlx, hly;
    if (j < th) { hlx = -1; hly = j; }
    else if (j < 2 * th) { hlx = 2 * tw; hly = j - th; }
    else if (j < 2 * th + 2 * tw) { hlx = j - 2 * th; hly = -1; }
    else { hlx = j - 2 * th - 2 * tw; hly = th; }
    gx = 2 * xh0 + hlx;
    gx = gx < 0 ? gx + N : (gx >= N ? gx - N : gx);
    gy = y0 + hly;
    gy = gy < 0 ? gy + N : (gy >= N ? gy - N : gy);
    pos = (hly + 1) * LW + (hlx + 1);
  };
  int rgx = 0, rgy = 0, rpos = 0;  // this thread's first ring site
  if (t < nh) ring_site(t, rgx, rgy, rpos);
  const double h = 0.5 * (p.tau_ideal / p.chi);
  const double axx = p.axx, H0 = p.H, dt = p.dt_ideal;

  // FP64 fields carry the raw charges of planes z and z+1 in registers (two
  // waves per SIMD).  FP32 fields keep those of plane z in thread-private
  // LDS slots (as stored, in FP32) and those of plane z+1 in registers from
  // their load to the end of the iteration, which keeps three waves per
  // SIMD; re-reading plane z's charges at kick time instead made the kernel
  // read 1.92x its algorithmic bytes (the re-reads miss the 4 MB L2).
  constexpr bool kCarry = sizeof(T) == 8;
  __shared__ T chq[kCarry ? 1 : 2][kCarry ? 1 : 6][kCarry ? 1 : kBlock];
  // The raw fields of the thread's first ring site of plane z+1 are loaded
  // before the kick of plane z, so that their latency overlaps the kick
  // (FP32: ten registers more, which the waves-per-SIMD bound of 3 holds at
  // 168 with three spilled; 256^3 FP32 347 -> 329 us, 2 waves 360 us)
  constexpr bool kEarlyRing = true;
  // The thread's two own sites of a plane are indexed by parity (pp), not by
  // x offset: a parity's plane base is the same for every lane, so the loads
  // and stores use one scalar base and 32-bit offsets.  The site of parity pp
  // sits at x = 2 xh + ((pp + y + z) & 1); its z +- 1 neighbours have parity
  // pp ^ 1 in the same half-column.
  double cen[2][4], nex[2][4];  // rotated phi at the own sites of planes z, z+1
  double chg[2][6];             // raw charges of plane z (kCarry only)
  T chn[2][6];                  // raw charges of plane z+1, as stored
  auto rotate_own = [&](int z, double out[2][4], T ch[2][6]) __attribute__((always_inline)) {
    if (!own) return;
    const uint32_t i = ((uint32_t)z * N + y) * NH + xh;
    // both sites' loads before the first rotation: one wait for the pair
    T raw[2][10];
#pragma unroll
    for (int pp = 0; pp < 2; pp++)
#pragma unroll
      for (int c = 0; c < 10; c++) raw[pp][c] = gplane<T, W>(p, cur, c, pp).raw(i);
#pragma unroll
    for (int pp = 0; pp < 2; pp++) {
      double f[10];
#pragma unroll
      for (int c = 0; c < 10; c++) f[c] = (double)raw[pp][c];
      rotate_site_fast(f, f + 4, f + 7, h);
#pragma unroll
      for (int c = 0; c < 4; c++) out[pp][c] = f[c];
#pragma unroll
      for (int c = 0; c < 6; c++) ch[pp][c] = (T)f[4 + c];
    }
  };
  auto fill_lds = [&](int z, const double o[2][4]) __attribute__((always_inline)) {
    if (own) {
#pragma unroll
      for (int pp = 0; pp < 2; pp++) {
        const int sx = (pp + y + z) & 1;
#pragma unroll
        for (int c = 0; c < 4; c++) lds[c][(ty + 1) * LW + (2 * tx + sx + 1)] = o[pp][c];
      }
    }
    auto ring_fill = [&](int gx, int gy, int pos) __attribute__((always_inline)) {
      double f[10];
      load_site<T, W>(p, cur, gx, gy, z, f);
      rotate_site_fast(f, f + 4, f + 7, h);
#pragma unroll
      for (int c = 0; c < 4; c++) lds[c][pos] = f[c];
    };
    if (!kEarlyRing && t < nh) ring_fill(rgx, rgy, rpos);
    for (int j = t + kBlock; j < nh; j += kBlock) {
      int gx, gy, pos;
      ring_site(j, gx, gy, pos);
      ring_fill(gx, gy, pos);
    }
  };

  rotate_own((z0 - 1 + N) % N, nex, chn);
  if (own) {
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int c = 0; c < 4; c++) lprv[s][c][t] = nex[s][c];
  }
  rotate_own(z0, cen, chn);
  // the charges of plane z0: registers (FP64) or the thread's LDS slots (FP32)
  if (own) {
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int c = 0; c < 6; c++) {
        if constexpr (kCarry) chg[s][c] = (double)chn[s][c];
        else chq[s][c][t] = chn[s][c];
      }
  }
  if (kEarlyRing && t < nh) {
    double f[10];
    load_site<T, W>(p, cur, rgx, rgy, z0, f);
    rotate_site_fast(f, f + 4, f + 7, h);
#pragma unroll
    for (int c = 0; c < 4; c++) lds[c][rpos] = f[c];
  }
  fill_lds(z0, cen);
  for (int k = 0; k < tz; k++) {
    const int z = z0 + k;
    const int zn = (z + 1) % N;
    rotate_own(zn, nex, chn);
    T rraw[10];  // the ring site's raw fields of plane z+1, as stored
    const bool ring1 = kEarlyRing && t < nh && k + 1 < tz;
    if (ring1) {
      const uint32_t ri = ((uint32_t)zn * N + rgy) * NH + (rgx >> 1);
      const int rpar = (rgx + rgy + zn) & 1;
#pragma unroll
      for (int c = 0; c < 10; c++) rraw[c] = gplane<T, W>(p, cur, c, rpar).raw(ri);
    }
    __syncthreads();
    if (own) {
      const uint32_t i = ((uint32_t)z * N + y) * NH + xh;
#pragma unroll
      for (int pp = 0; pp < 2; pp++) {
        // held to 1e-12 like the rotations: contraction into FMA allowed
#pragma clang fp contract(fast)
        const int lx = 2 * tx + ((pp + y + z) & 1) + 1, ly = ty + 1;
        // The kick of Stepper.cxx:112-163, summed per direction there, is
        // linear in the neighbours: with S the sum of the six neighbours'
        // rotated phi, the axial kick is axx (cp0 S_l - cp_l S0) - H cp_l
        // and the vector kick axx (cp_s1 S_s2 - cp_s2 S_s1), (s1, s2) the
        // cyclic successors of s (every direction shares axx).
        const double *zp = nex[pp ^ 1];
        double cp[4], S[4], A[3], V[3];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          cp[c] = cen[pp][c];
          S[c] = (lds[c][ly * LW + lx + 1] + lds[c][ly * LW + lx - 1]) +
                 (lds[c][(ly + 1) * LW + lx] + lds[c][(ly - 1) * LW + lx]) +
                 (zp[c] + lprv[pp ^ 1][c][t]);
        }
        const int par = pp;
        // the charges of plane z, unchanged by the first rotation
#pragma unroll
        for (int l = 0; l < 3; l++) {
          if constexpr (kCarry) {
            A[l] = chg[pp][l];
            V[l] = chg[pp][3 + l];
          } else {
            A[l] = (double)chq[pp][l][t];
            V[l] = (double)chq[pp][3 + l][t];
          }
        }
#pragma unroll
        for (int l = 0; l < 3; l++)
          A[l] += dt * (axx * (cp[0] * S[l + 1] - cp[l + 1] * S[0]) - H0 * cp[l + 1]);
#pragma unroll
        for (int s3 = 0; s3 < 3; s3++) {
          const int s1 = (s3 + 1) % 3, s2 = (s3 + 2) % 3;
          V[s3] += dt * (axx * (cp[1 + s1] * S[1 + s2] - cp[1 + s2] * S[1 + s1]));
        }
        rotate_site_fast(cp, A, V, h);
#pragma unroll
        for (int c = 0; c < 4; c++) gplane<T, W>(p, nxt, c, par).template put<NTS>(i, cp[c]);
#pragma unroll
        for (int c = 0; c < 3; c++) gplane<T, W>(p, nxt, 4 + c, par).template put<NTS>(i, A[c]);
#pragma unroll
        for (int c = 0; c < 3; c++) gplane<T, W>(p, nxt, 7 + c, par).template put<NTS>(i, V[c]);
      }
    }
    __syncthreads();
    if (ring1) {
      double rf[10];
#pragma unroll
      for (int c = 0; c < 10; c++) rf[c] = (double)rraw[c];
      rotate_site_fast(rf, rf + 4, rf + 7, h);
#pragma unroll
      for (int c = 0; c < 4; c++) lds[c][rpos] = rf[c];
    }
    if (k + 1 < tz) fill_lds(zn, nex);
#pragma unroll
    for (int s = 0; s < 2; s++) {
#pragma unroll
      for (int c = 0; c < 4; c++) {
        if (own) lprv[s][c][t] = cen[s][c];
        cen[s][c] = nex[s][c];
      }
      if (own) {
#pragma unroll
        for (int c = 0; c < 6; c++) {
          if constexpr (kCarry) chg[s][c] = (double)chn[s][c];
          else chq[s][c][t] = chn[s][c];
        }
      }
    }
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
                                                   int flip, EnergySlot *__restrict__ slots) {
  const int bid = xcd_block();
  const int r = blk_rep[bid];
  const ReplicaParams p = P[r];
  const int cur = live(S[r]) ^ flip;
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
// The end of a step (or of a single sweep): the clock, the stream counter
// and the live buffers.  The kernels of a step take the buffer flips of the
// launches before them as an argument (flip: bit 0 phi, bit 1 charges, XORed
// into the state's live bits) and this kernel applies the step's net flip.
__global__ void k_advance(ReplicaState *S, int nrep, double dt, int advance_time,
                          uint64_t nctr, int flip) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  if (advance_time) S[r].t += dt;
  S[r].sweep_ctr += nctr;
  S[r].cur_phi ^= flip & 1;
  S[r].cur_chg ^= (flip >> 1) & 1;
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
                            int z0, int nz, const double *aos) {
  const ReplicaParams &p = P[replica];
  const int cur = live(S[replica]);
  const int N = p.N;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // within the slab
  if (s >= (int64_t)nz * N * N) return;
  const int x = (int)(s % N);
  const int64_t yz = s / N + (int64_t)z0 * N;
  const int y = (int)(yz % N), z = (int)(yz / N);
  const int par = (x + y + z) & 1;
  const int64_t i = yz * p.NH + (x >> 1);
  for (int c = 0; c < 10; c++) plane<T>(p, cur, c, par)[i] = (T)aos[s * 10 + c];
}

template <typename T>
__global__ void k_rb_to_aos(const ReplicaParams *P, const ReplicaState *S, int replica,
                            int z0, int nz, double *aos) {
  const ReplicaParams &p = P[replica];
  const int cur = live(S[replica]);
  const int N = p.N;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // within the slab
  if (s >= (int64_t)nz * N * N) return;
  const int x = (int)(s % N);
  const int64_t yz = s / N + (int64_t)z0 * N;
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
__device__ __forceinline__ void site_obs_add(const ReplicaParams &p, int cur, int par, int64_t i,
                                             KSum acc[25]) {
  double f[10];
  for (int c = 0; c < 10; c++) f[c] = (double)plane<T>(p, cur, c, par)[i];
  double phi2 = 0.0;
  for (int l = 0; l < 4; l++) phi2 += f[l] * f[l];
  for (int c = 0; c < 10; c++) ksum_add(acc[c], f[c]);
  ksum_add(acc[10], phi2);
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
  ksum_add(acc[11], sigma);
  for (int l = 0; l < 4; l++) {
    ksum_add(acc[12 + l], n[l]);
    ksum_add(acc[16 + l], A[l]);
    ksum_add(acc[20 + l], V[l]);
  }
  ksum_add(acc[24], sigma * sigma);
}

// Per (x, z): compensated sum over y of the 25 site observables ->
// partXZ[ch][o][z][x]; per (x, y): sum over z -> partXY[ch][o][y][x].  The
// summed axis is split into nch chunks (measure_chunks) so that a 256^3
// lattice runs 8 N^2 threads instead of N^2; k_measure_fold adds the chunks.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_measure_cols(const ReplicaParams *P,
                                                         const ReplicaState *S, int replica,
                                                         int along_z, int nch, double *part) {
  const ReplicaParams &p = P[replica];
  const int cur = live(S[replica]);
  const int N = p.N;
  const int64_t N2 = (int64_t)N * N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N2 * nch) return;
  const int ch = (int)(t / N2);
  const int64_t col = t - (int64_t)ch * N2;
  const int x = (int)(col % N), w = (int)(col / N);
  const int v0 = (int)((int64_t)ch * N / nch), v1 = (int)((int64_t)(ch + 1) * N / nch);
  KSum acc[25];
  for (int o = 0; o < 25; o++) acc[o] = KSum{0., 0.};
  for (int v = v0; v < v1; v++) {
    const int y = along_z ? w : v;
    const int z = along_z ? v : w;
    const int par = (x + y + z) & 1;
    const int64_t i = ((int64_t)z * N + y) * p.NH + (x >> 1);
    site_obs_add<T>(p, cur, par, i, acc);
  }
  double *out = part + (int64_t)ch * 25 * N2;
  for (int k = 0; k < 25; k++) out[((int64_t)k * N + w) * N + x] = ksum_value(acc[k]);
}

// Folds the chunk partials into chunk 0: part[0][j] = compensated sum over
// ch of part[ch][j], j < 25 N^2 (one thread per element, coalesced).
__global__ void k_measure_fold(int64_t n, int nch, double *part) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  KSum acc{0., 0.};
  for (int ch = 0; ch < nch; ch++) ksum_add(acc, part[(int64_t)ch * n + j]);
  part[j] = ksum_value(acc);
}

// Walls, x direction: wall_x[k][n] = (compensated sum over z of
// partXZ[k][z][n]) / N^2.  A block takes 64 consecutive n (coalesced) of one
// k; its 8 thread groups sum z = g, g + 8, ... and the 8 partials are merged
// in group order.
__global__ __launch_bounds__(512) void k_measure_reduce_x(int N, const double *partXZ,
                                                          double *wall, double *wallphase) {
  __shared__ KSum part[8][64];
  const int nb = (N + 63) / 64;
  const int k = blockIdx.x / nb;
  const int n = (blockIdx.x % nb) * 64 + (threadIdx.x & 63);
  const int g = threadIdx.x >> 6;
  KSum acc{0., 0.};
  if (n < N)
    for (int m = g; m < N; m += 8) ksum_add(acc, partXZ[((int64_t)k * N + m) * N + n]);
  part[g][threadIdx.x & 63] = acc;
  __syncthreads();
  if (g == 0 && n < N) {
    for (int i = 1; i < 8; i++) ksum_merge(acc, part[i][threadIdx.x]);
    const double val = ksum_value(acc) / (double)(N * N);
    if (k < 11) wall[(int64_t)k * N + n] = val;
    else wallphase[(int64_t)(k - 11) * N + n] = val;
  }
}

// Walls, y and z directions: row sums, one wave per (d, k, n) row (lanes
// read consecutive x; lane partials merged by a fixed shuffle tree).
//   d = 1: wall_y[k][n] = sum_x partXY[k][n][x] / N^2
//   d = 2: wall_z[k][n] = sum_x partXZ[k][n][x] / N^2
__global__ __launch_bounds__(256) void k_measure_reduce_rows(int N, const double *partXZ,
                                                             const double *partXY,
                                                             double *wall, double *wallphase) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= 2 * 25 * N) return;  // whole waves exit together
  const int n = row % N, k = (row / N) % 25, d = 1 + row / (25 * N);
  const double *src = (d == 1 ? partXY : partXZ) + ((int64_t)k * N + n) * N;
  KSum acc{0., 0.};
  for (int x = lane; x < N; x += 64) ksum_add(acc, src[x]);
  for (int o = 32; o > 0; o >>= 1) {
    KSum t;
    t.s = __shfl_xor(acc.s, o, 64);
    t.c = __shfl_xor(acc.c, o, 64);
    ksum_merge(acc, t);
  }
  if (lane == 0) {
    const double val = ksum_value(acc) / (double)(N * N);
    if (k < 11) wall[((int64_t)d * 11 + k) * N + n] = val;
    else wallphase[((int64_t)d * 14 + (k - 11)) * N + n] = val;
  }
}

// measurer_fft::execute_row (measurer.h:33-41): FFTW's r2c forward
// transform divided by N, out[k] = (1/N) sum_x in[x] e^{-2 pi i k x / N},
// k = 0..N/2, as a direct DFT with exactly reduced twiddles and compensated
// sums.  One thread per (row, k); rows = 3 x 25 walls.
__global__ __launch_bounds__(256) void k_wall_dft(int N, const double *wall,
                                                  const double *wallphase, double *wall_k,
                                                  double *wallphase_k) {
  // twiddles cos/sin(2 pi m / N), m = 0 .. N-1, once per block
  __shared__ double tc[kMaxN], ts[kMaxN];
  for (int m = threadIdx.x; m < N; m += blockDim.x) sincospi(2.0 * m / N, &ts[m], &tc[m]);
  __syncthreads();
  const int nk = N / 2 + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * 25 * nk) return;
  const int k = t % nk;
  const int row = t / nk;  // d*25 + obs
  const int d = row / 25, o = row % 25;
  const double *in = o < 11 ? wall + ((int64_t)d * 11 + o) * N
                            : wallphase + ((int64_t)d * 14 + (o - 11)) * N;
  KSum re{0., 0.}, im{0., 0.};
  int m = 0;  // k x mod N
  for (int x = 0; x < N; x++) {
    ksum_add(re, in[x] * tc[m]);
    ksum_add(im, -(in[x] * ts[m]));
    m += k;
    if (m >= N) m -= N;
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
  k_wall_dft<<<(n + 255) / 256, 256, 0, st>>>(N, wall, wallphase, wall_k, wallphase_k);
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// W: some plane of the launch is 4 GiB or larger (64-bit element offsets).
// The same over the store mode as well (plan.stream).
#define LGV_DISPATCH_S(KERNEL, GRID, ST, ...)                                             \
  do {                                                                                    \
    if (plan.stream) {                                                                    \
      if (fp32) {                                                                         \
        if (plan.wide) KERNEL<float, true, true><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);   \
        else KERNEL<float, false, true><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);            \
      } else {                                                                            \
        if (plan.wide) KERNEL<double, true, true><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);  \
        else KERNEL<double, false, true><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);           \
      }                                                                                   \
    } else {                                                                              \
      if (fp32) {                                                                         \
        if (plan.wide) KERNEL<float, true, false><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);  \
        else KERNEL<float, false, false><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);           \
      } else {                                                                            \
        if (plan.wide) KERNEL<double, true, false><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__); \
        else KERNEL<double, false, false><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);          \
      }                                                                                   \
    }                                                                                     \
  } while (0)

#define LGV_DISPATCH(KERNEL, GRID, ST, ...)                                           \
  do {                                                                                \
    if (fp32) {                                                                       \
      if (plan.wide) KERNEL<float, true><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);       \
      else KERNEL<float, false><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);                \
    } else {                                                                          \
      if (plan.wide) KERNEL<double, true><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);      \
      else KERNEL<double, false><<<GRID, kBlock, 0, ST>>>(__VA_ARGS__);               \
    }                                                                                 \
  } while (0)

void launch_hb_phi(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                   int ieo, int ctr_off, int flip, MonSlot *mon, hipStream_t st) {
  LGV_DISPATCH(k_hb_phi, plan.nblocks, st, P, S, plan.d_blk_rep, plan.d_blk_first, ieo,
               ctr_off, flip, mon);
}

void launch_hb_march(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                     int ieo, int ctr_off, int flip, MonSlot *mon, hipStream_t st) {
  LGV_DISPATCH_S(k_hb_march, plan.nblocks, st, P, S, plan.d_blk_rep, plan.d_blk_first, ieo,
               ctr_off, flip, mon);
}

void launch_charge_march(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                         int ixyz, int ctr_off, int flip, MonSlot *mon, hipStream_t st) {
  // the row-per-wavefront x path costs registers (3 waves per SIMD instead of
  // 4), so it is compiled in only for launches whose replicas all use it
  const int *br = plan.d_blk_rep, *bf = plan.d_blk_first;
  const int g = plan.nblocks;
#define LGV_CHARGE(TT, WW, XX)                                                         \
  k_charge_march<TT, WW, XX><<<g, kBlock, 0, st>>>(P, S, br, bf, ixyz, ctr_off, flip, mon)
  if (plan.xwave) {
    if (fp32) { if (plan.wide) LGV_CHARGE(float, true, true); else LGV_CHARGE(float, false, true); }
    else { if (plan.wide) LGV_CHARGE(double, true, true); else LGV_CHARGE(double, false, true); }
  } else {
    if (fp32) { if (plan.wide) LGV_CHARGE(float, true, false); else LGV_CHARGE(float, false, false); }
    else { if (plan.wide) LGV_CHARGE(double, true, false); else LGV_CHARGE(double, false, false); }
  }
#undef LGV_CHARGE
}

void launch_ideal(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                  int flip, hipStream_t st) {
  LGV_DISPATCH_S(k_ideal_fused, plan.nblocks, st, P, S, plan.d_blk_rep, plan.d_blk_first, flip);
}

void launch_energy(bool fp32, const Plan &plan, const ReplicaParams *P,
                   const ReplicaState *S, int flip, EnergySlot *slots, hipStream_t st) {
  if (fp32)
    k_energy<float><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep, plan.d_blk_first,
                                                     flip, slots);
  else
    k_energy<double><<<plan.nblocks, kBlock, 0, st>>>(P, S, plan.d_blk_rep, plan.d_blk_first,
                                                      flip, slots);
}

void launch_energy_finish(int nrep, const Plan &plan, const ReplicaParams *P,
                          ReplicaState *S, const EnergySlot *slots, int which,
                          int ctr_off, MonSlot *mon_ideal, hipStream_t st) {
  k_energy_finish<<<nrep, 64, 0, st>>>(P, S, plan.d_blk_first, nrep, plan.nblocks, slots,
                                       which, ctr_off, mon_ideal);
}

void launch_advance(int nrep, ReplicaState *S, double dt, int advance_time, uint64_t nctr,
                    int flip, hipStream_t st) {
  k_advance<<<nblk(nrep, 64), 64, 0, st>>>(S, nrep, dt, advance_time, nctr, flip);
}

void launch_set_time(int nrep, int replica, ReplicaState *S, double t, hipStream_t st) {
  k_set_time<<<nblk(nrep, 64), 64, 0, st>>>(S, nrep, replica, t);
}

void launch_aos_to_rb(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                      int N, int z0, int nz, const double *aos, hipStream_t st) {
  const int64_t V = (int64_t)nz * N * N;
  if (fp32) k_aos_to_rb<float><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, z0, nz, aos);
  else k_aos_to_rb<double><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, z0, nz, aos);
}

void launch_rb_to_aos(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                      int N, int z0, int nz, double *aos, hipStream_t st) {
  const int64_t V = (int64_t)nz * N * N;
  if (fp32) k_rb_to_aos<float><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, z0, nz, aos);
  else k_rb_to_aos<double><<<nblk(V, 256), 256, 0, st>>>(P, S, replica, z0, nz, aos);
}

void launch_measure(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                    int N, double *partXZ, double *partXY, double *wall, double *wallphase,
                    hipStream_t st) {
  const int nch = measure_chunks(N);
  const int64_t threads = (int64_t)N * N * nch;
  if (fp32) {
    k_measure_cols<float><<<nblk(threads, kBlock), kBlock, 0, st>>>(P, S, replica, 0, nch, partXZ);
    k_measure_cols<float><<<nblk(threads, kBlock), kBlock, 0, st>>>(P, S, replica, 1, nch, partXY);
  } else {
    k_measure_cols<double><<<nblk(threads, kBlock), kBlock, 0, st>>>(P, S, replica, 0, nch, partXZ);
    k_measure_cols<double><<<nblk(threads, kBlock), kBlock, 0, st>>>(P, S, replica, 1, nch, partXY);
  }
  if (nch > 1) {
    const int64_t n = (int64_t)25 * N * N;
    k_measure_fold<<<nblk(n, 256), 256, 0, st>>>(n, nch, partXZ);
    k_measure_fold<<<nblk(n, 256), 256, 0, st>>>(n, nch, partXY);
  }
  k_measure_reduce_x<<<25 * ((N + 63) / 64), 512, 0, st>>>(N, partXZ, wall, wallphase);
  k_measure_reduce_rows<<<nblk(2 * 25 * N, 4), 256, 0, st>>>(N, partXZ, partXY, wall,
                                                             wallphase);
}

}  // namespace lgv
