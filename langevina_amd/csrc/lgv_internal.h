// lgv_internal.h -- structures shared by the kernels (lgv_kernels.hip) and
// the context / C-ABI layer (lgv_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace lgv {

constexpr int kBlock = 256;  // threads per block of every sweep kernel

// Read-only (for the kernels) description of one replica.  The step
// coefficients (rdtg, rms, dt_ideal, tau) are computed on the host with the
// reference's expressions for the dt in use.
struct ReplicaParams {
  void *fields[2];    // two buffers of 10 planes x 2 parities x half elements
                      // (T = double/float); ReplicaState::cur_phi / cur_chg
                      // select where the live phi / charge planes are, the
                      // out-of-place kernels write the other buffer
  int N, NH;
  int64_t half;       // N^3 / 2
  double mass0, dmassdt, lambda, H, chi;
  double axx;         // pow(1/h, 2), Stepper.cxx:91-93
  double rdtg;        // sqrt(2 dt_hb gamma), Stepper.cxx:301
  double rms;         // sqrt(2 dt sigma), Stepper.cxx:456
  double dt_ideal;    // dt / scounts[1]
  double tau_ideal;   // dt_ideal / 2.0
  uint64_t base_key;
  int tile_cap, tile_tz;   // ideal-kernel tile (see ideal_tile)
  int hb_cap, hb_tz;       // phi-march tile (see hb_tile; 0 = default)
  int chg_cap, chg_chunk;  // charge-march tile (see charge_tile; 0 = default)
};

// Device-written state of one replica (clock, counters, per-step data).
struct ReplicaState {
  double t;
  uint64_t sweep_ctr;
  int cur_phi;       // buffer (0/1) holding the live phi planes
  int cur_chg;       // buffer (0/1) holding the live charge planes
  int pad;
  double E_old, E_new;
};

// Tile of the fused ideal kernel: TXH half-columns x TY rows in x-y,
// marching over TZ planes in z.  txh_cap (8..64) and tz (4..64) are the
// tunables (0 = defaults 32 and 16); (TY+2)(2 TXH+2) <= kIdealLds.  The
// one-site halo ring (2 TY + 4 TXH sites) may exceed the block's threads
// (64 half-columns x 4 rows: 264); its remainder then takes a second pass.
constexpr int kIdealLds = 800;
// FP32 fields keep the raw charges of the current plane in LDS as well, so
// their plane grid is held to the tiles of width <= 32 half-columns (at most
// 660 entries) and three blocks still fit one CU's LDS.
constexpr int kIdealLdsF32 = 660;
__host__ __device__ constexpr int ideal_lds(bool fp32) { return fp32 ? kIdealLdsF32 : kIdealLds; }
__host__ __device__ inline void ideal_tile(int N, int txh_cap, int tz, int &TXH, int &TY,
                                           int &TZ) {
  const int NH = N / 2;
  const int cap = txh_cap > 0 ? txh_cap : 32;
  const int zc = tz > 0 ? tz : 16;
  TXH = NH < cap ? NH : cap;
  TY = kBlock / TXH;
  if (TY > N) TY = N;
  TZ = N < zc ? N : zc;
}

// Tile of the single-parity phi march: TXH half-columns x TY rows in x-y,
// chunks of TZ planes in z.  Defaults (cap, tz = 0): 64 x 4 x 16 from
// N = 128 on (wide rows keep the x-halo columns a small share of the lines
// each plane touches); cap and tz are tunables, valid while the packed grid
// fits (TY+2)(TXH+2) <= kHbMarchGrid and the ring 2 TY + 2 TXH <= kBlock.
__host__ __device__ inline void hb_tile(int N, int cap, int tz, int &TXH, int &TY, int &TZ) {
  const int NH = N / 2;
  const int c = cap > 0 ? cap : 64;
  TXH = NH < c ? NH : c;
  TY = 256 / TXH;
  if (TY > N) TY = N;
  const int zc = tz > 0 ? tz : 16;
  TZ = N < zc ? N : zc;
}

// Its packed LDS grid: (TY+2)(TXH+2) entries.
constexpr int kHbMarchGrid = 400;

// Tile of the charge march kernel: TXH half-columns x TQ lines of the
// cross-section, chunks of C sites along the direction (C even).  Defaults
// (cap, chunk = 0): TXH = min(NH, 64) and chunks of 16 from N = 128 on, 8
// below, where the shorter chunks give small lattices (and batches of them)
// enough blocks to fill the GPU (measured: 96-replica batch of 32^3-64^3,
// charge 268 -> 252 us; at 256^3 16 is faster).  cap and chunk are tunables.
__host__ __device__ inline void charge_tile(int N, int cap, int chunk, int &TXH, int &TQ,
                                            int &C) {
  const int NH = N / 2;
  const int c = cap > 0 ? cap : 64;
  TXH = NH < c ? NH : c;
  TQ = 256 / TXH;
  if (TQ > N) TQ = N;
  const int ch = chunk > 0 ? chunk : (N >= 128 ? 16 : 8);
  C = N < ch ? N : ch;
}

// Per-block monitor partials (o4_stepper_monitor counters).
struct MonSlot {
  long long down_counts, up_counts, up_yes, up_no;
  double down_dS, up_dS;
};

struct EnergySlot {
  double g_s, g_c, h_s, h_c, n_s, n_c;
};

// A launch over several replicas: block b belongs to replica blk_rep[b] and
// is its (b - blk_first[replica])-th block.
struct Plan {
  int nblocks = 0;
  int *d_blk_rep = nullptr;
  int *d_blk_first = nullptr;
  bool wide = false;    // some plane of the launch is >= 4 GiB (64-bit offsets)
  bool stream = false;  // outputs written with non-temporal stores (see kStreamBytes)
  bool xwave = false;   // charge launch: every replica has N = 128 or 256 (x rows per wave)
  std::vector<int> first;  // host copy of d_blk_first
};

// The phi march and the ideal kernel write their outputs with non-temporal
// stores when one buffer's fields (all replicas) exceed this: the outputs
// are next read by a later launch, after the whole lattice has been swept,
// so caching them only evicts the lines that neighbouring tiles still read.
// Measured at 256^3 FP64 (1.3 GB): ideal 602 -> 586 us, phi 175 -> 168 us;
// at 64^3-128^3 and for FP32 256^3 (0.7 GB) the plain stores are as fast or
// faster.
constexpr uint64_t kStreamBytes = 1ull << 30;

// FP32 fields above this (one buffer, all replicas) also take the phi
// z-march: its own launch is no faster than the one-thread-per-site kernel
// (256^3: 110 against 109 us), but it moves 1.09x instead of 1.36x the
// algorithmic bytes and leaves more of the charge and phi planes in the
// 256 MB last-level cache for the kernels that follow (256^3 FP32 step
// 3.72 -> 3.61 ms: charge 337 -> 317 us, ideal 429 -> 414 us).  Below it
// the size rule takes the per-site kernel (single 128^3-192^3 lattices run
// faster with it); the autotuner may choose the march instead (the
// 96-replica FP32 batch does).  A context of one lattice above this size
// also keeps the charge sweep after the phi sweeps rather than beside them.
constexpr uint64_t kMarchBytesFp32 = 600ull << 20;

// Largest side: N^3 must fit 32-bit site indices.
constexpr int kMaxN = 1536;

// Chunks of the measurement's column sums: about 2^19 column-chunks per
// pass (8 chunks at N = 256, 1 from N = 724 on).  The partial buffers hold
// 25 * measure_chunks(N) * N^2 <= 25 * max(2^19, N^2) doubles.
constexpr int64_t kMeasureColumns = int64_t(1) << 19;
__host__ __device__ inline int measure_chunks(int N) {
  const int64_t n2 = (int64_t)N * N;
  int c = 1;
  while (c < 8 && 2 * c * n2 <= kMeasureColumns) c *= 2;
  return c;
}

// Kernel launchers (lgv_kernels.hip).  fp32 selects the storage type.
// flip: the live-buffer flips of the earlier launches of the step, not yet
// in the replica state (bit 0 phi, bit 1 charges); launch_advance applies a
// step's net flip.
void launch_hb_phi(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                   int ieo, int ctr_off, int flip, MonSlot *mon, hipStream_t st);
void launch_hb_march(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                     int ieo, int ctr_off, int flip, MonSlot *mon, hipStream_t st);
void launch_charge_march(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                         int ixyz, int ctr_off, int flip, MonSlot *mon, hipStream_t st);
void launch_ideal(bool fp32, const Plan &plan, const ReplicaParams *P, ReplicaState *S,
                  int flip, hipStream_t st);
void launch_energy(bool fp32, const Plan &plan, const ReplicaParams *P,
                   const ReplicaState *S, int flip, EnergySlot *slots, hipStream_t st);
void launch_energy_finish(int nrep, const Plan &plan, const ReplicaParams *P,
                          ReplicaState *S, const EnergySlot *slots, int which,
                          int ctr_off, MonSlot *mon_ideal, hipStream_t st);
void launch_advance(int nrep, ReplicaState *S, double dt, int advance_time,
                    uint64_t nctr, int flip, hipStream_t st);
void launch_set_time(int nrep, int replica, ReplicaState *S, double t, hipStream_t st);
// Field transfers go through a staging buffer of at most kStageBytes, in
// slabs of whole z-planes.
constexpr int64_t kStageBytes = int64_t(2) << 30;
__host__ __device__ inline int stage_planes(int N) {
  const int64_t plane = (int64_t)N * N * 10 * (int64_t)sizeof(double);
  const int64_t n = kStageBytes / plane;
  return n < 1 ? 1 : (n > N ? N : (int)n);
}
void launch_aos_to_rb(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                      int N, int z0, int nz, const double *aos, hipStream_t st);
void launch_rb_to_aos(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                      int N, int z0, int nz, double *aos, hipStream_t st);
void launch_wall_dft(int N, const double *wall, const double *wallphase, double *wall_k,
                     double *wallphase_k, hipStream_t st);
void launch_measure(bool fp32, const ReplicaParams *P, const ReplicaState *S, int replica,
                    int N, double *partXZ, double *partXY, double *wall, double *wallphase,
                    hipStream_t st);

}  // namespace lgv
