/*
 * lgv.h -- C ABI of the MI355X ModelG hot path (liblgv_hip.so).
 *
 * Plain pointers and sizes only; every call returns LGV_OK (0) or a negative
 * status, with a message in lgv_last_error().  No exception crosses this
 * boundary.  One context drives one GPU and holds one or more independent
 * replicas (lattices with their own size, seed and couplings) that are
 * stepped together.
 *
 * Reference interface replaced (paths relative to the reference tree):
 *   Stepper::step / finalize                    ModelG/Stepper.h:6-11
 *   PV2HBSplit (scounts, nodiffuse, onlydiffuse) ModelG/Stepper.h:176-211
 *   IdealPV2 / EulerLangevinHB / ModelGChargeHB ModelG/Stepper.h:87-154
 *   ModelAData / ModelACoefficients (params)    ModelG/ModelA.h:71-111,185-269
 *   ModelA::solution (PETSc Vec of G_node)      ModelG/ModelA.h:274-336
 *   ModelA::initialize_gaussian_charges         ModelG/ModelA.h:463-525
 *   NoiseGenerator / ModelARndm                 ModelG/NoiseGenerator.h:62-160
 *   Measurer::measure (walls, OAverage)         ModelG/measurer.h:169-179
 *   o4_stepper_monitor                          ModelG/Stepper.h:17-83
 */
#ifndef LGV_H
#define LGV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGV_OK 0
#define LGV_ERR_ARG (-1)
#define LGV_ERR_HIP (-2)
#define LGV_ERR_STATE (-3)
#define LGV_ERR_RANGE (-4)

enum { LGV_FP64 = 0, LGV_FP32 = 1 };

/* ModelAData + ModelACoefficients + the seed of ModelAHandlerData
 * (ModelG/ModelA.h:71-111, 114-181, 185-248).  NY = NZ = NX as in the
 * reference; LX is read as an integer there (ModelA.h:229), 0 means LX = NX.
 * NX is even and at most 1536 (site indices are 32-bit on the device). */
typedef struct {
  int NX;
  int LX;
  double mass0, dmassdt;
  double lambda, H, chi;
  double gamma, diffusion;
  int seed;
  double initialtime;
} lgv_model_params;

typedef struct {
  int nreplicas;
  const lgv_model_params *replicas;
  int precision;        /* LGV_FP64 or LGV_FP32 (fields stored in FP32, all
                           arithmetic and reductions in FP64) */
  unsigned scounts[2];  /* PV2HBSplit {inner, outer}; {2,3} = PV2HBSplit23 */
  int nodiffuse;        /* PV2HBSplit23NoDiffuse  (SuperPions.cxx:212-215) */
  int onlydiffuse;      /* PV2HBSplit23OnlyDiffuse (SuperPions.cxx:216-220) */
  int track_energy;     /* compute IdealPV2's E_old/E_new and its monitor */
  int device;           /* HIP device ordinal */
} lgv_config;

/* o4_stepper_monitor, ModelG/Stepper.h:17-83 */
typedef struct {
  int64_t down_counts, up_counts, up_yes, up_no;
  double down_dS, up_dS;
} lgv_monitor;

enum { LGV_MON_PHI = 0, LGV_MON_CHARGE = 1, LGV_MON_IDEAL = 2 };

typedef struct lgv_ctx lgv_ctx;

const char *lgv_last_error(void);
int lgv_version(void);

/* ModelA::ModelA (grid + NoiseGenerator(seed)), ModelG/ModelA.h:311-336,
 * and the PV2HBSplit constructor, ModelG/Stepper.cxx:537-540.  Fields start
 * at zero (ModelA::initialize, ModelA.h:408-461). */
int lgv_create(const lgv_config *cfg, lgv_ctx **out);
/* Stepper::finalize + ModelA::finalize storage release. */
int lgv_destroy(lgv_ctx *ctx);

/* ModelA::solution in the o4fields layout [z][y][x][10] (G_node order
 * f[4], A[3], V[3]), FP64 host memory (ModelA.h:274-281, 384-402). */
int lgv_set_field(lgv_ctx *ctx, int replica, const double *aos);
int lgv_get_field(lgv_ctx *ctx, int replica, double *aos);

/* ModelATime::time (ModelA.h:32-68): the clock that m^2(t) reads.
 * replica < 0 sets every replica. */
int lgv_set_time(lgv_ctx *ctx, int replica, double t);
int lgv_get_time(lgv_ctx *ctx, int replica, double *t);

/* PV2HBSplit::step(dt), ModelG/Stepper.cxx:542-580, on every replica.
 * Stream-ordered; does not advance the clock (the caller does, as
 * run_event does at SuperPions.cxx:158-159). */
int lgv_step(lgv_ctx *ctx, double dt);
/* nsteps x { PV2HBSplit::step(dt); t += dt } -- the inner loop of
 * run_event (SuperPions.cxx:124-160) without the measurements. */
int lgv_run(lgv_ctx *ctx, double dt, int nsteps);
/* HIP graphs for lgv_run: 0 = off (kernels launched on the stream);
 * K >= 1 = one graph of K full steps (plus one of a single step for the
 * remainder of a run) captured on first use and replayed. */
int lgv_use_graph(lgv_ctx *ctx, int enable);

/* The sub-steppers on their own: IdealPV2::step (Stepper.cxx:23-66),
 * EulerLangevinHB::step (Stepper.cxx:287-389), ModelGChargeHB::step
 * (Stepper.cxx:449-522). */
int lgv_ideal_step(lgv_ctx *ctx, double dt);
int lgv_hb_phi_step(lgv_ctx *ctx, double dt);
int lgv_charge_step(lgv_ctx *ctx, double dt);

/* ModelA::initialize_gaussian_charges (ModelA.h:463-525): normal() pairs
 * from the replica's global engine, zero mode removed; host-generated. */
int lgv_init_gaussian_charges(lgv_ctx *ctx, int replica);
/* thermalize_event (SuperPions.cxx:21-76) without quench: charges, then
 * 6*nsteps phi sweeps at dt/6 with the clock held. */
int lgv_thermalize(lgv_ctx *ctx, double dt, int nsteps);

/* o4_stepper_monitor counters of one replica: out[LGV_MON_PHI],
 * out[LGV_MON_CHARGE], out[LGV_MON_IDEAL]. */
int lgv_monitor_get(lgv_ctx *ctx, int replica, lgv_monitor out[3]);
/* IdealPV2::getEnergy (Stepper.h:98-101): the last substep's E_old, E_new. */
int lgv_energy_get(lgv_ctx *ctx, int replica, double out[2]);

/* Random-number state: the replica's global engine (the two words that
 * NoiseGenerator::write stores in tag_<rank>.rng, NoiseGenerator.h:112-147),
 * the cached normal, and the per-site stream counter. */
typedef struct {
  uint64_t s0, s1;
  double saved_normal;
  int has_saved_normal;
  uint64_t sweep_ctr;
} lgv_rng_state;
int lgv_rng_get_state(lgv_ctx *ctx, int replica, lgv_rng_state *st);
int lgv_rng_set_state(lgv_ctx *ctx, int replica, const lgv_rng_state *st);

/* Measurer::computeSliceAverage / computeSliceAveragePhase and the
 * OAverage of computeDerivedObs (measurer.cxx:10-300) for one replica, in
 * FP64 with compensated sums: wall[3][11][N], wallphase[3][14][N]
 * (d = x, y, z; already divided by N^2), oavg[11]. */
int lgv_measure(lgv_ctx *ctx, int replica, double *wall, double *wallphase,
                double *oavg);

/* HIP-event timing of the kernel classes of a step on the context's stream
 * (the reference's PetscLogEvent "IdealStep" / "HBStep" / "QHBStep",
 * Stepper.cxx:547-577).  enable: 0 off, 1 every step, k > 1 every k-th step
 * of lgv_step / lgv_run from the next one on (the other steps run as without
 * timing, the charge sweep beside the phi sweeps where the context does so;
 * a timed step keeps one stream).  Enabling resets the totals; while
 * enabled, lgv_run issues kernels directly instead of replaying a graph.
 * The single sweeps (lgv_ideal_step, ...) are timed whenever enabled. */
enum { LGV_CLASS_IDEAL = 0, LGV_CLASS_PHI = 1, LGV_CLASS_CHARGE = 2, LGV_CLASS_ENERGY = 3,
       LGV_NCLASS = 4 };
int lgv_timing_enable(lgv_ctx *ctx, int enable);
int lgv_timing_get(lgv_ctx *ctx, double ms[LGV_NCLASS], long long launches[LGV_NCLASS]);

/* Measurer::measure in full (measurer.h:169-179, measurer.cxx:10-317):
 * walls, OAverage, the r2c transforms / N of every wall row (FFTW forward
 * sign, k = 0..N/2; computed on the device) and rotate_to_vev of the
 * k-space walls.  Any pointer may be NULL.  Shapes, d = x, y, z:
 *   wall[3][11][N]  wallphase[3][14][N]  oavg[11]
 *   wall_k[3][11][N/2+1][2]  wall_k_rotated[3][14][N/2+1][2]
 *   wallphase_k[3][14][N/2+1][2]
 * (the HDF5 rows of measurer_output.cxx:71-130). */
typedef struct {
  double *wall, *wallphase, *oavg;
  double *wall_k, *wall_k_rotated, *wallphase_k;
} lgv_measurement;
int lgv_measure_full(lgv_ctx *ctx, int replica, lgv_measurement *m);

/* ModelACoefficients::mass0 / dmassdt (ModelA.h:73-74), changed at run
 * time by thermalize_event's quench mode (SuperPions.cxx:31-75).
 * replica < 0 sets every replica. */
int lgv_set_mass(lgv_ctx *ctx, int replica, double mass0, double dmassdt);
/* Zeroes every o4_stepper_monitor counter (a fresh stepper object). */
int lgv_monitor_reset(lgv_ctx *ctx);

/* Tile of the fused ideal-step kernel for one replica (replica < 0: all):
 * at most txh_cap half-columns (2..64) in x, 256/txh rows in y, tz planes
 * per z-chunk.  Results do not depend on it; the autotuner picks it. */
int lgv_set_ideal_tile(lgv_ctx *ctx, int replica, int txh_cap, int tz);
/* Tiles of the two Metropolis sweep kernels for one replica (replica < 0:
 * all): the phi z-march (at most hb_cap half-columns, 4..64, x 256/hb_cap
 * rows, chunks of hb_tz planes) and the charge march (at most chg_cap
 * half-columns, 4..64, x 256/chg_cap lines, chunks of chg_chunk sites, even).
 * 0 keeps a default.  Results do not depend on them; the autotuner picks
 * them.  Monitor counts are kept across the change. */
int lgv_set_sweep_tiles(lgv_ctx *ctx, int replica, int hb_cap, int hb_tz, int chg_cap,
                        int chg_chunk);
/* The kernel of the phi half-sweeps (EulerLangevinHB::step,
 * ModelG/Stepper.cxx:287-389): LGV_PHI_SITE one thread per site,
 * LGV_PHI_MARCH the z-march with the other parity's plane in LDS,
 * LGV_PHI_AUTO the size rule (the march for FP64 fields above 1 GiB and FP32
 * fields above 600 MiB).  Results do not depend on it; the autotuner picks
 * it.  Monitor counts are kept across the change.  lgv_get_phi_kernel
 * reports the kernel in use (LGV_PHI_SITE or LGV_PHI_MARCH). */
enum { LGV_PHI_AUTO = 0, LGV_PHI_SITE = 1, LGV_PHI_MARCH = 2 };
int lgv_set_phi_kernel(lgv_ctx *ctx, int kernel);
int lgv_get_phi_kernel(lgv_ctx *ctx, int *kernel);

int lgv_synchronize(lgv_ctx *ctx);
/* Device address of a replica's field storage (10 component planes, each
 * split into even and odd half-lattices) and its element size in bytes. */
int lgv_field_device_ptr(lgv_ctx *ctx, int replica, void **ptr, int *elem_bytes);

/* ------------------------------------------------------------------------
 * Model A Euler--Maruyama stepper with a hand-written adjoint.  No
 * counterpart in the reference (its Model A is the Metropolis
 * EulerLangevinHB, Stepper.cxx:287-389; the forward-Euler recursion is the
 * one of ModelG/freeCase.ipynb cells 5, 7).  Fields are caller-owned FP64
 * device arrays phi[4][N][N][N] (a, z, y, x).  Noise is replayed from the
 * per-site counter streams, so the backward pass sees the forward noise.
 * The loss is mean_w(phibar) + N^3 var_w(phibar) over the last `window`
 * steps, phibar_n the volume mean of phi_0 after step n. */
typedef struct {
  int N;
  double m2, lambda, H, gamma, dt;
  int seed;       /* seed_seq{seed} word -> per-site stream keys */
  int noise_kind; /* 0: variance1 uniform (NoiseGenerator.h:105), 1: Box-Muller */
} lgv_em_params;
typedef struct lgv_em lgv_em;
const char *lgv_em_last_error(void);
int lgv_em_create(const lgv_em_params *p, int nsteps, int ckpt_every, int device,
                  lgv_em **out);
/* Launch on an external stream (e.g. torch's current stream); NULL = own. */
int lgv_em_set_stream(lgv_em *h, void *hip_stream);
int lgv_em_forward(lgv_em *h, const double *phi0, int window, double *phi_final,
                   double *phibar_host, double *loss);
/* d loss / d phi0 (device, same layout) and d loss / d (m2, H, lambda). */
int lgv_em_backward(lgv_em *h, double *grad_phi0, double grad_params[3]);
/* The observables of one state, 6 doubles: the volume means phibar_0..3,
 * the volume mean of |phi|^2 and V |phibar|^2 (the per-measurement
 * estimator whose series gives the susceptibility V (<|phibar|^2> -
 * <|phibar|>^2)); the volume averages of Measurer, measurer.h:169-179. */
#define LGV_EM_NOBS 6
/* The measurement series of the run loop (measure, then step:
 * SuperPions.cxx:124-160): the handle's nsteps steps from phi0, with the
 * observables of the states 0, every, 2 every, ... below nsteps written to
 * rows_host[ceil(nsteps / every)][LGV_EM_NOBS].  phi_final (device, may be
 * NULL) receives the last state.  Leaves the forward/backward state alone. */
int lgv_em_series(lgv_em *h, const double *phi0, int every, double *phi_final,
                  double *rows_host);
/* One step, number n (it selects the noise), from in to out (caller-owned
 * device arrays, in != out), stream-ordered and without a host
 * synchronisation.  obs (device memory, may be NULL) receives the
 * LGV_EM_NOBS observables of out. */
int lgv_em_step(lgv_em *h, const double *in, double *out, int n, double *obs);
/* lgv_em_step without a handle and without allocating anything: for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  No counterpart in the reference.  The
 * step's reduction slots live in one static device array per GPU (640 KiB,
 * held until the process ends), so N <= 256 (LGV_ERR_RANGE beyond);
 * calls that overlap on one GPU (two streams, two threads) race on it.
 * Same arguments and results as lgv_em_step with the handle's parameters,
 * on `device` (made current) and `hip_stream` (NULL = the null stream).
 * The first call on a device looks the array's address up (no stream
 * work); make that call before a graph capture. */
int lgv_em_step_noalloc(const lgv_em_params *p, const double *in, double *out, int n,
                        double *obs, int device, void *hip_stream);
int lgv_em_destroy(lgv_em *h);

#ifdef __cplusplus
}
#endif
#endif
