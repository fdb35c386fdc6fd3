/*
 * lgv_oracle.h -- CPU restatement of langevinA's ModelG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in langevina_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Layout: natural AoS lattice u[((k*N + j)*N + i)*10 + L], L = 0..3 phi_a,
 * 4..6 axial A_1..3, 7..9 vector V_1..3 -- the order of G_node
 * (ModelG/ModelA.h:274-281) and of the o4fields dataset.
 *
 * Two random-number modes:
 *   ORC_RNG_REFERENCE  one sequential xoroshiro128+ engine consumed in the
 *                      reference's loop order (ModelG/Stepper.cxx), i.e. the
 *                      single-rank reference trajectory;
 *   ORC_RNG_PER_SITE   per-site counter streams: every (sweep, site) pair
 *                      seeds its own xoroshiro128+ from a splitmix64 hash;
 *                      this is the mode the GPU implements bit-for-bit.
 */
#ifndef LGV_ORACLE_H
#define LGV_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- random numbers ---------------- */
typedef struct {
  uint64_t s0, s1;
} orc_xoro;

void orc_xoro_seed(orc_xoro *g, uint64_t seed);
uint64_t orc_xoro_next(orc_xoro *g);
double orc_canonical(uint64_t x);
double orc_uniform(orc_xoro *g);
double orc_variance1(orc_xoro *g);
void orc_seed_seq_generate(const uint32_t *v, size_t s, uint32_t *out, size_t n);
uint64_t orc_uniform_int(orc_xoro *g, uint64_t b);
void orc_shuffle(orc_xoro *g, int *a, int n);

typedef struct {
  orc_xoro g;
  double saved;
  int has_saved;
} orc_noise;
void orc_noise_init(orc_noise *ng, int base_seed);
double orc_normal(orc_noise *ng);

uint64_t orc_mix64(uint64_t z);
uint64_t orc_base_key(uint32_t seed32);
uint64_t orc_stream_key(uint64_t base, uint64_t ctr);
uint64_t orc_site_key(uint64_t skey, uint64_t site);
void orc_stream_moments(uint64_t base, uint64_t ctr, long n, int nout, double *mean,
                        double *cross);

/* ---------------- O(4) algebra ---------------- */
void orc_expmul(const double *v, double *w);
void orc_quatmul(const double *v1, const double *v2, double *w);
void orc_o4_rotation(const double *V, const double *A, double *phi);

/* ---------------- monitors ---------------- */
typedef struct {
  int64_t down_counts, up_counts, up_yes, up_no;
  double down_dS, up_dS;
} orc_monitor;

/* ---------------- lattice context ---------------- */
enum { ORC_RNG_REFERENCE = 0, ORC_RNG_PER_SITE = 1 };
enum { ORC_FP64 = 0, ORC_FP32 = 1 };

typedef struct {
  int N;
  double hx; /* lattice spacing LX/NX */
  double mass0, dmassdt, lambda, H, chi, gamma, diffusion;
  int seed;
  int rng_mode;
  int precision;
  unsigned scounts[2];
  int nodiffuse, onlydiffuse;
  /* state */
  double t;
  uint64_t sweep_ctr; /* per-site mode: next counter value */
  uint64_t base_key;
  orc_noise noise; /* the global engine (ModelARndm) */
  orc_monitor mon_phi, mon_charge, mon_ideal;
  double last_energy[2];
  int last_order[5];
} orc_ctx;

void orc_ctx_init(orc_ctx *c);
/* OpenMP threads for the per-site-stream mode (1 = the serial loops). */
void orc_set_threads(int n);
int orc_get_threads(void);
double orc_mass(const orc_ctx *c);
double orc_energy(const orc_ctx *c, const double *u);
void orc_rotate_phi(const orc_ctx *c, double *u, double tau);
void orc_ideal_kick(const orc_ctx *c, double *u, double dt);
void orc_ideal_step(orc_ctx *c, double *u, double dt);
void orc_hb_phi_halfsweep(orc_ctx *c, double *u, double dt, int ieo);
void orc_hb_phi_step(orc_ctx *c, double *u, double dt);
long orc_div_by_const_mismatches(double chi, long n, uint64_t seed);
void orc_one_site_chain(int per_site, long n, double h, double mass, double lambda, double H,
                        double dt, uint32_t seed, double out[2]);
void orc_charge_subsweep(orc_ctx *c, double *u, double rms, int dir, int eoA,
                         uint64_t ctr);
void orc_charge_step(orc_ctx *c, double *u, double dt);
double orc_update_charge_pair(double chi, double rms, double nA, double nB,
                              orc_xoro *g, orc_monitor *m);
void orc_pv2hbsplit_step(orc_ctx *c, double *u, double dt);
void orc_init_gaussian_charges(orc_ctx *c, double *u);
void orc_thermalize(orc_ctx *c, double *u, double dt, int nsteps);
void orc_round_fp32(double *u, size_t n);

/* ---------------- measurement ---------------- */
void orc_measure_walls(int N, const double *u, double *wall /*[3][11][N]*/,
                       double *wallphase /*[3][14][N]*/, double *oavg /*[11]*/);
void orc_dft_rows(int N, int rows, const double *in, double *out /*[rows][N/2+1][2]*/);
void orc_rotate_to_vev(int nk, const double *wallk /*[11][nk][2]*/,
                       double *rot /*[14][nk][2]*/);

/* ---------------- Model A Euler--Maruyama (new) ---------------- */
typedef struct {
  int N;
  double m2, lambda, H, gamma, dt;
  uint64_t base_key;
  int noise_kind; /* 0 = variance1 uniform, 1 = Box--Muller gaussian */
} orc_em_params;
void orc_em_noise(const orc_em_params *p, uint64_t step, double *xi /*[N^3][4]*/);
void orc_em_step(const orc_em_params *p, uint64_t step, const double *phi,
                 double *out /*[N^3][4]*/);
double orc_em_run(const orc_em_params *p, const double *phi0, int nsteps,
                  int window, double *phibar /*[nsteps]*/, double *phi_final);
void orc_em_series(const orc_em_params *p, const double *phi0, int nsteps, int every,
                   double *rows /*[ceil(nsteps/every)][6]*/, double *phi_final);
void orc_em_adjoint(const orc_em_params *p, const double *phi0, int nsteps,
                    int window, double *grad_phi0, double *grad_params /*[3] m2,H,lambda*/);

#ifdef __cplusplus
}
#endif
#endif
