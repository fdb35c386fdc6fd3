/*
 * lgv_oracle.c -- CPU restatement of langevinA's ModelG hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function names the reference lines it restates.  Arithmetic follows
 * the reference expression by expression (same operand order, no FMA: this
 * file is compiled with -ffp-contract=off), so that in ORC_RNG_REFERENCE
 * mode the trajectory is the single-rank reference trajectory.
 *
 * Pins (tests/test_oracle_*.py):
 *   xoroshiro128+, seed_seq, generate_canonical, uniform_real, normal,
 *   shuffle            -> bit-for-bit against oracle/_ref/ref_rng, built from
 *                         ModelG/xoroshiro128plus.h and libstdc++ itself;
 *   O(4) rotation      -> the three known answers of
 *                         ModelG/O4AlgebraSignTest.cxx:20-38, the 4x4 matrix
 *                         exponential of the so(4) generator (scipy), norm
 *                         preservation and inverse rotation;
 *   ideal kick         -> sum V conserved and sum A_l changes by
 *                         -dt*H*sum phi_{l+1} (lattice divergence), PV2
 *                         energy drift second order in dt;
 *   phi Metropolis     -> free-field closed form 1/(m^2+k^2) (freeCase.ipynb
 *                         cell 7 free2pt), detailed balance on a 2-site
 *                         brute-force enumeration;
 *   charge Metropolis  -> exact conservation of sum n, equilibrium variance
 *                         chi, transition-kernel histogram against the
 *                         closed-form Metropolis kernel;
 *   measurer           -> exact sums (math.fsum), DFT against numpy.fft.rfft,
 *                         the cos/sin test signal of ModelG/x2ktest.py;
 *   Euler--Maruyama    -> free forward-Euler two-point function
 *                         (freeCase.ipynb cells 5, 7 free2ptFE) and central
 *                         finite differences for the adjoint.
 *
 * orc_set_threads(n > 1) runs the per-site-stream mode's sweeps, rotations,
 * kick and energy sums over n OpenMP threads (z-planes split, as MPI ranks
 * split the reference's lattice).  Fields and accept/reject decisions are
 * unchanged; the monitor and energy sums are added in another order.  Only
 * the CPU throughput baseline uses it.
 */
#include "lgv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 1 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

/* ======================================================================
 * Random numbers
 * ==================================================================== */

static inline uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

/* xoroshiro128plus::seed(uint64_t), ModelG/xoroshiro128plus.h:86-91 */
void orc_xoro_seed(orc_xoro *g, uint64_t seed) {
  g->s0 = 0x01d353e5f3993bb0ULL + seed;
  g->s1 = 0x7b9c0df6cb193b20ULL * (seed + 1);
}

/* xoroshiro128plus::operator(), ModelG/xoroshiro128plus.h:97-107 */
uint64_t orc_xoro_next(orc_xoro *g) {
  const uint64_t s0 = g->s0;
  uint64_t s1 = g->s1;
  const uint64_t result = s0 + s1;
  s1 ^= s0;
  g->s0 = rotl64(s0, 55) ^ s1 ^ (s1 << 14);
  g->s1 = rotl64(s1, 36);
  return result;
}

/* std::generate_canonical<double,53>(xoroshiro128plus), libstdc++ 11
 * bits/random.tcc:3348-3379: one 64-bit draw (m = 1), sum = double(x),
 * ret = sum / 2^64, clamped below 1. */
double orc_canonical(uint64_t x) {
  double sum = (double)x;
  double ret = sum / 18446744073709551616.0;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

/* uniform_real_distribution(0,1): u*(1-0)+0, libstdc++ bits/random.h:1870;
 * NoiseGenerator::uniform, ModelG/NoiseGenerator.h:101 */
double orc_uniform(orc_xoro *g) {
  double u = orc_canonical(orc_xoro_next(g));
  return u * (1.0 - 0.0) + 0.0;
}

/* uniform_real_distribution(-sqrt3, sqrt3): u*(b-a)+a with b-a evaluated
 * per call; NoiseGenerator::variance1, ModelG/NoiseGenerator.h:68-69,105 */
double orc_variance1(orc_xoro *g) {
  const double a = -sqrt(3.), b = sqrt(3.);
  double u = orc_canonical(orc_xoro_next(g));
  return u * (b - a) + a;
}

/* std::seed_seq::generate, libstdc++ 11 bits/random.tcc:3250-3330
 * (the algorithm of [rand.util.seedseq]) */
void orc_seed_seq_generate(const uint32_t *v, size_t s, uint32_t *b, size_t n) {
  if (n == 0) return;
  for (size_t i = 0; i < n; i++) b[i] = 0x8b8b8b8bu;
  const size_t t = (n >= 623) ? 11 : (n >= 68) ? 7 : (n >= 39) ? 5
                 : (n >= 7) ? 3 : (n - 1) / 2;
  const size_t p = (n - t) / 2;
  const size_t q = p + t;
  const size_t m = (s + 1 > n) ? s + 1 : n;
  {
    uint32_t r1 = 1371501266u; /* 1664525 * T(0x8b8b8b8b) */
    uint32_t r2 = r1 + (uint32_t)s;
    b[p] += r1;
    b[q] = b[q] + r2;
    b[0] = r2;
  }
  for (size_t k = 1; k < m; ++k) {
    const size_t kn = k % n, kpn = (k + p) % n, kqn = (k + q) % n;
    uint32_t arg = b[kn] ^ b[kpn] ^ b[(k - 1) % n];
    uint32_t r1 = 1664525u * (arg ^ (arg >> 27));
    uint32_t r2 = r1 + (uint32_t)kn;
    if (k <= s) r2 += v[k - 1];
    b[kpn] = b[kpn] + r1;
    b[kqn] = b[kqn] + r2;
    b[kn] = r2;
  }
  for (size_t k = m; k < m + n; ++k) {
    const size_t kn = k % n, kpn = (k + p) % n, kqn = (k + q) % n;
    uint32_t arg = b[kn] + b[kpn] + b[(k - 1) % n];
    uint32_t r3 = 1566083941u * (arg ^ (arg >> 27));
    uint32_t r4 = r3 - (uint32_t)kn;
    b[kpn] ^= r3;
    b[kqn] ^= r4;
    b[kn] = r4;
  }
}

/* uniform_int_distribution<unsigned long>{0,b}(xoroshiro128plus): Lemire's
 * nearly divisionless downscaling with 128-bit products, libstdc++ 11
 * bits/uniform_int_dist.h:246-307 (_S_nd) */
uint64_t orc_uniform_int(orc_xoro *g, uint64_t b) {
  const uint64_t range = b + 1;
  unsigned __int128 product = (unsigned __int128)orc_xoro_next(g) * range;
  uint64_t low = (uint64_t)product;
  if (low < range) {
    const uint64_t threshold = (uint64_t)(-range) % range;
    while (low < threshold) {
      product = (unsigned __int128)orc_xoro_next(g) * range;
      low = (uint64_t)product;
    }
  }
  return (uint64_t)(product >> 64);
}

/* std::shuffle with a 64-bit engine, libstdc++ 11 bits/stl_algo.h:3731-3792:
 * the two-positions-per-draw path (taken whenever range^2 fits the engine) */
void orc_shuffle(orc_xoro *g, int *a, int n) {
  if (n <= 1) return;
  int i = 1;
  if (n % 2 == 0) {
    uint64_t d = orc_uniform_int(g, 1);
    int tmp = a[i]; a[i] = a[d]; a[d] = tmp;
    i++;
  }
  while (i != n) {
    const uint64_t r = (uint64_t)i + 1;
    const uint64_t x = orc_uniform_int(g, r * (r + 1) - 1);
    const uint64_t p0 = x / (r + 1), p1 = x % (r + 1);
    int tmp = a[i]; a[i] = a[p0]; a[p0] = tmp;
    i++;
    tmp = a[i]; a[i] = a[p1]; a[p1] = tmp;
    i++;
  }
}

/* NoiseGenerator(baseSeed), ModelG/NoiseGenerator.h:68-82 with one rank:
 * seed_seq{baseSeed}.generate(1 word) -> engine.seed(word) */
void orc_noise_init(orc_noise *ng, int base_seed) {
  uint32_t v = (uint32_t)base_seed, out = 0;
  orc_seed_seq_generate(&v, 1, &out, 1);
  orc_xoro_seed(&ng->g, (uint64_t)out);
  ng->has_saved = 0;
  ng->saved = 0.;
}

/* normal_distribution<double> (Marsaglia polar, pairs cached),
 * libstdc++ 11 bits/random.tcc:1802-1835 */
double orc_normal(orc_noise *ng) {
  double ret;
  if (ng->has_saved) {
    ng->has_saved = 0;
    ret = ng->saved;
  } else {
    double x, y, r2;
    do {
      x = 2.0 * orc_canonical(orc_xoro_next(&ng->g)) - 1.0;
      y = 2.0 * orc_canonical(orc_xoro_next(&ng->g)) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2 * log(r2) / r2);
    ng->saved = x * mult;
    ng->has_saved = 1;
    ret = y * mult;
  }
  return ret * 1.0 + 0.0;
}

/* Per-site counter streams (new; not in the reference).  splitmix64's
 * output function (Steele, Lea, Flood 2014; Vigna's public-domain
 * splitmix64.c) is the mixer.  A stream counter and a site index are hashed
 * into a 64-bit key; xoroshiro128plus::seed(key) then starts that site's
 * stream and draws proceed exactly as in the reference engine. */
uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
uint64_t orc_base_key(uint32_t seed32) {
  return orc_mix64((uint64_t)seed32 ^ 0x6a09e667f3bcc909ULL);
}
uint64_t orc_stream_key(uint64_t base, uint64_t ctr) {
  return orc_mix64(base + orc_mix64(ctr + 0x9e3779b97f4a7c15ULL));
}
uint64_t orc_site_key(uint64_t skey, uint64_t site) {
  return orc_mix64(skey ^ orc_mix64(site + 0x3c6ef372fe94f82bULL));
}

/* First moments of the keyed streams: mean[k] of the k-th canonical output
 * and cross[k] = E[u_k u_{k+1}] over the streams of sites 0..n-1 of one
 * counter (a statistical check of the per-site scheme). */
void orc_stream_moments(uint64_t base, uint64_t ctr, long n, int nout, double *mean,
                        double *cross) {
  const uint64_t skey = orc_stream_key(base, ctr);
  for (int k = 0; k < nout; k++) mean[k] = cross[k] = 0.;
  for (long i = 0; i < n; i++) {
    orc_xoro g;
    orc_xoro_seed(&g, orc_site_key(skey, (uint64_t)i));
    double prev = 0.;
    for (int k = 0; k < nout; k++) {
      const double u = orc_canonical(orc_xoro_next(&g));
      mean[k] += u;
      if (k > 0) cross[k - 1] += prev * u;
      prev = u;
    }
  }
  for (int k = 0; k < nout; k++) {
    mean[k] /= (double)n;
    cross[k] /= (double)n;
  }
}

/* ======================================================================
 * O(4) algebra, ModelG/O4AlgebraHelper.h
 * ==================================================================== */

/* norm, :17-19 */
static double orc_norm3(const double *v) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* expmul, :64-76 */
void orc_expmul(const double *v, double *w) {
  double abs = orc_norm3(v);
  if (abs > 1e-20) {
    w[0] = cos(abs);
    const double a = -(sin(abs) / abs);
    for (int i = 0; i < 3; ++i) w[1 + i] = a * v[i];
  } else {
    w[0] = 1;
    w[1] = 0.0;
    w[2] = 0.0;
    w[3] = 0.0;
  }
}

/* quatmul (with dot, cross, muladd), :40-62 */
void orc_quatmul(const double *v1, const double *v2, double *w) {
  double dot = 0;
  for (int i = 0; i < 3; ++i) dot += v1[1 + i] * v2[1 + i];
  w[0] = v1[0] * v2[0] - dot;
  const double *a = v1 + 1, *b = v2 + 1;
  w[1] = a[1] * b[2] - b[1] * a[2];
  w[2] = a[2] * b[0] - a[0] * b[2];
  w[3] = a[0] * b[1] - b[0] * a[1];
  for (int i = 0; i < 3; ++i) w[1 + i] += v2[0] * v1[1 + i];
  for (int i = 0; i < 3; ++i) w[1 + i] += v1[0] * v2[1 + i];
}

/* O4Rotation, :78-97 */
void orc_o4_rotation(const double *V, const double *A, double *phi) {
  double tmp1[4], tmp2[4], tmp3[4];
  for (int i = 0; i < 3; ++i) {
    tmp1[i] = -0.5 * (V[i] - A[i]);
    tmp2[i] = 0.5 * (V[i] + A[i]);
  }
  orc_expmul(tmp1, tmp3);
  orc_expmul(tmp2, tmp1);
  orc_quatmul(phi, tmp3, tmp2);
  orc_quatmul(tmp1, tmp2, phi);
}

/* ======================================================================
 * Lattice steppers, ModelG/Stepper.cxx
 * ==================================================================== */

#define NDOF 10
static inline size_t site_index(int N, int k, int j, int i) {
  return ((size_t)k * N + j) * N + i;
}
static inline int wrap(int a, int N) { return a < 0 ? a + N : (a >= N ? a - N : a); }

static inline float to_f32(double x) { return (float)x; }
void orc_round_fp32(double *u, size_t n) {
  for (size_t i = 0; i < n; i++) u[i] = (double)to_f32(u[i]);
}
static inline double store(const orc_ctx *c, double x) {
  return c->precision == ORC_FP32 ? (double)to_f32(x) : x;
}

void orc_ctx_init(orc_ctx *c) {
  orc_noise_init(&c->noise, c->seed);
  uint32_t v = (uint32_t)c->seed, out = 0;
  orc_seed_seq_generate(&v, 1, &out, 1);
  c->base_key = orc_base_key(out);
  c->sweep_ctr = 0;
  memset(&c->mon_phi, 0, sizeof(orc_monitor));
  memset(&c->mon_charge, 0, sizeof(orc_monitor));
  memset(&c->mon_ideal, 0, sizeof(orc_monitor));
  c->last_energy[0] = c->last_energy[1] = 0.;
  if (c->scounts[0] == 0) { c->scounts[0] = 2; c->scounts[1] = 3; }
}

/* ModelACoefficients::mass(t), ModelG/ModelA.h:84 */
double orc_mass(const orc_ctx *c) { return c->mass0 + c->dmassdt * c->t; }

/* o4_stepper_monitor increments, ModelG/Stepper.h:36-49 */
static void mon_down(orc_monitor *m, double dS) { m->down_counts++; m->down_dS += dS; }
static void mon_up_yes(orc_monitor *m, double dS) { m->up_counts++; m->up_yes++; m->up_dS += dS; }
static void mon_up_no(orc_monitor *m) { m->up_counts++; m->up_no++; m->up_dS += 0.; }

/* A stepper's random source: the global engine in reference mode, or a
 * per-site stream seeded from (counter, site) in per-site mode. */
typedef struct {
  orc_xoro *g;
  orc_xoro local;
} draw_src;
static orc_xoro *site_engine(orc_ctx *c, draw_src *d, uint64_t skey, uint64_t site) {
  if (c->rng_mode == ORC_RNG_REFERENCE) return &c->noise.g;
  orc_xoro_seed(&d->local, orc_site_key(skey, site));
  return &d->local;
}

/* IdealPV2::computeEnergy, ModelG/Stepper.cxx:204-268 (single rank; the
 * three partial sums are accumulated in k,j,i,s order) */
double orc_energy(const orc_ctx *c, const double *u) {
  const int N = c->N;
  const double H[4] = {c->H, 0., 0., 0.};
  double grad2 = 0, nab2 = 0, hEn = 0;
  const int nt = c->rng_mode == ORC_RNG_PER_SITE ? g_threads : 1;
#pragma omp parallel for num_threads(nt) if (nt > 1) reduction(+ : grad2, nab2, hEn) schedule(static)
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const double *p = u + site_index(N, k, j, i) * NDOF;
        const double *pz = u + site_index(N, wrap(k + 1, N), j, i) * NDOF;
        const double *py = u + site_index(N, k, wrap(j + 1, N), i) * NDOF;
        const double *px = u + site_index(N, k, j, wrap(i + 1, N)) * NDOF;
        for (int s = 0; s < 4; ++s) {
          double phimid = p[s];
          double d;
          d = pz[s] - phimid; grad2 += d * d;
          d = py[s] - phimid; grad2 += d * d;
          d = px[s] - phimid; grad2 += d * d;
          hEn += phimid * H[s];
        }
        for (int s = 0; s < 3; s++) nab2 += p[7 + s] * p[7 + s];
        for (int s = 0; s < 3; s++) nab2 += p[4 + s] * p[4 + s];
      }
  double e = 0;
  e += 0.5 / c->chi * nab2;
  e += 0.5 * grad2;
  e -= hEn;
  return e;
}

/* IdealPV2::rotatePhi, ModelG/Stepper.cxx:178-202 */
void orc_rotate_phi(const orc_ctx *c, double *u, double tau) {
  const size_t V = (size_t)c->N * c->N * c->N;
  const int nt = c->rng_mode == ORC_RNG_PER_SITE ? g_threads : 1;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
  for (size_t s = 0; s < V; s++) {
    double *p = u + s * NDOF;
    double vmu[3], amu[3];
    for (int a = 0; a < 3; a++) vmu[a] = -p[7 + a] / c->chi * tau;
    for (int a = 0; a < 3; a++) amu[a] = -p[4 + a] / c->chi * tau;
    orc_o4_rotation(vmu, amu, p);
  }
}

/* The charge kick of IdealPV2::step_no_reject, ModelG/Stepper.cxx:99-166.
 * Neighbours are read from a copy of the (rotated) fields, as the
 * reference reads them from the ghosted local vector. */
void orc_ideal_kick(const orc_ctx *c, double *u, double dt) {
  const int N = c->N;
  const size_t V = (size_t)N * N * N;
  double *f = (double *)malloc(V * 4 * sizeof(double));
  const int nt = c->rng_mode == ORC_RNG_PER_SITE ? g_threads : 1;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
  for (size_t s = 0; s < V; s++) memcpy(f + 4 * s, u + NDOF * s, 4 * sizeof(double));
  const double hinv = 1. / c->hx;
  const double axx = hinv * hinv, ayy = axx, azz = axx; /* pow(1/h,2) */
  const double H0 = c->H;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const double *cp = f + 4 * site_index(N, k, j, i);
        const double *xp = f + 4 * site_index(N, k, j, wrap(i + 1, N));
        const double *xm = f + 4 * site_index(N, k, j, wrap(i - 1, N));
        const double *yp = f + 4 * site_index(N, k, wrap(j + 1, N), i);
        const double *ym = f + 4 * site_index(N, k, wrap(j - 1, N), i);
        const double *zp = f + 4 * site_index(N, wrap(k + 1, N), j, i);
        const double *zm = f + 4 * site_index(N, wrap(k - 1, N), j, i);
        double *out = u + NDOF * site_index(N, k, j, i);
        for (int l = 0; l < 3; l++) {
          double advxx = (-xp[0] * cp[l + 1] + cp[0] * xp[l + 1] +
                          cp[0] * xm[l + 1] - xm[0] * cp[l + 1]) * axx;
          double advyy = (-yp[0] * cp[l + 1] + cp[0] * yp[l + 1] +
                          cp[0] * ym[l + 1] - ym[0] * cp[l + 1]) * ayy;
          double advzz = (-zp[0] * cp[l + 1] + cp[0] * zp[l + 1] +
                          cp[0] * zm[l + 1] - zm[0] * cp[l + 1]) * azz;
          out[4 + l] += dt * (advxx + advyy + advzz - H0 * cp[l + 1]);
        }
        for (int s = 0; s < 3; s++) {
          const int s1 = (s + 1) % 3, s2 = (s + 2) % 3;
          const int eps = 1; /* ((s-s1)(s1-s2)(s2-s))/2 == 1 for s = 0,1,2 */
          double advxx = eps * (xp[1 + s2] * cp[1 + s1] - cp[1 + s2] * xm[1 + s1] -
                                xp[1 + s1] * cp[1 + s2] + cp[1 + s1] * xm[1 + s2]) * axx;
          double advyy = eps * (yp[1 + s2] * cp[1 + s1] - cp[1 + s2] * ym[1 + s1] -
                                yp[1 + s1] * cp[1 + s2] + cp[1 + s1] * ym[1 + s2]) * ayy;
          double advzz = eps * (zp[1 + s2] * cp[1 + s1] - cp[1 + s2] * zm[1 + s1] -
                                zp[1 + s1] * cp[1 + s2] + cp[1 + s1] * zm[1 + s2]) * azz;
          out[7 + s] += dt * (advxx + advyy + advzz);
        }
      }
  free(f);
}

/* IdealPV2::step (accept_reject = false), ModelG/Stepper.cxx:23-66 with
 * step_no_reject :68-176.  FP32 mode rounds the stored fields once at the
 * end of the substep; the energies are those of the stored fields. */
void orc_ideal_step(orc_ctx *c, double *u, double dt) {
  const size_t n = (size_t)c->N * c->N * c->N * NDOF;
  uint64_t ctr = 0;
  if (c->rng_mode == ORC_RNG_PER_SITE) ctr = c->sweep_ctr++;
  double oldE = orc_energy(c, u);
  orc_rotate_phi(c, u, dt / 2.0);
  orc_ideal_kick(c, u, dt);
  orc_rotate_phi(c, u, dt / 2.0);
  if (c->precision == ORC_FP32) orc_round_fp32(u, n);
  double newE = orc_energy(c, u);
  double deltaE = newE - oldE;
  c->last_energy[0] = oldE;
  c->last_energy[1] = newE;
  if (deltaE > 0) {
    draw_src d;
    orc_xoro *g = site_engine(c, &d, orc_stream_key(c->base_key, ctr), 0);
    double prob = orc_uniform(g);
    if (prob > exp(-deltaE)) mon_up_no(&c->mon_ideal);
    else mon_up_yes(&c->mon_ideal, deltaE);
  } else {
    mon_down(&c->mon_ideal, deltaE);
  }
}

/* One site of EulerLangevinHB::step, ModelG/Stepper.cxx:330-383 */
static inline void hb_site(orc_ctx *c, double *u, int k, int j, int i, uint64_t skey,
                           double Lambda2, double lambda, double H, double rdtg,
                           orc_monitor *mon) {
  const int N = c->N;
  draw_src d;
  orc_xoro *g = site_engine(c, &d, skey, site_index(N, k, j, i));
  double *p = u + NDOF * site_index(N, k, j, i);
  const double *xp = u + NDOF * site_index(N, k, j, wrap(i + 1, N));
  const double *xm = u + NDOF * site_index(N, k, j, wrap(i - 1, N));
  const double *yp = u + NDOF * site_index(N, k, wrap(j + 1, N), i);
  const double *ym = u + NDOF * site_index(N, k, wrap(j - 1, N), i);
  const double *zp = u + NDOF * site_index(N, wrap(k + 1, N), j, i);
  const double *zm = u + NDOF * site_index(N, wrap(k - 1, N), j, i);
  double hphi_o = 0., hphi_n = 0., s_o = 0., s_n = 0.;
  double phi_o[4], phi_n[4];
  for (int l = 0; l < 4; l++) {
    double heff = (xp[l] + xm[l]) + (yp[l] + ym[l]) + (zp[l] + zm[l]);
    phi_o[l] = p[l];
    phi_n[l] = phi_o[l] + rdtg * orc_variance1(g);
    s_o += phi_o[l] * phi_o[l];
    s_n += phi_n[l] * phi_n[l];
    hphi_o += heff * phi_o[l];
    hphi_n += heff * phi_n[l];
  }
  hphi_o += (H * phi_o[0]);
  hphi_n += (H * phi_n[0]);
  double dS = -(hphi_n - hphi_o) + Lambda2 * (s_n - s_o) +
              lambda / 4. * (s_n * s_n - s_o * s_o);
  if (dS < 0) {
    for (int l = 0; l < 4; l++) p[l] = store(c, phi_n[l]);
    mon_down(mon, dS);
    return;
  }
  double r = orc_uniform(g);
  if (r < exp(-dS)) {
    for (int l = 0; l < 4; l++) p[l] = store(c, phi_n[l]);
    mon_up_yes(mon, dS);
  } else {
    mon_up_no(mon);
  }
}

static void mon_merge(orc_monitor *into, const orc_monitor *m) {
  into->down_counts += m->down_counts;
  into->up_counts += m->up_counts;
  into->up_yes += m->up_yes;
  into->up_no += m->up_no;
  into->down_dS += m->down_dS;
  into->up_dS += m->up_dS;
}

/* One parity of EulerLangevinHB::step, ModelG/Stepper.cxx:309-386 */
void orc_hb_phi_halfsweep(orc_ctx *c, double *u, double dt, int ieo) {
  const int N = c->N;
  const double Lambda2 = 0.5 * (2. * 3. + orc_mass(c));
  const double lambda = c->lambda;
  const double H = c->H;
  const double rdtg = sqrt(2. * dt * c->gamma);
  uint64_t skey = 0;
  if (c->rng_mode == ORC_RNG_PER_SITE) skey = orc_stream_key(c->base_key, c->sweep_ctr++);
  if (c->rng_mode == ORC_RNG_PER_SITE && g_threads > 1) {
#pragma omp parallel num_threads(g_threads)
    {
      orc_monitor m = {0};
#pragma omp for schedule(static)
      for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
          for (int i = 0; i < N; i++)
            if ((k + j + i) % 2 == ieo) hb_site(c, u, k, j, i, skey, Lambda2, lambda, H, rdtg, &m);
#pragma omp critical
      mon_merge(&c->mon_phi, &m);
    }
    return;
  }
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++)
        if ((k + j + i) % 2 == ieo)
          hb_site(c, u, k, j, i, skey, Lambda2, lambda, H, rdtg, &c->mon_phi);
}

/* A single site of the phi sweep in the fixed field heff = (h, 0, 0, 0):
 * n Metropolis updates of hb_site's arithmetic (Stepper.cxx:330-383) after
 * 10^4 discarded ones, drawing from one engine seeded with seed
 * (per_site = 0, the reference's layout) or from a fresh per-site stream
 * keyed by (update counter, site 5) as in the per-site mode.  Returns the
 * chain means of phi_0 and phi^2 (a check of the sampling against the
 * exact one-site distribution). */
void orc_one_site_chain(int per_site, long n, double h, double mass, double lambda, double H,
                        double dt, uint32_t seed, double out[2]) {
  const double Lambda2 = 0.5 * (2. * 3. + mass), rdtg = sqrt(2. * dt);
  double phi[4] = {0., 0., 0., 0.};
  orc_xoro glob;
  orc_xoro_seed(&glob, seed);
  const uint64_t base = orc_base_key(seed);
  double s0 = 0., s2 = 0.;
  const long burn = 10000;
  for (long t = 0; t < n + burn; t++) {
    orc_xoro loc, *g = &glob;
    if (per_site) {
      orc_xoro_seed(&loc, orc_site_key(orc_stream_key(base, (uint64_t)t), 5));
      g = &loc;
    }
    double hphi_o = 0., hphi_n = 0., so = 0., sn = 0., pn[4];
    for (int l = 0; l < 4; l++) {
      const double heff = l == 0 ? h : 0.;
      pn[l] = phi[l] + rdtg * orc_variance1(g);
      so += phi[l] * phi[l];
      sn += pn[l] * pn[l];
      hphi_o += heff * phi[l];
      hphi_n += heff * pn[l];
    }
    hphi_o += H * phi[0];
    hphi_n += H * pn[0];
    const double dS = -(hphi_n - hphi_o) + Lambda2 * (sn - so) + lambda / 4. * (sn * sn - so * so);
    if (dS < 0 || orc_uniform(g) < exp(-dS))
      for (int l = 0; l < 4; l++) phi[l] = pn[l];
    if (t >= burn) {
      s0 += phi[0];
      s2 += phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2] + phi[3] * phi[3];
    }
  }
  out[0] = s0 / n;
  out[1] = s2 / n;
}

/* The GPU charge sweep divides dS by 2 chi as q0 = x * inv, q0 + (x - q0 d)
 * * inv with one fused multiply-add each (inv = 1/d rounded).  Counts, over n
 * pseudo-random dividends of magnitude 2^-20 .. 2^20, how often that differs
 * from x / d. */
long orc_div_by_const_mismatches(double chi, long n, uint64_t seed) {
  const double d = 2. * chi, inv = 1. / d;
  orc_xoro g;
  orc_xoro_seed(&g, seed);
  long bad = 0;
  for (long i = 0; i < n; i++) {
    const uint64_t r = orc_xoro_next(&g);
    double x = ldexp((double)(r >> 11), -53) * ldexp(1.0, (int)(orc_xoro_next(&g) % 41) - 20);
    if (r & 1) x = -x;
    const double q0 = x * inv;
    if (fma(fma(-q0, d, x), inv, q0) != x / d) bad++;
  }
  return bad;
}

/* EulerLangevinHB::step, ModelG/Stepper.cxx:287-389 */
void orc_hb_phi_step(orc_ctx *c, double *u, double dt) {
  for (int ieo = 0; ieo < 2; ieo++) orc_hb_phi_halfsweep(c, u, dt, ieo);
}

/* modelg_update_charge_pair, ModelG/Stepper.cxx:419-441 */
double orc_update_charge_pair(double chi, double rms, double nA, double nB,
                                 orc_xoro *g, orc_monitor *m) {
  double q = rms * orc_variance1(g);
  double dS = ((nA - q) * (nA - q) + (nB + q) * (nB + q) - nA * nA - nB * nB) /
              (2. * chi);
  if (dS < 0) {
    mon_down(m, dS);
    return q;
  }
  double r = orc_uniform(g);
  if (r < exp(-dS)) {
    mon_up_yes(m, dS);
    return q;
  }
  mon_up_no(m);
  return 0.;
}

/* One (direction, parity) sub-sweep of ModelGChargeHB::step,
 * ModelG/Stepper.cxx:474-518 with g_face_cases :408-415.  The pairs
 * (A, A+e_dir) with A of parity eoA are disjoint for even N, so updating in
 * place equals the reference's snapshot + scatter-add: nA + (-q), nB + q. */
static inline void charge_site(orc_ctx *c, double *u, int k, int j, int i, int dir,
                               double rms, uint64_t skey, orc_monitor *mon) {
  const int N = c->N;
  const int iB = dir == 0, jB = dir == 1, kB = dir == 2;
  draw_src d;
  orc_xoro *g = site_engine(c, &d, skey, site_index(N, k, j, i));
  double *pA = u + NDOF * site_index(N, k, j, i);
  double *pB = u + NDOF * site_index(N, wrap(k + kB, N), wrap(j + jB, N), wrap(i + iB, N));
  for (int L = 4; L < NDOF; L++) {
    const double nA = pA[L], nB = pB[L];
    double q = orc_update_charge_pair(c->chi, rms, nA, nB, g, mon);
    pA[L] = store(c, nA + (-q));
    pB[L] = store(c, nB + q);
  }
}

void orc_charge_subsweep(orc_ctx *c, double *u, double rms, int dir, int eoA,
                         uint64_t ctr) {
  const int N = c->N;
  const uint64_t skey = orc_stream_key(c->base_key, ctr);
  if (c->rng_mode == ORC_RNG_PER_SITE && g_threads > 1) {
#pragma omp parallel num_threads(g_threads)
    {
      orc_monitor m = {0};
#pragma omp for schedule(static)
      for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
          for (int i = 0; i < N; i++)
            if ((k + j + i) % 2 == eoA) charge_site(c, u, k, j, i, dir, rms, skey, &m);
#pragma omp critical
      mon_merge(&c->mon_charge, &m);
    }
    return;
  }
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++)
        if ((k + j + i) % 2 == eoA) charge_site(c, u, k, j, i, dir, rms, skey, &c->mon_charge);
}

/* ModelGChargeHB::step, ModelG/Stepper.cxx:449-522 */
void orc_charge_step(orc_ctx *c, double *u, double dt) {
  const double rdtsigma = sqrt(2. * dt * (c->diffusion * c->chi));
  int orderxyz[3] = {0, 1, 2};
  int ordereo[2] = {0, 1};
  if (c->rng_mode == ORC_RNG_REFERENCE) {
    orc_shuffle(&c->noise.g, orderxyz, 3);
    orc_shuffle(&c->noise.g, ordereo, 2);
  } else {
    orc_xoro g;
    orc_xoro_seed(&g, orc_site_key(orc_stream_key(c->base_key, c->sweep_ctr++), 0));
    orc_shuffle(&g, orderxyz, 3);
    orc_shuffle(&g, ordereo, 2);
  }
  memcpy(c->last_order, orderxyz, sizeof orderxyz);
  memcpy(c->last_order + 3, ordereo, sizeof ordereo);
  for (int ixyz = 0; ixyz < 3; ixyz++)
    for (int ieo = 0; ieo < 2; ieo++) {
      uint64_t ctr = 0;
      if (c->rng_mode == ORC_RNG_PER_SITE) ctr = c->sweep_ctr++;
      orc_charge_subsweep(c, u, rdtsigma, orderxyz[ixyz], ordereo[ieo], ctr);
    }
}

/* PV2HBSplit::step, ModelG/Stepper.cxx:542-580 */
void orc_pv2hbsplit_step(orc_ctx *c, double *u, double dt) {
  const unsigned s0 = c->scounts[0], s1 = c->scounts[1];
  if (!c->onlydiffuse) {
    for (unsigned i1 = 0; i1 < s1; i1++) {
      orc_ideal_step(c, u, dt / (s1));
      for (unsigned i0 = 0; i0 < s0; i0++) orc_hb_phi_step(c, u, dt / (s0 * s1));
    }
  }
  if (!c->nodiffuse) orc_charge_step(c, u, dt);
}

/* ModelA::initialize_gaussian_charges, ModelG/ModelA.h:463-525: always the
 * global engine (normal() pairs), in k,j,i,L order, then the zero mode of
 * every charge is removed. */
void orc_init_gaussian_charges(orc_ctx *c, double *u) {
  const int N = c->N;
  const size_t V = (size_t)N * N * N;
  double sum[NDOF] = {0};
  const double chi = c->chi;
  for (size_t s = 0; s < V; s++)
    for (int L = 4; L < NDOF; L++) {
      u[s * NDOF + L] = sqrt(chi) * orc_normal(&c->noise);
      sum[L] += u[s * NDOF + L];
    }
  const double vol = (double)N * N * N;
  for (size_t s = 0; s < V; s++)
    for (int L = 4; L < NDOF; L++) u[s * NDOF + L] -= sum[L] / vol;
  if (c->precision == ORC_FP32) orc_round_fp32(u, V * NDOF);
}

/* thermalize_event, ModelG/SuperPions.cxx:21-76 (no quench): charges are
 * drawn, then 6*nsteps phi sweeps at dt/6 with the clock held. */
void orc_thermalize(orc_ctx *c, double *u, double dt, int nsteps) {
  orc_init_gaussian_charges(c, u);
  for (int i = 0; i < nsteps; i++) {
    const int substeps = 6;
    for (int j = 0; j < substeps; j++) orc_hb_phi_step(c, u, dt / substeps);
  }
}

/* ======================================================================
 * Measurer, ModelG/measurer.cxx
 * ==================================================================== */

/* contract_rho, measurer.cxx:101-108 (double case) */
static void contract_rho(const double *n, const double *rho, double *out) {
  out[0] = -(n[1] * rho[0]) - n[2] * rho[1] - n[3] * rho[2];
  out[1] = n[0] * rho[0] + n[3] * rho[4] - n[2] * rho[5];
  out[2] = n[0] * rho[1] - n[3] * rho[3] + n[1] * rho[5];
  out[3] = n[0] * rho[2] + n[2] * rho[3] - n[1] * rho[4];
}

/* computeSliceAverage :10-98, computeSliceAveragePhase :116-221 and the
 * OAverage part of computeDerivedObs :288-300.  wall[d][obs][n] with
 * d = 0,1,2 for x,y,z. */
void orc_measure_walls(int N, const double *u, double *wall, double *wallphase,
                       double *oavg) {
  const size_t W = (size_t)N;
  memset(wall, 0, 3 * 11 * W * sizeof(double));
  memset(wallphase, 0, 3 * 14 * W * sizeof(double));
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const double *p = u + NDOF * site_index(N, k, j, i);
        const int idx[3] = {i, j, k};
        double phi2 = 0.0;
        for (int l = 0; l < 4; l++) phi2 += p[l] * p[l];
        for (int d = 0; d < 3; d++) {
          for (int l = 0; l < 10; l++) wall[(d * 11 + l) * W + idx[d]] += p[l];
          wall[(d * 11 + 10) * W + idx[d]] += phi2;
        }
        /* phase observables */
        double sigma = sqrt(phi2), n[4], rho[6], drho[6], A[4], Vv[4];
        if (phi2 > 100. * 2.2250738585072014e-308) {
          for (int l = 0; l < 4; l++) n[l] = p[l] / sigma;
        } else {
          n[0] = 1; n[1] = n[2] = n[3] = 0;
        }
        for (int l = 0; l < 3; l++) { rho[l] = p[4 + l]; rho[l + 3] = p[7 + l]; }
        for (int l = 0; l < 3; l++) { drho[l] = rho[l + 3]; drho[l + 3] = rho[l]; }
        contract_rho(n, rho, A);
        contract_rho(n, drho, Vv);
        for (int d = 0; d < 3; d++) {
          double *w = wallphase + d * 14 * W;
          w[0 * W + idx[d]] += sigma;
          for (int l = 0; l < 4; l++) {
            w[(1 + l) * W + idx[d]] += n[l];
            w[(5 + l) * W + idx[d]] += A[l];
            w[(9 + l) * W + idx[d]] += Vv[l];
          }
          w[13 * W + idx[d]] += sigma * sigma;
        }
      }
  const double nn = (double)(N * N);
  for (size_t x = 0; x < 3 * 11 * W; x++) wall[x] /= nn;
  for (size_t x = 0; x < 3 * 14 * W; x++) wallphase[x] /= nn;
  for (int l = 0; l < 11; l++) {
    double s = 0.;
    for (int i = 0; i < N; i++) s += wall[l * W + i];
    oavg[l] = s / (double)N;
  }
}

/* measurer_fft::execute_row, measurer.h:33-41: the r2c transform with
 * FFTW's forward sign, out[k] = (1/N) sum_x in[x] exp(-2 pi i k x / N),
 * k = 0..N/2.  Evaluated as a direct DFT in long double. */
void orc_dft_rows(int N, int rows, const double *in, double *out) {
  const int nk = N / 2 + 1;
  const long double twopi = 6.283185307179586476925286766559L;
  for (int r = 0; r < rows; r++)
    for (int k = 0; k < nk; k++) {
      long double re = 0, im = 0;
      for (int x = 0; x < N; x++) {
        long double ang = twopi * (long double)((long)k * x % N) / N;
        re += in[r * N + x] * cosl(ang);
        im -= in[r * N + x] * sinl(ang);
      }
      out[(r * nk + k) * 2 + 0] = (double)(re / N);
      out[(r * nk + k) * 2 + 1] = (double)(im / N);
    }
}

/* rotate_to_vev, measurer.cxx:231-286 (complex arithmetic written out) */
typedef struct { double re, im; } cplx;
static cplx cmul(cplx a, cplx b) { cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static cplx cneg(cplx a) { cplx r = {-a.re, -a.im}; return r; }
static void contract_rho_c(const cplx *n, const cplx *rho, cplx *out) {
  out[0] = csub(csub(cneg(cmul(n[1], rho[0])), cmul(n[2], rho[1])), cmul(n[3], rho[2]));
  out[1] = csub(cadd(cmul(n[0], rho[0]), cmul(n[3], rho[4])), cmul(n[2], rho[5]));
  out[2] = cadd(csub(cmul(n[0], rho[1]), cmul(n[3], rho[3])), cmul(n[1], rho[5]));
  out[3] = csub(cadd(cmul(n[0], rho[2]), cmul(n[2], rho[3])), cmul(n[1], rho[4]));
}
void orc_rotate_to_vev(int nk, const double *wk, double *rot) {
#define WK(a, k) ((cplx){wk[((a) * nk + (k)) * 2], wk[((a) * nk + (k)) * 2 + 1]})
  cplx n[4];
  double norm = 0.;
  for (int a = 0; a < 4; a++) {
    n[a] = WK(a, 0);
    norm += n[a].re * n[a].re + n[a].im * n[a].im;
  }
  norm = sqrt(norm);
  if (norm < 2.2250738585072014e-308) {
    n[0] = (cplx){1, 0}; n[1] = n[2] = n[3] = (cplx){0, 0};
  } else {
    for (int a = 0; a < 4; a++) { n[a].re /= norm; n[a].im /= norm; }
  }
  for (int k = 0; k < nk; k++) {
    cplx phi[4], rho[6], drho[6], A[4], Vv[4], phir[4];
    for (int a = 0; a < 4; a++) phi[a] = WK(a, k);
    for (int ab = 0; ab < 6; ab++) rho[ab] = WK(4 + ab, k);
    cplx phi2k = WK(10, k);
    cplx phis = {0., 0.};
    for (int a = 0; a < 4; a++) phis = cadd(phis, cmul(n[a], phi[a]));
    for (int a = 0; a < 4; a++) phir[a] = csub(phi[a], cmul(n[a], phis));
    for (int l = 0; l < 3; l++) { drho[l] = rho[l + 3]; drho[l + 3] = rho[l]; }
    contract_rho_c(n, rho, A);
    contract_rho_c(n, drho, Vv);
#define RT(a, k) (rot + ((a) * nk + (k)) * 2)
    RT(0, k)[0] = phis.re; RT(0, k)[1] = phis.im;
    for (int a = 0; a < 4; a++) {
      RT(1 + a, k)[0] = phir[a].re; RT(1 + a, k)[1] = phir[a].im;
      RT(5 + a, k)[0] = A[a].re;    RT(5 + a, k)[1] = A[a].im;
      RT(9 + a, k)[0] = Vv[a].re;   RT(9 + a, k)[1] = Vv[a].im;
    }
    RT(13, k)[0] = phi2k.re; RT(13, k)[1] = phi2k.im;
  }
#undef RT
#undef WK
}

/* ======================================================================
 * Model A Euler--Maruyama stepper (new functionality; the reference has no
 * EM stepper).  phi <- phi - G dt (-Lap phi + m2 phi + lambda |phi|^2 phi
 * - H e0) + sqrt(2 G dt) xi, with the reference's lattice Laplacian
 * (sum of the six neighbours minus 6 phi) and noise xi from per-site
 * counter streams.  Layout phi[site][4].
 * ==================================================================== */
#define EM_TAG 0x4000000000000000ULL

void orc_em_noise(const orc_em_params *p, uint64_t step, double *xi) {
  const size_t V = (size_t)p->N * p->N * p->N;
  const uint64_t skey = orc_stream_key(p->base_key, EM_TAG | step);
  for (size_t s = 0; s < V; s++) {
    orc_xoro g;
    orc_xoro_seed(&g, orc_site_key(skey, s));
    if (p->noise_kind == 0) {
      for (int a = 0; a < 4; a++) xi[4 * s + a] = orc_variance1(&g);
    } else {
      for (int a = 0; a < 4; a += 2) {
        double u1 = orc_canonical(orc_xoro_next(&g));
        double u2 = orc_canonical(orc_xoro_next(&g));
        double r = sqrt(-2.0 * log(1.0 - u1));
        double th = 6.283185307179586 * u2;
        xi[4 * s + a] = r * cos(th);
        xi[4 * s + a + 1] = r * sin(th);
      }
    }
  }
}

void orc_em_step(const orc_em_params *p, uint64_t step, const double *phi, double *out) {
  const int N = p->N;
  const size_t V = (size_t)N * N * N;
  double *xi = (double *)malloc(V * 4 * sizeof(double));
  orc_em_noise(p, step, xi);
  const double gdt = p->gamma * p->dt;
  const double amp = sqrt(2.0 * gdt);
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const size_t s = site_index(N, k, j, i);
        const double *c = phi + 4 * s;
        const double *xp = phi + 4 * site_index(N, k, j, wrap(i + 1, N));
        const double *xm = phi + 4 * site_index(N, k, j, wrap(i - 1, N));
        const double *yp = phi + 4 * site_index(N, k, wrap(j + 1, N), i);
        const double *ym = phi + 4 * site_index(N, k, wrap(j - 1, N), i);
        const double *zp = phi + 4 * site_index(N, wrap(k + 1, N), j, i);
        const double *zm = phi + 4 * site_index(N, wrap(k - 1, N), j, i);
        double s2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
        for (int a = 0; a < 4; a++) {
          double nb = (xp[a] + xm[a]) + (yp[a] + ym[a]) + (zp[a] + zm[a]);
          double lap = nb - 6.0 * c[a];
          double drift = (-lap + p->m2 * c[a]) + (p->lambda * s2) * c[a];
          if (a == 0) drift = drift - p->H;
          out[4 * s + a] = (c[a] - gdt * drift) + amp * xi[4 * s + a];
        }
      }
  free(xi);
}

/* Runs nsteps EM steps from phi0; phibar[n] = volume mean of phi_0 after
 * step n+1.  Returns the loss  mean_w(phibar) + V * var_w(phibar)  over the
 * last `window` steps. */
double orc_em_run(const orc_em_params *p, const double *phi0, int nsteps, int window,
                  double *phibar, double *phi_final) {
  const size_t V = (size_t)p->N * p->N * p->N;
  double *a = (double *)malloc(V * 4 * sizeof(double));
  double *b = (double *)malloc(V * 4 * sizeof(double));
  memcpy(a, phi0, V * 4 * sizeof(double));
  for (int n = 0; n < nsteps; n++) {
    orc_em_step(p, (uint64_t)n, a, b);
    double *t = a; a = b; b = t;
    long double s = 0;
    for (size_t x = 0; x < V; x++) s += a[4 * x];
    phibar[n] = (double)(s / (long double)V);
  }
  if (phi_final) memcpy(phi_final, a, V * 4 * sizeof(double));
  free(a);
  free(b);
  long double m = 0, m2 = 0;
  for (int n = nsteps - window; n < nsteps; n++) { m += phibar[n]; m2 += (long double)phibar[n] * phibar[n]; }
  m /= window;
  m2 /= window;
  return (double)(m + (long double)V * (m2 - m * m));
}

/* The observables of one state (phi[V][4]): the volume means phibar_0..3,
 * the volume mean of |phi|^2 and V |phibar|^2 -- the volume averages a
 * measurement takes (Measurer, measurer.h:169-179), summed in long double. */
static void em_observe(size_t V, const double *phi, double *o) {
  long double s[5] = {0, 0, 0, 0, 0};
  for (size_t x = 0; x < V; x++) {
    const double *c = phi + 4 * x;
    for (int a = 0; a < 4; a++) s[a] += c[a];
    s[4] += ((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]) + c[3] * c[3];
  }
  for (int i = 0; i < 5; i++) o[i] = (double)(s[i] / (long double)V);
  o[5] = (double)V * (((o[0] * o[0] + o[1] * o[1]) + o[2] * o[2]) + o[3] * o[3]);
}

/* The run loop of SuperPions.cxx:124-160 (measure, then step) on the EM
 * stepper: rows[r] = observables of the state after r*every steps, for
 * r*every < nsteps. */
void orc_em_series(const orc_em_params *p, const double *phi0, int nsteps, int every,
                   double *rows /*[ceil(nsteps/every)][6]*/, double *phi_final) {
  const size_t V = (size_t)p->N * p->N * p->N;
  double *a = (double *)malloc(V * 4 * sizeof(double));
  double *b = (double *)malloc(V * 4 * sizeof(double));
  memcpy(a, phi0, V * 4 * sizeof(double));
  for (int n = 0; n < nsteps; n++) {
    if (n % every == 0) em_observe(V, a, rows + (size_t)(n / every) * 6);
    orc_em_step(p, (uint64_t)n, a, b);
    double *t = a; a = b; b = t;
  }
  if (phi_final) memcpy(phi_final, a, V * 4 * sizeof(double));
  free(a);
  free(b);
}

/* Reverse-mode derivative of orc_em_run with the noise held fixed: the
 * Jacobian of one step is symmetric, J = I - G dt (-Lap + m2 + lambda
 * (|phi|^2 + 2 phi phi^T)), so the adjoint step applies the same stencil.
 * All states are kept (small N only). */
void orc_em_adjoint(const orc_em_params *p, const double *phi0, int nsteps, int window,
                    double *grad_phi0, double *grad_params) {
  const int N = p->N;
  const size_t V = (size_t)N * N * N;
  double *states = (double *)malloc((size_t)(nsteps + 1) * V * 4 * sizeof(double));
  double *phibar = (double *)malloc(nsteps * sizeof(double));
  memcpy(states, phi0, V * 4 * sizeof(double));
  for (int n = 0; n < nsteps; n++) {
    orc_em_step(p, (uint64_t)n, states + (size_t)n * V * 4, states + (size_t)(n + 1) * V * 4);
    long double s = 0;
    const double *st = states + (size_t)(n + 1) * V * 4;
    for (size_t x = 0; x < V; x++) s += st[4 * x];
    phibar[n] = (double)(s / (long double)V);
  }
  long double mean = 0;
  for (int n = nsteps - window; n < nsteps; n++) mean += phibar[n];
  mean /= window;
  double *lam = (double *)calloc(V * 4, sizeof(double));
  double *nxt = (double *)malloc(V * 4 * sizeof(double));
  long double gm2 = 0, gH = 0, glam = 0;
  const double gdt = p->gamma * p->dt;
  for (int n = nsteps - 1; n >= 0; n--) {
    /* direct term of the loss on phi_{n+1} */
    if (n >= nsteps - window) {
      double dl = (1.0 / window + (double)V * 2.0 * (phibar[n] - (double)mean) / window) / (double)V;
      for (size_t x = 0; x < V; x++) lam[4 * x] += dl;
    }
    const double *ph = states + (size_t)n * V * 4;
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const size_t s = site_index(N, k, j, i);
          const double *c = ph + 4 * s;
          const double *l = lam + 4 * s;
          const double *xp = lam + 4 * site_index(N, k, j, wrap(i + 1, N));
          const double *xm = lam + 4 * site_index(N, k, j, wrap(i - 1, N));
          const double *yp = lam + 4 * site_index(N, k, wrap(j + 1, N), i);
          const double *ym = lam + 4 * site_index(N, k, wrap(j - 1, N), i);
          const double *zp = lam + 4 * site_index(N, wrap(k + 1, N), j, i);
          const double *zm = lam + 4 * site_index(N, wrap(k - 1, N), j, i);
          double s2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
          double pl = c[0] * l[0] + c[1] * l[1] + c[2] * l[2] + c[3] * l[3];
          for (int a = 0; a < 4; a++) {
            double nb = (xp[a] + xm[a]) + (yp[a] + ym[a]) + (zp[a] + zm[a]);
            double lapl = nb - 6.0 * l[a];
            double jt = (-lapl + p->m2 * l[a]) + p->lambda * (s2 * l[a] + 2.0 * c[a] * pl);
            nxt[4 * s + a] = l[a] - gdt * jt;
          }
          gm2 += -gdt * pl;
          glam += -gdt * s2 * pl;
          gH += gdt * l[0];
        }
    double *t = lam; lam = nxt; nxt = t;
  }
  memcpy(grad_phi0, lam, V * 4 * sizeof(double));
  grad_params[0] = (double)gm2;
  grad_params[1] = (double)gH;
  grad_params[2] = (double)glam;
  free(states); free(phibar); free(lam); free(nxt);
}
