// lgv_device.h -- device (and host-twin) building blocks of the ModelG
// kernels: the xoroshiro128+ engine and libstdc++ conversions, per-site
// counter keys, the red-black structure-of-arrays layout and the O(4)
// rotation.  Compiled with -ffp-contract=off: every floating-point
// expression rounds exactly where the reference's does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define LGV_HD __host__ __device__ __forceinline__

namespace lgv {

// ---------------------------------------------------------------- RNG

struct Xoro {
  uint64_t s0, s1;
};

// Tile (bx, by, bz) of march block b in a grid of exactly ntx x nty x nz
// blocks.  When nty is a multiple of 8, XCD b % 8 (the dispatcher deals
// workgroups round-robin over the 8 XCDs) sweeps the y band of tiles
// nty/8 * (b % 8) .. + nty/8, all x tiles, z chunk after z chunk: the tiles
// that share ring lines run together in one L2, and the eight XCDs stream
// from eight row bands of every plane.  Otherwise x, then y, then z.
__device__ __forceinline__ void yband_tile(int b, int ntx, int nty, int &bx, int &by, int &bz) {
  if ((nty & 7) == 0) {
    const int nyb = nty >> 3, i8 = b >> 3;
    bx = i8 % ntx;
    const int t8 = i8 / ntx;
    by = (b & 7) * nyb + t8 % nyb;
    bz = t8 / nyb;
  } else {
    bx = b % ntx;
    b /= ntx;
    by = b % nty;
    bz = b / nty;
  }
}

// Plan block of workgroup b for a one-thread-per-site launch that covers
// whole planes in bpp blocks each (bpp a multiple of 8): XCD b % 8 takes the
// (b % 8)-th eighth of every plane (a band of rows), plane after plane, so a
// site's y and z neighbours are read through the same L2.
__device__ __forceinline__ int band_block(int b, int bpp) {
  const int S = bpp >> 3, i = b >> 3, c = i / S;
  return (c * 8 + (b & 7)) * S + (i - c * S);
}

LGV_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoroshiro128plus::seed(uint64_t), ModelG/xoroshiro128plus.h:86-91
LGV_HD Xoro xoro_seeded(uint64_t seed) {
  Xoro g;
  g.s0 = 0x01d353e5f3993bb0ULL + seed;
  g.s1 = 0x7b9c0df6cb193b20ULL * (seed + 1);
  return g;
}

// 64-bit rotation right by r < 32 as two 32-bit funnel shifts
// (v_alignbit_b32), one instruction per half.
__device__ __forceinline__ uint64_t rotr64_lt32(uint64_t x, int r) {
  const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  return ((uint64_t)__builtin_amdgcn_alignbit(lo, hi, r) << 32) |
         __builtin_amdgcn_alignbit(hi, lo, r);
}

// xoroshiro128plus::operator(), ModelG/xoroshiro128plus.h:97-107; on the
// device the rotations by 55 and 36 are taken as rotations right by 9 and 28,
// and the three-way xor of the new s0 as one v_bitop3_b32 per half (truth
// table 0x96 = a ^ b ^ c): eleven instructions per step instead of twelve
// (256^3 FP32 charge sweep about -3 %).
LGV_HD uint64_t xoro_next(Xoro &g) {
  const uint64_t s0 = g.s0;
  uint64_t s1 = g.s1;
  const uint64_t result = s0 + s1;
  s1 ^= s0;
#ifdef __HIP_DEVICE_COMPILE__
  const uint64_t r = rotr64_lt32(s0, 9), t = s1 << 14;
  const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)r, (uint32_t)s1, (uint32_t)t, 0x96);
  const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(r >> 32), (uint32_t)(s1 >> 32),
                                                  (uint32_t)(t >> 32), 0x96);
  g.s0 = ((uint64_t)hi << 32) | lo;
  g.s1 = rotr64_lt32(s1, 28);
#else
  g.s0 = rotl64(s0, 55) ^ s1 ^ (s1 << 14);
  g.s1 = rotl64(s1, 36);
#endif
  return result;
}

// generate_canonical<double,53> over a 64-bit engine (libstdc++ 11
// bits/random.tcc:3348-3379): double(x) / 2^64, clamped below one.
// Multiplying by 2^-64 is the same exact scaling as dividing by 2^64.  On
// the device double(x) 2^-64 is formed as hi 2^-32 + lo 2^-64 (hi, lo the
// 32-bit halves) with one rounding -- an FMA whose product and addend are
// exact -- which is the same number, since scaling by a power of two
// commutes with rounding; the clamp is a min.
LGV_HD double canonical(uint64_t x) {
#ifdef __HIP_DEVICE_COMPILE__
  // the empty asm keeps the compiler from re-forming a 64-bit conversion of
  // x >> 32 (which it lowers with an extra add)
  uint32_t h = (uint32_t)(x >> 32);
  __asm__("" : "+v"(h));
  const double hi = (double)h * 0x1p-32;
  return fmin(fma((double)(uint32_t)x, 0x1p-64, hi), 0x1.fffffffffffffp-1);
#else
  double r = (double)x * 0x1p-64;
  return r >= 1.0 ? 0x1.fffffffffffffp-1 : r;
#endif
}

// uniform_real_distribution(-sqrt(3), sqrt(3)): u*(b-a)+a, b-a = 2 sqrt(3)
// (exact doubling); NoiseGenerator::variance1, NoiseGenerator.h:68-69,105
constexpr double kSqrt3 = 0x1.bb67ae8584caap+0;  // sqrt(3.) rounded
LGV_HD double variance1(Xoro &g) {
#ifdef __HIP_DEVICE_COMPILE__
  // The same number with the scaling moved into the multiply: d = double(x)
  // (hi 2^32 + lo, an FMA whose product and addend are exact, one
  // rounding), clamped at (1 - 2^-53) 2^64, times 2 sqrt(3) 2^-64 (exact):
  // d 2^-64 is canonical(x) exactly and a product's rounding commutes with
  // scaling by a power of two (no value here is subnormal).
  const uint64_t x = xoro_next(g);
  uint32_t h = (uint32_t)(x >> 32);
  __asm__("" : "+v"(h));
  const double d = fmin(fma((double)h, 0x1p32, (double)(uint32_t)x), 0x1.fffffffffffffp63);
  return d * ((kSqrt3 + kSqrt3) * 0x1p-64) + (-kSqrt3);
#else
  const double u = canonical(xoro_next(g));
  return u * (kSqrt3 + kSqrt3) + (-kSqrt3);
#endif
}
// uniform_real_distribution(0, 1): u*(1-0)+0 == u; NoiseGenerator::uniform
LGV_HD double uniform01(Xoro &g) { return canonical(xoro_next(g)); }

// Lemire downscaling of uniform_int_distribution{0,b} with a 64-bit engine,
// libstdc++ 11 bits/uniform_int_dist.h:246-307
LGV_HD uint64_t uniform_int(Xoro &g, uint64_t b) {
  const uint64_t range = b + 1;
  uint64_t x = xoro_next(g);
  uint64_t low = x * range;
#ifdef __HIP_DEVICE_COMPILE__
  uint64_t high = __umul64hi(x, range);
#else
  uint64_t high = (uint64_t)(((unsigned __int128)x * range) >> 64);
#endif
  if (low < range) {
    const uint64_t threshold = (0 - range) % range;
    while (low < threshold) {
      x = xoro_next(g);
      low = x * range;
#ifdef __HIP_DEVICE_COMPILE__
      high = __umul64hi(x, range);
#else
      high = (uint64_t)(((unsigned __int128)x * range) >> 64);
#endif
    }
  }
  return high;
}

// std::shuffle of {0,1,2} then {0,1} (libstdc++ 11 bits/stl_algo.h:
// 3731-3792, the two-positions-per-draw path); ModelGChargeHB::step,
// ModelG/Stepper.cxx:461-466.  order[0..2] = directions, order[3..4] = parities.
LGV_HD void shuffle_orders(Xoro &g, int order[5]) {
  order[0] = 0; order[1] = 1; order[2] = 2; order[3] = 0; order[4] = 1;
  {
    const uint64_t x = uniform_int(g, 2 * 3 - 1);
    const int p0 = (int)(x / 3), p1 = (int)(x % 3);
    int t = order[1]; order[1] = order[p0]; order[p0] = t;
    t = order[2]; order[2] = order[p1]; order[p1] = t;
  }
  {
    const int d = (int)uniform_int(g, 1);
    int t = order[3 + 1]; order[3 + 1] = order[3 + d]; order[3 + d] = t;
  }
}

// Per-site counter streams.  splitmix64's output function (Steele, Lea,
// Flood 2014) mixes a stream counter and then a site index into a 64-bit
// key; xoroshiro128plus::seed(key) starts that site's stream.
LGV_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
LGV_HD uint64_t base_key(uint32_t seed32) {
  return mix64((uint64_t)seed32 ^ 0x6a09e667f3bcc909ULL);
}
LGV_HD uint64_t stream_key(uint64_t base, uint64_t ctr) {
  return mix64(base + mix64(ctr + 0x9e3779b97f4a7c15ULL));
}
LGV_HD uint64_t site_key(uint64_t skey, uint64_t site) {
  return mix64(skey ^ mix64(site + 0x3c6ef372fe94f82bULL));
}

// ---------------------------------------------------------------- layout
//
// A replica of side N (even) stores component c (0..3 phi, 4..6 A, 7..9 V)
// of the site (x, y, z) with parity p = (x+y+z)&1 at
//   base + (c*2 + p) * half + (z*N + y) * (N/2) + x/2,   half = N^3/2.
// A half-sweep therefore reads one half-lattice and writes the other
// densely, and x-neighbours sit in the adjacent or the same column.

LGV_HD int site_x(int xh, int y, int z, int p) { return 2 * xh + ((p + y + z) & 1); }

// ---------------------------------------------------------------- O(4)
// IdealPV2::rotatePhi's per-site work (Stepper.cxx:186-198) with
// O4Rotation, expmul and quatmul (O4AlgebraHelper.h:54-97): with
// mu = -charge/chi * tau, phi <- L (phi R), R = expmul(-(mu_V - mu_A)/2),
// L = expmul((mu_V + mu_A)/2).
// With k = tau/chi the two generators are tmp1 = -0.5(mu_V - mu_A) =
// h (V - A) and tmp2 = 0.5(mu_V + mu_A) = -h (V + A), h = k/2.  expmul needs
// cos|v| and sin|v|/|v|; for |v| < 1/2 both are evaluated as Taylor series
// in y = |v|^2 (truncation below 1e-18; degree 4 below y = 2^-10, where
// the charges of a thermal lattice put almost every site), otherwise with
// sqrt/sincos.  Agrees
// with the reference's expression order to a few ulps (the reference's own
// libm calls are not bit-reproducible on the device anyway).
// The ideal step's arithmetic is held to 1e-12, not to bits, so these
// functions let the compiler contract a*b+c into fused multiply-adds (the
// rest of the library is built with -ffp-contract=off).
LGV_HD void cos_sinc(double y, double &c, double &s) {
#pragma clang fp contract(fast)
  if (__builtin_expect(y < 0x1p-10, 1)) {
    // the first omitted terms, y^5/10! and y^5/11!, are below 3e-22
    c = 1.0 + y * (-1.0 / 2 + y * (1.0 / 24 + y * (-1.0 / 720 + y * (1.0 / 40320))));
    s = 1.0 + y * (-1.0 / 6 + y * (1.0 / 120 + y * (-1.0 / 5040 + y * (1.0 / 362880))));
  } else if (__builtin_expect(y < 0.25, 1)) {
    c = 1.0 + y * (-1.0 / 2 + y * (1.0 / 24 + y * (-1.0 / 720 + y * (1.0 / 40320 +
        y * (-1.0 / 3628800 + y * (1.0 / 479001600 + y * (-1.0 / 87178291200.0)))))));
    s = 1.0 + y * (-1.0 / 6 + y * (1.0 / 120 + y * (-1.0 / 5040 + y * (1.0 / 362880 +
        y * (-1.0 / 39916800 + y * (1.0 / 6227020800.0 + y * (-1.0 / 1307674368000.0)))))));
  } else {
    const double a = sqrt(y);
    double sn, cs;
    sincos(a, &sn, &cs);
    c = cs;
    s = sn / a;
  }
}

// The quaternion product of quatmul with contraction allowed.
LGV_HD void quatmul_fma(const double v1[4], const double v2[4], double w[4]) {
#pragma clang fp contract(fast)
  w[0] = v1[0] * v2[0] - (v1[1] * v2[1] + v1[2] * v2[2] + v1[3] * v2[3]);
  w[1] = v1[0] * v2[1] + v2[0] * v1[1] + (v1[2] * v2[3] - v1[3] * v2[2]);
  w[2] = v1[0] * v2[2] + v2[0] * v1[2] + (v1[3] * v2[1] - v1[1] * v2[3]);
  w[3] = v1[0] * v2[3] + v2[0] * v1[3] + (v1[1] * v2[2] - v1[2] * v2[1]);
}

LGV_HD void rotate_site_fast(double phi[4], const double A[3], const double V[3], double h) {
#pragma clang fp contract(fast)
  double r[3], l[3];
  for (int i = 0; i < 3; i++) {
    r[i] = h * (V[i] - A[i]);
    l[i] = -h * (V[i] + A[i]);
  }
  double cR, sR, cL, sL;
  cos_sinc(r[0] * r[0] + r[1] * r[1] + r[2] * r[2], cR, sR);
  cos_sinc(l[0] * l[0] + l[1] * l[1] + l[2] * l[2], cL, sL);
  const double R[4] = {cR, -sR * r[0], -sR * r[1], -sR * r[2]};
  const double L[4] = {cL, -sL * l[0], -sL * l[1], -sL * l[2]};
  double t[4];
  quatmul_fma(phi, R, t);
  quatmul_fma(L, t, phi);
}

// ---------------------------------------------------------------- sums
// Compensated (TwoSum / Neumaier) accumulation for FP64 reductions.
struct KSum {
  double s, c;
};
LGV_HD void ksum_add(KSum &k, double x) {
  const double t = k.s + x;
  if (fabs(k.s) >= fabs(x)) k.c += (k.s - t) + x;
  else k.c += (x - t) + k.s;
  k.s = t;
}
LGV_HD void ksum_merge(KSum &a, const KSum &b) {
  ksum_add(a, b.s);
  a.c += b.c;
}
LGV_HD double ksum_value(const KSum &k) { return k.s + k.c; }

}  // namespace lgv
