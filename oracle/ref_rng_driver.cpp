// ref_rng_driver.cpp -- prints known-answer vectors from the reference's own
// random-number engine (ModelG/xoroshiro128plus.h, compiled from where it
// lies under the reference tree) driven through libstdc++'s distributions
// exactly as ModelG/NoiseGenerator.h:62-157 drives them.  NoiseGenerator.h
// itself needs PETSc headers and is not compiled; its constructor
// (seed_seq{base} -> seeds[rank] -> engine.seed) and its three
// distributions are spelled out here with the same standard-library calls.
//
// TEST INFRASTRUCTURE ONLY (builds into oracle/_ref/, never shipped).
//
// usage:
//   ref_rng engine  <seed> <n>            xoroshiro128plus::seed(seed), n outputs
//   ref_rng seedseq <base> <n>            std::seed_seq{base}.generate(n words)
//   ref_rng noise   <base> <pattern>      NoiseGenerator(base); pattern chars
//                                         u=uniform v=variance1 n=normal
//   ref_rng shuffle <s0> <s1> <count>     engine state "s0 s1"; count x
//                                         (shuffle{0,1,2}, shuffle{0,1})
//   ref_rng canon   <x>...                generate_canonical on given words
//   ref_rng state   <base> <ndraws>       NoiseGenerator(base), ndraws
//                                         uniforms, then the .rng text
#include "xoroshiro128plus.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

struct noise_twin {
  xoroshiro128plus rng;
  std::normal_distribution<double> normalDistribution;
  std::uniform_real_distribution<double> uniformDistributionRt3{-sqrt(3.), sqrt(3.)};
  std::uniform_real_distribution<double> uniformDistribution;
  explicit noise_twin(int baseSeed) {
    std::seed_seq sequence = {baseSeed};
    std::vector<unsigned int> seeds(1, 0);
    sequence.generate(seeds.begin(), seeds.end());
    rng.seed(seeds[0]);
  }
};

struct given_words {
  typedef uint64_t result_type;
  static constexpr uint64_t min() { return 0; }
  static constexpr uint64_t max() { return ~0ULL; }
  std::vector<uint64_t> w;
  size_t i = 0;
  uint64_t operator()() { return w[i++]; }
};

static void print_double(double x) { std::printf("%a\n", x); }

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::string cmd = argv[1];
  if (cmd == "engine" && argc == 4) {
    xoroshiro128plus g;
    g.seed(std::strtoull(argv[2], nullptr, 0));
    long n = std::atol(argv[3]);
    for (long i = 0; i < n; i++) std::printf("%llu\n", (unsigned long long)g());
    return 0;
  }
  if (cmd == "seedseq" && argc == 4) {
    std::seed_seq seq = {std::atoi(argv[2])};
    std::vector<unsigned int> out(std::atoi(argv[3]));
    seq.generate(out.begin(), out.end());
    for (auto v : out) std::printf("%u\n", v);
    return 0;
  }
  if (cmd == "noise" && argc == 4) {
    noise_twin ng(std::atoi(argv[2]));
    for (const char *p = argv[3]; *p; ++p) {
      if (*p == 'u') print_double(ng.uniformDistribution(ng.rng));
      else if (*p == 'v') print_double(ng.uniformDistributionRt3(ng.rng));
      else if (*p == 'n') print_double(ng.normalDistribution(ng.rng));
    }
    return 0;
  }
  if (cmd == "shuffle" && argc == 5) {
    xoroshiro128plus g;
    std::istringstream in(std::string(argv[2]) + " " + argv[3]);
    in >> g;
    long n = std::atol(argv[4]);
    for (long i = 0; i < n; i++) {
      std::array<int, 3> xyz{0, 1, 2};
      std::array<int, 2> eo{0, 1};
      std::shuffle(xyz.begin(), xyz.end(), g);
      std::shuffle(eo.begin(), eo.end(), g);
      std::printf("%d %d %d %d %d\n", xyz[0], xyz[1], xyz[2], eo[0], eo[1]);
    }
    return 0;
  }
  if (cmd == "canon") {
    given_words w;
    for (int i = 2; i < argc; i++) w.w.push_back(std::strtoull(argv[i], nullptr, 0));
    for (int i = 2; i < argc; i++)
      print_double(std::generate_canonical<double, std::numeric_limits<double>::digits>(w));
    return 0;
  }
  if (cmd == "state" && argc == 4) {
    noise_twin ng(std::atoi(argv[2]));
    long n = std::atol(argv[3]);
    for (long i = 0; i < n; i++) ng.uniformDistribution(ng.rng);
    std::cout << ng.rng << std::endl;
    return 0;
  }
  return 2;
}
