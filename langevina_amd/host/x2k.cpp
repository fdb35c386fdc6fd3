// x2k.cpp -- offline Fourier transform of a wall dataset (the tool of
// ModelG/x2k.cxx:43-121, same command line and output layout):
//
//   x2k.exe foo.h5 wally
//
// reads foo.h5:/wally [T, F, N] (row t, field f, lattice site x) and writes
// foo_out.h5:/wally_k [T, F, N/2+1, 2] with
//
//   wally_k[t, f, k, 0] + i wally_k[t, f, k, 1] = (1/N) sum_x e^{-2 pi i k x / N} wally[t, f, x],
//
// k = 0 .. N/2 (FFTW's r2c convention, divided by N).  foo_out.h5 is opened
// read-write if it exists, created otherwise; an existing wally_k in it is
// replaced.  ModelG/x2k.py runs this program for wallx, wally and wallz of
// every event file, then x2k_rotated.exe on the _out.h5 file.
//
// The transform is a direct sum with twiddles e^{-2 pi i m/N} tabulated at the
// exact index m = k x mod N (accuracy ~N ulp of the row's largest value): rows
// are short (N <= 1536) and the job is file-bound, so no FFT is needed.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "h5io.hpp"

using namespace lgv_host;

namespace {

class RealDFT {
 public:
  explicit RealDFT(size_t n) : n_(n), c_(n), s_(n) {
    const double w = 2.0 * M_PI / (double)n;
    for (size_t m = 0; m < n; m++) {
      c_[m] = std::cos(w * (double)m);
      s_[m] = std::sin(w * (double)m);
    }
  }
  size_t nk() const { return n_ / 2 + 1; }
  // re/im[k] = (1/N) sum_x e^{-2 pi i k x/N} in[x], k = 0 .. N/2
  void operator()(const double *in, double *re, double *im) const {
    for (size_t k = 0; k < nk(); k++) {
      double a = 0., b = 0.;
      size_t m = 0;  // k x mod N
      for (size_t x = 0; x < n_; x++) {
        a += in[x] * c_[m];
        b -= in[x] * s_[m];
        m += k;
        if (m >= n_) m -= n_;
      }
      re[k] = a / (double)n_;
      im[k] = b / (double)n_;
    }
  }

 private:
  size_t n_;
  std::vector<double> c_, s_;
};

int run(const std::string &fin, const std::string &dsetin) {
  const size_t dot = fin.find_last_of('.');
  if (dot == std::string::npos) {
    std::cerr << "x2k: input file name " << fin << " has no extension" << std::endl;
    return 1;
  }
  const std::string fout = fin.substr(0, dot) + "_out.h5";
  const std::string dsetout = dsetin + "_k";
  std::cout << "Creating dataset " << dsetout << " in " << fout << std::endl;

  H5File in_file(fin, H5File::Existing{}, false);
  Ntuple<2> in(dsetin, in_file.id);
  const auto N = in.getN();  // {fields, sites}
  if (N[1] == 0) throw std::runtime_error("dataset " + dsetin + " has no lattice sites");

  H5File out_file(fout, true);
  H5Api &h = H5Api::get();
  if (h.Lexists(out_file.id, dsetout.c_str(), H5P_DEFAULT) > 0 &&
      h.Ldelete(out_file.id, dsetout.c_str(), H5P_DEFAULT) < 0)
    throw std::runtime_error("cannot remove the old " + dsetout + " from " + fout);

  RealDFT dft(N[1]);
  const size_t nk = dft.nk();
  Ntuple<3> out({N[0], nk, 2}, dsetout, out_file.id);
  std::vector<double> re(nk), im(nk);
  for (size_t t = 0; t < in.nrows(); t++) {
    in.readrow(t);
    for (size_t f = 0; f < N[0]; f++) {
      dft(&in.row[in.at({f, 0})], re.data(), im.data());
      for (size_t k = 0; k < nk; k++) {
        out.Row({f, k, 0}) = re[k];
        out.Row({f, k, 1}) = im[k];
      }
    }
    out.fill();
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cout << "Usage x2k filename.h5 datasetname" << std::endl;
    return 0;
  }
  try {
    return run(argv[1], argv[2]);
  } catch (const std::exception &e) {
    std::cerr << "x2k: " << e.what() << std::endl;
    return 1;
  }
}
