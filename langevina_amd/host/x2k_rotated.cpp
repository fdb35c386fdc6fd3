// x2k_rotated.cpp -- offline rotation of a Fourier-transformed wall dataset
// into the frame of the vacuum expectation value (the tool of
// ModelG/x2k_rotated.cxx:35-150, same command line and output layout):
//
//   x2k_rotated.exe foo_out.h5 wally_k
//
// reads foo_out.h5:/wally_k [T, F >= 10, K, 2] (fields phi_0..3, A_1..3,
// V_1..3, ...; complex as (re, im)) and writes foo_out.h5:/wally_k_rotated
// [T, 13, K, 2].  Per row t the direction n = phi_a(k=0).re / |..| (n =
// (1,0,0,0) when the norm is below DBL_MIN) gives, for every k,
//
//   field 0       phi_s = sum_a n_a phi_a
//   fields 1..4   phi_r = phi - n phi_s                  (the part normal to n)
//   fields 5..8   A'_a  = sum_b rho_ab n_b
//   fields 9..12  V'_a  = sum_b (*rho)_ab n_b
//
// where rho is the antisymmetric charge tensor with rho_i0 = A_i,
// rho_ij = -eps_ijk V_k (the n_ab convention of measurer.cxx:101-114), and
// *rho is the same tensor with A and V exchanged.
#include <cfloat>
#include <cmath>
#include <complex>
#include <iostream>
#include <stdexcept>
#include <string>

#include "h5io.hpp"

using namespace lgv_host;
using cplx = std::complex<double>;

namespace {

// sum_b M_ab n_b for the antisymmetric M with M_i0 = a_i, M_ij = -eps_ijk v_k
void contract(const double n[4], const cplx a[3], const cplx v[3], cplx out[4]) {
  cplx M[4][4];
  for (int i = 0; i < 3; i++) {
    M[0][0] = 0.;
    M[i + 1][0] = a[i];
    M[0][i + 1] = -a[i];
  }
  for (int i = 0; i < 3; i++) {
    M[i + 1][i + 1] = 0.;
    const int j = (i + 1) % 3, k = (i + 2) % 3;  // (i, j, k) cyclic: eps_ijk = +1
    M[i + 1][j + 1] = -v[k];
    M[j + 1][i + 1] = v[k];
  }
  for (int r = 0; r < 4; r++) {
    cplx s = 0.;
    for (int c = 0; c < 4; c++) s += M[r][c] * n[c];
    out[r] = s;
  }
}

int run(const std::string &fname, const std::string &dsetin) {
  const std::string dsetout = dsetin + "_rotated";
  std::cout << "Creating dataset " << dsetout << " in " << fname << std::endl;

  H5File file(fname, H5File::Existing{}, true);
  H5Api &h = H5Api::get();
  if (h.Lexists(file.id, dsetin.c_str(), H5P_DEFAULT) <= 0) {
    std::cout << "The input dataset " << dsetin << " does not exist" << std::endl;
    return 1;
  }
  Ntuple<3> in(dsetin, file.id);
  const auto N = in.getN();  // {fields, K, 2}
  if (N[0] < 10 || N[2] != 2)
    throw std::runtime_error("dataset " + dsetin + " is not [T, F>=10, K, 2]");
  if (h.Lexists(file.id, dsetout.c_str(), H5P_DEFAULT) > 0 &&
      h.Ldelete(file.id, dsetout.c_str(), H5P_DEFAULT) < 0)
    throw std::runtime_error("cannot remove the old " + dsetout);

  const size_t K = N[1];
  Ntuple<3> out({13, K, 2}, dsetout, file.id);
  auto put = [&](size_t f, size_t k, cplx z) {
    out.Row({f, k, 0}) = z.real();
    out.Row({f, k, 1}) = z.imag();
  };
  for (size_t t = 0; t < in.nrows(); t++) {
    in.readrow(t);
    double n[4], norm2 = 0.;
    for (size_t a = 0; a < 4; a++) {
      n[a] = in.Row({a, 0, 0});
      norm2 += n[a] * n[a];
    }
    const double norm = std::sqrt(norm2);
    if (norm < DBL_MIN) {
      n[0] = 1.;
      n[1] = n[2] = n[3] = 0.;
    } else {
      for (double &c : n) c /= norm;
    }
    for (size_t k = 0; k < K; k++) {
      cplx phi[4], a[3], v[3];
      for (size_t c = 0; c < 4; c++) phi[c] = cplx(in.Row({c, k, 0}), in.Row({c, k, 1}));
      for (size_t c = 0; c < 3; c++) {
        a[c] = cplx(in.Row({4 + c, k, 0}), in.Row({4 + c, k, 1}));
        v[c] = cplx(in.Row({7 + c, k, 0}), in.Row({7 + c, k, 1}));
      }
      cplx phis = 0.;
      for (size_t c = 0; c < 4; c++) phis += n[c] * phi[c];
      put(0, k, phis);
      for (size_t c = 0; c < 4; c++) put(1 + c, k, phi[c] - n[c] * phis);
      cplx Ar[4], Vr[4];
      contract(n, a, v, Ar);
      contract(n, v, a, Vr);
      for (size_t c = 0; c < 4; c++) {
        put(5 + c, k, Ar[c]);
        put(9 + c, k, Vr[c]);
      }
    }
    out.fill();
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const char *usage = "Usage x2k_rotated filename_out.h5 datasetname";
  if (argc < 3) {
    std::cout << usage << std::endl;
    return 0;
  }
  const std::string fname(argv[1]);
  if (fname.rfind("_out.h5") == std::string::npos) {
    std::cout << usage << std::endl;
    return 0;
  }
  try {
    return run(fname, argv[2]);
  } catch (const std::exception &e) {
    std::cerr << "x2k_rotated: " << e.what() << std::endl;
    return 1;
  }
}
