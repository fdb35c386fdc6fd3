// superpions.cpp -- the run driver of ModelG/SuperPions.cxx over the HIP
// hot path: reads the same input.json, thermalizes, runs the
// measure-then-step loop, writes the same HDF5 measurement file and the
// same restart files (tag_save.h5:/o4fields and tag_<rank>.rng).
//
//   lgv_superpions -input input.json [-quit] [-log_view] [-device N]
//
// Differences that follow from the design: one process drives the whole
// lattice on one GPU (rank 0 only, so the .rng file is tag_0.rng), and the
// per-site random streams need the stream counter, which is stored next to
// it in tag_0.lgvctr.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "h5io.hpp"
#include "lgv.h"
#include "lgv_stepper.hpp"
#include "modeldata.hpp"

using namespace lgv_host;

namespace {

struct Model {
  ModelAData &data;
  lgv_ctx *ctx = nullptr;
  int N;
  explicit Model(ModelAData &d) : data(d), N(d.NX) {}
};

// Measurer + measurer_output_fasthdf5 (measurer.h:88-194,
// measurer_output.cxx:10-130): datasets phi[T,11], timeout[T,2],
// wall{x,y,z}[.,11,N] and wall{x,y,z}_phase[.,14,N] (created, never
// filled, as in the reference), wall{x,y,z}_k[T,11,N/2+1,2],
// wall{x,y,z}_k_rotated[T,14,N/2+1,2], wall{x,y,z}_phase_k[T,14,N/2+1,2].
class MeasurerOutput {
  Model &m;
  std::unique_ptr<H5File> file;
  std::unique_ptr<Ntuple<1>> scalars, timeout;
  std::unique_ptr<Ntuple<2>> wall[3], wall_phase[3];
  std::unique_ptr<Ntuple<3>> wall_k[3], wall_k_rot[3], wall_phase_k[3];

 public:
  MeasurerOutput(Model &model, const std::string &fname, bool append) : m(model) {
    file.reset(new H5File(fname, append));
    const size_t N = m.N, nk = N / 2 + 1;
    scalars.reset(new Ntuple<1>({11}, "phi", file->id));
    timeout.reset(new Ntuple<1>({2}, "timeout", file->id));
    const char *xyz[3] = {"x", "y", "z"};
    for (int d = 0; d < 3; d++)
      wall[d].reset(new Ntuple<2>({11, N}, std::string("wall") + xyz[d], file->id));
    for (int d = 0; d < 3; d++)
      wall_k[d].reset(new Ntuple<3>({11, nk, 2}, std::string("wall") + xyz[d] + "_k", file->id));
    for (int d = 0; d < 3; d++)
      wall_k_rot[d].reset(
          new Ntuple<3>({14, nk, 2}, std::string("wall") + xyz[d] + "_k_rotated", file->id));
    for (int d = 0; d < 3; d++)
      wall_phase[d].reset(new Ntuple<2>({14, N}, std::string("wall") + xyz[d] + "_phase", file->id));
    for (int d = 0; d < 3; d++)
      wall_phase_k[d].reset(
          new Ntuple<3>({14, nk, 2}, std::string("wall") + xyz[d] + "_phase_k", file->id));
  }

  void measure_and_save() {
    const size_t N = m.N, nk = N / 2 + 1;
    std::vector<double> oavg(11), wk(3 * 11 * nk * 2), wkr(3 * 14 * nk * 2),
        wpk(3 * 14 * nk * 2);
    lgv_measurement meas{nullptr, nullptr, oavg.data(), wk.data(), wkr.data(), wpk.data()};
    check(lgv_measure_full(m.ctx, 0, &meas), "lgv_measure_full");
    timeout->row[0] = m.data.atime.t();
    timeout->row[1] = m.data.mass();
    timeout->fill();
    scalars->row = oavg;
    scalars->fill();
    for (int d = 0; d < 3; d++) {
      std::copy(wk.begin() + d * 11 * nk * 2, wk.begin() + (d + 1) * 11 * nk * 2,
                wall_k[d]->row.begin());
      wall_k[d]->fill();
    }
    for (int d = 0; d < 3; d++) {
      std::copy(wkr.begin() + d * 14 * nk * 2, wkr.begin() + (d + 1) * 14 * nk * 2,
                wall_k_rot[d]->row.begin());
      wall_k_rot[d]->fill();
    }
    for (int d = 0; d < 3; d++) {
      std::copy(wpk.begin() + d * 14 * nk * 2, wpk.begin() + (d + 1) * 14 * nk * 2,
                wall_phase_k[d]->row.begin());
      wall_phase_k[d]->fill();
    }
  }
};

std::string rng_name(const std::string &tag) { return tag + "_0.rng"; }
std::string ctr_name(const std::string &tag) { return tag + "_0.lgvctr"; }

// NoiseGenerator::write (NoiseGenerator.h:112-128): "s0 s1"
void write_rng(Model &m) {
  lgv_rng_state st;
  check(lgv_rng_get_state(m.ctx, 0, &st), "lgv_rng_get_state");
  const std::string tag = m.data.ahandler.outputfiletag;
  std::ofstream out(rng_name(tag));
  if (out) {
    out << st.s0 << ' ' << st.s1;
  } else {
    std::cout << "urng::write Unable to open file for writing in URNG. Defaulting to seed."
              << std::endl;
  }
  std::ofstream c(ctr_name(tag));
  c << st.sweep_ctr << std::endl;
}

// NoiseGenerator::read (NoiseGenerator.h:132-147)
void read_rng(Model &m) {
  lgv_rng_state st;
  check(lgv_rng_get_state(m.ctx, 0, &st), "lgv_rng_get_state");
  const std::string tag = m.data.ahandler.outputfiletag;
  std::ifstream in(rng_name(tag));
  if (in) {
    in >> st.s0 >> st.s1;
    st.has_saved_normal = 0;
  } else {
    std::cout << "urng::read Unable to open file for reading in URNG. Defaulting to seed."
              << std::endl;
  }
  std::ifstream c(ctr_name(tag));
  if (c) c >> st.sweep_ctr;
  check(lgv_rng_set_state(m.ctx, 0, &st), "lgv_rng_set_state");
}

void write_fields(Model &m, const std::string &stub) {
  std::vector<double> u((size_t)m.N * m.N * m.N * 10);
  check(lgv_get_field(m.ctx, 0, u.data()), "lgv_get_field");
  write_o4fields(stub + "_save.h5", m.N, u.data());
}

// thermalize_event, SuperPions.cxx:21-76
lgv_monitor thermalize_event(Model &model) {
  const auto &ahandler = model.data.ahandler;
  auto &atime = model.data.atime;
  auto &co = model.data.acoefficients;
  const double mass0 = co.mass0, dmassdt = co.dmassdt;
  if (ahandler.quench_mode) {
    co.mass0 = ahandler.quench_mode_mass0;
    co.dmassdt = 0.;
    check(lgv_set_mass(model.ctx, -1, co.mass0, co.dmassdt), "lgv_set_mass");
    printf("Settinng up a quench initial condition with initial mass %e\n", co.mass0);
  }
  int nsteps = static_cast<int>(ahandler.thermalization_time / atime.dt());
  printf("Thermalizing event %d\n", ahandler.current_event);
  HipEulerLangevinHB thermalizer(model.ctx, &atime.time);
  lgv_monitor before[3];
  check(lgv_monitor_get(model.ctx, 0, before), "lgv_monitor_get");
  thermalizer.baseline[LGV_MON_PHI] = before[LGV_MON_PHI];
  check(lgv_init_gaussian_charges(model.ctx, 0), "lgv_init_gaussian_charges");
  for (int i = 0; i < nsteps; i++) {
    const int substeps = 6;
    for (int j = 0; j < substeps; j++) thermalizer.step(atime.dt() / substeps);
    printf("Thermalizing Event/Timestep %d/%d: step size = %g, time = %g, nsteps to "
           "thermalize = %d, mass %e \n",
           ahandler.current_event, i, (double)atime.dt(), (double)atime.t(), nsteps,
           model.data.mass());
  }
  const lgv_monitor therm = thermalizer.monitor(LGV_MON_PHI);
  thermalizer.finalize();
  if (ahandler.quench_mode) {
    printf("Finalizing  quench initial condition with initial mass %e\n", co.mass0);
    co.mass0 = mass0;
    co.dmassdt = dmassdt;
    check(lgv_set_mass(model.ctx, -1, co.mass0, co.dmassdt), "lgv_set_mass");
    printf("and final initial mass %e\n", co.mass0);
  }
  return therm;
}

// run_event, SuperPions.cxx:78-161
void run_event(Model &model, HipStepperBase &step, lgv_monitor &therm_total) {
  const auto &ahandler = model.data.ahandler;
  auto &atime = model.data.atime;
  atime.reset();
  check(lgv_set_time(model.ctx, -1, atime.t()), "lgv_set_time");
  if (!ahandler.restart) {
    const lgv_monitor t = thermalize_event(model);
    therm_total.down_counts += t.down_counts;
    therm_total.up_counts += t.up_counts;
    therm_total.up_yes += t.up_yes;
    therm_total.up_no += t.up_no;
    therm_total.down_dS += t.down_dS;
    therm_total.up_dS += t.up_dS;
    step.baseline[LGV_MON_PHI] = therm_total;
  }
  int steps = 0;
  std::string filename;
  if (ahandler.eventmode) {
    std::stringstream ns;
    ns << ahandler.outputfiletag << "_" << std::setw(4) << std::setfill('0')
       << ahandler.current_event << ".h5";
    filename = ns.str();
  } else {
    filename = ahandler.outputfiletag + ".h5";
  }
  MeasurerOutput out(model, filename, ahandler.restart);
  const double tiny = 1.e-10;
  while (atime.t() < atime.tfinal() - tiny) {
    if (steps % ahandler.saveFrequency == 0) {
      out.measure_and_save();
      printf("Event/Timestep %d/%d: step size = %g, time = %g, final = %g, mass = %e\n",
             ahandler.current_event, steps, (double)atime.dt(), (double)atime.t(),
             (double)atime.tfinal(), model.data.mass());
    }
    if (ahandler.writeFrequency > 0 && steps % ahandler.writeFrequency == 0) {
      std::ostringstream tString;
      tString << std::setprecision(4) << "_t_" << atime.t();
      write_fields(model, ahandler.outputfiletag + tString.str());
    }
    step.step(atime.dt());
    steps++;
    atime.time += atime.dt();
  }
}

}  // namespace

int main(int argc, char **argv) {
  std::string input;
  bool quit = false, log_view = false;
  int device = 0;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "-input") && i + 1 < argc) input = argv[++i];
    else if (!strcmp(argv[i], "-quit")) quit = true;
    else if (!strcmp(argv[i], "-log_view")) log_view = true;
    else if (!strcmp(argv[i], "-device") && i + 1 < argc) device = atoi(argv[++i]);
  }
  std::ifstream ifs(input);
  if (!ifs) {
    printf("Unable to open input file %s. Aborting...\n", input.c_str());
    return 0;
  }
  std::stringstream buf;
  buf << ifs.rdbuf();
  Json inputs;
  try {
    inputs = JsonReader(buf.str()).parse();
  } catch (const std::exception &e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  printf("Current version: langevina_amd lgv %d\n", lgv_version());
  std::unique_ptr<ModelAData> pdata;
  try {
    pdata.reset(new ModelAData(inputs));
  } catch (const std::exception &e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  ModelAData &data = *pdata;
  if (quit) return 0;

  const std::string &etype = data.ahandler.evolverType;
  int nodiffuse = 0, onlydiffuse = 0;
  if (etype == "PV2HBSplit23") {
  } else if (etype == "PV2HBSplit23NoDiffuse") {
    nodiffuse = 1;
  } else if (etype == "PV2HBSplit23OnlyDiffuse") {
    onlydiffuse = 1;
  } else {
    printf("Unrecognized stepper type %s. Aborting...\n", etype.c_str());
    return 0;
  }

  // ModelA::ModelA: grid + NoiseGenerator(seed), ModelA.h:311-336
  Model model(data);
  lgv_model_params mp{};
  mp.NX = data.NX;
  mp.LX = (int)data.LX;
  mp.mass0 = data.acoefficients.mass0;
  mp.dmassdt = data.acoefficients.dmassdt;
  mp.lambda = data.acoefficients.lambda;
  mp.H = data.acoefficients.H;
  mp.chi = data.acoefficients.chi;
  mp.gamma = data.acoefficients.gamma;
  mp.diffusion = data.acoefficients.diffusion;
  mp.seed = data.ahandler.seed;
  mp.initialtime = data.atime.initialtime;
  lgv_config cfg{};
  cfg.nreplicas = 1;
  cfg.replicas = &mp;
  cfg.precision = data.ahandler.precision == "fp32" ? LGV_FP32 : LGV_FP64;
  cfg.scounts[0] = 2;
  cfg.scounts[1] = 3;
  cfg.nodiffuse = nodiffuse;
  cfg.onlydiffuse = onlydiffuse;
  cfg.track_energy = 0;
  cfg.device = device;
  if (lgv_create(&cfg, &model.ctx) != LGV_OK) {
    fprintf(stderr, "lgv_create: %s\n", lgv_last_error());
    return 1;
  }
  try {
    if (data.ahandler.restart && !data.ahandler.eventmode) read_rng(model);
    // ModelA::initialize, ModelA.h:408-461
    if (data.ahandler.restart) {
      try {
        std::vector<double> u((size_t)model.N * model.N * model.N * 10);
        read_o4fields(data.ahandler.outputfiletag + "_save.h5", model.N, u.data());
        check(lgv_set_field(model.ctx, 0, u.data()), "lgv_set_field");
      } catch (const std::string &error) {
        std::cout << "Error in restart -- aborting!" << std::endl;
        std::cout << error << std::endl;
        std::abort();
      }
    }
    if (log_view) check(lgv_timing_enable(model.ctx, 1), "lgv_timing_enable");
    HipPV2HBSplit step(model.ctx, &data.atime.time);
    lgv_monitor therm_total{};
    auto &ah = data.ahandler;
    if (ah.eventmode) {
      for (int i = 0; i < ah.nevents; i++) {
        run_event(model, step, therm_total);
        ah.current_event++;
      }
    } else {
      run_event(model, step, therm_total);
    }
    step.finalize();
    // ModelA::finalize, ModelA.h:338-352
    if (!ah.eventmode) {
      write_rng(model);
      write_fields(model, ah.outputfiletag);
    }
    if (log_view) {
      double ms[LGV_NCLASS];
      long long n[LGV_NCLASS];
      check(lgv_timing_get(model.ctx, ms, n), "lgv_timing_get");
      const char *names[LGV_NCLASS] = {"IdealStep", "HBStep", "QHBStep", "Energy"};
      printf("---- kernel time on the GPU stream (HIP events) ----\n");
      for (int k = 0; k < LGV_NCLASS; k++)
        printf("%-10s %10lld launches %12.3f ms\n", names[k], n[k], ms[k]);
    }
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    lgv_destroy(model.ctx);
    return 1;
  }
  lgv_destroy(model.ctx);
  return 0;
}
