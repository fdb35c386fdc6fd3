// lgv_stepper.hpp -- C++ adapter of the reference's Stepper interface
// (ModelG/Stepper.h:6-11) over the C ABI of lgv.h.  A host program written
// against the reference's Stepper keeps its control flow: construct a
// stepper on the model, call step(dt) once per time step, finalize() at
// the end (which prints the o4_stepper_monitor statistics in the reference's
// format, Stepper.h:51-82).
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>

#include "lgv.h"

namespace lgv_host {

class Stepper {
 public:
  virtual bool step(const double &dt) = 0;
  virtual void finalize() = 0;
  virtual ~Stepper() = default;
};

inline lgv_monitor monitor_diff(const lgv_monitor &a, const lgv_monitor &b) {
  return lgv_monitor{a.down_counts - b.down_counts, a.up_counts - b.up_counts,
                     a.up_yes - b.up_yes,           a.up_no - b.up_no,
                     a.down_dS - b.down_dS,         a.up_dS - b.up_dS};
}

// o4_stepper_monitor::print, ModelG/Stepper.h:51-82
inline void print_monitor(FILE *fh, const std::string &description, const lgv_monitor &m) {
  fprintf(fh, "==== Statistics for %s ====\n", description.c_str());
  double tot = m.down_counts + m.up_counts;
  double prob = (m.up_yes + m.down_counts) / tot;
  fprintf(fh, "Probability of up acceptance = %f\n", prob);
  double mean_dS = m.up_dS / m.up_yes;
  prob = m.up_yes / tot;
  fprintf(fh, "dS >0: accepts = %ld; probability of up accept = %f; mean dS=%f\n",
          (long)m.up_yes, prob, mean_dS);
  prob = m.up_no / tot;
  fprintf(fh, "dS =0: rejects = %ld; probability of up reject = %f; mean dS=%f\n",
          (long)m.up_no, prob, 0.);
  prob = m.up_counts / tot;
  mean_dS = m.up_dS / m.up_counts;
  fprintf(fh, "dS>=0: counts  = %ld; probability of non-neg.  = %f; mean dS=%f\n",
          (long)m.up_counts, prob, mean_dS);
  prob = m.down_counts / tot;
  mean_dS = m.down_dS / m.down_counts;
  fprintf(fh, "dS <0: counts  = %ld; probability of down steps= %f; mean dS=%f\n",
          (long)m.down_counts, prob, mean_dS);
  fprintf(fh, "============================================\n");
}

inline void check(int rc, const char *what) {
  if (rc != LGV_OK) throw std::runtime_error(std::string(what) + ": " + lgv_last_error());
}

// The reference steppers read m^2(t) from the model's clock at every call
// (ModelA.h:216); the adapters push the caller's clock to the device before
// each step.  `baseline` holds monitor counts that belong to other stepper
// objects sharing the context (e.g. a thermalizer), subtracted when printing.
class HipStepperBase : public Stepper {
 protected:
  lgv_ctx *ctx_;
  const double *clock_;
  int replica_;

 public:
  lgv_monitor baseline[3] = {};
  HipStepperBase(lgv_ctx *ctx, const double *clock, int replica = 0)
      : ctx_(ctx), clock_(clock), replica_(replica) {}
  lgv_monitor monitor(int which) const {
    lgv_monitor m[3];
    check(lgv_monitor_get(ctx_, replica_, m), "lgv_monitor_get");
    return monitor_diff(m[which], baseline[which]);
  }
  void sync_clock() const {
    if (clock_) check(lgv_set_time(ctx_, -1, *clock_), "lgv_set_time");
  }
};

// PV2HBSplit (ModelG/Stepper.h:176-211); scounts, nodiffuse and onlydiffuse
// are fixed when the context is created.
class HipPV2HBSplit : public HipStepperBase {
 public:
  using HipStepperBase::HipStepperBase;
  bool step(const double &dt) override {
    sync_clock();
    return lgv_step(ctx_, dt) == LGV_OK;
  }
  void finalize() override {
    // IdealPV2::finalize prints only with accept_reject (false here)
    print_monitor(stdout, "Phi HB Steps", monitor(LGV_MON_PHI));
    print_monitor(stdout, "Charge HB Steps", monitor(LGV_MON_CHARGE));
  }
};

// EulerLangevinHB (ModelG/Stepper.h:125-137)
class HipEulerLangevinHB : public HipStepperBase {
 public:
  using HipStepperBase::HipStepperBase;
  bool step(const double &dt) override {
    sync_clock();
    return lgv_hb_phi_step(ctx_, dt) == LGV_OK;
  }
  void finalize() override { print_monitor(stdout, "Phi HB Steps", monitor(LGV_MON_PHI)); }
};

// ModelGChargeHB (ModelG/Stepper.h:141-154)
class HipModelGChargeHB : public HipStepperBase {
 public:
  using HipStepperBase::HipStepperBase;
  bool step(const double &dt) override {
    sync_clock();
    return lgv_charge_step(ctx_, dt) == LGV_OK;
  }
  void finalize() override {
    print_monitor(stdout, "Charge HB Steps", monitor(LGV_MON_CHARGE));
  }
};

// IdealPV2 without accept/reject (ModelG/Stepper.h:87-121)
class HipIdealPV2 : public HipStepperBase {
 public:
  using HipStepperBase::HipStepperBase;
  bool step(const double &dt) override {
    sync_clock();
    return lgv_ideal_step(ctx_, dt) == LGV_OK;
  }
  void finalize() override {}
  void getEnergy(double &oldE, double &newE) const {
    double e[2];
    check(lgv_energy_get(ctx_, replica_, e), "lgv_energy_get");
    oldE = e[0];
    newE = e[1];
  }
};

}  // namespace lgv_host
