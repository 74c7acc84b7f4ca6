// Operator plan: everything an evaluation needs that depends on geometry only.
#pragma once

#include <memory>
#include <vector>

#include "device.h"
#include "direct.h"

namespace hor {
struct TreePlan;  // kitc.hip
}

struct hor_plan {
  hor_ctx* ctx = nullptr;
  int kind = 0, method = 0, precision = 0;
  bool all_near = false, on_at = false;  // HOR_ALL_NEAR / HOR_ON_AT
  double DEL = 0;
  int ns = 0, nsp = 0, nt = 0, ntp = 0;
  int nsub = 0, subs_per_chunk = 0, nchunk = 0;
  std::vector<int> sperm_h, tperm_h;
  // sources, sorted SoA (stride nsp)
  hor::DBuf<double> sx, sn, sa, sd, sub_box, sn_orig;
  hor::DBuf<hor::SrcRec> rec;
  hor::DBuf<hor::SrcRec32> rec32;
  hor::DBuf<hor::SrcRecV> recv;  // SLDL_OFF records
  hor::DBuf<hor::SrcRecV32> recv32;
  hor::DBuf<hor::SrcRecP> recp;  // P_OFF records
  hor::DBuf<double> dlon_tab;  // on-surface stresslet table (kDlOnTable), DL_ON plans
  hor::DBuf<double> ae_tab;    // erf(rho)/rho and exp(-rho^2) table (kErfExpTable), other plans
  hor::DBuf<int> sperm;
  // targets, sorted SoA (stride ntp)
  hor::DBuf<int> tperm, tnear;
  hor::DBuf<double> tx, tcoef, tsub, tsub2, tx0, tn0, tb, n0_orig, wave_box, xi, partial;
  // staging for the host-pointer API
  hor::DBuf<double> in_dens, in_d0, out_buf;
  // statistics
  double direct_pairs = 0, near_pairs = 0, proxy_pairs = 0, moment_terms = 0;
  std::unique_ptr<hor::TreePlan> tree;
  hor_plan();
  ~hor_plan();
};

namespace hor {

bool target_is_near(const hor_plan* P, double b);
void target_coefs(int kind, bool nearp, double b, double DEL, double c[3]);
int setup_targets(hor_plan* P, const double* tx, const double* tb, const double* tx0, const double* tn0,
                  const double* sx_orig);
int setup_sources(hor_plan* P, const std::vector<int>& perm, const double* sx, const double* sn,
                  const double* sa);
DirectArgs direct_args(hor_plan* P);
double op_scale(int kind);
int prepare_targets(hor_plan* P, const double* dens_dev, const double* d0_dev);

// Evaluates into out_dev (caller order, 3*nt):
//   accumulate=0, base=NULL: out = coef*op;  accumulate=0: out = base + coef*op;
//   accumulate=1: out += coef*op.
int plan_eval_into(hor_plan* P, const double* dens_dev, const double* d0_dev, double* out_dev,
                   const double* base, double coef, int accumulate);

int tree_plan_create(hor_plan* P, const double* sx, const double* sn, const double* sa,
                     const double* tx, const double* tb, const double* tx0, const double* tn0,
                     const double* ctr, const hor_tree_params& tp);
// The user plan reads the moments the owner plan computes for both (same
// source tree and density; the owner is evaluated first).
int tree_plan_share_moments(hor_plan* owner, hor_plan* user, const double* sx);
// The double layer's subtraction coefficients of a treecode plan (a no-op for
// the single layer), before it reads another plan's moments.
int tree_plan_prepare_subtraction(hor_plan* P);
int tree_plan_eval(hor_plan* P, const double* dens_dev, const double* d0_dev, double* out_dev,
                   const double* base, double coef, int accumulate);
// HOR_FAR_VALU or HOR_FAR_MFMA (mixed-precision single-layer plans only).
int tree_plan_set_far_kernel(hor_plan* P, int kernel);
// Kernel time of each phase of a treecode plan's last evaluation: moments,
// near sums, far field (ms).
int tree_plan_phase_ms(hor_plan* P, double ms[3]);

}  // namespace hor
