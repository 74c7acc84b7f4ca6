// The reference's C++ sum interface (stokeslets.h, stresslets.h, KITC.h)
// served by libhor.so on the GPU.  See include/hor_dropin.h.
#include "hor_dropin.h"

#include <cstdint>
#include <cstring>
#include <list>
#include <mutex>
#include <stdexcept>
#include <string>

#include "hor.h"

namespace {

std::mutex g_mu;
hor_ctx* g_ctx = nullptr;
hor_tree_params g_tree = {6, 1000, 0.6};                      // KITC.h:14-17
double g_ctr[2][3] = {{0.0, 0.0, -0.5 - 1.0 / (16.0 * 16.0 * 16.0)}, {0.0, 0.0, 0.5}};  // utilities.h:20-26

void check(int rc, const char* where) {
  if (rc != HOR_OK) throw std::runtime_error(std::string(where) + ": " + hor_last_error());
}

hor_ctx* ctx() {
  if (!g_ctx) check(hor_ctx_create(0, &g_ctx), "hor_ctx_create");
  return g_ctx;
}

struct Flat {
  std::vector<double> sx, sn, sa, tx, tb, tx0, tn0;
};

// A cached plan and the exact geometry it was built from.  A call reuses the
// plan only when its own geometry equals the stored copy value for value
// (compared in place, straight from the caller's structures, without packing
// or hashing); densities are uploaded on every call.
struct CacheEntry {
  int kind = 0, method = 0;
  double DEL = 0;
  double ctr[3] = {0, 0, 0};
  hor_tree_params tree = {0, 0, 0};
  Flat F;
  hor_plan* plan = nullptr;
};

std::list<CacheEntry> g_cache;  // most recent first
const size_t kCacheSize = 16;

bool same_sources(const Flat& F, int N_quad, const vector<Surf_point>& S) {
  if (F.sa.size() != (size_t)N_quad) return false;
  for (int j = 0; j < N_quad; j++) {
    const Surf_point& p = S[j];
    if (F.sa[j] != p.Area) return false;
    for (int k = 0; k < 3; k++)
      if (F.sx[3 * j + k] != p.x[k] || F.sn[3 * j + k] != p.Nrml[k]) return false;
  }
  return true;
}

bool same_targets(const Flat& F, int N_target, const vector<Target_point>& T, bool off) {
  if (F.tx.size() != 3 * (size_t)N_target) return false;
  for (int i = 0; i < N_target; i++) {
    const Target_point& t = T[i];
    if (F.tx[3 * i] != t.x || F.tx[3 * i + 1] != t.y || F.tx[3 * i + 2] != t.z) return false;
    if (!off) continue;
    if (F.tb[i] != t.b || F.tx0[3 * i] != t.nrst_x || F.tx0[3 * i + 1] != t.nrst_y || F.tx0[3 * i + 2] != t.nrst_z)
      return false;
    for (int k = 0; k < 3; k++)
      if (F.tn0[3 * i + k] != t.nrst_Nrml[k]) return false;
  }
  return true;
}

bool same_setup(const CacheEntry& e, int kind, int method, double DEL, const double* ctr) {
  if (e.kind != kind || e.method != method || e.DEL != DEL) return false;
  if (method != HOR_TREE) return true;
  return e.ctr[0] == ctr[0] && e.ctr[1] == ctr[1] && e.ctr[2] == ctr[2] && e.tree.P == g_tree.P &&
         e.tree.N0 == g_tree.N0 && e.tree.theta == g_tree.theta;
}

void pack_sources(int N_quad, const vector<Surf_point>& S, Flat& F) {
  F.sx.resize(3 * (size_t)N_quad);
  F.sn.resize(3 * (size_t)N_quad);
  F.sa.resize(N_quad);
  for (int j = 0; j < N_quad; j++) {
    for (int k = 0; k < 3; k++) {
      F.sx[3 * j + k] = S[j].x[k];
      F.sn[3 * j + k] = S[j].Nrml[k];
    }
    F.sa[j] = S[j].Area;
  }
}

std::vector<double> pack_density(int N_quad, const vector<Surf_point>& S, bool use_g) {
  std::vector<double> d(3 * (size_t)N_quad);
  for (int j = 0; j < N_quad; j++)
    for (int k = 0; k < 3; k++) d[3 * j + k] = use_g ? S[j].g[k] : S[j].f[k];
  return d;
}

// Packs target positions; off-surface kinds also carry b and the nearest
// point and normal.
void pack_targets(int N_target, const vector<Target_point>& T, bool off, Flat& F) {
  F.tx.resize(3 * (size_t)N_target);
  for (int i = 0; i < N_target; i++) {
    F.tx[3 * i] = T[i].x;
    F.tx[3 * i + 1] = T[i].y;
    F.tx[3 * i + 2] = T[i].z;
  }
  F.tb.clear();
  F.tx0.clear();
  F.tn0.clear();
  if (!off) return;
  F.tb.resize(N_target);
  F.tx0.resize(3 * (size_t)N_target);
  F.tn0.resize(3 * (size_t)N_target);
  for (int i = 0; i < N_target; i++) {
    F.tb[i] = T[i].b;
    F.tx0[3 * i] = T[i].nrst_x;
    F.tx0[3 * i + 1] = T[i].nrst_y;
    F.tx0[3 * i + 2] = T[i].nrst_z;
    for (int k = 0; k < 3; k++) F.tn0[3 * i + k] = T[i].nrst_Nrml[k];
  }
}

// d0 of off-surface targets: nrst_f (SL) or nrst_g (DL).
std::vector<double> pack_d0(int N_target, const vector<Target_point>& T, bool use_g) {
  std::vector<double> d0(3 * (size_t)N_target);
  for (int i = 0; i < N_target; i++)
    for (int k = 0; k < 3; k++) d0[3 * i + k] = use_g ? T[i].nrst_g[k] : T[i].nrst_f[k];
  return d0;
}

// The cached plan of this geometry, or a new one (the least recently used
// entry is dropped beyond kCacheSize).
hor_plan* plan_for(int kind, int method, double DEL, const double* ctr, int N_quad, const vector<Surf_point>& S,
                   int N_target, const vector<Target_point>& T, bool off) {
  for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
    if (same_setup(*it, kind, method, DEL, ctr) && same_sources(it->F, N_quad, S) &&
        same_targets(it->F, N_target, T, off)) {
      g_cache.splice(g_cache.begin(), g_cache, it);
      return g_cache.front().plan;
    }
  CacheEntry e;
  e.kind = kind;
  e.method = method;
  e.DEL = DEL;
  if (ctr)
    for (int k = 0; k < 3; k++) e.ctr[k] = ctr[k];
  e.tree = g_tree;
  pack_sources(N_quad, S, e.F);
  pack_targets(N_target, T, off, e.F);
  const Flat& F = e.F;
  check(hor_plan_create(ctx(), kind, method, HOR_FP64, DEL, N_quad, F.sx.data(), F.sn.data(), F.sa.data(), N_target,
                        F.tx.data(), F.tb.empty() ? nullptr : F.tb.data(), F.tx0.empty() ? nullptr : F.tx0.data(),
                        F.tn0.empty() ? nullptr : F.tn0.data(), ctr, &g_tree, &e.plan),
        "hor_plan_create");
  g_cache.push_front(std::move(e));
  if (g_cache.size() > kCacheSize) {
    hor_plan_destroy(g_cache.back().plan);
    g_cache.pop_back();
  }
  return g_cache.front().plan;
}

// Batch evaluation: sources from Surfc (density f or g), targets from Target.
void batch(int kind, int method, int surf_num, double DEL, int N_quad, const vector<Surf_point>& Surfc,
           int N_target, const vector<Target_point>& Target, vector<double>& out) {
  std::lock_guard<std::mutex> lock(g_mu);
  bool dl = (kind == HOR_DL_ON || kind == HOR_DL_OFF);
  bool off = (kind == HOR_SL_OFF || kind == HOR_DL_OFF);
  const double* ctr = (method == HOR_TREE) ? g_ctr[surf_num == 1 ? 0 : 1] : nullptr;
  hor_plan* p = plan_for(kind, method, DEL, ctr, N_quad, Surfc, N_target, Target, off);
  std::vector<double> dens = pack_density(N_quad, Surfc, dl);
  std::vector<double> d0;
  if (off) d0 = pack_d0(N_target, Target, dl);
  if (out.size() < 3 * (size_t)N_target) out.resize(3 * (size_t)N_target);
  check(hor_plan_eval(p, dens.data(), off ? d0.data() : nullptr, out.data()), "hor_plan_eval");
}

// One target with explicit data (the per-target helpers).
void single(int kind, double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& dens,
            double x, double y, double z, double b, const double* x0, const double* n0, const double* d0,
            vector<double>& out) {
  std::lock_guard<std::mutex> lock(g_mu);
  vector<Target_point> T(1);
  T[0].x = x; T[0].y = y; T[0].z = z; T[0].b = b;
  T[0].nrst_x = x0[0]; T[0].nrst_y = x0[1]; T[0].nrst_z = x0[2];
  for (int k = 0; k < 3; k++) T[0].nrst_Nrml[k] = n0[k];
  hor_plan* p = plan_for(kind, HOR_DIRECT, DEL, nullptr, N_quad, Surfc, 1, T, true);
  if (dens.size() < 3 * (size_t)N_quad) throw std::runtime_error("density vector shorter than 3*N_quad");
  double r[3];
  check(hor_plan_eval(p, dens.data(), d0, r), "hor_plan_eval");
  if (out.size() < 3) out.resize(3);
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

}  // namespace

// ---- stokeslets.h -----------------------------------------------------------

void eval_Stokes_SL_offSurf(double, double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                            const vector<Target_point>& Target, vector<double>& SL) {
  batch(HOR_SL_OFF, HOR_DIRECT, 0, DEL, N_quad, Surfc, N_target, Target, SL);
}

void stokes_SL_7ord(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& f,
                    double f0_dot_n0, double x, double y, double z, double b, vector<double>& SL) {
  const double e1[3] = {1.0, 0.0, 0.0}, d0[3] = {f0_dot_n0, 0.0, 0.0};
  single(HOR_SL_OFF | HOR_ALL_NEAR, DEL, N_quad, Surfc, f, x, y, z, b, e1, e1, d0, SL);
}

void eval_Stokes_SL_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                           const vector<Target_point>& Target, vector<double>& SL_vec) {
  batch(HOR_SL_ON, HOR_DIRECT, 0, DEL, N_quad, Surfc, N_target, Target, SL_vec);
}

// The on-surface Stokeslet is the regularized one at lambda = 0
// (c(0) = (22/5, 8/5, 2/15)/sqrt(pi) reproduces utilities.cpp:240-242).
void stokes_SL_7ord_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& f,
                           double f0_dot_n0, double x, double y, double z, vector<double>& SL) {
  const double e1[3] = {1.0, 0.0, 0.0}, d0[3] = {f0_dot_n0, 0.0, 0.0};
  single(HOR_SL_OFF | HOR_ALL_NEAR, DEL, N_quad, Surfc, f, x, y, z, 0.0, e1, e1, d0, SL);
}

// ---- stresslets.h -----------------------------------------------------------

void eval_Stokes_DL_offSurf(double, double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                            vector<Target_point>& Target, vector<double>& DL_vec) {
  batch(HOR_DL_OFF, HOR_DIRECT, 0, DEL, N_quad, Surfc, N_target, Target, DL_vec);
}

void stokes_DL_7ord(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& g,
                    const vector<double>& g0, const vector<double>& x0, const vector<double>& n0, double x,
                    double y, double z, double b, vector<double>& DL) {
  single(HOR_DL_OFF | HOR_ALL_NEAR, DEL, N_quad, Surfc, g, x, y, z, b, x0.data(), n0.data(), g0.data(), DL);
}

void eval_Stokes_DL_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                           vector<Target_point>& Target, vector<double>& DL_vec) {
  batch(HOR_DL_ON, HOR_DIRECT, 0, DEL, N_quad, Surfc, N_target, Target, DL_vec);
}

void stokes_DL_7ord_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& g,
                           const vector<double>& g0, double x, double y, double z, vector<double>& DL) {
  const double e1[3] = {1.0, 0.0, 0.0}, p[3] = {x, y, z};
  single(HOR_DL_ON | HOR_ON_AT, DEL, N_quad, Surfc, g, x, y, z, 0.0, p, e1, g0.data(), DL);
}

// ---- KITC.h -----------------------------------------------------------------

void eval_Stokes_SL_onSurf_tree(int surf_num, double, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                int N_target, const vector<Target_point>& Target, vector<double>& SL) {
  batch(HOR_SL_ON, HOR_TREE, surf_num, DEL, N_quad, Surfc, N_target, Target, SL);
}

void eval_Stokes_SL_offSurf_tree(int surf_num, double, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                 int N_target, const vector<Target_point>& Target, vector<double>& SL) {
  batch(HOR_SL_OFF, HOR_TREE, surf_num, DEL, N_quad, Surfc, N_target, Target, SL);
}

void eval_Stokes_DL_onSurf_tree(int surf_num, double, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                int N_target, const vector<Target_point>& Target, vector<double>& DL) {
  batch(HOR_DL_ON, HOR_TREE, surf_num, DEL, N_quad, Surfc, N_target, Target, DL);
}

void eval_Stokes_DL_offSurf_tree(int surf_num, double, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                 int N_target, const vector<Target_point>& Target, vector<double>& SL) {
  batch(HOR_DL_OFF, HOR_TREE, surf_num, DEL, N_quad, Surfc, N_target, Target, SL);
}

// ---- side channel -----------------------------------------------------------

extern "C" void hor_kitc_set_params(int P, int N0, double theta) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_tree.P = P;
  g_tree.N0 = N0;
  g_tree.theta = theta;
}

extern "C" void hor_dropin_set_centre(int surf_num, const double ctr[3]) {
  std::lock_guard<std::mutex> lock(g_mu);
  int k = (surf_num == 1) ? 0 : 1;
  for (int d = 0; d < 3; d++) g_ctr[k][d] = ctr[d];
}

extern "C" void hor_dropin_release(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto& e : g_cache) hor_plan_destroy(e.plan);
  g_cache.clear();
  if (g_ctx) hor_ctx_destroy(g_ctx);
  g_ctx = nullptr;
}
