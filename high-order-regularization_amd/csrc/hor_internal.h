// Internal declarations shared by the host and device sources of libhor.so.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "hor.h"

namespace hor {

// Records the calling thread's last error text; returns code.
int set_error(int code, const char* fmt, ...);

// ---- host geometry (geometry.cpp)
int surface_generate(const hor_body& B, double h, double pou_deg, std::vector<double>& X,
                     std::vector<double>& Nr, std::vector<double>& Ar, std::vector<int>& Pl);
int nearest_data(const hor_body& B, double h, double DEL, int ns, const double* sx,
                 const int* splane, int nt, const double* tx, int* nrst, double* x0,
                 double* n0, double* b, int* i123, double* f0);
void surface_force(const hor_body& B, const double x0[3], const double n0[3], double f[3]);
int band_targets(const hor_body& B, const double* origin, double h, int ngrid, double band,
                 int ns, const double* sx, std::vector<int>& ijk, std::vector<double>& X,
                 std::vector<double>& b, std::vector<double>& X0, std::vector<double>& N0);
void traction_translating(const hor_body& B, const double U[3], int n, const double* x, double* f);
void spheroid_exact(const hor_body& B, int n, const double* x, double* u);
void spheroid_exact_pressure(const hor_body& B, int n, const double* x, double* p);

// ---- regularization coefficients (common.cpp): evaluate_In_024 +
// get_reg_coeff_7ord (highreg_stokes_2surf/utilities.cpp:163-175, 327-337),
// evaluated on the host with the C library's erfc/exp so that the GPU uses
// exactly the coefficients the reference computes.
void reg_coeff_7ord(double lam, double c[3]);

// ---- spatial ordering (common.cpp)
// Morton key of p inside the box [lo, lo+ext]^3 (21 bits per axis).
uint64_t morton_key(const double p[3], const double lo[3], double ext);
// Stable order of n points by (cls[i], Morton key); cls may be null.
std::vector<int> spatial_order(int n, const double* x, const int* cls);

// ---- treecode host structures (tree.cpp)
struct TreeNode {
  int lo, hi;          // member range (inclusive) in tree order
  double iv[6];        // x0,x1,y0,y1,z0,z1
  double c[3];         // centre used by the MAC
  double mac;          // (half diagonal)^2 / theta^2
  int child[8];
  int nchild;
  int parent;
  int level;
};

struct Tree {
  int P = 6, N0 = 1000;
  double theta = 0.6;
  std::vector<TreeNode> nodes;
  std::vector<int> perm;      // tree slot -> original source index
  std::vector<double> cheb;   // per node 3*(P+1): t1, t2, t3
};

// KITC tree build (KITC.cpp:30-348, 780-804).
void tree_build(Tree& T, int ns, const double* sx, const double ctr[3], int P, int N0,
                double theta);

// Interaction lists of each target: the nodes visited by the reference's
// recursive traversal (KITC.cpp:1066-1386), in visiting order, tagged far
// (proxy interaction) or near (leaf direct sum).
struct TargetLists {
  std::vector<int> start;     // nt+1 offsets into node/far
  std::vector<int> node;
  std::vector<uint8_t> far;
};
void tree_lists(const Tree& T, int nt, const double* tx, TargetLists& L);

// ---- k nearest neighbours (knn.cpp): nbr[i*k + a], sorted by distance,
// the point itself excluded.
int knn_graph(int n, const double* x, int k, std::vector<int>& nbr);

}  // namespace hor
