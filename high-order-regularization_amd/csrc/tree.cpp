// KITC treecode, host part: the octree over the sources and each target's
// interaction list.  Both depend on geometry only, so they are built once per
// plan (the reference rebuilds the tree on every operator call,
// KITC.cpp:1435-1436).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hor_internal.h"

namespace hor {

// Children of a node: midpoint split with the reference's child centres and
// MAC expressions (split_tree_node, KITC.cpp:116-243).  Octant o has bit 0 set
// for the upper x half, bit 1 for y, bit 2 for z (the reference's child order).
static void make_children(const TreeNode& p, double sqth, TreeNode ch[8]) {
  double x0 = p.iv[0], x1 = p.iv[1], y0 = p.iv[2], y1 = p.iv[3], z0 = p.iv[4], z1 = p.iv[5];
  double mx = (x0 + x1) / 2.0, my = (y0 + y1) / 2.0, mz = (z0 + z1) / 2.0;
  double xc[2] = {(x0 + mx) / 2.0, (x1 + mx) / 2.0};
  double yc[2] = {(y0 + my) / 2.0, (y1 + my) / 2.0};
  double zc[2] = {(z0 + mz) / 2.0, (z1 + mz) / 2.0};
  for (int o = 0; o < 8; o++) {
    int hx = o & 1, hy = (o >> 1) & 1, hz = (o >> 2) & 1;
    TreeNode& c = ch[o];
    std::memset(&c, 0, sizeof(c));
    c.iv[0] = hx ? mx : x0; c.iv[1] = hx ? x1 : mx;
    c.iv[2] = hy ? my : y0; c.iv[3] = hy ? y1 : my;
    c.iv[4] = hz ? mz : z0; c.iv[5] = hz ? z1 : mz;
    c.c[0] = xc[hx]; c.c[1] = yc[hy]; c.c[2] = zc[hz];
    // the reference measures from the centre to the corner on the far side
    // of the midpoint for a lower child and to the outer face for an upper one
    double ex = hx ? (x1 - xc[1]) : (mx - xc[0]);
    double ey = hy ? (y1 - yc[1]) : (my - yc[0]);
    double ez = hz ? (z1 - zc[1]) : (mz - zc[0]);
    c.mac = (ex * ex + ey * ey + ez * ez) / sqth;
  }
}

struct Builder {
  Tree& T;
  std::vector<double> px, py, pz;   // positions in tree order
  std::vector<int> oct;
  Builder(Tree& t) : T(t) {}

  // Stable partition of [lo, hi] by octant: the order the reference's
  // address-table swaps produce (KITC.cpp:244-319).
  void split(int ni) {
    TreeNode nd = T.nodes[ni];
    double mx = (nd.iv[0] + nd.iv[1]) / 2.0, my = (nd.iv[2] + nd.iv[3]) / 2.0,
           mz = (nd.iv[4] + nd.iv[5]) / 2.0;
    int n = nd.hi - nd.lo + 1;
    oct.resize(n);
    int cnt[8] = {0};
    for (int k = 0; k < n; k++) {
      int p = nd.lo + k;
      int o = (px[p] > mx ? 1 : 0) | (py[p] > my ? 2 : 0) | (pz[p] > mz ? 4 : 0);
      oct[k] = o;
      cnt[o]++;
    }
    int start[8], s = nd.lo;
    for (int o = 0; o < 8; o++) { start[o] = s; s += cnt[o]; }
    std::vector<double> tx(n), ty(n), tz(n);
    std::vector<int> tp(n);
    int pos[8];
    std::memcpy(pos, start, sizeof(pos));
    for (int k = 0; k < n; k++) {
      int d = pos[oct[k]]++ - nd.lo;
      int p = nd.lo + k;
      tx[d] = px[p]; ty[d] = py[p]; tz[d] = pz[p]; tp[d] = T.perm[p];
    }
    std::memcpy(&px[nd.lo], tx.data(), sizeof(double) * n);
    std::memcpy(&py[nd.lo], ty.data(), sizeof(double) * n);
    std::memcpy(&pz[nd.lo], tz.data(), sizeof(double) * n);
    std::memcpy(&T.perm[nd.lo], tp.data(), sizeof(int) * n);
    TreeNode ch[8];
    make_children(nd, T.theta * T.theta, ch);
    int nc = 0, kids[8];
    for (int o = 0; o < 8; o++) {
      if (cnt[o] < 1) continue;  // empty children are dropped
      ch[o].lo = start[o];
      ch[o].hi = start[o] + cnt[o] - 1;
      ch[o].parent = ni;
      ch[o].level = nd.level + 1;
      T.nodes.push_back(ch[o]);
      kids[nc++] = (int)T.nodes.size() - 1;
    }
    T.nodes[ni].nchild = nc;
    for (int k = 0; k < nc; k++) T.nodes[ni].child[k] = kids[k];
  }

  // build_tree_3D_Recursive (KITC.cpp:325-348): split iff n >= N0.
  void build(int ni) {
    int n = T.nodes[ni].hi - T.nodes[ni].lo + 1;
    if (n >= T.N0) {
      split(ni);
      int nc = T.nodes[ni].nchild;
      int kids[8];
      std::memcpy(kids, T.nodes[ni].child, sizeof(kids));
      for (int k = 0; k < nc; k++) build(kids[k]);
    }
  }
};

void tree_build(Tree& T, int ns, const double* sx, const double ctr[3], int P, int N0,
                double theta) {
  T.P = P;
  T.N0 = N0;
  T.theta = theta;
  T.nodes.clear();
  T.perm.resize(ns);
  Builder B(T);
  B.px.resize(ns); B.py.resize(ns); B.pz.resize(ns);
  for (int i = 0; i < ns; i++) {
    T.perm[i] = i;
    B.px[i] = sx[3 * i]; B.py[i] = sx[3 * i + 1]; B.pz[i] = sx[3 * i + 2];
  }
  // build_tree_init (KITC.cpp:30-69): the box [ctr-1-1e-8, ctr+1+1e-8]^3 and
  // the root MAC (3*10*10/4)/theta^2 computed in integer then double.
  TreeNode root;
  std::memset(&root, 0, sizeof(root));
  root.lo = 0;
  root.hi = ns - 1;
  double eps = 1.e-8;
  for (int d = 0; d < 3; d++) {
    root.iv[2 * d] = -1.0 - eps;
    root.iv[2 * d + 1] = 1.0 + eps;
    root.iv[2 * d] += ctr[d];
    root.iv[2 * d + 1] += ctr[d];
    root.c[d] = 0.5 * (root.iv[2 * d] + root.iv[2 * d + 1]);
  }
  root.mac = (3 * 10 * 10 / 4) / (theta * theta);
  root.parent = -1;
  T.nodes.push_back(root);
  B.build(0);
  // The sources of each childless node in Morton order of their positions
  // (the split keeps the input order within an octant): the near sums'
  // 64-source subtiles then have tight boxes.  Node ranges and members are
  // unchanged.
  for (const TreeNode& nd : T.nodes) {
    const int n = nd.hi - nd.lo + 1;
    if (nd.nchild || n < 2) continue;
    const double lo[3] = {nd.iv[0], nd.iv[2], nd.iv[4]};
    const double ext = std::max(nd.iv[1] - nd.iv[0], std::max(nd.iv[3] - nd.iv[2], nd.iv[5] - nd.iv[4]));
    std::vector<std::pair<uint64_t, int>> kp(n);
    for (int k = 0; k < n; k++) {
      const int o = T.perm[nd.lo + k];
      kp[k] = {morton_key(sx + 3 * (size_t)o, lo, ext), o};
    }
    std::stable_sort(kp.begin(), kp.end(),
                     [](const std::pair<uint64_t, int>& a, const std::pair<uint64_t, int>& b) { return a.first < b.first; });
    for (int k = 0; k < n; k++) T.perm[nd.lo + k] = kp[k].second;
  }
  // Cluster_Chev_Points (KITC.cpp:780-804): Chebyshev-Lobatto nodes per box.
  int np1 = P + 1;
  std::vector<double> t(np1);
  double hh = 3.14159265358979323846 / P;
  for (int i = 0; i < np1; i++) t[i] = std::cos(i * hh);
  T.cheb.assign((size_t)3 * np1 * T.nodes.size(), 0.0);
  for (size_t q = 0; q < T.nodes.size(); q++) {
    const double* iv = T.nodes[q].iv;
    double* cq = &T.cheb[3 * np1 * q];
    for (int i = 0; i < np1; i++) {
      cq[i] = iv[0] + (t[i] + 1) / 2 * (iv[1] - iv[0]);
      cq[np1 + i] = iv[2] + (t[i] + 1) / 2 * (iv[3] - iv[2]);
      cq[2 * np1 + i] = iv[4] + (t[i] + 1) / 2 * (iv[5] - iv[4]);
    }
  }
}

// Per-target traversal (eval_*_Sum_cluster, KITC.cpp:1066-1386): far when
// MAC < R^2, leaf when (members[1] - members[0]) < N0 (so a node of exactly
// N0 members is split yet summed directly), else the children in order.
static void visit(const Tree& T, int q, const double p[3], std::vector<int>& node,
                  std::vector<uint8_t>& far) {
  const TreeNode& nd = T.nodes[q];
  double tpx = p[0] - nd.c[0], tpy = p[1] - nd.c[1], tpz = p[2] - nd.c[2];
  double R2 = tpx * tpx + tpy * tpy + tpz * tpz;
  if (nd.mac < R2) {
    // The reference never computes root moments (KITC.cpp:1446 starts at
    // node 1), so a target far from the root box gets no contribution.
    if (q == 0) return;
    node.push_back(q);
    far.push_back(1);
  } else if ((nd.hi - nd.lo) < T.N0) {
    node.push_back(q);
    far.push_back(0);
  } else {
    for (int c = 0; c < nd.nchild; c++) visit(T, nd.child[c], p, node, far);
  }
}

void tree_lists(const Tree& T, int nt, const double* tx, TargetLists& L) {
  std::vector<std::vector<int>> nodes(nt);
  std::vector<std::vector<uint8_t>> fars(nt);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nt; i++) visit(T, 0, tx + 3 * i, nodes[i], fars[i]);
  L.start.assign(nt + 1, 0);
  for (int i = 0; i < nt; i++) L.start[i + 1] = L.start[i] + (int)nodes[i].size();
  L.node.resize(L.start[nt]);
  L.far.resize(L.start[nt]);
  for (int i = 0; i < nt; i++) {
    std::memcpy(&L.node[L.start[i]], nodes[i].data(), sizeof(int) * nodes[i].size());
    std::memcpy(&L.far[L.start[i]], fars[i].data(), nodes[i].size());
  }
}

}  // namespace hor
