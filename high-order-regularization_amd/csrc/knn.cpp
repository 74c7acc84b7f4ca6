// k nearest neighbours of every point of a point cloud (host, OpenMP): the
// message-passing graph of the learned preconditioner.  A k-d tree with
// leaves of at most 16 points; neighbours are sorted by (distance, index), the
// point itself excluded, so the graph is deterministic.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "hor_internal.h"

namespace hor {

namespace {

struct KdNode {
  int lo, hi;          // point range [lo, hi) in idx
  int left, right;     // children (-1 for a leaf)
  int dim;
  double split;
};

struct KdTree {
  const double* x = nullptr;
  std::vector<int> idx;
  std::vector<KdNode> nodes;

  int build(int lo, int hi) {
    KdNode nd{lo, hi, -1, -1, 0, 0.0};
    const int me = (int)nodes.size();
    nodes.push_back(nd);
    if (hi - lo <= 16) return me;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int p = lo; p < hi; p++)
      for (int d = 0; d < 3; d++) {
        mn[d] = std::min(mn[d], x[3 * idx[p] + d]);
        mx[d] = std::max(mx[d], x[3 * idx[p] + d]);
      }
    int dim = 0;
    for (int d = 1; d < 3; d++)
      if (mx[d] - mn[d] > mx[dim] - mn[dim]) dim = d;
    const int mid = (lo + hi) / 2;
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int a, int b) {
      double xa = x[3 * a + dim], xb = x[3 * b + dim];
      return xa < xb || (xa == xb && a < b);
    });
    nodes[me].dim = dim;
    nodes[me].split = x[3 * idx[mid] + dim];
    const int l = build(lo, mid);
    const int r = build(mid, hi);
    nodes[me].left = l;
    nodes[me].right = r;
    return me;
  }

  // best: sorted (d2, index) of the current k best, size <= k
  void query(int node, const double q[3], int self, int k, std::vector<std::pair<double, int>>& best) const {
    const KdNode& nd = nodes[node];
    if (nd.left < 0) {
      for (int p = nd.lo; p < nd.hi; p++) {
        const int j = idx[p];
        if (j == self) continue;
        const double dx = x[3 * j] - q[0], dy = x[3 * j + 1] - q[1], dz = x[3 * j + 2] - q[2];
        const std::pair<double, int> c(dx * dx + dy * dy + dz * dz, j);
        if ((int)best.size() < k) {
          best.insert(std::upper_bound(best.begin(), best.end(), c), c);
        } else if (c < best.back()) {
          best.pop_back();
          best.insert(std::upper_bound(best.begin(), best.end(), c), c);
        }
      }
      return;
    }
    const double diff = q[nd.dim] - nd.split;
    const int nearc = diff < 0 ? nd.left : nd.right;
    const int farc = diff < 0 ? nd.right : nd.left;
    query(nearc, q, self, k, best);
    if ((int)best.size() < k || diff * diff <= best.back().first) query(farc, q, self, k, best);
  }
};

}  // namespace

int knn_graph(int n, const double* x, int k, std::vector<int>& nbr) {
  if (n < 2 || k < 1) return set_error(HOR_E_ARG, "knn_graph: need n >= 2 and k >= 1 (n=%d, k=%d)", n, k);
  const int kk = std::min(k, n - 1);
  KdTree T;
  T.x = x;
  T.idx.resize(n);
  std::iota(T.idx.begin(), T.idx.end(), 0);
  T.nodes.reserve(2 * (size_t)n / 8 + 16);
  T.build(0, n);
  nbr.assign((size_t)n * k, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    std::vector<std::pair<double, int>> best;
    best.reserve(kk + 1);
    T.query(0, x + 3 * (size_t)i, i, kk, best);
    for (int a = 0; a < k; a++) nbr[(size_t)i * k + a] = best[std::min(a, kk - 1)].second;
  }
  return HOR_OK;
}

}  // namespace hor

extern "C" int hor_knn_graph(int n, const double* x, int k, int* nbr) {
  if (!x || !nbr) return hor::set_error(HOR_E_ARG, "hor_knn_graph: null");
  std::vector<int> v;
  int rc = hor::knn_graph(n, x, k, v);
  if (rc != HOR_OK) return rc;
  std::copy(v.begin(), v.end(), nbr);
  return HOR_OK;
}
