// OBB tree construction (host, FP64) for triangle meshes.
//
// Top-down build, one triangle per leaf.  Each node's axes are either the
// principal axes of its triangles' vertices (Jacobi eigen-decomposition of
// the 3x3 covariance) or the mesh axes, whichever gives the box of smaller
// surface area.  The node is split by the surface-area heuristic: triangle
// centroids are binned along each of the node's three axes, and the split
// minimising area(left) * n(left) + area(right) * n(right) is taken, with
// the boxes measured in the node's frame (the median split when every
// centroid falls in one bin).  Against the median split this cuts the box
// tests of link-vs-clutter traversals by about a third.  The leaf order is
// the depth-first order, so triangles of nearby leaves sit next to each
// other in memory.
//
// Stored boxes are exact bounding volumes in FP32: axes and centre are
// rounded to FP32 first, then the extents are measured in the dual basis of
// the rounded axes (so slightly non-orthonormal FP32 axes still bound every
// vertex) and rounded up.
//
// Large meshes are built in parallel: the top levels of the tree are built
// first and their open subtrees (disjoint ranges of the triangle index
// array) are built on host threads into separate arrays, then spliced in
// left-to-right order.  Every node is fitted and split exactly as in the
// serial build, so the tree and the leaf order are the same; only the node
// numbering differs (top levels first, then each subtree depth first).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "kg_types.h"
#include "kg_internal.h"

namespace {

void jacobi3(double A[3][3], double V[3][3], double w[3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = std::fabs(A[0][1]) + std::fabs(A[0][2]) + std::fabs(A[1][2]);
    double scale = std::fabs(A[0][0]) + std::fabs(A[1][1]) + std::fabs(A[2][2]);
    if (off <= 1e-30 * (scale + 1e-300)) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (std::fabs(A[p][q]) < 1e-300) continue;
        double theta = (A[q][q] - A[p][p]) / (2 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; k++) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; k++) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; i++) w[i] = A[i][i];
}

struct Builder {
  const double* V;
  const int32_t* T;
  const double* cen;        // centroids (3 per triangle), shared
  int* idx;                 // triangle order, shared: a builder reorders only its own ranges
  std::vector<ObbNode> nodes;
  std::vector<int> order;   // leaf order -> original triangle
  int depth = 0;

  // Surface area of the box of triangles [b,e) in the frame with axes A[k].
  double box_area(const double A[3][3], int b, int e) const {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = b; i < e; i++)
      for (int v = 0; v < 3; v++) {
        const double* p = V + 3 * T[3 * idx[i] + v];
        for (int k = 0; k < 3; k++) {
          const double s = A[k][0] * p[0] + A[k][1] * p[1] + A[k][2] * p[2];
          lo[k] = std::min(lo[k], s);
          hi[k] = std::max(hi[k], s);
        }
      }
    const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    return x * y + y * z + z * x;
  }

  // Split [b,e) by the binned surface-area heuristic over the node's axes R;
  // reorders idx[b,e) and returns the first index of the right part.
  int split(int b, int e, const float* R) {
    constexpr int NB = 32;
    const int n = e - b;
    double A[3][3];
    for (int k = 0; k < 3; k++)
      for (int r = 0; r < 3; r++) A[k][r] = R[k * 3 + r];
    auto proj = [&](int k, int t) { return A[k][0] * cen[3 * t] + A[k][1] * cen[3 * t + 1] + A[k][2] * cen[3 * t + 2]; };
    double best = 1e300;
    int bestk = -1, bestb = 0;
    for (int k = 0; k < 3; k++) {
      double cmin = 1e300, cmax = -1e300;
      for (int i = b; i < e; i++) {
        const double c = proj(k, idx[i]);
        cmin = std::min(cmin, c);
        cmax = std::max(cmax, c);
      }
      if (!(cmax > cmin)) continue;
      const double scale = NB / (cmax - cmin);
      double blo[NB][3], bhi[NB][3];
      int bn[NB] = {0};
      for (int j = 0; j < NB; j++)
        for (int m = 0; m < 3; m++) {
          blo[j][m] = 1e300;
          bhi[j][m] = -1e300;
        }
      for (int i = b; i < e; i++) {
        const int t = idx[i];
        const int j = std::min(NB - 1, (int)((proj(k, t) - cmin) * scale));
        bn[j]++;
        for (int v = 0; v < 3; v++) {
          const double* p = V + 3 * T[3 * t + v];
          for (int m = 0; m < 3; m++) {
            const double s = A[m][0] * p[0] + A[m][1] * p[1] + A[m][2] * p[2];
            blo[j][m] = std::min(blo[j][m], s);
            bhi[j][m] = std::max(bhi[j][m], s);
          }
        }
      }
      // suffix boxes, then sweep the left side
      double slo[NB + 1][3], shi[NB + 1][3];
      int sn[NB + 1];
      for (int m = 0; m < 3; m++) {
        slo[NB][m] = 1e300;
        shi[NB][m] = -1e300;
      }
      sn[NB] = 0;
      for (int j = NB - 1; j >= 0; j--) {
        for (int m = 0; m < 3; m++) {
          slo[j][m] = std::min(slo[j + 1][m], blo[j][m]);
          shi[j][m] = std::max(shi[j + 1][m], bhi[j][m]);
        }
        sn[j] = sn[j + 1] + bn[j];
      }
      double llo[3] = {1e300, 1e300, 1e300}, lhi[3] = {-1e300, -1e300, -1e300};
      int ln = 0;
      auto area = [](const double* lo, const double* hi) {
        const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return x * y + y * z + z * x;
      };
      for (int j = 0; j + 1 < NB; j++) {
        for (int m = 0; m < 3; m++) {
          llo[m] = std::min(llo[m], blo[j][m]);
          lhi[m] = std::max(lhi[m], bhi[j][m]);
        }
        ln += bn[j];
        if (ln == 0 || sn[j + 1] == 0) continue;
        const double c = area(llo, lhi) * ln + area(slo[j + 1], shi[j + 1]) * sn[j + 1];
        if (c < best) {
          best = c;
          bestk = k;
          bestb = j + 1;
        }
      }
    }
    if (bestk < 0) {
      // every centroid in one bin: median along the first axis
      const int mid = b + n / 2;
      std::nth_element(idx + b, idx + mid, idx + e, [&](int a, int c) {
        const double pa = proj(0, a), pc = proj(0, c);
        return pa < pc || (pa == pc && a < c);
      });
      return mid;
    }
    double cmin = 1e300, cmax = -1e300;
    for (int i = b; i < e; i++) {
      const double c = proj(bestk, idx[i]);
      cmin = std::min(cmin, c);
      cmax = std::max(cmax, c);
    }
    const double scale = NB / (cmax - cmin);
    auto left = [&](int t) { return std::min(NB - 1, (int)((proj(bestk, t) - cmin) * scale)) < bestb; };
    return (int)(std::stable_partition(idx + b, idx + e, left) - idx);
  }

  void fit(ObbNode& nd, int b, int e) {
    double mean[3] = {0, 0, 0};
    int nv = 0;
    for (int i = b; i < e; i++)
      for (int v = 0; v < 3; v++) {
        const double* p = V + 3 * T[3 * idx[i] + v];
        for (int k = 0; k < 3; k++) mean[k] += p[k];
        nv++;
      }
    for (int k = 0; k < 3; k++) mean[k] /= nv;
    double C[3][3] = {{0}};
    for (int i = b; i < e; i++)
      for (int v = 0; v < 3; v++) {
        const double* p = V + 3 * T[3 * idx[i] + v];
        double d[3] = {p[0] - mean[0], p[1] - mean[1], p[2] - mean[2]};
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) C[r][c] += d[r] * d[c];
      }
    double E[3][3], w[3];
    jacobi3(C, E, w);
    // sort axes by decreasing eigenvalue
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int c) { return w[a] > w[c]; });
    double ax[3][3];
    for (int k = 0; k < 3; k++)
      for (int r = 0; r < 3; r++) ax[k][r] = E[r][ord[k]];
    // right-handed: ax2 = ax0 x ax1
    double cx[3] = {ax[0][1] * ax[1][2] - ax[0][2] * ax[1][1], ax[0][2] * ax[1][0] - ax[0][0] * ax[1][2],
                    ax[0][0] * ax[1][1] - ax[0][1] * ax[1][0]};
    double n = std::sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    for (int r = 0; r < 3; r++) ax[2][r] = cx[r] / n;
    // keep the mesh axes instead when their box has the smaller surface area
    const double I3[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (box_area(I3, b, e) < box_area(ax, b, e))
      for (int k = 0; k < 3; k++)
        for (int r = 0; r < 3; r++) ax[k][r] = I3[k][r];
    // round axes to FP32
    double Rf[3][3];
    for (int k = 0; k < 3; k++)
      for (int r = 0; r < 3; r++) {
        nd.R[k * 3 + r] = (float)ax[k][r];
        Rf[k][r] = (double)nd.R[k * 3 + r];
      }
    // centre from the extent midpoints, rounded to FP32
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = b; i < e; i++)
      for (int v = 0; v < 3; v++) {
        const double* p = V + 3 * T[3 * idx[i] + v];
        for (int k = 0; k < 3; k++) {
          double s = Rf[k][0] * p[0] + Rf[k][1] * p[1] + Rf[k][2] * p[2];
          lo[k] = std::min(lo[k], s);
          hi[k] = std::max(hi[k], s);
        }
      }
    double c[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
      double mid = 0.5 * (lo[k] + hi[k]);
      for (int r = 0; r < 3; r++) c[r] += mid * Rf[k][r];
    }
    for (int r = 0; r < 3; r++) nd.c[r] = (float)c[r];
    // extents in the dual basis of the rounded axes: s = M^{-1}(p - c)
    double M[3][3];  // columns = axes
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) M[r][k] = Rf[k][r];
    double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                 M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                 M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    double Mi[3][3];
    Mi[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    Mi[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    Mi[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    Mi[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    Mi[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    Mi[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    Mi[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    Mi[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    Mi[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    double cf[3] = {nd.c[0], nd.c[1], nd.c[2]};
    double ext[3] = {0, 0, 0};
    for (int i = b; i < e; i++)
      for (int v = 0; v < 3; v++) {
        const double* p = V + 3 * T[3 * idx[i] + v];
        double d[3] = {p[0] - cf[0], p[1] - cf[1], p[2] - cf[2]};
        for (int k = 0; k < 3; k++) {
          double s = Mi[k][0] * d[0] + Mi[k][1] * d[1] + Mi[k][2] * d[2];
          ext[k] = std::max(ext[k], std::fabs(s));
        }
      }
    for (int k = 0; k < 3; k++) {
      double padded = ext[k] * (1.0 + 1e-12) + 1e-300;
      float f = (float)padded;
      if ((double)f < padded) f = std::nextafter(f, 1e30f);
      nd.e[k] = f;
    }
  }

  int build(int b, int e, int d) {
    int self = (int)nodes.size();
    nodes.push_back(ObbNode{});
    depth = std::max(depth, d);
    fit(nodes[self], b, e);
    if (e - b == 1) {
      nodes[self].child = ~(int)order.size();
      order.push_back(idx[b]);
      return self;
    }
    const int mid = split(b, e, nodes[self].R);
    int l = (int)nodes.size();
    nodes.push_back(ObbNode{});
    nodes.push_back(ObbNode{});
    nodes[self].child = l;
    build_into(l, b, mid, d + 1);
    build_into(l + 1, mid, e, d + 1);
    return self;
  }

  // Build the subtree for [b,e) into the preallocated slot `at`.
  void build_into(int at, int b, int e, int d) {
    depth = std::max(depth, d);
    fit(nodes[at], b, e);
    if (e - b == 1) {
      nodes[at].child = ~(int)order.size();
      order.push_back(idx[b]);
      return;
    }
    const int mid = split(b, e, nodes[at].R);
    int l = (int)nodes.size();
    nodes.push_back(ObbNode{});
    nodes.push_back(ObbNode{});
    nodes[at].child = l;
    build_into(l, b, mid, d + 1);
    build_into(l + 1, mid, e, d + 1);
  }
};

// An open subtree of the top levels: triangles [b,e) at depth d, its root
// going into node slot `at`.
struct Task {
  int at, b, e, d;
};

// Top levels: fit and split down to ranges of at most `cut` triangles, which
// become tasks (the serial build's recursion, stopped early).
void build_top(Builder& B, int at, int b, int e, int d, int cut, std::vector<Task>& tasks) {
  if (e - b <= cut) {
    tasks.push_back({at, b, e, d});
    return;
  }
  B.depth = std::max(B.depth, d);
  B.fit(B.nodes[at], b, e);
  const int mid = B.split(b, e, B.nodes[at].R);
  const int l = (int)B.nodes.size();
  B.nodes.push_back(ObbNode{});
  B.nodes.push_back(ObbNode{});
  B.nodes[at].child = l;
  build_top(B, l, b, mid, d + 1, cut, tasks);
  build_top(B, l + 1, mid, e, d + 1, cut, tasks);
}

}  // namespace

void kg_build_obb_tree(const double* V, int nv, const int32_t* T, int nt, BuiltTree& out) {
  (void)nv;
  std::vector<double> cen(3 * (size_t)nt);
  std::vector<int> idx(nt);
  for (int i = 0; i < nt; i++) {
    idx[i] = i;
    for (int k = 0; k < 3; k++)
      cen[3 * i + k] = (V[3 * T[3 * i] + k] + V[3 * T[3 * i + 1] + k] + V[3 * T[3 * i + 2] + k]) / 3.0;
  }
  Builder B;
  B.V = V;
  B.T = T;
  B.cen = cen.data();
  B.idx = idx.data();
  B.nodes.reserve(2 * (size_t)nt);
  B.order.reserve(nt);
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (nt < 8192 || hw == 1) {
    if (nt > 0) B.build(0, nt, 0);
  } else {
    B.nodes.push_back(ObbNode{});
    std::vector<Task> tasks;
    build_top(B, 0, 0, nt, 0, std::max(2048, nt / (int)(4 * hw)), tasks);
    std::vector<Builder> sub(tasks.size());
    std::vector<std::thread> th;
    std::atomic_int next(0);
    for (unsigned w = 0; w < hw; w++)
      th.emplace_back([&]() {
        for (int k = next.fetch_add(1); k < (int)tasks.size(); k = next.fetch_add(1)) {
          Builder& S = sub[k];
          S.V = V;
          S.T = T;
          S.cen = cen.data();
          S.idx = idx.data();
          S.nodes.reserve(2 * (size_t)(tasks[k].e - tasks[k].b));
          S.build(tasks[k].b, tasks[k].e, tasks[k].d);
        }
      });
    for (auto& t : th) t.join();
    // splice: subtree node 0 goes to its slot, nodes 1.. are appended
    for (size_t k = 0; k < tasks.size(); k++) {
      const Builder& S = sub[k];
      const int base = (int)B.nodes.size() - 1, obase = (int)B.order.size();
      auto remap = [&](ObbNode nd) {
        nd.child = nd.child >= 0 ? nd.child + base : ~(~nd.child + obase);
        return nd;
      };
      B.nodes[tasks[k].at] = remap(S.nodes[0]);
      for (size_t i = 1; i < S.nodes.size(); i++) B.nodes.push_back(remap(S.nodes[i]));
      B.order.insert(B.order.end(), S.order.begin(), S.order.end());
      B.depth = std::max(B.depth, S.depth);
    }
  }
  out.nodes.swap(B.nodes);
  out.order.swap(B.order);
  out.depth = B.depth;
}
