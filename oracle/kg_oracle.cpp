// FP64 CPU oracle for the collision / distance / feasibility path.
//
// TEST INFRASTRUCTURE ONLY.  It is loaded by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg as the checker; the product library never
// links or calls it.
//
// It restates, in double precision and compiled with -ffp-contract=off:
//   * Klampt's policy layer, line by line:
//       - collision mask defaults       Cpp/Planning/PlannerSettings.cpp:16-41
//       - world ID layout               Cpp/Modeling/World.cpp:47-53,156-180
//       - GetGeometries / AABB broad phase / set-level quick reject with its
//         swap-erase reordering / first hit
//                                       Cpp/Planning/PlannerSettings.cpp:214-331
//       - self-collision mask quirk (i,j)||(i,i)
//                                       Cpp/Planning/PlannerSettings.cpp:258
//       - IsFeasible: joint limits (closed interval), env before self
//                                       Cpp/Planning/RobotCSpace.cpp:610-637,786-823
//       - default self pairs: all but link-vs-parent
//                                       Cpp/Docs/Manual-FileTypes.md:212,
//                                       Cpp/Modeling/Robot.cpp:1274-1313
//       - interpolation x*(1-u)+y*u     Cpp/Modeling/Interpolate.cpp:13-14
//       - segment length (L2 / Linf)    Cpp/Modeling/Interpolate.cpp:208-274
//       - discretisation n=ceil(d/eps), u=i/n
//                                       Python/klampt/plan/robotcspace.py:216-228
//   * the KrisLibrary arithmetic (absent from the reference tree) from
//     published methods:
//       - FK: T_World[i] = T_World[p] * T0_Parent[i] * J(w,q), Rodrigues
//         rotation (Python/klampt/math/so3.py:308-322), column-major R;
//       - OBB overlap test with 15 separating axes (Gottschalk, Lin, Manocha,
//         "OBBTree", SIGGRAPH 1996) on local-frame AABB trees;
//       - triangle/triangle intersection by orientation predicates: two
//         closed triangles meet iff an edge of one meets the other (or they
//         are coplanar and overlap in their plane);
//       - triangle/triangle distance: min over 9 segment pairs and 6
//         vertex/triangle pairs (0 when intersecting), closest points by
//         Voronoi-region tests (Ericson, Real-Time Collision Detection, 5.1);
//       - cell-centred trilinear SDF (Python/klampt/src/geometry.h:322-347).
//
// Behaviour that only KrisLibrary could pin (element pair choice, margin
// combination, edge-checker sample order, SDF clamping) is *defined* here and
// documented at each function.

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <thread>
#include <vector>

namespace orc {

static const double kInf = std::numeric_limits<double>::infinity();

struct V3 {
  double x, y, z;
};
static inline V3 mk(double x, double y, double z) { return V3{x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator*(double s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Rigid transform, R column-major (R[c*3+r]).
struct Xf {
  double R[9];
  double t[3];
};
static inline V3 rot(const Xf& X, V3 p) {
  return mk(X.R[0] * p.x + X.R[3] * p.y + X.R[6] * p.z,
            X.R[1] * p.x + X.R[4] * p.y + X.R[7] * p.z,
            X.R[2] * p.x + X.R[5] * p.y + X.R[8] * p.z);
}
static inline V3 apply(const Xf& X, V3 p) { return rot(X, p) + mk(X.t[0], X.t[1], X.t[2]); }
static Xf compose(const Xf& A, const Xf& B) {  // A*B
  Xf C;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++)
      C.R[c * 3 + r] = A.R[0 * 3 + r] * B.R[c * 3 + 0] + A.R[1 * 3 + r] * B.R[c * 3 + 1] +
                       A.R[2 * 3 + r] * B.R[c * 3 + 2];
  V3 t = apply(A, mk(B.t[0], B.t[1], B.t[2]));
  C.t[0] = t.x; C.t[1] = t.y; C.t[2] = t.z;
  return C;
}
static Xf inverse(const Xf& A) {
  Xf I;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) I.R[c * 3 + r] = A.R[r * 3 + c];
  V3 t = rot(I, mk(A.t[0], A.t[1], A.t[2]));
  I.t[0] = -t.x; I.t[1] = -t.y; I.t[2] = -t.z;
  return I;
}
static Xf identity() {
  Xf X;
  for (int i = 0; i < 9; i++) X.R[i] = (i % 4 == 0) ? 1.0 : 0.0;
  X.t[0] = X.t[1] = X.t[2] = 0;
  return X;
}
static Xf from12(const double* x) {
  Xf X;
  for (int i = 0; i < 9; i++) X.R[i] = x[i];
  for (int i = 0; i < 3; i++) X.t[i] = x[9 + i];
  return X;
}

// Rodrigues: R = c I + (1-c) w w^T + s [w]x  (so3.rotation, so3.py:308-322)
static Xf joint_xf(int type, V3 w, double q) {
  Xf J = identity();
  if (type == 1) {  // prismatic: translation w*q
    J.t[0] = w.x * q; J.t[1] = w.y * q; J.t[2] = w.z * q;
    return J;
  }
  double c = std::cos(q), s = std::sin(q), v = 1.0 - c;
  double W[3] = {w.x, w.y, w.z};
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) J.R[col * 3 + row] = W[row] * W[col] * v;
  J.R[0] += c; J.R[4] += c; J.R[8] += c;
  // [w]x column-major: col0=(0,w2,-w1) col1=(-w2,0,w0) col2=(w1,-w0,0)
  J.R[1] += s * W[2]; J.R[2] -= s * W[1];
  J.R[3] -= s * W[2]; J.R[5] += s * W[0];
  J.R[6] += s * W[1]; J.R[7] -= s * W[0];
  return J;
}

// ------------------------------------------------------------------ meshes

struct Node {
  double lo[3], hi[3];
  int left, right, tri;  // tri >= 0 => leaf
};

// A node of an oriented-box tree given from outside (the engine's BVH, for
// counting traversal work on the same tree): points c + R diag(e) s,
// |s_k| <= 1, R column-major; child >= 0 is the first of two adjacent
// children, child < 0 a leaf holding triangle ~child.
struct ObbN {
  double c[3], e[3], R[9];
  int child;
};

struct Mesh {
  std::vector<V3> V;
  std::vector<std::array<int, 3>> T;
  std::vector<Node> nodes;
  std::vector<ObbN> obb;  // optional (or_mesh_set_obb)
  double lmin[3], lmax[3];
  bool empty() const { return T.empty(); }
};

static void build_rec(Mesh& m, std::vector<int>& idx, int b, int e, std::vector<V3>& cen, int self) {
  Node& nd = m.nodes[self];
  for (int k = 0; k < 3; k++) { nd.lo[k] = kInf; nd.hi[k] = -kInf; }
  for (int i = b; i < e; i++)
    for (int v = 0; v < 3; v++) {
      V3 p = m.V[m.T[idx[i]][v]];
      for (int k = 0; k < 3; k++) {
        nd.lo[k] = std::min(nd.lo[k], comp(p, k));
        nd.hi[k] = std::max(nd.hi[k], comp(p, k));
      }
    }
  if (e - b == 1) {
    nd.tri = idx[b]; nd.left = nd.right = -1;
    return;
  }
  double clo[3] = {kInf, kInf, kInf}, chi[3] = {-kInf, -kInf, -kInf};
  for (int i = b; i < e; i++)
    for (int k = 0; k < 3; k++) {
      clo[k] = std::min(clo[k], comp(cen[idx[i]], k));
      chi[k] = std::max(chi[k], comp(cen[idx[i]], k));
    }
  int ax = 0;
  for (int k = 1; k < 3; k++)
    if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
  int mid = (b + e) / 2;
  std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                   [&](int a, int c) { return comp(cen[a], ax) < comp(cen[c], ax); });
  int l = (int)m.nodes.size();
  m.nodes.push_back(Node{});
  int r = (int)m.nodes.size();
  m.nodes.push_back(Node{});
  m.nodes[self].left = l; m.nodes[self].right = r; m.nodes[self].tri = -1;
  build_rec(m, idx, b, mid, cen, l);
  build_rec(m, idx, mid, e, cen, r);
}

static void build(Mesh& m) {
  m.nodes.clear();
  for (int k = 0; k < 3; k++) { m.lmin[k] = kInf; m.lmax[k] = -kInf; }
  for (auto& p : m.V)
    for (int k = 0; k < 3; k++) {
      m.lmin[k] = std::min(m.lmin[k], comp(p, k));
      m.lmax[k] = std::max(m.lmax[k], comp(p, k));
    }
  if (m.T.empty()) return;
  std::vector<int> idx(m.T.size());
  std::vector<V3> cen(m.T.size());
  for (size_t i = 0; i < m.T.size(); i++) {
    idx[i] = (int)i;
    cen[i] = (1.0 / 3.0) * (m.V[m.T[i][0]] + m.V[m.T[i][1]] + m.V[m.T[i][2]]);
  }
  m.nodes.reserve(2 * m.T.size());
  m.nodes.push_back(Node{});
  build_rec(m, idx, 0, (int)m.T.size(), cen, 0);
}

// ------------------------------------------------------------------ predicates

// Signed volume (b-a)x(c-a).(d-a)
static inline double orient3d(V3 a, V3 b, V3 c, V3 d) { return dot(cross(b - a, c - a), d - a); }

static inline double orient2d(double ax, double ay, double bx, double by, double cx, double cy) {
  return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}

// 2D closed segment/segment intersection.
static bool seg_seg_2d(const double* p, const double* q, const double* a, const double* b) {
  double d1 = orient2d(a[0], a[1], b[0], b[1], p[0], p[1]);
  double d2 = orient2d(a[0], a[1], b[0], b[1], q[0], q[1]);
  double d3 = orient2d(p[0], p[1], q[0], q[1], a[0], a[1]);
  double d4 = orient2d(p[0], p[1], q[0], q[1], b[0], b[1]);
  if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0)))
    return true;
  auto on = [](const double* s, const double* e, const double* r, double o) {
    return o == 0 && std::min(s[0], e[0]) <= r[0] && r[0] <= std::max(s[0], e[0]) &&
           std::min(s[1], e[1]) <= r[1] && r[1] <= std::max(s[1], e[1]);
  };
  return on(a, b, p, d1) || on(a, b, q, d2) || on(p, q, a, d3) || on(p, q, b, d4);
}

static bool point_in_tri_2d(const double* p, const double* a, const double* b, const double* c) {
  double d1 = orient2d(a[0], a[1], b[0], b[1], p[0], p[1]);
  double d2 = orient2d(b[0], b[1], c[0], c[1], p[0], p[1]);
  double d3 = orient2d(c[0], c[1], a[0], a[1], p[0], p[1]);
  bool neg = d1 < 0 || d2 < 0 || d3 < 0, pos = d1 > 0 || d2 > 0 || d3 > 0;
  return !(neg && pos);
}

// Drop the dominant coordinate of the normal n.
static void proj2(V3 n, V3 p, double* out) {
  double ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
  if (ax >= ay && ax >= az) { out[0] = p.y; out[1] = p.z; }
  else if (ay >= az) { out[0] = p.x; out[1] = p.z; }
  else { out[0] = p.x; out[1] = p.y; }
}

// Closed segment PQ vs closed triangle ABC lying in a common plane with normal n.
static bool coplanar_seg_tri(V3 P, V3 Q, V3 A, V3 B, V3 C, V3 n) {
  double p[2], q[2], a[2], b[2], c[2];
  proj2(n, P, p); proj2(n, Q, q); proj2(n, A, a); proj2(n, B, b); proj2(n, C, c);
  if (point_in_tri_2d(p, a, b, c) || point_in_tri_2d(q, a, b, c)) return true;
  return seg_seg_2d(p, q, a, b) || seg_seg_2d(p, q, b, c) || seg_seg_2d(p, q, c, a);
}

// Closed segment PQ vs closed triangle ABC.
static bool seg_tri(V3 P, V3 Q, V3 A, V3 B, V3 C) {
  double sp = orient3d(A, B, C, P), sq = orient3d(A, B, C, Q);
  if ((sp > 0 && sq > 0) || (sp < 0 && sq < 0)) return false;
  if (sp == 0 && sq == 0) return coplanar_seg_tri(P, Q, A, B, C, cross(B - A, C - A));
  double s1 = orient3d(P, Q, A, B), s2 = orient3d(P, Q, B, C), s3 = orient3d(P, Q, C, A);
  return (s1 >= 0 && s2 >= 0 && s3 >= 0) || (s1 <= 0 && s2 <= 0 && s3 <= 0);
}

// Closed triangle/triangle intersection.
bool tri_tri(const V3* t1, const V3* t2) {
  // coplanar test first: all of t1 on t2's plane
  double o0 = orient3d(t2[0], t2[1], t2[2], t1[0]);
  double o1 = orient3d(t2[0], t2[1], t2[2], t1[1]);
  double o2 = orient3d(t2[0], t2[1], t2[2], t1[2]);
  if ((o0 > 0 && o1 > 0 && o2 > 0) || (o0 < 0 && o1 < 0 && o2 < 0)) return false;
  double r0 = orient3d(t1[0], t1[1], t1[2], t2[0]);
  double r1 = orient3d(t1[0], t1[1], t1[2], t2[1]);
  double r2 = orient3d(t1[0], t1[1], t1[2], t2[2]);
  if ((r0 > 0 && r1 > 0 && r2 > 0) || (r0 < 0 && r1 < 0 && r2 < 0)) return false;
  for (int e = 0; e < 3; e++) {
    if (seg_tri(t1[e], t1[(e + 1) % 3], t2[0], t2[1], t2[2])) return true;
    if (seg_tri(t2[e], t2[(e + 1) % 3], t1[0], t1[1], t1[2])) return true;
  }
  if (o0 == 0 && o1 == 0 && o2 == 0) {
    // coplanar: one triangle may contain the other entirely
    V3 n = cross(t2[1] - t2[0], t2[2] - t2[0]);
    double p[2], a[2], b[2], c[2];
    proj2(n, t2[0], a); proj2(n, t2[1], b); proj2(n, t2[2], c);
    proj2(n, t1[0], p);
    if (point_in_tri_2d(p, a, b, c)) return true;
    proj2(n, t1[0], a); proj2(n, t1[1], b); proj2(n, t1[2], c);
    proj2(n, t2[0], p);
    if (point_in_tri_2d(p, a, b, c)) return true;
  }
  return false;
}

// Closest points of segments P0P1, Q0Q1 (Ericson 5.1.9); returns squared distance.
static double seg_seg_dist2(V3 p0, V3 p1, V3 q0, V3 q1, V3* cp, V3* cq) {
  V3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
  double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  double s, t;
  if (a <= 0 && e <= 0) { s = t = 0; }
  else if (a <= 0) { s = 0; t = std::min(1.0, std::max(0.0, f / e)); }
  else {
    double c = dot(d1, r);
    if (e <= 0) { t = 0; s = std::min(1.0, std::max(0.0, -c / a)); }
    else {
      double b = dot(d1, d2), den = a * e - b * b;
      s = den != 0 ? std::min(1.0, std::max(0.0, (b * f - c * e) / den)) : 0.0;
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = std::min(1.0, std::max(0.0, -c / a)); }
      else if (t > 1) { t = 1; s = std::min(1.0, std::max(0.0, (b - c) / a)); }
    }
  }
  V3 c1 = p0 + s * d1, c2 = q0 + t * d2;
  if (cp) *cp = c1;
  if (cq) *cq = c2;
  V3 d = c1 - c2;
  return dot(d, d);
}

// Closest point on triangle ABC to P (Ericson 5.1.5).
static V3 closest_pt_tri(V3 p, V3 a, V3 b, V3 c) {
  V3 ab = b - a, ac = c - a, ap = p - a;
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0 && d2 <= 0) return a;
  V3 bp = p - b;
  double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0 && d4 <= d3) return b;
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) return a + (d1 / (d1 - d3)) * ab;
  V3 cp = p - c;
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0 && d5 <= d6) return c;
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) return a + (d2 / (d2 - d6)) * ac;
  double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0)
    return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
  double den = 1.0 / (va + vb + vc);
  return a + (vb * den) * ab + (vc * den) * ac;
}

// Triangle/triangle distance (0 when intersecting) with closest points.
double tri_tri_dist(const V3* t1, const V3* t2, V3* c1, V3* c2) {
  if (tri_tri(t1, t2)) {
    // report a point of the intersection: an edge/triangle crossing point
    if (c1) *c1 = t1[0];
    if (c2) *c2 = t1[0];
    return 0.0;
  }
  double best = kInf;
  V3 b1 = t1[0], b2 = t2[0];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      V3 p, q;
      double d = seg_seg_dist2(t1[i], t1[(i + 1) % 3], t2[j], t2[(j + 1) % 3], &p, &q);
      if (d < best) { best = d; b1 = p; b2 = q; }
    }
  for (int i = 0; i < 3; i++) {
    V3 q = closest_pt_tri(t1[i], t2[0], t2[1], t2[2]);
    V3 d = t1[i] - q;
    if (dot(d, d) < best) { best = dot(d, d); b1 = t1[i]; b2 = q; }
    q = closest_pt_tri(t2[i], t1[0], t1[1], t1[2]);
    d = t2[i] - q;
    if (dot(d, d) < best) { best = dot(d, d); b1 = q; b2 = t2[i]; }
  }
  if (c1) *c1 = b1;
  if (c2) *c2 = b2;
  return std::sqrt(best);
}

// Penetration depth of two intersecting triangles: the smallest overlap of
// their projections over the 17 candidate axes (2 normals, 9 edge crosses,
// 6 in-plane edge normals), i.e. the distance from the origin to the boundary
// of the Minkowski difference.  Used to classify the contact band.
double tri_tri_depth(const V3* t1, const V3* t2) {
  V3 e1[3] = {t1[1] - t1[0], t1[2] - t1[1], t1[0] - t1[2]};
  V3 e2[3] = {t2[1] - t2[0], t2[2] - t2[1], t2[0] - t2[2]};
  V3 n1 = cross(e1[0], e1[1]), n2 = cross(e2[0], e2[1]);
  V3 axes[17];
  int na = 0;
  axes[na++] = n1; axes[na++] = n2;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) axes[na++] = cross(e1[i], e2[j]);
  for (int i = 0; i < 3; i++) { axes[na++] = cross(n1, e1[i]); axes[na++] = cross(n2, e2[i]); }
  double best = kInf;
  for (int k = 0; k < na; k++) {
    double L = std::sqrt(dot(axes[k], axes[k]));
    if (L < 1e-300) continue;
    V3 a = (1.0 / L) * axes[k];
    double lo1 = kInf, hi1 = -kInf, lo2 = kInf, hi2 = -kInf;
    for (int v = 0; v < 3; v++) {
      double p = dot(a, t1[v]), q = dot(a, t2[v]);
      lo1 = std::min(lo1, p); hi1 = std::max(hi1, p);
      lo2 = std::min(lo2, q); hi2 = std::max(hi2, q);
    }
    double ov = std::min(hi1 - lo2, hi2 - lo1);
    best = std::min(best, ov);
  }
  return std::max(0.0, best);
}

// ------------------------------------------------------------------ BVH queries

// OBB overlap (Gottschalk 15-axis test).  Box A axis-aligned in frame A,
// box B given by centre/half extents in frame B; X maps frame B -> frame A.
// Returns true if separated by more than pad.
static bool obb_disjoint(const Node& na, const Node& nb, const Xf& X, double pad) {
  double a[3], b[3], ca[3], cbl[3];
  for (int k = 0; k < 3; k++) {
    a[k] = 0.5 * (na.hi[k] - na.lo[k]) + pad;
    ca[k] = 0.5 * (na.hi[k] + na.lo[k]);
    b[k] = 0.5 * (nb.hi[k] - nb.lo[k]);
    cbl[k] = 0.5 * (nb.hi[k] + nb.lo[k]);
  }
  V3 cb = apply(X, mk(cbl[0], cbl[1], cbl[2]));
  double T[3] = {cb.x - ca[0], cb.y - ca[1], cb.z - ca[2]};
  double R[3][3], Q[3][3];  // R[i][j] = A_i . B_j
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      R[i][j] = X.R[j * 3 + i];
      Q[i][j] = std::fabs(R[i][j]) + 1e-15;
    }
  for (int i = 0; i < 3; i++)
    if (std::fabs(T[i]) > a[i] + b[0] * Q[i][0] + b[1] * Q[i][1] + b[2] * Q[i][2]) return true;
  for (int j = 0; j < 3; j++) {
    double t = T[0] * R[0][j] + T[1] * R[1][j] + T[2] * R[2][j];
    if (std::fabs(t) > a[0] * Q[0][j] + a[1] * Q[1][j] + a[2] * Q[2][j] + b[j]) return true;
  }
  for (int i = 0; i < 3; i++) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; j++) {
      int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      double t = T[i2] * R[i1][j] - T[i1] * R[i2][j];
      double ra = a[i1] * Q[i2][j] + a[i2] * Q[i1][j];
      double rb = b[j1] * Q[i][j2] + b[j2] * Q[i][j1];
      if (std::fabs(t) > ra + rb) return true;
    }
  }
  return false;
}

// Lower bound on the distance between the two boxes: the larger of the
// bounding-sphere gap and the largest normalised projection gap over the 15
// axes (projections are 1-Lipschitz).
static double obb_lower_bound(const Node& na, const Node& nb, const Xf& X) {
  double a[3], b[3], ca[3], cbl[3];
  for (int k = 0; k < 3; k++) {
    a[k] = 0.5 * (na.hi[k] - na.lo[k]);
    ca[k] = 0.5 * (na.hi[k] + na.lo[k]);
    b[k] = 0.5 * (nb.hi[k] - nb.lo[k]);
    cbl[k] = 0.5 * (nb.hi[k] + nb.lo[k]);
  }
  V3 cb = apply(X, mk(cbl[0], cbl[1], cbl[2]));
  double T[3] = {cb.x - ca[0], cb.y - ca[1], cb.z - ca[2]};
  double R[3][3], Q[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      R[i][j] = X.R[j * 3 + i];
      Q[i][j] = std::fabs(R[i][j]);
    }
  // bounding-sphere bound (tight for far, diagonal pairs)
  double lb = std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]) -
              std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) -
              std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  lb = std::max(lb, 0.0);
  for (int i = 0; i < 3; i++)
    lb = std::max(lb, std::fabs(T[i]) - (a[i] + b[0] * Q[i][0] + b[1] * Q[i][1] + b[2] * Q[i][2]));
  for (int j = 0; j < 3; j++) {
    double t = T[0] * R[0][j] + T[1] * R[1][j] + T[2] * R[2][j];
    lb = std::max(lb, std::fabs(t) - (a[0] * Q[0][j] + a[1] * Q[1][j] + a[2] * Q[2][j] + b[j]));
  }
  for (int i = 0; i < 3; i++) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; j++) {
      int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      double L2 = 1.0 - R[i][j] * R[i][j];
      if (L2 < 1e-12) continue;
      double t = T[i2] * R[i1][j] - T[i1] * R[i2][j];
      double ra = a[i1] * Q[i2][j] + a[i2] * Q[i1][j];
      double rb = b[j1] * Q[i][j2] + b[j2] * Q[i][j1];
      lb = std::max(lb, (std::fabs(t) - ra - rb) / std::sqrt(L2));
    }
  }
  return lb;
}

static inline double node_vol(const Node& n) {
  return (n.hi[0] - n.lo[0]) * (n.hi[1] - n.lo[1]) * (n.hi[2] - n.lo[2]);
}

static inline void tri_world(const Mesh& m, int t, const Xf& X, V3* out) {
  for (int v = 0; v < 3; v++) out[v] = apply(X, m.V[m.T[t][v]]);
}

// Statistics of one traversal (used to define algorithmic bytes).
struct Counters {
  int64_t bv_tests = 0, tri_tests = 0;
};

// Mesh/mesh collision (margin 0).  Frames: XA, XB world transforms.
static bool mesh_collide(const Mesh& A, const Xf& XA, const Mesh& B, const Xf& XB, bool brute,
                         Counters* cnt = nullptr) {
  if (A.empty() || B.empty()) return false;
  if (brute) {
    V3 ta[3], tb[3];
    for (size_t i = 0; i < A.T.size(); i++) {
      tri_world(A, (int)i, XA, ta);
      for (size_t j = 0; j < B.T.size(); j++) {
        tri_world(B, (int)j, XB, tb);
        if (tri_tri(ta, tb)) return true;
      }
    }
    return false;
  }
  Xf X = compose(inverse(XA), XB);  // B local -> A local
  std::vector<std::pair<int, int>> st;
  st.reserve(256);
  st.push_back({0, 0});
  V3 ta[3], tb[3];
  while (!st.empty()) {
    auto [ia, ib] = st.back();
    st.pop_back();
    const Node& na = A.nodes[ia];
    const Node& nb = B.nodes[ib];
    if (cnt) cnt->bv_tests++;
    if (obb_disjoint(na, nb, X, 1e-12)) continue;
    bool la = na.tri >= 0, lb = nb.tri >= 0;
    if (la && lb) {
      if (cnt) cnt->tri_tests++;
      tri_world(A, na.tri, XA, ta);
      tri_world(B, nb.tri, XB, tb);
      if (tri_tri(ta, tb)) return true;
    } else if (lb || (!la && node_vol(na) >= node_vol(nb))) {
      st.push_back({na.right, ib});
      st.push_back({na.left, ib});
    } else {
      st.push_back({ia, nb.right});
      st.push_back({ia, nb.left});
    }
  }
  return false;
}

// Mesh/mesh distance (0 when touching), optionally with closest points.
static double mesh_distance(const Mesh& A, const Xf& XA, const Mesh& B, const Xf& XB, bool brute,
                            V3* cp1 = nullptr, V3* cp2 = nullptr, int* e1 = nullptr, int* e2 = nullptr) {
  if (A.empty() || B.empty()) return kInf;
  double best = kInf;
  V3 ta[3], tb[3], p, q;
  int bi = -1, bj = -1;
  V3 bp{0, 0, 0}, bq{0, 0, 0};
  if (brute) {
    for (size_t i = 0; i < A.T.size(); i++) {
      tri_world(A, (int)i, XA, ta);
      for (size_t j = 0; j < B.T.size(); j++) {
        tri_world(B, (int)j, XB, tb);
        double d = tri_tri_dist(ta, tb, &p, &q);
        if (d < best) { best = d; bi = (int)i; bj = (int)j; bp = p; bq = q; }
      }
    }
  } else {
    Xf X = compose(inverse(XA), XB);
    std::vector<std::pair<int, int>> st;
    st.push_back({0, 0});
    while (!st.empty()) {
      auto [ia, ib] = st.back();
      st.pop_back();
      const Node& na = A.nodes[ia];
      const Node& nb = B.nodes[ib];
      if (obb_lower_bound(na, nb, X) >= best) continue;
      bool la = na.tri >= 0, lb = nb.tri >= 0;
      if (la && lb) {
        tri_world(A, na.tri, XA, ta);
        tri_world(B, nb.tri, XB, tb);
        double d = tri_tri_dist(ta, tb, &p, &q);
        if (d < best) { best = d; bi = na.tri; bj = nb.tri; bp = p; bq = q; }
        if (best == 0) break;
      } else {
        int c0a, c0b, c1a, c1b;
        if (lb || (!la && node_vol(na) >= node_vol(nb))) {
          c0a = na.left; c1a = na.right; c0b = c1b = ib;
        } else {
          c0a = c1a = ia; c0b = nb.left; c1b = nb.right;
        }
        double d0 = obb_lower_bound(A.nodes[c0a], B.nodes[c0b], X);
        double d1 = obb_lower_bound(A.nodes[c1a], B.nodes[c1b], X);
        // visit the closer child first (pushed last)
        if (d0 <= d1) { st.push_back({c1a, c1b}); st.push_back({c0a, c0b}); }
        else { st.push_back({c0a, c0b}); st.push_back({c1a, c1b}); }
      }
    }
  }
  if (cp1) *cp1 = bp;
  if (cp2) *cp2 = bq;
  if (e1) *e1 = bi;
  if (e2) *e2 = bj;
  return best;
}

// Signed separation used for the contact band: +distance when disjoint,
// -(largest triangle-pair penetration depth) when intersecting.
static double mesh_separation(const Mesh& A, const Xf& XA, const Mesh& B, const Xf& XB) {
  if (A.empty() || B.empty()) return kInf;
  Xf X = compose(inverse(XA), XB);
  std::vector<std::pair<int, int>> st;
  st.push_back({0, 0});
  V3 ta[3], tb[3];
  double depth = -1;
  while (!st.empty()) {
    auto [ia, ib] = st.back();
    st.pop_back();
    const Node& na = A.nodes[ia];
    const Node& nb = B.nodes[ib];
    if (obb_disjoint(na, nb, X, 1e-12)) continue;
    bool la = na.tri >= 0, lb = nb.tri >= 0;
    if (la && lb) {
      tri_world(A, na.tri, XA, ta);
      tri_world(B, nb.tri, XB, tb);
      if (tri_tri(ta, tb)) depth = std::max(depth, tri_tri_depth(ta, tb));
    } else if (lb || (!la && node_vol(na) >= node_vol(nb))) {
      st.push_back({na.right, ib});
      st.push_back({na.left, ib});
    } else {
      st.push_back({ia, nb.right});
      st.push_back({ia, nb.left});
    }
  }
  if (depth >= 0) return -depth;
  return mesh_distance(A, XA, B, XB, false);
}

// World AABB of a mesh placed at X: the transformed local AABB (8 corners).
// (AnyCollisionGeometry3D::GetAABB lives in KrisLibrary; this definition is
// the one the GPU library uses too.)
// Work of one sequential traversal of two OBB trees, to define the
// algorithmic bytes of the collision path: box tests, environment-node
// visits (the root, then two nodes each time the environment side is split;
// a split of the link side re-uses the environment node it holds) and
// triangle tests.
struct ObbCount {
  int64_t box = 0, env_nodes = 0, tri = 0;
};

// 15-axis overlap test of node a (frame A) and node b (frame B), X: B -> A;
// slack: the smallest face-axis margin (how deeply they overlap).
static bool obbn_overlap(const ObbN& a, const ObbN& b, const Xf& X, double& slack) {
  // b's centre and axes in frame A, then in a's axes
  V3 cb = apply(X, mk(b.c[0], b.c[1], b.c[2]));
  V3 d = cb - mk(a.c[0], a.c[1], a.c[2]);
  double T[3], R[3][3], Q[3][3];
  for (int i = 0; i < 3; i++) {
    V3 ai = mk(a.R[3 * i], a.R[3 * i + 1], a.R[3 * i + 2]);
    T[i] = dot(ai, d);
    for (int j = 0; j < 3; j++) {
      V3 bj = rot(X, mk(b.R[3 * j], b.R[3 * j + 1], b.R[3 * j + 2]));
      R[i][j] = dot(ai, bj);
      Q[i][j] = std::fabs(R[i][j]) + 1e-12;
    }
  }
  const double* ea = a.e;
  const double* eb = b.e;
  slack = kInf;
  for (int i = 0; i < 3; i++) {
    const double rhs = ea[i] + eb[0] * Q[i][0] + eb[1] * Q[i][1] + eb[2] * Q[i][2];
    if (std::fabs(T[i]) > rhs) return false;
    slack = std::min(slack, rhs - std::fabs(T[i]));
  }
  for (int j = 0; j < 3; j++) {
    const double t = T[0] * R[0][j] + T[1] * R[1][j] + T[2] * R[2][j];
    const double rhs = ea[0] * Q[0][j] + ea[1] * Q[1][j] + ea[2] * Q[2][j] + eb[j];
    if (std::fabs(t) > rhs) return false;
    slack = std::min(slack, rhs - std::fabs(t));
  }
  for (int i = 0; i < 3; i++) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; j++) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      const double t = T[i2] * R[i1][j] - T[i1] * R[i2][j];
      const double ra = ea[i1] * Q[i2][j] + ea[i2] * Q[i1][j];
      const double rb = eb[j1] * Q[i][j2] + eb[j2] * Q[i][j1];
      if (std::fabs(t) > ra + rb) return false;
    }
  }
  return true;
}

// First-hit collide of link mesh A against mesh B on their OBB trees,
// depth first: the node of larger box volume is split (a leaf never), the
// more deeply overlapping child pair is explored first, the first
// intersecting triangle pair ends the query.  b_env: B is environment
// geometry (its node visits are counted).
static bool obb_collide_count(const Mesh& A, const Xf& XA, const Mesh& B, const Xf& XB, bool b_env, ObbCount& cnt) {
  if (A.empty() || B.empty()) return false;
  const Xf X = compose(inverse(XA), XB);  // B local -> A local
  auto vol = [](const ObbN& n) { return n.e[0] * n.e[1] * n.e[2]; };
  V3 ta[3], tb[3];
  std::vector<std::pair<int, int>> st;
  st.reserve(256);
  double sl;
  cnt.box++;
  if (b_env) cnt.env_nodes++;
  if (!obbn_overlap(A.obb[0], B.obb[0], X, sl)) return false;
  st.push_back({0, 0});
  while (!st.empty()) {
    const auto [ia, ib] = st.back();
    st.pop_back();
    const ObbN& na = A.obb[ia];
    const ObbN& nb = B.obb[ib];
    const bool la = na.child < 0, lb = nb.child < 0;
    if (la && lb) {
      cnt.tri++;
      tri_world(A, ~na.child, XA, ta);
      tri_world(B, ~nb.child, XB, tb);
      if (tri_tri(ta, tb)) return true;
      continue;
    }
    const bool splitA = lb || (!la && vol(na) >= vol(nb));
    std::pair<int, int> kid[2];
    double depth[2];
    int nk = 0;
    for (int k = 0; k < 2; k++) {
      const int ca = splitA ? na.child + k : ia, cb = splitA ? ib : nb.child + k;
      cnt.box++;
      if (b_env && !splitA) cnt.env_nodes++;
      if (obbn_overlap(A.obb[ca], B.obb[cb], X, sl)) {
        kid[nk] = {ca, cb};
        depth[nk] = sl;
        nk++;
      }
    }
    if (nk == 2 && depth[1] > depth[0]) std::swap(kid[0], kid[1]);
    for (int k = nk - 1; k >= 0; k--) st.push_back(kid[k]);  // deeper one on top
  }
  return false;
}

// ------------------------------------------------------------------ convex hulls
//
// Distance between convex hulls of point sets by GJK (Gilbert, Johnson and
// Keerthi 1988) with the closest point of each simplex to the origin found
// by Voronoi regions (Ericson, Real-Time Collision Detection 5.1.5-5.1.6),
// carrying barycentric weights so the closest points on both hulls follow.
// A hull is given by points and one or two rigid placements (two: the hull
// of both placed copies, the swept volume KineTrajOpt checks,
// kinetrajopt.py:603-611 setConvexHullGroup).

struct HullSet {
  const double* P;   // n x 3 local points
  int n;
  Xf X[2];
  int nx;            // placements (1 or 2)
};

static V3 hull_support(const HullSet& h, V3 d, int* which) {
  double best = -kInf;
  V3 out = mk(0, 0, 0);
  for (int k = 0; k < h.nx; k++) {
    const V3 dl = mk(dot(mk(h.X[k].R[0], h.X[k].R[1], h.X[k].R[2]), d),
                     dot(mk(h.X[k].R[3], h.X[k].R[4], h.X[k].R[5]), d),
                     dot(mk(h.X[k].R[6], h.X[k].R[7], h.X[k].R[8]), d));
    const double off = dot(mk(h.X[k].t[0], h.X[k].t[1], h.X[k].t[2]), d);  // the placement's translation
    for (int i = 0; i < h.n; i++) {
      const V3 p = mk(h.P[3 * i], h.P[3 * i + 1], h.P[3 * i + 2]);
      const double v = dot(p, dl) + off;
      if (v > best) {
        best = v;
        out = apply(h.X[k], p);
        if (which) *which = k;
      }
    }
  }
  return out;
}

// Closest point to the origin of the simplex w[0..n), as weights lam over
// its vertices; the simplex is reduced to the vertices with nonzero weight.
static V3 simplex_closest(V3* w, V3* a, V3* b, int& n, double* lam) {
  auto keep = [&](std::initializer_list<int> idx) {
    V3 w2[4], a2[4], b2[4];
    double l2[4];
    int m = 0;
    for (int i : idx) {
      w2[m] = w[i]; a2[m] = a[i]; b2[m] = b[i]; l2[m] = lam[i]; m++;
    }
    for (int i = 0; i < m; i++) { w[i] = w2[i]; a[i] = a2[i]; b[i] = b2[i]; lam[i] = l2[i]; }
    n = m;
  };
  if (n == 1) {
    lam[0] = 1;
    return w[0];
  }
  if (n == 2) {
    const V3 ab = w[1] - w[0];
    const double t = -dot(w[0], ab), den = dot(ab, ab);
    if (t <= 0) { lam[0] = 1; keep({0}); return w[0]; }
    if (t >= den) { lam[1] = 1; keep({1}); return w[0]; }
    lam[1] = t / den; lam[0] = 1 - lam[1];
    return lam[0] * w[0] + lam[1] * w[1];
  }
  if (n == 3) {
    // closest point on triangle (w0, w1, w2) to the origin, regions as in
    // closest_pt_tri with p = 0
    const V3 A = w[0], B = w[1], C = w[2];
    const V3 ab = B - A, ac = C - A, ap = mk(0, 0, 0) - A;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { lam[0] = 1; keep({0}); return A; }
    const V3 bp = mk(0, 0, 0) - B;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { lam[1] = 1; keep({1}); return B; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
      const double v = d1 / (d1 - d3);
      lam[0] = 1 - v; lam[1] = v; keep({0, 1});
      return A + v * ab;
    }
    const V3 cp = mk(0, 0, 0) - C;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { lam[2] = 1; keep({2}); return C; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
      const double v = d2 / (d2 - d6);
      lam[0] = 1 - v; lam[2] = v; keep({0, 2});
      return A + v * ac;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
      const double v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      lam[1] = 1 - v; lam[2] = v; keep({1, 2});
      return B + v * (C - B);
    }
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, ww = vc * den;
    lam[0] = 1 - v - ww; lam[1] = v; lam[2] = ww;
    return A + v * ab + ww * ac;
  }
  // tetrahedron: inside -> the origin; else the closest of the faces the
  // origin lies outside of
  const int F[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
  const int opp[4] = {3, 2, 1, 0};
  bool inside = true;
  double bestd = kInf;
  V3 best = mk(0, 0, 0);
  V3 bw[4], ba[4], bb[4];
  double bl[4];
  int bn = 0;
  for (int f = 0; f < 4; f++) {
    const V3 p0 = w[F[f][0]], p1 = w[F[f][1]], p2 = w[F[f][2]], q = w[opp[f]];
    const V3 nrm = cross(p1 - p0, p2 - p0);
    const double so = dot(mk(0, 0, 0) - p0, nrm), sq = dot(q - p0, nrm);
    // origin strictly on the fourth vertex's side of this face: not the
    // closest feature.  A flat tetrahedron (sq = 0) contains no interior
    // point, so every face is then examined.
    if (so * sq > 0) continue;
    inside = false;
    V3 tw[4] = {p0, p1, p2}, ta[4] = {a[F[f][0]], a[F[f][1]], a[F[f][2]]}, tb[4] = {b[F[f][0]], b[F[f][1]], b[F[f][2]]};
    double tl[4] = {0, 0, 0, 0};
    int tn = 3;
    const V3 c = simplex_closest(tw, ta, tb, tn, tl);
    const double dd = dot(c, c);
    if (dd < bestd) {
      bestd = dd;
      best = c;
      bn = tn;
      for (int i = 0; i < tn; i++) { bw[i] = tw[i]; ba[i] = ta[i]; bb[i] = tb[i]; bl[i] = tl[i]; }
    }
  }
  if (inside) {
    // barycentric weights of the origin
    const double vol = dot(w[1] - w[0], cross(w[2] - w[0], w[3] - w[0]));
    const V3 o = mk(0, 0, 0);
    lam[1] = dot(o - w[0], cross(w[2] - w[0], w[3] - w[0])) / vol;
    lam[2] = dot(w[1] - w[0], cross(o - w[0], w[3] - w[0])) / vol;
    lam[3] = dot(w[1] - w[0], cross(w[2] - w[0], o - w[0])) / vol;
    lam[0] = 1 - lam[1] - lam[2] - lam[3];
    return o;
  }
  n = bn;
  for (int i = 0; i < bn; i++) { w[i] = bw[i]; a[i] = ba[i]; b[i] = bb[i]; lam[i] = bl[i]; }
  return best;
}

// GJK distance between hulls A and B (0 when they overlap), with closest
// points (when separated).
static double hull_distance(const HullSet& A, const HullSet& B, V3* cpa, V3* cpb, int* iters) {
  V3 w[4], a[4], b[4];
  double lam[4] = {1, 0, 0, 0};
  int n = 1;
  a[0] = hull_support(A, mk(1, 0, 0), nullptr);
  b[0] = hull_support(B, mk(-1, 0, 0), nullptr);
  w[0] = a[0] - b[0];
  V3 v = w[0];
  double dist = std::sqrt(dot(v, v));
  int it = 0;
  for (; it < 200; it++) {
    const double vv = dot(v, v);
    if (vv <= 1e-30) { dist = 0; break; }
    const V3 na = hull_support(A, mk(-v.x, -v.y, -v.z), nullptr);
    const V3 nb = hull_support(B, v, nullptr);
    const V3 nw = na - nb;
    // converged: the new support point does not get closer than v by more
    // than the relative tolerance
    if (vv - dot(v, nw) <= 1e-14 * vv) break;
    bool dup = false;
    for (int i = 0; i < n; i++)
      if (w[i].x == nw.x && w[i].y == nw.y && w[i].z == nw.z) dup = true;
    if (dup) break;
    w[n] = nw; a[n] = na; b[n] = nb; lam[n] = 0; n++;
    v = simplex_closest(w, a, b, n, lam);
    if (n == 4) { dist = 0; break; }  // origin inside the tetrahedron
    const double nd = std::sqrt(dot(v, v));
    if (nd >= dist - 1e-15 * dist && it > 0) { dist = nd; break; }
    dist = nd;
  }
  if (iters) *iters = it;
  V3 pa = mk(0, 0, 0), pb = mk(0, 0, 0);
  for (int i = 0; i < n; i++) { pa = pa + lam[i] * a[i]; pb = pb + lam[i] * b[i]; }
  if (cpa) *cpa = pa;
  if (cpb) *cpb = pb;
  return dist;
}

static void world_aabb(const Mesh& m, const Xf& X, double* lo, double* hi) {
  double c[3], e[3];
  for (int k = 0; k < 3; k++) {
    c[k] = 0.5 * (m.lmin[k] + m.lmax[k]);
    e[k] = 0.5 * (m.lmax[k] - m.lmin[k]);
  }
  V3 wc = apply(X, mk(c[0], c[1], c[2]));
  for (int r = 0; r < 3; r++) {
    double ext = std::fabs(X.R[0 * 3 + r]) * e[0] + std::fabs(X.R[1 * 3 + r]) * e[1] +
                 std::fabs(X.R[2 * 3 + r]) * e[2];
    lo[r] = comp(wc, r) - ext;
    hi[r] = comp(wc, r) + ext;
  }
}

// ------------------------------------------------------------------ world

struct Robot {
  int n;
  std::vector<int> parents, jtype, geom;
  std::vector<Xf> T0;
  std::vector<V3> axes;
  std::vector<double> qmin, qmax;
  std::vector<uint8_t> selfcol;  // n*n, selfCollisions(i,j) != NULL
  std::vector<double> q;          // current config (used when it is "another robot")
};

struct World {
  std::vector<Mesh> geoms;
  std::vector<int> terrains;
  std::vector<std::pair<int, Xf>> objects;
  std::vector<Robot> robots;
  std::vector<uint8_t> mask;  // collisionEnabled, NumIDs^2
  int nids = 0;

  int NumIDs() const {
    int n = (int)terrains.size() + (int)objects.size() + (int)robots.size();
    for (auto& r : robots) n += r.n;
    return n;
  }
  int TerrainID(int i) const { return i; }
  int RigidObjectID(int i) const { return (int)terrains.size() + i; }
  int RobotID(int r) const {
    int id = (int)terrains.size() + (int)objects.size();
    for (int i = 0; i < r; i++) id += robots[i].n + 1;
    return id;
  }
  int RobotLinkID(int r, int l) const { return RobotID(r) + 1 + l; }
  bool en(int a, int b) const { return mask[(size_t)a * nids + b] != 0; }
};

static void fk(const Robot& R, const double* q, std::vector<Xf>& out) {
  out.resize(R.n);
  for (int i = 0; i < R.n; i++) {
    Xf J = joint_xf(R.jtype[i], R.axes[i], q[i]);
    Xf L = compose(R.T0[i], J);
    out[i] = R.parents[i] < 0 ? L : compose(out[R.parents[i]], L);
  }
}

// collisionEnabled defaults: PlannerSettings.cpp:16-41
static void init_default(World& w) {
  int n = w.NumIDs();
  w.nids = n;
  w.mask.assign((size_t)n * n, 1);
  for (int i = 0; i < n; i++) w.mask[(size_t)i * n + i] = 0;
  for (size_t r = 0; r < w.robots.size(); r++) {
    int k = w.RobotID((int)r);
    w.mask[(size_t)k * n + k] = 1;
    int base = w.RobotLinkID((int)r, 0);
    const Robot& R = w.robots[r];
    for (int j = 0; j < R.n; j++) {
      w.mask[(size_t)(base + j) * n + k] = 0;
      w.mask[(size_t)k * n + base + j] = 0;
    }
    for (int j = 0; j < R.n; j++)
      for (int l = 0; l < R.n; l++) w.mask[(size_t)(base + j) * n + base + l] = R.selfcol[j * R.n + l];
    for (int j = 0; j < R.n; j++)
      if (R.parents[j] == -1)
        for (size_t t = 0; t < w.terrains.size(); t++) {
          w.mask[(size_t)(base + j) * n + w.TerrainID((int)t)] = 0;
          w.mask[(size_t)w.TerrainID((int)t) * n + base + j] = 0;
        }
  }
}

// A geometry placed in the world for one query.
struct Placed {
  int id;          // world ID
  const Mesh* m;
  Xf X;
  double lo[3], hi[3];
};

static bool aabb_hit(const Placed& a, const Placed& b) {
  for (int k = 0; k < 3; k++)
    if (a.lo[k] > b.hi[k] || a.hi[k] < b.lo[k]) return false;
  return true;
}

static void place(Placed& p, int id, const Mesh* m, const Xf& X) {
  p.id = id; p.m = m; p.X = X;
  world_aabb(*m, X, p.lo, p.hi);
}

// GetGeometries for the robot being checked (links with non-empty geometry).
static void robot_links(const World& w, int r, const std::vector<Xf>& T, std::vector<Placed>& out) {
  const Robot& R = w.robots[r];
  for (int j = 0; j < R.n; j++) {
    if (R.geom[j] < 0 || w.geoms[R.geom[j]].empty()) continue;
    Placed p;
    place(p, w.RobotLinkID(r, j), &w.geoms[R.geom[j]], T[j]);
    out.push_back(p);
  }
}

// GetGeometries for idothers = terrains, rigid objects, other robots.
static void env_items(const World& w, int r, std::vector<Placed>& out) {
  for (size_t t = 0; t < w.terrains.size(); t++) {
    const Mesh& m = w.geoms[w.terrains[t]];
    if (m.empty()) continue;
    Placed p;
    place(p, w.TerrainID((int)t), &m, identity());
    out.push_back(p);
  }
  for (size_t o = 0; o < w.objects.size(); o++) {
    const Mesh& m = w.geoms[w.objects[o].first];
    if (m.empty()) continue;
    Placed p;
    place(p, w.RigidObjectID((int)o), &m, w.objects[o].second);
    out.push_back(p);
  }
  for (size_t r2 = 0; r2 < w.robots.size(); r2++) {
    if ((int)r2 == r) continue;
    std::vector<Xf> T;
    fk(w.robots[r2], w.robots[r2].q.data(), T);
    robot_links(w, (int)r2, T, out);
  }
}

// Planted policy mistakes, for tests that show the first-pair known-answer
// cases detect them (0 = Klampt's policy, the default; see or_set_mutation).
enum { MUT_NONE = 0, MUT_STABLE_ERASE = 1, MUT_LINKS_FIRST = 2, MUT_NO_REJECT = 3, MUT_ENV_MAJOR = 4 };
static int g_mutation = MUT_NONE;

// Set-level quick reject with swap-erase (PlannerSettings.cpp:290-317).
static void quick_reject(std::vector<Placed>& keep, const std::vector<Placed>& against) {
  double lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
  for (auto& a : against)
    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], a.lo[k]); hi[k] = std::max(hi[k], a.hi[k]); }
  for (size_t i = 0; i < keep.size(); i++) {
    bool hit = true;
    for (int k = 0; k < 3; k++)
      if (keep[i].lo[k] > hi[k] || keep[i].hi[k] < lo[k]) hit = false;
    if (!hit) {
      if (g_mutation == MUT_STABLE_ERASE) {
        keep.erase(keep.begin() + (long)i);
      } else {
        keep[i] = keep.back();
        keep.pop_back();
      }
      i--;
    }
  }
}

// Both reductions of CheckCollision(world, ids1, ids2): set 2 (the
// environment) against the union of set 1 (the links) first, then set 1
// against the union of what is left of set 2 (PlannerSettings.cpp:288-317).
static void reduce_sets(std::vector<Placed>& L1, std::vector<Placed>& L2) {
  if (g_mutation == MUT_NO_REJECT) return;
  if (g_mutation == MUT_LINKS_FIRST) {
    quick_reject(L1, L2);
    quick_reject(L2, L1);
    return;
  }
  quick_reject(L2, L1);
  quick_reject(L1, L2);
}

struct CheckResult {
  bool jl_ok = true;
  int env1 = -1, env2 = -1, self1 = -1, self2 = -1;
};

// SingleRobotCSpace::IsFeasible restated (RobotCSpace.cpp:786-823); both the
// env and the self check are always evaluated so both flags are reported.
static CheckResult check_config(const World& w, int r, const double* q, bool brute) {
  CheckResult res;
  const Robot& R = w.robots[r];
  for (int k = 0; k < R.n; k++)
    if (q[k] < R.qmin[k] || q[k] > R.qmax[k]) { res.jl_ok = false; break; }
  std::vector<Xf> T;
  fk(R, q, T);
  std::vector<Placed> L1, L2;
  robot_links(w, r, T, L1);
  env_items(w, r, L2);
  std::vector<Placed> S = L1;  // self check uses the unreduced link list
  // env: CheckCollision(world, ids1, ids2, tol=0)
  reduce_sets(L1, L2);
  const bool env_major = g_mutation == MUT_ENV_MAJOR;
  const size_t no = env_major ? L2.size() : L1.size(), ni = env_major ? L1.size() : L2.size();
  for (size_t o = 0; o < no && res.env1 < 0; o++)
    for (size_t k = 0; k < ni; k++) {
      const size_t i = env_major ? k : o, j = env_major ? o : k;
      if (!(w.en(L1[i].id, L2[j].id) || w.en(L2[j].id, L1[i].id))) continue;
      if (!aabb_hit(L1[i], L2[j])) continue;
      if (mesh_collide(*L1[i].m, L1[i].X, *L2[j].m, L2[j].X, brute)) {
        res.env1 = L1[i].id; res.env2 = L2[j].id;
        break;
      }
    }
  // self: CheckCollision(world, ids, tol=0) with the (i,j)||(i,i) quirk
  for (size_t i = 0; i < S.size() && res.self1 < 0; i++)
    for (size_t j = i + 1; j < S.size(); j++) {
      if (!(w.en(S[i].id, S[j].id) || w.en(S[i].id, S[i].id))) continue;
      if (!aabb_hit(S[i], S[j])) continue;
      if (mesh_collide(*S[i].m, S[i].X, *S[j].m, S[j].X, brute)) {
        res.self1 = S[i].id; res.self2 = S[j].id;
        break;
      }
    }
  return res;
}

// Fixed pair list for distance queries: (link, env item) pairs enabled by the
// mask in link-major order, then enabled self pairs (i<j).
static void pair_list(const World& w, int r, std::vector<std::pair<int, int>>& pairs) {
  const Robot& R = w.robots[r];
  std::vector<Xf> T;
  fk(R, R.q.data(), T);
  std::vector<Placed> L1, L2;
  robot_links(w, r, T, L1);
  env_items(w, r, L2);
  for (auto& a : L1)
    for (auto& b : L2)
      if (w.en(a.id, b.id) || w.en(b.id, a.id)) pairs.push_back({a.id, b.id});
  for (size_t i = 0; i < L1.size(); i++)
    for (size_t j = i + 1; j < L1.size(); j++)
      if (w.en(L1[i].id, L1[j].id) || w.en(L1[i].id, L1[i].id)) pairs.push_back({L1[i].id, L1[j].id});
}

template <class F>
static void parallel_for(int N, int nthreads, F f) {
  if (nthreads <= 1 || N < 2) {
    for (int i = 0; i < N; i++) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([&]() {
      for (;;) {
        int i = next.fetch_add(16);
        if (i >= N) break;
        for (int k = i; k < std::min(N, i + 16); k++) f(k);
      }
    });
  for (auto& t : th) t.join();
}

// Level-order bisection over the interior sample indices 1..n-1 of [0,n].
static void bisection_order(int n, std::vector<int>& order) {
  order.clear();
  if (n < 2) return;
  std::vector<std::pair<int, int>> q;
  q.push_back({0, n});
  for (size_t h = 0; h < q.size(); h++) {
    int lo = q[h].first, hi = q[h].second;
    if (hi - lo < 2) continue;
    int mid = lo + (hi - lo) / 2;
    order.push_back(mid);
    q.push_back({lo, mid});
    q.push_back({mid, hi});
  }
}

static int segment_samples(const double* a, const double* b, int n, double eps, int norm_inf) {
  double acc = 0;
  for (int k = 0; k < n; k++) {
    double d = a[k] - b[k];
    if (norm_inf) acc = std::max(acc, std::fabs(d));
    else acc += d * d;
  }
  double d = norm_inf ? acc : std::sqrt(acc);
  return (int)std::ceil(d / eps);
}

// ------------------------------------------------------------------ SDF

struct Grid {
  const double* v;  // C-order (i,j,k)
  int dims[3];
  double bmin[3], h[3];
};

// Cell-centred trilinear value and gradient.  Points are clamped to the box of
// cell centres (constant extension outside; the gradient along a clamped axis
// is 0).  geometry.h:322-347 fixes the cell-centre convention; the clamping is
// this library's definition.
static double sdf_eval(const Grid& g, V3 p, V3* grad) {
  double x[3] = {p.x, p.y, p.z};
  int i0[3];
  double f[3], gs[3];
  for (int a = 0; a < 3; a++) {
    double u = (x[a] - g.bmin[a]) / g.h[a] - 0.5;
    gs[a] = 1.0 / g.h[a];
    double um = (double)(g.dims[a] - 1);
    if (u <= 0) { u = 0; gs[a] = 0; }
    else if (u >= um) { u = um; gs[a] = 0; }
    int i = (int)std::floor(u);
    if (i > g.dims[a] - 2) i = g.dims[a] - 2;
    if (i < 0) i = 0;
    i0[a] = i;
    f[a] = g.dims[a] > 1 ? u - i : 0.0;
    if (g.dims[a] == 1) gs[a] = 0;
  }
  auto at = [&](int di, int dj, int dk) {
    int ii = std::min(i0[0] + di, g.dims[0] - 1);
    int jj = std::min(i0[1] + dj, g.dims[1] - 1);
    int kk = std::min(i0[2] + dk, g.dims[2] - 1);
    return g.v[((size_t)ii * g.dims[1] + jj) * g.dims[2] + kk];
  };
  double c000 = at(0, 0, 0), c001 = at(0, 0, 1), c010 = at(0, 1, 0), c011 = at(0, 1, 1);
  double c100 = at(1, 0, 0), c101 = at(1, 0, 1), c110 = at(1, 1, 0), c111 = at(1, 1, 1);
  double fx = f[0], fy = f[1], fz = f[2];
  double c00 = c000 * (1 - fz) + c001 * fz, c01 = c010 * (1 - fz) + c011 * fz;
  double c10 = c100 * (1 - fz) + c101 * fz, c11 = c110 * (1 - fz) + c111 * fz;
  double c0 = c00 * (1 - fy) + c01 * fy, c1 = c10 * (1 - fy) + c11 * fy;
  double val = c0 * (1 - fx) + c1 * fx;
  if (grad) {
    double dx = c1 - c0;
    double dy = (c01 - c00) * (1 - fx) + (c11 - c10) * fx;
    double dz0 = (c001 - c000) * (1 - fy) + (c011 - c010) * fy;
    double dz1 = (c101 - c100) * (1 - fy) + (c111 - c110) * fy;
    double dz = dz0 * (1 - fx) + dz1 * fx;
    *grad = mk(dx * gs[0], dy * gs[1], dz * gs[2]);
  }
  return val;
}

// ------------------------------------------------------------------ mesh -> SDF
//
// TriangleMesh -> ImplicitSurface (Geometry3D.convert("ImplicitSurface", h),
// Python/klampt/src/geometry.h:962-965).  KrisLibrary's fast-marching
// conversion is absent from the reference tree; this library defines the grid
// that method approximates, exactly:
//   value(i,j,k) = +-(Euclidean distance from the cell centre
//                  c = bmin + ((i,j,k) + 0.5) h to the closest triangle),
//   negative where the winding number of the mesh around c is nonzero.
// The winding number is counted along the +z line through c: a triangle whose
// xy projection contains (c.x, c.y) and which meets the line below c.z adds
// -1 when its normal points to +z and +1 when it points to -z.  Points on a
// projected edge u->v of the counter-clockwise-ordered triangle count when u
// precedes v in the order (y descending, then x ascending), which evaluates
// every line at the same infinitesimally shifted position, so an edge or a
// vertex shared by several triangles is counted once.  The edge functions are
// exact negations of each other for the two triangles of a shared edge.

static inline double edge_fn(double ux, double uy, double vx, double vy, double qx, double qy) {
  return (ux - qx) * (vy - qy) - (uy - qy) * (vx - qx);
}

static inline bool edge_inc(double w, double ux, double uy, double vx, double vy) {
  return w > 0 || (w == 0 && (uy > vy || (uy == vy && ux < vx)));
}

// Crossing of the +z line through (qx, qy) with triangle t: returns the
// winding increment (0: no crossing) and the crossing height in *zt.
static int z_crossing(const V3* t, double qx, double qy, double* zt) {
  V3 a = t[0], b = t[1], c = t[2];
  const double s = edge_fn(a.x, a.y, b.x, b.y, c.x, c.y);
  if (s == 0) return 0;
  const int w = s > 0 ? -1 : 1;
  if (s < 0) std::swap(b, c);
  const double wab = edge_fn(a.x, a.y, b.x, b.y, qx, qy);
  const double wbc = edge_fn(b.x, b.y, c.x, c.y, qx, qy);
  const double wca = edge_fn(c.x, c.y, a.x, a.y, qx, qy);
  if (!edge_inc(wab, a.x, a.y, b.x, b.y) || !edge_inc(wbc, b.x, b.y, c.x, c.y) ||
      !edge_inc(wca, c.x, c.y, a.x, a.y))
    return 0;
  const double den = (wab + wbc) + wca;
  if (den == 0) return 0;
  *zt = ((wbc * a.z + wca * b.z) + wab * c.z) / den;
  return w;
}

// First k with z_k = bmin_z + (k + 0.5) h > zt, in [0, nz].
static int first_above(double zt, double bz, double h, int nz) {
  double f = (zt - bz) / h - 0.5;
  if (!(f > -2.0)) f = -2.0;
  if (f > nz + 1.0) f = nz + 1.0;
  int k = (int)std::floor(f) + 1;
  if (k < 0) k = 0;
  if (k > nz) k = nz;
  while (k > 0 && bz + ((double)(k - 1) + 0.5) * h > zt) k--;
  while (k < nz && !(bz + ((double)k + 0.5) * h > zt)) k++;
  return k;
}

static double aabb_dist2(const Node& n, V3 p) {
  double d2 = 0;
  for (int k = 0; k < 3; k++) {
    double v = comp(p, k), e = 0;
    if (v < n.lo[k]) e = n.lo[k] - v;
    else if (v > n.hi[k]) e = v - n.hi[k];
    d2 += e * e;
  }
  return d2;
}

static double point_mesh_dist2(const Mesh& m, V3 p, bool brute) {
  double best = kInf;
  auto tri = [&](int t) {
    V3 q = closest_pt_tri(p, m.V[m.T[t][0]], m.V[m.T[t][1]], m.V[m.T[t][2]]);
    V3 d = p - q;
    best = std::min(best, dot(d, d));
  };
  if (brute || m.nodes.empty()) {
    for (size_t t = 0; t < m.T.size(); t++) tri((int)t);
    return best;
  }
  int stk[128], sp = 0;
  stk[sp++] = 0;
  while (sp) {
    const Node& n = m.nodes[stk[--sp]];
    if (aabb_dist2(n, p) >= best) continue;
    if (n.tri >= 0) { tri(n.tri); continue; }
    const double dl = aabb_dist2(m.nodes[n.left], p), dr = aabb_dist2(m.nodes[n.right], p);
    if (dl < dr) { stk[sp++] = n.right; stk[sp++] = n.left; }
    else { stk[sp++] = n.left; stk[sp++] = n.right; }
  }
  return best;
}

static void line_crossings(const Mesh& m, double qx, double qy, bool brute, std::vector<std::pair<double, int>>& out) {
  out.clear();
  auto tri = [&](int t) {
    V3 tv[3] = {m.V[m.T[t][0]], m.V[m.T[t][1]], m.V[m.T[t][2]]};
    double zt;
    int w = z_crossing(tv, qx, qy, &zt);
    if (w) out.push_back({zt, w});
  };
  if (brute || m.nodes.empty()) {
    for (size_t t = 0; t < m.T.size(); t++) tri((int)t);
    return;
  }
  int stk[128], sp = 0;
  stk[sp++] = 0;
  while (sp) {
    const Node& n = m.nodes[stk[--sp]];
    if (qx < n.lo[0] || qx > n.hi[0] || qy < n.lo[1] || qy > n.hi[1]) continue;
    if (n.tri >= 0) { tri(n.tri); continue; }
    stk[sp++] = n.left;
    stk[sp++] = n.right;
  }
}

}  // namespace orc

using namespace orc;

// ================================================================== C API

extern "C" {

void* or_world_new() { return new World(); }

// Test hook: plant one of the policy mistakes above in every later call
// (1 stable erase, 2 links reduced first, 3 no quick reject, 4 environment
// loop outermost); 0 restores Klampt's policy.  Not thread-safe with calls
// in flight.
void or_set_mutation(int m) { g_mutation = m; }
void or_world_free(void* w) { delete (World*)w; }

int or_add_mesh(void* wp, const double* V, int nv, const int* T, int nt) {
  World& w = *(World*)wp;
  Mesh m;
  m.V.resize(nv);
  for (int i = 0; i < nv; i++) m.V[i] = mk(V[3 * i], V[3 * i + 1], V[3 * i + 2]);
  m.T.resize(nt);
  for (int i = 0; i < nt; i++) m.T[i] = {T[3 * i], T[3 * i + 1], T[3 * i + 2]};
  build(m);
  w.geoms.push_back(std::move(m));
  return (int)w.geoms.size() - 1;
}

int or_add_terrain(void* wp, int g) {
  World& w = *(World*)wp;
  w.terrains.push_back(g);
  return (int)w.terrains.size() - 1;
}

int or_add_object(void* wp, int g, const double* xf12) {
  World& w = *(World*)wp;
  w.objects.push_back({g, from12(xf12)});
  return (int)w.objects.size() - 1;
}

int or_add_robot(void* wp, int n, const int* parents, const double* T0, const double* axes,
                 const int* jtype, const double* qmin, const double* qmax, const int* link_geom) {
  World& w = *(World*)wp;
  Robot R;
  R.n = n;
  R.parents.assign(parents, parents + n);
  R.jtype.assign(jtype, jtype + n);
  R.geom.assign(link_geom, link_geom + n);
  R.qmin.assign(qmin, qmin + n);
  R.qmax.assign(qmax, qmax + n);
  R.q.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    R.T0.push_back(from12(T0 + 12 * i));
    R.axes.push_back(mk(axes[3 * i], axes[3 * i + 1], axes[3 * i + 2]));
  }
  // InitAllSelfCollisions: all pairs except link vs parent (upper triangular)
  R.selfcol.assign((size_t)n * n, 0);
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (R.parents[j] != i && R.parents[i] != j) R.selfcol[i * n + j] = 1;
  w.robots.push_back(std::move(R));
  return (int)w.robots.size() - 1;
}

void or_set_self_collision(void* wp, int r, int i, int j, int enabled) {
  Robot& R = ((World*)wp)->robots[r];
  R.selfcol[i * R.n + j] = enabled ? 1 : 0;
}

void or_init_default(void* wp) { init_default(*(World*)wp); }

int or_num_ids(void* wp) { return ((World*)wp)->NumIDs(); }

void or_set_collision_enabled(void* wp, int a, int b, int en) {
  World& w = *(World*)wp;
  w.mask[(size_t)a * w.nids + b] = en ? 1 : 0;
}

int or_get_collision_enabled(void* wp, int a, int b) { return ((World*)wp)->en(a, b) ? 1 : 0; }

void or_set_robot_config(void* wp, int r, const double* q) {
  Robot& R = ((World*)wp)->robots[r];
  R.q.assign(q, q + R.n);
}

void or_fk(void* wp, int r, const double* q, double* out) {
  std::vector<Xf> T;
  fk(((World*)wp)->robots[r], q, T);
  for (size_t i = 0; i < T.size(); i++) {
    std::memcpy(out + 12 * i, T[i].R, 9 * sizeof(double));
    std::memcpy(out + 12 * i + 9, T[i].t, 3 * sizeof(double));
  }
}

// Batch feasibility.  first_pair (nullable) receives the Klampt-order first
// colliding (id1,id2): env first, then self; (-1,-1) if none.
void or_feasible_batch(void* wp, int r, const double* Q, int N, uint8_t* feas, uint8_t* env,
                       uint8_t* self, int32_t* first_pair, int brute, int nthreads) {
  World& w = *(World*)wp;
  int n = w.robots[r].n;
  parallel_for(N, nthreads, [&](int c) {
    CheckResult res = check_config(w, r, Q + (size_t)c * n, brute != 0);
    bool e = res.env1 >= 0, s = res.self1 >= 0;
    if (feas) feas[c] = (res.jl_ok && !e && !s) ? 1 : 0;
    if (env) env[c] = e ? 1 : 0;
    if (self) self[c] = s ? 1 : 0;
    if (first_pair) {
      first_pair[2 * c] = e ? res.env1 : res.self1;
      first_pair[2 * c + 1] = e ? res.env2 : res.self2;
    }
  });
}

// SingleRobotCSpace::IsFeasible with its early returns (RobotCSpace.cpp:786-823):
// joint limits, then env (first hit), then self (first hit).
static bool is_feasible(const World& w, int r, const double* q) {
  const Robot& R = w.robots[r];
  for (int k = 0; k < R.n; k++)
    if (q[k] < R.qmin[k] || q[k] > R.qmax[k]) return false;
  std::vector<Xf> T;
  fk(R, q, T);
  std::vector<Placed> L1, L2;
  robot_links(w, r, T, L1);
  env_items(w, r, L2);
  std::vector<Placed> S = L1;
  reduce_sets(L1, L2);
  for (size_t i = 0; i < L1.size(); i++)
    for (size_t j = 0; j < L2.size(); j++) {
      if (!(w.en(L1[i].id, L2[j].id) || w.en(L2[j].id, L1[i].id))) continue;
      if (!aabb_hit(L1[i], L2[j])) continue;
      if (mesh_collide(*L1[i].m, L1[i].X, *L2[j].m, L2[j].X, false)) return false;
    }
  for (size_t i = 0; i < S.size(); i++)
    for (size_t j = i + 1; j < S.size(); j++) {
      if (!(w.en(S[i].id, S[j].id) || w.en(S[i].id, S[i].id))) continue;
      if (!aabb_hit(S[i], S[j])) continue;
      if (mesh_collide(*S[i].m, S[i].X, *S[j].m, S[j].X, false)) return false;
    }
  return true;
}

void or_isfeasible_batch(void* wp, int r, const double* Q, int N, uint8_t* feas, int nthreads) {
  World& w = *(World*)wp;
  int n = w.robots[r].n;
  parallel_for(N, nthreads, [&](int c) { feas[c] = is_feasible(w, r, Q + (size_t)c * n) ? 1 : 0; });
}

// SingleRobotCSpace::IsFeasible with its early returns, traversing the OBB
// trees given by or_mesh_set_obb, with its work counted per configuration:
// out[c*5 + 0..4] = environment box tests, environment-node visits,
// environment triangle tests, self box tests, self triangle tests; feas[c]
// the answer.  Returns -1 (nothing computed) if a mesh of the check lacks
// an OBB tree.
static bool is_feasible_counted(const World& w, int r, const double* q, int64_t* out) {
  const Robot& R = w.robots[r];
  for (int k = 0; k < 5; k++) out[k] = 0;
  for (int k = 0; k < R.n; k++)
    if (q[k] < R.qmin[k] || q[k] > R.qmax[k]) return false;
  std::vector<Xf> T;
  fk(R, q, T);
  std::vector<Placed> L1, L2;
  robot_links(w, r, T, L1);
  env_items(w, r, L2);
  std::vector<Placed> S = L1;
  reduce_sets(L1, L2);
  ObbCount ce, cs;
  bool hit = false;
  for (size_t i = 0; i < L1.size() && !hit; i++)
    for (size_t j = 0; j < L2.size(); j++) {
      if (!(w.en(L1[i].id, L2[j].id) || w.en(L2[j].id, L1[i].id))) continue;
      if (!aabb_hit(L1[i], L2[j])) continue;
      if (obb_collide_count(*L1[i].m, L1[i].X, *L2[j].m, L2[j].X, true, ce)) {
        hit = true;
        break;
      }
    }
  for (size_t i = 0; i < S.size() && !hit; i++)
    for (size_t j = i + 1; j < S.size(); j++) {
      if (!(w.en(S[i].id, S[j].id) || w.en(S[i].id, S[i].id))) continue;
      if (!aabb_hit(S[i], S[j])) continue;
      if (obb_collide_count(*S[i].m, S[i].X, *S[j].m, S[j].X, false, cs)) {
        hit = true;
        break;
      }
    }
  out[0] = ce.box;
  out[1] = ce.env_nodes;
  out[2] = ce.tri;
  out[3] = cs.box;
  out[4] = cs.tri;
  return !hit;
}

// Distance between the hull of points A (na x 3, placed by XA[0] and, when
// nxa == 2, also by XA[1]: the hull of both copies) and the hull of points B
// (placed by XB); closest points cpa, cpb (world).  Returns the distance.
double or_hull_distance(const double* A, int na, const double* XA, int nxa, const double* B, int nb,
                        const double* XB, double* cpa, double* cpb, int* iters) {
  HullSet ha{A, na, {from12(XA), from12(nxa == 2 ? XA + 12 : XA)}, nxa == 2 ? 2 : 1};
  HullSet hb{B, nb, {from12(XB), from12(XB)}, 1};
  V3 pa, pb;
  const double d = hull_distance(ha, hb, &pa, &pb, iters);
  if (cpa) { cpa[0] = pa.x; cpa[1] = pa.y; cpa[2] = pa.z; }
  if (cpb) { cpb[0] = pb.x; cpb[1] = pb.y; cpb[2] = pb.z; }
  return d;
}

// Support point of the hull of points A placed by X (12) in direction d.
void or_hull_support(const double* A, int na, const double* X, const double* d, double* out) {
  HullSet h{A, na, {from12(X), from12(X)}, 1};
  const V3 p = hull_support(h, mk(d[0], d[1], d[2]), nullptr);
  out[0] = p.x; out[1] = p.y; out[2] = p.z;
}

int or_mesh_set_obb(void* wp, int g, const float* nodes, int nn) {
  World& w = *(World*)wp;
  Mesh& m = w.geoms[g];
  m.obb.resize(nn);
  for (int i = 0; i < nn; i++) {
    const float* p = nodes + 16 * (size_t)i;
    ObbN& o = m.obb[i];
    for (int k = 0; k < 3; k++) {
      o.c[k] = p[k];
      o.e[k] = p[3 + k];
    }
    for (int k = 0; k < 9; k++) o.R[k] = p[6 + k];
    std::memcpy(&o.child, p + 15, sizeof(int));
    if (o.child >= nn || (o.child < 0 && ~o.child >= (int)m.T.size())) return -1;
  }
  return 0;
}

int or_feasible_counted(void* wp, int r, const double* Q, int N, uint8_t* feas, int64_t* out, int nthreads) {
  World& w = *(World*)wp;
  const Robot& R = w.robots[r];
  for (int j = 0; j < R.n; j++)
    if (R.geom[j] >= 0 && !w.geoms[R.geom[j]].empty() && w.geoms[R.geom[j]].obb.empty()) return -1;
  for (int t : w.terrains)
    if (!w.geoms[t].empty() && w.geoms[t].obb.empty()) return -1;
  for (auto& o : w.objects)
    if (!w.geoms[o.first].empty() && w.geoms[o.first].obb.empty()) return -1;
  const int n = R.n;
  parallel_for(N, nthreads, [&](int c) {
    feas[c] = is_feasible_counted(w, r, Q + (size_t)c * n, out + 5 * (size_t)c) ? 1 : 0;
  });
  return 0;
}

int or_pair_list(void* wp, int r, int32_t* out, int maxp) {
  std::vector<std::pair<int, int>> P;
  pair_list(*(World*)wp, r, P);
  for (int i = 0; i < (int)P.size() && i < maxp; i++) { out[2 * i] = P[i].first; out[2 * i + 1] = P[i].second; }
  return (int)P.size();
}

// Per-pair distances for the fixed pair list (0 when touching).  With
// signed != 0 the contact-band separation is returned instead.
void or_distance_batch(void* wp, int r, const double* Q, int N, double* out, int brute, int signed_sep,
                       int nthreads) {
  World& w = *(World*)wp;
  std::vector<std::pair<int, int>> P;
  pair_list(w, r, P);
  int n = w.robots[r].n, np = (int)P.size();
  parallel_for(N, nthreads, [&](int c) {
    std::vector<Xf> T;
    fk(w.robots[r], Q + (size_t)c * n, T);
    std::vector<Placed> L1, L2;
    robot_links(w, r, T, L1);
    env_items(w, r, L2);
    auto find = [&](int id) -> const Placed* {
      for (auto& p : L1) if (p.id == id) return &p;
      for (auto& p : L2) if (p.id == id) return &p;
      return nullptr;
    };
    for (int k = 0; k < np; k++) {
      const Placed* a = find(P[k].first);
      const Placed* b = find(P[k].second);
      out[(size_t)c * np + k] = signed_sep ? mesh_separation(*a->m, a->X, *b->m, b->X)
                                           : mesh_distance(*a->m, a->X, *b->m, b->X, brute != 0);
    }
  });
}

// Edge checks.  Sample set: interior indices i=1..n-1 of n=ceil(d/eps), u=i/n,
// visited in level-order bisection; q(u) = a*(1-u) + b*u.  first_u receives u
// of the first infeasible sample in that order, or -1.
void or_edge_batch(void* wp, int r, const double* A, const double* B, int nseg, double eps, int norm_inf,
                   double* first_u, int32_t* nsamples, int32_t* nchecked, int nthreads) {
  World& w = *(World*)wp;
  int n = w.robots[r].n;
  parallel_for(nseg, nthreads, [&](int s) {
    const double* a = A + (size_t)s * n;
    const double* b = B + (size_t)s * n;
    int ns = segment_samples(a, b, n, eps, norm_inf);
    std::vector<int> order;
    bisection_order(ns, order);
    std::vector<double> q(n);
    double fu = -1;
    int checked = 0;
    for (int i : order) {
      double u = (double)i / (double)ns;
      for (int k = 0; k < n; k++) { q[k] = a[k] * (1.0 - u); q[k] += b[k] * u; }
      checked++;
      CheckResult res = check_config(w, r, q.data(), false);
      if (!res.jl_ok || res.env1 >= 0 || res.self1 >= 0) { fu = u; break; }
    }
    first_u[s] = fu;
    if (nsamples) nsamples[s] = ns;
    if (nchecked) nchecked[s] = checked;
  });
}

// Signed distance grid of a triangle mesh (definition above the helpers
// edge_fn .. line_crossings).  out: dims[0] x dims[1] x dims[2], C order;
// winding (nullable) receives the winding numbers.
void or_mesh_to_sdf(const double* V, int nv, const int* T, int nt, const double* bmin, double h, const int* dims,
                    double* out, int32_t* winding, int brute, int nthreads) {
  Mesh m;
  m.V.resize(nv);
  for (int i = 0; i < nv; i++) m.V[i] = mk(V[3 * i], V[3 * i + 1], V[3 * i + 2]);
  m.T.resize(nt);
  for (int i = 0; i < nt; i++) m.T[i] = {T[3 * i], T[3 * i + 1], T[3 * i + 2]};
  build(m);
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  parallel_for(nx * ny, nthreads, [&](int row) {
    const int i = row / ny, j = row % ny;
    const double x = bmin[0] + ((double)i + 0.5) * h, y = bmin[1] + ((double)j + 0.5) * h;
    std::vector<std::pair<double, int>> cr;
    line_crossings(m, x, y, brute != 0, cr);
    std::vector<int> diff(nz + 1, 0);
    for (auto& c : cr) diff[first_above(c.first, bmin[2], h, nz)] += c.second;
    int wind = 0;
    for (int k = 0; k < nz; k++) {
      wind += diff[k];
      const double z = bmin[2] + ((double)k + 0.5) * h;
      const double d = std::sqrt(point_mesh_dist2(m, mk(x, y, z), brute != 0));
      const size_t at = ((size_t)i * ny + j) * nz + k;
      out[at] = wind != 0 ? -d : d;
      if (winding) winding[at] = wind;
    }
  });
}

void or_bisection_order(int n, int32_t* out) {
  std::vector<int> o;
  bisection_order(n, o);
  for (size_t i = 0; i < o.size(); i++) out[i] = o[i];
}

// Single-pair queries on world-placed meshes.
int or_mesh_collide(void* wp, int ga, const double* xa, int gb, const double* xb, int brute) {
  World& w = *(World*)wp;
  return mesh_collide(w.geoms[ga], from12(xa), w.geoms[gb], from12(xb), brute != 0) ? 1 : 0;
}

double or_mesh_distance(void* wp, int ga, const double* xa, int gb, const double* xb, int brute, double* cp1,
                        double* cp2, int32_t* elems) {
  World& w = *(World*)wp;
  V3 p, q;
  int e1, e2;
  double d = mesh_distance(w.geoms[ga], from12(xa), w.geoms[gb], from12(xb), brute != 0, &p, &q, &e1, &e2);
  if (cp1) { cp1[0] = p.x; cp1[1] = p.y; cp1[2] = p.z; }
  if (cp2) { cp2[0] = q.x; cp2[1] = q.y; cp2[2] = q.z; }
  if (elems) { elems[0] = e1; elems[1] = e2; }
  return d;
}

double or_mesh_separation(void* wp, int ga, const double* xa, int gb, const double* xb) {
  World& w = *(World*)wp;
  return mesh_separation(w.geoms[ga], from12(xa), w.geoms[gb], from12(xb));
}

int or_tritri(const double* a, const double* b) {
  V3 t1[3], t2[3];
  for (int v = 0; v < 3; v++) {
    t1[v] = mk(a[3 * v], a[3 * v + 1], a[3 * v + 2]);
    t2[v] = mk(b[3 * v], b[3 * v + 1], b[3 * v + 2]);
  }
  return tri_tri(t1, t2) ? 1 : 0;
}

double or_tritri_dist(const double* a, const double* b) {
  V3 t1[3], t2[3];
  for (int v = 0; v < 3; v++) {
    t1[v] = mk(a[3 * v], a[3 * v + 1], a[3 * v + 2]);
    t2[v] = mk(b[3 * v], b[3 * v + 1], b[3 * v + 2]);
  }
  return tri_tri_dist(t1, t2, nullptr, nullptr);
}

double or_tritri_depth(const double* a, const double* b) {
  V3 t1[3], t2[3];
  for (int v = 0; v < 3; v++) {
    t1[v] = mk(a[3 * v], a[3 * v + 1], a[3 * v + 2]);
    t2[v] = mk(b[3 * v], b[3 * v + 1], b[3 * v + 2]);
  }
  return tri_tri_depth(t1, t2);
}

void or_world_aabb(void* wp, int g, const double* x, double* lo, double* hi) {
  world_aabb(((World*)wp)->geoms[g], from12(x), lo, hi);
}

// Cell-centred trilinear SDF value and gradient at world points.
void or_sdf_eval(const double* vals, const int* dims, const double* bmin, const double* bmax, const double* pts,
                 int npts, double* d, double* grad) {
  Grid g;
  g.v = vals;
  for (int a = 0; a < 3; a++) {
    g.dims[a] = dims[a];
    g.bmin[a] = bmin[a];
    g.h[a] = (bmax[a] - bmin[a]) / dims[a];
  }
  for (int i = 0; i < npts; i++) {
    V3 gr;
    d[i] = sdf_eval(g, mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), &gr);
    if (grad) { grad[3 * i] = gr.x; grad[3 * i + 1] = gr.y; grad[3 * i + 2] = gr.z; }
  }
}

// SDF hinge penalty per configuration, L(q) = sum_{links k, points p}
// max(0, dsafe - sdf(T_k(q) p))^2, and dL/dq by Klampt's distance-gradient
// rule: dL/dq = sum -2 max(0,dsafe-d) grad_d(x)^T J_pos(q, p)
// (geometry_ad.py:292 x kinematics_ad.py:56-59) with the explicit 3xn
// position Jacobian; or, with fd_h > 0, by central differences of L.
void or_sdf_penalty(void* wp, int r, const double* vals, const int* dims, const double* bmin, const double* bmax,
                    const double* Q, int N, const double* pts, const int32_t* pts_per_link, double dsafe,
                    double* loss, double* grad, double fd_h, int nthreads) {
  World& w = *(World*)wp;
  const Robot& R = w.robots[r];
  int n = R.n;
  Grid g;
  g.v = vals;
  for (int a = 0; a < 3; a++) {
    g.dims[a] = dims[a];
    g.bmin[a] = bmin[a];
    g.h[a] = (bmax[a] - bmin[a]) / dims[a];
  }
  std::vector<int> off(n + 1, 0);
  for (int k = 0; k < n; k++) off[k + 1] = off[k] + pts_per_link[k];
  auto eval_loss = [&](const double* q) {
    std::vector<Xf> T;
    fk(R, q, T);
    double L = 0;
    for (int k = 0; k < n; k++)
      for (int p = off[k]; p < off[k + 1]; p++) {
        V3 x = apply(T[k], mk(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]));
        double h = dsafe - sdf_eval(g, x, nullptr);
        if (h > 0) L += h * h;
      }
    return L;
  };
  parallel_for(N, nthreads, [&](int c) {
    const double* q = Q + (size_t)c * n;
    loss[c] = eval_loss(q);
    if (!grad) return;
    double* gq = grad + (size_t)c * n;
    if (fd_h > 0) {
      std::vector<double> qq(q, q + n);
      for (int j = 0; j < n; j++) {
        qq[j] = q[j] + fd_h;
        double lp = eval_loss(qq.data());
        qq[j] = q[j] - fd_h;
        double lm = eval_loss(qq.data());
        qq[j] = q[j];
        gq[j] = (lp - lm) / (2 * fd_h);
      }
      return;
    }
    std::vector<Xf> T;
    fk(R, q, T);
    for (int j = 0; j < n; j++) gq[j] = 0;
    for (int k = 0; k < n; k++)
      for (int p = off[k]; p < off[k + 1]; p++) {
        V3 x = apply(T[k], mk(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]));
        V3 gd;
        double h = dsafe - sdf_eval(g, x, &gd);
        if (h <= 0) continue;
        V3 dLdx = (-2.0 * h) * gd;
        // J_pos column j for every ancestor joint j of link k (including k)
        for (int j = k; j >= 0; j = R.parents[j]) {
          V3 wj = rot(T[j], R.axes[j]);
          V3 col = R.jtype[j] == 1 ? wj : cross(wj, x - mk(T[j].t[0], T[j].t[1], T[j].t[2]));
          gq[j] += dot(dLdx, col);
        }
      }
  });
}

}  // extern "C"
