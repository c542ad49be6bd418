// Triangle/triangle distance with closest points, usable from host and
// device code (FP32 or FP64).
//
// Distance = 0 when the closed triangles intersect (separating-axis test over
// the 17 candidate axes), else the minimum over the 9 edge/edge segment
// distances and the 6 vertex/triangle distances.  Segment closest points and
// the point/triangle closest point follow the Voronoi-region constructions of
// Ericson, Real-Time Collision Detection, 5.1.5 and 5.1.9.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace kgm {

template <typename R>
struct V {
  R x, y, z;
};
template <typename R>
__host__ __device__ inline V<R> mk(R x, R y, R z) {
  return V<R>{x, y, z};
}
template <typename R>
__host__ __device__ inline V<R> sub(V<R> a, V<R> b) {
  return mk<R>(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <typename R>
__host__ __device__ inline V<R> add(V<R> a, V<R> b) {
  return mk<R>(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <typename R>
__host__ __device__ inline V<R> mul(R s, V<R> a) {
  return mk<R>(s * a.x, s * a.y, s * a.z);
}
template <typename R>
__host__ __device__ inline R dot(V<R> a, V<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename R>
__host__ __device__ inline V<R> cross(V<R> a, V<R> b) {
  return mk<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename R>
__host__ __device__ inline R clamp01(R v) {
  return v < (R)0 ? (R)0 : (v > (R)1 ? (R)1 : v);
}

// Closest points of segments p0p1 and q0q1; returns the squared distance.
template <typename R>
__host__ __device__ inline R seg_seg(V<R> p0, V<R> p1, V<R> q0, V<R> q1, V<R>& cp, V<R>& cq) {
  const V<R> d1 = sub(p1, p0), d2 = sub(q1, q0), r = sub(p0, q0);
  const R a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  R s, t;
  if (a <= (R)0 && e <= (R)0) {
    s = t = (R)0;
  } else if (a <= (R)0) {
    s = (R)0;
    t = clamp01(f / e);
  } else {
    const R c = dot(d1, r);
    if (e <= (R)0) {
      t = (R)0;
      s = clamp01(-c / a);
    } else {
      const R b = dot(d1, d2), den = a * e - b * b;
      s = den != (R)0 ? clamp01((b * f - c * e) / den) : (R)0;
      t = (b * s + f) / e;
      if (t < (R)0) {
        t = (R)0;
        s = clamp01(-c / a);
      } else if (t > (R)1) {
        t = (R)1;
        s = clamp01((b - c) / a);
      }
    }
  }
  cp = add(p0, mul(s, d1));
  cq = add(q0, mul(t, d2));
  const V<R> d = sub(cp, cq);
  return dot(d, d);
}

// Closest point of triangle abc to p.
template <typename R>
__host__ __device__ inline V<R> closest_on_tri(V<R> p, V<R> a, V<R> b, V<R> c) {
  const V<R> ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const R d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= (R)0 && d2 <= (R)0) return a;
  const V<R> bp = sub(p, b);
  const R d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= (R)0 && d4 <= d3) return b;
  const R vc = d1 * d4 - d3 * d2;
  if (vc <= (R)0 && d1 >= (R)0 && d3 <= (R)0) return add(a, mul(d1 / (d1 - d3), ab));
  const V<R> cp = sub(p, c);
  const R d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= (R)0 && d5 <= d6) return c;
  const R vb = d5 * d2 - d1 * d6;
  if (vb <= (R)0 && d2 >= (R)0 && d6 <= (R)0) return add(a, mul(d2 / (d2 - d6), ac));
  const R va = d3 * d6 - d5 * d4;
  if (va <= (R)0 && (d4 - d3) >= (R)0 && (d5 - d6) >= (R)0)
    return add(b, mul((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(c, b)));
  const R den = (R)1 / (va + vb + vc);
  return add(a, add(mul(vb * den, ab), mul(vc * den, ac)));
}

// Separating-axis intersection test of closed triangles (no tolerance).
template <typename R>
__host__ __device__ inline bool tris_intersect(const V<R>* P, const V<R>* Q) {
  const V<R> eP[3] = {sub(P[1], P[0]), sub(P[2], P[1]), sub(P[0], P[2])};
  const V<R> eQ[3] = {sub(Q[1], Q[0]), sub(Q[2], Q[1]), sub(Q[0], Q[2])};
  const V<R> nP = cross(eP[0], eP[1]), nQ = cross(eQ[0], eQ[1]);
  auto separates = [&](V<R> a) {
    const R p0 = dot(a, P[0]), p1 = dot(a, P[1]), p2 = dot(a, P[2]);
    const R q0 = dot(a, Q[0]), q1 = dot(a, Q[1]), q2 = dot(a, Q[2]);
    const R pmin = fmin(p0, fmin(p1, p2)), pmax = fmax(p0, fmax(p1, p2));
    const R qmin = fmin(q0, fmin(q1, q2)), qmax = fmax(q0, fmax(q1, q2));
    return qmin > pmax || pmin > qmax;
  };
  if (separates(nP) || separates(nQ)) return false;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (separates(cross(eP[i], eQ[j]))) return false;
  for (int i = 0; i < 3; i++) {
    if (separates(cross(nP, eP[i]))) return false;
    if (separates(cross(nQ, eQ[i]))) return false;
  }
  return true;
}

// Distance between triangles P and Q with closest points cp (on P), cq (on Q).
// Coordinates are shifted to Q[0] internally for accuracy.
template <typename R>
__host__ __device__ inline R tri_dist(const V<R>* Pin, const V<R>* Qin, V<R>& cp, V<R>& cq) {
  const V<R> o = Qin[0];
  V<R> P[3], Q[3];
  for (int i = 0; i < 3; i++) {
    P[i] = sub(Pin[i], o);
    Q[i] = sub(Qin[i], o);
  }
  R best = (R)1e30;
  V<R> bp = P[0], bq = Q[0];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      V<R> a, b;
      const R d = seg_seg(P[i], P[(i + 1) % 3], Q[j], Q[(j + 1) % 3], a, b);
      if (d < best) {
        best = d;
        bp = a;
        bq = b;
      }
    }
  for (int i = 0; i < 3; i++) {
    V<R> c = closest_on_tri(P[i], Q[0], Q[1], Q[2]);
    V<R> d = sub(P[i], c);
    if (dot(d, d) < best) {
      best = dot(d, d);
      bp = P[i];
      bq = c;
    }
    c = closest_on_tri(Q[i], P[0], P[1], P[2]);
    d = sub(Q[i], c);
    if (dot(d, d) < best) {
      best = dot(d, d);
      bp = c;
      bq = Q[i];
    }
  }
  R dist = sqrt(best);
  if (tris_intersect(P, Q)) dist = (R)0;
  cp = add(bp, o);
  cq = add(bq, o);
  return dist;
}

}  // namespace kgm
