// HIP kernels for gfx950: forward kinematics, broad phase, BVH-vs-BVH
// collision traversal, and result assembly.
//
// Collision traversal (k_collide) runs one wavefront per workgroup over
// narrow-phase pair items (a link mesh against an environment mesh or another
// link), walking the two OBB trees: each step takes a node pair whose boxes
// overlap, tests both children of the larger-volume node against the other (15-axis
// separating-axis test), continues depth first with one overlapping child
// pair, pools the other, and queues leaf pairs for the 17-axis triangle test.
// A certain hit kills the remaining work of its configuration (any-hit), of
// its group (env / self), or of the items later in Klampt's loop order
// (ordered first-hit).
//
// Precision: the FP32 pass carries certified error bounds (Tol); a leaf test
// that cannot be decided within them marks its configuration for the FP64
// pass, which re-runs that configuration's items with FP64 transforms and
// FP64 triangles.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <cstdlib>
#include <mutex>
#include <type_traits>

#include "kg_internal.h"
#include "kg_edge.h"

#define KG_CPW_MAX 64
#define KG_NE 64  // configuration entries per wave (ring)
#define KG_COUNTER_STRIDE 32   // ints between the traversal's chunk counters (128 B)
#define KG_MIN_WAVES 64        // traversal waves kept for the smallest batches
#define KG_PREP_LDS_LINKS 8    // link boxes in LDS for robots of at most this many links
#define KG_PREP_LDS_BLOCK 128  // block size of that variant (48 KB of LDS)
#define KG_NONE64 0xffffffffffffffffull

namespace {

// ------------------------------------------------------------ FP64 transforms

struct X64 {
  double R[9];
  double t[3];
};

__device__ inline void x64_load(const double* p, X64& X) {
#pragma unroll
  for (int i = 0; i < 9; i++) X.R[i] = p[i];
#pragma unroll
  for (int i = 0; i < 3; i++) X.t[i] = p[9 + i];
}

__device__ inline void x64_store(double* p, const X64& X) {
#pragma unroll
  for (int i = 0; i < 9; i++) p[i] = X.R[i];
#pragma unroll
  for (int i = 0; i < 3; i++) p[9 + i] = X.t[i];
}

// Per-batch link frames are stored structure-of-arrays: component k (0..8 R
// column-major, 9..11 t) of frame row r at T[k * rows + r].  Single-plan
// batches use row l * N + c (consecutive configurations adjacent, so a
// wave's stores and loads are coalesced); heterogeneous batches use
// toff[c] + l.
__device__ inline size_t t_row(const int* toff, int N, int c, int l) {
  return toff ? (size_t)toff[c] + l : (size_t)l * N + c;
}
__device__ inline void t_load(const double* T, size_t rows, size_t row, X64& X) {
#pragma unroll
  for (int i = 0; i < 9; i++) X.R[i] = T[i * rows + row];
#pragma unroll
  for (int i = 0; i < 3; i++) X.t[i] = T[(9 + i) * rows + row];
}
__device__ inline void t_store(double* T, size_t rows, size_t row, const X64& X) {
#pragma unroll
  for (int i = 0; i < 9; i++) T[i * rows + row] = X.R[i];
#pragma unroll
  for (int i = 0; i < 3; i++) T[(9 + i) * rows + row] = X.t[i];
}

__device__ inline void x64_mul(const X64& A, const X64& B, X64& C) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int r = 0; r < 3; r++)
      C.R[c * 3 + r] = A.R[r] * B.R[c * 3] + A.R[3 + r] * B.R[c * 3 + 1] + A.R[6 + r] * B.R[c * 3 + 2];
#pragma unroll
  for (int r = 0; r < 3; r++) C.t[r] = A.R[r] * B.t[0] + A.R[3 + r] * B.t[1] + A.R[6 + r] * B.t[2] + A.t[r];
}

// Joint transform: Rodrigues rotation about w (revolute) or translation w*q.
__device__ inline void joint64(int type, const double* w, double q, X64& J) {
  if (type == 1) {
#pragma unroll
    for (int i = 0; i < 9; i++) J.R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    J.t[0] = w[0] * q; J.t[1] = w[1] * q; J.t[2] = w[2] * q;
    return;
  }
  double s, c;
  sincos(q, &s, &c);
  double v = 1.0 - c;
#pragma unroll
  for (int col = 0; col < 3; col++)
#pragma unroll
    for (int row = 0; row < 3; row++) J.R[col * 3 + row] = w[row] * w[col] * v;
  J.R[0] += c; J.R[4] += c; J.R[8] += c;
  J.R[1] += s * w[2]; J.R[2] -= s * w[1];
  J.R[3] -= s * w[2]; J.R[5] += s * w[0];
  J.R[6] += s * w[1]; J.R[7] -= s * w[0];
  J.t[0] = J.t[1] = J.t[2] = 0.0;
}

// FP64 world frame of link l from the joint values q[qb + k]: the joint
// frames T0_k J(q_k) composed from l up to its root (the broad phase composes
// the same product in the other association; the two differ by FP64
// rounding only).
__device__ inline void fk_link64(const PlanDev& PL, const float* Qf, const double* Qd, size_t qb, int l, X64& X) {
  auto local = [&](int k, X64& Lx) {
    const LinkDesc& L = PL.links[k];
    const double q = Qf ? (double)Qf[qb + k] : Qd[qb + k];
    X64 J, T0;
    joint64(L.jtype, L.axis, q, J);
    x64_load(L.T0, T0);
    x64_mul(T0, J, Lx);
  };
  local(l, X);
  for (int p = PL.links[l].parent; p >= 0; p = PL.links[p].parent) {
    X64 Lp, Y;
    local(p, Lp);
    x64_mul(Lp, X, Y);
    X = Y;
  }
}

// ------------------------------------------------------------ FK (FP32)

// One lane per configuration; the link table sits in LDS.
__global__ __launch_bounds__(256) void k_fk_f32(const LinkDesc* __restrict__ links, int n,
                                                 const float* __restrict__ Q, int N, float* __restrict__ Tout) {
  __shared__ float sT0[KG_MAX_LINKS][12];
  __shared__ float sAx[KG_MAX_LINKS][3];
  __shared__ int sPar[KG_MAX_LINKS], sTy[KG_MAX_LINKS];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    for (int k = 0; k < 12; k++) sT0[i][k] = (float)links[i].T0[k];
    for (int k = 0; k < 3; k++) sAx[i][k] = (float)links[i].axis[k];
    sPar[i] = links[i].parent;
    sTy[i] = links[i].jtype;
  }
  __syncthreads();
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float* out = Tout + (size_t)c * n * 12;
  for (int l = 0; l < n; l++) {
    float q = Q[(size_t)c * n + l];
    float J[12];
    const float* w = sAx[l];
    if (sTy[l] == 1) {
      for (int i = 0; i < 9; i++) J[i] = (i % 4 == 0) ? 1.f : 0.f;
      J[9] = w[0] * q; J[10] = w[1] * q; J[11] = w[2] * q;
    } else {
      float s, co;
      sincosf(q, &s, &co);
      float v = 1.f - co;
      for (int col = 0; col < 3; col++)
        for (int row = 0; row < 3; row++) J[col * 3 + row] = w[row] * w[col] * v;
      J[0] += co; J[4] += co; J[8] += co;
      J[1] += s * w[2]; J[2] -= s * w[1];
      J[3] -= s * w[2]; J[5] += s * w[0];
      J[6] += s * w[1]; J[7] -= s * w[0];
      J[9] = J[10] = J[11] = 0.f;
    }
    // L = T0 * J
    float L[12];
    const float* A = sT0[l];
    for (int col = 0; col < 3; col++)
      for (int r = 0; r < 3; r++)
        L[col * 3 + r] = A[r] * J[col * 3] + A[3 + r] * J[col * 3 + 1] + A[6 + r] * J[col * 3 + 2];
    for (int r = 0; r < 3; r++) L[9 + r] = A[r] * J[9] + A[3 + r] * J[10] + A[6 + r] * J[11] + A[9 + r];
    float* o = out + l * 12;
    int p = sPar[l];
    if (p < 0) {
      for (int i = 0; i < 12; i++) o[i] = L[i];
    } else {
      const float* P = out + p * 12;
      float Pv[12];
      for (int i = 0; i < 12; i++) Pv[i] = P[i];
      for (int col = 0; col < 3; col++)
        for (int r = 0; r < 3; r++)
          o[col * 3 + r] = Pv[r] * L[col * 3] + Pv[3 + r] * L[col * 3 + 1] + Pv[6 + r] * L[col * 3 + 2];
      for (int r = 0; r < 3; r++) o[9 + r] = Pv[r] * L[9] + Pv[3 + r] * L[10] + Pv[6 + r] * L[11] + Pv[9 + r];
    }
  }
}

// ------------------------------------------------------------ item records

// Error bounds of the FP32 narrow phase for an item whose transform has
// translation norm tnorm, between meshes of radii rA and rB.  A vertex p of
// mesh A (|p| <= rA) mapped by the FP32-rounded transform (R, t) lands within
// u(8 rA + 4|t|) of its exact image; a stored vertex of B (|v| <= rB) is
// within u rB of its FP64 value.  tau doubles the sum for safety margin; box
// tests multiply up to three rounded factors over coordinates bounded by
// |t| + rA + rB (including the inverted item transform).
__device__ inline void fp32_item_bounds(double tnorm, double rA, double rB, float& tau, float& pad) {
  const double u = 5.9604644775390625e-08;
  const double ev = u * (8.0 * rA + 4.0 * tnorm) + u * rB;
  tau = (float)(2.0 * ev);
  pad = (float)(2.0 * ev + 32.0 * u * (tnorm + rA + rB));
}

// Start record of an item: XA, XB are the FP64 world frames of meshes A and
// B (column-major R, then t); rel = XB^-1 XA rounded to FP32.
__device__ inline void write_item_rec(ItemRec* __restrict__ out, const double* XA, const double* XB, const GeomDesc& GA,
                                      const GeomDesc& GB) {
  float rel[12];
  const double d0 = XA[9] - XB[9], d1 = XA[10] - XB[10], d2 = XA[11] - XB[11];
#pragma unroll
  for (int col = 0; col < 3; col++)
#pragma unroll
    for (int r = 0; r < 3; r++)
      rel[col * 3 + r] = (float)(XB[r * 3] * XA[col * 3] + XB[r * 3 + 1] * XA[col * 3 + 1] + XB[r * 3 + 2] * XA[col * 3 + 2]);
  double tn = 0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const double v = XB[r * 3] * d0 + XB[r * 3 + 1] * d1 + XB[r * 3 + 2] * d2;
    rel[9 + r] = (float)v;
    tn += v * v;
  }
  float tau, pad;
  fp32_item_bounds(sqrt(tn), GA.rmax, GB.rmax, tau, pad);
  float4* q = reinterpret_cast<float4*>(out);
  q[0] = make_float4(rel[0], rel[1], rel[2], rel[3]);
  q[1] = make_float4(rel[4], rel[5], rel[6], rel[7]);
  q[2] = make_float4(rel[8], rel[9], rel[10], rel[11]);
  q[3] = make_float4(tau, pad, __int_as_float(GA.root), __int_as_float(GB.root));
}

// ------------------------------------------------------------ broad phase

__device__ inline bool box_hit(const double* alo, const double* ahi, const double* blo, const double* bhi) {
  return !(alo[0] > bhi[0] || ahi[0] < blo[0] || alo[1] > bhi[1] || ahi[1] < blo[1] || alo[2] > bhi[2] ||
           ahi[2] < blo[2]);
}

// Position of survivor j after Klampt's set-level quick reject, which erases
// a rejected entry by moving the current last entry into its place
// (PlannerSettings.cpp:290-317).  Survivors below m (the survivor count) keep
// their index; the k-th survivor counted from the end fills the k-th hole
// below m.  `rej(i)` says whether entry i was rejected.
template <class F>
__device__ int swap_erase_pos(int j, int E, int m, F rej) {
  if (j < m) return j;
  int k = 0;
  for (int i = j + 1; i < E; i++)
    if (!rej(i)) k++;
  int h = 0;
  for (int i = 0; i < m; i++)
    if (rej(i)) {
      if (h == k) return i;
      h++;
    }
  return -1;
}

// LDSN > 0: single-plan batches of at most LDSN links keep the link boxes in
// LDS (structure of arrays over the block's lanes) instead of HBM.
template <int MODE, int LDSN>
__device__ inline void prep_cfg(const PrepArgs& a, int c, double* sLB) {
  constexpr bool FILL = MODE != 0;
  constexpr int BS = LDSN > 0 ? KG_PREP_LDS_BLOCK : 256;
  const PlanDev& P = a.cfg_plan ? a.plans[a.cfg_plan[c]] : a.plan;
  const int n = P.nlinks;
  const size_t tb = a.toff ? (size_t)a.toff[c] : (size_t)c * n;
  const size_t qb = a.qoff ? (size_t)a.qoff[c] : (size_t)c * n;
  auto trow = [&](int l) { return t_row(a.toff, a.N, c, l); };
  double* LB = a.LB + tb * 6;
  // link box component r (0..2 lo, 3..5 hi) of link l
  auto lb = [&](int l, int r) -> double& {
    if (LDSN > 0) return sLB[(l * 6 + r) * BS + threadIdx.x];
    return LB[6 * l + r];
  };
  // box of link l against [blo, bhi] (closed intervals)
  auto box_hit_l = [&](int l, const double* blo, const double* bhi) {
    return !(lb(l, 0) > bhi[0] || lb(l, 3) < blo[0] || lb(l, 1) > bhi[1] || lb(l, 4) < blo[1] ||
             lb(l, 2) > bhi[2] || lb(l, 5) < blo[2]);
  };
  // union of the active links' boxes
  double ulo[3] = {1e300, 1e300, 1e300}, uhi[3] = {-1e300, -1e300, -1e300};
  uint64_t active = 0;
  if (MODE != 1) {
    bool jl = true;
    X64 Xprev;  // frame of link l-1, kept in registers for serial chains
    for (int l = 0; l < n; l++) {
      double q = a.Qf ? (double)a.Qf[qb + l] : a.Qd[qb + l];
      if (q < P.qmin[l] || q > P.qmax[l]) jl = false;
      const LinkDesc& L = P.links[l];
      X64 J, T0, Lx, X;
      joint64(L.jtype, L.axis, q, J);
      x64_load(L.T0, T0);
      x64_mul(T0, J, Lx);
      if (L.parent >= 0) {
        X64 Pp;
        if (L.parent == l - 1) Pp = Xprev;
        else t_load(a.T64, a.t_rows, trow(L.parent), Pp);
        x64_mul(Pp, Lx, X);
      } else {
        X = Lx;
      }
      Xprev = X;
      t_store(a.T64, a.t_rows, trow(l), X);
      if (L.geom >= 0) {
        const GeomDesc& g = a.pool.geoms[L.geom];
        active |= 1ull << l;
        for (int r = 0; r < 3; r++) {
          double wc = X.R[r] * g.lc[0] + X.R[3 + r] * g.lc[1] + X.R[6 + r] * g.lc[2] + X.t[r];
          double ext = fabs(X.R[r]) * g.le[0] + fabs(X.R[3 + r]) * g.le[1] + fabs(X.R[6 + r]) * g.le[2];
          lb(l, r) = wc - ext;
          lb(l, 3 + r) = wc + ext;
          ulo[r] = fmin(ulo[r], wc - ext);
          uhi[r] = fmax(uhi[r], wc + ext);
        }
      }
    }
    a.jl[c] = jl ? 1 : 0;
    if (a.mode == KG_MODE_ANY && !jl) {
      a.counts[c] = 0;
      return;
    }
  } else {
    if (a.mode == KG_MODE_ANY && !a.jl[c]) return;
  }
  const bool ordered = a.mode == KG_MODE_ORDERED;
  if (MODE == 1)
    for (int l = 0; l < n; l++)
      if (P.links[l].geom >= 0) {
        active |= 1ull << l;
        for (int r = 0; r < 3; r++) {
          ulo[r] = fmin(ulo[r], lb(l, r));
          uhi[r] = fmax(uhi[r], lb(l, 3 + r));
        }
      }
  // env entries surviving against the link union; union of survivors
  double elo[3] = {1e300, 1e300, 1e300}, ehi[3] = {-1e300, -1e300, -1e300};
  int m2 = 0;
  for (int e = 0; e < P.nenv; e++) {
    const EnvItem& E = P.env[e];
    if (box_hit(E.lo, E.hi, ulo, uhi)) {
      m2++;
      for (int r = 0; r < 3; r++) {
        elo[r] = fmin(elo[r], E.lo[r]);
        ehi[r] = fmax(ehi[r], E.hi[r]);
      }
    }
  }
  // links surviving against the env survivor union (positions in the active list)
  int nact = __popcll(active);
  uint64_t lrej = 0;  // bit = active-list position rejected
  {
    int pos = 0;
    for (int l = 0; l < n; l++)
      if (active >> l & 1) {
        if (!box_hit_l(l, elo, ehi)) lrej |= 1ull << pos;
        pos++;
      }
  }
  int m1 = nact - __popcll(lrej);
  int cnt = 0;
  int base = MODE == 1 ? a.offsets[c] : (MODE == 2 ? (a.ibase ? a.ibase[c] : c * a.cap) : 0);
  auto env_rej = [&](int i) { return !box_hit(P.env[i].lo, P.env[i].hi, ulo, uhi); };
  auto link_rej = [&](int i) { return (lrej >> i & 1) != 0; };
  // env pairs
  if (m2 > 0) {
    int pos = -1;
    for (int l = 0; l < n; l++) {
      if (!(active >> l & 1)) continue;
      pos++;
      if (lrej >> pos & 1) continue;
      int p1 = ordered ? swap_erase_pos(pos, nact, m1, link_rej) : 0;
      for (int e = 0; e < P.nenv; e++) {
        const EnvItem& E = P.env[e];
        if (!P.lemask[l * P.nenv + e]) continue;
        if (!box_hit(E.lo, E.hi, ulo, uhi)) continue;  // rejected env entry
        if (!box_hit_l(l, E.lo, E.hi)) continue;
        if (FILL) {
          int p2 = ordered ? swap_erase_pos(e, P.nenv, m2, env_rej) : 0;
          PairItem it;
          it.cfg = c;
          it.a = (int16_t)l;
          it.b = (int16_t)e;
          it.rank = ordered ? (p1 * 65536 + p2) : 0;
          a.items[base + cnt] = it;
          if (a.recs) {
            X64 XA;
            t_load(a.T64, a.t_rows, trow(l), XA);
            write_item_rec(a.recs + base + cnt, XA.R, E.X, a.pool.geoms[P.links[l].geom], a.pool.geoms[E.geom]);
          }
        }
        cnt++;
      }
    }
  }
  // self pairs
  for (int k = 0; k < P.nself; k++) {
    int i = P.selfpairs[2 * k], j = P.selfpairs[2 * k + 1];
    if (lb(i, 0) > lb(j, 3) || lb(i, 3) < lb(j, 0) || lb(i, 1) > lb(j, 4) || lb(i, 4) < lb(j, 1) ||
        lb(i, 2) > lb(j, 5) || lb(i, 5) < lb(j, 2))
      continue;
    if (FILL) {
      PairItem it;
      it.cfg = c;
      it.a = (int16_t)i;
      it.b = (int16_t)j;
      it.rank = (1 << KG_RANK_GROUP_BIT) | (ordered ? k : 0);
      a.items[base + cnt] = it;
      if (a.recs) {
        X64 XA, XB;
        t_load(a.T64, a.t_rows, trow(i), XA);
        t_load(a.T64, a.t_rows, trow(j), XB);
        write_item_rec(a.recs + base + cnt, XA.R, XB.R, a.pool.geoms[P.links[i].geom], a.pool.geoms[P.links[j].geom]);
      }
    }
    cnt++;
  }
  if (MODE != 1) a.counts[c] = cnt;
}

// One lane per configuration; with a device-side count the lanes loop over
// the configurations (the grid is then bounded by the launcher).
template <int MODE, int LDSN>
__global__ __launch_bounds__(256) void k_prep(PrepArgs a) {
  constexpr int BS = LDSN > 0 ? KG_PREP_LDS_BLOCK : 256;
  __shared__ double sLB[LDSN > 0 ? LDSN * 6 * BS : 1];
  const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (MODE != 1 && c0 < a.nreset) a.reset[c0] = 0;
  const int N = a.N_dev ? min(*a.N_dev, a.N) : a.N;
  for (int c = c0; c < N; c += gridDim.x * blockDim.x) prep_cfg<MODE, LDSN>(a, c, sLB);
}

// Broad phase with one lane per (configuration, link): groups of G lanes
// (G = 8, 16 or 32, at least the robot's link count) take one configuration
// each, of a single plan or of the configuration's own plan in a
// heterogeneous batch (fixed-capacity item layout).  Each lane computes its
// joint frame T0 J(q) in FP64; the group composes the world frames in LDS by
// pointer jumping over the parent links (ceil(log2 depth) rounds of all
// lanes: 3 for a 7-link chain, against 6 dependent products down the list);
// each lane then forms its link box, and the group emits the items: a link's
// environment pairs by its own lane, the self pairs round-robin over the
// group, each with its start record.  Item positions inside a configuration
// come from a counter in LDS; their order does not matter, since the ranks
// carry Klampt's loop order.  MULTI: per-configuration plans (the single
// plan is then read from the kernel arguments with scalar loads).
// one wavefront per block: a group never spans waves, so the block barriers
// only order the wave's own LDS traffic and no wave waits for another
#define KG_PREP_TB 64
template <int G, bool MULTI, bool LOOP>
__global__ __launch_bounds__(KG_PREP_TB) void k_prep_group(PrepArgs a) {
  constexpr int NG = KG_PREP_TB / G;  // configurations per block
  __shared__ double sF[NG][G][12];
  __shared__ double sB[NG][G][6];
  __shared__ int sAnc[NG][G];
  __shared__ int scnt[NG], sjl[NG];
  const int t = threadIdx.x, g = t / G, l = t % G;
  {
    const int r = blockIdx.x * blockDim.x + t;
    if (r < a.nreset) a.reset[r] = 0;
  }
  // LOOP (a device-side count): the blocks loop over the batch (block-uniform
  // trip count, so the barriers below are reached by the whole block); else
  // one group of configurations per block
  const int Nc = LOOP ? min(*a.N_dev, a.N) : a.N;
  const int cb0 = blockIdx.x * NG, step = LOOP ? (int)gridDim.x * NG : 1;
  for (int cb = cb0; LOOP ? cb < Nc : cb == cb0; cb += step) {
  const int c = cb + g;
  const bool valid = c < Nc;
  const PlanDev& P = MULTI ? a.plans[valid ? a.cfg_plan[c] : 0] : a.plan;
  const int n = valid ? P.nlinks : 0;
  const bool mine = l < n;
  if (l == 0) {
    scnt[g] = 0;
    sjl[g] = 1;
  }
  __syncthreads();
  // joint frames
  X64 X;
  int anc = -1;
  if (mine) {
    const size_t qb = MULTI ? (size_t)a.qoff[c] : (size_t)c * n;
    const double q = a.Qf ? (double)a.Qf[qb + l] : a.Qd[qb + l];
    if (q < P.qmin[l] || q > P.qmax[l]) sjl[g] = 0;
    const LinkDesc& L = P.links[l];
    X64 J, T0;
    joint64(L.jtype, L.axis, q, J);
    x64_load(L.T0, T0);
    x64_mul(T0, J, X);
    anc = L.parent;
    x64_store(sF[g][l], X);
    sAnc[g][l] = anc;
  }
  // world frames by pointer jumping: after each round X is the product of the
  // joint frames from the link `anc` (exclusive) down to this one, and `anc`
  // is twice as far up.  The round count, ceil(log2(depth)) over the batch's
  // plans, is a kernel argument, so the loop is uniform by construction.  A
  // data-dependent exit (a block-wide OR, or a wave vote) let the compiler,
  // which lowers the block barriers of a one-wave block to wave barriers,
  // split the loop into nested loops that lanes left at different rounds,
  // mixing frames of different rounds and voting over part of the wave.
  for (int round = 0; round < a.jump_rounds; round++) {
    X64 Y;
    int up = -1;
    __syncthreads();
    if (anc >= 0) {
      x64_load(sF[g][anc], Y);
      up = sAnc[g][anc];
    }
    __syncthreads();
    if (anc >= 0) {
      X64 Z;
      x64_mul(Y, X, Z);
      X = Z;
      anc = up;
      x64_store(sF[g][l], X);
      sAnc[g][l] = anc;
    }
  }
  if (mine) {
    if (a.T64) t_store(a.T64, a.t_rows, MULTI ? (size_t)a.toff[c] + l : (size_t)l * a.N + c, X);
    const LinkDesc& L = P.links[l];
    if (L.geom >= 0) {
      const GeomDesc& gd = a.pool.geoms[L.geom];
      for (int r = 0; r < 3; r++) {
        const double wc = X.R[r] * gd.lc[0] + X.R[3 + r] * gd.lc[1] + X.R[6 + r] * gd.lc[2] + X.t[r];
        const double ext = fabs(X.R[r]) * gd.le[0] + fabs(X.R[3 + r]) * gd.le[1] + fabs(X.R[6 + r]) * gd.le[2];
        sB[g][l][r] = wc - ext;
        sB[g][l][3 + r] = wc + ext;
      }
    }
  }
  // links with geometry, as a mask over the group's lanes
  const bool act_l = mine && P.links[l].geom >= 0;
  const uint64_t bal = __ballot(act_l);
  const int gshift = (t & 63) & ~(G - 1);
  const uint32_t gmask = G == 32 ? 0xffffffffu : ((1u << G) - 1);
  const uint32_t active = (uint32_t)(bal >> gshift) & gmask;
  // union of the active links' boxes: a min / max reduction over the group's
  // lanes (exact, as the per-link loop it replaces)
  double ulo[3], uhi[3];
  for (int r = 0; r < 3; r++) {
    ulo[r] = act_l ? sB[g][l][r] : 1e300;
    uhi[r] = act_l ? sB[g][l][3 + r] : -1e300;
  }
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1)
    for (int r = 0; r < 3; r++) {
      ulo[r] = fmin(ulo[r], __shfl_xor(ulo[r], off, G));
      uhi[r] = fmax(uhi[r], __shfl_xor(uhi[r], off, G));
    }
  __syncthreads();
  const bool jl = sjl[g] != 0;
  const bool go = valid && (a.mode != KG_MODE_ANY || jl);
  const bool ordered = a.mode == KG_MODE_ORDERED;
  if (go) {
    // environment entries surviving against the link union, their union
    double elo[3] = {1e300, 1e300, 1e300}, ehi[3] = {-1e300, -1e300, -1e300};
    int m2 = 0;
    for (int e = 0; e < P.nenv; e++) {
      const EnvItem& E = P.env[e];
      if (box_hit(E.lo, E.hi, ulo, uhi)) {
        m2++;
        for (int r = 0; r < 3; r++) {
          elo[r] = fmin(elo[r], E.lo[r]);
          ehi[r] = fmax(ehi[r], E.hi[r]);
        }
      }
    }
    // links surviving against the environment union (active-list positions):
    // each lane tests its own link, the group's verdicts are gathered by
    // ballot (every lane of the group is in this branch) and compacted
    const uint64_t rb = __ballot(act_l && !box_hit(sB[g][l], sB[g][l] + 3, elo, ehi));
    const uint32_t rej = (uint32_t)(rb >> gshift) & gmask;
    uint32_t lrej = 0;
    {
      int pos = 0;
      for (uint32_t m = active; m; m &= m - 1, pos++)
        if (rej >> (__ffs(m) - 1) & 1) lrej |= 1u << pos;
    }
    const int nact = __popc(active), m1 = nact - __popc(lrej);
    const int base = MULTI ? a.ibase[c] : c * a.cap;
    auto env_rej = [&](int i) { return !box_hit(P.env[i].lo, P.env[i].hi, ulo, uhi); };
    auto link_rej = [&](int i) { return (lrej >> i & 1) != 0; };
    // this link's environment pairs
    if (mine && m2 > 0 && (active >> l & 1)) {
      const int pos = __popc(active & ((1u << l) - 1));
      if (!(lrej >> pos & 1)) {
        const int p1 = ordered ? swap_erase_pos(pos, nact, m1, link_rej) : 0;
        const LinkDesc& L = P.links[l];
        for (int e = 0; e < P.nenv; e++) {
          const EnvItem& E = P.env[e];
          if (!P.lemask[l * P.nenv + e]) continue;
          if (!box_hit(E.lo, E.hi, ulo, uhi)) continue;
          if (!box_hit(sB[g][l], sB[g][l] + 3, E.lo, E.hi)) continue;
          const int p2 = ordered ? swap_erase_pos(e, P.nenv, m2, env_rej) : 0;
          const int idx = base + atomicAdd(&scnt[g], 1);
          PairItem it;
          it.cfg = c;
          it.a = (int16_t)l;
          it.b = (int16_t)e;
          it.rank = ordered ? (p1 * 65536 + p2) : 0;
          a.items[idx] = it;
          // the link frame from LDS (keeps it out of registers across the loop)
          if (a.recs) write_item_rec(a.recs + idx, sF[g][l], E.X, a.pool.geoms[L.geom], a.pool.geoms[E.geom]);
        }
      }
    }
    // self pairs, round-robin over the group
    for (int k = l; k < P.nself; k += G) {
      const int i = P.selfpairs[2 * k], j = P.selfpairs[2 * k + 1];
      if (!box_hit(sB[g][i], sB[g][i] + 3, sB[g][j], sB[g][j] + 3)) continue;
      const int idx = base + atomicAdd(&scnt[g], 1);
      PairItem it;
      it.cfg = c;
      it.a = (int16_t)i;
      it.b = (int16_t)j;
      it.rank = (1 << KG_RANK_GROUP_BIT) | (ordered ? k : 0);
      a.items[idx] = it;
      if (a.recs)
        write_item_rec(a.recs + idx, sF[g][i], sF[g][j], a.pool.geoms[P.links[i].geom],
                       a.pool.geoms[P.links[j].geom]);
    }
  }
  __syncthreads();
  if (valid && l == 0) {
    a.jl[c] = jl ? 1 : 0;
    a.counts[c] = go ? scnt[g] : 0;
  }
  }
}

// ------------------------------------------------------------ narrow phase

template <typename Real>
struct V3T {
  Real x, y, z;
};
template <typename Real>
__device__ inline V3T<Real> v3(Real x, Real y, Real z) {
  return V3T<Real>{x, y, z};
}
template <typename Real>
__device__ inline V3T<Real> vsub(V3T<Real> a, V3T<Real> b) {
  return v3<Real>(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <typename Real>
__device__ inline Real vdot(V3T<Real> a, V3T<Real> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename Real>
__device__ inline V3T<Real> vcross(V3T<Real> a, V3T<Real> b) {
  return v3<Real>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename Real>
__device__ inline Real vl1(V3T<Real> a) {
  return fabs(a.x) + fabs(a.y) + fabs(a.z);
}

enum { TT_SEP = 0, TT_HIT = 1, TT_UNSURE = 2 };

// One separating-axis candidate.  Returns true if the axis certainly
// separates; clears `sure` when the axis does not certainly overlap.
template <typename Real, bool EXACT>
__device__ inline bool sat_axis(V3T<Real> a, Real aerr, const V3T<Real>* P, const V3T<Real>* Q, Real tau, Real RD,
                                bool& sure) {
  Real p0 = vdot(a, P[0]), p1 = vdot(a, P[1]), p2 = vdot(a, P[2]);
  Real q0 = vdot(a, Q[0]), q1 = vdot(a, Q[1]), q2 = vdot(a, Q[2]);
  Real pmin = fmin(p0, fmin(p1, p2)), pmax = fmax(p0, fmax(p1, p2));
  Real qmin = fmin(q0, fmin(q1, q2)), qmax = fmax(q0, fmax(q1, q2));
  Real gap = fmax(qmin - pmax, pmin - qmax);
  if (EXACT) return gap > (Real)0;
  Real aa = vdot(a, a);
  if (gap > (Real)0 && gap * gap > tau * tau * aa) return true;
  Real al = sqrt(aa);
  if (!(-gap > tau * al + aerr * RD)) sure = false;
  return false;
}

// Closed triangle/triangle test by 17 separating axes (2 normals, 9 edge
// crosses, 6 in-plane edge normals), in coordinates shifted to q[0], spread
// over the 9 lanes sub = 0..8 of a lane group: every lane tests the edge
// cross eP(sub / 3) x eQ(sub % 3), then lanes 0..5 one in-plane edge normal
// each (nP x eP(sub), nQ x eQ(sub - 3)) and lanes 6, 7 the normals nP, nQ.
// Returns bit 0 when one of the lane's axes certainly separates and bit 1
// when one of them does not certainly overlap; the pair is separated if any
// lane of the group sets bit 0, else undecided (FP32) if any sets bit 1,
// else intersecting.  tau bounds the vertex position error (x2, both
// triangles) plus projection rounding; each axis' own rounding is bounded
// separately.  EXACT (FP64) sets bit 0 only.
template <typename Real, bool EXACT>
__device__ int tri_tri_split(const V3T<Real>* p, const V3T<Real>* q, Real tau, int sub) {
  V3T<Real> P[3], Q[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    P[i] = vsub(p[i], q[0]);
    Q[i] = vsub(q[i], q[0]);
  }
  const Real u = EXACT ? (Real)0 : (Real)5.9604645e-08;
  Real RD = 0;
  if (!EXACT) {
#pragma unroll
    for (int i = 0; i < 3; i++) RD = fmax(RD, fmax(vl1(P[i]), vl1(Q[i])));
    RD *= (Real)2;
    tau += (Real)8 * u * RD;
  }
  const V3T<Real> eP0 = vsub(P[1], P[0]), eP1 = vsub(P[2], P[1]), eP2 = vsub(P[0], P[2]);
  const V3T<Real> eQ0 = vsub(Q[1], Q[0]), eQ1 = vsub(Q[2], Q[1]), eQ2 = vsub(Q[0], Q[2]);
  const Real lP0 = EXACT ? 0 : vl1(eP0), lP1 = EXACT ? 0 : vl1(eP1), lP2 = EXACT ? 0 : vl1(eP2);
  const Real lQ0 = EXACT ? 0 : vl1(eQ0), lQ1 = EXACT ? 0 : vl1(eQ1), lQ2 = EXACT ? 0 : vl1(eQ2);
  auto pick = [](int k, V3T<Real> a0, V3T<Real> a1, V3T<Real> a2) { return k == 0 ? a0 : (k == 1 ? a1 : a2); };
  auto pickr = [](int k, Real a0, Real a1, Real a2) { return k == 0 ? a0 : (k == 1 ? a1 : a2); };
  bool sure = true;
  int out = 0;
  {
    const int i = sub / 3, j = sub - 3 * (sub / 3);
    const V3T<Real> a = vcross(pick(i, eP0, eP1, eP2), pick(j, eQ0, eQ1, eQ2));
    if (sat_axis<Real, EXACT>(a, (Real)8 * u * pickr(i, lP0, lP1, lP2) * pickr(j, lQ0, lQ1, lQ2), P, Q, tau, RD, sure))
      out = 1;
  }
  if (sub < 8) {
    const V3T<Real> nP = vcross(eP0, eP1), nQ = vcross(eQ0, eQ1);
    const bool onP = sub < 3 || sub == 6;
    const int k = sub < 3 ? sub : sub - 3;
    V3T<Real> a;
    Real err;
    if (sub < 6) {
      a = onP ? vcross(nP, pick(k, eP0, eP1, eP2)) : vcross(nQ, pick(k, eQ0, eQ1, eQ2));
      err = onP ? (Real)16 * u * lP0 * lP1 * pickr(k, lP0, lP1, lP2) : (Real)16 * u * lQ0 * lQ1 * pickr(k, lQ0, lQ1, lQ2);
    } else {
      a = onP ? nP : nQ;
      err = onP ? (Real)8 * u * lP0 * lP1 : (Real)8 * u * lQ0 * lQ1;
    }
    if (sat_axis<Real, EXACT>(a, err, P, Q, tau, RD, sure)) out = 1;
  }
  return out | (sure ? 0 : 2);
}

template <typename Real>
struct NodeT {
  Real c[3], e[3], R[9];
  int child;
};

__device__ inline void load_node(const ObbNode* __restrict__ nodes, int i, float* c, float* e, float* R, int& child) {
  const float4* p = reinterpret_cast<const float4*>(nodes + i);
  float4 a = p[0], b = p[1], d = p[2], f = p[3];
  c[0] = a.x; c[1] = a.y; c[2] = a.z; e[0] = a.w;
  e[1] = b.x; e[2] = b.y; R[0] = b.z; R[1] = b.w;
  R[2] = d.x; R[3] = d.y; R[4] = d.z; R[5] = d.w;
  R[6] = f.x; R[7] = f.y; R[8] = f.z;
  child = __float_as_int(f.w);
}

struct NodeF {
  float c[3], e[3], R[9];
  int ch;
};

__device__ inline NodeF load_nodef(const ObbNode* __restrict__ nodes, int i) {
  NodeF n;
  load_node(nodes, i, n.c, n.e, n.R, n.ch);
  return n;
}

__device__ inline NodeF sel(bool p, const NodeF& a, const NodeF& b) {
  NodeF r;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    r.c[i] = p ? a.c[i] : b.c[i];
    r.e[i] = p ? a.e[i] : b.e[i];
  }
#pragma unroll
  for (int i = 0; i < 9; i++) r.R[i] = p ? a.R[i] : b.R[i];
  r.ch = p ? a.ch : b.ch;
  return r;
}

// Frame of a task's two box tests.  S is the node whose children are tested
// (the split node), O the other node; rel' maps S's mesh frame to O's.
// With M = R_O^T R' and t0 = R_O^T (t' - c_O), a child box s of S has
// C = M R_s (its axes in O's axes) and T = M c_s + t0 (its centre), so the
// shared part is computed once per task instead of once per test.
template <typename Real>
struct TaskFrame {
  Real M[9];   // M[i*3+m] = O_i . (R' e_m)
  Real t0[3];
  Real eO[3];
  Real sumO;
};

template <typename Real>
__device__ inline void task_frame(const NodeF& O, const Real* rel, TaskFrame<Real>& F) {
  Real RO[9];
#pragma unroll
  for (int i = 0; i < 9; i++) RO[i] = (Real)O.R[i];
  Real d[3];
#pragma unroll
  for (int r = 0; r < 3; r++) d[r] = rel[9 + r] - (Real)O.c[r];
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int m = 0; m < 3; m++)
      F.M[i * 3 + m] = RO[i * 3] * rel[m * 3] + RO[i * 3 + 1] * rel[m * 3 + 1] + RO[i * 3 + 2] * rel[m * 3 + 2];
    F.t0[i] = RO[i * 3] * d[0] + RO[i * 3 + 1] * d[1] + RO[i * 3 + 2] * d[2];
    F.eO[i] = (Real)O.e[i];
  }
  F.sumO = F.eO[0] + F.eO[1] + F.eO[2];
}

// 15-axis separating-axis test of child box S against O in the task frame;
// every right-hand side padded by pad_abs + pad_rel * (sum of extents).
// fmargin: the smallest slack over the six face axes (how deeply the boxes
// overlap), used to order the descent.
template <typename Real>
__device__ inline bool frame_overlap(const TaskFrame<Real>& F, const NodeF& S, Real pad_abs, Real pad_rel,
                                     Real& fmargin) {
  Real C[3][3], Qa[3][3], T[3], eA[3];
#pragma unroll
  for (int j = 0; j < 3; j++) eA[j] = (Real)S.e[j];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    T[i] = F.M[i * 3] * (Real)S.c[0] + F.M[i * 3 + 1] * (Real)S.c[1] + F.M[i * 3 + 2] * (Real)S.c[2] + F.t0[i];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C[i][j] = F.M[i * 3] * (Real)S.R[j * 3] + F.M[i * 3 + 1] * (Real)S.R[j * 3 + 1] + F.M[i * 3 + 2] * (Real)S.R[j * 3 + 2];
      Qa[i][j] = fabs(C[i][j]);
    }
  }
  const Real* eB = F.eO;
  const Real pad = pad_abs + pad_rel * (eA[0] + eA[1] + eA[2] + F.sumO);
  bool sep = false;
  Real fm = (Real)3.0e38;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const Real rhs = eB[i] + eA[0] * Qa[i][0] + eA[1] * Qa[i][1] + eA[2] * Qa[i][2] + pad;
    sep |= fabs(T[i]) > rhs;
    fm = fmin(fm, rhs - fabs(T[i]));
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const Real t = T[0] * C[0][j] + T[1] * C[1][j] + T[2] * C[2][j];
    const Real rhs = eB[0] * Qa[0][j] + eB[1] * Qa[1][j] + eB[2] * Qa[2][j] + eA[j] + pad;
    sep |= fabs(t) > rhs;
    fm = fmin(fm, rhs - fabs(t));
  }
  fmargin = fm;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      const Real t = T[i2] * C[i1][j] - T[i1] * C[i2][j];
      const Real ra = eB[i1] * Qa[i2][j] + eB[i2] * Qa[i1][j];
      const Real rb = eA[j1] * Qa[i][j2] + eA[j2] * Qa[i][j1];
      sep |= fabs(t) > ra + rb + pad;
    }
  }
  return !sep;
}

template <typename Real>
__device__ inline void load_tri(const PoolDev& pool, int t, V3T<Real>* v);

template <>
__device__ inline void load_tri<float>(const PoolDev& pool, int t, V3T<float>* v) {
  float4 a = pool.tv0[t], b = pool.tv1[t], c = pool.tv2[t];
  v[0] = v3<float>(a.x, a.y, a.z);
  v[1] = v3<float>(b.x, b.y, b.z);
  v[2] = v3<float>(c.x, c.y, c.z);
}

template <>
__device__ inline void load_tri<double>(const PoolDev& pool, int t, V3T<double>* v) {
  const double* p = pool.tv64 + 9 * (size_t)t;
#pragma unroll
  for (int k = 0; k < 3; k++) v[k] = v3<double>(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
}

template <typename Real>
__device__ inline V3T<Real> xf_apply(const Real* X, V3T<Real> p) {
  return v3<Real>(X[0] * p.x + X[3] * p.y + X[6] * p.z + X[9], X[1] * p.x + X[4] * p.y + X[7] * p.z + X[10],
                  X[2] * p.x + X[5] * p.y + X[8] * p.z + X[11]);
}

__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }

// Position of the j-th (0-based) set bit of m; requires popcount(m) > j.
__device__ inline int nth_set_bit(uint64_t m, int j) {
  int pos = 0, c = __popc((uint32_t)m);
  if (j >= c) { j -= c; m >>= 32; pos += 32; }
  c = __popc((uint32_t)m & 0xffffu);
  if (j >= c) { j -= c; m >>= 16; pos += 16; }
  c = __popc((uint32_t)m & 0xffu);
  if (j >= c) { j -= c; m >>= 8; pos += 8; }
  c = __popc((uint32_t)m & 0xfu);
  if (j >= c) { j -= c; m >>= 4; pos += 4; }
  c = __popc((uint32_t)m & 0x3u);
  if (j >= c) { j -= c; m >>= 2; pos += 2; }
  if (j >= (int)(m & 1)) pos += 1;
  return pos;
}
__device__ inline int lane_prefix(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Rank/kill helpers.  rank word: group in bit 30, order below.
__device__ inline bool item_dead(int mode, const unsigned long long* best, int g, int r) {
  if (mode == KG_MODE_ANY) return best[0] != KG_NONE64 || best[1] != KG_NONE64;
  unsigned long long b = best[g];
  if (b == KG_NONE64) return false;
  if (mode == KG_MODE_GROUPS) return true;
  return (int)(b >> 32) < r;
}

// Per-slot error bound for the FP32 narrow phase.  A vertex p of mesh A
// (|p| <= rA) mapped by the FP32-rounded rel (R,t) to frame B lands within
// u*(8 rA + 4|t|) of its exact image; a stored vertex of B (|v| <= rB) is
// within u*rB of its FP64 value.  tau doubles the sum for safety margin.
template <typename Real, bool EXACT>
__device__ inline void slot_bounds(const TravArgs& a, double tnorm, double rA, double rB, float& tau, float& pad) {
  if (EXACT) {
    tau = 0.f;
    pad = a.tol.pad_bv;
    return;
  }
  fp32_item_bounds(tnorm, rA, rB, tau, pad);
}

// Result of configuration c from its kill state: best hit key and lowest
// uncertain rank per group (env, self).  A configuration whose answer an
// uncertain triangle test could still change goes to the re-check list
// unless this is the exact pass.
template <bool FPAIR>
__device__ inline void finalize_cfg(const FinArgs& a, int c, unsigned long long b0, unsigned long long b1, int u0,
                                    int u1) {
  const bool jl = a.jl[c] != 0;
  const bool h0 = b0 != KG_NONE64, h1 = b1 != KG_NONE64;
  bool sure;
  if (a.mode == KG_MODE_ANY) {
    sure = !jl || h0 || h1 || (u0 == KG_RANK_NONE && u1 == KG_RANK_NONE);
  } else if (a.mode == KG_MODE_GROUPS) {
    sure = (h0 || u0 == KG_RANK_NONE) && (h1 || u1 == KG_RANK_NONE);
  } else {
    const int r0 = h0 ? (int)(b0 >> 32) : KG_RANK_NONE;
    const int r1 = h1 ? (int)(b1 >> 32) : KG_RANK_NONE;
    sure = (u0 > r0 || u0 == KG_RANK_NONE) && (u1 > r1 || u1 == KG_RANK_NONE);
  }
  if ((!sure || a.force_recheck) && !a.final_pass) {
    const int slot = atomicAdd(a.redo_count, 1);
    a.redo_list[slot] = c;
    return;
  }
  // Exact pass: an answer can still be open only when the traversal dropped
  // work (spill region full).  What depends on it is reported on the
  // colliding side -- infeasible, the group's flag set, first pair (-2,-2)
  // -- and the configuration is counted, so no dropped work can turn into a
  // "feasible".
  const bool open0 = !h0 && u0 != KG_RANK_NONE, open1 = !h1 && u1 != KG_RANK_NONE;
  if (!sure) atomicAdd(a.uncertain, 1);
  if (a.feasible) a.feasible[c] = (jl && !h0 && !h1 && !open0 && !open1) ? 1 : 0;
  if (a.env_hit) a.env_hit[c] = (h0 || open0) ? 1 : 0;
  if (a.self_hit) a.self_hit[c] = (h1 || open1) ? 1 : 0;
  if (FPAIR && a.first_pair) {
    int id1 = -1, id2 = -1;
    if (!sure) {
      // the environment pair is known when a hit ranks before every open
      // environment item; the self pair only when the environment is free
      const int r0 = h0 ? (int)(b0 >> 32) : KG_RANK_NONE;
      const int r1 = h1 ? (int)(b1 >> 32) : KG_RANK_NONE;
      const bool env_known = u0 == KG_RANK_NONE || (h0 && u0 > r0);
      const bool self_known = u1 == KG_RANK_NONE || (h1 && u1 > r1);
      if (!(env_known && (h0 || self_known))) id1 = id2 = -2;
    }
    if (id1 != -2 && (h0 || h1)) {
      const PairItem it = a.items[(uint32_t)((h0 ? b0 : b1) & 0xffffffffu)];
      const PlanDev& PL = a.cfg_plan ? a.plans[a.cfg_plan[c]] : a.plan;
      id1 = PL.links[it.a].wid;
      id2 = h0 ? PL.env[it.b].wid : PL.links[it.b].wid;
    }
    a.first_pair[2 * c] = id1;
    a.first_pair[2 * c + 1] = id2;
  }
}

// Frame of pair item P: rel maps mesh A's frame to mesh B's (FP64 product,
// rounded to Real), with its error bounds and the two BVH roots.
template <typename Real, bool EXACT>
__device__ inline void item_frame(const TravArgs& a, const PairItem& P, Real* rel, float& tau, float& pad, int& rootA,
                                  int& rootB) {
  const PlanDev& PL = a.cfg_plan ? a.plans[a.cfg_plan[P.cfg]] : a.plan;
  const LinkDesc& LA = PL.links[P.a];
  const int g = (P.rank >> KG_RANK_GROUP_BIT) & 1;
  // both frames are copied into registers: a pointer that may select a
  // private array would place that array in scratch memory
  // the exact pass forms the frames from q (the broad phase may not have
  // stored them), the FP32 pass reads them from T64
  X64 FA, FB;
  const size_t qb = a.qoff ? (size_t)a.qoff[P.cfg] : (size_t)P.cfg * PL.nlinks;
  if (EXACT)
    fk_link64(PL, a.Qf, a.Qd, qb, P.a, FA);
  else
    t_load(a.T64, a.t_rows, t_row(a.toff, a.t_n, P.cfg, P.a), FA);
  const double* XA = FA.R;
  const double* XB = FB.R;
  int geomB;
  if (g == 0) {
    const double* XE = PL.env[P.b].X;
#pragma unroll
    for (int k = 0; k < 12; k++) FB.R[k] = XE[k];
    geomB = PL.env[P.b].geom;
  } else {
    if (EXACT)
      fk_link64(PL, a.Qf, a.Qd, qb, P.b, FB);
    else
      t_load(a.T64, a.t_rows, t_row(a.toff, a.t_n, P.cfg, P.b), FB);
    geomB = PL.links[P.b].geom;
  }
  const double d0 = XA[9] - XB[9], d1 = XA[10] - XB[10], d2 = XA[11] - XB[11];
#pragma unroll
  for (int col = 0; col < 3; col++)
#pragma unroll
    for (int r = 0; r < 3; r++)
      rel[col * 3 + r] = (Real)(XB[r * 3] * XA[col * 3] + XB[r * 3 + 1] * XA[col * 3 + 1] + XB[r * 3 + 2] * XA[col * 3 + 2]);
  double tn = 0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const double v = XB[r * 3] * d0 + XB[r * 3 + 1] * d1 + XB[r * 3 + 2] * d2;
    rel[9 + r] = (Real)v;
    tn += v * v;
  }
  const GeomDesc& GA = a.pool.geoms[LA.geom];
  const GeomDesc& GB = a.pool.geoms[geomB];
  rootA = GA.root;
  rootB = GB.root;
  slot_bounds<Real, EXACT>(a, sqrt(tn), GA.rmax, GB.rmax, tau, pad);
}

// The exact pass's verdict on one leaf pair (triangle ta of mesh A, tb of
// mesh B) of item P, for lane `sub` of its 9-lane group: the item frame
// formed in FP64 from q and the FP64 triangle test, as k_collide<double,
// true> computes them for the pair.  Bit 0: separated.
__device__ inline int exact_leaf(const TravArgs& a, const PairItem& P, const PoolDev& pool, int ta, int tb, int sub) {
  double rel[12];
  float tau, pad;
  int rootA, rootB;
  item_frame<double, true>(a, P, rel, tau, pad, rootA, rootB);
  V3T<double> tp[3], tq[3];
  load_tri<double>(pool, ta, tp);
  load_tri<double>(pool, tb, tq);
#pragma unroll
  for (int v = 0; v < 3; v++) tp[v] = xf_apply<double>(rel, tp[v]);
  return tri_tri_split<double, true>(tp, tq, 0.0, sub);
}

// Task word: slot(6) | side(1) | single(1) | X(28) | Y(28).
//   single = 1: test the pair (X as A-node, Y as B-node).
//   single = 0: X is the first of two sibling children; test (X+k, Y) when
//   side = 0 (A split) or (Y, X+k) when side = 1 (B split), k = 0,1.
// Leaf word: slot(6) | A-triangle(29) | B-triangle(29).
__device__ inline uint64_t mk_task(int slot, int side, int single, int x, int y) {
  return ((uint64_t)slot << 58) | ((uint64_t)side << 57) | ((uint64_t)single << 56) |
         ((uint64_t)(uint32_t)x << 28) | (uint64_t)(uint32_t)y;
}
__device__ inline uint64_t mk_leaf(int slot, int ta, int tb) {
  return ((uint64_t)slot << 58) | ((uint64_t)(uint32_t)ta << 29) | (uint64_t)(uint32_t)tb;
}

#define KG_SCAP 1024  // LDS pool entries of k_distance
#define KG_LQ 256
#define KG_NOTASK 0xffffffffffffffffull

// One wavefront per workgroup.  Each lane follows one depth-first path (its
// continuation, kept in registers); when a task yields two overlapping child
// pairs, the lane continues with one and pushes the other onto the wave's
// shared LIFO pool in LDS.  Lanes whose path ended pop the pool, and only when
// the pool is empty do they start new pair items, so every item is explored
// depth first while heavy items spread over idle lanes.  Leaf pairs are
// queued and tested wave-coherently.
//
// Configurations stream through a ring of KG_NE entries in LDS holding their
// kill state.  A wave fetches configurations from the global counter in
// chunks that shrink as the work runs out (at most cpw, about the remaining
// count over the grid size), starts their items as lanes and slots free up,
// and writes a configuration's result as soon as its last item finishes, so
// a heavy item never holds the wave's other lanes idle at a chunk boundary.
// FPAIR: the first-colliding-pair output is requested (ordered mode).
// REC: item start records exist (else each item's frame is formed here).
template <typename Real, bool EXACT, int SCAP, bool FPAIR, bool REC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SCAP == 256 && REC && !FPAIR ? 4 : 1))) void k_collide(
    TravArgs a, const ObbNode* __restrict__ nodes, const float4* __restrict__ tv0, const float4* __restrict__ tv1,
    const float4* __restrict__ tv2, const double* __restrict__ tv64, const PairItem* __restrict__ items,
    const ItemRec* __restrict__ recs) {
  // the pool, items and records are read-only here: separate restrict
  // arguments let the compiler treat their loads as unclobbered
  PoolDev rpool = a.pool;
  rpool.tv0 = tv0;
  rpool.tv1 = tv1;
  rpool.tv2 = tv2;
  rpool.tv64 = tv64;
  // the small-pool variant also halves the leaf queue and the ring so that
  // four waves fit per SIMD
  constexpr int LQ = SCAP >= 512 ? KG_LQ : KG_LQ / 2;
  constexpr int NE = SCAP >= 512 ? KG_NE : KG_NE / 2;
  __shared__ uint64_t stk[SCAP];
  __shared__ uint64_t lq[LQ];
  __shared__ Real s_rel[12][64];
  __shared__ float s_tau[64], s_pad[64];
  __shared__ int s_cfgk[64], s_rank[64], s_item[64], s_pend[64], s_unc[64];
  // configuration entries: kill state, open-item count (+1 while items of
  // the entry remain to be started), configuration index (-1 = free)
  __shared__ unsigned long long s_best[NE][2];
  __shared__ int s_uncmin[NE][2];
  __shared__ int e_cfg[NE], e_open[NE], e_beg[NE], e_end[NE];
  __shared__ int s_chunk;

  const int lane = threadIdx.x;
  const int mode = a.mode;
  const Real pad_rel = (Real)a.tol.pad_rel;
  uint64_t* gsp_base = a.spill + (size_t)blockIdx.x * a.spill_cap;
  const int ncfg = a.ncfg_dev ? *a.ncfg_dev : a.ncfg;
  if (EXACT && a.fin.xpairs) {
    // Undecided triangle pairs recorded by the FP32 pass, 7 per wave pass in
    // 9-lane groups: an intersecting one makes its configuration collide
    // (the FP32 pass finished it as if the pair were separated)
    const int np = min(*a.fin.xpair_count, a.fin.xpair_cap);
    const int li = lane / 9, sub = lane - 9 * li;
    for (int b = (int)blockIdx.x * 7; b < np; b += (int)gridDim.x * 7) {
      const int k = b + li;
      const bool act = li < 7 && k < np;
      int bits = 1, c = 0, g = 0;
      if (act) {
        const int4 x = a.fin.xpairs[k];
        const PairItem P = items[x.y];
        c = x.x;
        g = (P.rank >> KG_RANK_GROUP_BIT) & 1;
        // a configuration already decided by a certain hit (recorded before
        // the hit was found) cannot change: no test
        const uint8_t* flag = a.mode == KG_MODE_ANY ? a.fin.feasible : (g == 0 ? a.fin.env_hit : a.fin.self_hit);
        const bool moot = flag && flag[c] == (a.mode == KG_MODE_ANY ? 0 : 1);
        if (!moot) bits = exact_leaf(a, P, rpool, x.z, x.w, sub);
      }
      const uint64_t sepm = ballot64((bits & 1) != 0);
      if (act && sub == 0 && !(sepm & (0x1ffull << (9 * li)))) {
        if (a.fin.feasible) a.fin.feasible[c] = 0;
        if (g == 0 && a.fin.env_hit) a.fin.env_hit[c] = 1;
        if (g == 1 && a.fin.self_hit) a.fin.self_hit[c] = 1;
      }
    }
  }
  // nothing to do (typically the FP64 pass when no configuration was left
  // undecided): leave before the set-up, the whole block at once
  if (ncfg <= 0) return;
  // a small batch runs on about one wave per configuration (at least
  // KG_MIN_WAVES): every further wave would only probe the counters
  const int nw = min((int)gridDim.x, max(KG_MIN_WAVES, ncfg));
  if ((int)blockIdx.x >= nw) return;
  unsigned n_box = 0, n_tri = 0, n_items = 0;

  for (int i = lane; i < NE; i += 64) {
    e_cfg[i] = -1;
    e_open[i] = 0;
  }
  s_pend[lane] = 0;
  s_unc[lane] = 0;
  __syncthreads();

  auto cfg_of = [&](int k) { return a.cfg_list ? a.cfg_list[k] : k; };
  auto range = [&](int c, int& b, int& e) {
    if (a.offsets) {
      b = a.offsets[c];
      e = a.offsets[c + 1];
    } else {
      b = a.ibase ? a.ibase[c] : c * a.cap;
      e = b + a.counts[c];
    }
  };
  // ring positions (monotonic, wave-uniform): [head, tail) fetched entries,
  // fe = entry whose items are being started, [it, itend) its next items
  // static share: S0 configurations per wave, the rest dynamic from dyn0
  const int S0 = (int)(((long long)ncfg * a.static_pct / 100) / nw);
  const int dyn0 = S0 * nw;
  const int nsub = max(1, a.nsub);
  const int rlen = (ncfg - dyn0 + nsub - 1) / nsub;  // configurations per dynamic range
  const int wpr = max(1, nw / nsub);                 // waves per range
  int sub = (int)blockIdx.x % nsub, tried = 0;
  int head = 0, tail = 0, fe = 0, it = 0, itend = 0, last_k0 = dyn0 + sub * rlen, soff = 0;
  bool fe_open = false, exhausted = false;
  int sp = 0, gsp = 0, nl = 0;
  uint64_t cont = KG_NOTASK;  // this lane's depth-first continuation

  // an entry whose open count reaches 0 is done; its result is published
  // (and the entry freed) by finalize_done at the next fetch or at exit
  auto settle_entry = [&](int e, int delta) { atomicAdd(&e_open[e], delta); };
  auto finalize_done = [&]() {
    if (lane < NE) {
      const int cc = e_cfg[lane];
      if (cc >= 0 && e_open[lane] == 0) {
        finalize_cfg<FPAIR>(a.fin, cc, s_best[lane][0], s_best[lane][1], s_uncmin[lane][0], s_uncmin[lane][1]);
        e_cfg[lane] = -1;
      }
    }
    __syncthreads();
  };
  auto settle = [&](int slot, int delta) {
    if (atomicAdd(&s_pend[slot], delta) + delta == 0) {
      const int e = s_cfgk[slot];
      const int rk = s_rank[slot];
      const int g = (rk >> KG_RANK_GROUP_BIT) & 1, r = rk & ((1 << KG_RANK_GROUP_BIT) - 1);
      if (s_unc[slot] && !item_dead(mode, s_best[e], g, r)) atomicMin(&s_uncmin[e][g], r);
      settle_entry(e, -1);
    }
  };

  auto ensure_room = [&](int room) {
    if (sp + room <= SCAP) return;
    const int h = SCAP / 2;
    if (gsp + h <= a.spill_cap) {
      for (int i = lane; i < h; i += 64) gsp_base[gsp + i] = stk[i];
      gsp += h;
    } else {
      // spill region full: drop the oldest half, marking its items for the
      // exact re-check
      for (int i = lane; i < h; i += 64) {
        const int sl = (int)(stk[i] >> 58);
        s_unc[sl] = 1;
        settle(sl, -1);
      }
      if (lane == 0) atomicExch(a.status, 1);
    }
    __syncthreads();
    for (int i = lane; i < sp - h; i += 64) stk[i] = stk[i + h];
    sp -= h;
    __syncthreads();
  };

  // triangle tests of queued leaf pairs: lane groups of 9 (7 pairs per pass,
  // at most 3 passes per call) split each pair's axes between them
  auto run_leaves = [&]() {
    const int m = min(nl, 21);
    const int li = lane / 9, sub = lane - 9 * li;
    for (int b = 0; b < m; b += 7) {
      const int idx = b + li;
      const bool act = li < 7 && idx < m;
      int slot = 0, g = 0, r = 0, ck = 0, bits = 0, ta = 0, tb = 0;
      bool live = false;
      if (act) {
        const uint64_t e = lq[nl - 1 - idx];
        slot = (int)(e >> 58);
        ta = (int)((e >> 29) & 0x1fffffff);
        tb = (int)(e & 0x1fffffff);
        const int rk = s_rank[slot];
        g = (rk >> KG_RANK_GROUP_BIT) & 1;
        r = rk & ((1 << KG_RANK_GROUP_BIT) - 1);
        ck = s_cfgk[slot];
        live = !item_dead(mode, s_best[ck], g, r);
        if (live) {
          Real rel[12];
#pragma unroll
          for (int k = 0; k < 12; k++) rel[k] = s_rel[k][slot];
          V3T<Real> tp[3], tq[3];
          load_tri<Real>(rpool, ta, tp);
          load_tri<Real>(rpool, tb, tq);
#pragma unroll
          for (int v = 0; v < 3; v++) tp[v] = xf_apply<Real>(rel, tp[v]);
          bits = tri_tri_split<Real, EXACT>(tp, tq, (Real)s_tau[slot], sub);
        }
      }
      const uint64_t sepm = ballot64((bits & 1) != 0), unsm = ballot64((bits & 2) != 0);
      if (act && sub == 0) {
        if (live) {
          n_tri++;
          const uint64_t gm = 0x1ffull << (9 * li);
          if (!(sepm & gm)) {
            if (unsm & gm) {
              // undecided: recorded for the FP64 pass's exact test of this
              // pair (the item goes on as if it were separated), or, when
              // there is no room, the configuration is traversed again
              int k = a.fin.xpair_cap;
              if (!EXACT && a.fin.xpairs) k = atomicAdd(a.fin.xpair_count, 1);
              if (k < a.fin.xpair_cap)
                a.fin.xpairs[k] = make_int4(e_cfg[ck], s_item[slot], ta, tb);
              else
                s_unc[slot] = 1;
            } else {
              const unsigned long long key = ((unsigned long long)(uint32_t)r << 32) | (uint32_t)s_item[slot];
              atomicMin(&s_best[ck][g], key);
            }
          }
        }
        settle(slot, -1);
      }
    }
    nl -= m;
    __syncthreads();
  };

  // In a wave's tail (few lanes with work) the item-start and publishing
  // steps run only every tail_reps-th iteration: the iterations between
  // them pop the pool and test boxes, so the deep items that remain descend
  // several levels per scheduling step.
  bool quit = false;
  int rep = 0, reps = 1;
  for (;;) {
    const bool full = rep == 0;
    if (full) {
      // publish finished configurations and free their entries (the only
      // place results are written)
      finalize_done();
      if (quit) break;
    }
    // ---- lanes without a continuation pop the pool (bring spilled tasks
    // back first when it runs low)
    const bool hungry = cont == KG_NOTASK;
    const uint64_t hm = ballot64(hungry);
    const int nh = __popcll(hm);
    if (nh > sp && gsp > 0) {
      const int k = min(gsp, SCAP / 2);
      for (int i = sp - 1 - lane; i >= 0; i -= 64) stk[i + k] = stk[i];
      __syncthreads();
      for (int i = lane; i < k; i += 64) stk[i] = gsp_base[gsp - k + i];
      gsp -= k;
      sp += k;
      __syncthreads();
    }
    const int hp = lane_prefix(hm);
    const int npop = min(nh, sp);
    if (hungry && hp < npop) cont = stk[sp - 1 - hp];
    sp -= npop;
    if (full) {
    // ---- lanes still without work start new items on free slots
    const bool want = cont == KG_NOTASK;
    const uint64_t wm = ballot64(want);
    const uint64_t fm = ballot64(s_pend[lane] == 0);
    const int navail = min(__popcll(wm), __popcll(fm));
    if (navail > 0 && fe == tail && !exhausted) {
      // fetch configurations into the ring (wave-uniform): first this wave's
      // static share, then chunks from the global counter
      while (head < tail && __builtin_amdgcn_readfirstlane(e_cfg[head % NE]) < 0) head++;
      const int room = NE - (tail - head);
      if (room > 0) {
        int k0 = 0, nk = 0, kst = 1;
        if (soff < S0) {
          // a heterogeneous batch's static share is interleaved over the
          // waves (this wave's configurations blockIdx, blockIdx + nw, ...):
          // its configurations come robot by robot, and contiguous slices
          // would give some waves only the heaviest robot's (3% on the 64-robot
          // batch); a single robot's batch keeps contiguous slices, whose
          // items sit together in memory (interleaved: 3% slower)
          if (a.cfg_plan) {
            k0 = soff * nw + (int)blockIdx.x;
            kst = nw;
          } else {
            k0 = (int)blockIdx.x * S0 + soff;
          }
          nk = min(room, S0 - soff);
          soff += nk;
        } else {
          // the dynamic part is split into nsub ranges with a counter each
          // (own cache line); a wave starts on range blockIdx % nsub and
          // moves on when it runs dry
          for (;;) {
            if (tried >= nsub) {
              exhausted = true;
              break;
            }
            const int r0 = dyn0 + sub * rlen, r1 = min(ncfg, r0 + rlen);
            // A range that is empty, or that a wave already found drained,
            // is passed without an atomic: atomics on one address from all
            // waves queue behind each other (tens of ns each), so thousands
            // of waves each probing every drained counter would cost the
            // launch about 0.2 ms however little work it holds.  The flag is
            // a hint read through the caches (a stale 0 costs one probe).
            int* drained = a.chunk_drained + sub * KG_COUNTER_STRIDE;
            bool dry = r0 >= r1;
            if (!dry)
              dry = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(drained, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) != 0;
            if (!dry) {
              const int ch = min(room, max(1, min(a.cpw, (r1 - last_k0) / wpr)));
              if (lane == 0) s_chunk = r0 + atomicAdd(a.chunk_counter + sub * KG_COUNTER_STRIDE, ch);
              __syncthreads();
              k0 = __builtin_amdgcn_readfirstlane(s_chunk);
              __syncthreads();
              if (k0 < r1) {
                last_k0 = k0;
                nk = min(ch, r1 - k0);
                break;
              }
              if (lane == 0) __hip_atomic_store(drained, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            sub = sub + 1 == nsub ? 0 : sub + 1;
            tried++;
            last_k0 = dyn0 + sub * rlen;
          }
        }
        if (nk > 0) {
          if (lane < nk) {
            const int e = (tail + lane) % NE;
            const int cc = cfg_of(k0 + lane * kst);
            e_cfg[e] = cc;
            range(cc, e_beg[e], e_end[e]);
            e_open[e] = 1;
            s_best[e][0] = s_best[e][1] = KG_NONE64;
            s_uncmin[e][0] = s_uncmin[e][1] = KG_RANK_NONE;
          }
          tail += nk;
          __syncthreads();
        }
      }
    }
    int myitem = -1, myent = -1;
    const int rel_from = fe;
    {
      // the j-th wanting lane takes the j-th item in ring order; entries
      // whose items are all started are passed (their start hold released)
      const int wp = lane_prefix(wm);
      int assigned = 0;
      while (fe < tail) {
        if (!fe_open) {
          it = __builtin_amdgcn_readfirstlane(e_beg[fe % NE]);
          itend = __builtin_amdgcn_readfirstlane(e_end[fe % NE]);
          fe_open = true;
        }
        if (it >= itend) {
          fe++;
          fe_open = false;
          continue;
        }
        if (assigned >= navail) break;
        const int take = min(navail - assigned, itend - it);
        if (want && wp >= assigned && wp < assigned + take) {
          myitem = it + (wp - assigned);
          myent = fe % NE;
        }
        assigned += take;
        it += take;
      }
    }
    // slot for each new item: the j-th free slot (by lane index)
    int myslot = -1;
    if (myitem >= 0) myslot = nth_set_bit(fm, lane_prefix(wm));  // rank among wanting lanes < navail
    __syncthreads();
    if (myitem >= 0) {
      n_items++;
      atomicAdd(&e_open[myent], 1);
      const PairItem P = items[myitem];
      int rootA, rootB;
      if (!EXACT && REC) {
        const float4* q = reinterpret_cast<const float4*>(recs + myitem);
        const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        s_rel[0][myslot] = q0.x; s_rel[1][myslot] = q0.y; s_rel[2][myslot] = q0.z; s_rel[3][myslot] = q0.w;
        s_rel[4][myslot] = q1.x; s_rel[5][myslot] = q1.y; s_rel[6][myslot] = q1.z; s_rel[7][myslot] = q1.w;
        s_rel[8][myslot] = q2.x; s_rel[9][myslot] = q2.y; s_rel[10][myslot] = q2.z; s_rel[11][myslot] = q2.w;
        s_tau[myslot] = q3.x;
        s_pad[myslot] = q3.y;
        rootA = __float_as_int(q3.z);
        rootB = __float_as_int(q3.w);
      } else {
        Real rel[12];
        float tau, pad;
        item_frame<Real, EXACT>(a, P, rel, tau, pad, rootA, rootB);
#pragma unroll
        for (int k = 0; k < 12; k++) s_rel[k][myslot] = rel[k];
        s_tau[myslot] = tau;
        s_pad[myslot] = pad;
      }
      s_cfgk[myslot] = myent;
      s_rank[myslot] = P.rank;
      s_item[myslot] = myitem;
      s_pend[myslot] = 1;
      s_unc[myslot] = 0;
      cont = mk_task(myslot, 0, 1, rootA, rootB);
    }
    __syncthreads();
    // release the start hold of the entries passed above (after the items
    // started from them are counted)
    if (lane == 0)
      for (int e = rel_from; e < fe; e++) settle_entry(e % NE, -1);
    __syncthreads();
    }

    // ---- termination: no path, nothing pooled, no leaves, nothing to start
    const bool working = cont != KG_NOTASK;
    const uint64_t wk = ballot64(working);
    if (full) reps = __popcll(wk) < a.tail_lanes ? a.tail_reps : 1;
    rep = rep + 1 < reps ? rep + 1 : 0;
    if (!wk) {
      if (!full) {
        rep = 0;
        continue;
      }
      if (sp == 0 && gsp == 0) {
        if (nl > 0) {
          run_leaves();
        } else if (fe == tail && exhausted) {
          quit = true;
        }
      }
      continue;
    }

    // ---- process the continuation: up to two box tests
    uint64_t out0 = 0, out1 = 0, lf0 = 0, lf1 = 0;
    int nout = 0, nlf = 0, slot = 0;
    if (working) {
      const uint64_t task = cont;
      cont = KG_NOTASK;
      slot = (int)(task >> 58);
      const int side = (int)((task >> 57) & 1), single = (int)((task >> 56) & 1);
      const int x = (int)((task >> 28) & 0xfffffff), y = (int)(task & 0xfffffff);
      const int rk = s_rank[slot];
      const int g = (rk >> KG_RANK_GROUP_BIT) & 1, r = rk & ((1 << KG_RANK_GROUP_BIT) - 1);
      if (!item_dead(mode, s_best[s_cfgk[slot]], g, r)) {
        const NodeF O = load_nodef(nodes, y);
        const NodeF S0 = load_nodef(nodes, x);
        const NodeF S1 = load_nodef(nodes, single ? x : x + 1);
        // S (the split node) is in mesh A when A was split (or for a
        // single pair); otherwise the item transform is inverted
        const bool aIsS = single || side == 0;
        Real rel[12];
#pragma unroll
        for (int k = 0; k < 12; k++) rel[k] = s_rel[k][slot];
        if (!aIsS) {
          Real inv[12];
#pragma unroll
          for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r2 = 0; r2 < 3; r2++) inv[c * 3 + r2] = rel[r2 * 3 + c];
#pragma unroll
          for (int r2 = 0; r2 < 3; r2++)
            inv[9 + r2] = -(rel[r2 * 3] * rel[9] + rel[r2 * 3 + 1] * rel[10] + rel[r2 * 3 + 2] * rel[11]);
#pragma unroll
          for (int k = 0; k < 12; k++) rel[k] = inv[k];
        }
        TaskFrame<Real> F;
        task_frame<Real>(O, rel, F);
        const Real pad = (Real)s_pad[slot];
        const bool lO = O.ch < 0;
        Real m0 = 0;
#pragma unroll
        for (int k = 0; k < 2; k++) {
          if (k == 1 && single) break;
          const NodeF& S = k == 0 ? S0 : S1;  // k is a constant after unrolling
          const int nodeS = x + k;
          n_box++;
          Real fm;
          if (!frame_overlap<Real>(F, S, pad, pad_rel, fm)) continue;
          const bool lS = S.ch < 0;
          if (lS && lO) {
            const uint64_t w = aIsS ? mk_leaf(slot, ~S.ch, ~O.ch) : mk_leaf(slot, ~O.ch, ~S.ch);
            if (nlf == 0) lf0 = w; else lf1 = w;
            nlf++;
            continue;
          }
          // descend the box of larger volume (fewer box tests than larger
          // extents on thin and elongated boxes)
          const bool splitS = lO || (!lS && S.e[0] * S.e[1] * S.e[2] >= O.e[0] * O.e[1] * O.e[2]);
          // side 0: children of an A-node against a B-node; side 1: the reverse
          const int nside = (splitS == aIsS) ? 0 : 1;
          const uint64_t w = splitS ? mk_task(slot, nside, 0, S.ch, y) : mk_task(slot, nside, 0, O.ch, nodeS);
          // the more deeply overlapping pair is continued (an intersection,
          // which ends the item, tends to be found sooner), the other pooled
          if (nout == 0) {
            out0 = w;
            m0 = fm;
          } else if (fm > m0) {
            out1 = out0;
            out0 = w;
          } else {
            out1 = w;
          }
          nout++;
        }
      }
    }
    // keep the first child pair as this lane's continuation
    if (nout >= 1) cont = out0;

    // ---- queue leaves (flush batches first if the queue could overflow)
    {
      const uint64_t l1 = ballot64(nlf >= 1), l2 = ballot64(nlf >= 2);
      const int ltot = __popcll(l1) + __popcll(l2);
      if (nl + ltot > LQ) {
        __syncthreads();
        while (nl + ltot > LQ) run_leaves();
      }
      const int pos = nl + lane_prefix(l1) + lane_prefix(l2);
      if (nlf >= 1) lq[pos] = lf0;
      if (nlf >= 2) lq[pos + 1] = lf1;
      nl += ltot;
    }
    if (working) settle(slot, nout + nlf - 1);

    // ---- push the second child pair to the pool
    const uint64_t pm = ballot64(nout >= 2);
    const int ptot = __popcll(pm);
    ensure_room(ptot);
    if (nout >= 2) stk[sp + lane_prefix(pm)] = out1;
    sp += ptot;
    __syncthreads();
    if (nl >= a.leaf_min) run_leaves();
  }
  if (a.counters && (n_box | n_tri | n_items)) {
    // statistics: waves add to one of KG_CNT_COPIES lines (atomics on one
    // line from every wave would queue for tens of microseconds)
    unsigned long long* cnt = a.counters + (blockIdx.x % KG_CNT_COPIES) * (KG_COUNTER_STRIDE / 2);
    atomicAdd(&cnt[0], (unsigned long long)n_box);
    atomicAdd(&cnt[1], (unsigned long long)n_tri);
    atomicAdd(&cnt[2], (unsigned long long)n_items);
  }
}

__global__ void k_fill_int(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}


}  // namespace

// ------------------------------------------------------------ launchers

hipError_t kgl_fk_f32(const LinkDesc* links, int nlinks, const float* Q, int N, float* Tout, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  k_fk_f32<<<(N + 255) / 256, 256, 0, s>>>(links, nlinks, Q, N, Tout);
  return hipGetLastError();
}

hipError_t kgl_prep_count(const PrepArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  k_prep<0, 0><<<(a.N + 255) / 256, 256, 0, s>>>(a);
  k_fill_int<<<1, 1, 0, s>>>(a.counts + a.N, 1, 0);
  return hipGetLastError();
}

hipError_t kgl_scan(int* counts, int* offsets, int n, DevBuf& tmp, hipStream_t s) {
  size_t bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, counts, offsets, n, s);
  if (e != hipSuccess) return e;
  e = tmp.ensure(bytes);
  if (e != hipSuccess) return e;
  return hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, counts, offsets, n, s);
}

hipError_t kgl_prep_fill(const PrepArgs& a, hipStream_t s) {
  if (a.N <= 0) return hipSuccess;
  k_prep<1, 0><<<(a.N + 255) / 256, 256, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t kgl_prep_single(const PrepArgs& a, hipStream_t s) {
  const int n = a.N > a.nreset ? a.N : a.nreset;
  if (n <= 0) return hipSuccess;
  // a device-side count (at most a.N) bounds the grid: the blocks loop
  constexpr int kMaxBlocksDev = 4096;
  if (a.group > 0) {
    constexpr int TB = KG_PREP_TB;
    const int ng = TB / a.group;
    int blocks = (a.N + ng - 1) / ng;
    if (a.N_dev) blocks = min(blocks, kMaxBlocksDev);
    blocks = max(blocks, (a.nreset + TB - 1) / TB);
    const bool m = a.cfg_plan != nullptr;
    if (a.N_dev) {
      if (a.group == 8)
        m ? k_prep_group<8, true, true><<<blocks, TB, 0, s>>>(a) : k_prep_group<8, false, true><<<blocks, TB, 0, s>>>(a);
      else if (a.group == 16)
        m ? k_prep_group<16, true, true><<<blocks, TB, 0, s>>>(a)
          : k_prep_group<16, false, true><<<blocks, TB, 0, s>>>(a);
      else
        m ? k_prep_group<32, true, true><<<blocks, TB, 0, s>>>(a)
          : k_prep_group<32, false, true><<<blocks, TB, 0, s>>>(a);
    } else {
      if (a.group == 8)
        m ? k_prep_group<8, true, false><<<blocks, TB, 0, s>>>(a)
          : k_prep_group<8, false, false><<<blocks, TB, 0, s>>>(a);
      else if (a.group == 16)
        m ? k_prep_group<16, true, false><<<blocks, TB, 0, s>>>(a)
          : k_prep_group<16, false, false><<<blocks, TB, 0, s>>>(a);
      else
        m ? k_prep_group<32, true, false><<<blocks, TB, 0, s>>>(a)
          : k_prep_group<32, false, false><<<blocks, TB, 0, s>>>(a);
    }
  } else if (a.lds_links) {
    int blocks = (n + KG_PREP_LDS_BLOCK - 1) / KG_PREP_LDS_BLOCK;
    if (a.N_dev) blocks = max(min(blocks, kMaxBlocksDev), (a.nreset + KG_PREP_LDS_BLOCK - 1) / KG_PREP_LDS_BLOCK);
    k_prep<2, KG_PREP_LDS_LINKS><<<blocks, KG_PREP_LDS_BLOCK, 0, s>>>(a);
  } else {
    int blocks = (n + 255) / 256;
    if (a.N_dev) blocks = max(min(blocks, kMaxBlocksDev), (a.nreset + 255) / 256);
    k_prep<2, 0><<<blocks, 256, 0, s>>>(a);
  }
  return hipGetLastError();
}

template <typename Real, bool EXACT, int SCAP, bool FPAIR>
static void launch_rec(const TravArgs& a, hipStream_t s) {
  if (!EXACT && a.recs)
    k_collide<Real, EXACT, SCAP, FPAIR, !EXACT><<<a.grid, 64, 0, s>>>(a, a.pool.nodes, a.pool.tv0, a.pool.tv1,
                                                                    a.pool.tv2, a.pool.tv64, a.items, a.recs);
  else
    k_collide<Real, EXACT, SCAP, FPAIR, false><<<a.grid, 64, 0, s>>>(a, a.pool.nodes, a.pool.tv0, a.pool.tv1,
                                                                   a.pool.tv2, a.pool.tv64, a.items, a.recs);
}

template <typename Real, bool EXACT, int SCAP>
static void launch_scap(const TravArgs& a, hipStream_t s) {
  if (a.fin.first_pair)
    launch_rec<Real, EXACT, SCAP, true>(a, s);
  else
    launch_rec<Real, EXACT, SCAP, false>(a, s);
}

template <typename Real, bool EXACT>
static hipError_t launch_collide(const TravArgs& a, hipStream_t s) {
  if (a.scap == 256)
    launch_scap<Real, EXACT, 256>(a, s);
  else
    launch_scap<Real, EXACT, 512>(a, s);
  return hipGetLastError();
}

hipError_t kgl_collide_f32(const TravArgs& a, hipStream_t s) { return launch_collide<float, false>(a, s); }
hipError_t kgl_collide_f64(const TravArgs& a, hipStream_t s) { return launch_collide<double, true>(a, s); }


__global__ void k_status_copy(const int* __restrict__ st, int w_spill, int w_unc, int w_redo, int32_t* out) {
  if (threadIdx.x == 0) {
    out[0] = st[w_spill];
    out[1] = st[w_unc];
    out[2] = st[w_redo];
    out[3] = 0;
  }
}

hipError_t kgl_status_copy(const int* status, int w_spill, int w_unc, int w_redo, int32_t* out, hipStream_t s) {
  k_status_copy<<<1, 64, 0, s>>>(status, w_spill, w_unc, w_redo, out);
  return hipGetLastError();
}

__global__ void k_status_accum(const int* __restrict__ st, int* __restrict__ acc, int w_spill, int w_unc, int w_redo,
                               int first) {
  if (threadIdx.x == 0) {
    const int sp = st[w_spill], un = st[w_unc], rd = st[w_redo];
    acc[w_spill] = first ? sp : (acc[w_spill] | sp);
    acc[w_unc] = first ? un : acc[w_unc] + un;
    acc[w_redo] = first ? rd : acc[w_redo] + rd;
  }
}

hipError_t kgl_status_accum(const int* st, int* acc, int w_spill, int w_unc, int w_redo, int first, hipStream_t s) {
  k_status_accum<<<1, 64, 0, s>>>(st, acc, w_spill, w_unc, w_redo, first);
  return hipGetLastError();
}

// ------------------------------------------------------------ scratch of the allocation-free entry point
//
// kg_feasible_batch_dev_noalloc carves every workspace of a batch (status
// words, joint-limit flags, item counts, re-check list, pair items, start
// records, spill region) out of this array instead of allocating.  Each GPU
// holds its own copy, at its own address, until the process ends; host code
// never passes the array's name to a kernel (it is a host-side stand-in),
// only the per-device address kgl_noalloc_scratch looks up.
static __device__ __align__(256) unsigned char scratch[KG_NOALLOC_BYTES];

hipError_t kgl_noalloc_scratch(int device, void** ptr) {
  static std::mutex mu;
  static void* addr[KG_MAX_DEVICES] = {};
  if (device < 0 || device >= KG_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    // the caller has made `device` current
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

hipError_t kgl_fill_int(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill_int<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

int kgl_collide_grid(int exact, int num_cus, int scap) {
  int nb = 0;
  hipError_t e;
  if (scap == 256)
    e = exact ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_collide<double, true, 256, false, false>, 64, 0)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_collide<float, false, 256, false, true>, 64, 0);
  else
    e = exact ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_collide<double, true, 512, false, false>, 64, 0)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_collide<float, false, 512, false, true>, 64, 0);
  if (e != hipSuccess || nb < 1) nb = 8;
  return nb * num_cus;
}

// ------------------------------------------------------------ SDF queries

namespace {

// Cell-centred trilinear value and gradient (ImplicitSurface convention:
// samples at bmin + (i+1/2) h).  Points outside the box of cell centres are
// clamped to it; the gradient along a clamped axis is 0.
// The continuous cell coordinate and its fraction are formed in FP64 (a few
// operations per point): in FP32 their rounding, amplified by steep fields,
// would dominate the error; the lerps run in FP32.
__device__ inline float sdf_trilinear(const SdfDev& g, float x, float y, float z, float* gr) {
  const double u[3] = {((double)x - g.bmin64[0]) * g.inv_h64[0] - 0.5, ((double)y - g.bmin64[1]) * g.inv_h64[1] - 0.5,
                       ((double)z - g.bmin64[2]) * g.inv_h64[2] - 0.5};
  int i0[3], i1[3];
  float f[3], gs[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    gs[a] = g.inv_h[a];
    const double um = (double)(g.dims[a] - 1);
    double v = u[a];
    if (v <= 0.0) { v = 0.0; gs[a] = 0.f; }
    else if (v >= um) { v = um; gs[a] = 0.f; }
    int i = (int)floor(v);
    i = min(i, g.dims[a] - 2);
    i = max(i, 0);
    i0[a] = i;
    i1[a] = min(i + 1, g.dims[a] - 1);
    f[a] = g.dims[a] > 1 ? (float)(v - (double)i) : 0.f;
    if (g.dims[a] == 1) gs[a] = 0.f;
  }
  float c000, c001, c010, c011, c100, c101, c110, c111;
  if (g.v8) {
    // the cell's 8 corners, stored together
    const size_t cell = ((size_t)i0[0] * g.cdims[1] + i0[1]) * g.cdims[2] + i0[2];
    const float4 a = g.v8[2 * cell], b = g.v8[2 * cell + 1];
    c000 = a.x; c001 = a.y; c010 = a.z; c011 = a.w;
    c100 = b.x; c101 = b.y; c110 = b.z; c111 = b.w;
  } else {
    const int s1 = g.dims[2], s0 = g.dims[1] * g.dims[2];
    const float* v = g.v;
    c000 = v[i0[0] * s0 + i0[1] * s1 + i0[2]]; c001 = v[i0[0] * s0 + i0[1] * s1 + i1[2]];
    c010 = v[i0[0] * s0 + i1[1] * s1 + i0[2]]; c011 = v[i0[0] * s0 + i1[1] * s1 + i1[2]];
    c100 = v[i1[0] * s0 + i0[1] * s1 + i0[2]]; c101 = v[i1[0] * s0 + i0[1] * s1 + i1[2]];
    c110 = v[i1[0] * s0 + i1[1] * s1 + i0[2]]; c111 = v[i1[0] * s0 + i1[1] * s1 + i1[2]];
  }
  const float fx = f[0], fy = f[1], fz = f[2];
  const float c00 = c000 * (1.f - fz) + c001 * fz, c01 = c010 * (1.f - fz) + c011 * fz;
  const float c10 = c100 * (1.f - fz) + c101 * fz, c11 = c110 * (1.f - fz) + c111 * fz;
  const float c0 = c00 * (1.f - fy) + c01 * fy, c1 = c10 * (1.f - fy) + c11 * fy;
  if (gr) {
    const float dz0 = (c001 - c000) * (1.f - fy) + (c011 - c010) * fy;
    const float dz1 = (c101 - c100) * (1.f - fy) + (c111 - c110) * fy;
    gr[0] = (c1 - c0) * gs[0];
    gr[1] = ((c01 - c00) * (1.f - fx) + (c11 - c10) * fx) * gs[1];
    gr[2] = (dz0 * (1.f - fx) + dz1 * fx) * gs[2];
  }
  return c0 * (1.f - fx) + c1 * fx;
}

// World-frame query: map into the grid frame, rotate the gradient back.
__device__ inline float sdf_world(const SdfDev& g, float x, float y, float z, float* gr) {
  if (g.ident) return sdf_trilinear(g, x, y, z, gr);
  const float* X = g.Xinv;
  const float lx = X[0] * x + X[3] * y + X[6] * z + X[9];
  const float ly = X[1] * x + X[4] * y + X[7] * z + X[10];
  const float lz = X[2] * x + X[5] * y + X[8] * z + X[11];
  float gl[3];
  const float d = sdf_trilinear(g, lx, ly, lz, gr ? gl : nullptr);
  if (gr) {  // R = Xinv^T
    gr[0] = X[0] * gl[0] + X[1] * gl[1] + X[2] * gl[2];
    gr[1] = X[3] * gl[0] + X[4] * gl[1] + X[5] * gl[2];
    gr[2] = X[6] * gl[0] + X[7] * gl[1] + X[8] * gl[2];
  }
  return d;
}

__global__ __launch_bounds__(256) void k_sdf_eval(SdfDev g, const float* __restrict__ P, int n, float* __restrict__ d,
                                                  float* __restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gr[3];
  const float v = sdf_world(g, P[3 * i], P[3 * i + 1], P[3 * i + 2], grad ? gr : nullptr);
  d[i] = v;
  if (grad) {
    grad[3 * i] = gr[0];
    grad[3 * i + 1] = gr[1];
    grad[3 * i + 2] = gr[2];
  }
}

// SDF hinge penalty with a group of GL lanes per configuration (waypoint):
// the joint frames are formed by lanes 0..n-1 and composed into world frames
// in LDS by the group's first lane; the points are spread over the group
// (lane l takes points l, l + GL, ...), each lane accumulating per-link
// force and moment sums that are added into LDS; the loss is reduced over
// the group, and the group's first lane runs the FK-Jacobian adjoint
// (subtree sums added into the parent entries, children after parents):
//   dL/dq_j = w_j . (M_j - o_j x F_j) (revolute), w_j . F_j (prismatic).
// Dynamic LDS: gpb groups x n links x (12 frame + 6 sum) floats.
template <bool BWD, int GL>
__global__ __launch_bounds__(256) void k_sdf_pen_group(const LinkDesc* __restrict__ links, int n, SdfDev g,
                                                       const float* __restrict__ Q, int N,
                                                       const float* __restrict__ pts, const int* __restrict__ poff,
                                                       float dsafe, float* __restrict__ loss_out,
                                                       const float* __restrict__ dloss, float* __restrict__ dQ,
                                                       int gpb, int nb8) {
  extern __shared__ float dyn[];
  __shared__ float sT0[KG_MAX_LINKS][12];
  __shared__ float sAx[KG_MAX_LINKS][3];
  __shared__ int sPar[KG_MAX_LINKS], sTy[KG_MAX_LINKS], sOff[KG_MAX_LINKS + 1];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    for (int k = 0; k < 12; k++) sT0[i][k] = (float)links[i].T0[k];
    for (int k = 0; k < 3; k++) sAx[i][k] = (float)links[i].axis[k];
    sPar[i] = links[i].parent;
    sTy[i] = links[i].jtype;
  }
  for (int i = threadIdx.x; i <= n; i += blockDim.x) sOff[i] = poff[i];
  const int grp = threadIdx.x / GL, l = threadIdx.x % GL;
  // XCD-aware order: the hardware deals blocks to the 8 XCDs round-robin, so
  // block b runs on XCD b % 8; with nb8 > 0 the blocks of one XCD take a
  // contiguous run of waypoints (neighbouring waypoints of a trajectory gather
  // the same grid lines, which then stay in that XCD's L2)
  const int lb = nb8 ? (int)(blockIdx.x & 7) * nb8 + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  const int c = lb * gpb + grp;
  const bool valid = c < N;
  float* sT = dyn + (size_t)grp * n * 12;
  float* sFM = dyn + (size_t)gpb * n * 12 + (size_t)grp * n * 6;
  __syncthreads();
  // joint frames T0_k J(q_k)
  for (int k = l; k < n; k += GL) {
    for (int i = 0; i < 6; i++) sFM[k * 6 + i] = 0.f;
    if (!valid) continue;
    const float q = Q[(size_t)c * n + k];
    const float* w = sAx[k];
    float J[12];
    if (sTy[k] == 1) {
#pragma unroll
      for (int i = 0; i < 9; i++) J[i] = (i % 4 == 0) ? 1.f : 0.f;
      J[9] = w[0] * q; J[10] = w[1] * q; J[11] = w[2] * q;
    } else {
      float sn, co;
      sincosf(q, &sn, &co);
      const float v = 1.f - co;
#pragma unroll
      for (int col = 0; col < 3; col++)
#pragma unroll
        for (int row = 0; row < 3; row++) J[col * 3 + row] = w[row] * w[col] * v;
      J[0] += co; J[4] += co; J[8] += co;
      J[1] += sn * w[2]; J[2] -= sn * w[1];
      J[3] -= sn * w[2]; J[5] += sn * w[0];
      J[6] += sn * w[1]; J[7] -= sn * w[0];
      J[9] = J[10] = J[11] = 0.f;
    }
    const float* A = sT0[k];
    float* L = sT + k * 12;
#pragma unroll
    for (int col = 0; col < 3; col++)
#pragma unroll
      for (int r = 0; r < 3; r++) L[col * 3 + r] = A[r] * J[col * 3] + A[3 + r] * J[col * 3 + 1] + A[6 + r] * J[col * 3 + 2];
#pragma unroll
    for (int r = 0; r < 3; r++) L[9 + r] = A[r] * J[9] + A[3 + r] * J[10] + A[6 + r] * J[11] + A[9 + r];
  }
  __syncthreads();
  // world frames (parents precede their children)
  if (valid && l == 0)
    for (int k = 0; k < n; k++) {
      const int p = sPar[k];
      if (p < 0) continue;
      const float* P = sT + p * 12;
      float* T = sT + k * 12;
      float L[12];
#pragma unroll
      for (int i = 0; i < 12; i++) L[i] = T[i];
#pragma unroll
      for (int col = 0; col < 3; col++)
#pragma unroll
        for (int r = 0; r < 3; r++) T[col * 3 + r] = P[r] * L[col * 3] + P[3 + r] * L[col * 3 + 1] + P[6 + r] * L[col * 3 + 2];
#pragma unroll
      for (int r = 0; r < 3; r++) T[9 + r] = P[r] * L[9] + P[3 + r] * L[10] + P[6 + r] * L[11] + P[9 + r];
    }
  __syncthreads();
  float loss = 0.f;
  if (valid) {
    const int npts = sOff[n];
    int k = 0, cur = -1;
    float T[12];
    float f0 = 0.f, f1 = 0.f, f2 = 0.f, m0 = 0.f, m1 = 0.f, m2 = 0.f;
    auto flush = [&]() {
      if (BWD && cur >= 0) {
        float* F = sFM + cur * 6;
        atomicAdd(F + 0, f0); atomicAdd(F + 1, f1); atomicAdd(F + 2, f2);
        atomicAdd(F + 3, m0); atomicAdd(F + 4, m1); atomicAdd(F + 5, m2);
      }
      f0 = f1 = f2 = m0 = m1 = m2 = 0.f;
    };
    for (int i = l; i < npts; i += GL) {
      while (i >= sOff[k + 1]) k++;
      if (k != cur) {
        flush();
        cur = k;
#pragma unroll
        for (int j = 0; j < 12; j++) T[j] = sT[k * 12 + j];
      }
      const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
      const float x = T[0] * px + T[3] * py + T[6] * pz + T[9];
      const float y = T[1] * px + T[4] * py + T[7] * pz + T[10];
      const float z = T[2] * px + T[5] * py + T[8] * pz + T[11];
      float gr[3];
      const float d = sdf_world(g, x, y, z, BWD ? gr : nullptr);
      const float h = dsafe - d;
      if (h > 0.f) {
        loss += h * h;
        if (BWD) {
          const float s2 = -2.f * h;
          const float g0 = s2 * gr[0], g1 = s2 * gr[1], g2 = s2 * gr[2];
          f0 += g0; f1 += g1; f2 += g2;
          m0 += y * g2 - z * g1;
          m1 += z * g0 - x * g2;
          m2 += x * g1 - y * g0;
        }
      }
    }
    flush();
  }
#pragma unroll
  for (int off = GL / 2; off > 0; off >>= 1) loss += __shfl_xor(loss, off, GL);
  if (!BWD) {
    if (valid && l == 0) loss_out[c] = loss;
    return;
  }
  __syncthreads();
  if (!valid || l != 0) return;
  const float dl = dloss ? dloss[c] : 1.f;
  for (int j = n - 1; j >= 0; j--) {
    const float* T = sT + j * 12;
    float* FM = sFM + j * 6;
    const float* w = sAx[j];
    const float w0 = T[0] * w[0] + T[3] * w[1] + T[6] * w[2];
    const float w1 = T[1] * w[0] + T[4] * w[1] + T[7] * w[2];
    const float w2 = T[2] * w[0] + T[5] * w[1] + T[8] * w[2];
    const float o0 = T[9], o1 = T[10], o2 = T[11];
    const float F0 = FM[0], F1 = FM[1], F2 = FM[2], M0 = FM[3], M1 = FM[4], M2 = FM[5];
    float gq;
    if (sTy[j] == 1)
      gq = w0 * F0 + w1 * F1 + w2 * F2;
    else
      gq = w0 * (M0 - (o1 * F2 - o2 * F1)) + w1 * (M1 - (o2 * F0 - o0 * F2)) + w2 * (M2 - (o0 * F1 - o1 * F0));
    dQ[(size_t)c * n + j] = dl * gq;
    const int p = sPar[j];
    if (p >= 0) {
      float* PF = sFM + p * 6;
      for (int i = 0; i < 6; i++) PF[i] += FM[i];
    }
  }
  if (loss_out) loss_out[c] = loss;
}

}  // namespace

hipError_t kgl_sdf_eval(const SdfDev& g, const float* pts, int n, float* d, float* grad, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_sdf_eval<<<(n + 255) / 256, 256, 0, s>>>(g, pts, n, d, grad);
  return hipGetLastError();
}

namespace {
__global__ __launch_bounds__(256) void k_sdf_build8(const float* __restrict__ v, int dx, int dy, int dz,
                                                   float4* __restrict__ out) {
  const int cx = max(dx - 1, 1), cy = max(dy - 1, 1), cz = max(dz - 1, 1);
  const size_t ncell = (size_t)cx * cy * cz;
  for (size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x; cell < ncell;
       cell += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(cell % cz), j = (int)((cell / cz) % cy), i = (int)(cell / ((size_t)cz * cy));
    // the corner rule of sdf_trilinear: the upper index is clamped to the grid
    const int i1 = min(i + 1, dx - 1), j1 = min(j + 1, dy - 1), k1 = min(k + 1, dz - 1);
    const size_t s1 = dz, s0 = (size_t)dy * dz;
    out[2 * cell] = make_float4(v[i * s0 + j * s1 + k], v[i * s0 + j * s1 + k1], v[i * s0 + j1 * s1 + k],
                                v[i * s0 + j1 * s1 + k1]);
    out[2 * cell + 1] = make_float4(v[i1 * s0 + j * s1 + k], v[i1 * s0 + j * s1 + k1], v[i1 * s0 + j1 * s1 + k],
                                    v[i1 * s0 + j1 * s1 + k1]);
  }
}
}  // namespace

hipError_t kgl_sdf_build8(const float* v, const int dims[3], float4* out, hipStream_t s) {
  const size_t ncell = (size_t)std::max(dims[0] - 1, 1) * std::max(dims[1] - 1, 1) * std::max(dims[2] - 1, 1);
  const unsigned blocks = (unsigned)std::min<size_t>((ncell + 255) / 256, 65536);
  k_sdf_build8<<<blocks, 256, 0, s>>>(v, dims[0], dims[1], dims[2], out);
  return hipGetLastError();
}

hipError_t kgl_sdf_penalty(const LinkDesc* links, int n, const SdfDev& g, const float* Q, int N, const float* pts,
                           const int* poff, float dsafe, float* loss, const float* dloss, float* dQ, bool bwd,
                           int lanes, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  // GL lanes per configuration (8, 16, 32 or 64; 32 is 5% faster than 16 on
  // the 7-link arm, 448 points per waypoint); groups per block limited by
  // 48 KB of LDS
  const int GLr = (lanes == 8 || lanes == 16 || lanes == 64) ? lanes : 32;
  const size_t per_group = (size_t)n * 18 * sizeof(float);
  int gpb = 256 / GLr;
  while (gpb > 1 && gpb * per_group > 48 * 1024) gpb /= 2;
  const size_t lds = gpb * per_group;
  int blocks = (N + gpb - 1) / gpb;
  // XCD-aware block order for large batches: the grid is padded to a
  // multiple of 8 (padding blocks find no waypoint)
  int nb8 = 0;
  if (blocks >= 64) {
    nb8 = (blocks + 7) / 8;
    blocks = 8 * nb8;
  }
  auto go = [&](auto glc) {
    constexpr int GL = decltype(glc)::value;
    if (bwd)
      k_sdf_pen_group<true, GL><<<blocks, gpb * GL, lds, s>>>(links, n, g, Q, N, pts, poff, dsafe, loss, dloss, dQ, gpb,
                                                             nb8);
    else
      k_sdf_pen_group<false, GL><<<blocks, gpb * GL, lds, s>>>(links, n, g, Q, N, pts, poff, dsafe, loss, dloss, dQ,
                                                              gpb, nb8);
  };
  switch (GLr) {
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 64: go(std::integral_constant<int, 64>{}); break;
    case 16: go(std::integral_constant<int, 16>{}); break;
    default: go(std::integral_constant<int, 32>{}); break;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------ distance

#include "kg_geom_math.h"

namespace {

// Lower bound on the distance between box A (frame A, mapped by rel into
// frame B) and box B: the larger of the bounding-sphere gap and the largest
// normalised projection gap over the 15 separating axes (projections onto a
// unit axis are 1-Lipschitz).
template <typename Real>
__device__ inline Real obb_lower_bound(const NodeF& A, const NodeF& B, const Real* rel) {
  Real Ap[9], cp[3];
#pragma unroll
  for (int col = 0; col < 3; col++)
#pragma unroll
    for (int r = 0; r < 3; r++)
      Ap[col * 3 + r] = rel[r] * A.R[col * 3] + rel[3 + r] * A.R[col * 3 + 1] + rel[6 + r] * A.R[col * 3 + 2];
#pragma unroll
  for (int r = 0; r < 3; r++) cp[r] = rel[r] * A.c[0] + rel[3 + r] * A.c[1] + rel[6 + r] * A.c[2] + rel[9 + r] - B.c[r];
  Real C[3][3], Qa[3][3], T[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    T[i] = B.R[i * 3] * cp[0] + B.R[i * 3 + 1] * cp[1] + B.R[i * 3 + 2] * cp[2];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C[i][j] = B.R[i * 3] * Ap[j * 3] + B.R[i * 3 + 1] * Ap[j * 3 + 1] + B.R[i * 3 + 2] * Ap[j * 3 + 2];
      Qa[i][j] = fabs(C[i][j]);
    }
  }
  const Real eA0 = A.e[0], eA1 = A.e[1], eA2 = A.e[2], eB0 = B.e[0], eB1 = B.e[1], eB2 = B.e[2];
  const Real eA[3] = {eA0, eA1, eA2}, eB[3] = {eB0, eB1, eB2};
  // the axis through the two centres: |T| - (sum_i eB_i |T_i| + sum_j eA_j
  // |T . a_j|) / |T|, at least the bounding-sphere bound |T| - |eA| - |eB|
  // (Cauchy-Schwarz), and the best cheap axis for diagonally offset boxes
  const Real tt = T[0] * T[0] + T[1] * T[1] + T[2] * T[2];
  Real rc = eB0 * fabs(T[0]) + eB1 * fabs(T[1]) + eB2 * fabs(T[2]);
  Real lb = (Real)0;
#pragma unroll
  for (int i = 0; i < 3; i++) lb = fmax(lb, fabs(T[i]) - (eB[i] + eA[0] * Qa[i][0] + eA[1] * Qa[i][1] + eA[2] * Qa[i][2]));
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const Real t = T[0] * C[0][j] + T[1] * C[1][j] + T[2] * C[2][j];
    rc += eA[j] * fabs(t);
    lb = fmax(lb, fabs(t) - (eB[0] * Qa[0][j] + eB[1] * Qa[1][j] + eB[2] * Qa[2][j] + eA[j]));
  }
  if (tt > (Real)0) lb = fmax(lb, (tt - rc) * rsqrt(tt));
  // the axis from B's point nearest A's centre to that centre: the gap along
  // it is |d| minus A's half-width on it (B's support there is its nearest
  // point); and the same from A's side (B's centre in A's frame is -C^T T)
  {
    Real d[3], dd = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      d[i] = T[i] - fmin(fmax(T[i], -eB[i]), eB[i]);
      dd += d[i] * d[i];
    }
    if (dd > (Real)0) {
      Real ra = 0;
#pragma unroll
      for (int j = 0; j < 3; j++) ra += eA[j] * fabs(d[0] * C[0][j] + d[1] * C[1][j] + d[2] * C[2][j]);
      lb = fmax(lb, (dd - ra) * rsqrt(dd));
    }
  }
  {
    Real u[3], e[3], ee = 0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      u[j] = -(T[0] * C[0][j] + T[1] * C[1][j] + T[2] * C[2][j]);
      e[j] = u[j] - fmin(fmax(u[j], -eA[j]), eA[j]);
      ee += e[j] * e[j];
    }
    if (ee > (Real)0) {
      Real rb = 0;
#pragma unroll
      for (int i = 0; i < 3; i++) rb += eB[i] * fabs(e[0] * C[i][0] + e[1] * C[i][1] + e[2] * C[i][2]);
      lb = fmax(lb, (ee - rb) * rsqrt(ee));
    }
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      const Real l2 = (Real)1 - C[i][j] * C[i][j];
      if (l2 < (Real)1e-6) continue;
      const Real t = T[i2] * C[i1][j] - T[i1] * C[i2][j];
      const Real ra = eB[i1] * Qa[i2][j] + eB[i2] * Qa[i1][j];
      const Real rb = eA[j1] * Qa[i][j2] + eA[j2] * Qa[i][j1];
      lb = fmax(lb, (fabs(t) - ra - rb) * rsqrt(l2));
    }
  }
  return fmax(lb, (Real)0);
}

// Minimum distance per (configuration, pair) item; same pool scheme as
// k_collide (depth-first continuation per lane, shared LIFO pool, batched
// leaves).  A child pair is dropped when its lower bound exceeds the item's
// best distance by more than the FP32 error pad; the continuation is the
// child pair with the smaller bound.
__global__ __launch_bounds__(64) void k_distance(DistArgs a) {
  __shared__ uint64_t stk[KG_SCAP];
  __shared__ uint64_t lq[KG_LQ];
  __shared__ float s_rel[12][64];
  __shared__ float s_pad[64];
  __shared__ unsigned s_best[64], s_droplb[64];
  __shared__ int s_item[64], s_pend[64], s_ta[64], s_tb[64];
  __shared__ int s_chunk;

  const int lane = threadIdx.x;
  uint64_t* gsp_base = a.spill + (size_t)blockIdx.x * a.spill_cap;
  const int nitems = a.N * a.npairs;
  const unsigned stop = __float_as_uint(a.stop_below);
  const float rel1 = 1.0f + a.rel_err, abs_err = a.abs_err;
  unsigned n_box = 0, n_tri = 0;

  // Items stream through the wave: whenever lanes are free and a slot is
  // open a new item starts, from the current range of ipw items; a range is
  // fetched from the global counter only when the previous one is spent, so
  // no lane waits for the slowest item of a range.
  s_pend[lane] = 0;
  __syncthreads();
  int it = 0, i1 = 0, sp = 0, gsp = 0, nl = 0;
  bool exhausted = false;
  uint64_t cont = KG_NOTASK;
  {
    auto finish = [&](int slot, int delta) {
      if (atomicAdd(&s_pend[slot], delta) + delta == 0) {
        const int item = s_item[slot];
        const unsigned best = s_best[slot], dlb = s_droplb[slot];
        a.out[item] = __uint_as_float(min(best, dlb));
        if (dlb < best && best > stop) atomicAdd(a.uncertain, 1);
        if (a.out_tri) {
          a.out_tri[2 * item] = s_ta[slot];
          a.out_tri[2 * item + 1] = s_tb[slot];
        }
      }
    };
    auto ensure_room = [&](int room) {
      if (sp + room <= KG_SCAP) return;
      const int h = KG_SCAP / 2;
      if (gsp + h <= a.spill_cap) {
        for (int i = lane; i < h; i += 64) gsp_base[gsp + i] = stk[i];
        gsp += h;
        __syncthreads();
      } else {
        // overflow: drop the bottom half, keeping for each item the smallest
        // certified distance bound of its dropped tasks (box bound minus the
        // item's FP32 padding, at least 0)
        for (int i = lane; i < h; i += 64) {
          const uint64_t t = stk[i];
          const int sl = (int)(t >> 58);
          const int side = (int)((t >> 57) & 1), single = (int)((t >> 56) & 1);
          const int x = (int)((t >> 28) & 0xfffffff), y = (int)(t & 0xfffffff);
          float rel[12];
#pragma unroll
          for (int k = 0; k < 12; k++) rel[k] = s_rel[k][sl];
          const NodeF O = load_nodef(a.pool.nodes, y);
          const bool aIsS = single || side == 0;
          float lbm = 3e38f;
          for (int k = 0; k < (single ? 1 : 2); k++) {
            const NodeF S = load_nodef(a.pool.nodes, x + k);
            lbm = fminf(lbm, obb_lower_bound<float>(sel(aIsS, S, O), sel(aIsS, O, S), rel) - s_pad[sl]);
          }
          atomicMin(&s_droplb[sl], __float_as_uint(fmaxf(lbm, 0.f)));
        }
        if (lane == 0) atomicExch(a.status, 1);
        __syncthreads();
        for (int i = lane; i < h; i += 64) finish((int)(stk[i] >> 58), -1);
        __syncthreads();
      }
      for (int i = lane; i < sp - h; i += 64) stk[i] = stk[i + h];
      sp -= h;
      __syncthreads();
    };
    auto run_leaves = [&]() {
      const int m = min(nl, 64);
      float dl = 3e38f;
      int slot = 0, ta = 0, tb = 0;
      if (lane < m) {
        const uint64_t e = lq[nl - 1 - lane];
        slot = (int)(e >> 58);
        ta = (int)((e >> 29) & 0x1fffffff);
        tb = (int)(e & 0x1fffffff);
        if (s_best[slot] > stop) {
          n_tri++;
          float rel[12];
#pragma unroll
          for (int k = 0; k < 12; k++) rel[k] = s_rel[k][slot];
          V3T<float> tp[3], tq[3];
          load_tri<float>(a.pool, ta, tp);
          load_tri<float>(a.pool, tb, tq);
          kgm::V<float> P[3], Q[3], cp, cq;
#pragma unroll
          for (int v = 0; v < 3; v++) {
            const V3T<float> w = xf_apply<float>(rel, tp[v]);
            P[v] = kgm::mk(w.x, w.y, w.z);
            Q[v] = kgm::mk(tq[v].x, tq[v].y, tq[v].z);
          }
          dl = kgm::tri_dist<float>(P, Q, cp, cq);
          atomicMin(&s_best[slot], __float_as_uint(dl));
        }
      }
      __syncthreads();
      if (lane < m) {
        if (__float_as_uint(dl) == s_best[slot]) {
          s_ta[slot] = ta;
          s_tb[slot] = tb;
        }
      }
      __syncthreads();
      if (lane < m) finish(slot, -1);
      nl -= m;
      __syncthreads();
    };

    for (;;) {
      const bool hungry = cont == KG_NOTASK;
      const uint64_t hm = ballot64(hungry);
      const int nh = __popcll(hm);
      if (nh > sp && gsp > 0) {
        const int k = min(gsp, KG_SCAP / 2);
        for (int i = sp - 1 - lane; i >= 0; i -= 64) stk[i + k] = stk[i];
        __syncthreads();
        for (int i = lane; i < k; i += 64) stk[i] = gsp_base[gsp - k + i];
        gsp -= k;
        sp += k;
        __syncthreads();
      }
      const int hp = lane_prefix(hm);
      const int npop = min(nh, sp);
      if (hungry && hp < npop) cont = stk[sp - 1 - hp];
      sp -= npop;
      // new items on free slots, from the current range (the next range is
      // fetched once this one is spent)
      const bool want = cont == KG_NOTASK;
      const uint64_t wm = ballot64(want);
      const uint64_t fm = ballot64(s_pend[lane] == 0);
      const int nwant = min(__popcll(wm), __popcll(fm));
      if (nwant > 0 && it == i1 && !exhausted) {
        if (lane == 0) s_chunk = atomicAdd(a.chunk_counter, 1);
        __syncthreads();
        const int i0 = __builtin_amdgcn_readfirstlane(s_chunk) * a.ipw;
        __syncthreads();
        if (i0 >= nitems) {
          exhausted = true;
        } else {
          it = i0;
          i1 = min(i0 + a.ipw, nitems);
        }
      }
      const int navail = min(nwant, i1 - it);
      int myitem = -1, myslot = -1;
      const int wp = lane_prefix(wm);
      if (want && wp < navail) {
        myitem = it + wp;
        uint64_t f = fm;
        for (int q = 0; q < wp; q++) f &= f - 1;
        myslot = __ffsll((unsigned long long)f) - 1;
      }
      it += navail;
      __syncthreads();
      if (myitem >= 0) {
        const int c = myitem / a.npairs, pidx = myitem % a.npairs;
        const int la = a.pair_a[pidx], lb = a.pair_b[pidx], kind = a.pair_kind[pidx];
        const double* XA = a.T64 + ((size_t)c * a.plan.nlinks + la) * 12;
        const double* XB;
        int geomB;
        if (kind == 0) {
          XB = a.plan.env[lb].X;
          geomB = a.plan.env[lb].geom;
        } else {
          XB = a.T64 + ((size_t)c * a.plan.nlinks + lb) * 12;
          geomB = a.plan.links[lb].geom;
        }
        const double d0 = XA[9] - XB[9], d1 = XA[10] - XB[10], d2 = XA[11] - XB[11];
#pragma unroll
        for (int col = 0; col < 3; col++)
#pragma unroll
          for (int r = 0; r < 3; r++)
            s_rel[col * 3 + r][myslot] =
                (float)(XB[r * 3] * XA[col * 3] + XB[r * 3 + 1] * XA[col * 3 + 1] + XB[r * 3 + 2] * XA[col * 3 + 2]);
        double tn = 0;
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const double v = XB[r * 3] * d0 + XB[r * 3 + 1] * d1 + XB[r * 3 + 2] * d2;
          s_rel[9 + r][myslot] = (float)v;
          tn += v * v;
        }
        const GeomDesc& GA = a.pool.geoms[a.plan.links[la].geom];
        const GeomDesc& GB = a.pool.geoms[geomB];
        const double u = 5.9604644775390625e-08;
        const double ev = u * (8.0 * GA.rmax + 4.0 * sqrt(tn)) + u * GB.rmax;
        s_pad[myslot] = (float)(2.0 * ev + 16.0 * u * (sqrt(tn) + GA.rmax + GB.rmax));
        s_best[myslot] = __float_as_uint(a.upper);
        s_droplb[myslot] = 0x7f800000u;  // +inf: nothing dropped
        s_item[myslot] = myitem;
        s_ta[myslot] = -1;
        s_tb[myslot] = -1;
        s_pend[myslot] = 1;
        cont = mk_task(myslot, 0, 1, GA.root, GB.root);
      }
      __syncthreads();

      const bool working = cont != KG_NOTASK;
      if (!ballot64(working)) {
        if (sp == 0 && gsp == 0) {
          if (nl == 0) {
            if (exhausted && it == i1) break;
            continue;
          }
          run_leaves();
        }
        continue;
      }

      uint64_t keep = KG_NOTASK, push = KG_NOTASK, lf0 = 0, lf1 = 0;
      int nlf = 0, nout = 0, slot = 0;
      if (working) {
        const uint64_t task = cont;
        cont = KG_NOTASK;
        slot = (int)(task >> 58);
        const int side = (int)((task >> 57) & 1), single = (int)((task >> 56) & 1);
        const int x = (int)((task >> 28) & 0xfffffff), y = (int)(task & 0xfffffff);
        const float best = __uint_as_float(s_best[slot]);
        if (s_best[slot] > stop) {
          const NodeF O = load_nodef(a.pool.nodes, y);
          const NodeF S0 = load_nodef(a.pool.nodes, x);
          const NodeF S1 = load_nodef(a.pool.nodes, single ? x : x + 1);
          float rel[12];
#pragma unroll
          for (int k = 0; k < 12; k++) rel[k] = s_rel[k][slot];
          const float pad = s_pad[slot];
          const bool aIsS = single || side == 0;
          float lbk[2] = {3e38f, 3e38f};
          uint64_t tk[2] = {KG_NOTASK, KG_NOTASK};
#pragma unroll 1
          for (int k = 0; k < (single ? 1 : 2); k++) {
            const NodeF S = sel(k == 0, S0, S1);
            const NodeF A = sel(aIsS, S, O);
            const NodeF B = sel(aIsS, O, S);
            n_box++;
            const float lb = obb_lower_bound<float>(A, B, rel);
            // no closer pair below (certified bound lb - pad), or none closer
            // by more than the tolerances (DistanceQuerySettings relErr /
            // absErr: the result is then within them of the exact distance)
            const float lbc = lb - pad;
            if (lbc > best || lbc * rel1 > best || lbc + abs_err > best) continue;
            const bool la = A.ch < 0, lbf = B.ch < 0;
            if (la && lbf) {
              const uint64_t w = mk_leaf(slot, ~A.ch, ~B.ch);
              if (nlf == 0) lf0 = w; else lf1 = w;
              nlf++;
            } else {
              bool splitA;
              if (la) splitA = false;
              else if (lbf) splitA = true;
              else splitA = (A.e[0] + A.e[1] + A.e[2]) >= (B.e[0] + B.e[1] + B.e[2]);
              const int nodeS = x + k;
              const int nodeA = aIsS ? nodeS : y;
              const int nodeB = aIsS ? y : nodeS;
              const uint64_t w = splitA ? mk_task(slot, 0, 0, A.ch, nodeB) : mk_task(slot, 1, 0, B.ch, nodeA);
              if (k == 0) { tk[0] = w; lbk[0] = lb; } else { tk[1] = w; lbk[1] = lb; }
              nout++;
            }
          }
          // continue with the nearer child pair, pool the other
          const bool swap = lbk[1] < lbk[0];
          keep = swap ? tk[1] : tk[0];
          push = swap ? tk[0] : tk[1];
          if (keep == KG_NOTASK) {
            keep = push;
            push = KG_NOTASK;
          }
        }
      }
      cont = keep;
      {
        const uint64_t l1 = ballot64(nlf >= 1), l2 = ballot64(nlf >= 2);
        const int ltot = __popcll(l1) + __popcll(l2);
        if (nl + ltot > KG_LQ) {
          __syncthreads();
          while (nl + ltot > KG_LQ) run_leaves();
        }
        const int pos = nl + lane_prefix(l1) + lane_prefix(l2);
        if (nlf >= 1) lq[pos] = lf0;
        if (nlf >= 2) lq[pos + 1] = lf1;
        nl += ltot;
      }
      if (working) finish(slot, nout + nlf - 1);
      const bool haspush = push != KG_NOTASK;
      const uint64_t pm = ballot64(haspush);
      const int ptot = __popcll(pm);
      ensure_room(ptot);
      if (haspush) stk[sp + lane_prefix(pm)] = push;
      sp += ptot;
      __syncthreads();
      // triangle distances only once a full wave of leaf pairs is queued
      // (a step for one pair leaves 63 lanes idle for its whole length; the
      // drain above runs what is left): best distances tighten later, which
      // costs 2% more box tests against a third of the kernel's time
      if (nl >= 64) run_leaves();
    }
    __syncthreads();
  }
  if (a.counters && (n_box | n_tri)) {
    unsigned long long* cnt = a.counters + (blockIdx.x % KG_CNT_COPIES) * (KG_COUNTER_STRIDE / 2);
    atomicAdd(&cnt[0], (unsigned long long)n_box);
    atomicAdd(&cnt[1], (unsigned long long)n_tri);
  }
}

}  // namespace

hipError_t kgl_distance(const DistArgs& a, hipStream_t s) {
  if (a.N <= 0 || a.npairs <= 0) return hipSuccess;
  k_distance<<<a.grid, 64, 0, s>>>(a);
  return hipGetLastError();
}

int kgl_distance_grid(int num_cus) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_distance, 64, 0) != hipSuccess || nb < 1) nb = 8;
  return nb * num_cus;
}

namespace {
__global__ __launch_bounds__(256) void k_fk64(const LinkDesc* __restrict__ links, int n, const float* __restrict__ Qf,
                                             const double* __restrict__ Qd, int N, double* __restrict__ T64,
                                             int* reset, int nreset) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nreset) reset[c] = 0;
  if (c >= N) return;
  double* T = T64 + (size_t)c * n * 12;
  X64 Xprev;  // frame of link l-1, kept in registers for serial chains
  for (int l = 0; l < n; l++) {
    const double q = Qf ? (double)Qf[(size_t)c * n + l] : Qd[(size_t)c * n + l];
    const LinkDesc& L = links[l];
    X64 J, T0, Lx, X;
    joint64(L.jtype, L.axis, q, J);
    x64_load(L.T0, T0);
    x64_mul(T0, J, Lx);
    if (L.parent >= 0) {
      X64 Pp;
      if (L.parent == l - 1) Pp = Xprev;
      else x64_load(T + 12 * L.parent, Pp);
      x64_mul(Pp, Lx, X);
    } else {
      X = Lx;
    }
    Xprev = X;
    x64_store(T + 12 * l, X);
  }
}
}  // namespace

hipError_t kgl_fk64(const LinkDesc* links, int n, const float* Qf, const double* Qd, int N, double* T64, int* reset,
                    int nreset, hipStream_t s) {
  const int m = N > nreset ? N : nreset;
  if (m <= 0) return hipSuccess;
  k_fk64<<<(m + 255) / 256, 256, 0, s>>>(links, n, Qf, Qd, N, T64, reset, nreset);
  return hipGetLastError();
}

// ------------------------------------------------------------ edge checks

namespace {

// Sample configurations for one round: slot j belongs to live segment
// seg[s] (s found by binary search in the per-segment slot offsets) at
// bisection position pos0 + (j - soff[s]); q = a*(1-u) + b*u with u = i/n
// (Cpp/Modeling/Interpolate.cpp:13-14), i the sample index at that position.
// ---------------------------------------------------------------- edge checks
//
// Straight-line edge checks run as rounds on the device, with no host
// synchronization between them.  Each segment s has n_s steps and a cursor
// (bisection positions checked so far).  A round gives every live segment
// its next min(width, positions left) positions (width: the scheduled width,
// lowered so that the round fits the sample budget), lays the samples out
// contiguously (segment order), checks them with the feasibility pipeline,
// and then each segment either records its first infeasible position of the
// round (the first in bisection order overall, since the earlier rounds held
// none) or advances its cursor.

// Endpoint coordinate k of segment s (side 0: start a, 1: end b).
__device__ inline double edge_q(const EdgeArgs& a, int side, int s, int k) {
  const size_t i = (size_t)s * a.n + k;
  if (a.Ad) return side ? a.Bd[i] : a.Ad[i];
  return (double)(side ? a.Bf[i] : a.Af[i]);
}

// Segment steps n_s = ceil(d / eps) with d the L2 or L-infinity norm of
// b - a, in FP64 with separate roundings (the same expression as the host
// and the oracle).
__global__ __launch_bounds__(256) void k_edge_init(EdgeArgs a, int clear_ctl) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (clear_ctl && s < 8) a.ctl[s] = 0;
  if (s >= a.nseg) return;
  double acc = 0;
  for (int k = 0; k < a.n; k++) {
#pragma clang fp contract(off)
    const double d = edge_q(a, 0, s, k) - edge_q(a, 1, s, k);
    if (a.norm_inf)
      acc = fmax(acc, fabs(d));
    else
      acc += d * d;
  }
  const double dist = a.norm_inf ? acc : sqrt(acc);
  const double m = ceil(dist / a.eps);
  int ns;
  bool bad = false;
  if (!(m >= 0)) {  // NaN input
    ns = 0;
    bad = true;
  } else {
    ns = m > (double)(1 << 30) ? (1 << 30) : (int)m;
  }
  a.ns[s] = ns;
  a.cur[s] = 0;
  a.done[s] = (ns < 2 || bad) ? 1 : 0;
  a.first_u[s] = bad ? -2.0f : -1.0f;
  if (bad) atomicAdd(&a.ctl[6], 1);
  if (a.nsamples_out) a.nsamples_out[s] = ns;
}

// Round plan (one block): live segments, the round's width, each live
// segment's sample offset.  Thread t takes segments t, t + TB, ... (coalesced
// loads); the samples are laid out thread by thread (an exclusive scan over
// the threads' counts), which is all the later kernels need: the records are
// in increasing offset order.  The positions each segment has left are read
// from HBM once (loads of several segments in flight per thread) and kept in
// LDS for the first KG_EDGE_PLAN_LDS segments: the three passes over the
// segments otherwise wait on a chain of dependent loads per thread.
#define KG_EDGE_PLAN_TB 1024
#define KG_EDGE_PLAN_LDS 8192
__global__ __launch_bounds__(KG_EDGE_PLAN_TB) void k_edge_plan(EdgeArgs a, int width) {
  using Scan = hipcub::BlockScan<int, KG_EDGE_PLAN_TB>;
  __shared__ typename Scan::TempStorage tmp;
  __shared__ int s_nlive;
  __shared__ int s_left[KG_EDGE_PLAN_LDS];
  const int t = threadIdx.x;
  auto left_g = [&](int s) {  // positions of s this schedule may still check
    if (a.done[s]) return 0;
    const int m = min(a.ns[s] - 1, a.max_pos);
    return max(0, m - a.cur[s]);
  };
  auto left = [&](int s) { return s < KG_EDGE_PLAN_LDS ? s_left[s] : left_g(s); };
  int nl = 0;
#pragma unroll 8
  for (int s = t; s < a.nseg; s += KG_EDGE_PLAN_TB) {
    const int L = left_g(s);
    if (s < KG_EDGE_PLAN_LDS) s_left[s] = L;
    nl += L > 0;
  }
  int nlive_pre, nlive;
  Scan(tmp).ExclusiveSum(nl, nlive_pre, nlive);
  if (t == 0) s_nlive = nlive;
  __syncthreads();
  nlive = s_nlive;
  const int w = nlive > 0 ? max(1, min(width, a.budget / nlive)) : 0;
  int ns_loc = 0;
  for (int s = t; s < a.nseg; s += KG_EDGE_PLAN_TB) ns_loc += min(w, left(s));
  int off, total;
  __syncthreads();
  Scan(tmp).ExclusiveSum(ns_loc, off, total);
  int li = nlive_pre;
  for (int s = t; s < a.nseg; s += KG_EDGE_PLAN_TB) {
    const int k = min(w, left(s));
    if (k > 0) {
      a.seg[li] = s;
      a.soff[li] = off;
      li++;
      off += k;
    }
  }
  if (t == 0) {
    a.soff[nlive] = total;
    a.ctl[0] = nlive;
    a.ctl[1] = total;
  }
}

// Sample j of a wide round: segment by binary search over the offsets, its
// position cursor + (j - offset) in bisection order, q = a (1-u) + b u.
__global__ __launch_bounds__(256) void k_edge_samples_flat(EdgeArgs a) {
  const int nlive = a.ctl[0], total = a.ctl[1];
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x) {
    int lo = 0, hi = nlive - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.soff[mid] <= j) lo = mid; else hi = mid - 1;
    }
    const int s = a.seg[lo];
    const int ns = a.ns[s];
    const int i = kg_bisect_index(ns, a.cur[s] + (j - a.soff[lo]));
    const double u = (double)i / (double)ns;
    for (int d = 0; d < a.n; d++) {
#pragma clang fp contract(off)
      double q = edge_q(a, 0, s, d) * (1.0 - u);
      q += edge_q(a, 1, s, d) * u;
      a.Qd[(size_t)j * a.n + d] = q;
    }
  }
}

// The same samples for a narrow round (at most 64 positions per segment),
// one wave per live segment: no search, the segment's offset read directly.
__global__ __launch_bounds__(256) void k_edge_samples(EdgeArgs a) {
  const int nlive = a.ctl[0];
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int li = wave; li < nlive; li += nwaves) {
    const int s = a.seg[li];
    const int o = a.soff[li], k = a.soff[li + 1] - o;
    const int ns = a.ns[s], c0 = a.cur[s];
    for (int j = lane; j < k; j += 64) {
      const int i = kg_bisect_index(ns, c0 + j);
      const double u = (double)i / (double)ns;
      for (int d = 0; d < a.n; d++) {
#pragma clang fp contract(off)
        double q = edge_q(a, 0, s, d) * (1.0 - u);
        q += edge_q(a, 1, s, d) * u;
        a.Qd[(size_t)(o + j) * a.n + d] = q;
      }
    }
  }
}

// First infeasible sample of each live segment in the round, else the
// cursor advances; the round's status words are accumulated.  One wave per
// live segment: its slots 64 at a time, the first infeasible by ballot (the
// loop's bounds and exit are wave-uniform).
__global__ __launch_bounds__(256) void k_edge_reduce(EdgeArgs a) {
  const int nlive = a.ctl[0];
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctl[2] |= a.st[a.w_spill];
    a.ctl[3] += a.st[a.w_unc];
    a.ctl[4] += a.st[a.w_redo];
  }
  for (int t = wave; t < nlive; t += nwaves) {
    const int s = a.seg[t];
    const int b = a.soff[t], e = a.soff[t + 1];
    int first = -1;
    for (int j0 = b; j0 < e; j0 += 64) {
      const int j = j0 + lane;
      const uint64_t m = __ballot(j < e && !a.feas[j]);
      if (m) {
        first = j0 + __ffsll((unsigned long long)m) - 1;
        break;
      }
    }
    if (lane == 0) {
      if (first >= 0) {
        const int ns = a.ns[s];
        const int i = kg_bisect_index(ns, a.cur[s] + (first - b));
        a.first_u[s] = (float)((double)i / (double)ns);
        a.done[s] = 1;
      } else {
        a.cur[s] += e - b;
      }
    }
  }
}

// Segments the schedule did not finish (more positions than it covers) are
// reported undecided (-2) and counted.
__global__ __launch_bounds__(256) void k_edge_finish(EdgeArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nseg) return;
  if (!a.done[s] && a.cur[s] < a.ns[s] - 1) {
    a.first_u[s] = -2.0f;
    atomicAdd(&a.ctl[5], 1);
  }
}

// The call's status words, in the layout of a feasibility batch's.
__global__ void k_edge_status(EdgeArgs a) {
  if (threadIdx.x == 0) {
    a.out_status[a.w_spill] = a.ctl[2];
    a.out_status[a.w_unc] = a.ctl[3] + a.ctl[5] + a.ctl[6];
    a.out_status[a.w_redo] = a.ctl[4];
  }
}

}  // namespace

hipError_t kgl_edge_init(const EdgeArgs& a, bool clear_ctl, hipStream_t s) {
  k_edge_init<<<(max(a.nseg, 8) + 255) / 256, 256, 0, s>>>(a, clear_ctl ? 1 : 0);
  return hipGetLastError();
}

hipError_t kgl_edge_round(const EdgeArgs& a, int width, int phase, hipStream_t s) {
  // the reduction and a narrow round's samples: one wave per live segment
  // (4 per block); a wide round's samples: one thread each; grid-stride
  const int blocks = max(1, min(2048, (a.nseg + 3) / 4));
  if (phase == 0) {
    k_edge_plan<<<1, KG_EDGE_PLAN_TB, 0, s>>>(a, width);
    if (width <= 64)
      k_edge_samples<<<blocks, 256, 0, s>>>(a);
    else
      k_edge_samples_flat<<<max(1, min(1024, (a.budget + 255) / 256)), 256, 0, s>>>(a);
  } else {
    k_edge_reduce<<<blocks, 256, 0, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t kgl_edge_finish(const EdgeArgs& a, bool write_status, hipStream_t s) {
  k_edge_finish<<<(max(a.nseg, 1) + 255) / 256, 256, 0, s>>>(a);
  if (write_status) k_edge_status<<<1, 64, 0, s>>>(a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- mesh -> SDF
//
// TriangleMesh -> ImplicitSurface on a cell-centred grid (cell (i,j,k) centred
// at bmin + ((i,j,k) + 0.5) h):
//
// k_msdf_dist: one lane per cell, one wave per 4x4x4 brick of cells.  The
// wave runs one depth-first nearest-triangle search for its 64 cells: a node
// is expanded while, for some lane, the distance from that lane's cell centre
// to the node's box is below the lane's best triangle distance so far (so
// every lane still sees every node its own search needs); the child whose
// centre is nearer the brick's centre is visited first.  The traversal stack
// holds one entry per lane (64 deep).  FP32; the unsigned distance is
// written to out.
//
// k_msdf_sign: one lane per row (i, j) of cells along z, each lane with its
// own traversal (a line's boxes are not shared enough between neighbouring
// rows for a wave-wide walk to pay: that variant took 4x longer).  The +z
// line through the row's centres is tested against the tree's boxes (FP32,
// padded: a conservative cull) and against the triangles in FP64 with the
// crossing rule below; each crossing adds its winding increment at the first
// cell above it, and the lane negates the cells of its row whose running
// winding number is nonzero.
//
// Crossing rule (the same expressions as the oracle, no contraction, so the
// winding numbers are identical): a triangle whose xy projection has signed
// area s != 0 is ordered counter-clockwise; it is crossed when every edge
// function w(u,v,q) = (u.x-q.x)(v.y-q.y) - (u.y-q.y)(v.x-q.x) is > 0, or = 0
// on an edge u->v with u before v in the order (y descending, x ascending);
// the increment is -1 for s > 0 (normal towards +z), +1 for s < 0.

namespace {

__device__ inline void node_load(const ObbNode* nodes, int n, float c[3], float e[3], float R[9], int& child) {
  const float4* q = reinterpret_cast<const float4*>(nodes + n);
  const float4 a = q[0], b = q[1], d = q[2], f = q[3];
  c[0] = a.x; c[1] = a.y; c[2] = a.z; e[0] = a.w;
  e[1] = b.x; e[2] = b.y; R[0] = b.z; R[1] = b.w;
  R[2] = d.x; R[3] = d.y; R[4] = d.z; R[5] = d.w;
  R[6] = f.x; R[7] = f.y; R[8] = f.z;
  child = __float_as_int(f.w);
}

__global__ __launch_bounds__(64) void k_msdf_dist(MeshSdfArgs a, const ObbNode* __restrict__ nodes,
                                                  const float4* __restrict__ tv0, const float4* __restrict__ tv1,
                                                  const float4* __restrict__ tv2, float* __restrict__ out) {
  const int lane = threadIdx.x;
  const int nx = a.dims[0], ny = a.dims[1], nz = a.dims[2];
  const long bx = (nx + 3) >> 2, by = (ny + 3) >> 2, bz = (nz + 3) >> 2;
  const long nb = bx * by * bz;
  for (long b = blockIdx.x; b < nb; b += gridDim.x) {
    const int bk = (int)(b % bz);
    const long r = b / bz;
    const int bj = (int)(r % by), bi = (int)(r / by);
    const int i = bi * 4 + (lane >> 4), j = bj * 4 + ((lane >> 2) & 3), k = bk * 4 + (lane & 3);
    const bool act = i < nx && j < ny && k < nz;
    const kgm::V<float> p = kgm::mk<float>((float)(a.bmin[0] + ((double)min(i, nx - 1) + 0.5) * a.h),
                                           (float)(a.bmin[1] + ((double)min(j, ny - 1) + 0.5) * a.h),
                                           (float)(a.bmin[2] + ((double)min(k, nz - 1) + 0.5) * a.h));
    const float qx = (float)(a.bmin[0] + (bi * 4 + 2.0) * a.h), qy = (float)(a.bmin[1] + (bj * 4 + 2.0) * a.h),
                qz = (float)(a.bmin[2] + (bk * 4 + 2.0) * a.h);
    float best = 3.0e38f;
    // the wave walks the tree together: a node is expanded when any lane's
    // box bound is below that lane's best; the stack lives one entry per
    // lane (entry s in lane s), its depth is wave-uniform
    int stk = 0, sp = 0, n = a.root;
    for (;;) {
      n = __builtin_amdgcn_readfirstlane(n);
      float c[3], e[3], R[9];
      int child;
      node_load(nodes, n, c, e, R, child);
      const float dx = p.x - c[0], dy = p.y - c[1], dz = p.z - c[2];
      float lb = 0.f;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const float s = R[3 * q] * dx + R[3 * q + 1] * dy + R[3 * q + 2] * dz;
        const float ex = fmaxf(fabsf(s) - e[q], 0.f);
        lb += ex * ex;
      }
      // the box bound carries FP32 rounding: prune only when clearly farther
      const bool need = act && lb <= best * 1.00001f;
      if (__ballot(need)) {
        if (child < 0) {
          const int t = ~child;
          const float4 A = tv0[t], B = tv1[t], C = tv2[t];
          const kgm::V<float> cp = kgm::closest_on_tri(p, kgm::mk(A.x, A.y, A.z), kgm::mk(B.x, B.y, B.z),
                                                       kgm::mk(C.x, C.y, C.z));
          const kgm::V<float> d = kgm::sub(p, cp);
          best = fminf(best, kgm::dot(d, d));
        } else {
          const float4 c0 = reinterpret_cast<const float4*>(nodes + child)[0];
          const float4 c1 = reinterpret_cast<const float4*>(nodes + child + 1)[0];
          const float d0 = (qx - c0.x) * (qx - c0.x) + (qy - c0.y) * (qy - c0.y) + (qz - c0.z) * (qz - c0.z);
          const float d1 = (qx - c1.x) * (qx - c1.x) + (qy - c1.y) * (qy - c1.y) + (qz - c1.z) * (qz - c1.z);
          const bool first0 = d0 <= d1;
          if (lane == sp) stk = first0 ? child + 1 : child;
          sp++;
          n = first0 ? child : child + 1;
          continue;
        }
      }
      if (sp == 0) break;
      sp = __builtin_amdgcn_readfirstlane(sp - 1);
      n = __builtin_amdgcn_readlane(stk, sp);
    }
    if (act) out[((size_t)i * ny + j) * nz + k] = sqrtf(best);
  }
}

__device__ inline double msdf_edge(double ux, double uy, double vx, double vy, double qx, double qy) {
#pragma clang fp contract(off)
  return (ux - qx) * (vy - qy) - (uy - qy) * (vx - qx);
}

__device__ inline bool msdf_inc(double w, double ux, double uy, double vx, double vy) {
  return w > 0 || (w == 0 && (uy > vy || (uy == vy && ux < vx)));
}

// Winding increment of the +z line through (qx, qy) at triangle t (0: not
// crossed) and the crossing height.
__device__ inline int msdf_crossing(const double* t, double qx, double qy, double& zt) {
#pragma clang fp contract(off)
  double ax = t[0], ay = t[1], az = t[2];
  double bx = t[3], by = t[4], bz = t[5];
  double cx = t[6], cy = t[7], cz = t[8];
  const double s = msdf_edge(ax, ay, bx, by, cx, cy);
  if (s == 0) return 0;
  const int w = s > 0 ? -1 : 1;
  if (s < 0) {
    double tx = bx, ty = by, tz = bz;
    bx = cx; by = cy; bz = cz;
    cx = tx; cy = ty; cz = tz;
  }
  const double wab = msdf_edge(ax, ay, bx, by, qx, qy);
  const double wbc = msdf_edge(bx, by, cx, cy, qx, qy);
  const double wca = msdf_edge(cx, cy, ax, ay, qx, qy);
  if (!msdf_inc(wab, ax, ay, bx, by) || !msdf_inc(wbc, bx, by, cx, cy) || !msdf_inc(wca, cx, cy, ax, ay)) return 0;
  const double den = (wab + wbc) + wca;
  if (den == 0) return 0;
  zt = ((wbc * az + wca * bz) + wab * cz) / den;
  return w;
}

__device__ inline double msdf_centre(double b, int k, double h) {
#pragma clang fp contract(off)
  return b + ((double)k + 0.5) * h;
}

// First k in [0, nz] with centre z_k > zt.
__device__ inline int msdf_first_above(double zt, double bz, double h, int nz) {
#pragma clang fp contract(off)
  double f = (zt - bz) / h - 0.5;
  if (!(f > -2.0)) f = -2.0;
  if (f > nz + 1.0) f = nz + 1.0;
  int k = (int)floor(f) + 1;
  k = k < 0 ? 0 : (k > nz ? nz : k);
  while (k > 0 && msdf_centre(bz, k - 1, h) > zt) k--;
  while (k < nz && !(msdf_centre(bz, k, h) > zt)) k++;
  return k;
}

__global__ __launch_bounds__(64) void k_msdf_sign(MeshSdfArgs a) {
  __shared__ int stk[KG_MSDF_STK][64];
  const int lane = threadIdx.x;
  const int ny = a.dims[1], nz = a.dims[2];
  const long rows = (long)a.dims[0] * ny;
  for (long row = (long)blockIdx.x * 64 + lane; row < rows; row += (long)gridDim.x * 64) {
    const int i = (int)(row / ny), j = (int)(row % ny);
    const double qx = msdf_centre(a.bmin[0], i, a.h), qy = msdf_centre(a.bmin[1], j, a.h);
    const float qxf = (float)qx, qyf = (float)qy;
    int* diff = a.diff + (size_t)row * nz;
    int sp = 0;
    stk[sp++][lane] = a.root;
    while (sp > 0) {
      const int n = stk[--sp][lane];
      float c[3], e[3], R[9];
      int child;
      node_load(a.pool.nodes, n, c, e, R, child);
      // does the line (qx, qy, t) meet the padded box?
      const float dx = qxf - c[0], dy = qyf - c[1], dz = -c[2];
      const float pad = 1e-5f * (1.f + fabsf(qxf) + fabsf(qyf) + fabsf(c[0]) + fabsf(c[1]) + fabsf(c[2]));
      float tlo = -3.0e38f, thi = 3.0e38f;
      bool hit = true;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const float s0 = R[3 * q] * dx + R[3 * q + 1] * dy + R[3 * q + 2] * dz;
        const float r = R[3 * q + 2], ek = e[q] + pad;
        if (fabsf(r) < 1e-20f) {
          hit = hit && fabsf(s0) <= ek;
        } else {
          float t1 = (-ek - s0) / r, t2 = (ek - s0) / r;
          if (t1 > t2) {
            const float tt = t1;
            t1 = t2;
            t2 = tt;
          }
          tlo = fmaxf(tlo, t1);
          thi = fminf(thi, t2);
        }
      }
      if (!hit || tlo > thi) continue;
      if (child < 0) {
        const int t = ~child;
        double zt = 0;
        const int w = msdf_crossing(a.pool.tv64 + 9 * (size_t)t, qx, qy, zt);
        if (w) {
          const int k0 = msdf_first_above(zt, a.bmin[2], a.h, nz);
          if (k0 < nz) diff[k0] += w;
        }
        continue;
      }
      stk[sp++][lane] = child;
      stk[sp++][lane] = child + 1;
    }
    float* out = a.out + (size_t)row * nz;
    int wind = 0;
    for (int k = 0; k < nz; k++) {
      wind += diff[k];
      if (wind != 0) out[k] = -out[k];
    }
  }
}

}  // namespace

hipError_t kgl_msdf_dist(const MeshSdfArgs& a, hipStream_t s) {
  const long nb = (long)((a.dims[0] + 3) >> 2) * ((a.dims[1] + 3) >> 2) * ((a.dims[2] + 3) >> 2);
  if (nb <= 0) return hipSuccess;
  const int blocks = (int)(nb < (1l << 20) ? nb : (1l << 20));
  k_msdf_dist<<<blocks, 64, 0, s>>>(a, a.pool.nodes, a.pool.tv0, a.pool.tv1, a.pool.tv2, a.out);
  return hipGetLastError();
}

hipError_t kgl_msdf_sign(const MeshSdfArgs& a, hipStream_t s) {
  const long rows = (long)a.dims[0] * a.dims[1];
  if (rows <= 0) return hipSuccess;
  const long nb = (rows + 63) / 64;
  const int blocks = (int)(nb < (1l << 20) ? nb : (1l << 20));
  k_msdf_sign<<<blocks, 64, 0, s>>>(a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- point sets vs implicit surfaces
//
// PointCloud / mesh-vertex sets against an ImplicitSurface: every point is
// moved to world coordinates by the set's transform and evaluated in the
// surface's grid (trilinear, as k_sdf_eval); the minimum value and its point
// index are reduced over the wave, the block and the grid (one 64-bit
// atomicMin per block on an order-preserving key).

namespace {

struct PtsXf {
  float v[12];  // column-major R, then t
};

__device__ inline unsigned long long pts_key(float v, int i) {
  unsigned u = __float_as_uint(v);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)u << 32) | (unsigned)i;
}

__global__ __launch_bounds__(256) void k_pts_sdf_min(const float4* __restrict__ P, int n, PtsXf X, SdfDev g,
                                                     unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_min[4];
  unsigned long long best = ~0ull;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float4 p = P[i];
    const float* R = X.v;
    const float wx = R[0] * p.x + R[3] * p.y + R[6] * p.z + R[9];
    const float wy = R[1] * p.x + R[4] * p.y + R[7] * p.z + R[10];
    const float wz = R[2] * p.x + R[5] * p.y + R[8] * p.z + R[11];
    const unsigned long long k = pts_key(sdf_world(g, wx, wy, wz, nullptr), i);
    best = k < best ? k : best;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_xor(best, o);
    best = other < best ? other : best;
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_min[wave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = s_min[0];
    for (int w = 1; w < 4; w++) m = s_min[w] < m ? s_min[w] : m;
    if (m != ~0ull) atomicMin(out, m);
  }
}

}  // namespace

hipError_t kgl_pts_sdf_min(const float4* P, int n, const float X[12], const SdfDev& g, unsigned long long* out,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  PtsXf x;
  for (int i = 0; i < 12; i++) x.v[i] = X[i];
  const int blocks = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  k_pts_sdf_min<<<blocks, 256, 0, s>>>(P, n, x, g, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------- distance Jacobians
//
// For each (configuration, pair) of a batched distance call: the closest
// points of the closest triangle pair (FP64, from the FP64 triangles and
// frames) and the gradient of the distance in joint space,
//   dd/dq = nhat . (dp_a/dq - dp_b/dq),  nhat = (p_a - p_b) / |p_a - p_b|,
// with dp/dq_j = w_j x (p - o_j) for a revolute ancestor j of the point's
// link (w_j = R_j axis_j, o_j = t_j in world coordinates) and w_j for a
// prismatic one; environment points do not move.  Pairs with no triangle pair
// within the bound get d = +inf, zero points and a zero gradient; touching or
// intersecting pairs (d = 0) a zero gradient.

namespace {

template <typename G>
__device__ inline void dj_chain(const LinkDesc* links, const double* T, int l, const double p[3], const double nh[3],
                                double sgn, G* g) {
  for (int j = l; j >= 0; j = links[j].parent) {
    const double* X = T + 12 * j;
    const double* ax = links[j].axis;
    double w[3];
    for (int r = 0; r < 3; r++) w[r] = X[r] * ax[0] + X[3 + r] * ax[1] + X[6 + r] * ax[2];
    double v;
    if (links[j].jtype == 1) {
      v = nh[0] * w[0] + nh[1] * w[1] + nh[2] * w[2];
    } else {
      const double d0 = p[0] - X[9], d1 = p[1] - X[10], d2 = p[2] - X[11];
      const double c0 = w[1] * d2 - w[2] * d1, c1 = w[2] * d0 - w[0] * d2, c2 = w[0] * d1 - w[1] * d0;
      v = nh[0] * c0 + nh[1] * c1 + nh[2] * c2;
    }
    g[j] += (G)(sgn * v);
  }
}

// A point where segment p0p1 meets triangle T (FP64), if it does.
__device__ inline bool dj_seg_tri(const kgm::V<double>& p0, const kgm::V<double>& p1, const kgm::V<double>* T,
                                  kgm::V<double>& x) {
  const kgm::V<double> e1 = kgm::sub(T[1], T[0]), e2 = kgm::sub(T[2], T[0]);
  const kgm::V<double> n = kgm::cross(e1, e2);
  const double s0 = kgm::dot(n, kgm::sub(p0, T[0])), s1 = kgm::dot(n, kgm::sub(p1, T[0]));
  if ((s0 > 0 && s1 > 0) || (s0 < 0 && s1 < 0) || s0 == s1) return false;
  const double t = s0 / (s0 - s1);
  x = kgm::add(p0, kgm::mul(t, kgm::sub(p1, p0)));
  const kgm::V<double> r = kgm::sub(x, T[0]);
  const double d00 = kgm::dot(e1, e1), d01 = kgm::dot(e1, e2), d11 = kgm::dot(e2, e2);
  const double d20 = kgm::dot(r, e1), d21 = kgm::dot(r, e2);
  const double den = d00 * d11 - d01 * d01;
  if (!(den > 0)) return false;
  const double v = (d11 * d20 - d01 * d21) / den, w = (d00 * d21 - d01 * d20) / den;
  return v >= 0 && w >= 0 && v + w <= 1;
}

__global__ __launch_bounds__(256) void k_dist_jac(DistJacArgs a) {
  const long item = (long)blockIdx.x * 256 + threadIdx.x;
  if (item >= (long)a.N * a.npairs) return;
  const int c = (int)(item / a.npairs), p = (int)(item % a.npairs);
  float* g = a.grad_out + item * a.n;
  for (int j = 0; j < a.n; j++) g[j] = 0.f;
  float* cp = a.cp_out + item * 6;
  const float d = a.dist[item];
  const int ta = a.tri[2 * item], tb = a.tri[2 * item + 1];
  if (!(d < a.dcheck) || ta < 0 || tb < 0) {
    a.d_out[item] = __int_as_float(0x7f800000);
    for (int k = 0; k < 6; k++) cp[k] = 0.f;
    return;
  }
  const int la = a.pair_a[p], b = a.pair_b[p], kind = a.pair_kind[p];
  const double* T = a.T64 + (size_t)c * a.n * 12;
  const double* XA = T + 12 * la;
  const double* XB = kind == 0 ? a.plan.env[b].X : T + 12 * b;
  kgm::V<double> P[3], Q[3];
  for (int v = 0; v < 3; v++) {
    const double* u = a.pool.tv64 + 9 * (size_t)ta + 3 * v;
    const double* w = a.pool.tv64 + 9 * (size_t)tb + 3 * v;
    P[v] = kgm::mk(XA[0] * u[0] + XA[3] * u[1] + XA[6] * u[2] + XA[9], XA[1] * u[0] + XA[4] * u[1] + XA[7] * u[2] + XA[10],
                   XA[2] * u[0] + XA[5] * u[1] + XA[8] * u[2] + XA[11]);
    Q[v] = kgm::mk(XB[0] * w[0] + XB[3] * w[1] + XB[6] * w[2] + XB[9], XB[1] * w[0] + XB[4] * w[1] + XB[7] * w[2] + XB[10],
                   XB[2] * w[0] + XB[5] * w[1] + XB[8] * w[2] + XB[11]);
  }
  kgm::V<double> pa, pb;
  const double dd = kgm::tri_dist<double>(P, Q, pa, pb);
  if (!(dd > 0)) {
    // touching or intersecting: report one common point, where an edge of
    // one triangle meets the other (the closest-point pair is not defined)
    kgm::V<double> x;
    bool found = false;
    for (int e = 0; e < 3 && !found; e++) found = dj_seg_tri(P[e], P[(e + 1) % 3], Q, x);
    for (int e = 0; e < 3 && !found; e++) found = dj_seg_tri(Q[e], Q[(e + 1) % 3], P, x);
    if (!found) x = pa;
    pa = pb = x;
  }
  a.d_out[item] = d;
  cp[0] = (float)pa.x; cp[1] = (float)pa.y; cp[2] = (float)pa.z;
  cp[3] = (float)pb.x; cp[4] = (float)pb.y; cp[5] = (float)pb.z;
  if (!(dd > 0)) return;
  const double nh[3] = {(pa.x - pb.x) / dd, (pa.y - pb.y) / dd, (pa.z - pb.z) / dd};
  const double qa[3] = {pa.x, pa.y, pa.z}, qb[3] = {pb.x, pb.y, pb.z};
  dj_chain(a.plan.links, T, la, qa, nh, 1.0, g);
  if (kind == 1) dj_chain(a.plan.links, T, b, qb, nh, -1.0, g);
}

}  // namespace

hipError_t kgl_dist_jac(const DistJacArgs& a, hipStream_t s) {
  const long items = (long)a.N * a.npairs;
  if (items <= 0) return hipSuccess;
  k_dist_jac<<<(unsigned)((items + 255) / 256), 256, 0, s>>>(a);
  return hipGetLastError();
}

// ------------------------------------------------------------ KineTrajOpt sweep terms
//
// Per (segment k, link l, obstacle o): GJK distance (FP64) between the convex
// hull of link l's vertices placed at both ends of the segment (the swept
// volume of KineTrajOpt's setConvexHullGroup, kinetrajopt.py:603-611) and the
// obstacle's hull, then the add_sweep terms (kinetrajopt.py:556-583): the
// normal, the two placements' support points p0, p1 in -normal, alpha =
// |p1 - pswept| / (|p0 - pswept| + |p1 - pswept|) and normal . J at p0 (the
// segment's first configuration) and at p1 (its second).  One wave per
// query: the support maps are maxima over the vertices, taken by all lanes
// and reduced across the wave; the simplex logic runs uniformly on every
// lane (identical data), and lane 0 writes.
namespace {

struct SwV {
  double x, y, z;
};
__device__ inline SwV sw(double x, double y, double z) { return {x, y, z}; }
__device__ inline SwV operator-(SwV a, SwV b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline SwV operator+(SwV a, SwV b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline SwV operator*(double s, SwV a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline double swdot(SwV a, SwV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ inline SwV sw_apply(const double* X, const double* v) {
  return sw(X[0] * v[0] + X[3] * v[1] + X[6] * v[2] + X[9], X[1] * v[0] + X[4] * v[1] + X[7] * v[2] + X[10],
            X[2] * v[0] + X[5] * v[1] + X[8] * v[2] + X[11]);
}

// Support point of the copies of local points V[nv] placed by X0 (and X1 if
// given) in direction d: the index k * nv + i of the (copy k, vertex i) with
// the largest d . (R v + t), ties to the lowest copy, then vertex (the first
// maximum of the scan).  A NaN direction gives index 0.
__device__ inline int sw_support(const double* __restrict__ V, int nv, const double* X0, const double* X1, SwV d) {
  double best = -1.0 / 0.0;
  int bi = 0x7fffffff;
  for (int k = 0; k < (X1 ? 2 : 1); k++) {
    const double* X = k ? X1 : X0;
    // d in the copy's frame, and the translation's part
    const double l0 = X[0] * d.x + X[1] * d.y + X[2] * d.z;
    const double l1 = X[3] * d.x + X[4] * d.y + X[5] * d.z;
    const double l2 = X[6] * d.x + X[7] * d.y + X[8] * d.z;
    const double off = X[9] * d.x + X[10] * d.y + X[11] * d.z;
#pragma unroll 4
    for (int i = 0; i < nv; i++) {
      const double val = l0 * V[3 * i] + l1 * V[3 * i + 1] + l2 * V[3 * i + 2] + off;
      if (val > best) {
        best = val;
        bi = k * nv + i;
      }
    }
  }
  return bi == 0x7fffffff ? 0 : bi;
}

// The point of a support index.
__device__ inline SwV sw_point(const double* V, int nv, const double* X0, const double* X1, int bi) {
  const int k = X1 ? bi / nv : 0, i = bi - k * nv;
  return sw_apply(k ? X1 : X0, V + 3 * i);
}

// Vertex subsets of a simplex as bit masks: the number of members and the
// index of the j-th member, evaluated at compile time.
__host__ __device__ constexpr int sw_nbits(int s) { return (s & 1) + (s >> 1 & 1) + (s >> 2 & 1) + (s >> 3 & 1); }
__host__ __device__ constexpr int sw_bit(int s, int j) {
  return j == 0 ? ((s & 1) ? 0 : (s & 2) ? 1 : (s & 4) ? 2 : 3) : sw_bit(s & (s - 1), j - 1);
}

// One vertex subset S of w: the closest point of its affine hull (normal
// equations of its edge vectors) when all its barycentric weights are
// positive; it replaces the best so far when strictly nearer.  Every index
// is a compile-time constant, so the small arrays stay in registers.
template <int S>
__device__ __forceinline__ void sw_subset(const SwV* w, double& bestd, SwV& bestp, int& chosen, double* lam) {
  constexpr int K = sw_nbits(S), Q = K - 1;
  double mu[4] = {1, 0, 0, 0};
  if constexpr (K > 1) {
    const SwV p0 = w[sw_bit(S, 0)];
    SwV e[3];
#pragma unroll
    for (int j = 1; j < K; j++) e[j - 1] = w[sw_bit(S, j)] - p0;
    double G[3][3], r[3];
#pragma unroll
    for (int i = 0; i < Q; i++) {
      r[i] = -swdot(e[i], p0);
#pragma unroll
      for (int j = 0; j < Q; j++) G[i][j] = swdot(e[i], e[j]);
    }
    double x[3];
    double scale = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) scale = fmax(scale, G[i][i]);
    if constexpr (Q == 1) {
      if (!(G[0][0] > 0)) return;
      x[0] = r[0] / G[0][0];
    } else if constexpr (Q == 2) {
      const double det = G[0][0] * G[1][1] - G[0][1] * G[1][0];
      if (!(det > 1e-24 * scale * scale)) return;
      x[0] = (r[0] * G[1][1] - G[0][1] * r[1]) / det;
      x[1] = (G[0][0] * r[1] - r[0] * G[1][0]) / det;
    } else {
      const double c00 = G[1][1] * G[2][2] - G[1][2] * G[2][1];
      const double c01 = G[1][2] * G[2][0] - G[1][0] * G[2][2];
      const double c02 = G[1][0] * G[2][1] - G[1][1] * G[2][0];
      const double det = G[0][0] * c00 + G[0][1] * c01 + G[0][2] * c02;
      if (!(det > 1e-36 * scale * scale * scale)) return;
      x[0] = (r[0] * c00 + G[0][1] * (r[2] * G[1][2] - r[1] * G[2][2]) + G[0][2] * (r[1] * G[2][1] - r[2] * G[1][1])) / det;
      x[1] = (G[0][0] * (r[1] * G[2][2] - r[2] * G[1][2]) + r[0] * c01 + G[0][2] * (r[2] * G[1][0] - r[1] * G[2][0])) / det;
      x[2] = (G[0][0] * (r[2] * G[1][1] - r[1] * G[2][1]) + G[0][1] * (r[1] * G[2][0] - r[2] * G[1][0]) + r[0] * c02) / det;
    }
    double sum = 0;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < Q; i++) {
      mu[i + 1] = x[i];
      sum += x[i];
      ok = ok && x[i] > 0;
    }
    mu[0] = 1 - sum;
    if (!(ok && mu[0] > 0)) return;
  }
  SwV p = sw(0, 0, 0);
#pragma unroll
  for (int j = 0; j < K; j++) p = p + mu[j] * w[sw_bit(S, j)];
  const double dd = swdot(p, p);
  if (dd < bestd) {
    bestd = dd;
    bestp = p;
    chosen = S;
#pragma unroll
    for (int i = 0; i < 4; i++) lam[i] = 0;
#pragma unroll
    for (int j = 0; j < K; j++) lam[sw_bit(S, j)] = mu[j];
  }
}

// Closest point to the origin of the simplex w[0..m) (m <= 4): over the
// vertex subsets, in increasing mask order, the nearest closest point of a
// subset's affine hull that lies in the simplex.  lam receives the weights
// over w (0 outside the chosen subset), chosen the subset's mask.
__device__ __forceinline__ SwV sw_simplex_min(const SwV* w, int m, double* lam, int& chosen) {
  double bestd = 1.0 / 0.0;
  SwV bestp = w[0];
  chosen = 1;
  sw_subset<1>(w, bestd, bestp, chosen, lam);
  if (m >= 2) {
    sw_subset<2>(w, bestd, bestp, chosen, lam);
    sw_subset<3>(w, bestd, bestp, chosen, lam);
  }
  if (m >= 3) {
    sw_subset<4>(w, bestd, bestp, chosen, lam);
    sw_subset<5>(w, bestd, bestp, chosen, lam);
    sw_subset<6>(w, bestd, bestp, chosen, lam);
    sw_subset<7>(w, bestd, bestp, chosen, lam);
  }
  if (m >= 4) {
    sw_subset<8>(w, bestd, bestp, chosen, lam);
    sw_subset<9>(w, bestd, bestp, chosen, lam);
    sw_subset<10>(w, bestd, bestp, chosen, lam);
    sw_subset<11>(w, bestd, bestp, chosen, lam);
    sw_subset<12>(w, bestd, bestp, chosen, lam);
    sw_subset<13>(w, bestd, bestp, chosen, lam);
    sw_subset<14>(w, bestd, bestp, chosen, lam);
    sw_subset<15>(w, bestd, bestp, chosen, lam);
  }
  return bestp;
}

// GJK distance between the hull of A's copies (XA0, XA1 nullable) and the
// hull of B (XB); 0 when they overlap.  ca, cb: the closest points.  One
// lane runs one query: the simplex and its vertices' support indices live in
// registers, every runtime-indexed update written as selects over the four
// slots.
__device__ double sw_gjk(const double* VA, int na, const double* XA0, const double* XA1, const double* VB, int nb,
                         const double* XB, SwV& ca, SwV& cb) {
  SwV w[4];
  int ia[4] = {0, 0, 0, 0}, ib[4] = {0, 0, 0, 0};
  double lam[4] = {1, 0, 0, 0};
  int m = 1;
  w[0] = sw_apply(XA0, VA) - sw_apply(XB, VB);
  w[1] = w[2] = w[3] = w[0];
  SwV v = w[0];
  double dist = sqrt(swdot(v, v));
  for (int it = 0; it < 64; it++) {
    const double vv = swdot(v, v);
    if (vv <= 1e-30) {
      dist = 0;
      break;
    }
    const int ja = sw_support(VA, na, XA0, XA1, -1.0 * v);
    const int jb = sw_support(VB, nb, XB, nullptr, v);
    const SwV nw = sw_point(VA, na, XA0, XA1, ja) - sw_point(VB, nb, XB, nullptr, jb);
    // the new support point cannot bring v closer by more than the tolerance
    if (vv - swdot(v, nw) <= 1e-13 * vv) break;
    bool dup = false;
#pragma unroll
    for (int i = 0; i < 4; i++) dup = dup || (i < m && w[i].x == nw.x && w[i].y == nw.y && w[i].z == nw.z);
    if (dup) break;
#pragma unroll
    for (int i = 0; i < 4; i++)
      if (i == m) {
        w[i] = nw;
        ia[i] = ja;
        ib[i] = jb;
      }
    m++;
    int chosen;
    double l2[4] = {0, 0, 0, 0};
    const SwV nv = sw_simplex_min(w, m, l2, chosen);
    // keep the chosen vertices, in order
    SwV w2[4] = {w[0], w[0], w[0], w[0]};
    int ia2[4] = {0, 0, 0, 0}, ib2[4] = {0, 0, 0, 0};
    double lam2[4] = {lam[0], lam[1], lam[2], lam[3]};
    int k = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
      if (i < m && (chosen >> i & 1)) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (j == k) {
            w2[j] = w[i];
            ia2[j] = ia[i];
            ib2[j] = ib[i];
            lam2[j] = l2[i];
          }
        k++;
      }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      w[j] = w2[j];
      ia[j] = ia2[j];
      ib[j] = ib2[j];
      lam[j] = lam2[j];
    }
    m = k;
    if (m == 4) {
      dist = 0;
      break;
    }
    const double nd = sqrt(swdot(nv, nv));
    v = nv;
    if (nd >= dist) {
      dist = nd;
      break;
    }
    dist = nd;
  }
  ca = sw(0, 0, 0);
  cb = sw(0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (i < m) {
      ca = ca + lam[i] * sw_point(VA, na, XA0, XA1, ia[i]);
      cb = cb + lam[i] * sw_point(VB, nb, XB, nullptr, ib[i]);
    }
  return dist;
}

// One lane per (segment, link, obstacle) query.  The queries of a wave share
// their link and obstacle (each pair's N segments are padded to whole waves),
// so both hulls' addresses and their vertex scans are wave-uniform: their
// vertices come through scalar loads, and every load precedes the kernel's
// first store.
// Four waves per SIMD (128 VGPRs, a small spill) run the 4,096-segment batch
// 3-5% faster than two (174 VGPRs, no spill) or three.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_sweep(SweepArgs a, int wpl) {
  const int lo = (int)blockIdx.x / wpl;
  const int l = lo / a.O, o = lo - l * a.O;
  const int c = ((int)blockIdx.x - lo * wpl) * 64 + (int)threadIdx.x;
  if (c >= a.N) return;
  const long q = ((long)c * a.n + l) * a.O + o;
  const HullRef& H = a.lhull[l];
  const HullRef& B = a.ohull[o];
  const double* X0 = a.T0 + ((size_t)c * a.n + l) * 12;
  const double* X1 = a.T1 + ((size_t)c * a.n + l) * 12;
  SwV ps = sw(0, 0, 0), po = sw(0, 0, 0);
  double d = 1.0 / 0.0;
  if (H.nv > 0 && B.nv > 0) d = sw_gjk(H.V, H.nv, X0, X1, B.V, B.nv, B.X, ps, po);
  double* info = a.info_out + q * 16;
  double* coef = a.coef_out + q * 2 * a.n;
  if (!(d < a.dcheck) || !(d > 0)) {
    // beyond the bound (not a constraint), or overlapping hulls (no
    // separating normal): the distance, and no terms
    a.d_out[q] = d < a.dcheck ? d : 1.0 / 0.0;
    for (int k = 0; k < 16; k++) info[k] = 0;
    for (int j = 0; j < 2 * a.n; j++) coef[j] = 0;
    if (d < a.dcheck) {
      info[4] = ps.x; info[5] = ps.y; info[6] = ps.z;
      info[7] = po.x; info[8] = po.y; info[9] = po.z;
    }
    return;
  }
  const SwV nrm = (1.0 / d) * (ps - po);
  const SwV p0 = sw_point(H.V, H.nv, X0, nullptr, sw_support(H.V, H.nv, X0, nullptr, -1.0 * nrm));
  const SwV p1 = sw_point(H.V, H.nv, X1, nullptr, sw_support(H.V, H.nv, X1, nullptr, -1.0 * nrm));
  const SwV e0 = p0 - ps, e1 = p1 - ps;
  const double d0 = sqrt(swdot(e0, e0)), d1 = sqrt(swdot(e1, e1));
  const double alpha = d0 + d1 > 0 ? d1 / (d0 + d1) : 0.5;
  a.d_out[q] = d;
  info[0] = alpha;
  info[1] = nrm.x; info[2] = nrm.y; info[3] = nrm.z;
  info[4] = ps.x; info[5] = ps.y; info[6] = ps.z;
  info[7] = po.x; info[8] = po.y; info[9] = po.z;
  info[10] = p0.x; info[11] = p0.y; info[12] = p0.z;
  info[13] = p1.x; info[14] = p1.y; info[15] = p1.z;
  for (int j = 0; j < 2 * a.n; j++) coef[j] = 0;
  const double nh[3] = {nrm.x, nrm.y, nrm.z};
  const double q0[3] = {p0.x, p0.y, p0.z}, q1[3] = {p1.x, p1.y, p1.z};
  dj_chain(a.links, a.T0 + (size_t)c * a.n * 12, l, q0, nh, 1.0, coef);
  dj_chain(a.links, a.T1 + (size_t)c * a.n * 12, l, q1, nh, 1.0, coef + a.n);
}

}  // namespace

hipError_t kgl_sweep(const SweepArgs& a, hipStream_t s) {
  const long total = (long)a.N * a.n * a.O;
  if (total <= 0) return hipSuccess;
  // waves per (link, obstacle) pair: its N segments, one per lane
  const long wpl = ((long)a.N + 63) / 64;
  if (wpl * a.n * a.O > 0x7fffffffL) return hipErrorInvalidValue;
  k_sweep<<<(unsigned)(wpl * a.n * a.O), 64, 0, s>>>(a, (int)wpl);
  return hipGetLastError();
}
