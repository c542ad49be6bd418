// Data layout shared by the host code and the HIP kernels.
//
// HBM layout (one pool per context, all geometries concatenated):
//   ObbNode  nodes[]   64 B per node, one 64 B record per node; a node's
//                      children are adjacent (child, child+1); leaves hold one
//                      triangle (child = ~tri, tri in pool numbering).
//   float4   tv0[], tv1[], tv2[]   FP32 vertices of every triangle in BVH leaf
//                      order (three structure-of-arrays planes, 16 B each).
//   double   tv64[]    FP64 vertices (9 per triangle) for exact re-checks.
//   int32    tri_src[] original triangle index within its mesh.
#pragma once
#include <stdint.h>

#define KG_MAX_LINKS 64

// Oriented box: points c + sum_k s_k e_k R[:,k], |s_k| <= 1.  R column-major.
// The FP32 values are chosen so that the box contains the FP64 triangles
// exactly (centre and axes rounded first, extents computed in the dual basis
// of the rounded axes and rounded up).
struct ObbNode {
  float c[3];
  float e[3];
  float R[9];
  int child;  // >= 0: first child (pool index); < 0: leaf, triangle = ~child
};

struct GeomDesc {
  int root;       // pool index of the root node
  int nnodes;
  int tri0;       // first triangle (pool numbering)
  int ntris;
  int depth;      // max root-to-leaf edge count
  int pad_;
  double lc[3];   // local AABB centre (from the vertices, FP64)
  double le[3];   // local AABB half extents
  double rmax;    // max |vertex| in the local frame
};

// Per-robot link table (FP64 kinematics).
struct LinkDesc {
  double T0[12];  // column-major R then t
  double axis[3];
  int parent;
  int jtype;      // 0 revolute, 1 prismatic
  int geom;       // geometry index in the pool, -1 = empty
  int wid;        // world ID of the link
};

// A placed environment geometry (terrain, rigid object, other robot link).
struct EnvItem {
  double X[12];
  double lo[3], hi[3];  // world AABB
  int geom;
  int wid;
};

// One narrow-phase job: link `a` of config `cfg` against either env item
// `b` (kind 0) or link `b` of the same config (kind 1).
struct PairItem {
  int cfg;
  int16_t a;
  int16_t b;
  int rank;   // order inside the group (Klampt loop order); group in bit 30
};

// What the FP32 traversal needs to start a pair item, computed once per item
// by the per-link broad phase: the item transform (mesh A frame -> mesh B
// frame, FP64 product rounded to FP32), its error bounds and the two BVH
// roots.
struct __attribute__((aligned(16))) ItemRec {
  float rel[12];  // column-major R, then t
  float tau, pad;
  int rootA, rootB;
};

#define KG_RANK_GROUP_BIT 30
#define KG_RANK_NONE 0x7fffffff
