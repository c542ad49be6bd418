// Host-side objects behind the C-ABI handles and the kernel launch interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "kg_types.h"
#include "../../include/klampt_gpu.h"

struct BuiltTree {
  std::vector<ObbNode> nodes;  // local indices (child relative to node 0)
  std::vector<int> order;      // leaf order -> original triangle
  int depth = 0;
};
void kg_build_obb_tree(const double* V, int nv, const int32_t* T, int nt, BuiltTree& out);

// Copies of the traversal statistics (128-byte lines of the status words):
// a wave adds its counts to copy blockIdx % KG_CNT_COPIES.
#define KG_CNT_COPIES 16
// Triangle pairs the FP32 test left undecided, per batch, that the FP64
// pass decides one by one (beyond this count a configuration is traversed
// again in FP64 instead).
#define KG_XPAIR_CAP 1024

// Count of device-buffer frees in the process (a buffer growing frees its
// old block): a captured HIP graph holds workspace addresses, so its owner
// recaptures when the count has moved since the capture.  A first allocation
// invalidates no address and is not counted.
inline std::atomic<uint64_t> g_devbuf_gen{0};

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) {
      g_devbuf_gen.fetch_add(1);
      // kernels still queued on any stream may read the old block
      hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) return e;
      e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      bytes = 0;
    }
    size_t want = n < 256 ? 256 : n;
    want = (want + 4095) & ~(size_t)4095;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p) {
      g_devbuf_gen.fetch_add(1);
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

struct PoolDev {
  const ObbNode* nodes;
  const float4* tv0;
  const float4* tv1;
  const float4* tv2;
  const double* tv64;
  const GeomDesc* geoms;
  const int* trisrc;
};

struct SdfDev {
  const float* v;
  // nullable: per cell (i, j, k) of the grid's (dims - 1)^3 cells, its 8 corner
  // values, 2 float4 (c000 c001 c010 c011, c100 c101 c110 c111): one 32-byte
  // read per sample instead of 8 gathers over about 4 cache lines
  const float4* v8;
  int cdims[3];
  int dims[3];
  float bmin[3];
  float inv_h[3];  // 1/h
  float h[3];
  float Xinv[12];  // world -> grid frame (column-major R, then t)
  int ident;       // grid frame == world frame
  double bmin64[3];
  double inv_h64[3];
};

// Robot/world tables for one batched check, resident in HBM.
struct PlanDev {
  const LinkDesc* links;
  const double* qmin;
  const double* qmax;
  const EnvItem* env;
  const uint8_t* lemask;  // nlinks x nenv: link/env pair enabled
  const int* selfpairs;   // 2 x nself, link indices, Klampt loop order
  int nlinks;
  int nenv;
  int nself;
  int depth;              // longest root-to-link chain, in links
};

// Error-bound parameters of the FP32 narrow phase.
struct Tol {
  float ev;       // bound on a transformed vertex's position error (m)
  float tau;      // separation certainty threshold (normalised gap)
  float pad_bv;   // absolute bounding-volume padding
  float pad_rel;  // relative bounding-volume padding (x extents)
};

enum { KG_MODE_ANY = 0, KG_MODE_GROUPS = 1, KG_MODE_ORDERED = 2 };

// ---- kernel launchers (kg_kernels.hip)
hipError_t kgl_fk_f32(const LinkDesc* links, int nlinks, const float* Q, int N, float* Tout, hipStream_t s);

struct PrepArgs {
  const float* Qf;     // either Qf or Qd
  const double* Qd;
  int N;
  int mode;
  PlanDev plan;
  PoolDev pool;
  double* T64;         // N x nlinks x 12 (nullable for the per-link variant)
  double* LB;          // N x nlinks x 6 (world AABBs)
  uint8_t* jl;         // N: 1 if within joint limits
  int* counts;         // N+1
  int* offsets;        // N+1 (two-pass layout) or nullptr (fixed capacity)
  PairItem* items;
  int cap;             // items per configuration in the fixed-capacity layout
  int* reset;          // nreset ints zeroed by the prep kernel
  int nreset;
  // heterogeneous batches (all nullable): per-configuration plan index, Q
  // offset (floats), frame offset (links) and item base
  const PlanDev* plans;
  const int* cfg_plan;
  const int* qoff;
  const int* toff;
  const int* ibase;
  int lds_links;       // serial single-plan variant with the link boxes in LDS
  int group;           // > 0: one lane per link in groups of this many lanes (8, 16, 32)
  ItemRec* recs;       // nullable: item start records (written by the per-link variant only)
  size_t t_rows;       // frame rows of T64 (structure of arrays, see t_row)
  const int* N_dev;    // nullable: device-side configuration count (<= N, the capacity); the
                       // kernels then loop over the configurations on a bounded grid
  int jump_rounds;     // per-link variant: pointer-jumping rounds, ceil(log2(depth)) over
                       // the batch's plans
};
hipError_t kgl_prep_count(const PrepArgs& a, hipStream_t s);
hipError_t kgl_scan(int* counts, int* offsets, int n, DevBuf& tmp, hipStream_t s);
hipError_t kgl_prep_fill(const PrepArgs& a, hipStream_t s);
hipError_t kgl_prep_single(const PrepArgs& a, hipStream_t s);

struct FinArgs {
  int N;
  int mode;
  const uint8_t* jl;
  const PairItem* items;
  PlanDev plan;
  uint8_t* feasible;
  uint8_t* env_hit;
  uint8_t* self_hit;
  int* first_pair;
  int* redo_list;   // configs needing the FP64 pass
  int* redo_count;
  int final_pass;   // 1: uncertainty cannot remain
  int force_recheck;  // test hook: every configuration goes to the FP64 pass
  int* uncertain;     // count of configurations left open by dropped work (exact pass)
  const PlanDev* plans;  // heterogeneous batches (nullable)
  const int* cfg_plan;
  // nullable: undecided triangle pairs {configuration, item, triangle of A,
  // triangle of B} recorded by the FP32 pass for the exact test of the FP64
  // pass (any-hit and group modes only: a pair can then only add a hit)
  int4* xpairs;
  int* xpair_count;   // may exceed xpair_cap: the pairs beyond it were not recorded
  int xpair_cap;
};

struct TravArgs {
  PoolDev pool;
  PlanDev plan;
  const double* T64;
  const PairItem* items;
  const int* offsets;     // per config item range [offsets[c], offsets[c+1]), or nullptr:
  const int* counts;      //   [c*cap, c*cap + counts[c])
  int cap;
  const int* cfg_list;    // nullable: configs to process (else 0..N-1)
  int ncfg;
  const int* ncfg_dev;    // nullable: device-side configuration count
  int mode;
  Tol tol;
  int* chunk_counter;
  int* chunk_drained;         // per range (same stride): set once a wave found it drained
  uint64_t* spill;            // per-wave global overflow stacks
  int spill_cap;              // entries per wave
  int* status;                // [0]: spill overflow flag
  int cpw;                    // configs per chunk
  int leaf_min;               // lanes holding leaf pairs that trigger a triangle-test step
  int tail_lanes;             // fewer working lanes than this: tail cadence
  int tail_reps;              // iterations per item-start step in the tail
  int stack_cap;              // per-lane traversal stack entries (LDS)
  int grid;                   // number of waves launched
  unsigned long long* counters;  // [0] box tests, [1] triangle tests, [2] items (nullable)
  const PlanDev* plans;       // heterogeneous batches (nullable, see PrepArgs)
  const int* cfg_plan;
  const int* toff;
  const int* ibase;
  int scap;                   // LDS pool capacity variant (512 or 1024)
  const ItemRec* recs;        // per item start records (FP32 pass)
  int static_pct;             // percent of the configurations split statically over the waves
  int nsub;                   // dynamic ranges, each with a counter at chunk_counter + 32 r
  FinArgs fin;                // result assembly when a configuration's last item finishes
  size_t t_rows;              // frame rows of T64 and the batch size (t_row layout)
  int t_n;
  const float* Qf;            // joint values (FP64 pass forms its frames from them)
  const double* Qd;
  const int* qoff;            // heterogeneous batches: Q offset per configuration
};
hipError_t kgl_collide_f32(const TravArgs& a, hipStream_t s);

hipError_t kgl_collide_f64(const TravArgs& a, hipStream_t s);

struct DistArgs {
  PoolDev pool;
  PlanDev plan;
  const double* T64;
  int N;
  int npairs;
  const int* pair_a;     // link index
  const int* pair_b;     // env item (kind 0) or link index (kind 1)
  const int* pair_kind;
  float stop_below;      // an item stops once its distance is <= this
  float upper;           // initial bound (DistanceQuerySettings::upperBound)
  float rel_err, abs_err;  // DistanceQuerySettings relErr / absErr (0: exact)
  float* out;            // N x npairs
  int* out_tri;          // nullable: N x npairs x 2 closest triangle pair (pool numbering)
  int* chunk_counter;
  uint64_t* spill;
  int spill_cap;
  int* status;
  int grid;
  int ipw;               // items per chunk
  unsigned long long* counters;
  int* uncertain;        // count of items reported as a lower bound (dropped work)
};
hipError_t kgl_distance(const DistArgs& a, hipStream_t s);
int kgl_distance_grid(int num_cus);
hipError_t kgl_fk64(const LinkDesc* links, int n, const float* Qf, const double* Qd, int N, double* T64, int* reset,
                    int nreset, hipStream_t s);

// Device-side edge checks (rounds of kgl_edge_round, see kg_kernels.hip).
struct EdgeArgs {
  const float* Af;       // nseg x n endpoints, FP32 (or FP64 below)
  const float* Bf;
  const double* Ad;
  const double* Bd;
  int n, nseg;
  double eps;
  int norm_inf;
  int max_pos;           // bisection positions the round schedule covers per segment
  int budget;            // samples per round (capacity of the feasibility workspace)
  int* ns;               // nseg: steps n_s
  int* cur;              // nseg: positions checked
  uint8_t* done;         // nseg
  float* first_u;        // nseg: result
  int32_t* nsamples_out; // nullable: n_s per segment
  int* seg;              // nseg: live segments of the round
  int* soff;             // nseg + 1: their sample offsets
  int* ctl;              // 8 words: [0] live, [1] samples, [2] spill, [3] uncertain,
                         // [4] FP64 re-checks (accumulated), [5] undecided segments,
                         // [6] segments with a NaN coordinate
  double* Qd;            // budget x n samples
  const uint8_t* feas;   // budget: feasibility of the round's samples
  const int* st;         // status words of the feasibility pipeline
  int w_spill, w_unc, w_redo;
  int* out_status;       // status words of the call (same layout)
};
hipError_t kgl_edge_init(const EdgeArgs& a, bool clear_ctl, hipStream_t s);
// phase 0: plan + samples; phase 1: reduce
hipError_t kgl_edge_round(const EdgeArgs& a, int width, int phase, hipStream_t s);
hipError_t kgl_edge_finish(const EdgeArgs& a, bool write_status, hipStream_t s);

hipError_t kgl_fill_int(int* p, int n, int v, hipStream_t s);

// Scratch array of the allocation-free feasibility call: KG_NOALLOC_BYTES per
// device, its address on `device` (which must be current) looked up once.
#define KG_NOALLOC_BYTES (64u << 20)
#define KG_MAX_DEVICES 64
hipError_t kgl_noalloc_scratch(int device, void** ptr);

// TriangleMesh -> ImplicitSurface: the grid of cell-centred signed distances.
#define KG_MSDF_STK 64  // traversal stack entries (tree depth + 2 at most)
struct MeshSdfArgs {
  PoolDev pool;
  int root;          // pool index of the mesh's root node
  int dims[3];
  double bmin[3];    // grid corner, mesh frame
  double h;          // cell size
  float* out;        // dims[0] x dims[1] x dims[2], C order (k fastest)
  int* diff;         // same size: winding increments (zeroed before k_msdf_sign)
};
hipError_t kgl_msdf_dist(const MeshSdfArgs& a, hipStream_t s);
// Closest points and joint-space distance gradients of batched pair distances.
struct DistJacArgs {
  PoolDev pool;
  PlanDev plan;
  const double* T64;      // N x nlinks x 12 world frames (k_fk64 layout)
  int N, npairs, n;
  const int* pair_a;
  const int* pair_b;
  const int* pair_kind;
  const float* dist;      // N x npairs (k_distance)
  const int* tri;         // N x npairs x 2 closest triangle pair, pool numbering (-1: none within the bound)
  float dcheck;
  float* d_out;           // N x npairs
  float* cp_out;          // N x npairs x 6
  float* grad_out;        // N x npairs x n
};
hipError_t kgl_dist_jac(const DistJacArgs& a, hipStream_t s);
// A convex hull as the local points it spans, with a placement (obstacles;
// links take theirs from the frames).
struct HullRef {
  const double* V;  // nv x 3 (FP64)
  int nv;           // 0: no geometry
  int pad_;
  double X[12];     // column-major R, then t
};
struct SweepArgs {
  const LinkDesc* links;
  int n;                  // links
  int N;                  // segments
  int O;                  // obstacles
  const double* T0;       // N x n x 12 frames at each segment's first configuration
  const double* T1;       // ... and at its second
  const HullRef* lhull;   // n
  const HullRef* ohull;   // O
  double dcheck;
  double* d_out;          // N x n x O
  double* info_out;       // N x n x O x 16: alpha, normal, pswept, pobs, p0, p1
  double* coef_out;       // N x n x O x 2n: normal . J(q0, p0), normal . J(q1, p1)
};
hipError_t kgl_sweep(const SweepArgs& a, hipStream_t s);
// Minimum implicit-surface value over a point set: out = min over i of
// key(value(X * P[i]), i), key = order-preserving float bits << 32 | i
// (*out must hold ~0 before the call).
hipError_t kgl_pts_sdf_min(const float4* P, int n, const float X[12], const SdfDev& g, unsigned long long* out,
                           hipStream_t s);
hipError_t kgl_msdf_sign(const MeshSdfArgs& a, hipStream_t s);
int kgl_collide_grid(int exact, int num_cus, int scap);
hipError_t kgl_sdf_eval(const SdfDev& g, const float* pts, int n, float* d, float* grad, hipStream_t s);
// Fills out[2 * cell], out[2 * cell + 1] with the 8 corner values of each cell
// (layout of SdfDev::v8) from the grid values v.
hipError_t kgl_sdf_build8(const float* v, const int dims[3], float4* out, hipStream_t s);
hipError_t kgl_sdf_penalty(const LinkDesc* links, int n, const SdfDev& g, const float* Q, int N, const float* pts,
                           const int* poff, float dsafe, float* loss, const float* dloss, float* dQ, bool bwd,
                           int lanes, hipStream_t s);
// Copies the status words [0] spill flag, [1] uncertain count, [2] FP64
// re-checks of a batch into out[0..3] (out[3] = 0).
hipError_t kgl_status_copy(const int* status, int w_spill, int w_unc, int w_redo, int32_t* out, hipStream_t s);
// acc[w] = st[w] (first) or acc[w] OR= / += st[w] (spill flag / counts), for
// the status words w_spill, w_unc, w_redo of one part of a pipeline
hipError_t kgl_status_accum(const int* st, int* acc, int w_spill, int w_unc, int w_redo, int first, hipStream_t s);

// ---- C-ABI object layouts

inline kg_params kg_default_params() {
  kg_params p;
  p.cpw = 16;
  p.leaf_min = 1;
  p.scap = 256;  // leaves LDS for 4 traversal waves per SIMD
  p.waves_per_cu = 0;
  p.static_pct = 60;
  p.nsub = 8;
  p.tail_lanes = 16;
  p.tail_reps = 4;
  p.spill_cap = 16384;
  p.dist_ipw = 16;
  p.edge_w0 = 2;
  p.sdf_lanes = 32;
  p.prep_group = 1;
  p.force_recheck = 0;
  p.edge_budget_mb = 4096;
  p.edge_growth = 16;
  return p;
}

struct kg_ctx_s;
struct kg_geom_s;
struct kg_robot_s;
struct kg_world_s;

struct Xf64 {
  double R[9];
  double t[3];
};

struct kg_geom_s {
  kg_ctx_s* ctx;
  int kind;        // 0 triangle mesh, 1 SDF, 2 point cloud
  int index;       // GeomDesc index (mesh)
  Xf64 X;          // current transform
  // mesh data (kept for host-side queries)
  std::vector<double> V;
  std::vector<int32_t> T;
  // SDF
  int dims[3] = {0, 0, 0};
  double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  DevBuf dvals;
  DevBuf dpts;     // points (point cloud) or mesh vertices as float4, uploaded on first use
  DevBuf dv64;     // the same points in FP64 (convex-hull support maps), uploaded on first use
  DevBuf dv8;      // SDF: corner values per cell (SdfDev::v8), built on first use by the penalty
  bool alive = true;  // the handle has not been destroyed
  int refs = 1;       // the handle plus every robot link, terrain and rigid object slot holding it
  double margin = 0;  // collision margin (Geometry3D.setCollisionMargin)
};

struct kg_robot_s {
  kg_ctx_s* ctx;
  int n;
  std::vector<LinkDesc> links;
  std::vector<kg_geom_s*> link_geoms;
  std::vector<double> qmin, qmax, q;
  std::vector<uint8_t> selfcol;  // n x n, upper triangular (i<j)
  std::vector<Xf64> Tw;          // current link transforms
  DevBuf dlinks;
  bool dirty = true;
  int version = 0;               // bumped on any change affecting plans
};

struct RobotPlan {
  int built_version = -1;
  int nenv = 0, nself = 0;
  std::vector<EnvItem> env;
  std::vector<uint8_t> lemask;
  std::vector<int> selfpairs;
  std::vector<std::pair<int, int>> pairs;  // world ID pairs for distance output
  std::vector<int> pair_a, pair_b, pair_kind;  // link, env-or-link, 0 env / 1 self
  DevBuf dlinks, dqmin, dqmax, denv, dlemask, dself, dpa, dpb, dpk;
  PlanDev dev{};
  Tol tol32{}, tol64{};
  int maxdepth_a = 0, maxdepth_b = 0;
  // convex hulls of the links and obstacles (KineTrajOpt sweep terms)
  DevBuf dhull;
  int hull_version = -1;
  int nobs = 0;
};

struct kg_world_s {
  kg_ctx_s* ctx;
  std::vector<kg_geom_s*> terrains;
  std::vector<std::pair<kg_geom_s*, Xf64>> objects;
  std::vector<kg_robot_s*> robots;
  std::vector<uint8_t> mask;
  int nids = 0;
  int version = 0;
  std::vector<RobotPlan> plans;
};

struct kg_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // geometry pool, host mirror
  std::vector<ObbNode> nodes;
  std::vector<float4> tv0, tv1, tv2;
  std::vector<double> tv64;
  std::vector<int> trisrc;
  std::vector<GeomDesc> geoms;
  bool pool_dirty = true;
  DevBuf dnodes, dtv0, dtv1, dtv2, dtv64, dgeoms, dtrisrc;
  PoolDev pool{};
  int pool_version = 0;
  // workspace
  int* hstatus = nullptr;  // pinned copy of the last batch's status words
  bool has_batch = false;     // a feasibility batch has been issued
  hipEvent_t evStat = nullptr;  // recorded at the end of the last batch
  const int* stat_src = nullptr;  // device status words of the last batch
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int last_N = 0;
  // ordering of work across the streams callers name: evDone is recorded on
  // done_stream at the end of every call that issued GPU work
  hipEvent_t evDone = nullptr;
  // begin/end events of the FP32 traversal kernel of the most recent
  // feasibility batches (ring), for averaging its duration over many calls
  static constexpr int TRING = 64;
  hipEvent_t tr_beg[TRING] = {}, tr_end[TRING] = {};
  int64_t tr_count = 0;
  hipStream_t done_stream = nullptr;
  bool done_valid = false;
  bool last_slotted = false;  // the last call was a feasibility batch on a workspace slot
  // Workspace slots of device-pointer feasibility batches: a batch takes the
  // next slot and waits only for that slot's previous batch, so batches
  // issued on different streams run concurrently (each fills the idle tail
  // of the traversal before it).
  struct FeasSlot {
    DevBuf status, jl, counts, redo, items, recs, spill;
    hipEvent_t done = nullptr;
    bool valid = false;
  };
  static constexpr int FEAS_SLOTS = 3;
  FeasSlot fslot[FEAS_SLOTS];
  int fslot_next = 0;
  // Workspace slots of device-pointer edge batches, on the same terms: the
  // feasibility workspaces of the batch's rounds and its edge buffers
  // (segment state, round plan, samples, status).
  struct EdgeSlot {
    FeasSlot f;
    DevBuf ns, cur, done, seg, soff, ctl, q, feas, stat;
  };
  static constexpr int EDGE_SLOTS = 2;
  EdgeSlot eslot[EDGE_SLOTS];
  int eslot_next = 0;
  DevBuf wsQ, wsT64, wsLB, wsJL, wsCounts, wsOffsets, wsItems, wsScanTmp, wsBest, wsUnc, wsChunk, wsSpill,
      wsStatus, wsRedo, wsOut8a, wsOut8b, wsOut8c, wsOut32, wsF, wsG, wsPoff, wsPts, wsLoss, wsDQ,
      wsDl, wsEA, wsEB, wsESeg, wsESoff, wsENs, wsECur, wsECtl, wsEStat, wsEQ, wsEFeas, wsEU, wsEDone, wsRec,
      wsMsdf;
  int timing = 1;           // 0 none, 1 traversal kernel events, 2 also per-phase events
  bool last_phase_ev = false;
  int last_tr_slot = -1;    // kernel-event slot of the last feasibility batch (-1: none)
  int num_cus = 256;
  kg_params prm = kg_default_params();  // engine parameters (kg_ctx_set_params, kg_autotune)
  void* scratch = nullptr;   // this device's allocation-free scratch (kgl_noalloc_scratch)
  // heterogeneous-batch layout cache
  DevBuf mPlans, mCfgPlan, mQoff, mToff, mIbase;
  std::vector<uint64_t> mSig;
  std::vector<PlanDev> mPlansHost;
};
