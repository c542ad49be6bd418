// C-ABI implementation: handles, geometry pool, world plans and the batched
// feasibility pipeline.  Every public function maps a C++ exception or HIP
// error to a KG_E* code with a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/klampt_gpu.h"
#include "kg_internal.h"
#include "kg_edge.h"
#include "kg_geom_math.h"

hipError_t kgl_fill_int(int* p, int n, int v, hipStream_t s);
int kgl_collide_grid(int exact, int num_cus);



namespace {

thread_local std::string g_err;

struct KgError : std::runtime_error {
  int code;
  KgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw KgError(code, buf);
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(KG_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

template <class F>
int guard(F f) {
  try {
    f();
    return KG_OK;
  } catch (const KgError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return KG_E_LIMIT;
  } catch (const std::exception& e) {
    g_err = e.what();
    return KG_E_INVALID;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) fail(KG_E_INVALID, "%s", msg);
}

// ---------------------------------------------------------------- FP64 math

Xf64 xf_identity() {
  Xf64 X;
  for (int i = 0; i < 9; i++) X.R[i] = (i % 4 == 0) ? 1.0 : 0.0;
  X.t[0] = X.t[1] = X.t[2] = 0;
  return X;
}

Xf64 xf_from(const double* R, const double* t) {
  Xf64 X;
  for (int i = 0; i < 9; i++) X.R[i] = R[i];
  for (int i = 0; i < 3; i++) X.t[i] = t[i];
  return X;
}

Xf64 xf_mul(const Xf64& A, const Xf64& B) {
  Xf64 C;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++)
      C.R[c * 3 + r] = A.R[r] * B.R[c * 3] + A.R[3 + r] * B.R[c * 3 + 1] + A.R[6 + r] * B.R[c * 3 + 2];
  for (int r = 0; r < 3; r++) C.t[r] = A.R[r] * B.t[0] + A.R[3 + r] * B.t[1] + A.R[6 + r] * B.t[2] + A.t[r];
  return C;
}

Xf64 joint_xf(int type, const double* w, double q) {
  Xf64 J = xf_identity();
  if (type == 1) {
    J.t[0] = w[0] * q; J.t[1] = w[1] * q; J.t[2] = w[2] * q;
    return J;
  }
  double c = std::cos(q), s = std::sin(q), v = 1.0 - c;
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) J.R[col * 3 + row] = w[row] * w[col] * v;
  J.R[0] += c; J.R[4] += c; J.R[8] += c;
  J.R[1] += s * w[2]; J.R[2] -= s * w[1];
  J.R[3] -= s * w[2]; J.R[5] += s * w[0];
  J.R[6] += s * w[1]; J.R[7] -= s * w[0];
  return J;
}

void robot_fk(const kg_robot_s* r, const double* q, std::vector<Xf64>& T) {
  T.resize(r->n);
  for (int l = 0; l < r->n; l++) {
    const LinkDesc& L = r->links[l];
    Xf64 T0;
    for (int i = 0; i < 9; i++) T0.R[i] = L.T0[i];
    for (int i = 0; i < 3; i++) T0.t[i] = L.T0[9 + i];
    Xf64 Lx = xf_mul(T0, joint_xf(L.jtype, L.axis, q[l]));
    T[l] = L.parent < 0 ? Lx : xf_mul(T[L.parent], Lx);
  }
}

void world_aabb(const GeomDesc& g, const Xf64& X, double* lo, double* hi) {
  for (int r = 0; r < 3; r++) {
    double wc = X.R[r] * g.lc[0] + X.R[3 + r] * g.lc[1] + X.R[6 + r] * g.lc[2] + X.t[r];
    double ext = std::fabs(X.R[r]) * g.le[0] + std::fabs(X.R[3 + r]) * g.le[1] + std::fabs(X.R[6 + r]) * g.le[2];
    lo[r] = wc - ext;
    hi[r] = wc + ext;
  }
}

double norm3(const double* t) { return std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]); }

// ---------------------------------------------------------------- pool

// Tables (geometry pool, plans, batch layouts) change only when the scene or
// the batch shape changes.  Kernels queued earlier on any stream may still
// read the old contents, so the device is drained before the copy.
template <class T>
void upload(DevBuf& b, const std::vector<T>& v) {
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  hip_check(b.ensure(std::max<size_t>(sizeof(T), v.size() * sizeof(T))), "hipMalloc(pool)");
  if (!v.empty()) hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(pool)");
}

// Every call of a context shares its workspaces, so all of its GPU work forms
// one sequence whatever stream each call names.  A call on a stream other
// than the previous call's first makes its stream wait for that call's end
// event; when the call returns (or throws) its own end is recorded.  The
// event, not the stream, is kept, so a caller may destroy a stream after use.
struct StreamScope {
  kg_ctx_s* c;
  hipStream_t s;
  StreamScope(kg_ctx_s* c_, hipStream_t s_) : c(c_), s(s_) {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    if (!c->evDone)
      hip_check(hipEventCreateWithFlags(&c->evDone, hipEventDisableTiming | hipEventDisableSystemFence),
                "hipEventCreate");
    if (c->done_valid && c->done_stream != s) hip_check(hipStreamWaitEvent(s, c->evDone, 0), "hipStreamWaitEvent");
    // feasibility batches still running on other streams' slots (their
    // buffers are their own, but a plain call may read their status words or
    // replace tables they read)
    for (auto& f : c->fslot)
      if (f.valid) hip_check(hipStreamWaitEvent(s, f.done, 0), "hipStreamWaitEvent");
    for (auto& e : c->eslot)
      if (e.f.valid) hip_check(hipStreamWaitEvent(s, e.f.done, 0), "hipStreamWaitEvent");
  }
  ~StreamScope() {
    if (hipEventRecord(c->evDone, s) == hipSuccess) {
      c->done_stream = s;
      c->done_valid = true;
      c->last_slotted = false;
    }
  }
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;
};


void sync_pool(kg_ctx_s* c) {
  if (!c->pool_dirty) return;
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  // batches on other streams may still read the pool being replaced
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  upload(c->dnodes, c->nodes);
  upload(c->dtv0, c->tv0);
  upload(c->dtv1, c->tv1);
  upload(c->dtv2, c->tv2);
  upload(c->dtv64, c->tv64);
  upload(c->dgeoms, c->geoms);
  upload(c->dtrisrc, c->trisrc);
  c->pool.nodes = c->dnodes.as<ObbNode>();
  c->pool.tv0 = c->dtv0.as<float4>();
  c->pool.tv1 = c->dtv1.as<float4>();
  c->pool.tv2 = c->dtv2.as<float4>();
  c->pool.tv64 = c->dtv64.as<double>();
  c->pool.geoms = c->dgeoms.as<GeomDesc>();
  c->pool.trisrc = c->dtrisrc.as<int>();
  c->pool_dirty = false;
  c->pool_version++;
}

// Geometries are shared like Klampt's shared_ptr<AnyCollisionGeometry3D>:
// the handle and every slot holding the geometry (robot link, terrain, rigid
// object) own a reference; the last release frees it.
void geom_retain(kg_geom_s* g) {
  if (g) g->refs++;
}
void geom_release(kg_geom_s* g) {
  if (!g || --g->refs > 0) return;
  g->dvals.release();
  g->dpts.release();
  g->dv64.release();
  g->dv8.release();
  // mesh data stays in the context pool until the context is destroyed
  delete g;
}

bool geom_empty(const kg_geom_s* g) {
  return !g || g->kind != 0 || g->ctx->geoms[g->index].ntris == 0;
}

// ---------------------------------------------------------------- world mask

int robot_id(const kg_world_s* w, int r) {
  int id = (int)w->terrains.size() + (int)w->objects.size();
  for (int i = 0; i < r; i++) id += w->robots[i]->n + 1;
  return id;
}
int link_id(const kg_world_s* w, int r, int l) { return robot_id(w, r) + 1 + l; }
int num_ids(const kg_world_s* w) {
  int n = (int)w->terrains.size() + (int)w->objects.size() + (int)w->robots.size();
  for (auto* r : w->robots) n += r->n;
  return n;
}

struct Override {
  int a, b, en;
};
std::vector<std::pair<const kg_world_s*, std::vector<Override>>> g_overrides;  // per world

std::vector<Override>& overrides(const kg_world_s* w) {
  for (auto& p : g_overrides)
    if (p.first == w) return p.second;
  g_overrides.push_back({w, {}});
  return g_overrides.back().second;
}

// WorldPlannerSettings::InitializeDefault (PlannerSettings.cpp:16-41)
void rebuild_mask(kg_world_s* w) {
  int n = num_ids(w);
  w->nids = n;
  w->mask.assign((size_t)n * n, 1);
  for (int i = 0; i < n; i++) w->mask[(size_t)i * n + i] = 0;
  for (size_t r = 0; r < w->robots.size(); r++) {
    int k = robot_id(w, (int)r);
    w->mask[(size_t)k * n + k] = 1;
    int base = link_id(w, (int)r, 0);
    const kg_robot_s* R = w->robots[r];
    for (int j = 0; j < R->n; j++) {
      w->mask[(size_t)(base + j) * n + k] = 0;
      w->mask[(size_t)k * n + base + j] = 0;
    }
    for (int j = 0; j < R->n; j++)
      for (int l = 0; l < R->n; l++) w->mask[(size_t)(base + j) * n + base + l] = R->selfcol[j * R->n + l];
    for (int j = 0; j < R->n; j++)
      if (R->links[j].parent == -1)
        for (size_t t = 0; t < w->terrains.size(); t++) {
          w->mask[(size_t)(base + j) * n + t] = 0;
          w->mask[(size_t)t * n + base + j] = 0;
        }
  }
  for (auto& o : overrides(w)) w->mask[(size_t)o.a * n + o.b] = (uint8_t)o.en;
}

inline bool en(const kg_world_s* w, int a, int b) { return w->mask[(size_t)a * w->nids + b] != 0; }

int world_version(const kg_world_s* w) {
  int v = w->version * 7919;
  for (auto* r : w->robots) v = v * 31 + r->version;
  return v;
}

// Build (or reuse) the device tables for checking robot `ri` against the rest
// of the world.
RobotPlan& get_plan(kg_world_s* w, int ri) {
  need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
  kg_ctx_s* c = w->ctx;
  sync_pool(c);
  if (w->plans.size() < w->robots.size()) w->plans.resize(w->robots.size());
  RobotPlan& P = w->plans[ri];
  int ver = world_version(w) * 131 + c->pool_version;
  if (P.built_version == ver) return P;
  // batches on other streams may still read the tables being replaced
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  rebuild_mask(w);
  kg_robot_s* R = w->robots[ri];
  // links
  std::vector<LinkDesc> links = R->links;
  for (int l = 0; l < R->n; l++) {
    links[l].geom = geom_empty(R->link_geoms[l]) ? -1 : R->link_geoms[l]->index;
    links[l].wid = link_id(w, ri, l);
  }
  // env items: GetGeometries over terrains, rigid objects, other robots
  P.env.clear();
  auto add_env = [&](kg_geom_s* g, const Xf64& X, int wid) {
    if (geom_empty(g)) return;
    EnvItem E;
    for (int i = 0; i < 9; i++) E.X[i] = X.R[i];
    for (int i = 0; i < 3; i++) E.X[9 + i] = X.t[i];
    world_aabb(c->geoms[g->index], X, E.lo, E.hi);
    E.geom = g->index;
    E.wid = wid;
    P.env.push_back(E);
  };
  for (size_t t = 0; t < w->terrains.size(); t++) add_env(w->terrains[t], w->terrains[t]->X, (int)t);
  for (size_t o = 0; o < w->objects.size(); o++)
    add_env(w->objects[o].first, w->objects[o].second, (int)(w->terrains.size() + o));
  for (size_t r2 = 0; r2 < w->robots.size(); r2++) {
    if ((int)r2 == ri) continue;
    kg_robot_s* R2 = w->robots[r2];
    for (int l = 0; l < R2->n; l++) add_env(R2->link_geoms[l], R2->Tw[l], link_id(w, (int)r2, l));
  }
  P.nenv = (int)P.env.size();
  need(P.nenv < 32768, "too many environment geometries (limit 32767)");
  P.lemask.assign((size_t)R->n * std::max(1, P.nenv), 0);
  for (int l = 0; l < R->n; l++) {
    if (links[l].geom < 0) continue;
    for (int e = 0; e < P.nenv; e++) {
      int a = links[l].wid, b = P.env[e].wid;
      P.lemask[(size_t)l * P.nenv + e] = (en(w, a, b) || en(w, b, a)) ? 1 : 0;
    }
  }
  // self pairs: (i,j) i<j over links with geometry, mask (i,j)||(i,i)
  // (PlannerSettings.cpp:258)
  P.selfpairs.clear();
  for (int i = 0; i < R->n; i++) {
    if (links[i].geom < 0) continue;
    for (int j = i + 1; j < R->n; j++) {
      if (links[j].geom < 0) continue;
      int a = links[i].wid, b = links[j].wid;
      if (en(w, a, b) || en(w, a, a)) {
        P.selfpairs.push_back(i);
        P.selfpairs.push_back(j);
      }
    }
  }
  P.nself = (int)P.selfpairs.size() / 2;
  // fixed pair list for distance queries
  P.pairs.clear();
  P.pair_a.clear();
  P.pair_b.clear();
  P.pair_kind.clear();
  for (int l = 0; l < R->n; l++)
    for (int e = 0; e < P.nenv; e++)
      if (links[l].geom >= 0 && P.lemask[(size_t)l * P.nenv + e]) {
        P.pairs.push_back({links[l].wid, P.env[e].wid});
        P.pair_a.push_back(l);
        P.pair_b.push_back(e);
        P.pair_kind.push_back(0);
      }
  for (int k = 0; k < P.nself; k++) {
    int i = P.selfpairs[2 * k], j = P.selfpairs[2 * k + 1];
    P.pairs.push_back({links[i].wid, links[j].wid});
    P.pair_a.push_back(i);
    P.pair_b.push_back(j);
    P.pair_kind.push_back(1);
  }
  // error bounds of the FP32 narrow phase
  double reach = 0, Pmax = 0;
  for (int l = 0; l < R->n; l++) {
    reach += norm3(links[l].T0 + 9);
    if (links[l].jtype == 1) reach += std::max(std::fabs(R->qmin[l]), std::fabs(R->qmax[l]));
    if (links[l].geom >= 0) Pmax = std::max(Pmax, c->geoms[links[l].geom].rmax);
  }
  reach += Pmax;
  double L = 2 * reach + Pmax;
  int dA = 0, dB = 0;
  for (int l = 0; l < R->n; l++)
    if (links[l].geom >= 0) dA = std::max(dA, c->geoms[links[l].geom].depth);
  for (auto& E : P.env) {
    L = std::max(L, norm3(E.X + 9) + reach + Pmax);
    L = std::max(L, c->geoms[E.geom].rmax);
    dB = std::max(dB, c->geoms[E.geom].depth);
  }
  const double u = std::ldexp(1.0, -24);
  double ev = 2 * u * (10 * Pmax + 5 * L);
  P.tol32.ev = (float)ev;
  P.tol32.tau = (float)(2 * ev);
  P.tol32.pad_bv = (float)(2 * ev + 32 * u * L);
  P.tol32.pad_rel = 2e-6f;
  P.tol64.ev = 0;
  P.tol64.tau = 0;
  P.tol64.pad_bv = (float)(1e-9 + 1e-12 * L);
  P.tol64.pad_rel = 2e-6f;
  P.maxdepth_a = dA;
  P.maxdepth_b = std::max(dA, dB);
  // device tables
  std::vector<double> qmin = R->qmin, qmax = R->qmax;
  upload(P.dlinks, links);
  upload(P.dqmin, qmin);
  upload(P.dqmax, qmax);
  upload(P.denv, P.env);
  upload(P.dlemask, P.lemask);
  upload(P.dself, P.selfpairs);
  upload(P.dpa, P.pair_a);
  upload(P.dpb, P.pair_b);
  upload(P.dpk, P.pair_kind);
  P.dev.links = P.dlinks.as<LinkDesc>();
  P.dev.qmin = P.dqmin.as<double>();
  P.dev.qmax = P.dqmax.as<double>();
  P.dev.env = P.denv.as<EnvItem>();
  P.dev.lemask = P.dlemask.as<uint8_t>();
  P.dev.selfpairs = P.dself.as<int>();
  P.dev.nlinks = R->n;
  P.dev.nenv = P.nenv;
  P.dev.nself = P.nself;
  {
    // parents precede their children
    std::vector<int> dep(R->n);
    P.dev.depth = 0;
    for (int l = 0; l < R->n; l++) {
      dep[l] = R->links[l].parent < 0 ? 1 : dep[R->links[l].parent] + 1;
      P.dev.depth = std::max(P.dev.depth, dep[l]);
    }
  }
  P.built_version = ver;
  return P;
}

// Device tables of a world's plans.
void release_plans(kg_world_s* w) {
  for (auto& p : w->plans) {
    DevBuf* bufs[] = {&p.dlinks, &p.dqmin, &p.dqmax, &p.denv, &p.dlemask, &p.dself, &p.dpa, &p.dpb, &p.dpk};
    for (auto* b : bufs) b->release();
  }
}

struct BatchStats {
  int64_t redo = 0, items = 0, spills = 0;
  int64_t t_prep_ns = 0, t_trav32_ns = 0, t_trav64_ns = 0, box_tests = 0, tri_tests = 0, uncertain = 0, xpairs = 0;
};

// Status words of one batch (device, zeroed by the prep kernel):
//   [0] spill overflow  [1] FP64 re-check count  [2] FP32 chunk counter
//   [3] FP64 chunk counter  [4] undecided triangle pairs recorded for the
//   exact test  [10] uncertain results (work dropped on overflow);
// then the traversal's chunk counters, one per 128-byte line so that waves
// fetching from different ranges do not queue on one address, a "drained"
// flag per range (and one for the FP64 pass), again one per line, and
// KG_CNT_COPIES lines of u64 statistics (box tests, triangle tests, pair
// items), each wave adding to the copy blockIdx % KG_CNT_COPIES
enum { ST_SPILL = 0, ST_REDO = 1, ST_CHUNK32 = 2, ST_CHUNK64 = 3, ST_XPAIR = 4, ST_UNC = 10, ST_DYN = 32, ST_NSUB = 32,
       ST_DRAIN = ST_DYN + 32 * ST_NSUB, ST_DRAIN64 = ST_DRAIN + 32 * ST_NSUB, ST_CNT = ST_DRAIN64 + 32,
       ST_WORDS = ST_CNT + 32 * KG_CNT_COPIES };

// The re-check list of a batch of N configurations (N + 16 ints), then, at a
// 16-byte boundary, KG_XPAIR_CAP records of undecided triangle pairs for the
// FP64 pass's exact test.
inline size_t redo_head(size_t N) { return (N + 16 + 3) & ~(size_t)3; }
inline size_t redo_ints(size_t N) { return redo_head(N) + 4 * (size_t)KG_XPAIR_CAP; }
void ensure_status_host(kg_ctx_s* c) {
  if (!c->hstatus) {
    hip_check(hipHostMalloc((void**)&c->hstatus, ST_WORDS * sizeof(int), hipHostMallocDefault), "hipHostMalloc");
    // The events of a batch only order work on this device (stream waits)
    // or time it; none needs the system-scope fence a default event does
    // (an L2 writeback and invalidate on every record).  Host reads of
    // device results go through copies, which order themselves.
    constexpr unsigned kTimed = hipEventDisableSystemFence;
    constexpr unsigned kOrder = hipEventDisableTiming | hipEventDisableSystemFence;
    for (auto& e : c->ev) hip_check(hipEventCreateWithFlags(&e, kTimed), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&c->evStat, kOrder), "hipEventCreate");
    for (int i = 0; i < kg_ctx_s::TRING; i++) {
      hip_check(hipEventCreateWithFlags(&c->tr_beg[i], kTimed), "hipEventCreate");
      hip_check(hipEventCreateWithFlags(&c->tr_end[i], kTimed), "hipEventCreate");
    }
  }
}

// Status words of the last batch, copied to the host on request (waits for
// that batch; a batch issued later re-records evStat and stat_src).
const int* fetch_status(kg_ctx_s* c) {
  ensure_status_host(c);
  if (!c->stat_src) {
    std::memset(c->hstatus, 0, ST_WORDS * sizeof(int));
    return c->hstatus;
  }
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  hip_check(hipEventSynchronize(c->evStat), "sync");
  hip_check(hipMemcpy(c->hstatus, c->stat_src, ST_WORDS * sizeof(int), hipMemcpyDeviceToHost), "copy status");
  return c->hstatus;
}

// Layout of one batched feasibility call: a single robot plan, or one plan
// per configuration (heterogeneous batches, device tables in `multi`).
struct Layout {
  const RobotPlan* P = nullptr;     // single-plan batches
  bool multi = false;
  const PlanDev* plans = nullptr;   // device, heterogeneous batches
  const int* cfg_plan = nullptr;
  const int* qoff = nullptr;
  const int* toff = nullptr;
  const int* ibase = nullptr;
  size_t nrows = 0;                 // frame rows (links) over the batch
  size_t ncap = 0;                  // item capacity over the batch
  int cap = 0;                      // per-configuration capacity (single)
  int maxcap = 0;                   // largest per-configuration capacity
  int maxlinks = 0;                 // largest link count over the plans
  int maxdepth = 0;                 // longest root-to-link chain over the plans
  Tol tol32{}, tol64{};
};

// Workspaces of one batch given by the caller (the allocation-free entry
// point's scratch) instead of the context's grow-only buffers.  Only the
// fixed-capacity layout with the per-link broad phase uses them.
struct FeasWs {
  int* status;       // ST_WORDS
  uint8_t* jl;       // N
  int* counts;       // N + 1
  int* redo;         // redo_ints(N): re-check list, then the undecided pairs
  PairItem* items;   // N x cap
  ItemRec* recs;     // N x cap
  uint64_t* spill;   // spill region, split evenly over the traversal waves
  int spill_entries; // its size in entries
};

// N_dev (nullable): device-side configuration count, at most N (the capacity
// the workspaces are sized for); timed = false records no timing events;
// ws (nullable): caller-given workspaces, then nothing is allocated.
void feasible_core(kg_ctx_s* c, const Layout& L, const float* Qf, const double* Qd, int N, uint8_t* feas,
                   uint8_t* env, uint8_t* self, int32_t* fp, hipStream_t s, const int* N_dev = nullptr,
                   bool timed = true, const FeasWs* ws = nullptr) {
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  ensure_status_host(c);
  c->last_N = N;
  c->has_batch = true;
  if (N <= 0) {
    std::memset(c->hstatus, 0, ST_WORDS * sizeof(int));
    return;
  }
  const int mode = fp ? KG_MODE_ORDERED : ((env || self) ? KG_MODE_GROUPS : KG_MODE_ANY);
  const bool fixed = L.ncap * sizeof(PairItem) <= (2048ull << 20);
  need(fixed || !N_dev, "batch with a device-side count exceeds the item capacity limit");
  // robots of at most 32 links get the lane-per-link broad phase, which
  // also writes the item start records; otherwise the traversal forms each
  // item's frame when it starts it
  const int ml = L.multi ? L.maxlinks : L.P->dev.nlinks;
  int group = ml <= 8 ? 8 : (ml <= 16 ? 16 : (ml <= 32 ? 32 : 0));
  if (!c->prm.prep_group || !(fixed || L.multi)) group = 0;
  // prep_group 8 / 16 / 32: at least that many lanes per configuration
  if (group && c->prm.prep_group > group) group = c->prm.prep_group;
  // frames are stored only for the serial broad phase (the per-link one
  // writes start records, and the FP64 pass forms its frames from q); link
  // boxes only for the count / scan path
  const bool lds_links = !group && !L.multi && ml <= 8 && (fixed || L.multi);
  need(!ws || (fixed && group), "caller workspaces need the fixed-capacity, per-link broad phase");
  if (!ws) {
    if (!group) hip_check(c->wsT64.ensure(L.nrows * 12 * sizeof(double)), "hipMalloc(T64)");
    if (!group && !lds_links) hip_check(c->wsLB.ensure(L.nrows * 6 * sizeof(double)), "hipMalloc(LB)");
    hip_check(c->wsJL.ensure((size_t)N), "hipMalloc(jl)");
    hip_check(c->wsCounts.ensure((size_t)(N + 1) * sizeof(int)), "hipMalloc(counts)");
    hip_check(c->wsOffsets.ensure((size_t)(N + 1) * sizeof(int)), "hipMalloc(offsets)");
    hip_check(c->wsStatus.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
    hip_check(c->wsRedo.ensure(redo_ints(N) * sizeof(int)), "hipMalloc(redo)");
  }
  int* status = ws ? ws->status : c->wsStatus.as<int>();
  int* redo_list = ws ? ws->redo : c->wsRedo.as<int>();
  PrepArgs pa{};
  pa.Qf = Qf;
  pa.Qd = Qd;
  pa.N = N;
  pa.mode = mode;
  if (L.P) pa.plan = L.P->dev;
  pa.pool = c->pool;
  pa.T64 = ws ? nullptr : c->wsT64.as<double>();
  pa.t_rows = L.nrows;
  pa.LB = ws ? nullptr : c->wsLB.as<double>();
  pa.jl = ws ? ws->jl : c->wsJL.as<uint8_t>();
  pa.counts = ws ? ws->counts : c->wsCounts.as<int>();
  pa.reset = status;
  pa.nreset = ST_WORDS;
  pa.plans = L.plans;
  pa.cfg_plan = L.cfg_plan;
  pa.qoff = L.qoff;
  pa.toff = L.toff;
  pa.ibase = L.ibase;
  pa.N_dev = N_dev;
  {
    const int depth = L.multi ? L.maxdepth : L.P->dev.depth;
    while ((1 << pa.jump_rounds) < depth) pa.jump_rounds++;
  }
  // timing events cost a few microseconds each on the stream: per-phase ones
  // only at timing level 2, the traversal kernel's own at level 1 and up
  const bool phase_ev = timed && c->timing >= 2, kern_ev = timed && c->timing >= 1;
  c->last_phase_ev = phase_ev;
  c->last_tr_slot = -1;
  if (phase_ev) hip_check(hipEventRecord(c->ev[0], s), "hipEventRecord");
  if (fixed || L.multi) {
    need(fixed, "heterogeneous batch exceeds the item capacity limit (2 GiB of pair items)");
    if (!ws) hip_check(c->wsItems.ensure(std::max<size_t>(1, L.ncap) * sizeof(PairItem)), "hipMalloc(items)");
    pa.items = ws ? ws->items : c->wsItems.as<PairItem>();
    pa.cap = L.cap;
    pa.offsets = nullptr;
    pa.group = group;
    pa.lds_links = lds_links;
    if (pa.group) {
      if (!ws) hip_check(c->wsRec.ensure(std::max<size_t>(1, L.ncap) * sizeof(ItemRec)), "hipMalloc(records)");
      pa.recs = ws ? ws->recs : c->wsRec.as<ItemRec>();
      // with start records nothing reads the frames (the FP64 pass forms
      // its own from q)
      pa.T64 = nullptr;
    }
    hip_check(kgl_prep_single(pa, s), "prep");
  } else {
    // large environments: count, scan, then fill exactly
    pa.offsets = c->wsOffsets.as<int>();
    hip_check(kgl_fill_int(status, ST_WORDS, 0, s), "clear status");
    pa.nreset = 0;
    hip_check(kgl_prep_count(pa, s), "prep_count");
    hip_check(kgl_scan(pa.counts, pa.offsets, N + 1, c->wsScanTmp, s), "scan");
    int t = 0;
    hip_check(hipMemcpyAsync(&t, pa.offsets + N, sizeof(int), hipMemcpyDeviceToHost, s), "copy total");
    hip_check(hipStreamSynchronize(s), "sync");
    hip_check(c->wsItems.ensure(std::max<size_t>(1, (size_t)t) * sizeof(PairItem)), "hipMalloc(items)");
    pa.items = c->wsItems.as<PairItem>();
    hip_check(kgl_prep_fill(pa, s), "prep_fill");
  }
  if (phase_ev) hip_check(hipEventRecord(c->ev[1], s), "hipEventRecord");

  // defaults measured on 65,536 configurations of the 7-DOF arm against the
  // 200k-triangle clutter (kg_default_params)
  const kg_params& prm = c->prm;
  const int scap = prm.scap == 256 ? 256 : 512;
  const int wpc = prm.waves_per_cu;
  const int grid32 = wpc > 0 ? wpc * c->num_cus : kgl_collide_grid(0, c->num_cus, scap);
  const int grid64 = std::min(kgl_collide_grid(1, c->num_cus, scap), std::max(1, c->num_cus));
  const int spill_cap = ws ? ws->spill_entries / std::max(grid32, grid64) : prm.spill_cap;
  if (!ws)
    hip_check(c->wsSpill.ensure(std::max<size_t>(1, (size_t)std::max(grid32, grid64) * spill_cap) *
                                sizeof(uint64_t)),
              "hipMalloc(spill)");

  TravArgs ta{};
  ta.pool = c->pool;
  if (L.P) ta.plan = L.P->dev;
  ta.T64 = pa.T64;
  ta.t_rows = L.nrows;
  ta.t_n = N;
  ta.items = pa.items;
  ta.offsets = pa.offsets;
  ta.counts = pa.counts;
  ta.cap = pa.cap;
  ta.cfg_list = nullptr;
  ta.ncfg = N;
  ta.ncfg_dev = N_dev;
  ta.mode = mode;
  ta.tol = L.tol32;
  ta.chunk_counter = status + ST_DYN;
  ta.chunk_drained = status + ST_DRAIN;
  {
    const int nsub = std::max(1, std::min((int)ST_NSUB, prm.nsub));
    ta.nsub = N >= 64 * nsub ? nsub : 1;
  }
  ta.spill = ws ? ws->spill : c->wsSpill.as<uint64_t>();
  ta.spill_cap = spill_cap;
  ta.stack_cap = 0;
  ta.status = status + ST_SPILL;
  ta.cpw = prm.cpw;
  ta.leaf_min = prm.leaf_min;
  ta.tail_lanes = prm.tail_lanes;
  ta.tail_reps = prm.tail_reps;
  ta.grid = grid32;
  ta.counters = (unsigned long long*)(status + ST_CNT);
  ta.plans = L.plans;
  ta.cfg_plan = L.cfg_plan;
  ta.toff = L.toff;
  ta.ibase = L.ibase;
  ta.scap = scap;
  ta.static_pct = prm.static_pct;
  ta.recs = pa.recs;
  ta.Qf = Qf;
  ta.Qd = Qd;
  ta.qoff = L.qoff;
  // results are assembled inside the traversal as each configuration's last
  // item finishes; uncertain ones go to the re-check list
  FinArgs& fa = ta.fin;
  fa.N = N;
  fa.mode = mode;
  fa.jl = pa.jl;
  fa.items = pa.items;
  if (L.P) fa.plan = L.P->dev;
  fa.plans = L.plans;
  fa.cfg_plan = L.cfg_plan;
  fa.feasible = feas;
  fa.env_hit = env;
  fa.self_hit = self;
  fa.first_pair = fp;
  fa.redo_list = redo_list;
  fa.redo_count = status + ST_REDO;
  fa.final_pass = 0;
  fa.force_recheck = prm.force_recheck;
  fa.uncertain = status + ST_UNC;
  // triangle pairs the FP32 test leaves undecided are decided one by one by
  // the FP64 pass's exact test, which can only add hits (any-hit and group
  // modes); the ordered first-pair mode traverses such a configuration again
  if (mode != KG_MODE_ORDERED) {
    fa.xpairs = reinterpret_cast<int4*>(redo_list + redo_head(N));
    fa.xpair_count = status + ST_XPAIR;
    fa.xpair_cap = KG_XPAIR_CAP;
  }
  const int slot = (int)(c->tr_count % kg_ctx_s::TRING);
  if (kern_ev) hip_check(hipEventRecord(c->tr_beg[slot], s), "hipEventRecord");
  hip_check(kgl_collide_f32(ta, s), "collide_f32");
  if (kern_ev) {
    hip_check(hipEventRecord(c->tr_end[slot], s), "hipEventRecord");
    c->last_tr_slot = slot;
    c->tr_count++;
  }

  // FP64 re-check of the configurations the FP32 pass could not certify;
  // the count stays on the device (waves exit at once when it is zero)
  TravArgs tb = ta;
  tb.cfg_list = redo_list;
  tb.ncfg_dev = status + ST_REDO;
  tb.tol = L.tol64;
  tb.cpw = 1;
  tb.grid = grid64;
  tb.counters = nullptr;
  tb.chunk_counter = status + ST_CHUNK64;
  tb.chunk_drained = status + ST_DRAIN64;
  tb.static_pct = 0;
  tb.nsub = 1;
  tb.fin.final_pass = 1;
  if (phase_ev) hip_check(hipEventRecord(c->ev[4], s), "hipEventRecord");
  hip_check(kgl_collide_f64(tb, s), "collide_f64");
  if (phase_ev) hip_check(hipEventRecord(c->ev[5], s), "hipEventRecord");
  // the status words stay on the device until someone asks (fetch_status)
  c->stat_src = status;
  hip_check(hipEventRecord(c->evStat, s), "hipEventRecord");
}

int active_links(const kg_robot_s* r) {
  int nact = 0;
  for (int l = 0; l < r->n; l++)
    if (r->link_geoms[l] && !geom_empty(r->link_geoms[l])) nact++;
  return nact;
}

// Batched feasibility of one robot on device buffers (async on stream s).
// Collision margins turn a collide into a distance-threshold test; the
// batched traversal decides intersection only, so batches reject them.
void reject_margins(const kg_world_s* w) {
  auto has = [](const kg_geom_s* g) { return g && g->margin > 0; };
  bool any = false;
  for (auto* g : w->terrains) any |= has(g);
  for (auto& o : w->objects) any |= has(o.first);
  for (auto* r : w->robots)
    for (auto* g : r->link_geoms) any |= has(g);
  if (any)
    fail(KG_E_UNSUPPORTED,
         "collision margins are supported by single-pair queries (collide, withinDistance, distance), not by "
         "batched feasibility");
}

void feasible_dev(kg_world_s* w, int ri, const float* Qf, const double* Qd, int N, uint8_t* feas, uint8_t* env,
                  uint8_t* self, int32_t* fp, hipStream_t s) {
  reject_margins(w);
  RobotPlan& P = get_plan(w, ri);
  Layout L;
  L.P = &P;
  L.cap = active_links(w->robots[ri]) * P.nenv + P.nself;
  L.maxcap = L.cap;
  L.nrows = (size_t)std::max(N, 0) * P.dev.nlinks;
  L.ncap = (size_t)L.cap * std::max(N, 0);
  L.tol32 = P.tol32;
  L.tol64 = P.tol64;
  feasible_core(w->ctx, L, Qf, Qd, N, feas, env, self, fp, s);
}

// Read the status of the last batch (waits for it).
void read_stats(kg_world_s* w, BatchStats& st) {
  kg_ctx_s* c = w->ctx;
  st = BatchStats{};
  if (!c->hstatus || !c->has_batch || c->last_N <= 0) return;
  const int* h = fetch_status(c);
  st.spills = h[ST_SPILL];
  st.redo = h[ST_REDO];
  st.uncertain = h[ST_UNC];
  st.xpairs = h[ST_XPAIR];
  for (int k = 0; k < KG_CNT_COPIES; k++) {
    unsigned long long cnt[3];
    std::memcpy(cnt, h + ST_CNT + 32 * k, sizeof(cnt));
    st.box_tests += (int64_t)cnt[0];
    st.tri_tests += (int64_t)cnt[1];
    st.items += (int64_t)cnt[2];
  }
  float ms = 0;
  if (c->last_tr_slot >= 0) {
    if (hipEventElapsedTime(&ms, c->tr_beg[c->last_tr_slot], c->tr_end[c->last_tr_slot]) == hipSuccess)
      st.t_trav32_ns = (int64_t)(ms * 1e6);
    else
      (void)hipGetLastError();
  }
  if (!c->last_phase_ev) return;
  // feasibility batches record ev[0,1,4,5]; distance batches ev[2,3]
  auto span = [&](hipEvent_t a, hipEvent_t b, int64_t& out) {
    if (hipEventElapsedTime(&ms, a, b) == hipSuccess)
      out = (int64_t)(ms * 1e6);
    else
      (void)hipGetLastError();  // leave no error for a later launch check
  };
  if (c->last_tr_slot >= 0) {
    span(c->ev[0], c->ev[1], st.t_prep_ns);
    span(c->ev[4], c->ev[5], st.t_trav64_ns);
  } else {
    span(c->ev[2], c->ev[3], st.t_trav32_ns);
  }
}

const char* kOverflowMsg =
    "BVH traversal dropped work when its HBM spill region filled (kg_params.spill_cap); %d result(s) uncertain";

// Host-pointer calls: results that dropped work left uncertain are an error
// (the _dev calls report them through kg_ctx_status_dev instead).
void check_uncertain(kg_ctx_s* c) {
  const int unc = fetch_status(c)[ST_UNC];
  if (unc > 0) fail(KG_E_LIMIT, kOverflowMsg, unc);
}

void feasible_sync_check(kg_world_s* w) { check_uncertain(w->ctx); }

kg_world_s* solo_world(kg_robot_s* r);
void forget_solo_world(const kg_robot_s* r);
void forget_pair_world(const kg_ctx_s* c);

}  // namespace

// ================================================================== C ABI

extern "C" {

const char* kg_last_error(void) { return g_err.c_str(); }
int kg_version(void) { return 1; }

int kg_ctx_create(int device, kg_ctx* out) {
  return guard([&] {
    need(out != nullptr, "out is NULL");
    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    need(device >= 0 && device < ndev, "device index out of range");
    hip_check(hipSetDevice(device), "hipSetDevice");
    kg_ctx_s* c = new kg_ctx_s();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      hip_check(e, "hipStreamCreate");
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
      c->num_cus = cus;
    // set-up the allocation-free entry point relies on: this device's
    // scratch address, the pinned status words and the events
    try {
      hip_check(kgl_noalloc_scratch(device, &c->scratch), "hipGetSymbolAddress(scratch)");
      ensure_status_host(c);
      hip_check(hipEventCreateWithFlags(&c->evDone, hipEventDisableTiming | hipEventDisableSystemFence),
                "hipEventCreate");
    } catch (...) {
      (void)hipStreamDestroy(c->stream);
      delete c;
      throw;
    }
    *out = c;
  });
}

int kg_ctx_destroy(kg_ctx c) {
  return guard([&] {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    forget_pair_world(c);
    DevBuf* bufs[] = {&c->dnodes, &c->dtv0, &c->dtv1, &c->dtv2, &c->dtv64, &c->dgeoms, &c->dtrisrc,
                      &c->wsQ, &c->wsT64, &c->wsLB, &c->wsJL, &c->wsCounts, &c->wsOffsets, &c->wsItems,
                      &c->wsScanTmp, &c->wsBest, &c->wsUnc, &c->wsChunk, &c->wsSpill, &c->wsStatus, &c->wsRedo,
                      &c->wsOut8a, &c->wsOut8b, &c->wsOut8c, &c->wsOut32, &c->wsF, &c->wsG,
                      &c->wsPoff, &c->wsPts, &c->wsLoss, &c->wsDQ, &c->wsDl, &c->wsEA, &c->wsEB,
                      &c->wsESeg, &c->wsESoff, &c->wsENs, &c->wsECur, &c->wsECtl, &c->wsEStat, &c->wsEQ,
                      &c->wsEFeas, &c->wsEU, &c->wsEDone, &c->wsRec, &c->wsMsdf, &c->mPlans, &c->mCfgPlan,
                      &c->mQoff, &c->mToff, &c->mIbase};
    for (auto* b : bufs) b->release();
    if (c->hstatus) (void)hipHostFree(c->hstatus);
    for (auto& e : c->ev)
      if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < kg_ctx_s::TRING; i++) {
      if (c->tr_beg[i]) (void)hipEventDestroy(c->tr_beg[i]);
      if (c->tr_end[i]) (void)hipEventDestroy(c->tr_end[i]);
    }
    if (c->evStat) (void)hipEventDestroy(c->evStat);
    if (c->evDone) (void)hipEventDestroy(c->evDone);
    for (auto& f : c->fslot) {
      DevBuf* fb[] = {&f.status, &f.jl, &f.counts, &f.redo, &f.items, &f.recs, &f.spill};
      for (auto* b : fb) b->release();
      if (f.done) (void)hipEventDestroy(f.done);
    }
    for (auto& e : c->eslot) {
      DevBuf* eb[] = {&e.f.status, &e.f.jl, &e.f.counts, &e.f.redo, &e.f.items, &e.f.recs, &e.f.spill,
                      &e.ns, &e.cur, &e.done, &e.seg, &e.soff, &e.ctl, &e.q, &e.feas, &e.stat};
      for (auto* b : eb) b->release();
      if (e.f.done) (void)hipEventDestroy(e.f.done);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
  });
}

int kg_ctx_synchronize(kg_ctx c) {
  return guard([&] {
    need(c != nullptr, "ctx is NULL");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    if (c->done_valid) hip_check(hipEventSynchronize(c->evDone), "hipEventSynchronize");
    for (auto& f : c->fslot)
      if (f.valid) hip_check(hipEventSynchronize(f.done), "hipEventSynchronize");
    for (auto& e : c->eslot)
      if (e.f.valid) hip_check(hipEventSynchronize(e.f.done), "hipEventSynchronize");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

int kg_geom_create_trimesh(kg_ctx c, const double* V, int nv, const int32_t* T, int nt, kg_geom* out) {
  return guard([&] {
    need(c && out, "NULL argument");
    need(nv >= 0 && nt >= 0, "negative size");
    need(nt == 0 || (V && T), "NULL vertex or triangle array");
    for (int i = 0; i < 3 * nt; i++) need(T[i] >= 0 && T[i] < nv, "triangle index out of range");
    for (int i = 0; i < 3 * nv; i++) need(std::isfinite(V[i]), "non-finite vertex coordinate");
    kg_geom_s* g = new kg_geom_s();
    g->ctx = c;
    g->kind = 0;
    g->X = xf_identity();
    g->V.assign(V, V + 3 * (size_t)nv);
    g->T.assign(T, T + 3 * (size_t)nt);
    GeomDesc d{};
    d.ntris = nt;
    d.tri0 = (int)c->tv0.size();
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, rmax = 0;
    if (nv > 0)
      for (int k = 0; k < 3; k++) lo[k] = hi[k] = V[k];
    for (int i = 0; i < nv; i++) {
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], V[3 * i + k]);
        hi[k] = std::max(hi[k], V[3 * i + k]);
      }
      rmax = std::max(rmax, norm3(V + 3 * i));
    }
    for (int k = 0; k < 3; k++) {
      d.lc[k] = 0.5 * (lo[k] + hi[k]);
      d.le[k] = 0.5 * (hi[k] - lo[k]);
    }
    d.rmax = rmax;
    if (nt > 0) {
      BuiltTree bt;
      kg_build_obb_tree(V, nv, T, nt, bt);
      need(c->nodes.size() + bt.nodes.size() < (1u << 26), "geometry pool exceeds 2^26 BVH nodes");
      d.root = (int)c->nodes.size();
      d.nnodes = (int)bt.nodes.size();
      d.depth = bt.depth;
      for (auto nd : bt.nodes) {
        nd.child = nd.child >= 0 ? nd.child + d.root : ~(~nd.child + d.tri0);
        c->nodes.push_back(nd);
      }
      for (int li = 0; li < nt; li++) {
        int t = bt.order[li];
        const double* p0 = V + 3 * T[3 * t];
        const double* p1 = V + 3 * T[3 * t + 1];
        const double* p2 = V + 3 * T[3 * t + 2];
        c->tv0.push_back(make_float4((float)p0[0], (float)p0[1], (float)p0[2], 0.f));
        c->tv1.push_back(make_float4((float)p1[0], (float)p1[1], (float)p1[2], 0.f));
        c->tv2.push_back(make_float4((float)p2[0], (float)p2[1], (float)p2[2], 0.f));
        for (int k = 0; k < 3; k++) c->tv64.push_back(p0[k]);
        for (int k = 0; k < 3; k++) c->tv64.push_back(p1[k]);
        for (int k = 0; k < 3; k++) c->tv64.push_back(p2[k]);
        c->trisrc.push_back(t);
      }
    } else {
      d.root = -1;
    }
    g->index = (int)c->geoms.size();
    c->geoms.push_back(d);
    c->pool_dirty = true;
    *out = g;
  });
}

int kg_geom_get_bvh(kg_geom g, float* nodes, int max_nodes, int* nnodes) {
  return guard([&] {
    need(g && nnodes, "NULL argument");
    need(g->kind == 0, "not a triangle mesh");
    kg_ctx_s* c = g->ctx;
    const GeomDesc& d = c->geoms[g->index];
    const int nn = d.root < 0 ? 0 : d.nnodes;
    *nnodes = nn;
    if (!nodes) return;
    need(max_nodes >= nn, "node buffer too small");
    for (int i = 0; i < nn; i++) {
      const ObbNode& o = c->nodes[(size_t)d.root + i];
      float* p = nodes + 16 * (size_t)i;
      for (int k = 0; k < 3; k++) {
        p[k] = o.c[k];
        p[3 + k] = o.e[k];
      }
      for (int k = 0; k < 9; k++) p[6 + k] = o.R[k];
      const int child = o.child >= 0 ? o.child - d.root : ~c->trisrc[~o.child];
      std::memcpy(p + 15, &child, sizeof(int));
    }
  });
}

int kg_geom_create_sdf(kg_ctx c, const double bmin[3], const double bmax[3], const int32_t dims[3],
                       const float* values, kg_geom* out) {
  return guard([&] {
    need(c && bmin && bmax && dims && values && out, "NULL argument");
    for (int k = 0; k < 3; k++) {
      need(dims[k] >= 1, "SDF dims must be >= 1");
      need(bmax[k] > bmin[k], "SDF bmax must exceed bmin");
    }
    size_t nvals = (size_t)dims[0] * dims[1] * dims[2];
    kg_geom_s* g = new kg_geom_s();
    g->ctx = c;
    g->kind = 1;
    g->index = -1;
    g->X = xf_identity();
    for (int k = 0; k < 3; k++) {
      g->dims[k] = dims[k];
      g->bmin[k] = bmin[k];
      g->bmax[k] = bmax[k];
    }
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipError_t e = g->dvals.ensure(nvals * sizeof(float));
    if (e != hipSuccess) {
      delete g;
      hip_check(e, "hipMalloc(sdf)");
    }
    hip_check(hipMemcpy(g->dvals.p, values, nvals * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy(sdf)");
    *out = g;
  });
}

int kg_geom_destroy(kg_geom g) {
  return guard([&] {
    if (!g) return;
    need(g->alive, "geometry handle already destroyed");
    g->alive = false;
    geom_release(g);  // freed now, or when the last robot link / world slot lets go
  });
}

int kg_geom_create_pointcloud(kg_ctx c, const double* P, int n, kg_geom* out) {
  return guard([&] {
    need(c && out, "NULL argument");
    need(n >= 0, "negative size");
    need(n == 0 || P, "NULL point array");
    for (int i = 0; i < 3 * n; i++) need(std::isfinite(P[i]), "non-finite point coordinate");
    kg_geom_s* g = new kg_geom_s();
    g->ctx = c;
    g->kind = 2;
    g->index = -1;
    g->X = xf_identity();
    g->V.assign(P, P + 3 * (size_t)n);
    *out = g;
  });
}

int kg_geom_mesh_to_sdf(kg_geom g, double resolution, int pad, kg_geom* out) {
  return guard([&] {
    need(g && out, "NULL argument");
    need(g->kind == 0, "mesh_to_sdf converts a TriangleMesh geometry");
    need(std::isfinite(resolution) && resolution >= 0, "resolution must be finite and >= 0");
    need(pad >= 0 && pad <= 1024, "pad must be 0..1024 cells");
    kg_ctx_s* c = g->ctx;
    const GeomDesc& gd = c->geoms[g->index];
    need(gd.ntris > 0, "mesh is empty");
    need(gd.depth + 2 <= KG_MSDF_STK, "mesh tree is deeper than the conversion's traversal stack");
    const int nv = (int)(g->V.size() / 3), nt = (int)(g->T.size() / 3);
    double h = resolution;
    if (h == 0) {
      // default: the average triangle diameter (longest edge)
      double sum = 0;
      for (int t = 0; t < nt; t++) {
        double m = 0;
        for (int e = 0; e < 3; e++) {
          const double* p = &g->V[3 * g->T[3 * t + e]];
          const double* q = &g->V[3 * g->T[3 * t + (e + 1) % 3]];
          const double dd[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
          m = std::max(m, norm3(dd));
        }
        sum += m;
      }
      h = sum / nt;
    }
    need(h > 0, "resolution is 0 and the mesh's triangles have no extent");
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) lo[k] = hi[k] = g->V[k];
    for (int i = 0; i < nv; i++)
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], g->V[3 * i + k]);
        hi[k] = std::max(hi[k], g->V[3 * i + k]);
      }
    int dims[3];
    double bmin[3], bmax[3];
    size_t total = 1;
    for (int k = 0; k < 3; k++) {
      const double cells = std::ceil((hi[k] - lo[k]) / h);
      need(cells < 1e6, "grid too large for this resolution");
      dims[k] = std::max(1, (int)cells + 2 * pad);
      bmin[k] = lo[k] - pad * h;
      bmax[k] = bmin[k] + dims[k] * h;
      total *= (size_t)dims[k];
    }
    need(total < (1ull << 31), "grid has more than 2^31 cells");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    sync_pool(c);
    kg_geom_s* s = new kg_geom_s();
    s->ctx = c;
    s->kind = 1;
    s->index = -1;
    s->X = g->X;
    for (int k = 0; k < 3; k++) {
      s->dims[k] = dims[k];
      s->bmin[k] = bmin[k];
      s->bmax[k] = bmax[k];
    }
    hipError_t e = s->dvals.ensure(total * sizeof(float));
    if (e == hipSuccess) e = c->wsMsdf.ensure(total * sizeof(int));
    if (e != hipSuccess) {
      delete s;
      hip_check(e, "hipMalloc(mesh_to_sdf)");
    }
    MeshSdfArgs a{};
    a.pool = c->pool;
    a.root = gd.root;
    for (int k = 0; k < 3; k++) {
      a.dims[k] = dims[k];
      a.bmin[k] = bmin[k];
    }
    a.h = h;
    a.out = s->dvals.as<float>();
    a.diff = c->wsMsdf.as<int>();
    StreamScope scope(c, c->stream);
    e = kgl_fill_int(a.diff, (int)total, 0, c->stream);
    if (e == hipSuccess) e = kgl_msdf_dist(a, c->stream);
    if (e == hipSuccess) e = kgl_msdf_sign(a, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      delete s;
      hip_check(e, "mesh_to_sdf kernels");
    }
    *out = s;
  });
}

int kg_geom_get_sdf(kg_geom g, int32_t dims[3], double bmin[3], double bmax[3], float* values) {
  return guard([&] {
    need(g && dims && bmin && bmax, "NULL argument");
    need(g->kind == 1, "geometry is not an ImplicitSurface");
    for (int k = 0; k < 3; k++) {
      dims[k] = g->dims[k];
      bmin[k] = g->bmin[k];
      bmax[k] = g->bmax[k];
    }
    if (values) {
      const size_t n = (size_t)g->dims[0] * g->dims[1] * g->dims[2];
      hip_check(hipSetDevice(g->ctx->device), "hipSetDevice");
      hip_check(hipMemcpy(values, g->dvals.p, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy(sdf)");
    }
  });
}

int kg_geom_clone(kg_geom src, kg_geom* out) {
  return guard([&] {
    need(src && out, "NULL argument");
    kg_ctx_s* c = src->ctx;
    kg_geom_s* g = new kg_geom_s();
    g->ctx = c;
    g->kind = src->kind;
    g->index = src->index;  // meshes: the same pool descriptor, BVH and triangles
    g->X = src->X;
    g->margin = src->margin;
    g->V = src->V;
    g->T = src->T;
    if (src->kind == 1) {
      for (int k = 0; k < 3; k++) {
        g->dims[k] = src->dims[k];
        g->bmin[k] = src->bmin[k];
        g->bmax[k] = src->bmax[k];
      }
      size_t bytes = (size_t)src->dims[0] * src->dims[1] * src->dims[2] * sizeof(float);
      hipError_t e = hipSetDevice(c->device);
      if (e == hipSuccess) e = g->dvals.ensure(bytes);
      if (e == hipSuccess) e = hipMemcpy(g->dvals.p, src->dvals.p, bytes, hipMemcpyDeviceToDevice);
      if (e != hipSuccess) {
        delete g;
        hip_check(e, "kg_geom_clone(sdf)");
      }
    }
    *out = g;
  });
}

int kg_geom_set_transform(kg_geom g, const double R[9], const double t[3]) {
  return guard([&] {
    need(g && R && t, "NULL argument");
    g->X = xf_from(R, t);
  });
}

int kg_geom_get_transform(kg_geom g, double R[9], double t[3]) {
  return guard([&] {
    need(g && R && t, "NULL argument");
    std::memcpy(R, g->X.R, sizeof(g->X.R));
    std::memcpy(t, g->X.t, sizeof(g->X.t));
  });
}

int kg_geom_set_collision_margin(kg_geom g, double margin) {
  return guard([&] {
    need(g != nullptr, "geometry is NULL");
    need(margin >= 0 && std::isfinite(margin), "margin must be finite and >= 0");
    g->margin = margin;
  });
}

int kg_geom_get_collision_margin(kg_geom g, double* margin) {
  return guard([&] {
    need(g && margin, "NULL argument");
    *margin = g->margin;
  });
}

int kg_geom_get_bb(kg_geom g, double bmin[3], double bmax[3]) {
  return guard([&] {
    need(g && bmin && bmax, "NULL argument");
    if (g->kind == 1) {
      // transformed grid box
      GeomDesc d{};
      for (int k = 0; k < 3; k++) {
        d.lc[k] = 0.5 * (g->bmin[k] + g->bmax[k]);
        d.le[k] = 0.5 * (g->bmax[k] - g->bmin[k]);
      }
      world_aabb(d, g->X, bmin, bmax);
      return;
    }
    if (g->kind == 2) {
      // transformed points
      for (int k = 0; k < 3; k++) {
        bmin[k] = INFINITY;
        bmax[k] = -INFINITY;
      }
      for (size_t i = 0; i + 2 < g->V.size(); i += 3)
        for (int k = 0; k < 3; k++) {
          const double w = g->X.R[k] * g->V[i] + g->X.R[3 + k] * g->V[i + 1] + g->X.R[6 + k] * g->V[i + 2] + g->X.t[k];
          bmin[k] = std::min(bmin[k], w);
          bmax[k] = std::max(bmax[k], w);
        }
      return;
    }
    const GeomDesc& d = g->ctx->geoms[g->index];
    if (d.ntris == 0) {
      for (int k = 0; k < 3; k++) {
        bmin[k] = INFINITY;
        bmax[k] = -INFINITY;
      }
      return;
    }
    world_aabb(d, g->X, bmin, bmax);
  });
}

int kg_robot_create(kg_ctx c, int n, const int32_t* parents, const double* T0, const double* axes,
                    const int32_t* jtype, const double* qmin, const double* qmax, kg_robot* out) {
  return guard([&] {
    need(c && parents && T0 && axes && jtype && qmin && qmax && out, "NULL argument");
    need(n >= 1 && n <= KG_MAX_LINKS, "robot must have 1..64 links");
    kg_robot_s* r = new kg_robot_s();
    r->ctx = c;
    r->n = n;
    r->links.resize(n);
    for (int l = 0; l < n; l++) {
      if (!(parents[l] >= -1 && parents[l] < l)) {
        delete r;
        fail(KG_E_INVALID, "parents must be topologically sorted (parent < index)");
      }
      if (jtype[l] != 0 && jtype[l] != 1) {
        delete r;
        fail(KG_E_INVALID, "joint type must be 0 (revolute) or 1 (prismatic)");
      }
      LinkDesc& L = r->links[l];
      std::memcpy(L.T0, T0 + 12 * l, 12 * sizeof(double));
      std::memcpy(L.axis, axes + 3 * l, 3 * sizeof(double));
      L.parent = parents[l];
      L.jtype = jtype[l];
      L.geom = -1;
      L.wid = l + 1;
    }
    r->link_geoms.assign(n, nullptr);
    r->qmin.assign(qmin, qmin + n);
    r->qmax.assign(qmax, qmax + n);
    r->q.assign(n, 0.0);
    for (int l = 0; l < n; l++) r->q[l] = std::min(std::max(0.0, r->qmin[l]), r->qmax[l]);
    r->selfcol.assign((size_t)n * n, 0);
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++)
        if (parents[j] != i && parents[i] != j) r->selfcol[i * n + j] = 1;
    robot_fk(r, r->q.data(), r->Tw);
    *out = r;
  });
}

int kg_robot_destroy(kg_robot r) {
  return guard([&] {
    if (!r) return;
    forget_solo_world(r);
    for (auto* g : r->link_geoms) geom_release(g);
    r->dlinks.release();
    delete r;
  });
}

int kg_robot_set_link_geom(kg_robot r, int link, kg_geom g) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    need(link >= 0 && link < r->n, "link index out of range");
    need(!g || g->kind == 0, "link geometry must be a triangle mesh");
    need(!g || g->ctx == r->ctx, "geometry belongs to another context");
    need(!g || g->alive, "geometry handle was destroyed");
    geom_retain(g);
    geom_release(r->link_geoms[link]);
    r->link_geoms[link] = g;
    if (g) g->X = r->Tw[link];
    r->version++;
  });
}

int kg_robot_enable_self_collision(kg_robot r, int a, int b, int enabled) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    need(a >= 0 && a < r->n && b >= 0 && b < r->n && a != b, "invalid link pair");
    int i = std::min(a, b), j = std::max(a, b);
    r->selfcol[i * r->n + j] = enabled ? 1 : 0;
    r->version++;
  });
}

int kg_robot_set_self_pairs(kg_robot r, const int32_t* pairs, int npairs) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    need(npairs >= 0 && (npairs == 0 || pairs), "NULL pair list");
    for (int k = 0; k < npairs; k++) {
      const int a = pairs[2 * k], b = pairs[2 * k + 1];
      need(a >= 0 && a < r->n && b >= 0 && b < r->n && a != b, "invalid link pair");
    }
    std::fill(r->selfcol.begin(), r->selfcol.end(), 0);
    for (int k = 0; k < npairs; k++) {
      const int i = std::min(pairs[2 * k], pairs[2 * k + 1]), j = std::max(pairs[2 * k], pairs[2 * k + 1]);
      r->selfcol[i * r->n + j] = 1;
    }
    r->version++;
  });
}

int kg_robot_self_collision_enabled(kg_robot r, int a, int b, int* out) {
  return guard([&] {
    need(r && out, "NULL argument");
    need(a >= 0 && a < r->n && b >= 0 && b < r->n, "invalid link pair");
    int i = std::min(a, b), j = std::max(a, b);
    *out = (i != j && r->selfcol[i * r->n + j]) ? 1 : 0;
  });
}

int kg_robot_set_config(kg_robot r, const double* q) {
  return guard([&] {
    need(r && q, "NULL argument");
    for (int l = 0; l < r->n; l++) need(std::isfinite(q[l]), "non-finite configuration");
    r->q.assign(q, q + r->n);
    robot_fk(r, q, r->Tw);
    for (int l = 0; l < r->n; l++)
      if (r->link_geoms[l]) r->link_geoms[l]->X = r->Tw[l];
    r->version++;
  });
}

int kg_robot_get_config(kg_robot r, double* q) {
  return guard([&] {
    need(r && q, "NULL argument");
    std::memcpy(q, r->q.data(), r->n * sizeof(double));
  });
}

int kg_robot_get_link_transform(kg_robot r, int link, double R[9], double t[3]) {
  return guard([&] {
    need(r && R && t, "NULL argument");
    need(link >= 0 && link < r->n, "link index out of range");
    std::memcpy(R, r->Tw[link].R, 9 * sizeof(double));
    std::memcpy(t, r->Tw[link].t, 3 * sizeof(double));
  });
}

int kg_fk_batch_dev(kg_robot r, const float* dQ, int N, float* dT, void* stream) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    need(N >= 0, "negative batch size");
    need(N == 0 || (dQ && dT), "NULL buffer");
    hip_check(hipSetDevice(r->ctx->device), "hipSetDevice");
    if (r->dlinks.bytes == 0 || r->dirty) {
      upload(r->dlinks, r->links);
      r->dirty = false;
    }
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(r->ctx, s);
    hip_check(kgl_fk_f32(r->dlinks.as<LinkDesc>(), r->n, dQ, N, dT, s), "fk_f32");
  });
}

int kg_fk_batch(kg_robot r, const float* Q, int N, float* T) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    need(N >= 0, "negative batch size");
    if (N == 0) return;
    need(Q && T, "NULL buffer");
    kg_ctx_s* c = r->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    size_t qb = (size_t)N * r->n * sizeof(float), tb = qb * 12;
    hip_check(c->wsQ.ensure(qb), "hipMalloc");
    hip_check(c->wsF.ensure(tb), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsQ.p, Q, qb, hipMemcpyHostToDevice, c->stream), "copy Q");
    int rc = kg_fk_batch_dev(r, c->wsQ.as<float>(), N, c->wsF.as<float>(), c->stream);
    if (rc != KG_OK) fail(rc, "%s", g_err.c_str());
    hip_check(hipMemcpyAsync(T, c->wsF.p, tb, hipMemcpyDeviceToHost, c->stream), "copy T");
    hip_check(hipStreamSynchronize(c->stream), "sync");
  });
}

int kg_fk_batch64_dev(kg_robot r, const double* dQ, int N, double* dT, void* stream) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    need(N >= 0, "negative batch size");
    need(N == 0 || (dQ && dT), "NULL buffer");
    hip_check(hipSetDevice(r->ctx->device), "hipSetDevice");
    if (r->dlinks.bytes == 0 || r->dirty) {
      upload(r->dlinks, r->links);
      r->dirty = false;
    }
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(r->ctx, s);
    hip_check(kgl_fk64(r->dlinks.as<LinkDesc>(), r->n, nullptr, dQ, N, dT, nullptr, 0, s), "fk64");
  });
}

int kg_fk_batch64(kg_robot r, const double* Q, int N, double* T) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    need(N >= 0, "negative batch size");
    if (N == 0) return;
    need(Q && T, "NULL buffer");
    kg_ctx_s* c = r->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    size_t qb = (size_t)N * r->n * sizeof(double), tb = qb * 12;
    hip_check(c->wsG.ensure(qb), "hipMalloc");
    hip_check(c->wsT64.ensure(tb), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsG.p, Q, qb, hipMemcpyHostToDevice, c->stream), "copy Q");
    int rc = kg_fk_batch64_dev(r, c->wsG.as<double>(), N, c->wsT64.as<double>(), c->stream);
    if (rc != KG_OK) fail(rc, "%s", g_err.c_str());
    hip_check(hipMemcpyAsync(T, c->wsT64.p, tb, hipMemcpyDeviceToHost, c->stream), "copy T");
    hip_check(hipStreamSynchronize(c->stream), "sync");
  });
}

int kg_world_create(kg_ctx c, kg_world* out) {
  return guard([&] {
    need(c && out, "NULL argument");
    kg_world_s* w = new kg_world_s();
    w->ctx = c;
    rebuild_mask(w);
    *out = w;
  });
}

int kg_world_destroy(kg_world w) {
  return guard([&] {
    if (!w) return;
    release_plans(w);
    for (auto it = g_overrides.begin(); it != g_overrides.end(); ++it)
      if (it->first == w) {
        g_overrides.erase(it);
        break;
      }
    for (auto* g : w->terrains) geom_release(g);
    for (auto& o : w->objects) geom_release(o.first);
    delete w;
  });
}

static void world_changed(kg_world_s* w) {
  overrides(w).clear();
  w->version++;
  rebuild_mask(w);
}

int kg_world_add_terrain(kg_world w, kg_geom g, int* id_out) {
  return guard([&] {
    need(w && g, "NULL argument");
    need(g->kind == 0, "only triangle-mesh terrains are supported");
    need(g->ctx == w->ctx, "geometry belongs to another context");
    need(g->alive, "geometry handle was destroyed");
    geom_retain(g);
    w->terrains.push_back(g);
    world_changed(w);
    if (id_out) *id_out = (int)w->terrains.size() - 1;
  });
}

int kg_world_add_rigid_object(kg_world w, kg_geom g, const double R[9], const double t[3], int* id_out) {
  return guard([&] {
    need(w && g && R && t, "NULL argument");
    need(g->kind == 0, "only triangle-mesh rigid objects are supported");
    need(g->ctx == w->ctx, "geometry belongs to another context");
    need(g->alive, "geometry handle was destroyed");
    geom_retain(g);
    w->objects.push_back({g, xf_from(R, t)});
    world_changed(w);
    if (id_out) *id_out = (int)(w->terrains.size() + w->objects.size()) - 1;
  });
}

// A terrain's or rigid object's geometry replaced in place, as loading a file
// into an element's Geometry3D replaces the world's ManagedGeometry
// (Python/klampt/src/robotsim.cpp:1443-1453); the world ID and (for a rigid
// object) the placement stay.
int kg_world_set_terrain_geom(kg_world w, int ti, kg_geom g) {
  return guard([&] {
    need(w && g, "NULL argument");
    need(ti >= 0 && ti < (int)w->terrains.size(), "terrain index out of range");
    need(g->kind == 0, "only triangle-mesh terrains are supported");
    need(g->ctx == w->ctx, "geometry belongs to another context");
    need(g->alive, "geometry handle was destroyed");
    geom_retain(g);
    geom_release(w->terrains[ti]);
    w->terrains[ti] = g;
    w->version++;
  });
}

int kg_world_set_rigid_object_geom(kg_world w, int oi, kg_geom g) {
  return guard([&] {
    need(w && g, "NULL argument");
    need(oi >= 0 && oi < (int)w->objects.size(), "object index out of range");
    need(g->kind == 0, "only triangle-mesh rigid objects are supported");
    need(g->ctx == w->ctx, "geometry belongs to another context");
    need(g->alive, "geometry handle was destroyed");
    geom_retain(g);
    geom_release(w->objects[oi].first);
    w->objects[oi].first = g;
    w->version++;
  });
}

int kg_world_set_rigid_object_transform(kg_world w, int oi, const double R[9], const double t[3]) {
  return guard([&] {
    need(w && R && t, "NULL argument");
    need(oi >= 0 && oi < (int)w->objects.size(), "object index out of range");
    w->objects[oi].second = xf_from(R, t);
    w->version++;
  });
}

int kg_world_add_robot(kg_world w, kg_robot r, int* id_out) {
  return guard([&] {
    need(w && r, "NULL argument");
    need(r->ctx == w->ctx, "robot belongs to another context");
    for (auto* x : w->robots) need(x != r, "robot already in this world");
    w->robots.push_back(r);
    world_changed(w);
    if (id_out) *id_out = robot_id(w, (int)w->robots.size() - 1);
  });
}

int kg_world_num_ids(kg_world w) { return w ? num_ids(w) : 0; }

int kg_world_set_collision_enabled(kg_world w, int a, int b, int enabled) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    int n = num_ids(w);
    need(a >= 0 && a < n && b >= 0 && b < n, "world ID out of range");
    overrides(w).push_back({a, b, enabled ? 1 : 0});
    w->version++;
    rebuild_mask(w);
  });
}

int kg_world_get_collision_enabled(kg_world w, int a, int b, int* out) {
  return guard([&] {
    need(w && out, "NULL argument");
    rebuild_mask(w);
    need(a >= 0 && a < w->nids && b >= 0 && b < w->nids, "world ID out of range");
    *out = en(w, a, b) ? 1 : 0;
  });
}

int kg_world_pair_list(kg_world w, int ri, int32_t* pairs, int maxp, int* np) {
  return guard([&] {
    need(w && np, "NULL argument");
    RobotPlan& P = get_plan(w, ri);
    *np = (int)P.pairs.size();
    for (int i = 0; i < (int)P.pairs.size() && i < maxp && pairs; i++) {
      pairs[2 * i] = P.pairs[i].first;
      pairs[2 * i + 1] = P.pairs[i].second;
    }
  });
}

}  // extern "C"

namespace {
// True while s captures a HIP graph: a wait on work recorded before the
// capture cannot be part of the graph, so slotted calls then skip their slot
// waits (the caller orders the graph's replays against other work).
bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

// A slot's feasibility workspaces for a batch of N configurations and ncap
// pair items (grown as needed).
FeasWs slot_ws(kg_ctx_s* c, kg_ctx_s::FeasSlot& f, int N, size_t ncap) {
  const kg_params& prm = c->prm;
  const int scap = prm.scap == 256 ? 256 : 512;
  const int grid32 = prm.waves_per_cu > 0 ? prm.waves_per_cu * c->num_cus : kgl_collide_grid(0, c->num_cus, scap);
  const int grid64 = std::min(kgl_collide_grid(1, c->num_cus, scap), std::max(1, c->num_cus));
  const size_t spill_entries = (size_t)std::max(grid32, grid64) * prm.spill_cap;
  hip_check(f.status.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
  hip_check(f.jl.ensure((size_t)N), "hipMalloc(jl)");
  hip_check(f.counts.ensure((size_t)(N + 1) * sizeof(int)), "hipMalloc(counts)");
  hip_check(f.redo.ensure(redo_ints(N) * sizeof(int)), "hipMalloc(redo)");
  hip_check(f.items.ensure(std::max<size_t>(1, ncap) * sizeof(PairItem)), "hipMalloc(items)");
  hip_check(f.recs.ensure(std::max<size_t>(1, ncap) * sizeof(ItemRec)), "hipMalloc(records)");
  hip_check(f.spill.ensure(std::max<size_t>(1, spill_entries) * sizeof(uint64_t)), "hipMalloc(spill)");
  FeasWs ws;
  ws.status = f.status.as<int>();
  ws.jl = f.jl.as<uint8_t>();
  ws.counts = f.counts.as<int>();
  ws.redo = f.redo.as<int>();
  ws.items = f.items.as<PairItem>();
  ws.recs = f.recs.as<ItemRec>();
  ws.spill = f.spill.as<uint64_t>();
  ws.spill_entries = (int)std::min<size_t>(spill_entries, (size_t)INT32_MAX);
  return ws;
}

// A device-pointer feasibility batch on the context's next workspace slot
// (the fixed-capacity, per-link broad-phase path; anything else runs as a
// plain call in context order).  The batch waits for the previous batch on
// its slot and for the last plain call, not for batches on the other slots.
void feasible_dev_slot(kg_world_s* w, int ri, const float* Qf, const double* Qd, int N, uint8_t* feas, uint8_t* env,
                       uint8_t* self, int32_t* fp, hipStream_t s) {
  kg_ctx_s* c = w->ctx;
  need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  reject_margins(w);
  const int n = w->robots[ri]->n;
  // get_plan synchronizes the device before it replaces tables that batches
  // on other streams may read
  RobotPlan& P = get_plan(w, ri);
  Layout L;
  L.P = &P;
  L.cap = active_links(w->robots[ri]) * P.nenv + P.nself;
  L.maxcap = L.cap;
  L.nrows = (size_t)N * n;
  L.ncap = (size_t)L.cap * N;
  L.tol32 = P.tol32;
  L.tol64 = P.tol64;
  const bool slotted = N > 0 && c->prm.prep_group && n <= 32 && L.ncap * sizeof(PairItem) <= (2048ull << 20);
  if (!slotted) {
    StreamScope scope(c, s);
    feasible_core(c, L, Qf, Qd, N, feas, env, self, fp, s);
    return;
  }
  if (!c->evDone)
    hip_check(hipEventCreateWithFlags(&c->evDone, hipEventDisableTiming | hipEventDisableSystemFence),
              "hipEventCreate");
  kg_ctx_s::FeasSlot& f = c->fslot[c->fslot_next];
  c->fslot_next = (c->fslot_next + 1) % kg_ctx_s::FEAS_SLOTS;
  if (!f.done)
    hip_check(hipEventCreateWithFlags(&f.done, hipEventDisableTiming | hipEventDisableSystemFence),
              "hipEventCreate");
  // the slot's previous batch, and the last plain call (shared tables)
  const bool cap = capturing(s);
  if (f.valid && !cap) hip_check(hipStreamWaitEvent(s, f.done, 0), "hipStreamWaitEvent");
  if (c->done_valid && !c->last_slotted && c->done_stream != s && !cap)
    hip_check(hipStreamWaitEvent(s, c->evDone, 0), "hipStreamWaitEvent");
  FeasWs ws = slot_ws(c, f, N, L.ncap);
  feasible_core(c, L, Qf, Qd, N, feas, env, self, fp, s, nullptr, true, &ws);
  hip_check(hipEventRecord(f.done, s), "hipEventRecord");
  f.valid = true;
  if (hipEventRecord(c->evDone, s) == hipSuccess) {
    c->done_stream = s;
    c->done_valid = true;
    c->last_slotted = true;
  }
}
}  // namespace

extern "C" {

int kg_feasible_batch_dev(kg_world w, int ri, const float* dQ, int N, uint8_t* feas, uint8_t* env, uint8_t* self,
                          int32_t* fp, void* stream) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(N >= 0, "negative batch size");
    need(N == 0 || (dQ && feas), "NULL buffer");
    feasible_dev_slot(w, ri, dQ, nullptr, N, feas, env, self, fp, (hipStream_t)stream);
  });
}

int kg_feasible_batch_f64_dev(kg_world w, int ri, const double* dQ, int N, uint8_t* feas, uint8_t* env,
                              uint8_t* self, int32_t* fp, void* stream) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(N >= 0, "negative batch size");
    need(N == 0 || (dQ && feas), "NULL buffer");
    feasible_dev_slot(w, ri, nullptr, dQ, N, feas, env, self, fp, (hipStream_t)stream);
  });
}

// The allocation-free variant of kg_feasible_batch_dev: every workspace is
// carved from the device's scratch array (kgl_noalloc_scratch); the world's
// tables must already be on the device.
int kg_feasible_batch_dev_noalloc(kg_world w, int ri, const float* dQ, int N, uint8_t* feas, uint8_t* env,
                                  uint8_t* self, int32_t* fp, void* stream) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
    need(N >= 0, "negative batch size");
    need(N == 0 || (dQ && feas), "NULL buffer");
    kg_ctx_s* c = w->ctx;
    need(c->scratch != nullptr, "no scratch on this device");
    if (N == 0) return;
    reject_margins(w);
    // no table uploads here: the plan must be current
    need(!c->pool_dirty && ri < (int)w->plans.size() &&
             w->plans[ri].built_version == world_version(w) * 131 + c->pool_version,
         "the world changed since its tables were uploaded: run kg_feasible_batch_dev (or any batched call) on "
         "it once before the allocation-free call");
    RobotPlan& P = w->plans[ri];
    const int n = P.dev.nlinks;
    if (!(c->prm.prep_group && n <= 32))
      fail(KG_E_UNSUPPORTED, "the allocation-free call needs robots of at most 32 links and prep_group = 1");
    Layout L;
    L.P = &P;
    L.cap = active_links(w->robots[ri]) * P.nenv + P.nself;
    L.maxcap = L.cap;
    L.nrows = (size_t)N * n;
    L.ncap = (size_t)L.cap * N;
    L.tol32 = P.tol32;
    L.tol64 = P.tol64;
    // carve: status, spill (fixed 16 MiB), then the per-configuration arrays
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_status = al(ST_WORDS * sizeof(int)), b_spill = (size_t)16 << 20;
    const size_t b_jl = al((size_t)N), b_counts = al((size_t)(N + 1) * sizeof(int)),
                 b_redo = al(redo_ints(N) * sizeof(int)), b_items = al(L.ncap * sizeof(PairItem)),
                 b_recs = al(L.ncap * sizeof(ItemRec));
    const size_t total = b_status + b_spill + b_jl + b_counts + b_redo + b_items + b_recs;
    if (total > KG_NOALLOC_BYTES) {
      const size_t per = 1 + 2 * sizeof(int) + (size_t)L.cap * (sizeof(PairItem) + sizeof(ItemRec));
      fail(KG_E_LIMIT,
           "the allocation-free call holds %u bytes of scratch; %d configurations of this robot and world need "
           "%zu (about %zu at most)",
           (unsigned)KG_NOALLOC_BYTES, N, total,
           (size_t)(KG_NOALLOC_BYTES - b_status - b_spill - 5 * 256 - (redo_ints(0) + 4) * sizeof(int)) / per);
    }
    unsigned char* p = (unsigned char*)c->scratch;
    FeasWs ws;
    ws.status = (int*)p;
    p += b_status;
    ws.spill = (uint64_t*)p;
    ws.spill_entries = (int)(b_spill / sizeof(uint64_t));
    p += b_spill;
    ws.jl = p;
    p += b_jl;
    ws.counts = (int*)p;
    p += b_counts;
    ws.redo = (int*)p;
    p += b_redo;
    ws.items = (PairItem*)p;
    p += b_items;
    ws.recs = (ItemRec*)p;
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(c, s);
    feasible_core(c, L, dQ, nullptr, N, feas, env, self, fp, s, nullptr, false, &ws);
  });
}
}  // extern "C"

namespace {
// Host-pointer feasibility with float (Qf) or double (Qd) configurations.
void feasible_host(kg_world_s* w, int ri, const float* Qf, const double* Qd, int N, uint8_t* feas, uint8_t* env,
                   uint8_t* self, int32_t* fp) {
    need(w != nullptr, "world is NULL");
    need(N >= 0, "negative batch size");
    need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
    if (N == 0) return;
    need((Qf || Qd) && feas, "NULL buffer");
    kg_ctx_s* c = w->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    int n = w->robots[ri]->n;
    const size_t qb = (size_t)N * n * (Qd ? sizeof(double) : sizeof(float));
    hip_check(c->wsQ.ensure(qb), "hipMalloc");
    hip_check(c->wsOut8a.ensure(N), "hipMalloc");
    hip_check(c->wsOut8b.ensure(N), "hipMalloc");
    hip_check(c->wsOut8c.ensure(N), "hipMalloc");
    hip_check(c->wsOut32.ensure((size_t)N * 2 * sizeof(int)), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsQ.p, Qd ? (const void*)Qd : (const void*)Qf, qb, hipMemcpyHostToDevice, c->stream),
              "copy Q");
    uint8_t* dfe = c->wsOut8a.as<uint8_t>();
    uint8_t* den = env ? c->wsOut8b.as<uint8_t>() : nullptr;
    uint8_t* dse = self ? c->wsOut8c.as<uint8_t>() : nullptr;
    int32_t* dfp = fp ? c->wsOut32.as<int32_t>() : nullptr;
    feasible_dev(w, ri, Qd ? nullptr : c->wsQ.as<float>(), Qd ? c->wsQ.as<double>() : nullptr, N, dfe, den, dse, dfp,
                 c->stream);
    hip_check(hipMemcpyAsync(feas, dfe, N, hipMemcpyDeviceToHost, c->stream), "copy out");
    if (env) hip_check(hipMemcpyAsync(env, den, N, hipMemcpyDeviceToHost, c->stream), "copy out");
    if (self) hip_check(hipMemcpyAsync(self, dse, N, hipMemcpyDeviceToHost, c->stream), "copy out");
    if (fp) hip_check(hipMemcpyAsync(fp, dfp, (size_t)N * 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    feasible_sync_check(w);
}
}  // namespace

extern "C" {

int kg_feasible_batch(kg_world w, int ri, const float* Q, int N, uint8_t* feas, uint8_t* env, uint8_t* self,
                      int32_t* fp) {
  return guard([&] { feasible_host(w, ri, Q, nullptr, N, feas, env, self, fp); });
}

int kg_feasible_batch_f64(kg_world w, int ri, const double* Q, int N, uint8_t* feas, uint8_t* env, uint8_t* self,
                          int32_t* fp) {
  return guard([&] {
    need(Q != nullptr || N == 0, "NULL buffer");
    feasible_host(w, ri, nullptr, Q, N, feas, env, self, fp);
  });
}

int kg_world_last_stats(kg_world w, int64_t* out, int nstats) {
  return guard([&] {
    need(w && out, "NULL argument");
    BatchStats st;
    read_stats(w, st);
    int64_t v[10] = {st.redo, st.items, st.spills, st.t_prep_ns, st.t_trav32_ns, st.t_trav64_ns, st.box_tests,
                     st.tri_tests, st.uncertain, st.xpairs};
    for (int i = 0; i < nstats && i < 10; i++) out[i] = v[i];
  });
}

// Sum of the FP32 traversal kernel durations of the feasibility batches
// issued since the last reset (at most the last TRING of them; *count says
// how many were summed).  Waits for them to finish.
int kg_ctx_kernel_times(kg_ctx c, int reset, int64_t* total_ns, int32_t* count) {
  return guard([&] {
    need(c && total_ns && count, "NULL argument");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const int n = (int)std::min<int64_t>(c->tr_count, kg_ctx_s::TRING);
    double ms_sum = 0;
    for (int k = 0; k < n; k++) {
      int slot = (int)((c->tr_count - 1 - k) % kg_ctx_s::TRING);
      hip_check(hipEventSynchronize(c->tr_end[slot]), "hipEventSynchronize");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, c->tr_beg[slot], c->tr_end[slot]), "hipEventElapsedTime");
      ms_sum += ms;
    }
    *total_ns = (int64_t)(ms_sum * 1e6);
    *count = n;
    if (reset) c->tr_count = 0;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- robot-only world

namespace {
std::vector<std::pair<const kg_robot_s*, kg_world_s*>> g_solo;

kg_world_s* solo_world(kg_robot_s* r) {
  for (auto& p : g_solo)
    if (p.first == r) return p.second;
  kg_world_s* w = new kg_world_s();
  w->ctx = r->ctx;
  w->robots.push_back(r);
  rebuild_mask(w);
  g_solo.push_back({r, w});
  return w;
}

// The internal one-robot world of a robot (kg_robot_self_collides) goes
// with the robot, so a later robot at the same address gets a fresh one.
void forget_solo_world(const kg_robot_s* r) {
  for (auto it = g_solo.begin(); it != g_solo.end(); ++it)
    if (it->first == r) {
      release_plans(it->second);
      delete it->second;
      g_solo.erase(it);
      return;
    }
}
}  // namespace

extern "C" int kg_robot_self_collides(kg_robot r, int* out) {
  return guard([&] {
    need(r && out, "NULL argument");
    kg_world_s* w = solo_world(r);
    std::vector<float> q(r->n);
    // the current configuration in FP64
    kg_ctx_s* c = r->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    hip_check(c->wsG.ensure(r->n * sizeof(double) + 64), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsG.p, r->q.data(), r->n * sizeof(double), hipMemcpyHostToDevice, c->stream),
              "copy q");
    hip_check(c->wsOut8a.ensure(64), "hipMalloc");
    hip_check(c->wsOut8b.ensure(64), "hipMalloc");
    hip_check(c->wsOut8c.ensure(64), "hipMalloc");
    feasible_dev(w, 0, nullptr, c->wsG.as<double>(), 1, c->wsOut8a.as<uint8_t>(), c->wsOut8b.as<uint8_t>(),
                 c->wsOut8c.as<uint8_t>(), nullptr, c->stream);
    uint8_t h = 0;
    hip_check(hipMemcpyAsync(&h, c->wsOut8c.p, 1, hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    feasible_sync_check(w);
    *out = h;
  });
}

// ---------------------------------------------------------------- single pairs
//
// A pair query (a at X_a, b at X_b) runs through the batch pipeline on an
// internal world: a one-link robot whose T0_Parent is X_a (q = 0 gives the
// joint transform I exactly, so the link frame is X_a bit for bit) and a rigid
// object b at X_b.
namespace {
struct PairWorld {
  kg_robot_s* r = nullptr;
  kg_world_s* w = nullptr;
};
// One entry per context, shared by every context of the process: a list
// (entries stay in place as others come and go) under a mutex; each entry is
// used only by its own context's calls.
std::list<std::pair<const kg_ctx_s*, PairWorld>> g_pairworlds;
std::mutex g_pairworlds_mu;

PairWorld& pair_world(kg_ctx_s* c, kg_geom_s* a, kg_geom_s* b) {
  PairWorld* pw = nullptr;
  std::lock_guard<std::mutex> lock(g_pairworlds_mu);
  for (auto& p : g_pairworlds)
    if (p.first == c) pw = &p.second;
  if (!pw) {
    g_pairworlds.push_back({c, PairWorld{}});
    pw = &g_pairworlds.back().second;
    kg_robot_s* r = new kg_robot_s();
    r->ctx = c;
    r->n = 1;
    r->links.resize(1);
    LinkDesc& L = r->links[0];
    for (int i = 0; i < 12; i++) L.T0[i] = (i % 4 == 0 && i < 9) ? 1.0 : 0.0;
    L.axis[0] = 0; L.axis[1] = 0; L.axis[2] = 1;
    L.parent = -1;
    L.jtype = 0;
    L.geom = -1;
    L.wid = 1;
    r->link_geoms.assign(1, nullptr);
    r->qmin.assign(1, 0.0);
    r->qmax.assign(1, 0.0);
    r->q.assign(1, 0.0);
    r->selfcol.assign(1, 0);
    r->Tw.assign(1, xf_identity());
    kg_world_s* w = new kg_world_s();
    w->ctx = c;
    w->robots.push_back(r);
    w->objects.push_back({nullptr, xf_identity()});
    pw->r = r;
    pw->w = w;
  }
  for (int i = 0; i < 9; i++) pw->r->links[0].T0[i] = a->X.R[i];
  for (int i = 0; i < 3; i++) pw->r->links[0].T0[9 + i] = a->X.t[i];
  pw->r->link_geoms[0] = a;
  pw->r->Tw[0] = a->X;
  pw->r->version++;
  pw->w->objects[0] = {b, b->X};
  pw->w->version++;
  rebuild_mask(pw->w);
  return *pw;
}

// The context's internal pair world goes with the context (its geometry
// slots hold no references: they are set anew before every query).
void forget_pair_world(const kg_ctx_s* c) {
  std::lock_guard<std::mutex> lock(g_pairworlds_mu);
  for (auto it = g_pairworlds.begin(); it != g_pairworlds.end(); ++it)
    if (it->first == c) {
      release_plans(it->second.w);
      delete it->second.w;
      it->second.r->dlinks.release();
      delete it->second.r;
      g_pairworlds.erase(it);
      return;
    }
}
}  // namespace

namespace {
void pair_distance(kg_geom_s* a, kg_geom_s* b, float stop_below, double upper, kg_dist_result* out, float* dgpu,
                   double rel_err = 0, double abs_err = 0);
SdfDev sdf_dev(kg_geom_s* g);

// Point sets against implicit surfaces: a PointCloud's points or a
// TriangleMesh's vertices, each moved by its geometry's transform and
// evaluated in the surface's grid.  A pair qualifies when exactly one side is
// an ImplicitSurface; `pts` is then the other side.
bool point_set_pair(kg_geom_s* a, kg_geom_s* b, kg_geom_s*& pts, kg_geom_s*& sdf) {
  if (a->kind == 1 && b->kind != 1) {
    sdf = a;
    pts = b;
    return true;
  }
  if (b->kind == 1 && a->kind != 1) {
    sdf = b;
    pts = a;
    return true;
  }
  return false;
}

bool point_set_empty(const kg_geom_s* g) { return g->kind == 0 ? geom_empty(g) : g->V.empty(); }

const float4* point_set(kg_geom_s* g, int& n) {
  n = (int)(g->V.size() / 3);
  if (!g->dpts.p) {
    std::vector<float4> P(n);
    for (int i = 0; i < n; i++) P[i] = make_float4((float)g->V[3 * i], (float)g->V[3 * i + 1], (float)g->V[3 * i + 2], 0.f);
    hip_check(g->dpts.ensure((size_t)n * sizeof(float4)), "hipMalloc");
    hip_check(hipMemcpy(g->dpts.p, P.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice), "copy points");
  }
  return g->dpts.as<float4>();
}

// Smallest implicit-surface value over the point set; *index = its point.
double pts_sdf_min(kg_geom_s* pts, kg_geom_s* sdf, int* index) {
  kg_ctx_s* c = pts->ctx;
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  int n = 0;
  const float4* P = point_set(pts, n);
  StreamScope scope(c, c->stream);
  float X[12];
  for (int i = 0; i < 9; i++) X[i] = (float)pts->X.R[i];
  for (int i = 0; i < 3; i++) X[9 + i] = (float)pts->X.t[i];
  hip_check(c->wsG.ensure(64), "hipMalloc");
  unsigned long long key = ~0ull;
  hip_check(hipMemcpyAsync(c->wsG.p, &key, sizeof(key), hipMemcpyHostToDevice, c->stream), "copy");
  hip_check(kgl_pts_sdf_min(P, n, X, sdf_dev(sdf), c->wsG.as<unsigned long long>(), c->stream), "pts_sdf_min");
  hip_check(hipMemcpyAsync(&key, c->wsG.p, sizeof(key), hipMemcpyDeviceToHost, c->stream), "copy");
  hip_check(hipStreamSynchronize(c->stream), "sync");
  uint32_t u = (uint32_t)(key >> 32);
  u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float v;
  std::memcpy(&v, &u, sizeof(v));
  *index = (int)(key & 0xffffffffu);
  return v;
}
}  // namespace

extern "C" int kg_geom_collide(kg_geom a, kg_geom b, int* out) {
  return guard([&] {
    need(a && b && out, "NULL argument");
    need(a->ctx == b->ctx, "geometries belong to different contexts");
    *out = 0;
    if (a->kind != 0 || b->kind != 0) {
      kg_geom_s *p = nullptr, *f = nullptr;
      if (!point_set_pair(a, b, p, f))
        fail(KG_E_UNSUPPORTED,
             "collide is implemented for TriangleMesh pairs and PointCloud / TriangleMesh against ImplicitSurface");
      if (point_set_empty(p)) return;
      int idx = 0;
      *out = pts_sdf_min(p, f, &idx) <= a->margin + b->margin ? 1 : 0;
      return;
    }
    if (geom_empty(a) || geom_empty(b)) return;  // Geometry3D.collides: empty -> False
    const double m = a->margin + b->margin;
    if (m > 0) {
      // fattened geometries collide when within the summed margins
      kg_dist_result r;
      float dg = 0;
      pair_distance(a, b, (float)m, INFINITY, &r, &dg);
      *out = r.d <= m ? 1 : 0;
      return;
    }
    kg_ctx_s* c = a->ctx;
    PairWorld& pw = pair_world(c, a, b);
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    hip_check(c->wsG.ensure(64), "hipMalloc");
    double q0 = 0.0;
    hip_check(hipMemcpyAsync(c->wsG.p, &q0, sizeof(double), hipMemcpyHostToDevice, c->stream), "copy q");
    hip_check(c->wsOut8a.ensure(64), "hipMalloc");
    hip_check(c->wsOut8b.ensure(64), "hipMalloc");
    feasible_dev(pw.w, 0, nullptr, c->wsG.as<double>(), 1, c->wsOut8a.as<uint8_t>(), c->wsOut8b.as<uint8_t>(),
                 nullptr, nullptr, c->stream);
    uint8_t h = 0;
    hip_check(hipMemcpyAsync(&h, c->wsOut8b.p, 1, hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    feasible_sync_check(pw.w);
    *out = h;
  });
}

// ---------------------------------------------------------------- SDF
namespace {

SdfDev sdf_dev(kg_geom_s* g) {
  need(g && g->kind == 1, "geometry is not an ImplicitSurface");
  SdfDev d{};
  d.v = g->dvals.as<float>();
  d.v8 = g->dv8.as<float4>();  // present once a penalty call has built it
  for (int k = 0; k < 3; k++) {
    d.cdims[k] = std::max(g->dims[k] - 1, 1);
    d.dims[k] = g->dims[k];
    const double h = (g->bmax[k] - g->bmin[k]) / g->dims[k];
    d.bmin[k] = (float)g->bmin[k];
    d.h[k] = (float)h;
    d.inv_h[k] = (float)(1.0 / h);
    d.bmin64[k] = g->bmin[k];
    d.inv_h64[k] = 1.0 / h;
  }
  // Xinv = X^-1
  const Xf64& X = g->X;
  bool ident = true;
  for (int i = 0; i < 9; i++) ident = ident && X.R[i] == ((i % 4 == 0) ? 1.0 : 0.0);
  for (int i = 0; i < 3; i++) ident = ident && X.t[i] == 0.0;
  d.ident = ident ? 1 : 0;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) d.Xinv[c * 3 + r] = (float)X.R[r * 3 + c];
  for (int r = 0; r < 3; r++)
    d.Xinv[9 + r] = (float)(-(X.R[r * 3] * X.t[0] + X.R[r * 3 + 1] * X.t[1] + X.R[r * 3 + 2] * X.t[2]));
  return d;
}

void robot_links_dev(kg_robot_s* r) {
  if (r->dlinks.bytes == 0 || r->dirty) {
    upload(r->dlinks, r->links);
    r->dirty = false;
  }
}

// Penalty on device buffers; pts_per_link is a host array.
void sdf_penalty_dev(kg_robot_s* r, kg_geom_s* sdf, const float* dQ, int N, const float* dpts, const int32_t* ppl,
                     float dsafe, float* dloss, const float* ddl, float* ddQ, bool bwd, hipStream_t s) {
  need(r != nullptr, "robot is NULL");
  need(sdf != nullptr && sdf->kind == 1, "sdf must be an ImplicitSurface geometry");
  need(ppl != nullptr, "pts_per_link is NULL");
  need(N >= 0, "negative batch size");
  if (N == 0) return;
  need(dQ && dpts, "NULL buffer");
  need(bwd ? ddQ != nullptr : dloss != nullptr, "NULL output");
  kg_ctx_s* c = r->ctx;
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  const int n = r->n;
  std::vector<int> off(n + 1, 0);
  for (int k = 0; k < n; k++) {
    need(ppl[k] >= 0, "negative point count");
    off[k + 1] = off[k] + ppl[k];
  }
  robot_links_dev(r);
  hip_check(c->wsPoff.ensure((n + 1) * sizeof(int)), "hipMalloc");
  hip_check(hipMemcpyAsync(c->wsPoff.p, off.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, s), "copy offsets");
  if (!sdf->dv8.p) {
    // the corner-per-cell copy the penalty kernels read (8x the grid: 0.5 GB
    // for 256^3 cells; grids whose copy would pass 4 GiB keep the gathers)
    const size_t ncell =
        (size_t)std::max(sdf->dims[0] - 1, 1) * std::max(sdf->dims[1] - 1, 1) * std::max(sdf->dims[2] - 1, 1);
    if (ncell * 2 * sizeof(float4) <= (4ull << 30)) {
      hip_check(sdf->dv8.ensure(ncell * 2 * sizeof(float4)), "hipMalloc(sdf corners)");
      hip_check(kgl_sdf_build8(sdf->dvals.as<float>(), sdf->dims, sdf->dv8.as<float4>(), s), "sdf corners");
    }
  }
  SdfDev g = sdf_dev(sdf);
  hip_check(kgl_sdf_penalty(r->dlinks.as<LinkDesc>(), n, g, dQ, N, dpts, c->wsPoff.as<int>(), dsafe, dloss, ddl, ddQ, bwd,
                            c->prm.sdf_lanes, s),
            "sdf_penalty");
}

int total_points(const kg_robot_s* r, const int32_t* ppl) {
  int t = 0;
  for (int k = 0; k < r->n; k++) t += ppl[k];
  return t;
}

}  // namespace

extern "C" int kg_sdf_eval(kg_geom sdf, const float* pts, int n, float* d, float* grad) {
  return guard([&] {
    need(sdf && sdf->kind == 1, "geometry is not an ImplicitSurface");
    need(n >= 0, "negative size");
    if (n == 0) return;
    need(pts && d, "NULL buffer");
    kg_ctx_s* c = sdf->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    hip_check(c->wsPts.ensure((size_t)n * 3 * sizeof(float)), "hipMalloc");
    hip_check(c->wsLoss.ensure((size_t)n * sizeof(float)), "hipMalloc");
    hip_check(c->wsDQ.ensure((size_t)n * 3 * sizeof(float)), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsPts.p, pts, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy");
    hip_check(kgl_sdf_eval(sdf_dev(sdf), c->wsPts.as<float>(), n, c->wsLoss.as<float>(), grad ? c->wsDQ.as<float>() : nullptr,
                           c->stream),
              "sdf_eval");
    hip_check(hipMemcpyAsync(d, c->wsLoss.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
    if (grad)
      hip_check(hipMemcpyAsync(grad, c->wsDQ.p, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
  });
}

extern "C" int kg_geom_distance_point(kg_geom g, const double pt[3], const kg_dist_settings* settings,
                                      kg_dist_result* out) {
  return guard([&] {
    need(g && pt && out, "NULL argument");
    (void)settings;
    if (g->kind != 1) fail(KG_E_UNSUPPORTED, "distance_point is implemented for ImplicitSurface geometries");
    float p[3] = {(float)pt[0], (float)pt[1], (float)pt[2]}, d = 0, gr[3] = {0, 0, 0};
    int rc = kg_sdf_eval(g, p, 1, &d, gr);
    if (rc != KG_OK) fail(rc, "%s", g_err.c_str());
    std::memset(out, 0, sizeof(*out));
    out->d = d;
    // closest point on the zero level set along the gradient; gradients as
    // in Python/klampt/src/robotsim.cpp:1768-1773 (grad2 = -grad1)
    out->hasClosestPoints = 1;
    out->hasGradients = 1;
    for (int k = 0; k < 3; k++) {
      out->grad1[k] = gr[k];
      out->grad2[k] = -gr[k];
      out->cp1[k] = pt[k] - d * gr[k];
      out->cp2[k] = pt[k];
    }
    out->elem1 = out->elem2 = -1;
  });
}

extern "C" int kg_sdf_penalty_fwd_dev(kg_robot r, kg_geom sdf, const float* dQ, int N, const float* dpts,
                                      const int32_t* ppl, float dsafe, float* dloss, void* stream) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(r->ctx, s);
    sdf_penalty_dev(r, sdf, dQ, N, dpts, ppl, dsafe, dloss, nullptr, nullptr, false, s);
  });
}

extern "C" int kg_sdf_penalty_bwd_dev(kg_robot r, kg_geom sdf, const float* dQ, int N, const float* dpts,
                                      const int32_t* ppl, float dsafe, const float* ddl, float* ddQ, void* stream) {
  return guard([&] {
    need(r != nullptr, "robot is NULL");
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(r->ctx, s);
    sdf_penalty_dev(r, sdf, dQ, N, dpts, ppl, dsafe, nullptr, ddl, ddQ, true, s);
  });
}

static void sdf_penalty_host(kg_robot_s* r, kg_geom_s* sdf, const float* Q, int N, const float* pts, const int32_t* ppl,
                             float dsafe, float* loss, const float* dl, float* dQ, bool bwd) {
  need(r && sdf && ppl, "NULL argument");
  need(N >= 0, "negative batch size");
  if (N == 0) return;
  need(Q && pts, "NULL buffer");
  kg_ctx_s* c = r->ctx;
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  StreamScope scope(c, c->stream);
  const int n = r->n, P = total_points(r, ppl);
  hip_check(c->wsQ.ensure((size_t)N * n * sizeof(float)), "hipMalloc");
  hip_check(c->wsPts.ensure((size_t)std::max(1, P) * 3 * sizeof(float)), "hipMalloc");
  hip_check(c->wsLoss.ensure((size_t)N * sizeof(float)), "hipMalloc");
  hip_check(c->wsDQ.ensure((size_t)N * n * sizeof(float)), "hipMalloc");
  hip_check(c->wsDl.ensure((size_t)N * sizeof(float)), "hipMalloc");
  hip_check(hipMemcpyAsync(c->wsQ.p, Q, (size_t)N * n * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy");
  if (P > 0) hip_check(hipMemcpyAsync(c->wsPts.p, pts, (size_t)P * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy");
  if (bwd && dl) hip_check(hipMemcpyAsync(c->wsDl.p, dl, (size_t)N * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy");
  sdf_penalty_dev(r, sdf, c->wsQ.as<float>(), N, c->wsPts.as<float>(), ppl, dsafe, bwd ? nullptr : c->wsLoss.as<float>(),
                  (bwd && dl) ? c->wsDl.as<float>() : nullptr, bwd ? c->wsDQ.as<float>() : nullptr, bwd, c->stream);
  if (bwd)
    hip_check(hipMemcpyAsync(dQ, c->wsDQ.p, (size_t)N * n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
  else
    hip_check(hipMemcpyAsync(loss, c->wsLoss.p, (size_t)N * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
  hip_check(hipStreamSynchronize(c->stream), "sync");
}

extern "C" int kg_sdf_penalty_fwd(kg_robot r, kg_geom sdf, const float* Q, int N, const float* pts,
                                  const int32_t* ppl, float dsafe, float* loss) {
  return guard([&] {
    need(loss != nullptr || N == 0, "loss is NULL");
    sdf_penalty_host(r, sdf, Q, N, pts, ppl, dsafe, loss, nullptr, nullptr, false);
  });
}

extern "C" int kg_sdf_penalty_bwd(kg_robot r, kg_geom sdf, const float* Q, int N, const float* pts,
                                  const int32_t* ppl, float dsafe, const float* dl, float* dQ) {
  return guard([&] {
    need(dQ != nullptr || N == 0, "dQ is NULL");
    sdf_penalty_host(r, sdf, Q, N, pts, ppl, dsafe, nullptr, dl, dQ, true);
  });
}

// ---------------------------------------------------------------- distance
namespace {

// Per-pair minimum distances for robot `ri` on device buffers (async).
void distance_dev(kg_world_s* w, int ri, const float* Qf, const double* Qd, int N, float* out, int* out_tri,
                  float stop_below, float upper, hipStream_t s, float rel_err = 0.f, float abs_err = 0.f) {
  kg_ctx_s* c = w->ctx;
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  RobotPlan& P = get_plan(w, ri);
  const int npairs = (int)P.pairs.size();
  if (N <= 0 || npairs == 0) return;
  const int n = P.dev.nlinks;
  hip_check(c->wsT64.ensure((size_t)N * n * 12 * sizeof(double)), "hipMalloc(T64)");
  hip_check(c->wsStatus.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
  int* status = c->wsStatus.as<int>();
  hip_check(kgl_fk64(P.dev.links, n, Qf, Qd, N, c->wsT64.as<double>(), status, ST_WORDS, s), "fk64");
  const int grid = kgl_distance_grid(c->num_cus);
  const int spill_cap = c->prm.spill_cap;
  hip_check(c->wsSpill.ensure(std::max<size_t>(1, (size_t)grid * spill_cap) * sizeof(uint64_t)), "hipMalloc(spill)");
  DistArgs da{};
  da.pool = c->pool;
  da.plan = P.dev;
  da.T64 = c->wsT64.as<double>();
  da.N = N;
  da.npairs = npairs;
  da.pair_a = P.dpa.as<int>();
  da.pair_b = P.dpb.as<int>();
  da.pair_kind = P.dpk.as<int>();
  da.stop_below = stop_below;
  da.upper = upper;
  da.rel_err = rel_err;
  da.abs_err = abs_err;
  da.out = out;
  da.out_tri = out_tri;
  da.chunk_counter = status + ST_CHUNK32;
  da.spill = c->wsSpill.as<uint64_t>();
  da.spill_cap = spill_cap;
  da.status = status + ST_SPILL;
  da.grid = grid;
  da.ipw = c->prm.dist_ipw;
  da.counters = (unsigned long long*)(status + ST_CNT);
  da.uncertain = status + ST_UNC;
  const bool phase_ev = c->timing >= 2;
  c->last_phase_ev = phase_ev;
  c->last_tr_slot = -1;
  if (phase_ev) hip_check(hipEventRecord(c->ev[2], s), "hipEventRecord");
  hip_check(kgl_distance(da, s), "distance");
  if (phase_ev) hip_check(hipEventRecord(c->ev[3], s), "hipEventRecord");
  c->last_N = N;
  c->has_batch = true;
  c->stat_src = status;
  hip_check(hipEventRecord(c->evStat, s), "hipEventRecord");
}


// FP64 distance and closest points of one triangle pair (original indices)
double pair_tri_dist(const kg_geom_s* a, int ta, const kg_geom_s* b, int tb, double* cp1, double* cp2) {
  kgm::V<double> P[3], Q[3];
  for (int v = 0; v < 3; v++) {
    const double* pa = &a->V[3 * a->T[3 * ta + v]];
    const double* pb = &b->V[3 * b->T[3 * tb + v]];
    double wa[3], wb[3];
    for (int r = 0; r < 3; r++) {
      wa[r] = a->X.R[r] * pa[0] + a->X.R[3 + r] * pa[1] + a->X.R[6 + r] * pa[2] + a->X.t[r];
      wb[r] = b->X.R[r] * pb[0] + b->X.R[3 + r] * pb[1] + b->X.R[6 + r] * pb[2] + b->X.t[r];
    }
    P[v] = kgm::mk(wa[0], wa[1], wa[2]);
    Q[v] = kgm::mk(wb[0], wb[1], wb[2]);
  }
  kgm::V<double> c1, c2;
  const double d = kgm::tri_dist<double>(P, Q, c1, c2);
  cp1[0] = c1.x; cp1[1] = c1.y; cp1[2] = c1.z;
  cp2[0] = c2.x; cp2[1] = c2.y; cp2[2] = c2.z;
  return d;
}

// Single pair: distance with closest points via the pair world.
void pair_distance(kg_geom_s* a, kg_geom_s* b, float stop_below, double upper, kg_dist_result* out, float* dgpu,
                   double rel_err, double abs_err) {
  kg_ctx_s* c = a->ctx;
  ensure_status_host(c);
  PairWorld& pw = pair_world(c, a, b);
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  StreamScope scope(c, c->stream);
  hip_check(c->wsG.ensure(64), "hipMalloc");
  hip_check(c->wsOut32.ensure(64), "hipMalloc");
  double q0 = 0.0;
  hip_check(hipMemcpyAsync(c->wsG.p, &q0, sizeof(double), hipMemcpyHostToDevice, c->stream), "copy q");
  float* dout = (float*)c->wsOut32.p;
  int* dtri = (int*)((char*)c->wsOut32.p + 16);
  const float up = std::isfinite(upper) ? (float)upper : 3.0e38f;
  distance_dev(pw.w, 0, nullptr, c->wsG.as<double>(), 1, dout, dtri, stop_below, up, c->stream, (float)rel_err,
               (float)abs_err);
  float d = 0;
  int tri[2] = {-1, -1};
  hip_check(hipMemcpyAsync(&d, dout, sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
  hip_check(hipMemcpyAsync(tri, dtri, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream), "copy");
  hip_check(hipStreamSynchronize(c->stream), "sync");
  check_uncertain(c);
  if (dgpu) *dgpu = d;
  std::memset(out, 0, sizeof(*out));
  out->d = d;
  out->elem1 = out->elem2 = -1;
  if (tri[0] >= 0 && tri[1] >= 0) {
    const int ta = c->trisrc[tri[0]], tb = c->trisrc[tri[1]];
    double cp1[3], cp2[3];
    const double d64 = pair_tri_dist(a, ta, b, tb, cp1, cp2);
    out->d = d64;
    out->hasClosestPoints = 1;
    out->elem1 = ta;
    out->elem2 = tb;
    for (int k = 0; k < 3; k++) {
      out->cp1[k] = cp1[k];
      out->cp2[k] = cp2[k];
    }
    if (d64 > 0) {
      // grad1 points from the other object toward this one
      // (Python/klampt/src/geometry.h:650-662: move object1 along grad1 by -d to touch)
      out->hasGradients = 1;
      for (int k = 0; k < 3; k++) {
        out->grad1[k] = (cp1[k] - cp2[k]) / d64;
        out->grad2[k] = -out->grad1[k];
      }
    }
  }
}

}  // namespace

extern "C" int kg_distance_batch_dev(kg_world w, int ri, const float* dQ, int N, float* dout, void* stream) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(N >= 0, "negative batch size");
    need(N == 0 || (dQ && dout), "NULL buffer");
    ensure_status_host(w->ctx);
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(w->ctx, s);
    distance_dev(w, ri, dQ, nullptr, N, dout, nullptr, 0.0f, 3.0e38f, s);
  });
}

namespace {
void distance_jacobian_dev(kg_world_s* w, int ri, const float* dQ, int N, float dcheck, float* d_out, float* cp_out,
                           float* grad_out, hipStream_t s) {
  kg_ctx_s* c = w->ctx;
  need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
  need(dcheck > 0, "dcheck must be > 0");
  RobotPlan& P = get_plan(w, ri);
  const int np = (int)P.pairs.size();
  if (N <= 0 || np == 0) return;
  hip_check(c->wsDl.ensure((size_t)N * np * sizeof(float)), "hipMalloc");
  hip_check(c->wsEA.ensure((size_t)N * np * 2 * sizeof(int)), "hipMalloc");
  distance_dev(w, ri, dQ, nullptr, N, c->wsDl.as<float>(), c->wsEA.as<int>(), 0.0f, dcheck, s);
  DistJacArgs a{};
  a.pool = c->pool;
  a.plan = P.dev;
  a.T64 = c->wsT64.as<double>();
  a.N = N;
  a.npairs = np;
  a.n = P.dev.nlinks;
  a.pair_a = P.dpa.as<int>();
  a.pair_b = P.dpb.as<int>();
  a.pair_kind = P.dpk.as<int>();
  a.dist = c->wsDl.as<float>();
  a.tri = c->wsEA.as<int>();
  a.dcheck = dcheck;
  a.d_out = d_out;
  a.cp_out = cp_out;
  a.grad_out = grad_out;
  hip_check(kgl_dist_jac(a, s), "dist_jac");
}
}  // namespace

// ---------------------------------------------------------------- KineTrajOpt sweep terms
namespace {

// FP64 points of a mesh or point cloud on the device (its convex hull's
// support map reads them); nullptr for an SDF or empty geometry.
const double* hull_points(kg_geom_s* g, int& n) {
  n = 0;
  if (!g || g->kind == 1 || g->V.empty() || (g->kind == 0 && geom_empty(g))) return nullptr;
  n = (int)(g->V.size() / 3);
  if (!g->dv64.p) {
    hip_check(g->dv64.ensure(g->V.size() * sizeof(double)), "hipMalloc");
    hip_check(hipMemcpy(g->dv64.p, g->V.data(), g->V.size() * sizeof(double), hipMemcpyHostToDevice), "copy points");
  }
  return g->dv64.as<double>();
}

// Hull table of a robot's plan: its links' hulls (link frames come from FK),
// then the obstacles in world ID order, terrains then rigid objects, each at
// its placement.  Rebuilt with the plan.
void ensure_hulls(kg_world_s* w, int ri, RobotPlan& P) {
  if (P.hull_version == P.built_version) return;
  kg_robot_s* R = w->robots[ri];
  std::vector<HullRef> T;
  for (int l = 0; l < R->n; l++) {
    HullRef h{};
    h.V = hull_points(R->link_geoms[l], h.nv);
    T.push_back(h);
  }
  auto add_obs = [&](kg_geom_s* g, const Xf64& X) {
    HullRef h{};
    h.V = hull_points(g, h.nv);
    for (int i = 0; i < 9; i++) h.X[i] = X.R[i];
    for (int i = 0; i < 3; i++) h.X[9 + i] = X.t[i];
    T.push_back(h);
  };
  for (kg_geom_s* g : w->terrains) add_obs(g, g->X);
  for (auto& o : w->objects) add_obs(o.first, o.second);
  P.nobs = (int)T.size() - R->n;
  upload(P.dhull, T);
  P.hull_version = P.built_version;
}

void sweep_dev(kg_world_s* w, int ri, const double* dQ0, const double* dQ1, int N, double dcheck, double* d_out,
               double* info_out, double* coef_out, hipStream_t s) {
  need(w != nullptr, "world is NULL");
  need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
  need(N >= 0, "negative batch size");
  need(dcheck > 0, "dcheck must be positive");
  kg_ctx_s* c = w->ctx;
  RobotPlan& P = get_plan(w, ri);
  ensure_hulls(w, ri, P);
  const int n = w->robots[ri]->n;
  if (N == 0 || P.nobs == 0) return;
  need(dQ0 && dQ1 && d_out && info_out && coef_out, "NULL buffer");
  hip_check(c->wsT64.ensure((size_t)2 * N * n * 12 * sizeof(double)), "hipMalloc(T64)");
  double* T0 = c->wsT64.as<double>();
  double* T1 = T0 + (size_t)N * n * 12;
  hip_check(kgl_fk64(P.dev.links, n, nullptr, dQ0, N, T0, nullptr, 0, s), "fk64");
  hip_check(kgl_fk64(P.dev.links, n, nullptr, dQ1, N, T1, nullptr, 0, s), "fk64");
  SweepArgs a{};
  a.links = P.dev.links;
  a.n = n;
  a.N = N;
  a.O = P.nobs;
  a.T0 = T0;
  a.T1 = T1;
  a.lhull = P.dhull.as<HullRef>();
  a.ohull = P.dhull.as<HullRef>() + n;
  a.dcheck = dcheck;
  a.d_out = d_out;
  a.info_out = info_out;
  a.coef_out = coef_out;
  hip_check(kgl_sweep(a, s), "sweep");
}

}  // namespace

extern "C" int kg_world_num_obstacles(kg_world w, int* out) {
  return guard([&] {
    need(w && out, "NULL argument");
    *out = (int)(w->terrains.size() + w->objects.size());
  });
}

extern "C" int kg_sweep_batch_dev(kg_world w, int ri, const double* dQ0, const double* dQ1, int N, double dcheck,
                                  double* d_out, double* info_out, double* coef_out, void* stream) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    hip_check(hipSetDevice(w->ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(w->ctx, s);
    sweep_dev(w, ri, dQ0, dQ1, N, dcheck, d_out, info_out, coef_out, s);
  });
}

extern "C" int kg_sweep_batch(kg_world w, int ri, const double* Q0, const double* Q1, int N, double dcheck,
                              double* d_out, double* info_out, double* coef_out) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
    need(N >= 0, "negative batch size");
    if (N == 0) return;
    need(Q0 && Q1 && d_out && info_out && coef_out, "NULL buffer");
    kg_ctx_s* c = w->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    const int n = w->robots[ri]->n;
    const size_t O = w->terrains.size() + w->objects.size();
    const size_t nq = (size_t)N * n, nr = (size_t)N * n * O;
    if (nr == 0) return;
    hip_check(c->wsEA.ensure(2 * nq * sizeof(double)), "hipMalloc");
    hip_check(c->wsEB.ensure(nr * (1 + 16 + 2 * (size_t)n) * sizeof(double)), "hipMalloc");
    double* dQ = c->wsEA.as<double>();
    double* dd = c->wsEB.as<double>();
    double* di = dd + nr;
    double* dc = di + nr * 16;
    hip_check(hipMemcpyAsync(dQ, Q0, nq * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    hip_check(hipMemcpyAsync(dQ + nq, Q1, nq * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    sweep_dev(w, ri, dQ, dQ + nq, N, dcheck, dd, di, dc, c->stream);
    hip_check(hipMemcpyAsync(d_out, dd, nr * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipMemcpyAsync(info_out, di, nr * 16 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipMemcpyAsync(coef_out, dc, nr * 2 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
  });
}

extern "C" int kg_distance_jacobian_batch_dev(kg_world w, int ri, const float* dQ, int N, float dcheck, float* d_out,
                                              float* cp_out, float* grad_out, void* stream) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(N >= 0, "negative batch size");
    need(N == 0 || (dQ && d_out && cp_out && grad_out), "NULL buffer");
    ensure_status_host(w->ctx);
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(w->ctx, s);
    distance_jacobian_dev(w, ri, dQ, N, dcheck, d_out, cp_out, grad_out, s);
  });
}

extern "C" int kg_distance_jacobian_batch(kg_world w, int ri, const float* Q, int N, float dcheck, float* d_out,
                                          float* cp_out, float* grad_out) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(N >= 0, "negative batch size");
    need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
    if (N == 0) return;
    need(Q && d_out && cp_out && grad_out, "NULL buffer");
    kg_ctx_s* c = w->ctx;
    ensure_status_host(c);
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    RobotPlan& P = get_plan(w, ri);
    const int n = w->robots[ri]->n, np = (int)P.pairs.size();
    if (np == 0) return;
    hip_check(c->wsQ.ensure((size_t)N * n * sizeof(float)), "hipMalloc");
    hip_check(c->wsF.ensure((size_t)N * np * sizeof(float)), "hipMalloc");
    hip_check(c->wsEB.ensure((size_t)N * np * 6 * sizeof(float)), "hipMalloc");
    hip_check(c->wsEQ.ensure((size_t)N * np * n * sizeof(float)), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsQ.p, Q, (size_t)N * n * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy");
    distance_jacobian_dev(w, ri, c->wsQ.as<float>(), N, dcheck, c->wsF.as<float>(), c->wsEB.as<float>(),
                          c->wsEQ.as<float>(), c->stream);
    hip_check(hipMemcpyAsync(d_out, c->wsF.p, (size_t)N * np * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipMemcpyAsync(cp_out, c->wsEB.p, (size_t)N * np * 6 * sizeof(float), hipMemcpyDeviceToHost, c->stream),
              "copy");
    hip_check(hipMemcpyAsync(grad_out, c->wsEQ.p, (size_t)N * np * n * sizeof(float), hipMemcpyDeviceToHost, c->stream),
              "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    check_uncertain(c);
  });
}

extern "C" int kg_distance_batch(kg_world w, int ri, const float* Q, int N, float* out) {
  return guard([&] {
    need(w != nullptr, "world is NULL");
    need(N >= 0, "negative batch size");
    need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
    if (N == 0) return;
    need(Q && out, "NULL buffer");
    kg_ctx_s* c = w->ctx;
    ensure_status_host(c);
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    RobotPlan& P = get_plan(w, ri);
    const int n = w->robots[ri]->n, np = (int)P.pairs.size();
    if (np == 0) return;
    hip_check(c->wsQ.ensure((size_t)N * n * sizeof(float)), "hipMalloc");
    hip_check(c->wsF.ensure((size_t)N * np * sizeof(float)), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsQ.p, Q, (size_t)N * n * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy");
    distance_dev(w, ri, c->wsQ.as<float>(), nullptr, N, c->wsF.as<float>(), nullptr, 0.0f, 3.0e38f, c->stream);
    hip_check(hipMemcpyAsync(out, c->wsF.p, (size_t)N * np * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    check_uncertain(c);
  });
}

extern "C" int kg_geom_distance(kg_geom a, kg_geom b, const kg_dist_settings* settings, kg_dist_result* out) {
  return guard([&] {
    need(a && b && out, "NULL argument");
    need(a->ctx == b->ctx, "geometries belong to different contexts");
    if (a->kind != 0 || b->kind != 0) {
      kg_geom_s *p = nullptr, *f = nullptr;
      if (!point_set_pair(a, b, p, f) || p->kind != 2)
        fail(KG_E_UNSUPPORTED, "distance is implemented for TriangleMesh pairs and PointCloud / ImplicitSurface");
      if (point_set_empty(p))
        fail(KG_E_EMPTY, p == a ? "Geometry3D.distance: Geometry is empty"
                                : "Geometry3D.distance: Other geometry is empty");
      int idx = 0;
      const double v = pts_sdf_min(p, f, &idx);
      // the minimising point in world coordinates, and the surface gradient there
      const double* q = &p->V[3 * (size_t)idx];
      double w[3];
      for (int k = 0; k < 3; k++) w[k] = p->X.R[k] * q[0] + p->X.R[3 + k] * q[1] + p->X.R[6 + k] * q[2] + p->X.t[k];
      float wf[3] = {(float)w[0], (float)w[1], (float)w[2]}, dv = 0, gr[3] = {0, 0, 0};
      const int rc = kg_sdf_eval(f, wf, 1, &dv, gr);
      if (rc != KG_OK) fail(rc, "%s", g_err.c_str());
      std::memset(out, 0, sizeof(*out));
      out->d = v - (a->margin + b->margin);
      out->hasClosestPoints = 1;
      out->hasGradients = 1;
      const bool pa = p == a;
      for (int k = 0; k < 3; k++) {
        const double surf = w[k] - v * gr[k];
        out->cp1[k] = pa ? w[k] : surf;
        out->cp2[k] = pa ? surf : w[k];
        out->grad1[k] = pa ? gr[k] : -gr[k];
        out->grad2[k] = pa ? -gr[k] : gr[k];
      }
      out->elem1 = pa ? idx : -1;
      out->elem2 = pa ? -1 : idx;
      return;
    }
    if (geom_empty(a)) fail(KG_E_EMPTY, "Geometry3D.distance: Geometry is empty");
    if (geom_empty(b)) fail(KG_E_EMPTY, "Geometry3D.distance: Other geometry is empty");
    const double m = a->margin + b->margin;
    const double upper = settings ? settings->upperBound + m : INFINITY;
    const double rel = settings ? settings->relErr : 0, abse = settings ? settings->absErr : 0;
    need(rel >= 0 && abse >= 0, "relErr and absErr must be >= 0");
    pair_distance(a, b, 0.0f, upper, out, nullptr, rel, abse);
    // margins shrink the distance (negative when the fattened geometries overlap)
    out->d -= m;
  });
}

extern "C" int kg_geom_within_distance(kg_geom a, kg_geom b, double tol, int* out) {
  return guard([&] {
    need(a && b && out, "NULL argument");
    need(tol >= 0, "tolerance must be >= 0");
    *out = 0;
    if (a->kind != 0 || b->kind != 0) {
      kg_geom_s *p = nullptr, *f = nullptr;
      if (!point_set_pair(a, b, p, f))
        fail(KG_E_UNSUPPORTED,
             "withinDistance is implemented for TriangleMesh pairs and PointCloud / TriangleMesh against "
             "ImplicitSurface");
      if (point_set_empty(p)) return;
      int idx = 0;
      *out = pts_sdf_min(p, f, &idx) <= tol + a->margin + b->margin ? 1 : 0;
      return;
    }
    if (geom_empty(a) || geom_empty(b)) return;
    const double thr = tol + a->margin + b->margin;
    kg_dist_result r;
    float dg = 0;
    pair_distance(a, b, (float)thr, INFINITY, &r, &dg);
    *out = r.d <= thr ? 1 : 0;
  });
}

// ---------------------------------------------------------------- edge checks
namespace {

// Straight-line edge checks on device buffers (SingleRobotCSpace::PathChecker
// -> EpsilonEdgeChecker, Cpp/Planning/RobotCSpace.cpp:835-837).  Rounds of
// plan / samples / feasibility pipeline / reduce run on the device with no
// host synchronization; the host only launches a schedule fixed in advance:
// round r is w0 g^r positions wide (kg_params.edge_w0, edge_growth), at most the sample
// budget over the live segments, and the schedule has enough rounds to cover
// max_pos positions of every segment even at the narrowest width the budget
// allows.  Most colliding segments fail at their first bisection samples
// (on the 8192 arm segments in the 200k soup, 3 of 4 at the midpoint), while
// free segments must check every sample: a narrow first round wastes little
// on the former, and fast-growing ones cover the latter in few rounds.  A
// round costs at least the traversal's critical path (0.1-0.2 ms there,
// against 2-4 ns per sample), so few wide rounds beat many narrow ones:
// widths 2, 32, 512 take 3 rounds to the 6 of 1, 4, 16, ...  A last round
// that would be no wider than the one before it is folded into that one
// (2, 32, 596 for a joint-range diagonal of 630 positions: the 4th round
// would only cover 84 positions, and every round costs its launches).
// Segments needing more positions than max_pos come out -2 (undecided) and
// are counted in the status words.  A batch of more segments than the
// budget's samples runs as consecutive chunks.
void edge_core(kg_world_s* w, int ri, const float* Af, const float* Bf, const double* Ad, const double* Bd, int nseg,
               double eps, int norm_inf, int max_pos, float* first_u, int32_t* nsamples, hipStream_t s,
               kg_ctx_s::EdgeSlot* es = nullptr) {
  kg_ctx_s* c = w->ctx;
  need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
  need(nseg >= 0, "negative segment count");
  need(eps > 0 && std::isfinite(eps), "eps must be positive");  // motionplanning.cpp:727-728
  need(max_pos >= 0, "negative position bound");
  if (nseg == 0) return;
  reject_margins(w);
  ensure_status_host(c);
  RobotPlan& P = get_plan(w, ri);
  const int n = P.dev.nlinks;
  const kg_params& prm = c->prm;
  // feasibility workspace bytes per sample (fixed-capacity item layout)
  const int cap = active_links(w->robots[ri]) * P.nenv + P.nself;
  const bool group = prm.prep_group && n <= 32;  // the broad phase feasible_core picks
  size_t per = (size_t)n * sizeof(double) + 1 + 1 + 3 * sizeof(int) + (size_t)cap * sizeof(PairItem) +
               (group ? (size_t)cap * sizeof(ItemRec) : (size_t)n * 18 * sizeof(double));
  const size_t budget_bytes = (size_t)prm.edge_budget_mb << 20;
  int B = (int)std::max<size_t>(1, std::min<size_t>(budget_bytes / per, (size_t)1 << 24));
  // the item layout of one round must stay under the traversal's 2 GiB limit
  B = (int)std::min<size_t>((size_t)B, std::max<size_t>(1, ((size_t)2048 << 20) / std::max<size_t>(1, (size_t)cap * sizeof(PairItem))));
  const int chunk = std::min(nseg, B);

  // the edge buffers: the slot's, or the context's for a plain call
  DevBuf& bNs = es ? es->ns : c->wsENs;
  DevBuf& bCur = es ? es->cur : c->wsECur;
  DevBuf& bDone = es ? es->done : c->wsEDone;
  DevBuf& bSeg = es ? es->seg : c->wsESeg;
  DevBuf& bSoff = es ? es->soff : c->wsESoff;
  DevBuf& bCtl = es ? es->ctl : c->wsECtl;
  DevBuf& bQ = es ? es->q : c->wsEQ;
  DevBuf& bFeas = es ? es->feas : c->wsEFeas;
  DevBuf& bStat = es ? es->stat : c->wsEStat;
  hip_check(bNs.ensure((size_t)nseg * sizeof(int)), "hipMalloc");
  hip_check(bCur.ensure((size_t)nseg * sizeof(int)), "hipMalloc");
  hip_check(bDone.ensure((size_t)nseg), "hipMalloc");
  hip_check(bSeg.ensure((size_t)chunk * sizeof(int)), "hipMalloc");
  hip_check(bSoff.ensure((size_t)(chunk + 1) * sizeof(int)), "hipMalloc");
  hip_check(bCtl.ensure(8 * sizeof(int)), "hipMalloc");
  hip_check(bQ.ensure((size_t)B * n * sizeof(double)), "hipMalloc");
  hip_check(bFeas.ensure((size_t)B), "hipMalloc");
  hip_check(bStat.ensure(ST_WORDS * sizeof(int)), "hipMalloc");
  FeasWs ws;
  if (es) {
    need(group, "edge workspace slots need the per-link broad phase");
    ws = slot_ws(c, es->f, B, (size_t)cap * B);
  } else {
    hip_check(c->wsStatus.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
  }

  Layout L;
  L.P = &P;
  L.cap = cap;
  L.maxcap = cap;
  L.nrows = (size_t)B * n;
  L.ncap = (size_t)cap * B;
  L.tol32 = P.tol32;
  L.tol64 = P.tol64;

  EdgeArgs ea{};
  ea.n = n;
  ea.eps = eps;
  ea.norm_inf = norm_inf ? 1 : 0;
  ea.max_pos = max_pos;
  ea.budget = B;
  ea.seg = bSeg.as<int>();
  ea.soff = bSoff.as<int>();
  ea.ctl = bCtl.as<int>();
  ea.Qd = bQ.as<double>();
  ea.feas = bFeas.as<uint8_t>();
  ea.st = es ? ws.status : c->wsStatus.as<int>();
  ea.w_spill = ST_SPILL;
  ea.w_unc = ST_UNC;
  ea.w_redo = ST_REDO;
  ea.out_status = bStat.as<int>();
  // the control words accumulate over the chunks: cleared once, by the first
  // chunk's init kernel
  for (int s0 = 0; s0 < nseg; s0 += chunk) {
    const int ns = std::min(chunk, nseg - s0);
    ea.nseg = ns;
    ea.Af = Af ? Af + (size_t)s0 * n : nullptr;
    ea.Bf = Bf ? Bf + (size_t)s0 * n : nullptr;
    ea.Ad = Ad ? Ad + (size_t)s0 * n : nullptr;
    ea.Bd = Bd ? Bd + (size_t)s0 * n : nullptr;
    ea.ns = bNs.as<int>() + s0;
    ea.cur = bCur.as<int>() + s0;
    ea.done = bDone.as<uint8_t>() + s0;
    ea.first_u = first_u + s0;
    ea.nsamples_out = nsamples ? nsamples + s0 : nullptr;
    ea.ctl = bCtl.as<int>();
    hip_check(kgl_edge_init(ea, s0 == 0, s), "edge_init");
    // schedule: widths w0 g^r; each round gives every live segment at least
    // min(width, B / ns) positions
    const int floor_w = std::max(1, B / ns);
    int64_t covered = 0;
    int64_t wr = prm.edge_w0;
    for (int r = 0; covered < max_pos; r++, wr = std::min<int64_t>(wr * prm.edge_growth, 1 << 20)) {
      int64_t wf = wr;
      if (covered + wf < max_pos && max_pos - covered - wf <= wf) wf = max_pos - covered;
      const int width = (int)std::min<int64_t>(wf, 1 << 20);
      covered += std::min(width, floor_w);
      hip_check(kgl_edge_round(ea, width, 0, s), "edge_plan");
      feasible_core(c, L, nullptr, ea.Qd, B, bFeas.as<uint8_t>(), nullptr, nullptr, nullptr, s, ea.ctl + 1, false,
                    es ? &ws : nullptr);
      hip_check(kgl_edge_round(ea, width, 1, s), "edge_reduce");
    }
    hip_check(kgl_edge_finish(ea, s0 + chunk >= nseg, s), "edge_finish");
  }
  c->last_N = nseg;
  c->has_batch = true;
  c->last_tr_slot = -1;
  c->last_phase_ev = false;
  c->stat_src = bStat.as<int>();
  hip_check(hipEventRecord(c->evStat, s), "hipEventRecord");
}

// Positions every segment of a batch may need: max_steps - 1, or from the
// joint ranges when max_steps is 0 (no segment between two configurations
// within the limits is longer than the range box's diagonal).
int edge_max_pos(const kg_robot_s* r, double eps, int norm_inf, int max_steps) {
  if (max_steps > 0) return max_steps - 1;
  double acc = 0;
  for (int k = 0; k < r->n; k++) {
    const double d = r->qmax[k] - r->qmin[k];
    if (norm_inf) acc = std::max(acc, std::fabs(d));
    else acc += d * d;
  }
  const double D = norm_inf ? acc : std::sqrt(acc);
  const double m = std::ceil(D / eps) + 1;
  need(std::isfinite(m) && m < 1e9,
       "unbounded joint range: give max_steps (the most steps a segment of the batch may need)");
  return (int)m;
}

// Host-pointer edge checks: exact steps per segment on the host (the
// schedule covers the longest), the device rounds, one copy back.
void edge_host(kg_world_s* w, int ri, const float* Af, const float* Bf, const double* Ad, const double* Bd, int nseg,
               double eps, int norm_inf, float* first_u) {
  need(w != nullptr, "world is NULL");
  need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
  need(nseg >= 0, "negative segment count");
  need(eps > 0 && std::isfinite(eps), "eps must be positive");  // motionplanning.cpp:727-728
  if (nseg == 0) return;
  need((Af || Ad) && (Bf || Bd) && first_u, "NULL buffer");
  kg_ctx_s* c = w->ctx;
  const int n = w->robots[ri]->n;
  int maxm = 0;
  for (int s = 0; s < nseg; s++) {
#pragma clang fp contract(off)
    double acc = 0;
    for (int k = 0; k < n; k++) {
      const size_t i = (size_t)s * n + k;
      const double d = (Ad ? Ad[i] : (double)Af[i]) - (Bd ? Bd[i] : (double)Bf[i]);
      if (norm_inf) acc = std::max(acc, std::fabs(d));
      else acc += d * d;
    }
    const double m = std::ceil((norm_inf ? acc : std::sqrt(acc)) / eps);
    need(m < 1e7, "segment needs more than 1e7 samples");  // also rejects NaN
    maxm = std::max(maxm, (int)m);
  }
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  StreamScope scope(c, c->stream);
  const size_t eb = Ad ? sizeof(double) : sizeof(float);
  hip_check(c->wsEA.ensure((size_t)nseg * n * eb), "hipMalloc");
  hip_check(c->wsEB.ensure((size_t)nseg * n * eb), "hipMalloc");
  hip_check(c->wsEU.ensure((size_t)nseg * sizeof(float)), "hipMalloc");
  hip_check(hipMemcpyAsync(c->wsEA.p, Ad ? (const void*)Ad : (const void*)Af, (size_t)nseg * n * eb,
                           hipMemcpyHostToDevice, c->stream), "copy");
  hip_check(hipMemcpyAsync(c->wsEB.p, Bd ? (const void*)Bd : (const void*)Bf, (size_t)nseg * n * eb,
                           hipMemcpyHostToDevice, c->stream), "copy");
  // one position of margin over the host's step counts
  edge_core(w, ri, Ad ? nullptr : c->wsEA.as<float>(), Bd ? nullptr : c->wsEB.as<float>(),
            Ad ? c->wsEA.as<double>() : nullptr, Ad ? c->wsEB.as<double>() : nullptr, nseg, eps, norm_inf, maxm,
            c->wsEU.as<float>(), nullptr, c->stream);
  hip_check(hipMemcpyAsync(first_u, c->wsEU.p, (size_t)nseg * sizeof(float), hipMemcpyDeviceToHost, c->stream),
            "copy");
  hip_check(hipStreamSynchronize(c->stream), "sync");
  check_uncertain(c);
}

}  // namespace

extern "C" int kg_edge_batch(kg_world w, int ri, const float* A, const float* B, int nseg, double eps, int norm_inf,
                             float* first_u) {
  return guard([&] { edge_host(w, ri, A, B, nullptr, nullptr, nseg, eps, norm_inf, first_u); });
}

extern "C" int kg_edge_batch_f64(kg_world w, int ri, const double* A, const double* B, int nseg, double eps,
                                 int norm_inf, float* first_u) {
  return guard([&] { edge_host(w, ri, nullptr, nullptr, A, B, nseg, eps, norm_inf, first_u); });
}

namespace {
void edge_dev(kg_world w, int ri, const float* Af, const float* Bf, const double* Ad, const double* Bd, int nseg,
              double eps, int norm_inf, int max_steps, float* first_u, int32_t* nsamples, void* stream) {
  need(w != nullptr, "world is NULL");
  need(ri >= 0 && ri < (int)w->robots.size(), "robot index out of range");
  need(nseg >= 0, "negative segment count");
  need(max_steps >= 0, "negative max_steps");
  need(eps > 0 && std::isfinite(eps), "eps must be positive");
  if (nseg == 0) return;
  need((Af || Ad) && (Bf || Bd) && first_u, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  kg_ctx_s* c = w->ctx;
  const int mp = edge_max_pos(w->robots[ri], eps, norm_inf, max_steps);
  // on the context's next edge slot (per-link broad phase): the batch waits
  // for the previous batch on its slot and for the last plain call, not for
  // batches on other slots, so batches issued on different streams overlap
  if (!(c->prm.prep_group && w->robots[ri]->n <= 32)) {
    StreamScope scope(c, s);
    edge_core(w, ri, Af, Bf, Ad, Bd, nseg, eps, norm_inf, mp, first_u, nsamples, s);
    return;
  }
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  if (!c->evDone)
    hip_check(hipEventCreateWithFlags(&c->evDone, hipEventDisableTiming | hipEventDisableSystemFence),
              "hipEventCreate");
  kg_ctx_s::EdgeSlot& es = c->eslot[c->eslot_next];
  c->eslot_next = (c->eslot_next + 1) % kg_ctx_s::EDGE_SLOTS;
  if (!es.f.done)
    hip_check(hipEventCreateWithFlags(&es.f.done, hipEventDisableTiming | hipEventDisableSystemFence),
              "hipEventCreate");
  const bool cap = capturing(s);
  if (es.f.valid && !cap) hip_check(hipStreamWaitEvent(s, es.f.done, 0), "hipStreamWaitEvent");
  if (c->done_valid && !c->last_slotted && c->done_stream != s && !cap)
    hip_check(hipStreamWaitEvent(s, c->evDone, 0), "hipStreamWaitEvent");
  edge_core(w, ri, Af, Bf, Ad, Bd, nseg, eps, norm_inf, mp, first_u, nsamples, s, &es);
  hip_check(hipEventRecord(es.f.done, s), "hipEventRecord");
  es.f.valid = true;
  if (hipEventRecord(c->evDone, s) == hipSuccess) {
    c->done_stream = s;
    c->done_valid = true;
    c->last_slotted = true;
  }
}
}  // namespace

extern "C" int kg_edge_batch_dev(kg_world w, int ri, const float* dA, const float* dB, int nseg, double eps,
                                 int norm_inf, int max_steps, float* d_first_u, int32_t* d_nsamples, void* stream) {
  return guard([&] { edge_dev(w, ri, dA, dB, nullptr, nullptr, nseg, eps, norm_inf, max_steps, d_first_u, d_nsamples,
                              stream); });
}

extern "C" int kg_edge_batch_f64_dev(kg_world w, int ri, const double* dA, const double* dB, int nseg, double eps,
                                     int norm_inf, int max_steps, float* d_first_u, int32_t* d_nsamples,
                                     void* stream) {
  return guard([&] { edge_dev(w, ri, nullptr, nullptr, dA, dB, nseg, eps, norm_inf, max_steps, d_first_u, d_nsamples,
                              stream); });
}

extern "C" int kg_edge_sample_index(int nsteps, int k) {
  if (nsteps < 2 || k < 0 || k >= nsteps - 1) return -1;
  return kg_bisect_index(nsteps, k);
}

// ---------------------------------------------------------------- heterogeneous batches
namespace {

struct MultiBatch {
  std::vector<RobotPlan*> plans;
  std::vector<int> nlinks;
  size_t N = 0;
};

// Build (or reuse) the per-configuration layout tables of a heterogeneous
// batch: plan index, Q offset, frame offset, item base.
Layout multi_layout(kg_ctx_s* c, int nb, kg_world_s* const* worlds, const int32_t* ris, const int32_t* counts,
                    hipStream_t s, size_t* Nout) {
  std::vector<PlanDev> devplans(nb);
  std::vector<uint64_t> sig;
  size_t N = 0, rows = 0, caps = 0;
  float pad_rel = 0, pad_bv64 = 0;
  Tol t32{}, t64{};
  std::vector<int> caps_b(nb), n_b(nb);
  for (int b = 0; b < nb; b++) {
    need(worlds[b] != nullptr, "NULL world");
    need(worlds[b]->ctx == c, "worlds belong to different contexts");
    reject_margins(worlds[b]);
    need(counts[b] >= 0, "negative configuration count");
    RobotPlan& P = get_plan(worlds[b], ris[b]);
    devplans[b] = P.dev;
    n_b[b] = P.dev.nlinks;
    caps_b[b] = active_links(worlds[b]->robots[ris[b]]) * P.nenv + P.nself;
    N += counts[b];
    rows += (size_t)counts[b] * n_b[b];
    caps += (size_t)counts[b] * caps_b[b];
    pad_rel = std::max(pad_rel, P.tol32.pad_rel);
    pad_bv64 = std::max(pad_bv64, P.tol64.pad_bv);
    t32 = P.tol32;
    sig.push_back((uint64_t)(uintptr_t)worlds[b]);
    sig.push_back(((uint64_t)ris[b] << 32) | (uint32_t)counts[b]);
    sig.push_back((uint64_t)P.built_version);
  }
  need(N < (1ull << 31) && rows < (1ull << 31) && caps < (1ull << 31), "heterogeneous batch too large");
  t32.pad_rel = pad_rel;
  t64 = t32;
  t64.ev = 0;
  t64.tau = 0;
  t64.pad_bv = pad_bv64;
  if (sig != c->mSig) {
    std::vector<int> cfg_plan(N), qoff(N), toff(N), ibase(N);
    size_t k = 0, q = 0, t = 0, ib = 0;
    for (int b = 0; b < nb; b++)
      for (int i = 0; i < counts[b]; i++, k++) {
        cfg_plan[k] = b;
        qoff[k] = (int)q;
        toff[k] = (int)t;
        ibase[k] = (int)ib;
        q += n_b[b];
        t += n_b[b];
        ib += caps_b[b];
      }
    upload(c->mCfgPlan, cfg_plan);
    upload(c->mQoff, qoff);
    upload(c->mToff, toff);
    upload(c->mIbase, ibase);
    c->mSig = sig;
  }
  if (devplans.size() != c->mPlansHost.size() ||
      std::memcmp(devplans.data(), c->mPlansHost.data(), devplans.size() * sizeof(PlanDev)) != 0) {
    upload(c->mPlans, devplans);
    c->mPlansHost = devplans;
  }
  Layout L;
  L.multi = true;
  L.plans = c->mPlans.as<PlanDev>();
  L.cfg_plan = c->mCfgPlan.as<int>();
  L.qoff = c->mQoff.as<int>();
  L.toff = c->mToff.as<int>();
  L.ibase = c->mIbase.as<int>();
  L.nrows = rows;
  L.ncap = caps;
  L.maxcap = *std::max_element(caps_b.begin(), caps_b.end());
  L.maxlinks = *std::max_element(n_b.begin(), n_b.end());
  for (const PlanDev& d : devplans) L.maxdepth = std::max(L.maxdepth, d.depth);
  L.tol32 = t32;
  L.tol64 = t64;
  (void)s;
  *Nout = N;
  return L;
}

}  // namespace

extern "C" int kg_feasible_multi_dev(int nb, const kg_world* worlds, const int32_t* ris, const int32_t* counts,
                                     const float* dQ, uint8_t* dfeas, void* stream) {
  return guard([&] {
    need(nb >= 1 && worlds && ris && counts, "NULL argument");
    kg_ctx_s* c = worlds[0]->ctx;
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(c, s);
    size_t N = 0;
    Layout L = multi_layout(c, nb, worlds, ris, counts, s, &N);
    need(N == 0 || (dQ && dfeas), "NULL buffer");
    feasible_core(c, L, dQ, nullptr, (int)N, dfeas, nullptr, nullptr, nullptr, s);
  });
}

extern "C" int kg_feasible_multi(int nb, const kg_world* worlds, const int32_t* ris, const int32_t* counts,
                                 const float* Q, uint8_t* feas) {
  return guard([&] {
    need(nb >= 1 && worlds && ris && counts, "NULL argument");
    kg_ctx_s* c = worlds[0]->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    StreamScope scope(c, c->stream);
    size_t N = 0, nq = 0;
    Layout L = multi_layout(c, nb, worlds, ris, counts, c->stream, &N);
    for (int b = 0; b < nb; b++) nq += (size_t)counts[b] * worlds[b]->robots[ris[b]]->n;
    if (N == 0) return;
    need(Q && feas, "NULL buffer");
    hip_check(c->wsQ.ensure(nq * sizeof(float)), "hipMalloc");
    hip_check(c->wsOut8a.ensure(N), "hipMalloc");
    hip_check(hipMemcpyAsync(c->wsQ.p, Q, nq * sizeof(float), hipMemcpyHostToDevice, c->stream), "copy Q");
    feasible_core(c, L, c->wsQ.as<float>(), nullptr, (int)N, c->wsOut8a.as<uint8_t>(), nullptr, nullptr, nullptr,
                  c->stream);
    hip_check(hipMemcpyAsync(feas, c->wsOut8a.p, N, hipMemcpyDeviceToHost, c->stream), "copy out");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    check_uncertain(c);
  });
}

// Autotuner: times the traversal on the given heterogeneous batch (device Q)
// for each candidate setting and keeps the fastest on the context.
extern "C" int kg_autotune(int nb, const kg_world* worlds, const int32_t* ris, const int32_t* counts, const float* dQ,
                           int32_t* chosen, float* best_ms) {
  return guard([&] {
    need(nb >= 1 && worlds && ris && counts && dQ, "NULL argument");
    kg_ctx_s* c = worlds[0]->ctx;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    // dQ may have been written on any stream of the caller
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    StreamScope scope(c, c->stream);
    hipStream_t s = c->stream;
    size_t N = 0;
    Layout L = multi_layout(c, nb, worlds, ris, counts, s, &N);
    if (N == 0) return;
    hip_check(c->wsOut8b.ensure(N), "hipMalloc");
    // the timing below reads the per-phase events (timing level 2)
    struct Level {
      kg_ctx_s* c;
      int saved;
      kg_params prm;
      bool done = false;
      ~Level() {
        c->timing = saved;
        if (!done) c->prm = prm;  // an error leaves the parameters as they were
      }
    } level{c, c->timing, c->prm};
    c->timing = 2;
    const int cpws[] = {4, 8, 16, 32};
    const int leafs[] = {1, 8};
    const int scaps[] = {256, 512};
    const int wpcs[] = {0};
    float best = 1e30f;
    int bc[4] = {16, 1, 256, 0};
    for (int scap : scaps)
      for (int wpc : wpcs)
        for (int cpw : cpws)
          for (int leaf : leafs) {
            c->prm.cpw = cpw;
            c->prm.leaf_min = leaf;
            c->prm.scap = scap;
            c->prm.waves_per_cu = wpc;
            float tot = 0;
            for (int rep = 0; rep < 3; rep++) {
              feasible_core(c, L, dQ, nullptr, (int)N, c->wsOut8b.as<uint8_t>(), nullptr, nullptr, nullptr, s);
              hip_check(hipStreamSynchronize(s), "sync");
              float ms = 0;
              hip_check(hipEventElapsedTime(&ms, c->ev[0], c->ev[5]), "hipEventElapsedTime");
              if (rep > 0) tot += ms;
            }
            if (tot < best) {
              best = tot;
              bc[0] = cpw;
              bc[1] = leaf;
              bc[2] = scap;
              bc[3] = wpc;
            }
          }
    c->prm.cpw = bc[0];
    c->prm.leaf_min = bc[1];
    c->prm.scap = bc[2];
    c->prm.waves_per_cu = bc[3];
    level.done = true;
    if (chosen)
      for (int i = 0; i < 4; i++) chosen[i] = bc[i];
    if (best_ms) *best_ms = best / 2;
  });
}

extern "C" int kg_ctx_set_timing(kg_ctx c, int level) {
  return guard([&] {
    need(c != nullptr, "NULL context");
    need(level >= 0 && level <= 2, "timing level must be 0, 1 or 2");
    c->timing = level;
  });
}

extern "C" int kg_ctx_set_tuning(kg_ctx c, int cpw, int leaf_min, int scap, int waves_per_cu) {
  return guard([&] {
    need(c != nullptr, "ctx is NULL");
    need(cpw >= 0 && cpw <= 64 && leaf_min >= 0 && leaf_min <= 64, "tuning value out of range");
    need(scap == 0 || scap == 256 || scap == 512, "scap must be 0, 256 or 512");
    need(waves_per_cu >= 0 && waves_per_cu <= 64, "waves_per_cu must be 0..64");
    const kg_params d = kg_default_params();
    c->prm.cpw = cpw ? cpw : d.cpw;
    c->prm.leaf_min = leaf_min ? leaf_min : d.leaf_min;
    c->prm.scap = scap ? scap : d.scap;
    c->prm.waves_per_cu = waves_per_cu;
  });
}

extern "C" int kg_ctx_get_params(kg_ctx c, kg_params* out) {
  return guard([&] {
    need(c && out, "NULL argument");
    *out = c->prm;
  });
}

extern "C" int kg_ctx_set_params(kg_ctx c, const kg_params* p) {
  return guard([&] {
    need(c && p, "NULL argument");
    auto in = [](int v, int lo, int hi) { return v >= lo && v <= hi; };
    need(in(p->cpw, 1, 64), "cpw must be 1..64");
    need(in(p->leaf_min, 1, 64), "leaf_min must be 1..64");
    need(p->scap == 256 || p->scap == 512, "scap must be 256 or 512");
    need(in(p->waves_per_cu, 0, 64), "waves_per_cu must be 0..64");
    need(in(p->static_pct, 0, 100), "static_pct must be 0..100");
    need(in(p->nsub, 1, ST_NSUB), "nsub must be 1..32");
    need(in(p->tail_lanes, 0, 64), "tail_lanes must be 0..64");
    need(in(p->tail_reps, 1, 8), "tail_reps must be 1..8");
    need(in(p->spill_cap, 0, 1 << 24), "spill_cap must be 0..2^24");
    need(in(p->dist_ipw, 1, 64), "dist_ipw must be 1..64");
    need(in(p->edge_w0, 1, 1 << 20), "edge_w0 must be 1..2^20");
    need(p->sdf_lanes == 8 || p->sdf_lanes == 16 || p->sdf_lanes == 32 || p->sdf_lanes == 64,
         "sdf_lanes must be 8, 16, 32 or 64");
    need(p->prep_group == 0 || p->prep_group == 1 || p->prep_group == 8 || p->prep_group == 16 ||
             p->prep_group == 32,
         "prep_group must be 0, 1, 8, 16 or 32");
    need(in(p->force_recheck, 0, 1), "force_recheck must be 0 or 1");
    need(in(p->edge_budget_mb, 16, 1 << 20), "edge_budget_mb must be 16..2^20");
    need(in(p->edge_growth, 2, 1024), "edge_growth must be 2..1024");
    c->prm = *p;
  });
}

// ---------------------------------------------------------------- serving graph
//
// The batching form of the engine: a heterogeneous feasibility batch (as
// kg_feasible_multi_dev) and the edge batches of several robots (as
// kg_edge_batch_dev each) on fixed device buffers, issued once to size every
// workspace and then captured into one HIP graph that kg_serve_run replays.
// Many SingleRobotCSpace::IsFeasible and PathChecker(a,b)->IsVisible() calls
// of a planner's iteration (Cpp/Planning/RobotCSpace.cpp:786,835-837) become
// one graph launch; the edge rounds compact the live segments on the device.
struct kg_serve_s {
  kg_ctx_s* c = nullptr;
  std::vector<kg_world_s*> fw, ew;
  std::vector<int32_t> fri, fcnt, eri, ecnt;
  const float* dQ = nullptr;
  uint8_t* dfeas = nullptr;
  const float* dA = nullptr;
  const float* dB = nullptr;
  float* dU = nullptr;
  double eps = 0.01;
  int norm_inf = 0;
  int max_steps = 0;
  DevBuf stat;  // ST_WORDS ints: the status words of the whole pipeline
  DevBuf estat;  // ST_WORDS ints: the edge parts' status (their own branch)
  // the feasibility batch's workspaces when it runs beside the edges
  kg_ctx_s::FeasSlot fws;
  hipStream_t cs = nullptr, cs2 = nullptr;
  hipEvent_t eFork = nullptr, eJoin = nullptr;
  bool fork = false;  // two branches: the feasibility batch beside the edge batches
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // what the captured graph baked in
  uint64_t gen = 0;
  kg_params prm{};
  std::vector<int> versions;
  int captures = 0;
};

namespace {

void serve_drop_graph(kg_serve_s* sv) {
  if (sv->exec) (void)hipGraphExecDestroy(sv->exec);
  if (sv->graph) (void)hipGraphDestroy(sv->graph);
  sv->exec = nullptr;
  sv->graph = nullptr;
}

// Everything the pipeline owns (after its streams have drained).
void serve_release(kg_serve_s* sv) {
  if (sv->cs) (void)hipStreamSynchronize(sv->cs);
  if (sv->cs2) (void)hipStreamSynchronize(sv->cs2);
  serve_drop_graph(sv);
  if (sv->cs) (void)hipStreamDestroy(sv->cs);
  if (sv->cs2) (void)hipStreamDestroy(sv->cs2);
  if (sv->eFork) (void)hipEventDestroy(sv->eFork);
  if (sv->eJoin) (void)hipEventDestroy(sv->eJoin);
  sv->cs = sv->cs2 = nullptr;
  sv->eFork = sv->eJoin = nullptr;
  DevBuf* bufs[] = {&sv->stat, &sv->estat, &sv->fws.status, &sv->fws.jl, &sv->fws.counts, &sv->fws.redo,
                    &sv->fws.items, &sv->fws.recs, &sv->fws.spill};
  for (auto* b : bufs) b->release();
}

// The pipeline on stream s.  With sv->fork, two branches: the feasibility
// batch on s (on the pipeline's own workspaces) and the edge batches on cs2
// (the context's), forked and joined with events, so that the edge rounds,
// bound by the traversal's critical path, overlap the batch; otherwise one
// after the other.  Each branch folds its status words into its own
// accumulator; the edge branch's is added after the join.
void serve_issue(kg_serve_s* sv, hipStream_t s) {
  kg_ctx_s* c = sv->c;
  const bool fork = sv->fork;
  hipStream_t se = fork ? sv->cs2 : s;
  if (fork) {
    hip_check(hipEventRecord(sv->eFork, s), "hipEventRecord");
    hip_check(hipStreamWaitEvent(se, sv->eFork, 0), "hipStreamWaitEvent");
  }
  bool have_feas = false;
  if (!sv->fw.empty()) {
    size_t N = 0;
    Layout L = multi_layout(c, (int)sv->fw.size(), sv->fw.data(), sv->fri.data(), sv->fcnt.data(), s, &N);
    need(L.ncap * sizeof(PairItem) <= (2048ull << 20), "serving feasibility batch exceeds the item capacity limit");
    if (N > 0) {
      FeasWs ws;
      if (fork) {
        kg_ctx_s::FeasSlot& f = sv->fws;
        ws.status = f.status.as<int>();
        ws.jl = f.jl.as<uint8_t>();
        ws.counts = f.counts.as<int>();
        ws.redo = f.redo.as<int>();
        ws.items = f.items.as<PairItem>();
        ws.recs = f.recs.as<ItemRec>();
        ws.spill = f.spill.as<uint64_t>();
        ws.spill_entries = (int)std::min<size_t>(f.spill.bytes / sizeof(uint64_t), (size_t)INT32_MAX);
      }
      feasible_core(c, L, sv->dQ, nullptr, (int)N, sv->dfeas, nullptr, nullptr, nullptr, s, nullptr, false,
                    fork ? &ws : nullptr);
      hip_check(kgl_status_accum(fork ? ws.status : c->wsStatus.as<int>(), sv->stat.as<int>(), ST_SPILL, ST_UNC,
                                 ST_REDO, 1, s),
                "status");
      have_feas = true;
    }
  }
  bool have_edges = false;
  size_t qo = 0, so = 0;
  for (size_t g = 0; g < sv->ew.size(); g++) {
    kg_world_s* w = sv->ew[g];
    const int n = w->robots[sv->eri[g]]->n;
    const int cnt = sv->ecnt[g];
    if (cnt > 0) {
      edge_core(w, sv->eri[g], sv->dA + qo, sv->dB + qo, nullptr, nullptr, cnt, sv->eps, sv->norm_inf,
                edge_max_pos(w->robots[sv->eri[g]], sv->eps, sv->norm_inf, sv->max_steps), sv->dU + so, nullptr,
                se);
      hip_check(kgl_status_accum(c->wsEStat.as<int>(), sv->estat.as<int>(), ST_SPILL, ST_UNC, ST_REDO,
                                 have_edges ? 0 : 1, se),
                "status");
      have_edges = true;
    }
    qo += (size_t)cnt * n;
    so += (size_t)cnt;
  }
  if (fork) {
    hip_check(hipEventRecord(sv->eJoin, se), "hipEventRecord");
    hip_check(hipStreamWaitEvent(s, sv->eJoin, 0), "hipStreamWaitEvent");
  }
  if (have_edges)
    hip_check(kgl_status_accum(sv->estat.as<int>(), sv->stat.as<int>(), ST_SPILL, ST_UNC, ST_REDO, have_feas ? 0 : 1,
                               s),
              "status");
  if (!have_feas && !have_edges) hip_check(kgl_fill_int(sv->stat.as<int>(), ST_WORDS, 0, s), "clear status");
}

// Whether the feasibility batch can run beside the edges (its own workspaces:
// the fixed-capacity, per-link broad-phase path), and those workspaces sized.
void serve_size_branches(kg_serve_s* sv) {
  kg_ctx_s* c = sv->c;
  sv->fork = false;
  size_t N = 0, nseg = 0;
  for (int x : sv->fcnt) N += x;
  for (int x : sv->ecnt) nseg += x;
  if (N == 0 || nseg == 0 || !c->prm.prep_group) return;
  Layout L = multi_layout(c, (int)sv->fw.size(), sv->fw.data(), sv->fri.data(), sv->fcnt.data(), c->stream, &N);
  if (L.maxlinks > 32 || L.ncap * sizeof(PairItem) > (2048ull << 20)) return;
  const kg_params& prm = c->prm;
  const int scap = prm.scap == 256 ? 256 : 512;
  const int grid32 = prm.waves_per_cu > 0 ? prm.waves_per_cu * c->num_cus : kgl_collide_grid(0, c->num_cus, scap);
  const int grid64 = std::min(kgl_collide_grid(1, c->num_cus, scap), std::max(1, c->num_cus));
  const size_t spill_entries = (size_t)std::max(grid32, grid64) * prm.spill_cap;
  kg_ctx_s::FeasSlot& f = sv->fws;
  hip_check(f.status.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
  hip_check(f.jl.ensure(N), "hipMalloc(jl)");
  hip_check(f.counts.ensure((N + 1) * sizeof(int)), "hipMalloc(counts)");
  hip_check(f.redo.ensure(redo_ints(N) * sizeof(int)), "hipMalloc(redo)");
  hip_check(f.items.ensure(std::max<size_t>(1, L.ncap) * sizeof(PairItem)), "hipMalloc(items)");
  hip_check(f.recs.ensure(std::max<size_t>(1, L.ncap) * sizeof(ItemRec)), "hipMalloc(records)");
  hip_check(f.spill.ensure(std::max<size_t>(1, spill_entries) * sizeof(uint64_t)), "hipMalloc(spill)");
  sv->fork = true;
}

// Plan versions of every (world, robot) of the pipeline; rebuilding a plan
// here (the world changed) moves g_devbuf_gen when its tables grow.
std::vector<int> serve_versions(kg_serve_s* sv) {
  std::vector<int> v;
  for (size_t b = 0; b < sv->fw.size(); b++) v.push_back(get_plan(sv->fw[b], sv->fri[b]).built_version);
  for (size_t g = 0; g < sv->ew.size(); g++) v.push_back(get_plan(sv->ew[g], sv->eri[g]).built_version);
  return v;
}

// Warm-up issue (sizes every workspace), then the capture.  A capture during
// which a workspace grew (the warm-up and the capture disagree) is dropped
// and taken again.
void serve_capture(kg_serve_s* sv) {
  kg_ctx_s* c = sv->c;
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  serve_drop_graph(sv);
  for (int attempt = 0;; attempt++) {
    need(attempt < 3, "serving pipeline: workspaces kept growing during capture");
    serve_size_branches(sv);
    {
      StreamScope scope(c, sv->cs);
      serve_issue(sv, sv->cs);
    }
    hip_check(hipStreamSynchronize(sv->cs2), "sync");
    hip_check(hipStreamSynchronize(sv->cs), "sync");
    const std::vector<int> ver = serve_versions(sv);
    const uint64_t g0 = g_devbuf_gen.load();
    hip_check(hipStreamBeginCapture(sv->cs, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    try {
      serve_issue(sv, sv->cs);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(sv->cs, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hip_check(hipStreamEndCapture(sv->cs, &sv->graph), "hipStreamEndCapture");
    if (g_devbuf_gen.load() != g0) {
      serve_drop_graph(sv);
      continue;
    }
    hip_check(hipGraphInstantiate(&sv->exec, sv->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    sv->gen = g0;
    sv->prm = c->prm;
    sv->versions = ver;
    sv->captures++;
    // the capture stream's last work (the warm-up) is done; the graph's
    // launches are ordered by kg_serve_run's stream scope
    return;
  }
}

}  // namespace

extern "C" int kg_serve_create(int nbatch, const kg_world* worlds, const int32_t* robot_indices,
                               const int32_t* counts, const float* dQ, uint8_t* d_feasible, int nedge,
                               const kg_world* edge_worlds, const int32_t* edge_robots, const int32_t* edge_counts,
                               const float* dA, const float* dB, double eps, int norm_inf, int max_steps,
                               float* d_first_u, kg_serve* out) {
  return guard([&] {
    need(out != nullptr, "NULL argument");
    *out = nullptr;
    need(nbatch >= 0 && nedge >= 0 && nbatch + nedge >= 1, "empty pipeline");
    need(nbatch == 0 || (worlds && robot_indices && counts), "NULL feasibility batch arrays");
    need(nedge == 0 || (edge_worlds && edge_robots && edge_counts), "NULL edge batch arrays");
    need(eps > 0 && std::isfinite(eps), "eps must be positive");  // motionplanning.cpp:727-728
    need(max_steps >= 0, "negative max_steps");
    kg_ctx_s* c = nbatch ? worlds[0]->ctx : edge_worlds[0]->ctx;
    std::unique_ptr<kg_serve_s> sv(new kg_serve_s);
    sv->c = c;
    size_t N = 0;
    for (int b = 0; b < nbatch; b++) {
      need(worlds[b] && worlds[b]->ctx == c, "worlds belong to different contexts");
      need(robot_indices[b] >= 0 && robot_indices[b] < (int)worlds[b]->robots.size(), "robot index out of range");
      need(counts[b] >= 0, "negative configuration count");
      sv->fw.push_back(worlds[b]);
      sv->fri.push_back(robot_indices[b]);
      sv->fcnt.push_back(counts[b]);
      N += counts[b];
    }
    need(N == 0 || (dQ && d_feasible), "NULL feasibility buffer");
    size_t nseg = 0;
    for (int g = 0; g < nedge; g++) {
      need(edge_worlds[g] && edge_worlds[g]->ctx == c, "worlds belong to different contexts");
      need(edge_robots[g] >= 0 && edge_robots[g] < (int)edge_worlds[g]->robots.size(), "robot index out of range");
      need(edge_counts[g] >= 0, "negative segment count");
      sv->ew.push_back(edge_worlds[g]);
      sv->eri.push_back(edge_robots[g]);
      sv->ecnt.push_back(edge_counts[g]);
      nseg += edge_counts[g];
      // rejects an unbounded joint range when max_steps is 0
      (void)edge_max_pos(edge_worlds[g]->robots[edge_robots[g]], eps, norm_inf, max_steps);
    }
    need(nseg == 0 || (dA && dB && d_first_u), "NULL edge buffer");
    sv->dQ = dQ;
    sv->dfeas = d_feasible;
    sv->dA = dA;
    sv->dB = dB;
    sv->dU = d_first_u;
    sv->eps = eps;
    sv->norm_inf = norm_inf ? 1 : 0;
    sv->max_steps = max_steps;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hip_check(sv->stat.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
    hip_check(sv->estat.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
    hip_check(hipStreamCreateWithFlags(&sv->cs, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&sv->cs2, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&sv->eFork, hipEventDisableTiming | hipEventDisableSystemFence), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&sv->eJoin, hipEventDisableTiming | hipEventDisableSystemFence), "hipEventCreate");
    try {
      serve_capture(sv.get());
    } catch (...) {
      serve_release(sv.get());
      throw;
    }
    *out = sv.release();
  });
}

extern "C" int kg_serve_run(kg_serve sv, void* stream) {
  return guard([&] {
    need(sv != nullptr, "NULL argument");
    kg_ctx_s* c = sv->c;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    // recapture when what the graph baked in has changed: a plan (the world
    // or its geometry), the engine parameters, or a workspace address
    const std::vector<int> ver = serve_versions(sv);
    if (ver != sv->versions || std::memcmp(&sv->prm, &c->prm, sizeof(kg_params)) != 0 ||
        g_devbuf_gen.load() != sv->gen)
      serve_capture(sv);
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(c, s);
    hip_check(hipGraphLaunch(sv->exec, s), "hipGraphLaunch");
    c->stat_src = sv->stat.as<int>();
    c->has_batch = true;
    c->last_tr_slot = -1;
    c->last_phase_ev = false;
    hip_check(hipEventRecord(c->evStat, s), "hipEventRecord");
  });
}

extern "C" int kg_serve_captures(kg_serve sv, int32_t* out) {
  return guard([&] {
    need(sv && out, "NULL argument");
    *out = sv->captures;
  });
}

extern "C" int kg_serve_destroy(kg_serve sv) {
  return guard([&] {
    if (!sv) return;
    (void)hipSetDevice(sv->c->device);
    if (sv->c->done_valid) (void)hipEventSynchronize(sv->c->evDone);
    serve_release(sv);
    delete sv;
  });
}

extern "C" int kg_ctx_status_dev(kg_ctx c, int32_t* d_out, void* stream) {
  return guard([&] {
    need(c && d_out, "NULL argument");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    StreamScope scope(c, s);
    if (!c->stat_src) {
      // nothing issued yet: all zero
      hip_check(c->wsStatus.ensure(ST_WORDS * sizeof(int)), "hipMalloc(status)");
      hip_check(kgl_fill_int(c->wsStatus.as<int>(), ST_WORDS, 0, s), "clear status");
      c->stat_src = c->wsStatus.as<int>();
    }
    hip_check(kgl_status_copy(c->stat_src, ST_SPILL, ST_UNC, ST_REDO, d_out, s), "status copy");
  });
}
