/* klampt_gpu.h — C ABI of the MI355X collision / distance / feasibility engine.
 *
 * Plain pointers and sizes only.  Every function returns KG_OK (0) or a
 * negative KG_E* code; the message of the last error on the calling thread
 * is available from kg_last_error().  Handles are not thread-safe.  Host
 * arrays passed in are copied; batch outputs are written to caller memory
 * and host-pointer calls return when they are complete (they run on a
 * private stream of the context).  Functions suffixed _dev take device
 * pointers and a hipStream_t (as void*; NULL is the default stream) and run
 * asynchronously on it: their inputs must be ready in stream order, and
 * their outputs are ready in that stream's order.  All calls of one context
 * share its workspaces and execute in call order, whatever streams they
 * name: a call on a different stream waits for the previous call's work.
 *
 * Transforms: R is a column-major 3x3 rotation (9 doubles) and t a 3-vector,
 * as in Klampt (Python/klampt/math/so3.py:1-16).  A link transform array is
 * 12 values per link: R (9, column-major) then t (3).
 *
 * Reference interfaces replaced are cited per entry point (paths relative to
 * the Klampt source tree).
 */
#ifndef KLAMPT_GPU_H
#define KLAMPT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KG_OK 0
#define KG_E_INVALID -1     /* bad argument (mirrors PyException on bad input) */
#define KG_E_HIP -2         /* HIP runtime error */
#define KG_E_UNSUPPORTED -3 /* unsupported geometry pair (Klampt: Inf distance -> raise) */
#define KG_E_EMPTY -4       /* empty geometry where a value is required */
#define KG_E_LIMIT -5       /* a documented size limit was exceeded */

typedef struct kg_ctx_s* kg_ctx;
typedef struct kg_geom_s* kg_geom;
typedef struct kg_robot_s* kg_robot;
typedef struct kg_world_s* kg_world;
typedef struct kg_serve_s* kg_serve;

/* DistanceQuerySettings (Python/klampt/src/geometry.h:631-638) */
typedef struct {
  double relErr, absErr, upperBound;
} kg_dist_settings;

/* DistanceQueryResult (Python/klampt/src/geometry.h:663-673) */
typedef struct {
  double d;
  int hasClosestPoints, hasGradients;
  double cp1[3], cp2[3];
  double grad1[3], grad2[3];
  int elem1, elem2;
} kg_dist_result;

/* Last error message of the calling thread (the text SWIG's %exception
 * raises, Python/klampt/src/robotsim.i:41-53). */
const char* kg_last_error(void);
int kg_version(void);

/* ---- context ----------------------------------------------------------- */
int kg_ctx_create(int device, kg_ctx* out);
int kg_ctx_destroy(kg_ctx ctx);
int kg_ctx_synchronize(kg_ctx ctx);

/* Engine parameters of a context: the traversal schedule, its capacities and
 * a test hook, with their defaults in parentheses.  kg_ctx_get_params reads
 * the current values (the defaults after kg_ctx_create), kg_ctx_set_params
 * checks and replaces all of them (KG_E_INVALID if one is out of range, and
 * then nothing changes), kg_autotune writes the fields it tunes.  No Klampt
 * counterpart: Klampt's collision checks have no schedule to set. */
typedef struct {
  int32_t cpw;           /* configurations per chunk a traversal wave takes, 1..64 (16) */
  int32_t leaf_min;      /* queued leaf pairs that start a triangle-test step, 1..64 (1) */
  int32_t scap;          /* LDS task-pool entries per traversal wave: 256 or 512 (256) */
  int32_t waves_per_cu;  /* traversal waves per CU, 0 = the occupancy limit (0) */
  int32_t static_pct;    /* percent of a batch dealt to the waves up front, 0..100 (60) */
  int32_t nsub;          /* ranges of the rest of the batch, one counter each, 1..32 (8) */
  int32_t tail_lanes;    /* a wave with fewer working lanes runs the tail cadence, 0..64 (16) */
  int32_t tail_reps;     /* box-test iterations per scheduling step in the tail, 1..8 (4) */
  int32_t spill_cap;     /* HBM spill entries per traversal wave, 0..2^24 (16384); when a
                            wave's spill region is full its oldest pooled tasks are dropped
                            and what depends on them is reported as uncertain (below) */
  int32_t dist_ipw;      /* distance items a traversal wave takes per fetch, 1..64 (16) */
  int32_t edge_w0;       /* bisection positions in the first edge-check round, 1..2^20 (2);
                            round r is at most w0 edge_growth^r wide */
  int32_t sdf_lanes;     /* lanes per waypoint of the SDF penalty kernels: 8, 16, 32, 64 (32) */
  int32_t prep_group;    /* 1: lane-per-link broad phase where it applies (groups of 8, 16
                            or 32 lanes, the fewest that hold the robot's links), 8 / 16 / 32:
                            groups of at least that many lanes, 0: one lane per
                            configuration (1) */
  int32_t force_recheck; /* test hook: every configuration goes through the FP64 pass (0) */
  int32_t edge_budget_mb; /* device workspace of one edge-check round, MiB, 16..2^20 (4096):
                             it bounds the samples checked per round */
  int32_t edge_growth;    /* width factor between edge-check rounds, 2..1024 (16) */
} kg_params;
int kg_ctx_get_params(kg_ctx ctx, kg_params* out);
int kg_ctx_set_params(kg_ctx ctx, const kg_params* params);

/* Status of the last batched call issued on this context (feasibility,
 * distance, distance Jacobian, edge round), copied device-to-device into
 * d_out[0..3] on `stream` in stream order, with no host synchronization (so a
 * _dev caller or a captured graph can test it on the device):
 *   [0] 1 if a traversal wave's spill region filled (tasks were dropped);
 *   [1] results left uncertain by dropped tasks: configurations reported
 *       infeasible (their open flags set, first pair (-2,-2)) without proof,
 *       or pair distances reported as a certified lower bound;
 *   [2] configurations traversed again in FP64;  [3] 0.
 * When [1] is 0 every result of the call is exact.  Host-pointer calls fail
 * with KG_E_LIMIT instead when [1] is not 0.  No Klampt counterpart. */
int kg_ctx_status_dev(kg_ctx ctx, int32_t* d_out, void* stream);

/* ---- geometry ----------------------------------------------------------
 * Triangle mesh: Geometry3D(TriangleMesh) with its collision data
 * (Python/klampt/src/robotsim.cpp:2886-2949 vertices/indices; the BVH built
 * once, as AnyCollisionGeometry3D::InitCollisionData does).  verts: nv x 3
 * doubles, tris: nt x 3 int32. */
int kg_geom_create_trimesh(kg_ctx ctx, const double* verts, int nverts, const int32_t* tris, int ntris,
                           kg_geom* out);
/* The mesh's collision data: its OBB tree (built once at creation, as
 * AnyCollisionGeometry3D::InitCollisionData builds PQP's), nnodes x 16 floats
 * per node: centre (3), half extents (3), axes (9, column-major), then the
 * child as int32 bits -- >= 0: the first of the two adjacent children (node
 * index), < 0: a leaf holding triangle ~child (the mesh's own numbering).
 * Node 0 is the root.  nodes may be NULL to ask for the count only.  Host
 * memory only (no GPU work).  Klampt does not expose PQP's tree; this is for
 * measurement and tests. */
int kg_geom_get_bvh(kg_geom g, float* nodes, int max_nodes, int* nnodes);
/* ImplicitSurface (Python/klampt/src/geometry.h:322-391): cell-centred values,
 * C-order (i,j,k), dims[0] x dims[1] x dims[2], > 0 outside. */
int kg_geom_create_sdf(kg_ctx ctx, const double bmin[3], const double bmax[3], const int32_t dims[3],
                       const float* values, kg_geom* out);
/* Releases the handle.  Like Klampt's shared_ptr geometries, a geometry still
 * held by a robot link, terrain or rigid object stays alive until the last of
 * them lets it go (kg_robot_set_link_geom with another geometry,
 * kg_world_set_*_geom, kg_robot_destroy, kg_world_destroy); the handle itself
 * must not be used again. */
int kg_geom_destroy(kg_geom g);
/* PointCloud (Python/klampt/src/geometry.h PointCloud; the points only): pts
 * npts x 3 doubles in the geometry's local frame.  Point clouds take part in
 * the single-pair queries against ImplicitSurface geometries
 * (Cpp/docs/Manual-Geometry.md:134,293-294,341-342): collide = some point's
 * surface value <= the summed margins, withinDistance = <= tol + margins,
 * distance = the smallest value minus the margins, with that point, the
 * surface point along the gradient, and the gradient.  A TriangleMesh against
 * an ImplicitSurface collides (withinDistance) by its vertices the same way;
 * its distance is not available (as in the reference's table). */
int kg_geom_create_pointcloud(kg_ctx ctx, const double* pts, int npts, kg_geom* out);
/* TriangleMesh -> ImplicitSurface: Geometry3D.convert("ImplicitSurface",
 * resolution) (Python/klampt/src/geometry.h:944-1005, robotsim.cpp:1602-1646).
 * Cell-centred grid of spacing `resolution` (0: the average longest triangle
 * edge) covering the mesh's vertex bounds plus `pad` cells on every side:
 * dims[k] = ceil((hi[k]-lo[k]) / h) + 2 pad (at least 1), bmin = lo - pad h.
 * Value at a cell centre: the Euclidean distance to the closest triangle,
 * negated where the mesh's winding number around the centre (counted along
 * +z) is nonzero.  Klampt computes this grid by fast marching from the
 * surface; here every cell holds the distance that method approximates.  The
 * new geometry starts with the mesh's current transform. */
int kg_geom_mesh_to_sdf(kg_geom mesh, double resolution, int pad, kg_geom* out);
/* ImplicitSurface grid: dims, bounds, and (when values != NULL) the
 * dims[0]*dims[1]*dims[2] values, C order, copied to host memory
 * (Geometry3D.getImplicitSurface, geometry.h:322-391). */
int kg_geom_get_sdf(kg_geom sdf, int32_t dims[3], double bmin[3], double bmax[3], float* values);
/* A second handle on the same geometry data, as ManagedGeometry::Load hands a
 * copy of an already loaded file's geometry to every later loader of that
 * file (Cpp/Modeling/ManagedGeometry.cpp:111-164): a mesh shares its BVH and
 * device triangles with src (nothing is rebuilt or re-uploaded); an SDF gets
 * its own device copy of the grid.  The new handle starts with src's current
 * transform and collision margin and changes them independently; destroying
 * either handle leaves the other valid. */
int kg_geom_clone(kg_geom src, kg_geom* out);
/* Geometry3D.setCurrentTransform / getCurrentTransform
 * (Python/klampt/src/robotsim.cpp:1465-1473) */
int kg_geom_set_transform(kg_geom g, const double R[9], const double t[3]);
int kg_geom_get_transform(kg_geom g, double R[9], double t[3]);
/* Geometry3D.setCollisionMargin / getCollisionMargin
 * (Python/klampt/src/robotsim.cpp:1561-1573): a nonnegative margin that
 * fattens the geometry for collision testing (Cpp/docs/Manual-Geometry.md,
 * "Collision geometries").  kg_geom_collide then tests distance <= the summed
 * margins, kg_geom_within_distance adds them to the tolerance and
 * kg_geom_distance subtracts them (negative when the fattened geometries
 * overlap; defined by this build).  Batched feasibility calls reject worlds
 * with margins (KG_E_UNSUPPORTED). */
int kg_geom_set_collision_margin(kg_geom g, double margin);
int kg_geom_get_collision_margin(kg_geom g, double* margin);
/* Geometry3D.getBB (Python/klampt/src/robotsim.cpp:1575-1586): world AABB of
 * the transformed local bounding box. */
int kg_geom_get_bb(kg_geom g, double bmin[3], double bmax[3]);
/* Geometry3D.collides (Python/klampt/src/robotsim.cpp:1656-1662); empty
 * geometry -> 0 (no error). */
int kg_geom_collide(kg_geom a, kg_geom b, int* out);
/* Geometry3D.withinDistance (Python/klampt/src/robotsim.cpp:1685-1692) */
int kg_geom_within_distance(kg_geom a, kg_geom b, double tol, int* out);
/* Geometry3D.distance_ext (Python/klampt/src/robotsim.cpp:1777-1819):
 * unsigned TriangleMesh/TriangleMesh distance, 0 when touching.  settings
 * (nullable) as DistanceQuerySettings: upperBound caps the search (a result
 * above it is not searched for), relErr / absErr let the search stop early
 * once a subtree cannot improve the result by more than either tolerance:
 * the result d then satisfies D <= d <= max(D (1 + relErr), D + absErr) for
 * the exact distance D (both 0: exact). */
int kg_geom_distance(kg_geom a, kg_geom b, const kg_dist_settings* settings, kg_dist_result* out);
/* Geometry3D.distance_point_ext on an ImplicitSurface
 * (Python/klampt/src/robotsim.cpp:1740-1775, geometry.h:1052-1063): signed
 * trilinear value and gradient. */
int kg_geom_distance_point(kg_geom g, const double pt[3], const kg_dist_settings* settings, kg_dist_result* out);
/* Batched ImplicitSurface point queries (world points, FP32): d[n], grad[n][3]
 * (grad nullable). */
int kg_sdf_eval(kg_geom sdf, const float* pts, int n, float* d, float* grad);

/* ---- robot -------------------------------------------------------------
 * RobotModel kinematic tree (Cpp/Modeling/Robot.h:83-163; links set up as in
 * Cpp/Modeling/Robot.cpp:957-981).  parents[n] (topologically sorted, -1 =
 * world), T0[n][12] (T0_Parent), axes[n][3], jtype[n] (0 revolute, 1
 * prismatic), qmin/qmax[n].  Self-collision pairs default to all pairs except
 * link vs parent (Cpp/docs/Manual-FileTypes.md:212). */
int kg_robot_create(kg_ctx ctx, int nlinks, const int32_t* parents, const double* T0, const double* axes,
                    const int32_t* jtype, const double* qmin, const double* qmax, kg_robot* out);
int kg_robot_destroy(kg_robot r);
/* RobotModelLink.geometry() assignment; g may be NULL (empty geometry). */
int kg_robot_set_link_geom(kg_robot r, int link, kg_geom g);
/* RobotModel.enableSelfCollision (Python/klampt/src/robotsim.cpp:5504-5521) */
int kg_robot_enable_self_collision(kg_robot r, int link1, int link2, int enabled);
int kg_robot_self_collision_enabled(kg_robot r, int link1, int link2, int* out);
/* Replaces the self-collision pair set by the npairs pairs[2k], pairs[2k+1]
 * (the .rob "selfcollision" list, Cpp/docs/Manual-FileTypes.md:212). */
int kg_robot_set_self_pairs(kg_robot r, const int32_t* pairs, int npairs);
/* RobotModel.setConfig (Python/klampt/src/robotsim.cpp:5311-5321):
 * UpdateConfig (FK) + UpdateGeometry. */
int kg_robot_set_config(kg_robot r, const double* q);
int kg_robot_get_config(kg_robot r, double* q);
/* RobotModelLink.getTransform after setConfig. */
int kg_robot_get_link_transform(kg_robot r, int link, double R[9], double t[3]);
/* RobotModel.selfCollides (Python/klampt/src/robotsim.cpp:5529-5539). */
int kg_robot_self_collides(kg_robot r, int* out);
/* Batched RobotModel::UpdateConfig in FP32: Q[N][n] -> T[N][n][12]. */
int kg_fk_batch(kg_robot r, const float* Q, int N, float* T_out);
int kg_fk_batch_dev(kg_robot r, const float* dQ, int N, float* dT_out, void* stream);
/* The same in FP64 (the arithmetic of the feasibility broad phase). */
int kg_fk_batch64(kg_robot r, const double* Q, int N, double* T_out);
int kg_fk_batch64_dev(kg_robot r, const double* dQ, int N, double* dT_out, void* stream);

/* ---- world -------------------------------------------------------------
 * WorldModel ID layout (Cpp/Modeling/World.cpp:47-53,156-180): terrains, then
 * rigid objects, then per robot [robot, link0..]; the collision mask is
 * WorldPlannerSettings::InitializeDefault (Cpp/Planning/PlannerSettings.cpp:16-41),
 * rebuilt (and user overrides cleared) whenever an item is added. */
int kg_world_create(kg_ctx ctx, kg_world* out);
int kg_world_destroy(kg_world w);
int kg_world_add_terrain(kg_world w, kg_geom g, int* id_out);
int kg_world_add_rigid_object(kg_world w, kg_geom g, const double R[9], const double t[3], int* id_out);
int kg_world_set_rigid_object_transform(kg_world w, int object_index, const double R[9], const double t[3]);
/* Replace a terrain's or rigid object's geometry in place (world ID and
 * placement kept), as loading a file into the element's Geometry3D replaces
 * the world's ManagedGeometry (Python/klampt/src/robotsim.cpp:1443-1453). */
int kg_world_set_terrain_geom(kg_world w, int terrain_index, kg_geom g);
int kg_world_set_rigid_object_geom(kg_world w, int object_index, kg_geom g);
int kg_world_add_robot(kg_world w, kg_robot r, int* id_out);
int kg_world_num_ids(kg_world w);
/* WorldPlannerSettings::collisionEnabled(id1,id2) */
int kg_world_set_collision_enabled(kg_world w, int id1, int id2, int enabled);
int kg_world_get_collision_enabled(kg_world w, int id1, int id2, int* out);
/* Pair list used by kg_distance_batch for robot `robot_index`: mask-enabled
 * (link, env item) pairs in link-major order, then enabled self pairs (i<j);
 * world IDs, 2 per pair. */
int kg_world_pair_list(kg_world w, int robot_index, int32_t* pairs, int max_pairs, int* npairs);

/* SingleRobotCSpace::IsFeasible over a batch
 * (Cpp/Planning/RobotCSpace.cpp:786-823; broad phase and order of
 * WorldPlannerSettings::CheckCollision, Cpp/Planning/PlannerSettings.cpp:214-331).
 * Q[N][n] float.  feasible[N] required; env_hit, self_hit, first_pair (N x 2
 * world IDs of the first colliding pair in Klampt's loop order, env before
 * self, -1 if none) are optional (NULL).  Asking only for `feasible` enables
 * whole-configuration early exit. */
int kg_feasible_batch(kg_world w, int robot_index, const float* Q, int N, uint8_t* feasible, uint8_t* env_hit,
                      uint8_t* self_hit, int32_t* first_pair);
int kg_feasible_batch_dev(kg_world w, int robot_index, const float* dQ, int N, uint8_t* d_feasible,
                          uint8_t* d_env_hit, uint8_t* d_self_hit, int32_t* d_first_pair, void* stream);
/* The same with double-precision configurations, as Klampt's Config is
 * (SingleRobotCSpace::IsFeasible(const Config&), RobotCSpace.cpp:786): joint
 * limits are tested on the doubles as given (a value on a limit is inside),
 * and the link frames are formed from them in FP64. */
int kg_feasible_batch_f64(kg_world w, int robot_index, const double* Q, int N, uint8_t* feasible, uint8_t* env_hit,
                          uint8_t* self_hit, int32_t* first_pair);
int kg_feasible_batch_f64_dev(kg_world w, int robot_index, const double* dQ, int N, uint8_t* d_feasible,
                              uint8_t* d_env_hit, uint8_t* d_self_hit, int32_t* d_first_pair, void* stream);

/* kg_feasible_batch_dev that allocates nothing.  No counterpart in Klampt:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * workspace of the batch comes from a fixed scratch array of 64 MiB that the
 * library holds on each GPU it runs on until the process ends (16 MiB of it
 * is the traversal's spill region; overflow is reported as for the other
 * calls, kg_ctx_status_dev).  Limits: robots of at most 32 links, and N at
 * most what fits -- about 29,000 configurations of a 7-link arm against one
 * terrain (22 pair items each); KG_E_LIMIT names the bound for the world.
 * The world's tables must already be on the device: KG_E_INVALID if the
 * world changed since its last batched call.  All calls on one GPU share
 * the scratch array and it keeps the last call's contents: calls that
 * overlap on one GPU (two contexts, or a stream not ordered after the
 * previous call) race. */
int kg_feasible_batch_dev_noalloc(kg_world w, int robot_index, const float* dQ, int N, uint8_t* d_feasible,
                                  uint8_t* d_env_hit, uint8_t* d_self_hit, int32_t* d_first_pair, void* stream);

/* Pair distances with closest points and joint-space gradients: the
 * point-collision step of KineTrajOpt
 * (Python/klampt/plan/kinetrajopt/kinetrajopt.py:530-556: distance_ext per
 * link / obstacle pair with upperBound = dcheck, then the normal and
 * robot.positionJacobian at the link's closest point), batched over
 * configurations and the kg_world_pair_list pairs.  Per configuration c and
 * pair p (row-major):
 *   d_out[c][p]        distance, +inf when no point pair is within dcheck;
 *   cp_out[c][p][6]    closest points in world coordinates: on the pair's
 *                      robot link, then on the other side (0 when d = inf;
 *                      when d = 0, twice a point common to the two closest
 *                      triangles);
 *   grad_out[c][p][n]  dd/dq = nhat . (J_a(cp_a) - J_b(cp_b)) with
 *                      nhat = (cp_a - cp_b) / |cp_a - cp_b|; J_b = 0 for
 *                      environment geometry; 0 when d = 0 or d = inf. */
int kg_distance_jacobian_batch(kg_world w, int robot_index, const float* Q, int N, float dcheck, float* d_out,
                               float* cp_out, float* grad_out);
int kg_distance_jacobian_batch_dev(kg_world w, int robot_index, const float* dQ, int N, float dcheck, float* d_out,
                                   float* cp_out, float* grad_out, void* stream);
/* KineTrajOpt sweep terms (Python/klampt/plan/kinetrajopt/kinetrajopt.py:
 * 530-634, find_collision_pair with compute_sweep_collision): for each segment
 * k (configurations Q0[k] -> Q1[k], FP64 as Klampt's Config), link l and
 * obstacle o, the distance between the convex hull of link l's geometry
 * placed at both configurations (setConvexHullGroup of the two copies of
 * convert('ConvexHull'), :603-611) and the convex hull of the obstacle, by GJK
 * in FP64, and the add_sweep terms (:556-583).  Obstacles are the world's
 * terrains then its rigid objects, in world ID order, as convex hulls of their
 * points at their placements (kinetrajopt.py:495-501 takes the terrains; the
 * rigid objects follow them here); kg_world_num_obstacles counts them.  Per
 * query q = (k * nlinks + l) * nobs + o:
 *   d_out[q]          distance; +inf when not below dcheck (DistanceQuerySettings
 *                     upperBound = dcheck, kept only if d < dcheck) or when
 *                     either side has no geometry; 0 when the hulls overlap
 *                     (then no terms: Klampt's penetration depth is not formed);
 *   info_out[q][16]   alpha, normal (from the obstacle to the swept hull),
 *                     pswept and pobs (the closest points), p0 and p1 (the
 *                     support points of the hull at Q0[k] and at Q1[k] in
 *                     -normal); zeros without terms;
 *   coef_out[q][2][n] normal . J_l(Q0[k], p0) and normal . J_l(Q1[k], p1) (the
 *                     lin_coef1 / lin_coef2 of the sweep constraint, :454-461,
 *                     whose weights are alpha and 1 - alpha).
 * With Q0[k] == Q1[k] the swept hull is the link's hull at that
 * configuration: the point terms of the hull geometries. */
int kg_world_num_obstacles(kg_world w, int* out);
int kg_sweep_batch(kg_world w, int robot_index, const double* Q0, const double* Q1, int N, double dcheck,
                   double* d_out, double* info_out, double* coef_out);
int kg_sweep_batch_dev(kg_world w, int robot_index, const double* dQ0, const double* dQ1, int N, double dcheck,
                       double* d_out, double* info_out, double* coef_out, void* stream);
/* Per-pair minimum distance for the kg_world_pair_list pairs (AnyCollisionQuery
 * ::Distance, Cpp/Planning/PlannerSettings.cpp:109-115): out[N][npairs], 0 when
 * touching. */
int kg_distance_batch(kg_world w, int robot_index, const float* Q, int N, float* pair_dist);
int kg_distance_batch_dev(kg_world w, int robot_index, const float* dQ, int N, float* d_pair_dist, void* stream);

/* Straight-line edge checks (SingleRobotCSpace::PathChecker ->
 * EpsilonEdgeChecker, Cpp/Planning/RobotCSpace.cpp:835-837; sample rule
 * n = ceil(d/eps), u = i/n of Python/klampt/plan/robotcspace.py:216-228;
 * interpolation Cpp/Modeling/Interpolate.cpp:13-14; d the L2 norm
 * (norm_inf = 0, Klampt's default) or the L-infinity norm).  Interior samples
 * i = 1..n-1 are checked in level-order bisection; first_u[s] is u of the
 * first infeasible sample in that order, -1 if all are feasible.  A[nseg][n]
 * and B[nseg][n] are the segment endpoints (FP32, or FP64 for the _f64
 * variants, which interpolate Klampt's double Config exactly).
 * The checks run as rounds on the device with no host synchronization
 * between them (kg_params.edge_w0, edge_budget_mb).  Host-pointer calls
 * cover every segment.  The _dev calls cover segments of at most max_steps
 * steps n (0: the steps of the joint-range diagonal, which bounds every
 * segment between configurations within the joint limits; KG_E_INVALID if a
 * joint range is unbounded); a longer segment's first_u is -2 and it is
 * counted in kg_ctx_status_dev [1].  d_nsamples (nullable) receives n per
 * segment.  The _dev calls of a robot of at most 32 links run on a ring of
 * two workspace slots per context, as kg_feasible_batch_dev does on its
 * three: a call waits for the previous call on its slot and for the last
 * plain call, so calls issued on different streams run concurrently (two
 * batches in flight fill the GPU the rounds of one leave idle).  They
 * allocate nothing once two calls of the same robot with as many segments
 * have run (one per slot), so they can be captured in a HIP graph; a
 * captured call holds one slot, and its replays race with _dev edge calls
 * issued on other streams that the caller does not order after them. */
int kg_edge_batch(kg_world w, int robot_index, const float* A, const float* B, int nseg, double eps, int norm_inf,
                  float* first_u);
int kg_edge_batch_f64(kg_world w, int robot_index, const double* A, const double* B, int nseg, double eps,
                      int norm_inf, float* first_u);
int kg_edge_batch_dev(kg_world w, int robot_index, const float* dA, const float* dB, int nseg, double eps,
                      int norm_inf, int max_steps, float* d_first_u, int32_t* d_nsamples, void* stream);
int kg_edge_batch_f64_dev(kg_world w, int robot_index, const double* dA, const double* dB, int nseg, double eps,
                          int norm_inf, int max_steps, float* d_first_u, int32_t* d_nsamples, void* stream);
/* The k-th interior sample index (0 <= k < nsteps-1) in the level-order
 * bisection order of a segment of nsteps steps; -1 if out of range.  Host
 * only (no GPU needed). */
int kg_edge_sample_index(int nsteps, int k);

/* SDF hinge penalty L(q) = sum_{links k} sum_{p in pts_k} max(0, dsafe - sdf(T_k(q) p))^2
 * and its gradient by an FK-Jacobian adjoint (the distance-gradient rule of
 * Python/klampt/math/autodiff/geometry_ad.py:250-293 chained with
 * kinematics_ad.py:53-59).  local_pts: concatenated per-link points
 * (pts_per_link[n]), link frames.  bwd recomputes FK and the SDF gathers. */
int kg_sdf_penalty_fwd(kg_robot r, kg_geom sdf, const float* Q, int N, const float* local_pts,
                       const int32_t* pts_per_link, float dsafe, float* loss);
int kg_sdf_penalty_bwd(kg_robot r, kg_geom sdf, const float* Q, int N, const float* local_pts,
                       const int32_t* pts_per_link, float dsafe, const float* dloss, float* dQ);
int kg_sdf_penalty_fwd_dev(kg_robot r, kg_geom sdf, const float* dQ, int N, const float* d_local_pts,
                           const int32_t* pts_per_link, float dsafe, float* d_loss, void* stream);
int kg_sdf_penalty_bwd_dev(kg_robot r, kg_geom sdf, const float* dQ, int N, const float* d_local_pts,
                           const int32_t* pts_per_link, float dsafe, const float* d_dloss, float* d_dQ,
                           void* stream);

/* Heterogeneous batch: robot robot_indices[b] of worlds[b] at counts[b]
 * configurations each (b < nbatch), checked as IsFeasible in one broad-phase
 * launch and one persistent traversal launch.  Q is the concatenation of the
 * per-batch configuration arrays (counts[b] x n_b floats), feasible the
 * concatenation of the outputs.  No Klampt counterpart: it batches many
 * SingleRobotCSpace::IsFeasible calls (Cpp/Planning/RobotCSpace.cpp:786). */
int kg_feasible_multi(int nbatch, const kg_world* worlds, const int32_t* robot_indices, const int32_t* counts,
                      const float* Q, uint8_t* feasible);
int kg_feasible_multi_dev(int nbatch, const kg_world* worlds, const int32_t* robot_indices, const int32_t* counts,
                          const float* dQ, uint8_t* d_feasible, void* stream);
/* Serving pipeline: a heterogeneous feasibility batch (as kg_feasible_multi_dev:
 * robot robot_indices[b] of worlds[b] at counts[b] configurations, dQ /
 * d_feasible the concatenations) and nedge edge batches (as kg_edge_batch_dev:
 * edge_counts[g] segments of robot edge_robots[g] of edge_worlds[g], dA / dB /
 * d_first_u the concatenations, one eps, norm and max_steps for all), captured
 * by the library into one HIP graph.  No Klampt counterpart: it batches the
 * SingleRobotCSpace::IsFeasible and PathChecker(a,b)->IsVisible() calls of a
 * planner iteration (Cpp/Planning/RobotCSpace.cpp:786,835-837) into one graph
 * launch.  The buffers are the caller's and stay fixed: write new inputs into
 * them, then kg_serve_run replays the graph on `stream` (ordered after the
 * context's earlier calls, as any _dev call).  The pipeline is issued once at
 * creation to size the context's workspaces, then captured.  kg_serve_run
 * recaptures first when the graph is stale: a world or geometry of the
 * pipeline changed, kg_params changed (kg_autotune, kg_ctx_set_params), or a
 * context workspace was reallocated by a larger call of another shape.
 * kg_ctx_status_dev after a run reports the whole pipeline (spill flags ORed,
 * uncertain / undecided and re-check counts summed).  A feasibility batch
 * whose pair items exceed the fixed-capacity layout (2 GiB) is rejected
 * (KG_E_INVALID); the pipelines and calls of one context share its
 * workspaces, so they run one at a time in stream order.  The worlds, their
 * robots and geometries and the buffers must outlive the pipeline
 * (kg_serve_destroy first).  kg_serve_captures reports how many captures the
 * pipeline has taken. */
int kg_serve_create(int nbatch, const kg_world* worlds, const int32_t* robot_indices, const int32_t* counts,
                    const float* dQ, uint8_t* d_feasible, int nedge, const kg_world* edge_worlds,
                    const int32_t* edge_robots, const int32_t* edge_counts, const float* dA, const float* dB,
                    double eps, int norm_inf, int max_steps, float* d_first_u, kg_serve* out);
int kg_serve_run(kg_serve serve, void* stream);
int kg_serve_captures(kg_serve serve, int32_t* out);
int kg_serve_destroy(kg_serve serve);
/* Autotuner: times the traversal settings (configurations per wave chunk,
 * leaf batch threshold, LDS pool capacity, waves per CU) on the given batch
 * (device Q; the device is synchronized first, so dQ may come from any
 * stream) and keeps the fastest on the context.  chosen[4] (nullable)
 * receives {cpw, leaf_min, scap, waves_per_cu (0 = occupancy limit)}. */
int kg_autotune(int nbatch, const kg_world* worlds, const int32_t* robot_indices, const int32_t* counts,
                const float* dQ, int32_t* chosen, float* best_ms);
/* Sets the four tuned fields of the context's kg_params; 0 restores a field's
 * default (scap: 0, 256 or 512). */
int kg_ctx_set_tuning(kg_ctx ctx, int cpw, int leaf_min, int scap, int waves_per_cu);

/* Counters of the last batched call on this world's context (waits for it):
 * [0] configurations traversed again in FP64, [1] narrow-phase pair items, [2] 1
 * if a spill region filled, [3] broad-phase time (ns), [4] FP32 traversal
 * kernel time (ns), [5] FP64 re-check kernel time (ns) -- HIP events on the
 * launch stream; [3] and [5] need timing level 2, [4] level 1 or 2
 * (kg_ctx_set_timing), else 0 -- [6] box tests, [7] triangle tests of the
 * FP32 traversal, [8] uncertain results (kg_ctx_status_dev [1]), [9]
 * triangle pairs the FP32 test left undecided, each decided by the FP64
 * pass's exact test alone (any-hit and group modes; in the ordered
 * first-pair mode such a configuration is traversed again, [0]). */
int kg_world_last_stats(kg_world w, int64_t* stats, int nstats);
/* Sum of the FP32 traversal kernel durations (HIP events on the launch
 * stream) of the feasibility batches issued on this context since the last
 * reset, at most the 64 most recent; count receives how many were summed.
 * Waits for those batches.  No Klampt counterpart (measurement). */
int kg_ctx_kernel_times(kg_ctx ctx, int reset, int64_t* total_ns, int32_t* count);
/* Which HIP events the batched calls of this context record on their stream
 * (each costs a few microseconds of stream time): 0 none, 1 (default) the
 * FP32 traversal kernel's begin/end (kg_ctx_kernel_times, stats [4]), 2 also
 * the broad phase and the FP64 pass (stats [3], [5]).  No Klampt
 * counterpart (measurement). */
int kg_ctx_set_timing(kg_ctx ctx, int level);

#ifdef __cplusplus
}
#endif
#endif /* KLAMPT_GPU_H */
