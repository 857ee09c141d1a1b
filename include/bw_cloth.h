/*
 * bw_cloth.h -- C ABI of the MI355X-native Baraff-Witkin implicit cloth step.
 *
 * Drop-in boundary for the reference's plain cloth method (cloth_type
 * "fem", class cBaraffCloth).  The reference binds its cloths through the
 * C++ virtual interface cBaseCloth (src/sim/cloth/BaseCloth.h:19-80) and the
 * factory BuildCloth (src/sim/cloth/ClothBuilder.cpp:17-54); each entry point
 * below names the member it replaces.  Plain pointers and sizes only; every
 * call returns a status (BW_OK = 0) instead of exiting the process, and
 * bw_last_error() describes the last failure on the calling thread.
 *
 * Configuration is the reference's cloth JSON (keys read by cBaraffCloth::Init,
 * BaraffCloth.cpp:22-83, and cBaseCloth::Init, BaseCloth.cpp:44-55), plus
 * optional keys of this library:
 *   "precision":        "mixed" (default: FP32 matrix and PCG recurrences,
 *                        FP64 dot products, FP64 solution, FP64 true-residual
 *                        replacement and convergence checks) or "fp64"
 *                        (everything FP64)
 *   "solver_tolerance": relative residual ||b - W dx|| / ||b|| (default 1e-6)
 *   "solver_max_iters": PCG iteration cap per solve (default 20000)
 *   "solver_mode":      "tolerance" (default) or "reference" (exactly
 *                        max_iters-1 FP64 PCG iterations, BaraffCloth.cpp:529)
 *   "delaunay_seed":    seed of geometry_type "delaunay" (the reference draws
 *                        its points unseeded, TriangulatorDelaunay.cpp:66-69)
 *   "delaunay_min_spacing": minimum distance between random points of
 *                        geometry_type "delaunay" (default: the reference's
 *                        sqrt(w h / (n pi)), TriangulatorDelaunay.cpp:60-61)
 *   "pin_top_edge":     true pins every vertex of the cloth's top side
 * Positions are 3*V doubles, xyz interleaved, in the reference's vertex order.
 */
#ifndef BW_CLOTH_H
#define BW_CLOTH_H

#ifdef __cplusplus
extern "C" {
#endif

#define BW_OK 0
#define BW_ERR_ARG -1      /* bad argument or configuration */
#define BW_ERR_DEVICE -2   /* HIP runtime error */
#define BW_ERR_NUMERIC -3  /* NaN/Inf or solver breakdown */
#define BW_ERR_STATE -4    /* call not valid in this state */
#define BW_ERR_CAPACITY -5 /* over a size or memory budget */

typedef struct bw_cloth bw_cloth;
typedef struct bw_batch bw_batch;
typedef struct bw_body bw_body;

typedef struct bw_step_stats
{
    int pcg_iterations;   /* PCG iterations of the last step */
    int outer_iterations; /* FP64 true-residual checks of the last step */
    double rel_residual;  /* final FP64 true residual / ||b|| on free dofs */
    double step_us;       /* host wall time of the last step */
    int matrix_fp64;      /* "mixed": 1 when the last solve stalled on the FP32
                             matrix and finished on the FP64 one */
    int reserved;
} bw_step_stats;

/* ---- single cloth: the cBaseCloth interface --------------------------- */

/* cBaraffCloth::Init(const Json::Value&) (BaraffCloth.cpp:22-83) via the
   factory key "fem" (ClothBuilder.cpp:17-54).  x0 (3V, may be NULL) is then
   applied as SetPos(x0) + SetPrePos(x0) (BaseCloth.cpp:97-109). */
int bw_create(const char *json_config, const double *x0, bw_cloth **out);
/* A cloth on a given triangle mesh: rest positions (3V doubles), triangles
   (3T vertex ids, the winding the reference's mesh builders produce) and the
   cloth JSON with "geometry_type": "mesh" (cloth_size / subdivision may be
   omitted).  Edges are numbered at their first appearance in the triangle
   list and UV is the rest xy, as cTriangulator::DelaunayTriangulation
   builds them (TriangulatorDelaunay.cpp:9-160); this replaces the
   reference's mesh ingest for this path (enable_cloth_from_obj,
   BaseCloth.cpp:301-357, reads the mesh from an OBJ file instead).  x0 may
   be NULL. */
int bw_create_mesh(const char *json_config, const double *rest_pos, int nv, const int *tri,
                   int nt, const double *x0, bw_cloth **out);
/* The mesh of "geometry_type": "delaunay" for cloth_size {height, width},
   n = subdivision[0] * subdivision[1] points, delaunay_seed and
   delaunay_min_spacing (0: the reference's spacing rule), before the
   init transform, in the winding the cloth uses.  This is
   cTriangulator::DelaunayTriangulation (TriangulatorDelaunay.cpp:45-160)
   with a seeded point stream; it has n points, or more when the rectangle's
   boundary points alone exceed n.  *nv and *nt receive the counts;
   rest_pos (3 nv doubles, z = 0, room for cap_vertices vertices) and tri
   (3 nt ids, room for cap_triangles) are filled when not NULL, and
   BW_ERR_CAPACITY is returned when they are too small.  Host only: needs
   no device. */
int bw_delaunay_mesh(double height, double width, int n, unsigned long long seed,
                     double min_spacing, double *rest_pos, int cap_vertices, int *tri,
                     int cap_triangles, int *nv, int *nt);
/* destructor of the object (shared_ptr release in the reference) */
void bw_destroy(bw_cloth *c);

/* cBaraffCloth::UpdatePos(double dt) (BaraffCloth.cpp:90-163); like the
   reference, dt is ignored and "default_timestep" is used. */
int bw_update_pos(bw_cloth *c, double dt);
/* cBaseObject::Reset -> cBaseCloth::Reset (BaseCloth.cpp:57-64): initial
   positions, and the pending user force cleared (ClearForce, :66-73) */
int bw_reset(bw_cloth *c);
/* cBaseCloth::GetPos (BaseCloth.cpp:112): copies 3V doubles */
int bw_get_pos(const bw_cloth *c, double *x);
/* mXpre (BaseCloth.h:63) */
int bw_get_pre_pos(const bw_cloth *c, double *x);
/* velocity (mXcur - mXpre) / dt (BaraffCloth.cpp:283-284) */
int bw_get_velocity(const bw_cloth *c, double *v);
/* cBaseCloth::SetPos / SetPrePos (BaseCloth.cpp:97-109) */
int bw_set_pos(bw_cloth *c, const double *x);
int bw_set_pre_pos(bw_cloth *c, const double *x);
/* cBaseCloth::GetInitPos (BaseCloth.cpp:570) */
int bw_get_init_pos(const bw_cloth *c, double *x);
/* cBaseCloth::GetDefaultTimestep (BaseCloth.h:39) */
double bw_get_default_timestep(const bw_cloth *c);
/* cBaseObject::GetNumOfVertices/Triangles/Edges (BaseObject.cpp) */
int bw_num_vertices(const bw_cloth *c);
int bw_num_triangles(const bw_cloth *c);
int bw_num_edges(const bw_cloth *c);
/* Get/Add/ClearConstraintStaticVertices (BaseCloth.cpp:553-568); the get
   returns the count and copies at most cap ids */
int bw_get_constraint_static_vertices(const bw_cloth *c, int *ids, int cap);
int bw_add_constraint_static_vertices(bw_cloth *c, const int *ids, int n);
int bw_clear_constraint_static_vertices(bw_cloth *c);
/* mesh readback: tTriangle ids (T x 3) and tEdge {id0, id1, tri0, tri1,
   boundary} (E x 5) as built by cTriangulator (Triangulator.cpp:141-390) */
int bw_get_triangles(const bw_cloth *c, int *tri3);
int bw_get_edges(const bw_cloth *c, int *edge5);
/* Render-side readback: cBaseCloth::Update's normal refresh after a substep
   (BaseCloth.cpp:585-590: cBaseObject::UpdateTriangleNormal and
   UpdateVertexNormalFromTriangleNormal, BaseObject.cpp:119-160) at the
   current positions, computed on the device.  vertex_normals (3V doubles):
   normalised sum of the incident triangles' normals weighted by their rest
   areas; triangle_normals (3T doubles): normalised (x1-x0) x (x2-x1).
   Either may be NULL. */
int bw_get_normals(const bw_cloth *c, double *vertex_normals, double *triangle_normals);
/* Contacts for the next substep only, which clears them, as the scene's
   collision detection hands them over before a step
   (cSimScene::PerformCollisionDetection, SimScene.cpp:174-190):
   cBaseObject::SetPointTriangleCollisionInfo (BaseObject.cpp:199-203) with
   tPointTriangleCollisionInfo records (sim/collision/CollisionInfo.h:45-61).
   n contacts; ids: n x {cloth_is_point (obj0_is_cloth_obj1_is_rb: the point
   is a cloth vertex and the triangle a body's), vertex (mVertexId0),
   triangle (mTriangleId1)}; normal: n x 3 (mOuterNormal); bary: n x 3
   (mBary); depth: n (mDepth).  The step adds the collision Hessian to W
   (cBaraffCloth::PrepareCollisionHessian, BaraffCloth.cpp:928-990, :293)
   and the repulsion impulses to the solution (cBaraffCloth::Repulsion,
   :626-926, :147), then forgets the contacts (:156-157).  The cloth-side
   index (the vertex when cloth_is_point, else the triangle) must be in
   range and the depth below 3e-3 (the reference asserts 3e-3 - depth > 0,
   :832); the body-side index is not used.  n = 0 clears.  A "fem_gpu"
   cloth keeps and ignores them, as the reference's GPU form does. */
int bw_set_point_triangle_contacts(bw_cloth *c, int n, const int *ids, const double *normal,
                                   const double *bary, const float *depth);
/* The same for edge-edge contacts: cBaseObject::SetEdgeEdgeCollisionInfo
   (BaseObject.cpp:204-207) with tEdgeEdgeCollisionInfo records
   (CollisionInfo.h:62-66).  ids: n x {cloth_is_edge0, edge0 (mEdgeId0),
   edge1 (mEdgeId1)}; normal: n x 3; bary: n x 2 (mBary0, mBary1); depth: n.
   The cloth's edge (edge0 when cloth_is_edge0, else edge1) must be in
   range.  Edge-edge contacts give repulsion impulses only. */
int bw_set_edge_edge_contacts(bw_cloth *c, int n, const int *ids, const double *normal,
                              const float *bary, const float *depth);
/* The contacts set for the next substep (by the calls above or by
   bw_detect_contacts), in the layouts above; each returns the count and
   copies at most cap records (cBaseObject's mPointTriangleCollisionInfo and
   mEdgeEdgeCollisionInfo, BaseObject.h:101-102). */
int bw_get_point_triangle_contacts(const bw_cloth *c, int cap, int *ids, double *normal,
                                   double *bary, float *depth);
int bw_get_edge_edge_contacts(const bw_cloth *c, int cap, int *ids, double *normal, float *bary,
                              float *depth);

/* A kinematic body for collision detection (cKinematicBody,
   src/sim/kinematic/KinematicBody.cpp): a triangle mesh (positions 3*nv,
   triangles 3*nt) whose edges are numbered as cObjUtil::BuildEdge numbers
   them (ascending {min, max}; ObjUtil.cpp:141-224) and whose triangle and
   vertex normals are those of the positions given here (cKinematicBody::Init
   computes them once, KinematicBody.cpp:114-116).  BW_ERR_ARG for an index
   out of range or an edge shared by more than two triangles. */
int bw_body_create(const double *pos, int nv, const int *tri, int nt, bw_body **out);
void bw_body_destroy(bw_body *b);
/* move the body (cKinematicBody::SetMeshPos, KinematicBody.cpp:189-198);
   its normals keep their creation values, as the reference's do */
int bw_body_set_pos(bw_body *b, const double *pos);
int bw_body_num_edges(const bw_body *b);
/* edges as {id0, id1, tri0, tri1 (-1 on the boundary)}, E x 4 */
int bw_body_get_edges(const bw_body *b, int *edge4);
/* Collision detection of the cloth at its current positions against the
   body, on the device (cBVHCollisionDetecter::PerformCD for the pair,
   BVHCollisionDetecter.cpp:360-651; cSimScene::PerformCollisionDetection
   hands the result to the cloth, SimScene.cpp:174-190): every cloth
   vertex / body triangle, body vertex / cloth triangle and edge / edge pair
   closer than 3e-3, in the reference's order, replaces the cloth's contacts
   for its next substep.  cloth_first = 1 when the cloth's object id is
   below the body's (the cloth is obj0 of the pair).  The counts of
   point-triangle and edge-edge contacts go to n_point_triangle and
   n_edge_edge (either may be NULL). */
int bw_detect_contacts(bw_cloth *c, const bw_body *b, int cloth_first, int *n_point_triangle,
                       int *n_edge_edge);
/* User perturbation (the scene's mouse drag, tPerturb, sim/Perturb.h), for
   the plain method; a "fem_gpu" cloth returns BW_ERR_STATE.
   cBaseCloth::ApplyPerturb(tPerturb*) (BaseCloth.cpp:76-87): a third of
   `force` (tPerturb::GetPerturbForce, xyz) is added to the user force of
   each vertex of `triangle` (tPerturb::mAffectedTriId).  The user force
   (mUserForce) enters the next substep's rhs as dt^2 (g + f_user + f)
   (BaraffCloth.cpp:282) and that substep clears it (:154). */
int bw_apply_perturb(bw_cloth *c, int triangle, const double force[3]);
/* cBaraffCloth::ApplyUserPerturbForceOnce(tPerturb*) (BaraffCloth.cpp:165-178):
   the first vertex of `triangle` moves to `goal` (tPerturb::GetGoalPos, xyz;
   current and previous position) and is fixed in the next substep's solve
   like a static vertex (mDragVertexIdx, :291-292), which then releases it
   (:155).  The scene calls it before each frame's update while the user
   drags (cSimScene::UpdateObjects, SimScene.cpp:148-169). */
int bw_apply_user_perturb_force_once(bw_cloth *c, int triangle, const double goal[3]);
/* cBaseCloth::mYDragForce (BaseCloth.h:66; the ImGui slider of
   BaraffCloth.cpp:384, default 0): every substep adds +F/n_up to the y force
   of the n_up vertices on the cloth's top UV edge and -F/n_low on the bottom
   edge (cBaseCloth::ApplyYDragForce, BaseCloth.cpp:592-623, called by
   UpdatePos, BaraffCloth.cpp:129).  Kept until changed. */
int bw_set_y_drag_force(bw_cloth *c, float force);

/* ELL column structure in reference numbering: row v lists, ascending,
   {v} U every interior-edge stencil containing v, as
   cBaraffClothGpu::InitELLMatrix builds it (BaraffClothGpu.cpp:438-500), with
   no 32-column limit.  row_ptr has V+1 entries; returns the block count, and
   only the count when row_ptr or cols is NULL. */
int bw_get_ell_map(const bw_cloth *c, int *row_ptr, int *cols);
/* counters of the last step (no reference counterpart: gProfRec phase times
   are ImGui-only there, BaraffCloth.cpp:393-396) */
int bw_get_last_stats(const bw_cloth *c, bw_step_stats *st);

/* Assemble the implicit system at the current state without solving and
   copy it out: W (nnzb x 3 x 3, row-major blocks in the order of
   bw_get_ell_map) and rhs (3V), both FP64.  Rows and columns of static
   vertices are zero, as the reference's solve filters them
   (BaraffCloth.cpp:469-506, 515-522).  No reference counterpart: the
   reference forms W and b inside SolveForDx (BaraffCloth.cpp:262-298) and
   never exposes them.  Returns nnzb. */
int bw_assemble_system(bw_cloth *c, double *W, double *rhs);

/* last error message of the calling thread */
const char *bw_last_error(void);

/* ---- batched stepping (no reference counterpart: many independent cloths
   of uneven size stepped together; from the second step on, a batch of 8
   or more cloths solves the ones that needed the most launches in chains
   of their own on separate streams, with results bit-identical to one
   chain) -------------------------------------------------------------- */

/* n cloths from n JSON configs; x0s[i] may be NULL */
int bw_batch_create(const char *const *json_configs, const double *const *x0s,
                    int n, bw_batch **out);
void bw_batch_destroy(bw_batch *b);
int bw_batch_size(const bw_batch *b);
/* nsteps substeps of every member */
int bw_batch_step(bw_batch *b, int nsteps);
int bw_batch_num_vertices(const bw_batch *b, int i);
/* true 3x3 blocks of member i's matrix (the ELL column structure's entries,
   bw_get_ell_map's count), for bandwidth accounting; no reference
   counterpart */
int bw_batch_num_blocks(const bw_batch *b, int i);
int bw_batch_num_triangles(const bw_batch *b, int i);
int bw_batch_get_pos(const bw_batch *b, int i, double *x);
int bw_batch_set_state(bw_batch *b, int i, const double *x, const double *xpre);
int bw_batch_get_stats(const bw_batch *b, int i, bw_step_stats *st);
/* bw_get_normals for member i */
int bw_batch_get_normals(bw_batch *b, int i, double *vertex_normals, double *triangle_normals);
/* bw_set_point_triangle_contacts / bw_set_edge_edge_contacts for member i's
   next step */
int bw_batch_set_point_triangle_contacts(bw_batch *b, int i, int n, const int *ids,
                                         const double *normal, const double *bary,
                                         const float *depth);
int bw_batch_set_edge_edge_contacts(bw_batch *b, int i, int n, const int *ids,
                                    const double *normal, const float *bary, const float *depth);
/* bw_apply_perturb, bw_apply_user_perturb_force_once and
   bw_set_y_drag_force for member i */
int bw_batch_apply_perturb(bw_batch *b, int i, int triangle, const double force[3]);
int bw_batch_apply_user_perturb_force_once(bw_batch *b, int i, int triangle,
                                           const double goal[3]);
int bw_batch_set_y_drag_force(bw_batch *b, int i, float force);
/* bw_detect_contacts for member i */
int bw_batch_detect_contacts(bw_batch *bt, int i, const bw_body *b, int cloth_first,
                             int *n_point_triangle, int *n_edge_edge);
/* No reference counterpart.  bw_batch_step for callers that cannot
   allocate (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc): the same kernels, launched directly on the batch's
   stream, with the solver's scratch (two solver states per cloth, two slots
   of FP64 partial sums per 64-row slice, group sums of large cloths) in a
   statically sized device array of 4 MiB per GPU instead of the batch's own
   memory; it allocates no device or host memory (the batch, created
   beforehand, holds everything else).  The solver states are carried in and
   out, so results are those of bw_batch_step with the same SpMV variant:
   the autotuner's pick if bw_batch_step or bw_batch_autotune ran before,
   else the variant the BW_SPMV_* environment variables fix or the default;
   it never tunes.  Limits: tolerance-mode "fem" batches (BW_ERR_STATE
   otherwise, and while contacts or user perturbations are pending);
   BW_ERR_CAPACITY when the batch's scratch exceeds 4 MiB (beyond about
   52,000 slices, 3.3M vertices).  Every call on a GPU shares that GPU's
   array: calls that overlap on one GPU race. */
int bw_batch_step_noalloc(bw_batch *b, int nsteps);
/* use HIP graph replay for the per-iteration launch chains (default on) */
int bw_batch_set_graphs(bw_batch *b, int enable);
/* Measurement (no reference counterpart): time `reps` back-to-back launches
   of the batch's dominant solver kernel, the PCG SpMV (direction update
   d = P r + beta d_prev fused into q = W d and the partial d.q), with HIP
   events on the batch's stream, every cloth active.  avg_us gets the mean
   time per launch, alg_bytes the algorithmic bytes of one launch: 36 + 4
   (FP32 matrix) or 72 + 4 (FP64) bytes per true 3x3 block plus the vector
   values per vertex, each read or written once (tolerance modes: P r and
   d_prev read, d and q written, 12 values; reference mode: r, P, d_prev
   read, d and q written, 15 values).  With the staged SpMV variant the
   column index is a 2-byte position (36 + 2 bytes per block) and each
   entry of the slices' column lists adds 4 bytes.  The system is
   assembled at the current state first; the solver's vectors are
   overwritten (positions are not). */
int bw_batch_time_solver_kernel(bw_batch *b, int reps, double *avg_us, double *alg_bytes);
/* Measurement aid (no reference counterpart): an empty kernel with the grid
   and workgroup size of that SpMV launch (one workgroup per slice), launched
   reps times back to back on the batch's stream; avg_us gets the mean time
   per launch, the floor that dispatch, completion and the kernel boundary
   put under any launch of that shape which waits for the one before. */
int bw_batch_time_launch_floor(bw_batch *b, int reps, double *avg_us);
/* Autotuner (no reference counterpart): time the SpMV variants on the
   batch's own system and keep the fastest, or the first variant in the
   order below (fewer lanes, gathered, cached first) that times within 3%
   of it, so that near-equal variants do not swap between runs; the first
   bw_batch_step does this by itself.  A variant is 1, 2 or 4 lanes per matrix row; the
   direction vector gathered from memory per slot, or staged in LDS once
   per distinct column of a 64-row slice; matrix loads with or without the
   nontemporal hint.  lanes gets the lanes per row chosen, staged and
   nontemporal (each may be NULL) 1 if that form was chosen, us (12
   doubles, may be NULL) the mean launch time of variant
   v = (0, 1, 2 for 1, 2, 4 lanes) + 3 * staged + 6 * nontemporal (0 for a
   variant not timed: the staged form needs every slice to have at most
   2048 distinct columns).  BW_SPMV_LPR=1|2|4, BW_SPMV_STAGE=0|1 and
   BW_SPMV_NT=0|1 in the environment fix those choices. */
int bw_batch_autotune(bw_batch *b, int *lanes, int *staged, int *nontemporal, double *us);

/* ---- serving (no reference counterpart): step requests of different
   cloths arrive at any time, are admitted against a device-memory budget,
   stay resident on the device for their steps, and are solved together with
   everything else resident (continuous batching at the granularity of solver
   chunks; a request's results are the same alone and in any mix).  All
   requests of one server use the same precision and the tolerance solver. -- */

typedef struct bw_server bw_server;

typedef struct bw_request_info
{
    int status;              /* BW_OK, or the failed request's error code */
    int done;                /* all steps taken, or failed */
    int num_vertices;
    int steps, steps_done;
    long long pcg_iterations; /* over all steps */
    double max_rel_residual;  /* worst final residual over the steps */
    double t_submit, t_admit, t_first_step, t_done; /* seconds, server clock */
    double device_bytes;      /* device memory the request held */
    double admitted_bytes;    /* what admission reserved against the budget
                                 before allocating: an upper bound of
                                 device_bytes computed from the mesh */
} bw_request_info;

typedef struct bw_server_counters
{
    int queued, resident, completed, max_resident_seen;
    double bytes_in_use;
    long long chunks, launches; /* multi-request solver chunks and launches */
} bw_server_counters;

/* budget_bytes: device memory for resident requests (e.g. 200e9);
   max_resident: state slots (<= 65536); precision: "mixed" or "fp64" */
int bw_server_create(double budget_bytes, int max_resident, const char *precision,
                     bw_server **out);
/* stops the server; requests still running are abandoned */
void bw_server_destroy(bw_server *s);
/* enqueue nsteps substeps of the cloth in json (the cloth JSON of
   bw_create); x0 (nx0 = 3V doubles) may be NULL.  Returns at once; errors
   in the configuration are reported by the request's status. */
int bw_server_submit(bw_server *s, const char *json, const double *x0, int nx0, int nsteps,
                     long long *id);
/* BW_OK once the request is done (or failed), BW_ERR_STATE on timeout */
int bw_server_wait(bw_server *s, long long id, double timeout_s);
int bw_server_get_info(bw_server *s, long long id, bw_request_info *info);
/* final positions (3V doubles) of a request that completed with BW_OK */
int bw_server_get_pos(bw_server *s, long long id, double *x);
/* forget a finished request */
int bw_server_release(bw_server *s, long long id);
int bw_server_get_counters(bw_server *s, bw_server_counters *c);
/* seconds on the server's clock */
double bw_server_now(bw_server *s);

#ifdef __cplusplus
}
#endif
#endif
