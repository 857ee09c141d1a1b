// Device-resident batch of independent cloths and the host driver of one
// implicit substep (cBaraffCloth::UpdatePos, BaraffCloth.cpp:90-163).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "collide.h"
#include "device.h"
#include "model.h"

namespace bw
{

struct StepStats
{
    int pcg_iterations = 0;
    int outer_iterations = 0; // FP64 true-residual checks (reliable updates)
    double rel_residual = 0;
    double step_us = 0;
    int matrix_fp64 = 0; // mixed form: the last solve moved to the FP64 matrix
};

// how a batch solves: the reference's own PCG with exactly max_iters-1
// iterations (FP64), or the tolerance-mode PCG machine in FP32 (mixed) or FP64
enum class SolverKind
{
    RefPCG64,
    CG32,
    CG64,
    GpuForm // cloth_type "fem_gpu": the reference's GPU form (gpuform.hip)
};

// Contacts handed to one cloth for its next substep, as the scene's
// collision detection hands them to each object
// (cBaseObject::SetPointTriangleCollisionInfo / SetEdgeEdgeCollisionInfo,
// BaseObject.cpp:199-207; records tPointTriangleCollisionInfo and
// tEdgeEdgeCollisionInfo, sim/collision/CollisionInfo.h:45-66)
struct Contacts
{
    std::vector<int> pt_ids;       // 3 per contact: cloth_is_point, vertex, triangle
    std::vector<double> pt_normal; // 3
    std::vector<double> pt_bary;   // 3
    std::vector<float> pt_depth;
    std::vector<int> ee_ids;       // 3 per contact: cloth_is_edge0, edge0, edge1
    std::vector<double> ee_normal; // 3
    std::vector<float> ee_bary;    // 2: mBary0, mBary1
    std::vector<float> ee_depth;
    int size() const { return (int)(pt_depth.size() + ee_depth.size()); }
};

// Optional resources a Batch borrows instead of creating: the stream its
// kernels run on, and device storage for its solver states (2 CgState per
// cloth), so that a server can keep many batches' states in one table.
struct BatchOptions
{
    hipStream_t stream = nullptr;
    CgState *cg_states = nullptr;
    // the batch's solver launches are shared with other batches (a server's
    // rounds): group sums pay off from smaller cloths
    bool shared_launches = false;
};

class Batch
{
public:
    explicit Batch(std::vector<ClothModel> models, const BatchOptions &opt = BatchOptions());
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    int size() const { return (int)models_.size(); }
    const ClothModel &model(int i) const { return models_[i]; }
    SolverKind kind() const { return kind_; }
    // pins changed on the host model: re-upload the pinned mask
    void refresh_pins(int i, const std::vector<int> &pins);

    void step(int nsteps);
    // step() for callers that cannot allocate: the tolerance solvers only,
    // no contacts or user inputs pending, direct launches (no graphs), no
    // autotuning, and the solver's scratch (states, partial and group sums)
    // in the statically sized device array of noalloc.hip instead of the
    // batch's own memory; allocates no device or host memory
    void step_noalloc(int nsteps);
    // One step driven from outside (tolerance solvers): begin_step queues the
    // element, assembly and solver-reset kernels; the caller then runs the
    // solver launches (launch_cg_step_multi) until every cloth's state says
    // done, and end_step takes those states (host copy, 2 per cloth), fills
    // the step counters and queues the position update.
    void begin_step();
    void end_step(const CgState *states, double step_us);
    size_t device_bytes() const { return device_bytes_; }
    // choose the SpMV variant by timing (done by the first step()); returns
    // the lanes per row picked; tune_us() has the times of variant v =
    // lanes index (1, 2, 4) + 3 * staged + 6 * nontemporal (0: not timed;
    // the staged form needs every slice to have at most HALO_CAP columns)
    static constexpr int TUNE_VARIANTS = 12;
    int autotune();
    // the SpMV choices BW_SPMV_LPR / BW_SPMV_STAGE / BW_SPMV_NT fix (0 / -1:
    // free)
    void fixed_variant(int &lpr, int &stage, int &nt) const;
    const double *tune_us() const { return tune_us_; }
    int spmv_lpr() const { return d_.spmv_lpr; }
    bool spmv_stage() const { return d_.spmv_stage != 0; }
    bool spmv_nt() const { return d_.spmv_nt != 0; }
    int halo_max() const { return d_.halo_max; }
    int halo_lds() const { return d_.halo_lds; }
    // assemble W and rhs at the current state (no solve) and copy them out in
    // the model's block order
    void assemble_system(int i, double *W, double *rhs);
    void get_pos(int i, double *x) const;
    void get_pre_pos(int i, double *x) const;
    void set_pos(int i, const double *x);
    void set_pre_pos(int i, const double *x);
    // GPU form: forget the warm start (the reference zeroes its solution on
    // Reset, BaraffClothGpu.cpp:293-299)
    void clear_warm_start(int i);
    // normals at the current positions (cBaseObject::UpdateTriangleNormal /
    // UpdateVertexNormalFromTriangleNormal): vertex normals 3 nv, triangle
    // normals 3 nt, in the model's order; either may be null
    void normals(int i, double *vertex_normals, double *triangle_normals);
    // contacts for cloth i's next substep, which clears them (the collision
    // Hessian enters W, the repulsion impulses the solution; the GPU form
    // ignores them, as the reference's does); n = 0 clears one kind
    void set_point_triangle_contacts(int i, int n, const int *ids, const double *normal,
                                     const double *bary, const float *depth);
    void set_edge_edge_contacts(int i, int n, const int *ids, const double *normal,
                                const float *bary, const float *depth);
    int num_contacts(int i) const { return contacts_[i].size(); }
    // User inputs of cloth i's next substep, which clears them (the plain
    // method only).  apply_perturb: cBaseCloth::ApplyPerturb
    // (BaseCloth.cpp:76-87), a third of `force` added to the user force of
    // each vertex of triangle `tri` (mUserForce, which enters the rhs
    // dt^2 (g + f_user + f), BaraffCloth.cpp:282).
    // apply_user_perturb_force_once: cBaraffCloth::ApplyUserPerturbForceOnce
    // (BaraffCloth.cpp:165-178), the triangle's first vertex moved to `goal`
    // (current and previous position) and fixed in the next solve
    // (mDragVertexIdx, :291-292).  set_y_drag_force: cBaseCloth::mYDragForce,
    // applied by every substep (ApplyYDragForce, BaseCloth.cpp:592-623).
    // clear_user_force: cBaseCloth::ClearForce's user-force part (Reset).
    void apply_perturb(int i, int tri, const double force[3]);
    void apply_user_perturb_force_once(int i, int tri, const double goal[3]);
    void set_y_drag_force(int i, float f);
    void clear_user_force(int i);
    const Contacts &contacts(int i) const { return contacts_[i]; }
    // collision detection of cloth i at its current positions against a
    // kinematic body (cBVHCollisionDetecter::PerformCD for the pair,
    // BVHCollisionDetecter.cpp:360-651); the contacts replace cloth i's for
    // its next substep.  cloth_first: the cloth is obj0 of the pair (its
    // object id is the lower one).  Returns the point-triangle and
    // edge-edge counts.
    void detect_contacts(int i, const Body &body, bool cloth_first, int *n_pt, int *n_ee);
    const StepStats &stats(int i) const { return stats_[i]; }
    void set_graphs(bool on) { use_graphs_ = on && !graphs_disabled_; }
    // mean microseconds per launch of the dominant solver kernel over reps
    // launches, every cloth active; algorithmic bytes per launch
    double time_solver_kernel(int reps, double *alg_bytes);
    // mean microseconds per launch of an empty kernel with the grid and
    // workgroup size of that kernel, launched back to back on the batch's
    // stream: the floor under the solver kernel's launch time
    double time_launch_floor(int reps);

    const DeviceBatch &device() const { return d_; }

private:
    void build_layout();
    void upload_state(int i, const double *src, double *dst_dev);
    // cloth i (all for -1) starts its next solve on the FP32 matrix
    void mark_fresh(int i);
    void download(int i, const double *src_dev, double *dst) const;
    struct GraphEntry
    {
        int length, slices, first;
        hipGraphExec_t exec;
        unsigned long long last_use;
    };
    // A chain of solver launches on one stream over a set of the cloths (see
    // plan_lanes).  lanes_[0] runs on the batch's stream with its list and
    // state buffers; further lanes own theirs.
    struct Lane
    {
        hipStream_t stream = nullptr;
        hipEvent_t ev = nullptr; // after the lane's last queued state copy
        bool owned = false;      // stream, event and host buffers are the lane's
        int lds_reserve = 0;     // dynamic LDS bytes per solver workgroup
        SolverEntry *spmv_list = nullptr, *upd_list = nullptr; // device
        SolverEntry *host_spmv = nullptr, *host_upd = nullptr; // pinned
        CgState *host_cg = nullptr;                    // pinned, 2 per cloth
        // first slice of the contiguous range the next launches cover, or
        // -1 when they run from the slice lists
        int launch_first = -1;
        std::vector<GraphEntry> graphs;
        std::vector<int> members;      // cloths of this solve
        std::vector<char> active;      // per cloth: still solving
        std::vector<long long> bounds; // chunk boundaries (launch counts)
        size_t next = 0;
        long long launched = 0;
        int blind = 0;                 // chunks run without a predicted boundary
        bool busy = false, done = false;
    };
    std::vector<Lane> lanes_;
    hipEvent_t lanes_start_ = nullptr;
    bool lanes_allowed_ = true; // false inside bw_batch_step_noalloc
    std::vector<int> lane_order_;
    int plan_lanes();
    void open_lane(Lane &L);
    void close_lane(Lane &L);
    void lane_chunk(Lane &L, bool sync);
    bool lane_update(Lane &L, long long cap);
    void clear_graphs();
    // states to the lane's host copy; sync: wait for it, else record L.ev
    void read_states(Lane &L, bool sync = true);
    // one chunk of solver launches; slices: the k_cg_spmv grid
    void launch_chain(Lane &L, int length, int slices);
    void run_chunk(Lane &L, int length, int slices); // as replayed graphs when enabled
    void run_graph(Lane &L, int length, int slices);
    // chunk boundaries of this solve of the lane's cloths into L.bounds
    void chunk_schedule(Lane &L);
    std::vector<std::pair<long long, int>> sched_pairs_;
    int upload_lists(Lane &L);
    void solve_ref();
    void solve_cg();
    // contacts: device lists for the next step (false if there are none)
    bool upload_contacts();
    // user inputs: upload the next step's user forces and drag pins before
    // it, and take the drag pins back and forget the inputs after it
    void prepare_user_inputs();
    void finish_user_inputs();
    std::vector<std::vector<double>> user_force_; // per cloth: 3 nv, or empty
    std::vector<int> drag_vertex_;                // per cloth: -1 or the vertex
    std::vector<float> ydrag_;                    // per cloth
    std::vector<char> fuser_dirty_;               // d_.fuser rows of the cloth hold values
    std::vector<char> drag_pinned_;               // row_pinned set for the drag vertex
    void clear_contacts();

    std::vector<ClothModel> models_;
    std::vector<int> row_base_, slice_base_, nslices_;
    std::vector<int> vert_base_, tri_base_; // first global vertex / triangle per cloth
    double *normals_dev_ = nullptr;         // 3 per triangle, then 3 per vertex (lazy)
    std::vector<std::vector<int>> row_of_; // per cloth: vertex -> row within the cloth
    std::vector<StepStats> stats_;
    DeviceBatch d_;
    hipStream_t stream_ = nullptr;
    BatchOptions opt_;
    bool own_stream_ = false;
    bool tuned_ = false;
    double tune_us_[TUNE_VARIANTS] = {};
    long long n_halo_ = 0; // entries of the staged SpMV's column lists
    size_t device_bytes_ = 0;
    SolverKind kind_ = SolverKind::CG32;
    InstState *host_state_ = nullptr; // pinned, reference PCG state
    CgState *host_cg_ = nullptr;      // pinned, 2 per cloth
    GfState *host_gf_ = nullptr;      // pinned, 2 per cloth
    void solve_gpu_form();
    bool use_graphs_ = true;
    bool graphs_disabled_ = false; // BW_GRAPHS=0 in the environment
    int last_launches_ = 0;
    std::vector<long long> pred_launches_;      // per cloth, from the last solve
    SolverEntry *host_spmv_list_ = nullptr, *host_upd_list_ = nullptr; // pinned (lane 0)
    std::vector<int4> slice_info_; // host copy of d_.slice_info (list entries carry it)
    unsigned long long graph_clock_ = 0;
    std::vector<void *> allocations_;
    template <typename T> T *dalloc(size_t n);
    std::vector<int> host_slice_off_;
    std::vector<Contacts> contacts_;
    int contact_total_ = 0;          // contacts over all cloths
    void *contact_dev_ = nullptr;    // device lists of the next step
    size_t contact_cap_ = 0;
    std::vector<char> contact_host_; // their host staging (kept until the step ends)
    // detection: global vertex -> row, 2 rows per edge (all cloths), scratch
    int *vrow_dev_ = nullptr, *erow_dev_ = nullptr;
    std::vector<int> edge_base_;
    void *detect_count_dev_ = nullptr, *detect_hit_dev_ = nullptr;
    size_t detect_count_cap_ = 0, detect_hit_cap_ = 0;
    int contact_blocks_ = 0, contact_groups_ = 0;
    const long long *c_widx_ = nullptr;
    const int *c_diag_ = nullptr, *c_wgl_ = nullptr, *c_lvl_ = nullptr;
    const int *c_wgr_ = nullptr, *c_rows_ = nullptr, *c_chl_ = nullptr;
    const double *c_h9_ = nullptr, *c_rec_ = nullptr;
    const int4 *c_ids_ = nullptr;
};

} // namespace bw
