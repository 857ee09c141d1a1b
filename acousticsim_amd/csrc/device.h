// Shared host/device declarations: the batched device layout and the kernel
// launchers.  One "batch" holds any number of independent cloths; every
// array below is indexed by global (batch-wide) ids.
//
// Matrix layout: sliced ELL with slice height SLICE = 64.  Rows of one cloth
// occupy whole slices.  Entry (slice s, slot k, row l) has index
// e = slice_off[s] + k*64 + l and column col[e]; its 3x3 block is stored as
// nine planes of 64 values: W[9*slice_off[s] + (9*k + q)*64 + l] for
// q = 3*a + b, so a load of one q over 64 rows is one contiguous run.
// (Block-major FP32 storage read with 16-byte loads measured no faster.)
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace bw
{

constexpr int SLICE = 64;
// column-index entries after the last slice's: the gathered SpMV loads the
// first chunk of a row's column indices (up to 8 slots) before it knows the
// row's width, so those loads stay inside the allocation
constexpr int COL_PAD = 8 * SLICE;
// longest per-slice column list the staged SpMV keeps in LDS (4 values of
// the solver's precision per entry: 32 KB in FP32, 64 KB in FP64)
constexpr int HALO_CAP = 2048;

// Global-address-space views of device pointers.  Pointers that kernels read
// from device memory (a table of batches) are generic to the compiler, and
// generic (flat) accesses cost more than global ones; every buffer here is
// global memory, so the solver kernels access all of them through G<T>.
// (The host compilation pass of these sources has no such address space.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T> using G = __attribute__((address_space(1))) T;
#else
template <typename T> using G = T;
#endif
template <typename T> __host__ __device__ inline G<T> *glob(T *p) { return (G<T> *)p; }
// slices per solver workgroup (SB * 64 threads); a cloth's slices are padded
// to a multiple of it so that no workgroup spans two cloths
#ifndef BW_SB
#define BW_SB 4
#endif
// FP64 partial sums per workgroup
constexpr int PARTS = 5;
// slices per group sum of the tolerance-mode solver (cloths of more than
// group_min slices are reduced through group sums)
constexpr int GROUP = 64;

struct InstParams
{
    double dt, sk, sm, k0, k1, k2, grav;
    double tol;       // relative residual tolerance
    int max_iter;     // PCG iteration cap (reference mode: max_iters - 1)
    int ref_mode;     // 1: reference semantics (no tolerance stop)
    double rel2;      // reliable update when r.r drops below rel2 * (r.r at the last one)
    int max_since;    // ... or after this many iterations
    int slice_begin, slice_count;
    int blk_begin, blk_count; // solver workgroups of the cloth
    int group_begin, group_count; // group sums of the cloth (0: none)
    // mixed form: a true residual forced by max_since whose r.r is above
    // stall * (r.r at the last replacement) switches the cloth's SpMV to the
    // FP64 matrix for the rest of the solve (0: never)
    double stall;
    // a cloth whose last `carry` solves (consecutively) ended on the FP64
    // matrix starts its next solve on it
    int carry;
    float gf_a, gf_b; // GPU form: Rayleigh a, b as the reference's float members
};

// Per-cloth state of the reference-semantics PCG (pcg.hip).  Each field has
// one writing kernel.
struct InstState
{
    double rho[2];      // r.(P r) of iterations of parity 0 / 1
    double rr, bb, rr64;
    int it_a;           // next iteration (written by the update kernel)
    int it_b;           // current iteration (written by the SpMV kernel)
    int done, total_iter, breakdown, pad;
};

// Per-cloth state of the tolerance-mode PCG machine (cg.hip); two slots:
// k_cg_spmv reads slot 0 and writes slot 1, k_cg_update the reverse.
struct CgState
{
    double rho;     // r.(P r) of the current direction
    double rr_ref;  // true r.r at the last residual replacement
    double bb;      // b.b
    double rr_true; // true r.r at the last check
    double rr;      // recursive r.r
    int start;      // 1: the warm start's residual is still to be formed
    int first;      // 1: next direction is P r (no beta term)
    int iter;       // direction index (selects the direction buffer)
    int total;      // iterations this solve
    int since;      // iterations since the last residual replacement
    int updates;    // true residuals formed this solve
    int done, breakdown;
    int pending;     // 1: the last k_cg_spmv formed a true residual
    int force_check; // 1: form a true residual next
    int final_check; // 1: stop after the next true residual
    int w64;         // 1: the SpMV reads the FP64 matrix (mixed form, after a stall)
    int stalls;      // consecutive solves of the cloth before this one that
                     // ended on the FP64 matrix
};

// Per-cloth state of the GPU-form PCG (gpuform.hip); two slots like CgState
struct GfState
{
    double rmr; // r.(M^-1 r) of the current direction
    double rr;  // r.r after the last update
    double bb;  // b.b over unpinned rows
    int it;     // iterations done
    int done;
};

// An entry of a solver launch's list: the slice (k_cg_spmv) or solver
// workgroup (k_cg_update) to run, with the slice_info entry of that slice
// (of the workgroup's first slice), so that the workgroup's first load
// gives it both; padding has every field negative (info.z INT_MIN).
struct __align__(16) SolverEntry
{
    int4 info;
    int id, pad0, pad1, pad2;
};

struct DeviceBatch
{
    int n_inst = 0, n_slices = 0, n_rows = 0;
    int xcd_map = 1;   // solver workgroups remapped so each XCD owns a contiguous range
    int spmv_lpr = 4;  // k_cg_spmv variant: lanes per row (1, 2 or 4)
    int spmv_stage = 0; // k_cg_spmv variant: directions staged in LDS (1) or gathered (0)
    int spmv_nt = 0;    // k_cg_spmv variant: matrix loads with the nontemporal hint (1)
    long long n_entries = 0;
    int n_tri = 0, n_ie = 0, n_vert = 0, n_blocks = 0;
    int n_blocks_wg = 0; // solver workgroups
    // per solver workgroup / slice / row
    int *blk_inst = nullptr;
    int *slice_off = nullptr, *slice_w = nullptr, *slice_inst = nullptr;
    int *row_vertex = nullptr, *row_w = nullptr;
    uint8_t *row_pinned = nullptr;
    double *row_mass = nullptr;
    // matrix
    int *col = nullptr;
    // the staged SpMV: per slice the ascending distinct column rows of its
    // entries (halo_row[halo_off[s] ...], n_slices + 1 offsets), each
    // entry's position in its slice's list (lcol, indexed like col), and the
    // longest list (halo_max; 0: the staged variant is unavailable)
    int *halo_off = nullptr, *halo_row = nullptr;
    uint16_t *lcol = nullptr;
    int halo_max = 0;
    // LDS entries the staged variant provides per workgroup: a slice whose
    // list is longer takes the gathered product instead (a few slices with
    // long lists would otherwise size every workgroup's LDS, and LDS bounds
    // how many workgroups a CU holds)
    int halo_lds = 0;
    double *W64 = nullptr;
    float *W32 = nullptr;
    // assembly inputs (global vertex ids map to physical rows)
    int *vert_row = nullptr;                       // n_vert
    int *blk_ptr = nullptr;                        // per vertex, n_vert + 1 (CSR over blocks)
    int *blk_col = nullptr;                        // physical column row per block
    int *s_ptr = nullptr, *s_list = nullptr;       // stretch contributions per block
    int *b_ptr = nullptr, *b_list = nullptr;       // bending contributions per block
    int *vt_ptr = nullptr, *vt_list = nullptr;     // per vertex incident triangles
    int *ve_ptr = nullptr, *ve_list = nullptr;     // per vertex incident interior edges
    int *tri_rows = nullptr;                       // 3 per triangle (physical rows)
    int *tri_inst = nullptr;
    float *tri_coef = nullptr, *tri_area = nullptr;
    int *ie_rows = nullptr;                        // 4 per interior edge
    double *ie_factor = nullptr;
    double *tri_ing = nullptr, *ie_ing = nullptr;  // 25 doubles per element
    // state and vectors (3 per row)
    double *x = nullptr, *xpre = nullptr, *dx = nullptr, *b = nullptr, *P64 = nullptr;
    // user force of the next step (3 per row; the reference's mUserForce),
    // allocated on first use; null: zero everywhere
    double *fuser = nullptr;
    double *r64 = nullptr, *d64[2] = {nullptr, nullptr}, *q64 = nullptr;
    float *P32 = nullptr;
    double *partial = nullptr; // PARTS per solver workgroup
    InstParams *params = nullptr;
    InstState *state = nullptr;
    // tolerance-mode PCG: state (2 slots per cloth), partials (2 slots x
    // PARTS per slice), FP32 vectors of the mixed form (the FP64
    // form uses r64, z64, q64, d64); z = P r
    CgState *cg = nullptr;
    double *cg_partial = nullptr;
    // group sums (2 slots x PARTS per group of GROUP slices) of large cloths,
    // and the cloth of each group
    double *cg_group = nullptr;
    int *group_inst = nullptr;
    int n_groups = 0;
    float *r32 = nullptr, *z32 = nullptr, *q32 = nullptr, *d32[2] = {nullptr, nullptr};
    // mixed form: the solution's increments since the last true residual,
    // accumulated in FP32 and folded into the FP64 dx by the k_cg_update
    // that follows each true residual (the solution is dx + dx32)
    float *dx32 = nullptr;
    // per cloth: 1 until the next k_cg_reset, which then starts the cloth
    // on the FP32 matrix (else a cloth keeps the FP64 matrix it moved to)
    unsigned char *cg_fresh = nullptr;
    double *z64 = nullptr;
    // GPU form (gpuform.hip): bending scalar per ELL entry, per-triangle
    // FP32 force and Hessian (90), block-Jacobi inverse (9 per row), the
    // persistent FP32 solution, pinned targets, FP32 mass and gravity per
    // row, two state slots per cloth
    float *gf_qb = nullptr, *gf_tri = nullptr, *gf_pinv = nullptr, *gf_dx = nullptr;
    float *gf_target = nullptr, *gf_mass = nullptr, *gf_grav = nullptr;
    GfState *gf_state = nullptr;
    // optional lists of the slices (k_cg_spmv) and solver workgroups
    // (k_cg_update) a launch runs; null: the contiguous range of slices from
    // slice_first (a multiple of BW_SB) that the grid covers
    SolverEntry *spmv_list = nullptr, *upd_list = nullptr;
    int slice_first = 0;
    int lds_pad = 0; // dynamic LDS bytes the solver launches reserve and leave unused
    // per slice, what the solver kernels need before anything else, in one
    // load: {cloth, first partial set of the cloth's solver sums, their
    // count (negative: -count group sums), the slice's first ELL entry}
    int4 *slice_info = nullptr;
};

// Workgroup b of a G-workgroup launch runs on XCD b % 8 (the dispatcher deals
// workgroups round-robin over the 8 XCDs).  The logical block returned here
// gives XCD x the contiguous range [x*q + min(x, G%8), ...) of G/8 (+1)
// blocks, so neighbouring slices of a cloth, which gather the same vector
// entries, share one XCD's L2.
__device__ inline int solver_block(int xcd_map)
{
    const int b = blockIdx.x;
    if (!xcd_map)
        return b;
    const int G = gridDim.x, q = G / 8, rem = G % 8, x = b % 8;
    return x * q + (x < rem ? x : rem) + b / 8;
}

// assembly (assemble.hip)
void launch_elements(const DeviceBatch &d, hipStream_t s);
void launch_assemble_rows(const DeviceBatch &d, hipStream_t s);
// solver (pcg.hip)
void launch_reset_states(const DeviceBatch &d, hipStream_t s);
void launch_pcg_init64(const DeviceBatch &d, hipStream_t s);
void launch_pcg_iter64(const DeviceBatch &d, hipStream_t s);
// mode 0: FP64 solve start; 2: counters only
void launch_residual64(const DeviceBatch &d, int mode, hipStream_t s);
void launch_update_positions(const DeviceBatch &d, hipStream_t s);
// tolerance-mode PCG (cg.hip); phase 0: k_cg_spmv, 1: k_cg_update, each
// followed by k_cg_group when the batch has group sums.  slices: the
// k_cg_spmv grid (list length, or n_slices without lists); k_cg_update runs
// ceil(slices / BW_SB) workgroups.
void launch_cg_reset(const DeviceBatch &d, hipStream_t s);
void launch_cg_step(const DeviceBatch &d, bool fp32, int phase, int slices, hipStream_t s);
// k_cg_spmv alone (kernel timing)
void launch_cg_spmv_only(const DeviceBatch &d, bool fp32, hipStream_t s);
// the same over many batches: `table` holds the batches (device memory),
// spmv_list one (batch, slice) entry per k_cg_spmv workgroup, upd_list one
// (batch, solver workgroup) entry per k_cg_update workgroup, grp_list one
// (batch, group) entry per group-sum wavefront (grp_grid 0: none); batch < 0
// pads; halo_max > 0 runs the staged SpMV variant (every batch in the table
// must have its lists) with LDS for that many entries; nt: matrix loads with
// the nontemporal hint
void launch_cg_step_multi(const DeviceBatch *table, const int2 *spmv_list, int spmv_grid,
                          const int2 *upd_list, int upd_grid, const int2 *grp_list, int grp_grid,
                          bool fp32, int phase, int lpr, int halo_max, bool nt, int xcd_map,
                          hipStream_t s);
void launch_cg_timing_setup(const DeviceBatch &d, bool fp32, hipStream_t s);
// an empty kernel of grid x block threads (the dependent-launch floor)
void launch_empty(int grid, int block, hipStream_t s);
// group sums of the solver partials in `slot` for cloths that have them
void launch_group_sums(const DeviceBatch &d, int slot, hipStream_t s);
// GPU form (gpuform.hip): pins, FP32 elements and rows; then the 100-
// iteration block-Jacobi solve and the position update
void launch_gf_assemble(const DeviceBatch &d, hipStream_t s);
void launch_gf_solve(const DeviceBatch &d, hipStream_t s);
// render readback (normals.hip): triangle normals (3 per global triangle)
// and vertex normals (3 per global vertex) at the current positions
void launch_normals(const DeviceBatch &d, double *tn, double *vn, hipStream_t s);

// collision response (contact.hip): doubles per repulsion record
// {N[3], w[3], mass, den, d, coef, dt*K, dt}
constexpr int CONTACT_REC = 12;
// W += H for n summed 3x3 blocks; w_index = index of the block's entry 0 in
// W64/W32 (entry q at + q*SLICE); diag_row = the row of a diagonal block
// (its Jacobi diagonal is refreshed) or -1
void launch_contact_hessian(const DeviceBatch &d, int n, const long long *w_index,
                            const int *diag_row, const double *h9, hipStream_t s);
// repulsion impulses on dx: workgroup g runs chunks [wg_chunk[g],
// wg_chunk[g+1]), chunk c the levels [chunk_level[c], chunk_level[c+1]),
// level L the contacts [level_ptr[L], level_ptr[L+1]); a chunk holds at most
// REPULSION_CHUNK contacts unless it is one wider level.  ids = {type (0
// edge-edge, 1 cloth vertex, 2 cloth triangle), slots...}, a slot indexing
// the workgroup's touched rows rows[wg_rows[g] ... wg_rows[g+1]); up to
// REPULSION_LDS_ROWS of them are kept in LDS
constexpr int REPULSION_LDS_ROWS = 1536;
constexpr int REPULSION_CHUNK = 192;
void launch_repulsion(const DeviceBatch &d, int n_groups, const int *wg_chunk,
                      const int *chunk_level, const int *level_ptr, const int *wg_rows,
                      const int *rows, const int4 *ids, const double *rec, hipStream_t s);
// measurement helpers: set the solver residual to 1 on every row; one SpMV
// launch of the solver's precision
void launch_fill_residual(const DeviceBatch &d, bool fp32, hipStream_t s);
void launch_spmv_only(const DeviceBatch &d, bool fp32, hipStream_t s);

} // namespace bw
