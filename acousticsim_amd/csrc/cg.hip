// Tolerance-mode solver: Jacobi-PCG (the recurrences of cBaraffCloth::Solve,
// BaraffCloth.cpp:507-611) as a per-cloth state machine on the device, with
// reliable updates.
//
// One iteration is two launches, k_cg_spmv then k_cg_update.  Every cloth's
// state, and the FP64 partial sums (one set per 64-row slice), have two
// slots: k_cg_spmv reads slot 0 and writes slot 1, k_cg_update reads slot 1
// and writes slot 0, so no workgroup reads what another workgroup of the same
// launch writes, and no grid-wide synchronisation is needed.  Every
// wavefront of a cloth sums the cloth's partials of the previous launch in
// the same fixed order, so all of them derive identical scalars and take the
// same branch, and a cloth's arithmetic does not depend on the rest of the
// batch.  That redundant sum costs every workgroup a pass over all of its
// cloth's partials, which grows with the square of the cloth's size; cloths
// of more than group_min slices therefore get a third, tiny launch after each
// of the two, k_cg_group, which sums each group of 64 slice partials once,
// and their workgroups read the group sums instead (a 1.5M-vertex cloth
// reads 366 sums rather than 23k partials per workgroup).
//
//   k_cg_spmv    one 64-row slice per workgroup, L = 1, 2 or 4 lanes per
//                row (lane phase j takes ELL slots j, j+L, ...; the partial
//                row sums meet by lane shuffles; more lanes shorten the
//                serial chain of dependent loads per lane, which bounds a
//                cloth solved alone), the direction gathered per slot or
//                staged in LDS per distinct column of the slice, matrix
//                loads cached or nontemporal: twelve variants, autotuned.
//                Prologue: reduce r.z and r.r of the previous update;
//                beta = rho_new / rho; d = z + beta d_prev, formed on the
//                fly for every gathered (or staged) column; q = W d;
//                partial d.q.
//                When a check is due it computes the true residual instead,
//                r = b - W64 dx in FP64, z = P r (partials r.z, r.r, b.b),
//                and the following k_cg_update idles for that cloth.
//   k_cg_update  four slices per workgroup, one lane per row: alpha =
//                rho / d.q; dx += alpha d, r -= alpha q, z = P r;
//                partials r.z, r.r.
// z = P r is stored so that the SpMV gathers two vectors per column (z and
// d_prev) rather than three.  Both kernels take an optional list of the
// slices (workgroups) to run, so that the host can shrink the grid to the
// cloths still solving (entries < 0 are padding).
//
// The recurrences are Hestenes-Stiefel PCG, the reference's own.  Single-
// reduction rearrangements (Chronopoulos-Gear, or r.(P r) from an alpha^2
// q.Pq recurrence) would save a launch per iteration but stall with FP32
// recurrences on stiff cloths; this form does not.
//
// "mixed": W (FP32 copy) and r, z, q, d are FP32; every dot product and
// scalar is FP64; the solution is dx (FP64) + dx32, where dx32 accumulates
// the iterations' increments in FP32 and is folded into dx after each true
// residual (which reads both), so an iteration moves 24 rather than 48
// bytes of solution per row and the FP64 residual check still sees the
// exact iterate it judges.  Reliable updates (Sleijpen and
// van der Vorst, 1996) replace the recursive residual by the true FP64
// residual each time r.r has dropped by rel2 since the last replacement,
// after max_since iterations, and before convergence is accepted; the Krylov
// recurrence then continues from the replaced residual
// (beta = r_new.z_new / rho_old).  Convergence is always decided on the FP64
// true residual: ||b - W64 dx|| <= tol ||b||.  "fp64" runs the same machine
// with T = double.
#include <algorithm>
#include <climits>
#include <cstdlib>

#include "device.h"
#include "reduce.h"

namespace bw
{
namespace
{

// field-wise state copies: whole-struct copies make the compiler keep a
// per-thread copy of the state in LDS or scratch
template <typename SP> __device__ inline CgState load_state(SP p)
{
    CgState s;
    s.rho = p->rho;
    s.rr_ref = p->rr_ref;
    s.bb = p->bb;
    s.rr_true = p->rr_true;
    s.rr = p->rr;
    s.start = p->start;
    s.first = p->first;
    s.iter = p->iter;
    s.total = p->total;
    s.since = p->since;
    s.updates = p->updates;
    s.done = p->done;
    s.breakdown = p->breakdown;
    s.pending = p->pending;
    s.force_check = p->force_check;
    s.final_check = p->final_check;
    s.w64 = p->w64;
    s.stalls = p->stalls;
    return s;
}

template <typename SP> __device__ inline void store_state(SP p, const CgState &s)
{
    p->rho = s.rho;
    p->rr_ref = s.rr_ref;
    p->bb = s.bb;
    p->rr_true = s.rr_true;
    p->rr = s.rr;
    p->start = s.start;
    p->first = s.first;
    p->iter = s.iter;
    p->total = s.total;
    p->since = s.since;
    p->updates = s.updates;
    p->done = s.done;
    p->breakdown = s.breakdown;
    p->pending = s.pending;
    p->force_check = s.force_check;
    p->final_check = s.final_check;
    p->w64 = s.w64;
    p->stalls = s.stalls;
}

// SpMV variants: L lanes per row (1, 2 or 4; a workgroup of L wavefronts
// covers one 64-row slice, lane phase j takes ELL slots j, j+L, ...), slot
// loop unrolled 4/L times.  Fewer lanes per row keep more independent loads
// per lane in flight; more lanes shorten each lane's chain of dependent
// loads, which is what bounds a cloth solved alone.  The batch picks the
// fastest on its own matrix (Batch::autotune).

// The solver's arithmetic is written out operation by operation, fused
// multiply-adds included, and this file is compiled without floating-point
// contraction: every kernel and variant that computes the same quantity
// rounds it the same way, whatever the surrounding code (a cloth's result
// cannot depend on which launch shape or kernel variant computed it).
__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// d = z + beta d_prev, z = P r
template <typename T> __device__ inline T dir_value(T z, T dold, T beta, bool first)
{
    return first ? z : fma_t(beta, dold, z);
}

// one row of a 3x3 block times a 3-vector: (m0 v0 + m1 v1) + m2 v2 with the
// two additions fused
template <typename T> __device__ inline T block_row(T m0, T m1, T m2, T v0, T v1, T v2)
{
    return fma_t(m2, v2, fma_t(m1, v1, m0 * v0));
}

// sum over the L lane phases of a row (lanes l, l+64/L, ...)
template <int L, typename T> __device__ inline T phase_sum(T v)
{
    for (int off = SLICE / L; off < 64; off <<= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// the solver vectors of precision T
template <typename T> struct Vecs;
template <> struct Vecs<float>
{
    __device__ static const float *W(const DeviceBatch &d) { return d.W32; }
    __device__ static const float *P(const DeviceBatch &d) { return d.P32; }
    __device__ static float *r(const DeviceBatch &d) { return d.r32; }
    __device__ static float *z(const DeviceBatch &d) { return d.z32; }
    __device__ static float *q(const DeviceBatch &d) { return d.q32; }
    __device__ static float *d0(const DeviceBatch &d) { return d.d32[0]; }
    __device__ static float *d1(const DeviceBatch &d) { return d.d32[1]; }
};
template <> struct Vecs<double>
{
    __device__ static const double *W(const DeviceBatch &d) { return d.W64; }
    __device__ static const double *P(const DeviceBatch &d) { return d.P64; }
    __device__ static double *r(const DeviceBatch &d) { return d.r64; }
    __device__ static double *z(const DeviceBatch &d) { return d.z64; }
    __device__ static double *q(const DeviceBatch &d) { return d.q64; }
    __device__ static double *d0(const DeviceBatch &d) { return d.d64[0]; }
    __device__ static double *d1(const DeviceBatch &d) { return d.d64[1]; }
};

// Sums of components comp[] of the previous launch's partials of a cloth
// (slot 0: written by k_cg_update, 1: by k_cg_spmv), in every wavefront.
// Large cloths read their group sums, which k_cg_group formed after that
// launch; the others read their per-slice partials.  si is the slice's
// entry of d.slice_info, so the loads need nothing but the slice id and go
// out together with the state's.
template <int N>
__device__ inline void solver_sums(const DeviceBatch &d, int4 si, int slot, const int (&comp)[N],
                                   double (&out)[N])
{
    if (si.z < 0)
        wave_sets_sum<N>(glob(d.cg_group) + (size_t)slot * PARTS * d.n_groups, si.y, -si.z, comp,
                         out);
    else
        wave_sets_sum<N>(glob(d.cg_partial) + (size_t)slot * PARTS * d.n_slices, si.y, si.z, comp,
                         out);
}

// solver_sums in two phases: sums_issue puts the loads in flight (cloths of
// up to 64 * E slices or groups: one round, issued before the kernel waits
// on the cloth's state, so that the two round trips overlap), sums_finish
// adds them, or sums the larger cloths' sets the usual way.  E costs 2 N E
// registers while the loads are in flight: the variants for launches of up
// to 2,048 slices take 17 (1,088 sets), the SpMV variants for whole batches,
// whose occupancy matters, none, and the update for whole batches, which
// needs few registers, 5.
template <int N, int E> struct EarlySums
{
    double a[E > 0 ? E : 1][N];
    int nb;
    bool one;
};

template <int N, int E>
__device__ inline void sums_issue(const DeviceBatch &d, int4 si, int slot, const int (&comp)[N],
                                  EarlySums<N, E> &e)
{
    const bool grp = si.z < 0;
    e.nb = grp ? -si.z : si.z;
    e.one = E > 0 && e.nb <= 64 * E;
    if constexpr (E > 0)
        if (e.one)
        {
            const size_t n = grp ? (size_t)d.n_groups : (size_t)d.n_slices; // sets per slot
            const G<double> *sets =
                (grp ? glob(d.cg_group) : glob(d.cg_partial)) + (size_t)slot * PARTS * n;
            sets_load<N, E>(sets, si.y, e.nb, comp, e.a);
        }
}

template <int N, int E>
__device__ inline void sums_finish(const DeviceBatch &d, int4 si, int slot, const int (&comp)[N],
                                   const EarlySums<N, E> &e, double (&out)[N])
{
    if constexpr (E > 0)
        if (e.one)
        {
            sets_add<N, E>(e.a, e.nb, out);
            return;
        }
    solver_sums<N>(d, si, slot, comp, out);
}

// Loads of data that one launch reads once (matrix planes, column indices),
// with the nontemporal hint when NT: a matrix larger than the Infinity Cache
// then streams past the caches instead of evicting the direction vectors
// that the gathers and the staging read many times.  A matrix that fits the
// Infinity Cache is read from it by the next launch without the hint, so the
// hint is one more autotuned dimension of the SpMV.
template <bool NT, typename V> __device__ inline V stream_load(const G<V> *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// q = W d over the rows of one slice for lane phase j: W of type TW (the
// solver's matrix, or W64 in the mixed form after a stall), d = z + beta
// d_prev formed on the fly in T, sums in TW.
// The slots go in chunks of U per lane, software-pipelined: the column
// indices of the next chunk are loaded after the gathers and blocks of the
// current one, so each chunk waits on one memory round trip (gathers and
// blocks together) rather than two (indices, then gathers), and the
// indices' loads, issued last, stay in flight across the chunk's arithmetic.
template <typename TW, typename T, int L, bool NT, int UW = 1>
__device__ inline void slice_product(const G<TW> *__restrict__ Wl, const G<int> *__restrict__ cl,
                                     int w, int j, const G<T> *__restrict__ z,
                                     const G<T> *__restrict__ dold, T bt, bool first, TW (&y)[3])
{
    // slots per lane and chunk: UW for the FP64 matrix under FP32 vectors
    constexpr int U = sizeof(TW) == sizeof(T) ? 4 / L : UW;
    constexpr int STEP = U * L;
    y[0] = y[1] = y[2] = 0;
    int c[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        c[u] = j + u * L < w ? stream_load<NT>(cl + (j + u * L) * SLICE) : -1;
    for (int k0 = j; k0 < w; k0 += STEP)
    {
        T zv[U][3], dv[U][3];
        TW m[U][9];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (c[u] >= 0)
            {
                const size_t c3 = 3 * (size_t)c[u];
                for (int a = 0; a < 3; a++)
                {
                    zv[u][a] = z[c3 + a];
                    dv[u][a] = dold[c3 + a];
                }
                const G<TW> *wk = Wl + 9 * (size_t)(k0 + u * L) * SLICE;
                for (int e = 0; e < 9; e++)
                    m[u][e] = stream_load<NT>(wk + e * SLICE);
            }
        int cn[U];
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const int k = k0 + STEP + u * L;
            cn[u] = k < w ? stream_load<NT>(cl + k * SLICE) : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            if (c[u] >= 0)
            {
                const TW v0 = (TW)dir_value<T>(zv[u][0], dv[u][0], bt, first);
                const TW v1 = (TW)dir_value<T>(zv[u][1], dv[u][1], bt, first);
                const TW v2 = (TW)dir_value<T>(zv[u][2], dv[u][2], bt, first);
                y[0] += block_row<TW>(m[u][0], m[u][1], m[u][2], v0, v1, v2);
                y[1] += block_row<TW>(m[u][3], m[u][4], m[u][5], v0, v1, v2);
                y[2] += block_row<TW>(m[u][6], m[u][7], m[u][8], v0, v1, v2);
            }
            c[u] = cn[u];
        }
    }
    for (int a = 0; a < 3; a++)
        y[a] = phase_sum<L>(y[a]);
}

// column indices of the gathered product's first chunk (slots j, j + L, ...,
// U of them), loaded before the prologue: the chain column -> gather is the
// product's critical path, the blocks go out with the gathers.  The loads
// do not wait for the row's width: slots past it read the next rows' (or
// the padding after the last slice: U L <= 8 slots of COL_PAD), and
// gathered_mask drops them once the width is known.
template <int L, int U, bool NT>
__device__ inline void gathered_columns(const G<int> *__restrict__ cl, int j, int (&c)[U])
{
#pragma unroll
    for (int u = 0; u < U; u++)
        c[u] = stream_load<NT>(cl + (j + u * L) * SLICE);
}

template <int L, int U> __device__ inline void gathered_mask(int j, int w, int (&c)[U])
{
#pragma unroll
    for (int u = 0; u < U; u++)
        if (j + u * L >= w)
            c[u] = -1;
}

// slice_product with the column indices of its first chunk already loaded,
// in chunks of U slots per lane: the same arithmetic in the same order for
// any U (a lane adds its slots j, j + L, ... in turn), so U trades registers
// for memory round trips (a row of w slots waits on ceil(w / (U L)) of them)
template <typename T, int L, int U, bool NT>
__device__ inline void slice_product_from(const G<T> *__restrict__ Wl, const G<int> *__restrict__ cl,
                                          int w, int j, const G<T> *__restrict__ z,
                                          const G<T> *__restrict__ dold, T bt, bool first,
                                          int (&c)[U], T (&y)[3])
{
    constexpr int STEP = U * L;
    y[0] = y[1] = y[2] = 0;
    for (int k0 = j; k0 < w; k0 += STEP)
    {
        T zv[U][3], dv[U][3];
        T m[U][9];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (c[u] >= 0)
            {
                const size_t c3 = 3 * (size_t)c[u];
                for (int a = 0; a < 3; a++)
                {
                    zv[u][a] = z[c3 + a];
                    dv[u][a] = dold[c3 + a];
                }
                const G<T> *wk = Wl + 9 * (size_t)(k0 + u * L) * SLICE;
                for (int e = 0; e < 9; e++)
                    m[u][e] = stream_load<NT>(wk + e * SLICE);
            }
        int cn[U];
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const int k = k0 + STEP + u * L;
            cn[u] = k < w ? stream_load<NT>(cl + k * SLICE) : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            if (c[u] >= 0)
            {
                const T v0 = dir_value<T>(zv[u][0], dv[u][0], bt, first);
                const T v1 = dir_value<T>(zv[u][1], dv[u][1], bt, first);
                const T v2 = dir_value<T>(zv[u][2], dv[u][2], bt, first);
                y[0] += block_row<T>(m[u][0], m[u][1], m[u][2], v0, v1, v2);
                y[1] += block_row<T>(m[u][3], m[u][4], m[u][5], v0, v1, v2);
                y[2] += block_row<T>(m[u][6], m[u][7], m[u][8], v0, v1, v2);
            }
            c[u] = cn[u];
        }
    }
    for (int a = 0; a < 3; a++)
        y[a] = phase_sum<L>(y[a]);
}

// The staged variant: the workgroup first forms d = z + beta d_prev once for
// every distinct column of its slice (the slice's halo list, ascending rows,
// so neighbouring lanes load neighbouring rows) into LDS, 4 values per
// column; the slot loop then reads each column's d from LDS through its
// 16-bit position in the list.  A column shared by several rows of the
// slice is loaded from memory once instead of once per row, and the loop's
// only memory loads are the matrix planes and the 2-byte positions, which
// do not wait on one another.
// Staging in rounds of R = 8/L entries per lane (512 per workgroup; 4/L
// measured 1% slower on the 1.5M-point cloth, 2/L 7% slower): the first
// round's column rows are loaded at the very start of the kernel and its z
// and d_prev right after the state load (none of them depends on beta), so
// that they are in flight during the prologue; each later round loads its
// rows, then all of its
// gathers, then writes, so a round costs two memory round trips rather than
// two per entry.
template <typename T, int L> struct StageRound
{
    static constexpr int R = 8 / L;
    int row[R];
    T zv[R][3], dv[R][3];
};

template <typename T, int L>
__device__ inline void stage_round_rows(const G<int> *__restrict__ hr, int i0, int H,
                                        StageRound<T, L> &b)
{
    constexpr int R = StageRound<T, L>::R;
#pragma unroll
    for (int u = 0; u < R; u++)
    {
        const int i = i0 + u * L * 64;
        b.row[u] = i < H ? hr[i] : -1;
    }
}

template <typename T, int L>
__device__ inline void stage_round_gather(const G<T> *__restrict__ z,
                                          const G<T> *__restrict__ dold, StageRound<T, L> &b)
{
    constexpr int R = StageRound<T, L>::R;
#pragma unroll
    for (int u = 0; u < R; u++)
        if (b.row[u] >= 0)
        {
            const size_t c3 = 3 * (size_t)b.row[u];
            for (int a = 0; a < 3; a++)
            {
                b.zv[u][a] = z[c3 + a];
                b.dv[u][a] = dold[c3 + a];
            }
        }
}

template <typename T, int L>
__device__ inline void stage_round_store(int i0, T bt, bool first, const StageRound<T, L> &b,
                                         T *vs)
{
    constexpr int R = StageRound<T, L>::R;
#pragma unroll
    for (int u = 0; u < R; u++)
        if (b.row[u] >= 0)
        {
            const int i = i0 + u * L * 64;
            for (int a = 0; a < 3; a++)
                vs[4 * i + a] = dir_value<T>(b.zv[u][a], b.dv[u][a], bt, first);
        }
}

// the slice's column list: offset and length
__device__ inline int2 stage_list(const DeviceBatch &d, int sl)
{
    const int h0 = glob(d.halo_off)[sl];
    return make_int2(h0, glob(d.halo_off)[sl + 1] - h0);
}

// the rest of the staging after the first round (b, loaded early), and the
// barrier after which the slot loop reads the LDS
template <typename T, int L>
__device__ inline void stage_directions(const DeviceBatch &d, int2 list, const G<T> *__restrict__ z,
                                        const G<T> *__restrict__ dold, T bt, bool first,
                                        const StageRound<T, L> &b, T *vs)
{
    constexpr int PER = StageRound<T, L>::R * L * 64; // entries per round
    stage_round_store<T, L>(threadIdx.x, bt, first, b, vs);
    const G<int> *hr = glob(d.halo_row) + list.x;
    for (int i0 = threadIdx.x + PER; i0 - (int)threadIdx.x < list.y; i0 += PER)
    {
        StageRound<T, L> nb;
        stage_round_rows<T, L>(hr, i0, list.y, nb);
        stage_round_gather<T, L>(z, dold, nb);
        stage_round_store<T, L>(i0, bt, first, nb, vs);
    }
    __syncthreads();
}

// blocks and positions of the chunk of slots k0, k0 + L, ... (U of them)
template <typename TW, int L, int U, bool NT>
__device__ inline void staged_chunk(const G<TW> *__restrict__ Wl,
                                    const G<uint16_t> *__restrict__ lc, int k0, int w,
                                    TW (&m)[U][9], int (&c)[U])
{
#pragma unroll
    for (int u = 0; u < U; u++)
    {
        const int k = k0 + u * L;
        c[u] = k < w ? (int)stream_load<NT>(lc + k * SLICE) : -1;
        if (k < w)
            for (int e = 0; e < 9; e++)
                m[u][e] = stream_load<NT>(Wl + 9 * (size_t)k * SLICE + e * SLICE);
    }
}

// the staged product given its first chunk (m, c), which the caller loaded
// early so that it is in flight while the directions are staged
template <typename TW, typename T, int L, int U, bool NT>
__device__ inline void staged_product_from(const DeviceBatch &d, int sl,
                                           const G<TW> *__restrict__ Wl,
                                           const G<uint16_t> *__restrict__ lc, int w, int j,
                                           const G<T> *__restrict__ z,
                                           const G<T> *__restrict__ dold, T bt, bool first,
                                           T *vs, const StageRound<T, L> &sb, int2 list,
                                           TW (&m)[U][9], int (&c)[U], TW (&y)[3])
{
    constexpr int STEP = U * L;
    y[0] = y[1] = y[2] = 0;
    stage_directions<T, L>(d, list, z, dold, bt, first, sb, vs);
    for (int k0 = j; k0 < w; k0 += STEP)
    {
#pragma unroll
        for (int u = 0; u < U; u++)
            if (c[u] >= 0)
            {
                const int cc = c[u];
                const TW v0 = (TW)vs[4 * cc], v1 = (TW)vs[4 * cc + 1], v2 = (TW)vs[4 * cc + 2];
                y[0] += block_row<TW>(m[u][0], m[u][1], m[u][2], v0, v1, v2);
                y[1] += block_row<TW>(m[u][3], m[u][4], m[u][5], v0, v1, v2);
                y[2] += block_row<TW>(m[u][6], m[u][7], m[u][8], v0, v1, v2);
            }
        staged_chunk<TW, L, U, NT>(Wl, lc, k0 + STEP, w, m, c);
    }
    for (int a = 0; a < 3; a++)
        y[a] = phase_sum<L>(y[a]);
}

// k_cg_spmv on slice sl of batch d; S: the staged variant (vs: its LDS);
// UG: slots per lane and chunk of the gathered product
template <typename T, int L, bool S, bool NT, int UG = 4 / L>
__device__ inline void spmv_body(const DeviceBatch &d, const int sl, const int4 si,
                                 const G<T> *__restrict__ W,
                                 const G<T> *__restrict__ P, G<T> *__restrict__ r,
                                 G<T> *__restrict__ z, G<T> *__restrict__ q,
                                 G<T> *__restrict__ dbuf0, G<T> *__restrict__ dbuf1, double *lds,
                                 T *vs)
{
    const int lane = threadIdx.x & 63;
    constexpr int RPW = SLICE / L;                                  // rows per wavefront
    const int row = (threadIdx.x >> 6) * RPW + (lane & (RPW - 1)); // row within the slice
    const int j = lane / RPW;                                        // lane phase
    const int p = sl * SLICE + row;
    const size_t p3 = 3 * (size_t)p;
    // Loads in dependent rounds, as few as the data allow: (1) the slice's
    // info (si: the caller's load, from the launch list's entry or from
    // slice_info); (2) the row's width, the first chunk's column indices
    // (staged: the positions and blocks and the first staging round's
    // rows), the cloth's parameters and state and, in the variants for
    // launches of up to 2,048 slices, its solver sums and the row's own
    // vectors; (3) the direction gathers with their blocks.  A cloth solved
    // alone is bound by these round trips, not by bandwidth.
    const int w = glob(d.row_w)[p];
    const int inst = si.x;
    const size_t base = (size_t)si.w;
    // the variants for launches of up to 2,048 slices (a cloth solved
    // alone: latency-bound) also load the row's own z and d_prev (from both
    // direction buffers: the state that picks one is not loaded yet) now
    // rather than after the product, and the solver sums with the state
    constexpr bool SMALL = !S && UG * L == 8;
    T zo[3], dpo[2][3];
    if constexpr (SMALL)
        if (j == 0)
            for (int a = 0; a < 3; a++)
            {
                zo[a] = z[p3 + a];
                dpo[0][a] = dbuf0[p3 + a];
                dpo[1][a] = dbuf1[p3 + a];
            }
    constexpr int U = S ? 4 / L : UG;
    T m0[U][9];
    int c0[U];
    StageRound<T, L> sb;
    int2 list = make_int2(0, 0);
    if constexpr (S)
    {
        staged_chunk<T, L, U, NT>(W + 9 * base + row, glob(d.lcol) + base + row, j, w, m0, c0);
        // the first staging round's column rows; its gathers of z and
        // d_prev follow the state load, which selects d_prev's buffer
        list = stage_list(d, sl);
        stage_round_rows<T, L>(glob(d.halo_row) + list.x, threadIdx.x, list.y, sb);
    }
    else
        gathered_columns<L, U, NT>(glob(d.col) + base + row, j, c0);
    // the solver sums of the previous update (not used by the first launch
    // of a solve, whose partials are stale but allocated), in flight with
    // the state (b.b only after a true residual, below)
    const int sum_comp[2] = {P_RZ, P_RR};
    constexpr int E = SMALL ? 17 : 0;
    EarlySums<2, E> es;
    sums_issue<2, E>(d, si, 0, sum_comp, es);
    const InstParams ip = glob(d.params)[inst];
    const bool lead = sl == ip.slice_begin && threadIdx.x == 0;
    CgState st = load_state(glob(d.cg) + 2 * inst);
    G<CgState> *st_out = glob(d.cg) + 2 * inst + 1;
    double sums[2];
    sums_finish<2, E>(d, si, 0, sum_comp, es, sums);
    if (st.done)
    {
        if (lead)
            store_state(st_out, st);
        return;
    }
    if constexpr (S)
        stage_round_gather<T, L>(z, (st.iter & 1) ? dbuf0 : dbuf1, sb);
    G<double> *part_out = glob(d.cg_partial) + (size_t)PARTS * d.n_slices;
    const double tol2 = ip.tol * ip.tol;
    bool residual = false;
    double beta = 0.0;
    if (st.start)
        residual = true; // r0 = b - W dx0 (BaraffCloth.cpp:515)
    else
    {
        const double *v = sums;
        const double rz = v[0], rr = v[1];
        if (st.pending)
        {
            // the previous k_cg_spmv computed the true residual
            double bb[1];
            const int bb_comp[1] = {P_BB};
            solver_sums<1>(d, si, 0, bb_comp, bb);
            st.pending = 0;
            st.rr_true = rr;
            st.bb = bb[0];
            if (rr <= tol2 * bb[0] || st.final_check || st.total >= ip.max_iter || !isfinite(rr))
            {
                st.done = 1;
                if (lead)
                    store_state(st_out, st);
                return;
            }
            // FP32 matrix: cond(W) times the FP32 rounding of W bounds
            // what the FP32-matrix Krylov process can resolve; when a check
            // forced by max_since finds the true residual stalled, the rest
            // of the solve reads W64 (direction restarted), which converges
            // like FP64
            if (ip.stall > 0.0 && !st.w64 && st.since >= ip.max_since &&
                rr > ip.stall * st.rr_ref)
            {
                st.w64 = 1;
                st.first = 1;
            }
            st.rr_ref = rr;
            st.since = 0;
        }
        else
        {
            st.rr = rr;
            if (st.force_check || st.total >= ip.max_iter || !isfinite(rr) ||
                (st.since > 0 && (rr <= tol2 * st.bb || rr <= ip.rel2 * st.rr_ref)) ||
                st.since >= ip.max_since)
                residual = true;
        }
        if (!residual)
        {
            beta = st.first ? 0.0 : rz / st.rho;
            st.rho = rz;
        }
    }
    const G<int> *cl = glob(d.col) + base + row;
    if (residual)
    {
        // true residual in FP64: r = b - W64 dx (dx + dx32 in the mixed form)
        const G<double> *Wl = glob(d.W64) + 9 * base + row;
        const G<double> *dx = glob(d.dx);
        const G<float> *dx32 = glob(d.dx32);
        double y[3] = {0.0, 0.0, 0.0};
        for (int k = j; k < w; k += L)
        {
            const size_t c3 = 3 * (size_t)cl[k * SLICE];
            const G<double> *wk = Wl + 9 * (size_t)k * SLICE;
            double x0 = dx[c3], x1 = dx[c3 + 1], x2 = dx[c3 + 2];
            if constexpr (sizeof(T) == sizeof(float))
            {
                x0 += (double)dx32[c3];
                x1 += (double)dx32[c3 + 1];
                x2 += (double)dx32[c3 + 2];
            }
            for (int a = 0; a < 3; a++)
                y[a] += block_row<double>(wk[(3 * a) * SLICE], wk[(3 * a + 1) * SLICE],
                                          wk[(3 * a + 2) * SLICE], x0, x1, x2);
        }
        for (int a = 0; a < 3; a++)
            y[a] = phase_sum<L>(y[a]);
        double v[3] = {0.0, 0.0, 0.0};
        if (j == 0)
            for (int a = 0; a < 3; a++)
            {
                const double bv = glob(d.b)[p3 + a];
                const double rv = bv - y[a];
                const T rt = (T)rv;
                const T zt = P[p3 + a] * rt;
                r[p3 + a] = rt;
                z[p3 + a] = zt;
                v[0] = fma_t((double)rt, (double)zt, v[0]);
                v[1] = fma_t(rv, rv, v[1]);
                v[2] = fma_t(bv, bv, v[2]);
            }
        block_sum<L, 3>(v, lds);
        const int comp[3] = {P_RZ, P_RR, P_BB};
        write_partials<3>(part_out, sl, comp, v);
        st.pending = 1;
        st.force_check = 0;
        st.updates += 1;
        if (st.start)
        {
            st.start = 0;
            st.first = 1;
        }
        if (lead)
            store_state(st_out, st);
        return;
    }
    const bool first = st.first != 0;
    const T bt = (T)beta;
    const G<T> *dold = (st.iter & 1) ? dbuf0 : dbuf1;
    G<T> *dnew = (st.iter & 1) ? dbuf1 : dbuf0;
    T y0, y1, y2;
    bool wide = false; // the FP64 matrix under FP32 vectors
    if constexpr (sizeof(T) == sizeof(float))
        wide = st.w64 != 0;
    // staged variant: a slice whose column list exceeds the LDS provided
    // (uniform over the workgroup) takes the gathered product
    bool staged = false;
    if constexpr (S)
        staged = list.y <= d.halo_lds;
    if (staged)
    {
        const G<uint16_t> *lc = glob(d.lcol) + base + row;
        if (wide)
        {
            double y[3], mw[1][9];
            int cw[1];
            const G<double> *Wl = glob(d.W64) + 9 * base + row;
            staged_chunk<double, L, 1, NT>(Wl, lc, j, w, mw, cw);
            staged_product_from<double, T, L, 1, NT>(d, sl, Wl, lc, w, j, z, dold, bt, first, vs,
                                                     sb, list, mw, cw, y);
            y0 = (T)y[0];
            y1 = (T)y[1];
            y2 = (T)y[2];
        }
        else
        {
            T y[3];
            staged_product_from<T, T, L, U, NT>(d, sl, W + 9 * base + row, lc, w, j, z, dold, bt,
                                                first, vs, sb, list, m0, c0, y);
            y0 = y[0];
            y1 = y[1];
            y2 = y[2];
        }
    }
    else if (wide)
    {
        double y[3];
        // the FP64 matrix in chunks of 4 / L slots per lane in the variants
        // for launches of up to 2,048 slices (a stiff cloth solved alone
        // switches to it: each chunk is a round trip), else 1 to keep the
        // registers of the variants for whole batches
        slice_product<double, T, L, NT, SMALL ? 4 / L : 1>(glob(d.W64) + 9 * base + row, cl, w, j,
                                                            z, dold, bt, first, y);
        y0 = (T)y[0];
        y1 = (T)y[1];
        y2 = (T)y[2];
    }
    else
    {
        T y[3];
        // the staged variant's slices with lists too long for its LDS take
        // the gathered product; their c0 holds list positions, not columns
        if constexpr (S)
            slice_product<T, T, L, NT>(W + 9 * base + row, cl, w, j, z, dold, bt, first, y);
        else
        {
            gathered_mask<L, U>(j, w, c0);
            slice_product_from<T, L, U, NT>(W + 9 * base + row, cl, w, j, z, dold, bt, first, c0,
                                            y);
        }
        y0 = y[0];
        y1 = y[1];
        y2 = y[2];
    }
    double v[1] = {0.0};
    if (j == 0)
    {
        T dn[3];
        for (int a = 0; a < 3; a++)
            if constexpr (SMALL)
                dn[a] = dir_value<T>(zo[a], (st.iter & 1) ? dpo[0][a] : dpo[1][a], bt, first);
            else
                dn[a] = dir_value<T>(z[p3 + a], dold[p3 + a], bt, first);
        dnew[p3] = dn[0];
        dnew[p3 + 1] = dn[1];
        dnew[p3 + 2] = dn[2];
        q[p3] = y0;
        q[p3 + 1] = y1;
        q[p3 + 2] = y2;
        v[0] = block_row<double>((double)dn[0], (double)dn[1], (double)dn[2], (double)y0,
                                 (double)y1, (double)y2);
    }
    block_sum<L, 1>(v, lds);
    const int comp[1] = {P_DQ};
    write_partials<1>(part_out, sl, comp, v);
    if (lead)
        store_state(st_out, st);
}

// The row vectors an update reads: loaded before the prologue by the
// prefetching variant (PF: both direction buffers, since the state that
// picks one is not known yet; 12 more bytes per row, for small grids that
// latency bounds), else after it.
template <typename T> struct UpdateRow
{
    T r[3], q[3], P[3], d0[3], d1[3];
    float x32[3];
    double x64[3];
};

template <typename T>
__device__ inline void load_update_row(const DeviceBatch &d, size_t p3, const G<T> *P,
                                       const G<T> *r, const G<T> *q, const G<T> *dbuf0,
                                       const G<T> *dbuf1, int which, UpdateRow<T> &u)
{
    for (int a = 0; a < 3; a++)
    {
        u.r[a] = r[p3 + a];
        u.q[a] = q[p3 + a];
        u.P[a] = P[p3 + a];
        if (which != 1)
            u.d0[a] = dbuf0[p3 + a];
        if (which != 0)
            u.d1[a] = dbuf1[p3 + a];
        if constexpr (sizeof(T) == sizeof(float))
            u.x32[a] = glob(d.dx32)[p3 + a];
        else
            u.x64[a] = glob(d.dx)[p3 + a];
    }
}

// k_cg_update on solver workgroup blk (SB slices) of batch d
template <typename T, int SB, bool PF>
__device__ inline void update_body(const DeviceBatch &d, const int blk, const int4 si,
                                   const G<T> *__restrict__ P, G<T> *__restrict__ r,
                                   G<T> *__restrict__ z, const G<T> *__restrict__ q,
                                   const G<T> *__restrict__ dbuf0, const G<T> *__restrict__ dbuf1)
{
    // the SB slices of a workgroup belong to one cloth: si is its first
    // slice's slice_info entry
    const int inst = si.x;
    const int lane = threadIdx.x & 63;
    const int sl = blk * SB + (threadIdx.x >> 6);
    const size_t p3 = 3 * (size_t)(sl * SLICE + lane);
    // d.Wd of the preceding k_cg_spmv, in flight with the state (unused when
    // the cloth is done or that launch formed a true residual)
    const int dq_comp[1] = {P_DQ};
    constexpr int E = PF ? 17 : 5;
    EarlySums<1, E> es;
    sums_issue<1, E>(d, si, 1, dq_comp, es);
    UpdateRow<T> row;
    if constexpr (PF)
        load_update_row<T>(d, p3, P, r, q, dbuf0, dbuf1, 2, row);
    const InstParams ip = glob(d.params)[inst];
    const bool lead = blk == ip.blk_begin && threadIdx.x == 0;
    CgState st = load_state(glob(d.cg) + 2 * inst + 1);
    G<CgState> *st_out = glob(d.cg) + 2 * inst;
    const G<double> *part_in = glob(d.cg_partial) + (size_t)PARTS * d.n_slices;
    G<double> *part_out = glob(d.cg_partial);
    const size_t part0 = (size_t)PARTS * SB * blk; // this workgroup's slices
    double v1[1];
    sums_finish<1, E>(d, si, 1, dq_comp, es, v1);
    if (st.done)
    {
        if (lead)
            store_state(st_out, st);
        return;
    }
    if (st.pending)
    {
        // idle: hand the true-residual partials on to the next k_cg_spmv,
        // and fold the FP32 increments into dx (no kernel reads them now)
        if (threadIdx.x < PARTS * SB)
            part_out[part0 + threadIdx.x] = part_in[part0 + threadIdx.x];
        if constexpr (sizeof(T) == sizeof(float))
        {
            const size_t q3 = 3 * (size_t)((blk * SB + (threadIdx.x >> 6)) * SLICE + (threadIdx.x & 63));
            G<double> *dx = glob(d.dx);
            G<float> *dx32 = glob(d.dx32);
            for (int a = 0; a < 3; a++)
            {
                dx[q3 + a] = dx[q3 + a] + (double)dx32[q3 + a];
                dx32[q3 + a] = 0.0f;
            }
        }
        if (lead)
            store_state(st_out, st);
        return;
    }
    const double dq = v1[0];
    if (!(dq > 0.0) || !isfinite(dq))
    {
        // d.Wd == 0 (BaraffCloth.cpp:547-551: the reference stops before
        // updating x) or a loss of definiteness: stop after a final true
        // residual, which the next k_cg_spmv computes
        st.final_check = 1;
        st.force_check = 1;
        st.breakdown = dq == 0.0 ? 0 : 1;
        if (threadIdx.x < PARTS * SB)
            part_out[part0 + threadIdx.x] = 0.0;
        if (lead)
            store_state(st_out, st);
        return;
    }
    const double alpha = st.rho / dq;
    const int which = (st.iter & 1) ? 1 : 0; // the direction buffer k_cg_spmv wrote
    if constexpr (!PF)
        load_update_row<T>(d, p3, P, r, q, dbuf0, dbuf1, which, row);
    T dn[3];
    for (int a = 0; a < 3; a++)
        dn[a] = which ? row.d1[a] : row.d0[a];
    G<double> *dx = glob(d.dx);
    G<float> *dx32 = glob(d.dx32);
    double v[2] = {0.0, 0.0};
    for (int a = 0; a < 3; a++)
    {
        if constexpr (sizeof(T) == sizeof(float))
            dx32[p3 + a] = row.x32[a] + (float)(alpha * (double)dn[a]);
        else
            dx[p3 + a] = fma_t(alpha, (double)dn[a], row.x64[a]);
        const T rn = (T)fma_t(-alpha, (double)row.q[a], (double)row.r[a]);
        const T zn = row.P[a] * rn;
        r[p3 + a] = rn;
        z[p3 + a] = zn;
        v[0] = fma_t((double)rn, (double)zn, v[0]);
        v[1] = fma_t((double)rn, (double)rn, v[1]);
    }
    // one partial set per slice (wavefront)
    v[0] = wave_sum(v[0]);
    v[1] = wave_sum(v[1]);
    if (lane == 0)
    {
        part_out[(size_t)PARTS * sl + P_RZ] = v[0];
        part_out[(size_t)PARTS * sl + P_RR] = v[1];
    }
    st.iter += 1;
    st.total += 1;
    st.since += 1;
    st.first = 0;
    if (lead)
        store_state(st_out, st);
}

// the staged variant's LDS: 4 values per halo entry (HALO_CAP at most),
// sized at launch
extern __shared__ __align__(16) unsigned char stage_lds[];

// A padding entry of a launch list.  Padding has every field negative (the
// count INT_MIN); the test reads every field of the entry, so that all of
// its loads go out before the branch, in one round, rather than the rest
// of the entry after it.
__device__ inline bool entry_padding(const SolverEntry &e)
{
    return (e.id | e.info.x | e.info.y | e.info.w) < 0 || e.info.z == INT_MIN;
}

template <typename T, int L, bool S, bool NT, int UG = 4 / L>
__global__ void __launch_bounds__(L * 64)
k_cg_spmv(DeviceBatch d, const T *__restrict__ W, const T *__restrict__ P, T *__restrict__ r,
          T *__restrict__ z, T *__restrict__ q, T *__restrict__ dbuf0, T *__restrict__ dbuf1)
{
    __shared__ double lds[8 * L];
    // one slice per workgroup, from the list of active slices when given,
    // else from the contiguous range at slice_first
    int sl = solver_block(d.xcd_map);
    int4 si;
    if (d.spmv_list)
    {
        const SolverEntry e = d.spmv_list[sl];
        if (entry_padding(e))
            return;
        sl = e.id;
        si = e.info;
    }
    else
    {
        sl += d.slice_first;
        si = glob(d.slice_info)[sl];
    }
    spmv_body<T, L, S, NT, UG>(d, sl, si, glob(W), glob(P), glob(r), glob(z), glob(q),
                               glob(dbuf0), glob(dbuf1), lds, reinterpret_cast<T *>(stage_lds));
}

template <typename T, int SB, bool PF>
__global__ void __launch_bounds__(SB * 64)
k_cg_update(DeviceBatch d, const T *__restrict__ P, T *__restrict__ r, T *__restrict__ z,
            const T *__restrict__ q, const T *__restrict__ dbuf0, const T *__restrict__ dbuf1)
{
    int blk = solver_block(d.xcd_map);
    int4 si;
    if (d.upd_list)
    {
        const SolverEntry e = d.upd_list[blk];
        if (entry_padding(e))
            return;
        blk = e.id;
        si = e.info;
    }
    else
    {
        blk += d.slice_first / SB;
        si = glob(d.slice_info)[blk * SB];
    }
    update_body<T, SB, PF>(d, blk, si, glob(P), glob(r), glob(z), glob(q), glob(dbuf0),
                           glob(dbuf1));
}

// group sums of the partials in `slot` (just written by the launch before):
// one wavefront per group of GROUP consecutive slices of a large cloth,
// lane l taking slice l of the group, then the wavefront butterfly
__device__ inline void group_body(const DeviceBatch &d, int g, int slot)
{
    const int inst = glob(d.group_inst)[g];
    const InstParams ip = glob(d.params)[inst];
    const int k = g - ip.group_begin;
    const int s0 = ip.slice_begin + k * GROUP;
    const int n = min(GROUP, ip.slice_count - k * GROUP);
    const int lane = threadIdx.x & 63;
    const G<double> *part = glob(d.cg_partial) + (size_t)slot * PARTS * d.n_slices;
    double v[PARTS];
    for (int j = 0; j < PARTS; j++)
        v[j] = lane < n ? part[PARTS * (size_t)(s0 + lane) + j] : 0.0;
    for (int j = 0; j < PARTS; j++)
        v[j] = wave_sum(v[j]);
    G<double> *out = glob(d.cg_group) + (size_t)slot * PARTS * d.n_groups + PARTS * (size_t)g;
    if (lane == 0)
        for (int j = 0; j < PARTS; j++)
            out[j] = v[j];
}

__global__ void __launch_bounds__(64) k_cg_group(DeviceBatch d, int slot)
{
    group_body(d, blockIdx.x, slot);
}

// Many batches in one launch: entry (batch, slice) of `list` per workgroup,
// batches in the device table `table`; entries with batch < 0 are padding.
template <typename T, int L, bool S, bool NT>
__global__ void __launch_bounds__(L * 64)
k_cg_spmv_multi(const DeviceBatch *__restrict__ table, const int2 *__restrict__ list, int xcd_map)
{
    __shared__ double lds[8 * L];
    const int2 e = list[solver_block(xcd_map)];
    if (e.x < 0)
        return;
    const DeviceBatch &d = table[e.x];
    spmv_body<T, L, S, NT>(d, e.y, glob(d.slice_info)[e.y], glob(Vecs<T>::W(d)), glob(Vecs<T>::P(d)), glob(Vecs<T>::r(d)),
                       glob(Vecs<T>::z(d)), glob(Vecs<T>::q(d)), glob(Vecs<T>::d0(d)),
                       glob(Vecs<T>::d1(d)), lds, reinterpret_cast<T *>(stage_lds));
}

template <typename T, int SB, bool PF>
__global__ void __launch_bounds__(SB * 64)
k_cg_update_multi(const DeviceBatch *__restrict__ table, const int2 *__restrict__ list,
                  int xcd_map)
{
    const int2 e = list[solver_block(xcd_map)];
    if (e.x < 0)
        return;
    const DeviceBatch &d = table[e.x];
    update_body<T, SB, PF>(d, e.y, glob(d.slice_info)[e.y * SB], glob(Vecs<T>::P(d)), glob(Vecs<T>::r(d)), glob(Vecs<T>::z(d)),
                       glob(Vecs<T>::q(d)), glob(Vecs<T>::d0(d)), glob(Vecs<T>::d1(d)));
}

__global__ void __launch_bounds__(64)
k_cg_group_multi(const DeviceBatch *__restrict__ table, const int2 *__restrict__ list, int slot)
{
    const int2 e = list[blockIdx.x];
    if (e.x < 0)
        return;
    group_body(table[e.x], e.y, slot);
}

// Solver states for a new solve, one thread per cloth: both slots to the
// solve start (true residual of the warm start first).  A cloth whose last
// params[i].carry solves in a row moved to the FP64 matrix starts on it: it
// stalls step after step, and finding that out again costs every solve up
// to max_since iterations; one whose solves stall now and then returns to
// the FP32 matrix.  fresh[i] (set by the host when the cloth's state or pins
// are set from outside) starts it on the FP32 matrix with no history, as a
// new cloth.
__global__ void k_cg_reset(CgState *st, unsigned char *fresh, const InstParams *params, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const CgState &a = st[2 * i], &b = st[2 * i + 1];
    const int stalls = fresh[i] ? 0 : (a.w64 | b.w64) ? max(a.stalls, b.stalls) + 1 : 0;
    fresh[i] = 0;
    CgState z = {};
    z.start = 1;
    z.first = 1;
    z.stalls = stalls;
    z.w64 = stalls >= params[i].carry ? 1 : 0;
    st[2 * i] = z;
    st[2 * i + 1] = z;
}

// timing set-up: every cloth mid-solve, at a k_cg_spmv with residual 1
// Does nothing: launched back to back on a stream, each launch waits for the
// previous one, so the time per launch is what dispatch, completion and the
// kernel boundary's cache maintenance cost a grid of that shape, the floor of
// any dependent launch of it.
__global__ void k_empty() {}

template <typename T> __global__ void k_cg_timing_setup(DeviceBatch d, T *r, T *z)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < 3 * (size_t)d.n_rows)
    {
        r[i] = (T)1;
        z[i] = (T)1;
    }
    if (i < (size_t)d.n_slices)
        for (int slot = 0; slot < 2; slot++)
        {
            double *part = d.cg_partial + (size_t)slot * PARTS * d.n_slices + PARTS * i;
            part[P_RZ] = 1.0;
            part[P_RR] = 1.0;
            part[P_BB] = 1.0;
            part[P_DQ] = 2.0;
        }
    if (i < (size_t)d.n_groups)
        for (int slot = 0; slot < 2; slot++)
        {
            double *grp = d.cg_group + (size_t)slot * PARTS * d.n_groups + PARTS * i;
            grp[P_RZ] = 1.0;
            grp[P_RR] = 1.0;
            grp[P_BB] = 1.0;
            grp[P_DQ] = 2.0;
        }
    if (i < 2 * (size_t)d.n_inst)
    {
        CgState s = {};
        s.rho = 1.0;
        d.cg[i] = s;
    }
}

constexpr int SB = BW_SB;

} // namespace

void launch_cg_reset(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_reset, dim3((d.n_inst + 63) / 64), dim3(64), 0, s, d.cg, d.cg_fresh,
                       d.params, d.n_inst);
}

// bytes of LDS the staged variant needs for lists of up to halo_max entries
template <typename T> size_t stage_bytes(int halo_max) { return 4 * sizeof(T) * (size_t)halo_max; }

// Launches of at most this many slices (two wavefronts per SIMD with 1 lane
// per row) run the gathered product in chunks of 8 / L slots per lane
// instead of 4 / L: more registers, which such a grid does not need for
// occupancy, and half the dependent gather rounds per row, which bound it.
// Same arithmetic in the same order.  FP64 launches of up to half as many
// (that variant holds one wavefront per SIMD).  BW_ILP_SLICES overrides.
static int ilp_slices()
{
    static const int v = [] {
        const char *e = std::getenv("BW_ILP_SLICES");
        return e ? std::max(0, std::atoi(e)) : 2048;
    }();
    return v;
}

template <typename T, int L, bool NT>
void launch_spmv_nt(const DeviceBatch &d, int slices, const T *W, const T *P, T *r, T *z, T *q,
                    T *d0, T *d1, hipStream_t s)
{
    if (!d.spmv_stage && slices <= (sizeof(T) == sizeof(float) ? ilp_slices() : ilp_slices() / 2))
        hipLaunchKernelGGL((k_cg_spmv<T, L, false, NT, 8 / L>), dim3(slices), dim3(L * 64),
                           d.lds_pad, s, d, W, P, r, z, q, d0, d1);
    else if (d.spmv_stage)
        hipLaunchKernelGGL((k_cg_spmv<T, L, true, NT>), dim3(slices), dim3(L * 64),
                           stage_bytes<T>(d.halo_lds), s, d, W, P, r, z, q, d0, d1);
    else
        hipLaunchKernelGGL((k_cg_spmv<T, L, false, NT>), dim3(slices), dim3(L * 64), d.lds_pad, s,
                           d, W, P, r, z, q, d0, d1);
}

template <typename T, int L>
void launch_spmv(const DeviceBatch &d, int slices, const T *W, const T *P, T *r, T *z, T *q,
                 T *d0, T *d1, hipStream_t s)
{
    if (d.spmv_nt)
        launch_spmv_nt<T, L, true>(d, slices, W, P, r, z, q, d0, d1, s);
    else
        launch_spmv_nt<T, L, false>(d, slices, W, P, r, z, q, d0, d1, s);
}

template <typename T>
void launch_spmv_lpr(const DeviceBatch &d, int slices, const T *W, const T *P, T *r, T *z, T *q,
                     T *d0, T *d1, hipStream_t s)
{
    if (d.spmv_lpr == 1)
        launch_spmv<T, 1>(d, slices, W, P, r, z, q, d0, d1, s);
    else if (d.spmv_lpr == 2)
        launch_spmv<T, 2>(d, slices, W, P, r, z, q, d0, d1, s);
    else
        launch_spmv<T, 4>(d, slices, W, P, r, z, q, d0, d1, s);
}

template <typename T, bool S, bool NT>
void launch_spmv_multi_v(const DeviceBatch *table, const int2 *list, int grid, int lpr,
                         int xcd_map, size_t lds, hipStream_t s)
{
    if (lpr == 1)
        hipLaunchKernelGGL((k_cg_spmv_multi<T, 1, S, NT>), dim3(grid), dim3(64), lds, s, table,
                           list, xcd_map);
    else if (lpr == 2)
        hipLaunchKernelGGL((k_cg_spmv_multi<T, 2, S, NT>), dim3(grid), dim3(128), lds, s, table,
                           list, xcd_map);
    else
        hipLaunchKernelGGL((k_cg_spmv_multi<T, 4, S, NT>), dim3(grid), dim3(256), lds, s, table,
                           list, xcd_map);
}

template <typename T>
void launch_spmv_multi(const DeviceBatch *table, const int2 *list, int grid, int lpr,
                       int halo_max, bool nt, int xcd_map, hipStream_t s)
{
    const size_t lds = stage_bytes<T>(halo_max);
    if (halo_max > 0 && nt)
        launch_spmv_multi_v<T, true, true>(table, list, grid, lpr, xcd_map, lds, s);
    else if (halo_max > 0)
        launch_spmv_multi_v<T, true, false>(table, list, grid, lpr, xcd_map, lds, s);
    else if (nt)
        launch_spmv_multi_v<T, false, true>(table, list, grid, lpr, xcd_map, 0, s);
    else
        launch_spmv_multi_v<T, false, false>(table, list, grid, lpr, xcd_map, 0, s);
}

void launch_group_sums(const DeviceBatch &d, int slot, hipStream_t s)
{
    if (d.n_groups > 0)
        hipLaunchKernelGGL(k_cg_group, dim3(d.n_groups), dim3(64), 0, s, d, slot);
}

void launch_cg_spmv_only(const DeviceBatch &d, bool fp32, hipStream_t s)
{
    if (fp32)
        launch_spmv_lpr<float>(d, d.n_slices, d.W32, d.P32, d.r32, d.z32, d.q32, d.d32[0],
                               d.d32[1], s);
    else
        launch_spmv_lpr<double>(d, d.n_slices, d.W64, d.P64, d.r64, d.z64, d.q64, d.d64[0],
                                d.d64[1], s);
}

// Grids of at most this many slices run the update kernel that loads its rows
// before its prologue (latency over bytes; BW_PF_SLICES overrides)
static int prefetch_slices()
{
    static const int v = [] {
        const char *e = std::getenv("BW_PF_SLICES");
        return e ? std::max(0, std::atoi(e)) : 2048;
    }();
    return v;
}

template <typename T>
void launch_update(const DeviceBatch &d, int slices, const T *P, T *r, T *z, const T *q,
                   const T *d0, const T *d1, hipStream_t s)
{
    const dim3 upd_grid((slices + SB - 1) / SB), upd_block(SB * 64);
    if (slices <= prefetch_slices())
        hipLaunchKernelGGL((k_cg_update<T, SB, true>), upd_grid, upd_block, d.lds_pad, s, d, P, r, z, q,
                           d0, d1);
    else
        hipLaunchKernelGGL((k_cg_update<T, SB, false>), upd_grid, upd_block, d.lds_pad, s, d, P, r, z, q,
                           d0, d1);
}

void launch_cg_step(const DeviceBatch &d, bool fp32, int phase, int slices, hipStream_t s)
{
    if (fp32)
    {
        if (phase == 0)
            launch_spmv_lpr<float>(d, slices, d.W32, d.P32, d.r32, d.z32, d.q32, d.d32[0], d.d32[1],
                                   s);
        else
            launch_update<float>(d, slices, d.P32, d.r32, d.z32, d.q32, d.d32[0], d.d32[1], s);
    }
    else
    {
        if (phase == 0)
            launch_spmv_lpr<double>(d, slices, d.W64, d.P64, d.r64, d.z64, d.q64, d.d64[0],
                                    d.d64[1], s);
        else
            launch_update<double>(d, slices, d.P64, d.r64, d.z64, d.q64, d.d64[0], d.d64[1], s);
    }
    if (d.n_groups > 0)
        hipLaunchKernelGGL(k_cg_group, dim3(d.n_groups), dim3(64), 0, s, d, phase == 0 ? 1 : 0);
}

void launch_cg_step_multi(const DeviceBatch *table, const int2 *spmv_list, int spmv_grid,
                          const int2 *upd_list, int upd_grid, const int2 *grp_list, int grp_grid,
                          bool fp32, int phase, int lpr, int halo_max, bool nt, int xcd_map,
                          hipStream_t s)
{
    if (fp32)
    {
        if (phase == 0)
            launch_spmv_multi<float>(table, spmv_list, spmv_grid, lpr, halo_max, nt, xcd_map, s);
        else if (upd_grid * SB <= prefetch_slices())
            hipLaunchKernelGGL((k_cg_update_multi<float, SB, true>), dim3(upd_grid), dim3(SB * 64),
                               0, s, table, upd_list, xcd_map);
        else
            hipLaunchKernelGGL((k_cg_update_multi<float, SB, false>), dim3(upd_grid),
                               dim3(SB * 64), 0, s, table, upd_list, xcd_map);
    }
    else
    {
        if (phase == 0)
            launch_spmv_multi<double>(table, spmv_list, spmv_grid, lpr, halo_max, nt, xcd_map, s);
        else if (upd_grid * SB <= prefetch_slices())
            hipLaunchKernelGGL((k_cg_update_multi<double, SB, true>), dim3(upd_grid),
                               dim3(SB * 64), 0, s, table, upd_list, xcd_map);
        else
            hipLaunchKernelGGL((k_cg_update_multi<double, SB, false>), dim3(upd_grid),
                               dim3(SB * 64), 0, s, table, upd_list, xcd_map);
    }
    if (grp_grid > 0)
        hipLaunchKernelGGL(k_cg_group_multi, dim3(grp_grid), dim3(64), 0, s, table, grp_list,
                           phase == 0 ? 1 : 0);
}

void launch_empty(int grid, int block, hipStream_t s)
{
    hipLaunchKernelGGL(k_empty, dim3(grid), dim3(block), 0, s);
}

void launch_cg_timing_setup(const DeviceBatch &d, bool fp32, hipStream_t s)
{
    const size_t n = std::max<size_t>(3 * (size_t)d.n_rows, 2 * (size_t)d.n_inst);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(k_cg_timing_setup<float>, grid, dim3(256), 0, s, d, d.r32, d.z32);
    else
        hipLaunchKernelGGL(k_cg_timing_setup<double>, grid, dim3(256), 0, s, d, d.r64, d.z64);
}

} // namespace bw
