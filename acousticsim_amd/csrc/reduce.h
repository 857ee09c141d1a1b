// Deterministic workgroup reductions shared by the solver kernels.
//
// A workgroup is SB wavefronts of 64 lanes.  Sums run in a fixed order: a
// butterfly inside each wavefront, then wavefront 0..SB-1 in order.  A
// cloth's partials (one per workgroup) are summed by every workgroup of the
// cloth in the same fixed order, so all of them obtain identical scalars.
#pragma once
#include "device.h"

namespace bw
{

// partial-sum components (PARTS per workgroup)
enum Part
{
    P_DQ = 0,   // d.Wd  (PCG) / w.u (single-reduction CG)
    P_RZ = 1,   // r.(P r)
    P_RR = 2,   // r.r in the solver's precision
    P_BB = 3,   // b.b
    P_RR64 = 4, // FP64 true residual r64.r64
};

// v of the lane that DPP control CTRL selects (every lane active); a
// double moves as its two halves
template <int CTRL> __device__ inline double dpp_lane(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// v of lane ^ 32 (M = 32) or lane ^ 16 (M = 16) through gfx950's lane swap
// of two registers (halves of the wavefront, or pairs of 16-lane rows).
// Given the same register twice, one of the two results holds the partner
// lane's value in every lane; which one is read off the same swap of the
// lane index.
template <int M> __device__ inline double swap_lane(double v)
{
    const int lane = threadIdx.x & 63;
    const int lo = __double2loint(v), hi = __double2hiint(v);
    if constexpr (M == 32)
    {
        const auto li = __builtin_amdgcn_permlane32_swap(lane, lane, false, false);
        const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        return (int)li[0] == (lane ^ 32) ? __hiloint2double(h[0], l[0])
                                         : __hiloint2double(h[1], l[1]);
    }
    else
    {
        const auto li = __builtin_amdgcn_permlane16_swap(lane, lane, false, false);
        const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        return (int)li[0] == (lane ^ 16) ? __hiloint2double(h[0], l[0])
                                         : __hiloint2double(h[1], l[1]);
    }
}

// Sum over the 64 lanes of a wavefront (all active), the same value in every
// lane: the butterfly v += v[lane ^ m] for m = 32, 16, 8, 4, 2, 1, with the
// partner values moved by VALU lane operations rather than LDS permutes
// (the sums sit on every solver launch's dependent chain): the lane swaps
// for 32 and 16, a rotation by 8 inside each 16-lane row, the mirror inside
// each 8 lanes followed by the reversal inside each 4 for 4, and quad
// permutations for 2 and 1.
__device__ inline double wave_sum(double v)
{
    v += swap_lane<32>(v);
    v += swap_lane<16>(v);
    v += dpp_lane<0x128>(v);                 // row_ror:8
    v += dpp_lane<0x1B>(dpp_lane<0x141>(v)); // row_half_mirror, then quad_perm [3,2,1,0]
    v += dpp_lane<0x4E>(v);                  // quad_perm [2,3,0,1]
    v += dpp_lane<0xB1>(v);                  // quad_perm [1,0,3,2]
    return v;
}

// block-wide sums of N values (N <= 8), result in every thread
template <int SB, int N> __device__ inline void block_sum(double (&v)[N], double *lds)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = 0; j < N; j++)
        v[j] = wave_sum(v[j]);
    if (SB == 1)
        return;
    __syncthreads();
    if (lane == 0)
        for (int j = 0; j < N; j++)
            lds[wave * 8 + j] = v[j];
    __syncthreads();
    for (int j = 0; j < N; j++)
    {
        double s = lds[j];
        for (int w = 1; w < SB; w++)
            s += lds[w * 8 + j];
        v[j] = s;
    }
    __syncthreads();
}

// store this workgroup's (block-summed) values as its partials
template <int N, typename PP>
__device__ inline void write_partials(PP part, int blk, const int (&comp)[N], const double (&v)[N])
{
    if (threadIdx.x == 0)
        for (int j = 0; j < N; j++)
            part[PARTS * (size_t)blk + comp[j]] = v[j];
}

// Sum of components comp[] over partial sets b0 .. b0+nb-1, computed by
// every wavefront on its own (identical in all of them, whatever the
// workgroup size): lane l adds sets l, l+64, l+128, ... in that order, then
// the wavefront butterfly.  Up to 64 * ONE sets (ONE = 17 per lane for one
// or two components: a cloth of 1,088 slices, 263 x 263 vertices) all of a
// lane's loads go out at once, so the sum costs one memory round trip; more
// sets go F at a time per lane (4 for three components, 8 for fewer), to
// bound the registers.  Both paths add in the same order.
template <int N, typename PP>
__device__ inline void wave_sets_sum(PP sets, int b0, int nb, const int (&comp)[N], double (&out)[N])
{
    constexpr int F = N >= 3 ? 4 : 8;
    constexpr int ONE = N >= 3 ? 8 : 17;
    const int lane = threadIdx.x & 63;
    for (int j = 0; j < N; j++)
        out[j] = 0.0;
    if (nb <= 64 * ONE)
    {
        double a[ONE][N];
#pragma unroll
        for (int u = 0; u < ONE; u++)
            if (lane + 64 * u < nb)
                for (int j = 0; j < N; j++)
                    a[u][j] = sets[PARTS * (size_t)(b0 + lane + 64 * u) + comp[j]];
#pragma unroll
        for (int u = 0; u < ONE; u++)
            if (lane + 64 * u < nb)
                for (int j = 0; j < N; j++)
                    out[j] += a[u][j];
    }
    else
    {
        int i = lane;
        for (; i + 64 * (F - 1) < nb; i += 64 * F)
        {
            double a[F][N];
            for (int u = 0; u < F; u++)
                for (int j = 0; j < N; j++)
                    a[u][j] = sets[PARTS * (size_t)(b0 + i + 64 * u) + comp[j]];
            for (int u = 0; u < F; u++)
                for (int j = 0; j < N; j++)
                    out[j] += a[u][j];
        }
        for (; i < nb; i += 64)
            for (int j = 0; j < N; j++)
                out[j] += sets[PARTS * (size_t)(b0 + i) + comp[j]];
    }
    for (int j = 0; j < N; j++)
        out[j] = wave_sum(out[j]);
}

// wave_sets_sum's one-round form for up to 64 * E sets (E <= 17) in two
// phases, so that a kernel can put the loads in flight before it waits on
// other loads (the cloth's state): sets_load issues them all (lane l: sets
// l, l + 64, ...; lanes past the end read the last set again, so no load
// sits behind a per-lane branch), sets_add adds them in wave_sets_sum's
// order, with the same result bit for bit.
template <int N, int E, typename PP>
__device__ inline void sets_load(PP sets, int b0, int nb, const int (&comp)[N], double (&a)[E][N])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < E; u++)
        if (64 * u < nb)
        {
            const int s = b0 + min(lane + 64 * u, nb - 1);
            for (int j = 0; j < N; j++)
                a[u][j] = sets[PARTS * (size_t)s + comp[j]];
        }
}

template <int N, int E>
__device__ inline void sets_add(const double (&a)[E][N], int nb, double (&out)[N])
{
    const int lane = threadIdx.x & 63;
    for (int j = 0; j < N; j++)
        out[j] = 0.0;
#pragma unroll
    for (int u = 0; u < E; u++)
        if (64 * u < nb)
        {
            const bool in = lane + 64 * u < nb;
            for (int j = 0; j < N; j++)
            {
                const double t = out[j] + a[u][j];
                out[j] = in ? t : out[j];
            }
        }
    for (int j = 0; j < N; j++)
        out[j] = wave_sum(out[j]);
}

// sum the cloth's partials (workgroups b0 .. b0+nb-1), result in every thread
template <int SB, int N, typename PP>
__device__ inline void reduce_cloth(PP part, int b0, int nb, const int (&comp)[N],
                                    double (&out)[N], double *lds)
{
    for (int j = 0; j < N; j++)
        out[j] = 0.0;
    for (int i = threadIdx.x; i < nb; i += SB * 64)
        for (int j = 0; j < N; j++)
            out[j] += part[PARTS * (size_t)(b0 + i) + comp[j]];
    block_sum<SB, N>(out, lds);
}

} // namespace bw
