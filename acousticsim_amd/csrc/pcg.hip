// Batched Jacobi-PCG on the sliced-ELL matrix (cBaraffCloth::Solve,
// BaraffCloth.cpp:507-611, restated for the device).
//
// Work decomposition: one workgroup = SB slices of 64 rows (one wavefront
// per slice, one lane per block row).  Each cloth owns whole workgroups, so
// workgroups never mix cloths.
//
// No kernel communicates across workgroups.  Every kernel writes one FP64
// partial sum per workgroup for each dot product it forms (a fixed shuffle
// tree inside the wavefront, then a fixed order across the SB wavefronts);
// the NEXT kernel starts by summing the cloth's partials in a fixed order,
// redundantly in every workgroup of the cloth, and derives the scalars from
// them.  The kernel boundary makes the partials visible, there are no atomics
// and no fences, and a cloth's arithmetic does not depend on what else is in
// the batch.  Per-cloth fields of InstState are each written by one kernel
// and read only by later kernels (writes come from the cloth's first
// workgroup).
//
// Per PCG iteration k:
//   k_pcg_spmv(k)   prologue (k >= 1): rho_k = sum r.(P r), rr_k = sum r.r
//                   from the partials of k_pcg_update(k-1); beta = rho_k /
//                   rho_{k-1}; stop if converged.  Body: d_k = P r + beta
//                   d_{k-1} (formed on the fly for every gathered column,
//                   stored for the own row), q = W d_k, partial d_k.q.
//   k_pcg_update(k) prologue: alpha = rho_k / sum d.q (stop on d.q == 0, as
//                   the reference does, :547-551).  Body: x += alpha d,
//                   r -= alpha q, partials r.(P r) and r.r.
// Direction buffers alternate by iteration parity.
//
// This is the solver of "reference" solver mode (FP64, exactly max_iters-1
// iterations, the reference's own recurrences and break conditions).  The
// tolerance modes use the state machine of cg.hip, which runs the same
// recurrences with a tolerance stop and reliable updates.
#include "device.h"
#include "reduce.h"

namespace bw
{
namespace
{

template <typename T>
__device__ inline T dir_value(T P, T r, T dold, T beta, bool first)
{
    return first ? P * r : P * r + beta * dold;
}

template <typename T, int SB>
__global__ void __launch_bounds__(SB * 64)
k_pcg_spmv(DeviceBatch d, const T *__restrict__ W, const T *__restrict__ P,
           const T *__restrict__ r, T *__restrict__ dbuf0, T *__restrict__ dbuf1,
           T *__restrict__ q)
{
    __shared__ double lds[8 * SB];
    const int blk = solver_block(d.xcd_map);
    const int inst = d.blk_inst[blk];
    InstState *st = d.state + inst;
    if (st->done)
        return;
    const InstParams &ip = d.params[inst];
    const int k = st->it_a;
    double beta = 0.0;
    if (k > 0)
    {
        double v[2];
        const int comp[2] = {P_RZ, P_RR};
        reduce_cloth<SB, 2>(d.partial, ip.blk_begin, ip.blk_count, comp, v, lds);
        const double rho = v[0], rr = v[1];
        beta = rho / st->rho[(k - 1) & 1];
        const bool done = k >= ip.max_iter || beta == 0.0; // :529, :585-590
        if (blk == ip.blk_begin && threadIdx.x == 0)
        {
            st->rho[k & 1] = rho;
            st->rr = rr;
            st->it_b = k;
            if (done)
                st->done = 1;
        }
        if (done)
            return;
    }
    else if (blk == ip.blk_begin && threadIdx.x == 0)
        st->it_b = 0;
    const bool first = k == 0;
    const T bt = (T)beta;
    const T *dold = (k & 1) ? dbuf0 : dbuf1;
    T *dnew = (k & 1) ? dbuf1 : dbuf0;
    const int lane = threadIdx.x & 63;
    const int s = blk * SB + (threadIdx.x >> 6);
    const int p = s * SLICE + lane;
    const size_t p3 = 3 * (size_t)p;
    T dn[3];
    for (int a = 0; a < 3; a++)
        dn[a] = dir_value<T>(P[p3 + a], r[p3 + a], dold[p3 + a], bt, first);
    const int w = d.row_w[p];
    const size_t base = (size_t)d.slice_off[s];
    const T *Wl = W + 9 * base + lane;
    const int *cl = d.col + base + lane;
    T y0 = 0, y1 = 0, y2 = 0;
#pragma unroll 2
    for (int kk = 0; kk < w; kk++)
    {
        const int c = cl[kk * SLICE];
        const size_t c3 = 3 * (size_t)c;
        const T v0 = dir_value<T>(P[c3], r[c3], dold[c3], bt, first);
        const T v1 = dir_value<T>(P[c3 + 1], r[c3 + 1], dold[c3 + 1], bt, first);
        const T v2 = dir_value<T>(P[c3 + 2], r[c3 + 2], dold[c3 + 2], bt, first);
        const T *wk = Wl + 9 * (size_t)kk * SLICE;
        y0 += wk[0 * SLICE] * v0 + wk[1 * SLICE] * v1 + wk[2 * SLICE] * v2;
        y1 += wk[3 * SLICE] * v0 + wk[4 * SLICE] * v1 + wk[5 * SLICE] * v2;
        y2 += wk[6 * SLICE] * v0 + wk[7 * SLICE] * v1 + wk[8 * SLICE] * v2;
    }
    dnew[p3] = dn[0];
    dnew[p3 + 1] = dn[1];
    dnew[p3 + 2] = dn[2];
    q[p3] = y0;
    q[p3 + 1] = y1;
    q[p3 + 2] = y2;
    double v[1] = {(double)dn[0] * (double)y0 + (double)dn[1] * (double)y1 +
                   (double)dn[2] * (double)y2};
    block_sum<SB, 1>(v, lds);
    const int comp[1] = {P_DQ};
    write_partials<1>(d.partial, blk, comp, v);
}

template <typename T, int SB>
__global__ void __launch_bounds__(SB * 64)
k_pcg_update(DeviceBatch d, const T *__restrict__ P, T *__restrict__ x, T *__restrict__ r,
             const T *__restrict__ dbuf0, const T *__restrict__ dbuf1, const T *__restrict__ q)
{
    __shared__ double lds[8 * SB];
    const int blk = solver_block(d.xcd_map);
    const int inst = d.blk_inst[blk];
    InstState *st = d.state + inst;
    if (st->done)
        return;
    const InstParams &ip = d.params[inst];
    const int k = st->it_b;
    double v[1];
    {
        const int comp[1] = {P_DQ};
        reduce_cloth<SB, 1>(d.partial, ip.blk_begin, ip.blk_count, comp, v, lds);
    }
    const double dq = v[0];
    const bool lead = blk == ip.blk_begin && threadIdx.x == 0;
    if (dq == 0.0 || !isfinite(dq))
    {
        if (lead)
        {
            st->done = 1;
            st->breakdown = dq == 0.0 ? 0 : 1;
        }
        return;
    }
    const double alpha = st->rho[k & 1] / dq;
    if (lead)
    {
        st->it_a = k + 1;
        st->total_iter += 1;
    }
    const T *dn = (k & 1) ? dbuf1 : dbuf0;
    const int s = blk * SB + (threadIdx.x >> 6);
    const size_t p3 = 3 * (size_t)(s * SLICE + (threadIdx.x & 63));
    double vals[2] = {0.0, 0.0};
    for (int a = 0; a < 3; a++)
    {
        const T xn = (T)((double)x[p3 + a] + alpha * (double)dn[p3 + a]);
        const T rn = (T)((double)r[p3 + a] - alpha * (double)q[p3 + a]);
        x[p3 + a] = xn;
        r[p3 + a] = rn;
        const T z = P[p3 + a] * rn;
        vals[0] += (double)rn * (double)z;
        vals[1] += (double)rn * (double)rn;
    }
    block_sum<SB, 2>(vals, lds);
    const int comp[2] = {P_RZ, P_RR};
    write_partials<2>(d.partial, blk, comp, vals);
}

// r64 = b - W64 dx.  mode 0: solver start (r64 kept, partials rz, rr, bb);
// mode 2: residual for the step counters only.
template <int SB> __global__ void __launch_bounds__(SB * 64) k_residual64(DeviceBatch d, int mode)
{
    __shared__ double lds[8 * SB];
    const int blk = solver_block(d.xcd_map);
    const int lane = threadIdx.x & 63;
    const int s = blk * SB + (threadIdx.x >> 6);
    const int p = s * SLICE + lane;
    const size_t p3 = 3 * (size_t)p;
    const int w = d.row_w[p];
    const size_t base = (size_t)d.slice_off[s];
    const double *Wl = d.W64 + 9 * base + lane;
    const int *cl = d.col + base + lane;
    double y[3] = {0.0, 0.0, 0.0};
    for (int kk = 0; kk < w; kk++)
    {
        const size_t c3 = 3 * (size_t)cl[kk * SLICE];
        const double *wk = Wl + 9 * (size_t)kk * SLICE;
        const double x0 = d.dx[c3], x1 = d.dx[c3 + 1], x2 = d.dx[c3 + 2];
        for (int a = 0; a < 3; a++)
            y[a] += wk[(3 * a) * SLICE] * x0 + wk[(3 * a + 1) * SLICE] * x1 +
                    wk[(3 * a + 2) * SLICE] * x2;
    }
    double v[4] = {0.0, 0.0, 0.0, 0.0}; // rz, rr, bb, rr64
    for (int a = 0; a < 3; a++)
    {
        const double bv = d.b[p3 + a];
        const double rv = bv - y[a];
        v[2] += bv * bv;
        v[3] += rv * rv;
        if (mode == 0)
        {
            d.r64[p3 + a] = rv;
            v[0] += rv * (d.P64[p3 + a] * rv);
            v[1] += rv * rv;
        }
    }
    block_sum<SB, 4>(v, lds);
    const int comp[4] = {P_RZ, P_RR, P_BB, P_RR64};
    write_partials<4>(d.partial, blk, comp, v);
}

// One workgroup per cloth: reduce the residual partials.
//   mode 0 (solve start): rho_0; done at once when max_iters <= 1
//   mode 2 (counters):    record ||r64|| and ||b||
template <int SB> __global__ void __launch_bounds__(SB * 64) k_decide(DeviceBatch d, int mode)
{
    __shared__ double lds[8 * SB];
    const int inst = blockIdx.x;
    InstState *st = d.state + inst;
    const InstParams &ip = d.params[inst];
    double v[4];
    const int comp[4] = {P_RZ, P_RR, P_BB, P_RR64};
    reduce_cloth<SB, 4>(d.partial, ip.blk_begin, ip.blk_count, comp, v, lds);
    if (threadIdx.x != 0)
        return;
    st->bb = v[2];
    st->rr64 = v[3];
    if (mode == 2)
        return;
    st->it_a = 0;
    st->it_b = 0;
    st->rho[0] = v[0];
    st->rr = v[1];
    st->done = ip.max_iter <= 0;
}

// x_pre <- x ; x <- x_pre + dx (BaraffCloth.cpp:150-153)
__global__ void __launch_bounds__(256) k_update_positions(DeviceBatch d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n_rows || d.row_vertex[p] < 0)
        return;
    const size_t p3 = 3 * (size_t)p;
    for (int a = 0; a < 3; a++)
    {
        const double x = d.x[p3 + a];
        d.xpre[p3 + a] = x;
        d.x[p3 + a] = x + d.dx[p3 + a];
    }
}

// per-step solver state reset (a kernel rather than a memset, so that it may
// sit inside a captured graph)
__global__ void k_reset_states(InstState *st, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    InstState z = {};
    st[i] = z;
}

template <typename T> __global__ void k_fill(T *v, size_t n, T value)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n)
        v[i] = value;
}

constexpr int SB = BW_SB;

} // namespace

void launch_reset_states(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_states, dim3((d.n_inst + 63) / 64), dim3(64), 0, s, d.state,
                       d.n_inst);
}

void launch_pcg_init64(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_residual64<SB>, dim3(d.n_blocks_wg), dim3(SB * 64), 0, s, d, 0);
    hipLaunchKernelGGL(k_decide<SB>, dim3(d.n_inst), dim3(SB * 64), 0, s, d, 0);
}

void launch_pcg_iter64(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL((k_pcg_spmv<double, SB>), dim3(d.n_blocks_wg), dim3(SB * 64), 0, s, d,
                       d.W64, d.P64, d.r64, d.d64[0], d.d64[1], d.q64);
    hipLaunchKernelGGL((k_pcg_update<double, SB>), dim3(d.n_blocks_wg), dim3(SB * 64), 0, s, d,
                       d.P64, d.dx, d.r64, d.d64[0], d.d64[1], d.q64);
}

void launch_residual64(const DeviceBatch &d, int mode, hipStream_t s)
{
    hipLaunchKernelGGL(k_residual64<SB>, dim3(d.n_blocks_wg), dim3(SB * 64), 0, s, d, mode);
    hipLaunchKernelGGL(k_decide<SB>, dim3(d.n_inst), dim3(SB * 64), 0, s, d, mode);
}

void launch_update_positions(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_positions, dim3((d.n_rows + 255) / 256), dim3(256), 0, s, d);
}

void launch_fill_residual(const DeviceBatch &d, bool, hipStream_t s)
{
    const size_t n = 3 * (size_t)d.n_rows;
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_fill<double>, grid, dim3(256), 0, s, d.r64, n, 1.0);
}

void launch_spmv_only(const DeviceBatch &d, bool, hipStream_t s)
{
    hipLaunchKernelGGL((k_pcg_spmv<double, SB>), dim3(d.n_blocks_wg), dim3(SB * 64), 0, s, d,
                       d.W64, d.P64, d.r64, d.d64[0], d.d64[1], d.q64);
}

} // namespace bw
