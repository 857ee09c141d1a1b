// Batched Jacobi-PCG on the sliced-ELL matrix (cBaraffCloth::Solve,
// BaraffCloth.cpp:507-611, restated for the device).
//
// One wavefront per 64-row slice, one lane per block row.  Each cloth of the
// batch owns whole slices and its own scalars (InstState); converged cloths
// skip their slices.  Dot products are reduced deterministically: each slice
// reduces its 64 lanes in a fixed shuffle tree and writes one FP64 partial;
// the last slice of a cloth to arrive sums that cloth's partials in a fixed
// order and updates its scalars.  A cloth's results therefore do not depend
// on what else is in the batch.
//
// Per iteration two kernels run:
//   k_pcg_spmv   d = P r + beta d_prev (formed on the fly for every gathered
//                column, written for the own row), q = W d, partial d.q;
//                the last slice sets alpha = rho / (d.q)
//   k_pcg_update x += alpha d, r -= alpha q, partials r.(P r) and r.r; the
//                last slice sets beta, rho and the convergence flag
// The direction buffer alternates between two arrays by iteration parity.
//
// Precision: "fp64" runs everything in FP64.  "mixed" runs the inner PCG on
// the FP32 copy of W with FP32 vectors but FP64 dots and scalars, inside an
// FP64 residual refinement: r64 = b - W64 dx, inner solve W32 e = r64 to a
// relative reduction eta, dx += e, repeat until ||r64|| <= tol ||b||.
#include "device.h"

namespace bw
{
namespace
{

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// last-arrival protocol: returns true in the last slice of the cloth
__device__ inline bool last_arrival(InstState *st, int nslices)
{
    __threadfence();
    unsigned int prev = 0;
    if (threadIdx.x == 0)
        prev = atomicAdd(&st->count, 1u);
    prev = __shfl(prev, 0, 64);
    const bool last = prev == (unsigned int)(nslices - 1);
    if (last)
        __threadfence();
    return last;
}

__device__ inline double reduce_partials(const double *part, int comp, int sb, int n)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 64)
        acc += part[4 * (size_t)(sb + i) + comp];
    return wave_sum(acc);
}

template <typename T>
__device__ inline T dir_value(T P, T r, T dold, T beta, bool first)
{
    return first ? P * r : P * r + beta * dold;
}

template <typename T>
__global__ void __launch_bounds__(64)
k_pcg_spmv(DeviceBatch d, const T *__restrict__ W, const T *__restrict__ P,
           const T *__restrict__ r, T *__restrict__ dbuf0, T *__restrict__ dbuf1,
           T *__restrict__ q)
{
    const int s = blockIdx.x;
    const int inst = d.slice_inst[s];
    InstState *st = d.state + inst;
    if (st->done)
        return;
    const int it = st->iter;
    const bool first = it == 0;
    const T beta = (T)st->beta;
    const T *dold = (it & 1) ? dbuf1 : dbuf0;
    T *dnew = (it & 1) ? dbuf0 : dbuf1;
    const int lane = threadIdx.x;
    const int p = s * SLICE + lane;
    const size_t p3 = 3 * (size_t)p;
    T dn[3];
    for (int a = 0; a < 3; a++)
        dn[a] = dir_value<T>(P[p3 + a], r[p3 + a], dold[p3 + a], beta, first);
    const int w = d.row_w[p];
    const size_t base = (size_t)d.slice_off[s];
    const T *Wl = W + 9 * base + lane;
    const int *cl = d.col + base + lane;
    T y0 = 0, y1 = 0, y2 = 0;
    for (int k = 0; k < w; k++)
    {
        const int c = cl[k * SLICE];
        const size_t c3 = 3 * (size_t)c;
        const T v0 = dir_value<T>(P[c3], r[c3], dold[c3], beta, first);
        const T v1 = dir_value<T>(P[c3 + 1], r[c3 + 1], dold[c3 + 1], beta, first);
        const T v2 = dir_value<T>(P[c3 + 2], r[c3 + 2], dold[c3 + 2], beta, first);
        const T *wk = Wl + 9 * (size_t)k * SLICE;
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
    double part = (double)dn[0] * (double)y0 + (double)dn[1] * (double)y1 + (double)dn[2] * (double)y2;
    part = wave_sum(part);
    if (lane == 0)
        d.partial[4 * (size_t)s] = part;
    const InstParams &ip = d.params[inst];
    if (last_arrival(st, ip.slice_count))
    {
        const double dq = reduce_partials(d.partial, 0, ip.slice_begin, ip.slice_count);
        if (lane == 0)
        {
            st->dq = dq;
            if (dq == 0.0 || !isfinite(dq))
            {
                // the reference breaks out of its loop on d.Ad == 0 (:547-551)
                st->done = 1;
                st->breakdown = dq == 0.0 ? 0 : 1;
            }
            else
                st->alpha = st->rho / dq;
            st->count = 0;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(64)
k_pcg_update(DeviceBatch d, const T *__restrict__ P, T *__restrict__ x, T *__restrict__ r,
             const T *__restrict__ dbuf0, const T *__restrict__ dbuf1, const T *__restrict__ q)
{
    const int s = blockIdx.x;
    const int inst = d.slice_inst[s];
    InstState *st = d.state + inst;
    if (st->done)
        return;
    const int it = st->iter;
    const double alpha = st->alpha;
    const T *dn = (it & 1) ? dbuf0 : dbuf1;
    const int lane = threadIdx.x;
    const size_t p3 = 3 * (size_t)(s * SLICE + lane);
    double rz = 0.0, rr = 0.0;
    for (int a = 0; a < 3; a++)
    {
        const T xn = (T)((double)x[p3 + a] + alpha * (double)dn[p3 + a]);
        const T rn = (T)((double)r[p3 + a] - alpha * (double)q[p3 + a]);
        x[p3 + a] = xn;
        r[p3 + a] = rn;
        const T z = P[p3 + a] * rn;
        rz += (double)rn * (double)z;
        rr += (double)rn * (double)rn;
    }
    rz = wave_sum(rz);
    rr = wave_sum(rr);
    if (lane == 0)
    {
        d.partial[4 * (size_t)s + 1] = rz;
        d.partial[4 * (size_t)s + 2] = rr;
    }
    const InstParams &ip = d.params[inst];
    if (last_arrival(st, ip.slice_count))
    {
        const double rz_t = reduce_partials(d.partial, 1, ip.slice_begin, ip.slice_count);
        const double rr_t = reduce_partials(d.partial, 2, ip.slice_begin, ip.slice_count);
        if (lane == 0)
        {
            const int it1 = it + 1;
            st->iter = it1;
            st->total_iter += 1;
            const double beta = rz_t / st->rho;
            st->beta = beta;
            st->rho = rz_t;
            st->rr = rr_t;
            bool done;
            if (ip.ref_mode)
                done = it1 >= ip.max_iter || beta == 0.0; // :585-590
            else
                done = rr_t <= st->inner_tol2 || st->total_iter >= ip.max_iter || !isfinite(rr_t);
            if (done)
                st->done = 1;
            st->count = 0;
        }
    }
}

// r = b - W64 dx (the warm start), first direction data; FP64 mode
__global__ void __launch_bounds__(64) k_pcg_init64(DeviceBatch d)
{
    const int s = blockIdx.x;
    const int inst = d.slice_inst[s];
    InstState *st = d.state + inst;
    const int lane = threadIdx.x;
    const int p = s * SLICE + lane;
    const size_t p3 = 3 * (size_t)p;
    const int w = d.row_w[p];
    const size_t base = (size_t)d.slice_off[s];
    const double *Wl = d.W64 + 9 * base + lane;
    const int *cl = d.col + base + lane;
    double y[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < w; k++)
    {
        const size_t c3 = 3 * (size_t)cl[k * SLICE];
        const double *wk = Wl + 9 * (size_t)k * SLICE;
        for (int a = 0; a < 3; a++)
            y[a] += wk[(3 * a) * SLICE] * d.dx[c3] + wk[(3 * a + 1) * SLICE] * d.dx[c3 + 1] +
                    wk[(3 * a + 2) * SLICE] * d.dx[c3 + 2];
    }
    double rz = 0.0, rr = 0.0, bb = 0.0;
    for (int a = 0; a < 3; a++)
    {
        const double bv = d.b[p3 + a];
        const double rv = bv - y[a];
        d.r64[p3 + a] = rv;
        rz += rv * (d.P64[p3 + a] * rv);
        rr += rv * rv;
        bb += bv * bv;
    }
    rz = wave_sum(rz);
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    if (lane == 0)
    {
        d.partial[4 * (size_t)s + 1] = rz;
        d.partial[4 * (size_t)s + 2] = rr;
        d.partial[4 * (size_t)s + 3] = bb;
    }
    const InstParams &ip = d.params[inst];
    if (last_arrival(st, ip.slice_count))
    {
        const double rz_t = reduce_partials(d.partial, 1, ip.slice_begin, ip.slice_count);
        const double rr_t = reduce_partials(d.partial, 2, ip.slice_begin, ip.slice_count);
        const double bb_t = reduce_partials(d.partial, 3, ip.slice_begin, ip.slice_count);
        if (lane == 0)
        {
            st->rho = rz_t;
            st->rr = rr_t;
            st->bb = bb_t;
            st->iter = 0;
            st->beta = 0.0;
            st->inner_tol2 = ip.tol * ip.tol * bb_t;
            bool done = ip.ref_mode ? ip.max_iter <= 0 : rr_t <= st->inner_tol2;
            st->done = done ? 1 : 0;
            st->count = 0;
        }
    }
}

// FP64 residual r64 = b - W64 dx.  With start_inner it also decides the
// refinement: converged, or start an inner FP32 solve on r32 = float(r64).
__global__ void __launch_bounds__(64) k_residual64(DeviceBatch d, int start_inner)
{
    const int s = blockIdx.x;
    const int inst = d.slice_inst[s];
    InstState *st = d.state + inst;
    if (start_inner && st->outer_done)
        return;
    const int lane = threadIdx.x;
    const int p = s * SLICE + lane;
    const size_t p3 = 3 * (size_t)p;
    const int w = d.row_w[p];
    const size_t base = (size_t)d.slice_off[s];
    const double *Wl = d.W64 + 9 * base + lane;
    const int *cl = d.col + base + lane;
    double y[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < w; k++)
    {
        const size_t c3 = 3 * (size_t)cl[k * SLICE];
        const double *wk = Wl + 9 * (size_t)k * SLICE;
        for (int a = 0; a < 3; a++)
            y[a] += wk[(3 * a) * SLICE] * d.dx[c3] + wk[(3 * a + 1) * SLICE] * d.dx[c3 + 1] +
                    wk[(3 * a + 2) * SLICE] * d.dx[c3 + 2];
    }
    double rr64 = 0.0, bb = 0.0, rz = 0.0, rr32 = 0.0;
    for (int a = 0; a < 3; a++)
    {
        const double bv = d.b[p3 + a];
        const double rv = bv - y[a];
        rr64 += rv * rv;
        bb += bv * bv;
        if (start_inner)
        {
            const float rf = (float)rv;
            d.r32[p3 + a] = rf;
            d.e32[p3 + a] = 0.0f;
            const float z = d.P32[p3 + a] * rf;
            rz += (double)rf * (double)z;
            rr32 += (double)rf * (double)rf;
        }
    }
    rr64 = wave_sum(rr64);
    bb = wave_sum(bb);
    rz = wave_sum(rz);
    rr32 = wave_sum(rr32);
    if (lane == 0)
    {
        d.partial[4 * (size_t)s + 0] = rr64;
        d.partial[4 * (size_t)s + 1] = rz;
        d.partial[4 * (size_t)s + 2] = rr32;
        d.partial[4 * (size_t)s + 3] = bb;
    }
    const InstParams &ip = d.params[inst];
    if (last_arrival(st, ip.slice_count))
    {
        const double rr64_t = reduce_partials(d.partial, 0, ip.slice_begin, ip.slice_count);
        const double rz_t = reduce_partials(d.partial, 1, ip.slice_begin, ip.slice_count);
        const double rr32_t = reduce_partials(d.partial, 2, ip.slice_begin, ip.slice_count);
        const double bb_t = reduce_partials(d.partial, 3, ip.slice_begin, ip.slice_count);
        if (lane == 0)
        {
            st->rr64 = rr64_t;
            st->bb = bb_t;
            if (start_inner)
            {
                const double tol2 = ip.tol * ip.tol * bb_t;
                if (rr64_t <= tol2 || st->outer >= ip.max_outer || st->total_iter >= ip.max_iter ||
                    !isfinite(rr64_t))
                {
                    st->outer_done = 1;
                    st->done = 1;
                }
                else
                {
                    st->outer += 1;
                    st->iter = 0;
                    st->beta = 0.0;
                    st->rho = rz_t;
                    st->rr = rr32_t;
                    // inner target: the outer tolerance with a margin, but no
                    // deeper than FP32 arithmetic can carry
                    double eta = 0.5 * ip.tol * sqrt(bb_t / rr64_t);
                    if (eta < 2e-5)
                        eta = 2e-5;
                    st->inner_tol2 = eta * eta * rr32_t;
                    st->done = rr32_t <= st->inner_tol2 ? 1 : 0;
                }
            }
            st->count = 0;
        }
    }
}

// dx += e after an inner solve
__global__ void __launch_bounds__(256) k_correct(DeviceBatch d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n_rows)
        return;
    const int s = p / SLICE;
    if (d.state[d.slice_inst[s]].outer_done)
        return;
    const size_t p3 = 3 * (size_t)p;
    for (int a = 0; a < 3; a++)
        d.dx[p3 + a] = d.dx[p3 + a] + (double)d.e32[p3 + a];
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

} // namespace

void launch_fill_residual(const DeviceBatch &d, bool fp32, hipStream_t s)
{
    const size_t n = 3 * (size_t)d.n_rows;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(k_fill<float>, grid, dim3(256), 0, s, d.r32, n, 1.0f);
    else
        hipLaunchKernelGGL(k_fill<double>, grid, dim3(256), 0, s, d.r64, n, 1.0);
}

void launch_spmv_only(const DeviceBatch &d, bool fp32, hipStream_t s)
{
    if (fp32)
        hipLaunchKernelGGL(k_pcg_spmv<float>, dim3(d.n_slices), dim3(64), 0, s, d, d.W32, d.P32,
                           d.r32, d.d32[0], d.d32[1], d.q32);
    else
        hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(d.n_slices), dim3(64), 0, s, d, d.W64,
                           d.P64, d.r64, d.d64[0], d.d64[1], d.q64);
}

void launch_reset_states(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_states, dim3((d.n_inst + 63) / 64), dim3(64), 0, s, d.state,
                       d.n_inst);
}

void launch_pcg_init64(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_init64, dim3(d.n_slices), dim3(64), 0, s, d);
}

void launch_pcg_iter64(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(d.n_slices), dim3(64), 0, s, d, d.W64, d.P64,
                       d.r64, d.d64[0], d.d64[1], d.q64);
    hipLaunchKernelGGL(k_pcg_update<double>, dim3(d.n_slices), dim3(64), 0, s, d, d.P64, d.dx,
                       d.r64, d.d64[0], d.d64[1], d.q64);
}

void launch_pcg_iter32(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_spmv<float>, dim3(d.n_slices), dim3(64), 0, s, d, d.W32, d.P32,
                       d.r32, d.d32[0], d.d32[1], d.q32);
    hipLaunchKernelGGL(k_pcg_update<float>, dim3(d.n_slices), dim3(64), 0, s, d, d.P32, d.e32,
                       d.r32, d.d32[0], d.d32[1], d.q32);
}

void launch_residual64(const DeviceBatch &d, int start_inner, hipStream_t s)
{
    hipLaunchKernelGGL(k_residual64, dim3(d.n_slices), dim3(64), 0, s, d, start_inner);
}

void launch_correct(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_correct, dim3((d.n_rows + 255) / 256), dim3(256), 0, s, d);
}

void launch_update_positions(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_positions, dim3((d.n_rows + 255) / 256), dim3(256), 0, s, d);
}

} // namespace bw
