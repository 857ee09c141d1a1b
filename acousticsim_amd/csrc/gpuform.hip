// The reference's GPU form (cloth_type "fem_gpu", cBaraffClothGpu,
// BaraffClothGpu.cpp / .cu, PCGSolver.cu) as an A/B comparison mode on this
// library's layout: FP32 stretch and assembly, the constant quadratic-bending
// matrix, and a block-Jacobi PCG of exactly 100 iterations (absolute
// residual 1e-10 checked every 10), warm-started from the previous step's
// solution, with pinned vertices reset to their targets before each step.
//
// Per step:
//   k_gf_pins      pinned rows: x = x_pre = their FP32 target
//                  (UpdateConstraintVertexPos, BaraffClothGpu.cpp:404-432)
//   k_gf_tri       per triangle, FP32: stretch force (3 x 3: force on each
//                  corner) and Hessian (9 blocks) of the warp/weft and bias
//                  parts (UpdateStretch_K_fint, BaraffClothGpu.cu:126-190;
//                  the Dv-uses-|Fu| slip kept)
//   k_gf_rows      per block row, FP32: K = sum of the row's stretch blocks
//                  (contribution lists, fixed order) - qb * I3; W = (dt b -
//                  dt^2) K + m (1 + dt a) I on the diagonal; f = stretch +
//                  (-Kb x); rhs = dt^2 (g + f) + m (x - x_pre); the inverse
//                  of the diagonal block (cofactors, CudaMatrix.h:362-391)
//                  (.cu:228-300, AddBendingHessianAndIntForce :447-478)
//   k_gf_init      r0 = b - W x0 over all columns (the reference's pinned-
//                  column skip in Calc_r0_d0_r0Minvr0 is a no-op), pinned
//                  rows 0; partials r.(M^-1 r), r.r
//   k_gf_spmv/k_gf_update, 100 times: d = M^-1 r + beta d_prev formed on the
//                  fly, z = W d without pinned columns, alpha = r.M^-1 r /
//                  d.z, r -= alpha z, dx += alpha d
//   k_gf_finish    x_pre = x; x += dx
// Dot products are FP64 sums of per-slice partials in a fixed order (the
// reference adds FP32 atomics in arrival order).
#include "device.h"
#include "reduce.h"

namespace bw
{
namespace
{

// Fu, Fv of one part and the force / Hessian accumulation
// (CalcFuAndFv, CalcFint, CalcHessian, BaraffClothGpu.cu:7-105)
__device__ inline void gf_part(float Ku, float Kv, const float *cu, const float *cv,
                               const float3 *x, float area, float (&f)[9], float (&H)[81])
{
    float Fu[3], Fv[3];
    const float xs[3][3] = {{x[0].x, x[0].y, x[0].z}, {x[1].x, x[1].y, x[1].z},
                            {x[2].x, x[2].y, x[2].z}};
    for (int a = 0; a < 3; a++)
    {
        Fu[a] = cu[0] * xs[0][a] + cu[1] * xs[1][a] + cu[2] * xs[2][a];
        Fv[a] = cv[0] * xs[0][a] + cv[1] * xs[1][a] + cv[2] * xs[2][a];
    }
    const float nu = sqrtf(Fu[0] * Fu[0] + Fu[1] * Fu[1] + Fu[2] * Fu[2]);
    const float nv = sqrtf(Fv[0] * Fv[0] + Fv[1] * Fv[1] + Fv[2] * Fv[2]);
    float Un[3], Vn[3];
    for (int a = 0; a < 3; a++)
    {
        Un[a] = nu > 0 ? Fu[a] / nu : Fu[a];
        Vn[a] = nv > 0 ? Fv[a] / nv : Fv[a];
    }
    const float Cu = nu - 1, Cv = nv - 1;
    for (int a = 0; a < 3; a++)
    {
        const float ub = -area * Ku * Cu * Un[a], vb = -area * Kv * Cv * Vn[a];
        for (int i = 0; i < 3; i++)
            f[3 * i + a] += ub * cu[i] + vb * cv[i];
    }
    // Du, Dv; Dv scales by |Fu| as the reference does
    const float iu = 1.0f / nu;
    float Du[9], Dv[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
        {
            Du[3 * a + b] = iu * Un[a] * Un[b] + (a == b && nu > 1 ? 1 - iu : 0.0f);
            Dv[3 * a + b] = iu * Vn[a] * Vn[b] + (a == b && nu > 1 ? 1 - iu : 0.0f);
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
        {
            const float su = Ku * cu[i] * cu[j], sv = Kv * cv[i] * cv[j];
            for (int q = 0; q < 9; q++)
                H[9 * (3 * i + j) + q] += -area * (su * Du[q] + sv * Dv[q]);
        }
}

__global__ void __launch_bounds__(128) k_gf_tri(DeviceBatch d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.n_tri)
        return;
    const InstParams &ip = d.params[d.tri_inst[t]];
    float3 x[3];
    for (int k = 0; k < 3; k++)
    {
        const double *p = d.x + 3 * (size_t)d.tri_rows[3 * (size_t)t + k];
        x[k] = make_float3((float)p[0], (float)p[1], (float)p[2]);
    }
    const float *cf = d.tri_coef + 12 * (size_t)t;
    const float area = d.tri_area[t];
    float f[9] = {}, H[81] = {};
    gf_part((float)ip.k0, (float)ip.k1, cf, cf + 3, x, area, f, H);
    gf_part((float)ip.k2, (float)ip.k2, cf + 6, cf + 9, x, area, f, H);
    float *out = d.gf_tri + 90 * (size_t)t;
    for (int q = 0; q < 9; q++)
        out[q] = f[q];
    for (int q = 0; q < 81; q++)
        out[9 + q] = H[q];
}

// one lane per row, one wavefront per slice
__global__ void __launch_bounds__(64) k_gf_rows(DeviceBatch d)
{
    const int s = blockIdx.x, lane = threadIdx.x;
    const int p = s * SLICE + lane;
    const int vg = d.row_vertex[p];
    const size_t base = (size_t)d.slice_off[s];
    const int w = d.slice_w[s];
    float *Wl = d.W32 + 9 * base + lane;
    if (vg < 0)
    {
        // padding row: identity diagonal keeps the inverse finite
        for (int k = 0; k < w; k++)
            for (int q = 0; q < 9; q++)
                Wl[(9 * (size_t)k + q) * SLICE] = 0.0f;
        for (int q = 0; q < 9; q++)
            d.gf_pinv[9 * (size_t)p + q] = (q % 4 == 0) ? 1.0f : 0.0f;
        for (int a = 0; a < 3; a++)
            d.b[3 * (size_t)p + a] = 0.0;
        return;
    }
    const InstParams &ip = d.params[d.slice_inst[s]];
    const float dt = (float)ip.dt, alpha = ip.gf_a, beta = ip.gf_b;
    const float m = d.gf_mass[p];
    const float kscale = dt * beta - dt * dt;
    const int b0 = d.blk_ptr[vg], nb = d.blk_ptr[vg + 1] - b0;
    float fb[3] = {0, 0, 0};
    float diag[9] = {};
    for (int k = 0; k < w; k++)
    {
        const size_t e = base + (size_t)k * SLICE + lane;
        float K[9] = {};
        if (k < nb)
        {
            const int blk = b0 + k;
            for (int c = d.s_ptr[blk]; c < d.s_ptr[blk + 1]; c++)
            {
                const int code = d.s_list[c];
                const int t = code >> 4, i = (code >> 2) & 3, j = code & 3;
                const float *H = d.gf_tri + 90 * (size_t)t + 9 + 9 * (3 * i + j);
                for (int q = 0; q < 9; q++)
                    K[q] += H[q];
            }
            const float qb = d.gf_qb[e];
            K[0] -= qb;
            K[4] -= qb;
            K[8] -= qb;
            // bending force -Kb x
            const double *xc = d.x + 3 * (size_t)d.col[e];
            for (int a = 0; a < 3; a++)
                fb[a] += -1.0f * qb * (float)xc[a];
        }
        float Wb[9];
        for (int q = 0; q < 9; q++)
            Wb[q] = kscale * K[q];
        if (k < nb && d.col[e] == p)
        {
            const float md = m * (1 + dt * alpha);
            Wb[0] += md;
            Wb[4] += md;
            Wb[8] += md;
            for (int q = 0; q < 9; q++)
                diag[q] = Wb[q];
        }
        for (int q = 0; q < 9; q++)
            Wl[(9 * (size_t)k + q) * SLICE] = Wb[q];
    }
    // stretch force, gathered from the incident triangles in order
    float f[3] = {0, 0, 0};
    for (int c = d.vt_ptr[vg]; c < d.vt_ptr[vg + 1]; c++)
    {
        const int code = d.vt_list[c];
        const float *ft = d.gf_tri + 90 * (size_t)(code >> 2) + 3 * (code & 3);
        for (int a = 0; a < 3; a++)
            f[a] += ft[a];
    }
    for (int a = 0; a < 3; a++)
        f[a] += fb[a];
    const double *x = d.x + 3 * (size_t)p, *xp = d.xpre + 3 * (size_t)p;
    const float g[3] = {0.0f, d.gf_grav[p], 0.0f};
    for (int a = 0; a < 3; a++)
    {
        const float rhs = dt * dt * (g[a] + f[a]) + ((float)x[a] - (float)xp[a]) * m;
        d.b[3 * (size_t)p + a] = (double)rhs;
    }
    // block-Jacobi preconditioner: cofactor inverse of the diagonal block
    const float a = diag[0], b = diag[1], c = diag[2], dd = diag[3], e = diag[4], f2 = diag[5],
                g2 = diag[6], h = diag[7], i = diag[8];
    float inv[9] = {e * i - f2 * h, c * h - b * i, b * f2 - c * e,
                    f2 * g2 - dd * i, a * i - c * g2, c * dd - a * f2,
                    dd * h - e * g2, b * g2 - a * h, a * e - b * dd};
    const float pre = 1.0f / (a * (e * i - f2 * h) - b * (dd * i - f2 * g2) + c * (dd * h - e * g2));
    for (int q = 0; q < 9; q++)
        d.gf_pinv[9 * (size_t)p + q] = inv[q] * pre;
}

__device__ inline void mat3vec(const G<float> *M, const float (&v)[3], float (&o)[3])
{
    for (int a = 0; a < 3; a++)
        o[a] = M[3 * a] * v[0] + M[3 * a + 1] * v[1] + M[3 * a + 2] * v[2];
}

__device__ inline GfState load_gf(const GfState *p)
{
    GfState s;
    s.rmr = p->rmr;
    s.rr = p->rr;
    s.bb = p->bb;
    s.it = p->it;
    s.done = p->done;
    return s;
}

__device__ inline void store_gf(GfState *p, const GfState &s)
{
    p->rmr = s.rmr;
    p->rr = s.rr;
    p->bb = s.bb;
    p->it = s.it;
    p->done = s.done;
}

// group sums of a slot for large cloths are formed by k_cg_group between
// launches; the sums here read them when present
template <int N>
__device__ inline void gf_sums(const DeviceBatch &d, const InstParams &ip, int slot,
                               const int (&comp)[N], double (&out)[N])
{
    if (ip.group_count > 0)
        wave_sets_sum<N>(glob(d.cg_group) + (size_t)slot * PARTS * d.n_groups, ip.group_begin,
                         ip.group_count, comp, out);
    else
        wave_sets_sum<N>(glob(d.cg_partial) + (size_t)slot * PARTS * d.n_slices, ip.slice_begin,
                         ip.slice_count, comp, out);
}

__device__ inline void gf_write(double *part, int sl, int comp, double v)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0)
        part[PARTS * (size_t)sl + comp] = v;
}

__global__ void __launch_bounds__(64) k_gf_pins(DeviceBatch d)
{
    const int p = blockIdx.x * SLICE + threadIdx.x;
    if (p >= d.n_rows || !d.row_pinned[p])
        return;
    for (int a = 0; a < 3; a++)
    {
        const double t = (double)d.gf_target[3 * (size_t)p + a];
        d.x[3 * (size_t)p + a] = t;
        d.xpre[3 * (size_t)p + a] = t;
    }
}

// r0 = b - W dx0, pinned rows 0; partials r.(M^-1 r), r.r and b.b over
// unpinned rows into slot 0
__global__ void __launch_bounds__(64) k_gf_init(DeviceBatch d)
{
    const int s = blockIdx.x, lane = threadIdx.x, p = s * SLICE + lane;
    const int inst = d.slice_inst[s];
    const InstParams &ip = d.params[inst];
    const size_t base = (size_t)d.slice_off[s];
    const int w = d.row_w[p];
    float y[3] = {0, 0, 0};
    for (int k = 0; k < w; k++)
    {
        const size_t e = base + (size_t)k * SLICE + lane;
        const float *Wk = d.W32 + 9 * base + (9 * (size_t)k) * SLICE + lane;
        const float *xc = d.gf_dx + 3 * (size_t)d.col[e];
        for (int a = 0; a < 3; a++)
            y[a] += Wk[(3 * a) * SLICE] * xc[0] + Wk[(3 * a + 1) * SLICE] * xc[1] +
                    Wk[(3 * a + 2) * SLICE] * xc[2];
    }
    float r[3];
    for (int a = 0; a < 3; a++)
        r[a] = d.row_pinned[p] ? 0.0f : (float)d.b[3 * (size_t)p + a] - y[a];
    float mr[3];
    mat3vec(glob(d.gf_pinv) + 9 * (size_t)p, r, mr);
    for (int a = 0; a < 3; a++)
        d.r32[3 * (size_t)p + a] = r[a];
    double *part = d.cg_partial;
    gf_write(part, s, P_RZ, (double)r[0] * mr[0] + (double)r[1] * mr[1] + (double)r[2] * mr[2]);
    gf_write(part, s, P_RR, (double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2]);
    double bb = 0.0;
    if (!d.row_pinned[p])
        for (int a = 0; a < 3; a++)
            bb += d.b[3 * (size_t)p + a] * d.b[3 * (size_t)p + a];
    gf_write(part, s, P_BB, bb);
    if (s == ip.slice_begin && lane == 0)
    {
        GfState z = {};
        store_gf(d.gf_state + 2 * inst, z);
    }
}

// iteration it: prologue decides beta and the stop (slot 0 state and sums);
// body d = M^-1 r + beta d_prev, z = W d over unpinned columns, partial d.z
__global__ void __launch_bounds__(64) k_gf_spmv(DeviceBatch d)
{
    const int s = blockIdx.x, lane = threadIdx.x, p = s * SLICE + lane;
    const int inst = d.slice_inst[s];
    const InstParams &ip = d.params[inst];
    const bool lead = s == ip.slice_begin && lane == 0;
    GfState st = load_gf(d.gf_state + 2 * inst);
    GfState *out = d.gf_state + 2 * inst + 1;
    if (st.done)
    {
        if (lead)
            store_gf(out, st);
        return;
    }
    double v[3];
    const int comp[3] = {P_RZ, P_RR, P_BB};
    gf_sums<3>(d, ip, 0, comp, v);
    if (st.it == 0)
        st.bb = v[2];
    double beta = 0.0;
    if (st.it > 0)
    {
        const int j = st.it - 1; // the iteration just finished
        if ((j % 10 == 0 || j == 99) && sqrt(v[1]) < 1e-10)
            st.done = 1;
        if (st.it >= 100)
            st.done = 1;
        beta = st.rmr == 0.0 ? 0.0 : v[0] / st.rmr;
    }
    st.rmr = v[0];
    st.rr = v[1];
    if (st.done)
    {
        if (lead)
            store_gf(out, st);
        return;
    }
    const float bt = (float)beta;
    const float *dold = (st.it & 1) ? d.d32[0] : d.d32[1];
    float *dnew = (st.it & 1) ? d.d32[1] : d.d32[0];
    auto dir = [&](int row, float (&o)[3]) {
        float rr[3] = {d.r32[3 * (size_t)row], d.r32[3 * (size_t)row + 1],
                       d.r32[3 * (size_t)row + 2]};
        float mr[3];
        mat3vec(glob(d.gf_pinv) + 9 * (size_t)row, rr, mr);
        for (int a = 0; a < 3; a++)
            o[a] = st.it == 0 ? mr[a] : mr[a] + bt * dold[3 * (size_t)row + a];
    };
    const size_t base = (size_t)d.slice_off[s];
    const int w = d.row_w[p];
    float z[3] = {0, 0, 0};
    if (!d.row_pinned[p])
        for (int k = 0; k < w; k++)
        {
            const size_t e = base + (size_t)k * SLICE + lane;
            const int c = d.col[e];
            if (d.row_pinned[c])
                continue;
            float dc[3];
            dir(c, dc);
            const float *Wk = d.W32 + 9 * base + (9 * (size_t)k) * SLICE + lane;
            for (int a = 0; a < 3; a++)
                z[a] += Wk[(3 * a) * SLICE] * dc[0] + Wk[(3 * a + 1) * SLICE] * dc[1] +
                        Wk[(3 * a + 2) * SLICE] * dc[2];
        }
    float dn[3];
    dir(p, dn);
    for (int a = 0; a < 3; a++)
    {
        dnew[3 * (size_t)p + a] = dn[a];
        d.q32[3 * (size_t)p + a] = z[a];
    }
    gf_write(d.cg_partial + (size_t)PARTS * d.n_slices, s, P_DQ,
             (double)dn[0] * z[0] + (double)dn[1] * z[1] + (double)dn[2] * z[2]);
    if (lead)
        store_gf(out, st);
}

// alpha = r.M^-1 r / d.z; r -= alpha z; dx += alpha d; partials into slot 0
__global__ void __launch_bounds__(64) k_gf_update(DeviceBatch d)
{
    const int s = blockIdx.x, lane = threadIdx.x, p = s * SLICE + lane;
    const int inst = d.slice_inst[s];
    const InstParams &ip = d.params[inst];
    const bool lead = s == ip.slice_begin && lane == 0;
    GfState st = load_gf(d.gf_state + 2 * inst + 1);
    GfState *out = d.gf_state + 2 * inst;
    if (st.done)
    {
        if (lead)
            store_gf(out, st);
        return;
    }
    double v[1];
    const int comp[1] = {P_DQ};
    gf_sums<1>(d, ip, 1, comp, v);
    const float alpha = v[0] == 0.0 ? 0.0f : (float)(st.rmr / v[0]);
    const float *dn = (st.it & 1) ? d.d32[1] : d.d32[0];
    float r[3];
    for (int a = 0; a < 3; a++)
    {
        const size_t i = 3 * (size_t)p + a;
        r[a] = d.r32[i] - alpha * d.q32[i];
        d.r32[i] = r[a];
        d.gf_dx[i] = d.gf_dx[i] + alpha * dn[i];
    }
    float mr[3];
    mat3vec(glob(d.gf_pinv) + 9 * (size_t)p, r, mr);
    double *part = d.cg_partial;
    gf_write(part, s, P_RZ, (double)r[0] * mr[0] + (double)r[1] * mr[1] + (double)r[2] * mr[2]);
    gf_write(part, s, P_RR, (double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2]);
    st.it += 1;
    if (lead)
        store_gf(out, st);
}

// x_pre = x; x += dx (the solution stays as the next step's warm start)
__global__ void __launch_bounds__(256) k_gf_finish(DeviceBatch d)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= 3 * (size_t)d.n_rows)
        return;
    d.xpre[i] = d.x[i];
    d.x[i] = d.x[i] + (double)d.gf_dx[i];
}

} // namespace

void launch_gf_assemble(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_gf_pins, dim3(d.n_slices), dim3(64), 0, s, d);
    if (d.n_tri > 0)
        hipLaunchKernelGGL(k_gf_tri, dim3((d.n_tri + 127) / 128), dim3(128), 0, s, d);
    hipLaunchKernelGGL(k_gf_rows, dim3(d.n_slices), dim3(64), 0, s, d);
}

void launch_gf_solve(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_gf_init, dim3(d.n_slices), dim3(64), 0, s, d);
    launch_group_sums(d, 0, s);
    // 100 iterations, then one more prologue that records the last sums
    for (int it = 0; it <= 100; it++)
    {
        hipLaunchKernelGGL(k_gf_spmv, dim3(d.n_slices), dim3(64), 0, s, d);
        launch_group_sums(d, 1, s);
        if (it == 100)
            break;
        hipLaunchKernelGGL(k_gf_update, dim3(d.n_slices), dim3(64), 0, s, d);
        launch_group_sums(d, 0, s);
    }
    hipLaunchKernelGGL(k_gf_finish, dim3((unsigned)((3 * (size_t)d.n_rows + 255) / 256)), dim3(256),
                       0, s, d);
}

} // namespace bw
