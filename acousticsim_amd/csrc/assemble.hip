// Per-step assembly on the device, FP64, deterministic (gather, no atomics).
//
//   k_stretch_elements : one thread per triangle.  Baraff-Witkin stretch on
//     warp/weft and on the +-45 degree bias directions (cBaraffMaterial::Update,
//     BaraffMaterial.cpp:173-256): per direction set it stores the unit
//     vectors n_u, n_v, 1/|Fu| and the identity term, and the element force.
//   k_bend_elements : one thread per interior edge.  Dihedral angle beta and
//     its gradient g (cDihedralMaterial::Update, DihedralBending.cpp:378-447);
//     it stores l = factor * g, g and factor * beta.
//   k_assemble_rows : one wavefront per ELL slice, one lane per block row.
//     Each block W(v, c) = sk * (K_s + K_b) (+ sm * m_v I on the diagonal) is
//     formed by summing its element contributions in ascending element order,
//     exactly the order in which the reference's setFromTriplets sums
//     duplicate triplets (BaraffMaterial.cpp:297-351, DihedralBending.cpp:449-489),
//     so W equals the reference's W = (1 + h a) M + h (b - h) K
//     (BaraffCloth.cpp:274-275) entry for entry.  The same lane sums the
//     vertex force and forms rhs = h^2 (g + f_user + f) + h M (x - x_pre) / h
//     (BaraffCloth.cpp:282-283), the Jacobi diagonal and the warm start
//     dx0 = x - x_pre (BaraffCloth.cpp:284).  Rows and columns of pinned
//     vertices are zeroed, which is the reference's filter in ApplyMatmul
//     (BaraffCloth.cpp:469-506).
//
// This file is compiled with floating-point contraction off: every product
// and sum is rounded where the reference rounds it.
#include "device.h"

namespace bw
{
namespace
{

__device__ inline double dot3(const double *a, const double *b)
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
__device__ inline void cross3(const double *a, const double *b, double *o)
{
    const double t0 = a[1] * b[2] - a[2] * b[1];
    const double t1 = a[2] * b[0] - a[0] * b[2];
    const double t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
}
__device__ inline void normalized3(const double *a, double *o)
{
    const double z = dot3(a, a);
    if (z > 0)
    {
        const double s = sqrt(z);
        o[0] = a[0] / s;
        o[1] = a[1] / s;
        o[2] = a[2] / s;
    }
    else
    {
        o[0] = a[0];
        o[1] = a[1];
        o[2] = a[2];
    }
}
__device__ inline void load3(const double *x, int row, double *o)
{
    o[0] = x[3 * (size_t)row];
    o[1] = x[3 * (size_t)row + 1];
    o[2] = x[3 * (size_t)row + 2];
}

constexpr int ING = 25; // doubles per element record

__global__ void __launch_bounds__(256)
k_stretch_elements(int n_tri, const int *__restrict__ tri_rows, const int *__restrict__ tri_inst,
                   const float *__restrict__ coef, const float *__restrict__ area,
                   const InstParams *__restrict__ params, const double *__restrict__ X,
                   double *__restrict__ ing)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tri)
        return;
    double x[3][3];
    for (int i = 0; i < 3; i++)
        load3(X, tri_rows[3 * t + i], x[i]);
    const InstParams &P = params[tri_inst[t]];
    const double A = (double)area[t];
    double f[3][3];
    double *out = ing + (size_t)ING * t;
    for (int set = 0; set < 2; set++)
    {
        const double Ku = set == 0 ? P.k0 : P.k2, Kv = set == 0 ? P.k1 : P.k2;
        double cu[3], cv[3];
        for (int i = 0; i < 3; i++)
        {
            cu[i] = (double)coef[12 * t + 6 * set + i];
            cv[i] = (double)coef[12 * t + 6 * set + 3 + i];
        }
        double Fu[3], Fv[3];
        for (int a = 0; a < 3; a++)
        {
            Fu[a] = (cu[0] * x[0][a] + cu[1] * x[1][a]) + cu[2] * x[2][a];
            Fv[a] = (cv[0] * x[0][a] + cv[1] * x[1][a]) + cv[2] * x[2][a];
        }
        const double nFu = sqrt(dot3(Fu, Fu));
        const double nFv = sqrt(dot3(Fv, Fv));
        const double Cu = nFu - 1, Cv = nFv - 1;
        double nu[3], nv[3];
        normalized3(Fu, nu);
        normalized3(Fv, nv);
        // CalcFint (BaraffMaterial.cpp:113-125)
        const double su = ((-A) * Ku) * Cu, sv = ((-A) * Kv) * Cv;
        const double ub[3] = {su * nu[0], su * nu[1], su * nu[2]};
        const double vb[3] = {sv * nv[0], sv * nv[1], sv * nv[2]};
        for (int i = 0; i < 3; i++)
            for (int a = 0; a < 3; a++)
            {
                const double fi = ub[a] * cu[i] + vb[a] * cv[i];
                f[i][a] = set == 0 ? (0.0 + fi) : f[i][a] + fi;
            }
        // CalcHessian ingredients; both D_u and D_v use |Fu| (BaraffMaterial.cpp:141-142)
        const double inv = 1.0 / nFu;
        // identity term (1 - 1/|Fu|) I when |Fu| > 1, else a negative marker
        const double add = nFu > 1 ? 1 - 1.0 / nFu : -1.0;
        double *o = out + 8 * set;
        o[0] = nu[0];
        o[1] = nu[1];
        o[2] = nu[2];
        o[3] = nv[0];
        o[4] = nv[1];
        o[5] = nv[2];
        o[6] = inv;
        o[7] = add;
    }
    for (int i = 0; i < 3; i++)
        for (int a = 0; a < 3; a++)
            out[16 + 3 * i + a] = f[i][a];
}

// (I - n n^T) / |v| (DihedralBending.cpp:115-119)
__device__ inline void iprime(const double *vec, double M[3][3])
{
    double n[3];
    normalized3(vec, n);
    const double len = sqrt(dot3(vec, vec));
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            M[a][b] = ((a == b ? 1.0 : 0.0) - n[b] * n[a]) / len;
}

__global__ void __launch_bounds__(256)
k_bend_elements(int n_ie, const int *__restrict__ ie_rows, const double *__restrict__ factor,
                const double *__restrict__ X, double *__restrict__ ing)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_ie)
        return;
    double v0[3], v1[3], v2[3], v3[3];
    load3(X, ie_rows[4 * e], v0);
    load3(X, ie_rows[4 * e + 1], v1);
    load3(X, ie_rows[4 * e + 2], v2);
    load3(X, ie_rows[4 * e + 3], v3);
    double t0[3], t1[3], n0[3], n1[3], ed[3];
    for (int a = 0; a < 3; a++)
    {
        t0[a] = v1[a] - v0[a];
        t1[a] = v2[a] - v1[a];
    }
    cross3(t0, t1, n0);
    for (int a = 0; a < 3; a++)
    {
        t0[a] = v0[a] - v1[a];
        t1[a] = v3[a] - v0[a];
        ed[a] = v1[a] - v0[a];
    }
    cross3(t0, t1, n1);
    double n0b[3], n1b[3], eb[3], c01[3];
    normalized3(n0, n0b);
    normalized3(n1, n1b);
    normalized3(ed, eb);
    const double xx = dot3(n0b, n1b);
    cross3(n0b, n1b, c01);
    const double yy = dot3(eb, c01);
    const double beta = atan2(yy, xx);
    // CalcDBetaDnormal_and_De (DihedralBending.cpp:163-223)
    const double sign_sinb = sin(beta) > 0 ? 1.0 : -1.0;
    double I0[3][3], I1[3][3];
    iprime(n0, I0);
    iprime(n1, I1);
    const double m = 1 - xx * xx;
    double deno = sqrt(m > 0 ? m : 0.0);
    if (deno < 1e-10)
        deno = 1e-10;
    const double fct = -sign_sinb * 1.0 / deno;
    double dn0[3], dn1[3];
    for (int a = 0; a < 3; a++)
    {
        dn0[a] = ((fct * I0[0][a]) * n1b[0] + (fct * I0[1][a]) * n1b[1]) + (fct * I0[2][a]) * n1b[2];
        dn1[a] = ((fct * I1[0][a]) * n0b[0] + (fct * I1[1][a]) * n0b[1]) + (fct * I1[2][a]) * n0b[2];
    }
    // CalcDBetaDx (DihedralBending.cpp:259-264)
    double g[12], a0[3], a1[3], c0[3], c1[3];
    for (int a = 0; a < 3; a++)
    {
        a0[a] = v1[a] - v2[a];
        a1[a] = v3[a] - v1[a];
    }
    cross3(a0, dn0, c0);
    cross3(a1, dn1, c1);
    for (int a = 0; a < 3; a++)
        g[a] = c0[a] + c1[a];
    for (int a = 0; a < 3; a++)
    {
        a0[a] = v2[a] - v0[a];
        a1[a] = v0[a] - v3[a];
    }
    cross3(a0, dn0, c0);
    cross3(a1, dn1, c1);
    for (int a = 0; a < 3; a++)
        g[3 + a] = c0[a] + c1[a];
    for (int a = 0; a < 3; a++)
    {
        a0[a] = v0[a] - v1[a];
        a1[a] = v1[a] - v0[a];
    }
    cross3(a0, dn0, g + 6);
    cross3(a1, dn1, g + 9);
    const double F = factor[e];
    double *out = ing + (size_t)ING * e;
    for (int k = 0; k < 12; k++)
    {
        out[k] = F * g[k];
        out[12 + k] = g[k];
    }
    out[24] = F * beta;
}

__global__ void __launch_bounds__(SLICE)
k_assemble_rows(DeviceBatch d)
{
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    const int p = s * SLICE + lane;
    const int width = d.slice_w[s];
    const size_t base = (size_t)d.slice_off[s];
    const int vg = d.row_vertex[p];
    const InstParams &P = d.params[d.slice_inst[s]];
    double *W64 = d.W64 + 9 * base + lane;
    float *W32 = d.W32 + 9 * base + lane;
    if (vg < 0)
    {
        for (int k = 0; k < width; k++)
            for (int q = 0; q < 9; q++)
            {
                W64[(9 * k + q) * SLICE] = 0.0;
                W32[(9 * k + q) * SLICE] = 0.0f;
            }
        return;
    }
    const bool pin_p = d.row_pinned[p] != 0;
    const double mass = d.row_mass[p];
    const int b0 = d.blk_ptr[vg], nb = d.blk_ptr[vg + 1] - b0;
    double diag[3] = {1.0, 1.0, 1.0};
    for (int k = 0; k < width; k++)
    {
        double w[9];
        if (k < nb)
        {
            const int blk = b0 + k;
            const int c = d.blk_col[blk];
            double ks[9], kb[9];
            for (int q = 0; q < 9; q++)
            {
                ks[q] = 0.0;
                kb[q] = 0.0;
            }
            // stretch: (0 + H_warpweft) + H_bias per triangle, triangles ascending
            for (int it = d.s_ptr[blk]; it < d.s_ptr[blk + 1]; it++)
            {
                const int code = d.s_list[it];
                const int t = code >> 4, i = (code >> 2) & 3, j = code & 3;
                const double *ing = d.tri_ing + (size_t)ING * t;
                const float *cf = d.tri_coef + 12 * (size_t)t;
                const double A = (double)d.tri_area[t];
                const InstParams &Pt = d.params[d.tri_inst[t]];
                double h[9];
                for (int set = 0; set < 2; set++)
                {
                    const double Ku = set == 0 ? Pt.k0 : Pt.k2, Kv = set == 0 ? Pt.k1 : Pt.k2;
                    const double su2 = (Ku * (double)cf[6 * set + i]) * (double)cf[6 * set + j];
                    const double sv2 = (Kv * (double)cf[6 * set + 3 + i]) * (double)cf[6 * set + 3 + j];
                    const double *o = ing + 8 * set;
                    const double inv = o[6], add = o[7];
                    double lu[3], lv[3];
                    for (int a = 0; a < 3; a++)
                    {
                        lu[a] = inv * o[a];
                        lv[a] = inv * o[3 + a];
                    }
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++)
                        {
                            double du = o[b] * lu[a];
                            double dv = o[3 + b] * lv[a];
                            if (add >= 0)
                            {
                                const double ad = a == b ? add : add * 0.0;
                                du = du + ad;
                                dv = dv + ad;
                            }
                            const double hh = (-A) * (su2 * du + sv2 * dv);
                            h[3 * a + b] = set == 0 ? (0.0 + hh) : h[3 * a + b] + hh;
                        }
                }
                for (int q = 0; q < 9; q++)
                    ks[q] = ks[q] + h[q];
            }
            // bending: float(-(l_a g_b)) per edge, edges ascending, negated total
            for (int it = d.b_ptr[blk]; it < d.b_ptr[blk + 1]; it++)
            {
                const int code = d.b_list[it];
                const int e = code >> 4, i = (code >> 2) & 3, j = code & 3;
                const double *ing = d.ie_ing + (size_t)ING * e;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                    {
                        const float hf = (float)(-(ing[3 * i + a] * ing[12 + 3 * j + b]));
                        kb[3 * a + b] = kb[3 * a + b] + (double)hf;
                    }
            }
            for (int q = 0; q < 9; q++)
                w[q] = P.sk * (ks[q] + (-1 * kb[q]));
            if (c == p)
            {
                const double mm = P.sm * mass;
                w[0] = mm + w[0];
                w[4] = mm + w[4];
                w[8] = mm + w[8];
                diag[0] = w[0];
                diag[1] = w[4];
                diag[2] = w[8];
            }
            if (pin_p || d.row_pinned[c])
                for (int q = 0; q < 9; q++)
                    w[q] = 0.0;
        }
        else
        {
            for (int q = 0; q < 9; q++)
                w[q] = 0.0;
        }
        for (int q = 0; q < 9; q++)
        {
            W64[(9 * k + q) * SLICE] = w[q];
            W32[(9 * k + q) * SLICE] = (float)w[q];
        }
    }
    // force: stretch sum over incident triangles, then bending over incident
    // interior edges (CalcTotalForce BaraffMaterial.cpp:267-282,
    // AssignForceAndMatrix DihedralBending.cpp:449-489)
    double fs[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0};
    for (int it = d.vt_ptr[vg]; it < d.vt_ptr[vg + 1]; it++)
    {
        const int code = d.vt_list[it];
        const double *ing = d.tri_ing + (size_t)ING * (code >> 2) + 16 + 3 * (code & 3);
        for (int a = 0; a < 3; a++)
            fs[a] = fs[a] + ing[a];
    }
    for (int it = d.ve_ptr[vg]; it < d.ve_ptr[vg + 1]; it++)
    {
        const int code = d.ve_list[it];
        const double *ing = d.ie_ing + (size_t)ING * (code >> 2);
        const int i = code & 3;
        for (int a = 0; a < 3; a++)
            fb[a] = fb[a] + (ing[24] * ing[12 + 3 * i + a]) / 2;
    }
    const double dt = P.dt, dt2 = dt * dt;
    const size_t r3 = 3 * (size_t)p;
    for (int a = 0; a < 3; a++)
    {
        const double f = fs[a] + fb[a];
        const double grav = a == 1 ? -mass * 9.8 * P.grav : 0.0;
        const double x = d.x[r3 + a], xp = d.xpre[r3 + a];
        const double u = d.fuser ? d.fuser[r3 + a] : 0.0;
        const double t1 = dt2 * ((grav + u) + f);
        const double t2 = ((dt * mass) * (x - xp)) / dt;
        d.b[r3 + a] = pin_p ? 0.0 : t1 + t2;
        d.dx[r3 + a] = pin_p ? 0.0 : x - xp;
        if (d.dx32)
            d.dx32[r3 + a] = 0.0f;
        const double pinv = pin_p ? 0.0 : 1.0 / diag[a];
        d.P64[r3 + a] = pinv;
        d.P32[r3 + a] = (float)pinv;
    }
}

} // namespace

void launch_elements(const DeviceBatch &d, hipStream_t s)
{
    if (d.n_tri > 0)
        hipLaunchKernelGGL(k_stretch_elements, dim3((d.n_tri + 255) / 256), dim3(256), 0, s,
                           d.n_tri, d.tri_rows, d.tri_inst, d.tri_coef, d.tri_area, d.params,
                           d.x, d.tri_ing);
    if (d.n_ie > 0)
        hipLaunchKernelGGL(k_bend_elements, dim3((d.n_ie + 255) / 256), dim3(256), 0, s, d.n_ie,
                           d.ie_rows, d.ie_factor, d.x, d.ie_ing);
}

void launch_assemble_rows(const DeviceBatch &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_assemble_rows, dim3(d.n_slices), dim3(SLICE), 0, s, d);
}

} // namespace bw
