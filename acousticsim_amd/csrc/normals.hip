// Render-side readback: triangle and vertex normals at the current positions
// (cBaseObject::UpdateTriangleNormal and UpdateVertexNormalFromTriangleNormal,
// BaseObject.cpp:119-160, which cBaseCloth::Update runs after every substep,
// BaseCloth.cpp:585-590).
//
//   triangle t: n_t = normalize((x1 - x0) x (x2 - x1))
//   vertex v:   n_v = normalize(sum over triangles t containing v, in
//                     ascending t, of n_t * A_t), A_t the float rest area
//
// normalize(u) = u / sqrt(|u|^2) when |u|^2 > 0, else u, with |u|^2 summed as
// (x^2 + z^2) + y^2 (the pairwise order of the reference's 4-wide vector,
// whose fourth component is 0).  Contraction is off for this file so that
// products round as on the reference's CPU.
#include "device.h"

namespace bw
{
namespace
{

__device__ inline void normalize3(double &x, double &y, double &z)
{
    const double n2 = (x * x + z * z) + y * y;
    if (n2 > 0.0)
    {
        const double n = sqrt(n2);
        x = x / n;
        y = y / n;
        z = z / n;
    }
}

__global__ void __launch_bounds__(256)
k_triangle_normals(int n_tri, const int *__restrict__ tri_rows, const double *__restrict__ X,
                   double *__restrict__ tn)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tri)
        return;
    const double *p0 = X + 3 * (size_t)tri_rows[3 * (size_t)t];
    const double *p1 = X + 3 * (size_t)tri_rows[3 * (size_t)t + 1];
    const double *p2 = X + 3 * (size_t)tri_rows[3 * (size_t)t + 2];
    const double a0 = p1[0] - p0[0], a1 = p1[1] - p0[1], a2 = p1[2] - p0[2];
    const double b0 = p2[0] - p1[0], b1 = p2[1] - p1[1], b2 = p2[2] - p1[2];
    double x = a1 * b2 - a2 * b1, y = a2 * b0 - a0 * b2, z = a0 * b1 - a1 * b0;
    normalize3(x, y, z);
    tn[3 * (size_t)t] = x;
    tn[3 * (size_t)t + 1] = y;
    tn[3 * (size_t)t + 2] = z;
}

__global__ void __launch_bounds__(256)
k_vertex_normals(int n_vert, const int *__restrict__ vt_ptr, const int *__restrict__ vt_list,
                 const float *__restrict__ tri_area, const double *__restrict__ tn,
                 double *__restrict__ vn)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vert)
        return;
    double x = 0.0, y = 0.0, z = 0.0;
    for (int k = vt_ptr[v]; k < vt_ptr[v + 1]; k++)
    {
        const int t = vt_list[k] >> 2;
        const double a = (double)tri_area[t];
        x = x + tn[3 * (size_t)t] * a;
        y = y + tn[3 * (size_t)t + 1] * a;
        z = z + tn[3 * (size_t)t + 2] * a;
    }
    normalize3(x, y, z);
    vn[3 * (size_t)v] = x;
    vn[3 * (size_t)v + 1] = y;
    vn[3 * (size_t)v + 2] = z;
}

} // namespace

void launch_normals(const DeviceBatch &d, double *tn, double *vn, hipStream_t s)
{
    if (d.n_tri > 0)
        hipLaunchKernelGGL(k_triangle_normals, dim3((d.n_tri + 255) / 256), dim3(256), 0, s,
                           d.n_tri, d.tri_rows, d.x, tn);
    if (d.n_vert > 0)
        hipLaunchKernelGGL(k_vertex_normals, dim3((d.n_vert + 255) / 256), dim3(256), 0, s,
                           d.n_vert, d.vt_ptr, d.vt_list, d.tri_area, tn, vn);
}

} // namespace bw
