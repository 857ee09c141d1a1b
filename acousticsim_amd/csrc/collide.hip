// Cloth-body collision detection (cBVHCollisionDetecter::PerformCD,
// BVHCollisionDetecter.cpp:360-651) for one cloth against one kinematic body.
//
// The reference's broadphase grows every triangle's box by 5e-3
// (ObjectBVH.cpp:71-97) and pairs the triangles whose boxes meet, so every
// pair closer than 1e-2 reaches the narrow phase, which keeps those whose
// computed distance is below 3e-3.  The contacts are therefore all pairs
// below 3e-3; here each thread takes one cloth vertex, triangle or edge and
// scans the body's triangles, vertices or edges through LDS in tiles,
// skipping those whose box lies more than 4e-3 away, and applies the
// reference's narrow-phase arithmetic (float region tests, double points) to
// the rest; a workgroup skips every tile whose box (kept per tile of 256
// body items) lies more than 4e-3 from the box of its own cloth items.  Two
// passes: counts, then writes at host-scanned offsets, so the
// output order is fixed by the thread index and the scan order.
//
// Contraction is off for this file (the narrow phase rounds as the
// reference's x86-64 build).
#include "collide.h"
#include "device.h"

namespace bw
{
namespace
{

constexpr int TILE = DETECT_TILE;
constexpr double REJECT = 4e-3; // beyond it no computed distance is below 3e-3
constexpr double CONTACT = 3e-3;

__device__ inline double d3(const double *a, const double *b)
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

__device__ inline void sub3(const double *a, const double *b, double *o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

// ClosestPtPointTriangle (BVHCollisionDetecter.cpp:17-83)
__device__ void closest_pt_tri(const double *p, const double *a, const double *b, const double *c,
                               float *bary, double *out)
{
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    sub3(b, a, ab);
    sub3(c, a, ac);
    sub3(p, a, ap);
    const float d1 = (float)d3(ab, ap), d2 = (float)d3(ac, ap);
    bary[0] = bary[1] = bary[2] = 0.0f;
    if (d1 <= 0.0f && d2 <= 0.0f)
    {
        bary[0] = 1;
        for (int k = 0; k < 3; k++)
            out[k] = a[k];
        return;
    }
    sub3(p, b, bp);
    const float d3f = (float)d3(ab, bp), d4 = (float)d3(ac, bp);
    if (d3f >= 0.0f && d4 <= d3f)
    {
        bary[1] = 1;
        for (int k = 0; k < 3; k++)
            out[k] = b[k];
        return;
    }
    const float vc = d1 * d4 - d3f * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3f <= 0.0f)
    {
        const float v = d1 / (d1 - d3f);
        bary[0] = 1 - v;
        bary[1] = v;
        for (int k = 0; k < 3; k++)
            out[k] = a[k] + v * ab[k];
        return;
    }
    sub3(p, c, cp);
    const float d5 = (float)d3(ab, cp), d6 = (float)d3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6)
    {
        bary[2] = 1;
        for (int k = 0; k < 3; k++)
            out[k] = c[k];
        return;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f)
    {
        const float w = d2 / (d2 - d6);
        bary[0] = 1 - w;
        bary[2] = w;
        for (int k = 0; k < 3; k++)
            out[k] = a[k] + w * ac[k];
        return;
    }
    const float va = d3f * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3f) >= 0.0f && (d5 - d6) >= 0.0f)
    {
        const float w = (d4 - d3f) / ((d4 - d3f) + (d5 - d6));
        bary[1] = 1 - w;
        bary[2] = w;
        for (int k = 0; k < 3; k++)
            out[k] = b[k] + w * (c[k] - b[k]);
        return;
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    bary[0] = 1 - v - w;
    bary[1] = v;
    bary[2] = w;
    for (int k = 0; k < 3; k++)
        out[k] = a[k] + ab[k] * v + ac[k] * w;
}

// cMathUtil::Clamp(v, 0, 1) in double (MathUtil.cpp:31-34)
__device__ inline float clamp01(float v)
{
    const double m = (1.0 < (double)v) ? 1.0 : (double)v;
    return (float)((0.0 < m) ? m : 0.0);
}

// EdgeEdgeDist's parameters (BVHCollisionDetecter.cpp:104-177)
__device__ void edge_edge_params(const double *p1, const double *q1, const double *p2,
                                 const double *q2, float &s, float &t)
{
    double dd1[3], dd2[3], r[3];
    sub3(q1, p1, dd1);
    sub3(q2, p2, dd2);
    sub3(p1, p2, r);
    const float a = (float)d3(dd1, dd1), e = (float)d3(dd2, dd2), f = (float)d3(dd2, r);
    const float EPSILON = 1e-10f;
    s = 0.0f;
    t = 0.0f;
    if (a <= EPSILON)
    {
        s = 0.0f;
        t = clamp01(f / e);
    }
    else
    {
        const float c = (float)d3(dd1, r);
        if (e <= EPSILON)
        {
            t = 0.0f;
            s = clamp01(-c / a);
        }
        else
        {
            const float bb = (float)d3(dd1, dd2);
            const float denom = a * e - bb * bb;
            s = denom != 0.0f ? clamp01((bb * f - c * e) / denom) : 0.0f;
            t = (bb * s + f) / e;
            if (t < 0.0f)
            {
                t = 0.0f;
                s = clamp01(-c / a);
            }
            else if (t > 1.0f)
            {
                t = 1.0f;
                s = clamp01((bb - c) / a);
            }
        }
    }
}

__device__ inline bool outside(const double *p, const double *lo, const double *hi)
{
    return p[0] < lo[0] - REJECT || p[0] > hi[0] + REJECT || p[1] < lo[1] - REJECT ||
           p[1] > hi[1] + REJECT || p[2] < lo[2] - REJECT || p[2] > hi[2] + REJECT;
}

// the workgroup's box: union of its threads' boxes (lo = +huge, hi = -huge
// for threads without work), reduced through LDS
__device__ void block_box(double (*red)[TILE], const double *lo, const double *hi, double *blo,
                          double *bhi)
{
    const int t = threadIdx.x;
    for (int a = 0; a < 3; a++)
    {
        red[a][t] = lo[a];
        red[3 + a][t] = hi[a];
    }
    __syncthreads();
    for (int w = TILE / 2; w > 0; w >>= 1)
    {
        if (t < w)
            for (int a = 0; a < 3; a++)
            {
                red[a][t] = fmin(red[a][t], red[a][t + w]);
                red[3 + a][t] = fmax(red[3 + a][t], red[3 + a][t + w]);
            }
        __syncthreads();
    }
    for (int a = 0; a < 3; a++)
    {
        blo[a] = red[a][0];
        bhi[a] = red[3 + a][0];
    }
    __syncthreads();
}

// a tile of body items wholly outside the workgroup's box (grown by the
// reject distance) holds no pair for it
__device__ inline bool tile_apart(const double *tb, const double *blo, const double *bhi)
{
    return tb[0] > bhi[0] + REJECT || tb[3] < blo[0] - REJECT || tb[1] > bhi[1] + REJECT ||
           tb[4] < blo[1] - REJECT || tb[2] > bhi[2] + REJECT || tb[5] < blo[2] - REJECT;
}

__device__ inline void box3(const double *a, const double *b, const double *c, double *lo,
                            double *hi)
{
    for (int k = 0; k < 3; k++)
    {
        lo[k] = fmin(fmin(a[k], b[k]), c[k]);
        hi[k] = fmax(fmax(a[k], b[k]), c[k]);
    }
}

// cloth vertex against the body's triangles (obj0_is_cloth_obj1_is_rb):
// outer normal = the body triangle's, depth = (p - closest) . normal
__global__ void __launch_bounds__(TILE)
k_detect_points(ClothView c, int nbt, const double *__restrict__ bx, const int *__restrict__ btri,
                const double *__restrict__ btn, const double *__restrict__ tile_box,
                int *__restrict__ count, const int *__restrict__ offset, Hit *__restrict__ out)
{
    __shared__ double s_v[TILE][9];
    __shared__ double s_box[TILE][6];
    __shared__ double s_red[6][TILE];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = v < c.nv && c.vt_ptr[c.vbase + v + 1] > c.vt_ptr[c.vbase + v];
    double p[3] = {0, 0, 0};
    if (active)
    {
        const double *x = c.X + 3 * (size_t)c.vrow[v];
        p[0] = x[0];
        p[1] = x[1];
        p[2] = x[2];
    }
    double blo[3], bhi[3];
    {
        const double big = 1e300;
        const double lo[3] = {active ? p[0] : big, active ? p[1] : big, active ? p[2] : big};
        const double hi[3] = {active ? p[0] : -big, active ? p[1] : -big, active ? p[2] : -big};
        block_box(s_red, lo, hi, blo, bhi);
    }
    int n = 0;
    const int o = (out && active) ? offset[v] : 0;
    for (int t0 = 0; t0 < nbt; t0 += TILE)
    {
        if (tile_apart(tile_box + 6 * (size_t)(t0 / TILE), blo, bhi))
            continue; // the same decision in every thread
        __syncthreads();
        const int k = threadIdx.x;
        if (t0 + k < nbt)
        {
            for (int j = 0; j < 3; j++)
                for (int a = 0; a < 3; a++)
                    s_v[k][3 * j + a] = bx[3 * (size_t)btri[3 * (size_t)(t0 + k) + j] + a];
            box3(s_v[k], s_v[k] + 3, s_v[k] + 6, s_box[k], s_box[k] + 3);
        }
        __syncthreads();
        const int m = min(TILE, nbt - t0);
        if (active)
            for (int q = 0; q < m; q++)
            {
                if (outside(p, s_box[q], s_box[q] + 3))
                    continue;
                float bary[3];
                double cl[3], dv[3];
                closest_pt_tri(p, s_v[q], s_v[q] + 3, s_v[q] + 6, bary, cl);
                sub3(p, cl, dv);
                if (sqrt(d3(dv, dv)) < CONTACT)
                {
                    if (out)
                    {
                        Hit h;
                        const double *nn = btn + 3 * (size_t)(t0 + q);
                        h.n[0] = nn[0];
                        h.n[1] = nn[1];
                        h.n[2] = nn[2];
                        h.depth = (float)d3(dv, h.n);
                        h.bary[0] = bary[0];
                        h.bary[1] = bary[1];
                        h.bary[2] = bary[2];
                        h.flag = 1;
                        h.key0 = v;
                        h.key1 = t0 + q;
                        h.pad = 0;
                        out[o + n] = h;
                    }
                    n++;
                }
            }
    }
    if (!out && v < c.nv)
        count[v] = active ? n : 0;
}

// body vertex against the cloth's triangles (obj0_is_cloth_obj1_is_rb
// false): outer normal = the cloth triangle's, turned towards the body
// vertex's normal; depth = (closest - p) . normal
__global__ void __launch_bounds__(TILE)
k_detect_body_points(ClothView c, int nbv, const double *__restrict__ bx,
                     const double *__restrict__ bvn, const unsigned char *__restrict__ b_in,
                     const double *__restrict__ tile_box, int *__restrict__ count,
                     const int *__restrict__ offset, Hit *__restrict__ out)
{
    __shared__ double s_p[TILE][3];
    __shared__ double s_n[TILE][3];
    __shared__ int s_in[TILE];
    __shared__ double s_red[6][TILE];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = t < c.nt;
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, cc[3] = {0, 0, 0}, lo[3], hi[3], tn[3] = {0, 0, 0};
    if (active)
    {
        const int *r = c.tri_rows + 3 * (size_t)(c.tbase + t);
        for (int k = 0; k < 3; k++)
        {
            a[k] = c.X[3 * (size_t)r[0] + k];
            b[k] = c.X[3 * (size_t)r[1] + k];
            cc[k] = c.X[3 * (size_t)r[2] + k];
            tn[k] = c.tn[3 * (size_t)(c.tbase + t) + k];
        }
    }
    box3(a, b, cc, lo, hi);
    double blo[3], bhi[3];
    {
        const double big = 1e300;
        const double l2[3] = {active ? lo[0] : big, active ? lo[1] : big, active ? lo[2] : big};
        const double h2[3] = {active ? hi[0] : -big, active ? hi[1] : -big, active ? hi[2] : -big};
        block_box(s_red, l2, h2, blo, bhi);
    }
    int n = 0;
    const int o = (out && active) ? offset[t] : 0;
    for (int p0 = 0; p0 < nbv; p0 += TILE)
    {
        if (tile_apart(tile_box + 6 * (size_t)(p0 / TILE), blo, bhi))
            continue; // the same decision in every thread
        __syncthreads();
        const int k = threadIdx.x;
        if (p0 + k < nbv)
        {
            for (int j = 0; j < 3; j++)
            {
                s_p[k][j] = bx[3 * (size_t)(p0 + k) + j];
                s_n[k][j] = bvn[3 * (size_t)(p0 + k) + j];
            }
            s_in[k] = b_in[p0 + k];
        }
        __syncthreads();
        const int m = min(TILE, nbv - p0);
        if (active)
            for (int q = 0; q < m; q++)
            {
                if (!s_in[q] || outside(s_p[q], lo, hi))
                    continue;
                float bary[3];
                double cl[3], dv[3], nn[3] = {tn[0], tn[1], tn[2]};
                closest_pt_tri(s_p[q], a, b, cc, bary, cl);
                if (d3(nn, s_n[q]) < 0)
                    for (int j = 0; j < 3; j++)
                        nn[j] = nn[j] * -1;
                sub3(cl, s_p[q], dv);
                if (sqrt(d3(dv, dv)) < CONTACT)
                {
                    if (out)
                    {
                        Hit h;
                        h.n[0] = nn[0];
                        h.n[1] = nn[1];
                        h.n[2] = nn[2];
                        h.depth = (float)d3(dv, nn);
                        h.bary[0] = bary[0];
                        h.bary[1] = bary[1];
                        h.bary[2] = bary[2];
                        h.flag = 0;
                        h.key0 = p0 + q;
                        h.key1 = t;
                        h.pad = 0;
                        out[o + n] = h;
                    }
                    n++;
                }
            }
    }
    if (!out && t < c.nt)
        count[t] = n;
}

// cloth edge against the body's edges: closest points of the two segments,
// outer normal = the normalised cross product of the edges (obj0's x obj1's)
// turned towards the body edge's side normal, distance vector from the
// body's point to the cloth's
__global__ void __launch_bounds__(TILE)
k_detect_edges(ClothView c, int nbe, const double *__restrict__ bx, const int *__restrict__ bedge,
               const double *__restrict__ bedge_n, const double *__restrict__ tile_box,
               int cloth_first, int *__restrict__ count, const int *__restrict__ offset,
               Hit *__restrict__ out)
{
    __shared__ double s_e[TILE][6];
    __shared__ double s_n[TILE][3];
    __shared__ double s_red[6][TILE];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = e < c.ne;
    double ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
    if (active)
        for (int k = 0; k < 3; k++)
        {
            ca[k] = c.X[3 * (size_t)c.edge_rows[2 * (size_t)e] + k];
            cb[k] = c.X[3 * (size_t)c.edge_rows[2 * (size_t)e + 1] + k];
        }
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++)
    {
        lo[k] = fmin(ca[k], cb[k]) - REJECT;
        hi[k] = fmax(ca[k], cb[k]) + REJECT;
    }
    double blo[3], bhi[3];
    {
        // lo / hi are already grown by the reject distance
        const double big = 1e300;
        const double l2[3] = {active ? lo[0] + REJECT : big, active ? lo[1] + REJECT : big,
                              active ? lo[2] + REJECT : big};
        const double h2[3] = {active ? hi[0] - REJECT : -big, active ? hi[1] - REJECT : -big,
                              active ? hi[2] - REJECT : -big};
        block_box(s_red, l2, h2, blo, bhi);
    }
    int n = 0;
    const int o = (out && active) ? offset[e] : 0;
    for (int e0 = 0; e0 < nbe; e0 += TILE)
    {
        if (tile_apart(tile_box + 6 * (size_t)(e0 / TILE), blo, bhi))
            continue; // the same decision in every thread
        __syncthreads();
        const int k = threadIdx.x;
        if (e0 + k < nbe)
        {
            const int *ev = bedge + 4 * (size_t)(e0 + k);
            for (int j = 0; j < 3; j++)
            {
                s_e[k][j] = bx[3 * (size_t)ev[0] + j];
                s_e[k][3 + j] = bx[3 * (size_t)ev[1] + j];
                s_n[k][j] = bedge_n[3 * (size_t)(e0 + k) + j];
            }
        }
        __syncthreads();
        const int m = min(TILE, nbe - e0);
        if (active)
            for (int q = 0; q < m; q++)
            {
                const double *ba = s_e[q], *bb = s_e[q] + 3;
                bool far = false;
                for (int j = 0; j < 3; j++)
                    far |= fmax(ba[j], bb[j]) < lo[j] || fmin(ba[j], bb[j]) > hi[j];
                if (far)
                    continue;
                const double *p0 = cloth_first ? ca : ba, *p1 = cloth_first ? cb : bb;
                const double *p2 = cloth_first ? ba : ca, *p3 = cloth_first ? bb : cb;
                float s, t;
                edge_edge_params(p0, p1, p2, p3, s, t);
                double e0p[3], e1p[3], u[3], w[3], cr[3], dv[3];
                for (int j = 0; j < 3; j++)
                {
                    e0p[j] = (double)(1 - s) * p0[j] + (double)s * p1[j];
                    e1p[j] = (double)(1 - t) * p2[j] + (double)t * p3[j];
                }
                sub3(p1, p0, u);
                sub3(p3, p2, w);
                cr[0] = u[1] * w[2] - u[2] * w[1];
                cr[1] = u[2] * w[0] - u[0] * w[2];
                cr[2] = u[0] * w[1] - u[1] * w[0];
                const double z = d3(cr, cr);
                if (z > 0)
                {
                    const double nz = sqrt(z);
                    for (int j = 0; j < 3; j++)
                        cr[j] = cr[j] / nz;
                }
                if (d3(s_n[q], cr) < 0)
                    for (int j = 0; j < 3; j++)
                        cr[j] = cr[j] * -1;
                if (cloth_first)
                    sub3(e0p, e1p, dv);
                else
                    sub3(e1p, e0p, dv);
                if (sqrt(d3(dv, dv)) < CONTACT)
                {
                    if (out)
                    {
                        Hit h;
                        h.n[0] = cr[0];
                        h.n[1] = cr[1];
                        h.n[2] = cr[2];
                        h.depth = (float)d3(dv, cr);
                        h.bary[0] = s;
                        h.bary[1] = t;
                        h.bary[2] = 0.0f;
                        h.flag = cloth_first ? 1 : 0;
                        h.key0 = cloth_first ? e : e0 + q;
                        h.key1 = cloth_first ? e0 + q : e;
                        h.pad = 0;
                        out[o + n] = h;
                    }
                    n++;
                }
            }
    }
    if (!out && e < c.ne)
        count[e] = n;
}

} // namespace

void launch_detect_points(const ClothView &c, const Body &b, int *count, const int *offset,
                          Hit *out, hipStream_t s)
{
    if (c.nv > 0)
        hipLaunchKernelGGL(k_detect_points, dim3((c.nv + TILE - 1) / TILE), dim3(TILE), 0, s, c,
                           b.nt, b.x_dev, b.tri_dev, b.tn_dev, b.tile_box_dev, count, offset, out);
}

void launch_detect_body_points(const ClothView &c, const Body &b, int *count, const int *offset,
                               Hit *out, hipStream_t s)
{
    if (c.nt > 0)
        hipLaunchKernelGGL(k_detect_body_points, dim3((c.nt + TILE - 1) / TILE), dim3(TILE), 0, s,
                           c, b.nv, b.x_dev, b.vn_dev, b.in_tri_dev,
                           b.tile_box_dev + 6 * (size_t)b.tiles(b.nt), count, offset, out);
}

void launch_detect_edges(const ClothView &c, const Body &b, int cloth_first, int *count,
                         const int *offset, Hit *out, hipStream_t s)
{
    if (c.ne > 0)
        hipLaunchKernelGGL(k_detect_edges, dim3((c.ne + TILE - 1) / TILE), dim3(TILE), 0, s, c,
                           b.ne, b.x_dev, b.edge_dev, b.edge_n_dev,
                           b.tile_box_dev + 6 * (size_t)(b.tiles(b.nt) + b.tiles(b.nv)),
                           cloth_first, count, offset, out);
}

} // namespace bw
