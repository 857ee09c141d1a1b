// Kinematic body set-up (host; contraction off so that the normals round as
// the reference's).  See collide.h.
#include "collide.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>

namespace bw
{
namespace
{
void hcheck(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error: ") + what + ": " + hipGetErrorString(e));
}
template <typename T> T *to_device(const std::vector<T> &v)
{
    T *p = nullptr;
    hcheck(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)), "hipMalloc body");
    if (!v.empty())
        hcheck(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice),
               "hipMemcpy body");
    return p;
}
// Vector3d dot: Eigen's vectorised sum (a0 b0 + a1 b1) + a2 b2
double d3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
// normalise a 4-wide normal (w = 0): squared norm (x^2 + z^2) + y^2
void normalize4(double *u)
{
    const double n2 = (u[0] * u[0] + u[2] * u[2]) + u[1] * u[1];
    if (n2 > 0)
    {
        const double n = std::sqrt(n2);
        for (int a = 0; a < 3; a++)
            u[a] = u[a] / n;
    }
}
} // namespace

namespace
{
// boxes of tiles of DETECT_TILE consecutive items, each item's points given
// by pts(item, k) for k < npts
template <typename F> std::vector<double> tile_boxes(int n, int npts, F pts)
{
    const int nt = (n + DETECT_TILE - 1) / DETECT_TILE;
    std::vector<double> box(6 * (size_t)nt);
    for (int t = 0; t < nt; t++)
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = t * DETECT_TILE; i < std::min(n, (t + 1) * DETECT_TILE); i++)
            for (int k = 0; k < npts; k++)
            {
                const double *p = pts(i, k);
                for (int a = 0; a < 3; a++)
                {
                    lo[a] = std::min(lo[a], p[a]);
                    hi[a] = std::max(hi[a], p[a]);
                }
            }
        for (int a = 0; a < 3; a++)
        {
            box[6 * (size_t)t + a] = lo[a];
            box[6 * (size_t)t + 3 + a] = hi[a];
        }
    }
    return box;
}
} // namespace

void Body::refresh_tiles()
{
    tile_box_tri = tile_boxes(nt, 3, [&](int i, int k) { return &x[3 * (size_t)tri[3 * (size_t)i + k]]; });
    tile_box_vert = tile_boxes(nv, 1, [&](int i, int) { return &x[3 * (size_t)i]; });
    tile_box_edge = tile_boxes(ne, 2, [&](int i, int k) { return &x[3 * (size_t)edge[4 * (size_t)i + k]]; });
    std::vector<double> all(tile_box_tri);
    all.insert(all.end(), tile_box_vert.begin(), tile_box_vert.end());
    all.insert(all.end(), tile_box_edge.begin(), tile_box_edge.end());
    if (!tile_box_dev)
        hcheck(hipMalloc(&tile_box_dev, all.size() * sizeof(double)), "hipMalloc body");
    hcheck(hipMemcpy(tile_box_dev, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice),
           "hipMemcpy body");
}

Body::Body(const double *xyz, int nv_, const int *tri_, int nt_) : nv(nv_), nt(nt_)
{
    if (nv < 3 || nt < 1 || !xyz || !tri_)
        throw std::invalid_argument("body: needs at least 3 vertices and 1 triangle");
    x.assign(xyz, xyz + 3 * (size_t)nv);
    tri.assign(tri_, tri_ + 3 * (size_t)nt);
    for (int v : tri)
        if (v < 0 || v >= nv)
            throw std::invalid_argument("body: triangle vertex out of range");
    for (double v : x)
        if (!std::isfinite(v))
            throw std::invalid_argument("body: non-finite position");
    // edges: sides sorted by ({min, max}, side)
    std::vector<std::array<int, 3>> sd(3 * (size_t)nt);
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
        {
            const int a = tri[3 * t + k], c = tri[3 * t + (k + 1) % 3];
            sd[3 * t + k] = {std::min(a, c), std::max(a, c), 3 * t + k};
        }
    std::sort(sd.begin(), sd.end());
    for (size_t i = 0; i < sd.size();)
    {
        size_t j = i + 1;
        while (j < sd.size() && sd[j][0] == sd[i][0] && sd[j][1] == sd[i][1])
            j++;
        if (j - i > 2)
            throw std::invalid_argument("body: an edge with more than two triangles");
        edge.insert(edge.end(),
                    {sd[i][0], sd[i][1], sd[i][2] / 3, j - i == 2 ? sd[i + 1][2] / 3 : -1});
        i = j;
    }
    ne = (int)edge.size() / 4;
    // normals at creation (BaseObject.cpp:119-160, 291-302)
    tn.assign(3 * (size_t)nt, 0.0);
    vn.assign(3 * (size_t)nv, 0.0);
    std::vector<float> area(nt);
    for (int t = 0; t < nt; t++)
    {
        const double *p0 = &x[3 * (size_t)tri[3 * t]], *p1 = &x[3 * (size_t)tri[3 * t + 1]];
        const double *p2 = &x[3 * (size_t)tri[3 * t + 2]];
        double u[3], v[3], w[3];
        for (int a = 0; a < 3; a++)
        {
            u[a] = p1[a] - p0[a];
            v[a] = p2[a] - p1[a];
        }
        double *n = &tn[3 * (size_t)t];
        n[0] = u[1] * v[2] - u[2] * v[1];
        n[1] = u[2] * v[0] - u[0] * v[2];
        n[2] = u[0] * v[1] - u[1] * v[0];
        normalize4(n);
        for (int a = 0; a < 3; a++)
            v[a] = p2[a] - p0[a];
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
        area[t] = (float)(0.5 * std::sqrt(d3(w, w)));
    }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
        {
            double *v = &vn[3 * (size_t)tri[3 * t + k]];
            for (int a = 0; a < 3; a++)
                v[a] = v[a] + tn[3 * (size_t)t + a] * (double)area[t];
        }
    for (int v = 0; v < nv; v++)
        normalize4(&vn[3 * (size_t)v]);
    // each edge's side normal: the normalised sum of its triangles' normals
    // (BVHCollisionDetecter.cpp:565-576, 592-604)
    edge_n.assign(3 * (size_t)ne, 0.0);
    for (int e = 0; e < ne; e++)
    {
        double on[3];
        for (int a = 0; a < 3; a++)
            on[a] = tn[3 * (size_t)edge[4 * e + 2] + a];
        if (edge[4 * e + 3] != -1)
            for (int a = 0; a < 3; a++)
                on[a] = on[a] + tn[3 * (size_t)edge[4 * e + 3] + a];
        const double z = d3(on, on);
        if (z > 0)
        {
            const double n = std::sqrt(z);
            for (int a = 0; a < 3; a++)
                on[a] = on[a] / n;
        }
        for (int a = 0; a < 3; a++)
            edge_n[3 * (size_t)e + a] = on[a];
    }
    in_tri.assign(nv, 0);
    for (int v : tri)
        in_tri[v] = 1;
    x_dev = to_device(x);
    tri_dev = to_device(tri);
    edge_dev = to_device(edge);
    tn_dev = to_device(tn);
    vn_dev = to_device(vn);
    edge_n_dev = to_device(edge_n);
    in_tri_dev = to_device(in_tri);
    refresh_tiles();
}

Body::~Body()
{
    (void)hipFree(x_dev);
    (void)hipFree(tri_dev);
    (void)hipFree(edge_dev);
    (void)hipFree(tn_dev);
    (void)hipFree(vn_dev);
    (void)hipFree(edge_n_dev);
    (void)hipFree(in_tri_dev);
    (void)hipFree(tile_box_dev);
}

void Body::set_pos(const double *xyz)
{
    if (!xyz)
        throw std::invalid_argument("body: NULL positions");
    for (size_t k = 0; k < 3 * (size_t)nv; k++)
        if (!std::isfinite(xyz[k]))
            throw std::invalid_argument("body: non-finite position");
    x.assign(xyz, xyz + 3 * (size_t)nv);
    hcheck(hipMemcpy(x_dev, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice),
           "hipMemcpy body");
    refresh_tiles();
}

} // namespace bw
