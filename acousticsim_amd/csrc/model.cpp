#include "model.h"
#include "mesh_gen.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <stdexcept>

namespace bw
{

// ---------------------------------------------------------------- config --

ClothConfig ClothConfig::from_json(const Json &j)
{
    ClothConfig c;
    c.geometry_type = j.at("geometry_type").as_string();
    if (c.geometry_type != "regular_triangle" && c.geometry_type != "delaunay" &&
        c.geometry_type != "mesh")
        throw std::runtime_error("config: geometry_type \"" + c.geometry_type +
                                 "\" is not supported (regular_triangle, delaunay, mesh)");
    c.delaunay_seed = (uint64_t)j.get_double("delaunay_seed", 0);
    c.delaunay_min_spacing = j.get_double("delaunay_min_spacing", 0);
    if (!(c.delaunay_min_spacing >= 0))
        throw std::runtime_error("config: delaunay_min_spacing must be >= 0");
    c.pin_top_edge = j.has("pin_top_edge") && j.at("pin_top_edge").as_bool();
    c.from_obj = j.has("enable_cloth_from_obj") && j.at("enable_cloth_from_obj").as_bool();
    if (c.from_obj)
    {
        c.obj_path = j.at("cloth_obj_path").as_string();
        c.relocate_obj = j.has("enable_relocate_obj") && j.at("enable_relocate_obj").as_bool();
        if (j.has("cloth_obj_translation"))
        {
            auto t = j.at("cloth_obj_translation").as_vector();
            for (size_t k = 0; k < 3 && k < t.size(); k++)
                c.obj_translation[k] = t[k];
        }
    }
    // sizes come with the mesh arrays or the OBJ file
    const bool mesh = c.geometry_type == "mesh" || c.from_obj;
    auto size = mesh && !j.has("cloth_size") ? std::vector<double>{1.0, 1.0}
                                             : j.at("cloth_size").as_vector();
    auto sub = mesh && !j.has("subdivision") ? std::vector<double>{1.0, 1.0}
                                             : j.at("subdivision").as_vector();
    if (size.size() < 2 || sub.size() < 2)
        throw std::runtime_error("config: cloth_size and subdivision need 2 values");
    c.cloth_size[0] = size[0];
    c.cloth_size[1] = size[1];
    // range-checked as doubles before the cast (a cast of an out-of-range
    // double is undefined); a grid's vertex, triangle and edge counts must
    // fit in int, which 2^28 cells bound (edges < 3 cells + sides + 1)
    for (int k = 0; k < 2; k++)
        if (!(sub[k] >= 1.0 && sub[k] <= (double)(1 << 28)))
            throw std::runtime_error("config: subdivision must be in [1, 2^28]");
    c.subdivision[0] = (int)sub[0];
    c.subdivision[1] = (int)sub[1];
    if ((long long)c.subdivision[0] * c.subdivision[1] > (1LL << 28))
        throw std::runtime_error("config: subdivision[0] * subdivision[1] must be at most 2^28");
    if (!(size[0] > 0) || !(size[1] > 0) || !std::isfinite(size[0]) || !std::isfinite(size[1]))
        throw std::runtime_error("config: cloth_size must be positive and finite");
    if (j.has("cloth_init_pos"))
    {
        auto p = j.at("cloth_init_pos").as_vector();
        for (size_t k = 0; k < 3 && k < p.size(); k++)
            c.init_pos[k] = p[k];
    }
    if (j.has("cloth_init_orientation"))
    {
        auto p = j.at("cloth_init_orientation").as_vector();
        for (size_t k = 0; k < 3 && k < p.size(); k++)
            c.init_euler[k] = p[k];
    }
    c.uv_rotation_deg = (float)j.get_double("cloth_uv_rotation", 0);
    int rot = (int)c.uv_rotation_deg;
    if (rot != 0 && rot != 45 && rot != 90)
        throw std::runtime_error("config: cloth_uv_rotation must be 0, 45 or 90");
    c.density = (float)j.at("cloth_density").as_double();
    auto sk = j.at("cloth_stretch_stiffness").as_vector();
    if (sk.size() < 3)
        throw std::runtime_error("config: cloth_stretch_stiffness needs 3 values");
    for (int k = 0; k < 3; k++)
        c.stretch_k[k] = (float)sk[k];
    auto bk = j.at("cloth_bending_stiffness").as_vector();
    if (bk.empty())
        throw std::runtime_error("config: cloth_bending_stiffness is empty");
    c.bending_k = (float)bk[0];
    c.bending_uv[0] = (float)bk[0];
    c.bending_uv[1] = bk.size() > 1 ? (float)bk[1] : (float)bk[0];
    c.rayleigh_a = (float)j.at("rayleigh_damping_a").as_double();
    c.rayleigh_b = (float)j.at("rayleigh_damping_b").as_double();
    c.dt = j.at("default_timestep").as_double();
    if (!(c.dt > 0))
        throw std::runtime_error("config: default_timestep must be positive");
    c.gravity = j.at("enable_gravity").as_double() != 0;
    c.max_iters = j.at("max_iters").as_int();
    if (j.has("constraints") && j.at("constraints").has("static_point"))
    {
        for (auto &p : j.at("constraints").at("static_point").arr)
        {
            auto v = p.as_vector();
            if (v.size() != 2)
                throw std::runtime_error("config: static_point entries need 2 values");
            c.pin_uv.push_back({(float)v[0], (float)v[1]});
        }
    }
    std::string prec = j.get_string("precision", "mixed");
    if (prec == "mixed")
        c.precision = Precision::Mixed;
    else if (prec == "fp64")
        c.precision = Precision::FP64;
    else
        throw std::runtime_error("config: precision must be \"mixed\" or \"fp64\"");
    std::string mode = j.get_string("solver_mode", "tolerance");
    if (mode == "tolerance")
        c.solve_mode = SolveMode::Tolerance;
    else if (mode == "reference")
        c.solve_mode = SolveMode::Reference;
    else
        throw std::runtime_error("config: solver_mode must be \"tolerance\" or \"reference\"");
    c.tolerance = j.get_double("solver_tolerance", 1e-6);
    c.solver_max_iters = (int)j.get_double("solver_max_iters", 20000);
    if (!(c.tolerance >= 0) || c.solver_max_iters < 1)
        throw std::runtime_error("config: bad solver_tolerance / solver_max_iters");
    if (c.solve_mode == SolveMode::Reference && c.max_iters < 1)
        throw std::runtime_error("config: max_iters must be >= 1");
    return c;
}

// -------------------------------------------------------------- geometry --

namespace
{

// Regular grid with alternating diagonals (Triangulator.cpp:141-457).
// Vertex (i, j) -> id i*(hd+1)+j at (i*uw - width/2, j*uh - height/2).
// Cell (row, col) -> triangles 2*(row*hd+col)+{0,1}; the diagonal runs
// lu->rd when (row+col) is even, ru->ld otherwise.  Edge slots per row of
// cells: hd "top" slots, then left/skew/right interleaved (left of col c+1
// is the right of col c), then the next row's top slots are this row's
// bottoms.
// triangle -> edge ids (Triangulator.cpp:11-31): edge (tri[k], tri[k+1]) of
// every triangle, found through the edges listed by their smaller vertex
void triangle_edge_ids(ClothModel &m)
{
    std::vector<int> ptr(m.nv + 1, 0), hi(m.ne), id(m.ne);
    for (const Edge &e : m.edges)
        ptr[std::min(e.v0, e.v1) + 1]++;
    for (int v = 0; v < m.nv; v++)
        ptr[v + 1] += ptr[v];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < m.ne; e++)
    {
        const int lo = std::min(m.edges[e].v0, m.edges[e].v1);
        hi[fill[lo]] = std::max(m.edges[e].v0, m.edges[e].v1);
        id[fill[lo]++] = e;
    }
    m.tri_e.resize(m.nt);
    for (int t = 0; t < m.nt; t++)
        for (int k = 0; k < 3; k++)
        {
            const int a = m.tri[t][k], b = m.tri[t][(k + 1) % 3];
            const int lo = std::min(a, b), h = std::max(a, b);
            int found = -1;
            for (int q = ptr[lo]; q < ptr[lo + 1]; q++)
                if (hi[q] == h)
                    found = id[q];
            if (found < 0)
                throw std::runtime_error("config: triangle side is not an edge of the mesh");
            m.tri_e[t][k] = found;
        }
}

void apply_init_transform(const ClothConfig &c, ClothModel &m);

// Edges of a triangle list as cTriangulator::DelaunayTriangulation files
// them (add_edge_into_edge_array, TriangulatorDelaunay.cpp:9-43): sides
// (v0,v1), (v1,v2), (v2,v0) of each triangle in order, an edge numbered at
// its first appearance with {min, max} vertices, tri0 = that triangle, tri1
// = the second one holding it; boundary when there is none.
void edges_from_triangles(ClothModel &m)
{
    const size_t n3 = 3 * (size_t)m.nt;
    std::vector<std::pair<uint64_t, uint32_t>> rec(n3); // (vertex pair, side index 3t+k)
    for (int t = 0; t < m.nt; t++)
        for (int k = 0; k < 3; k++)
        {
            const int a = m.tri[t][k], b = m.tri[t][(k + 1) % 3];
            if (a == b)
                throw std::runtime_error("config: degenerate triangle in the mesh");
            const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            rec[3 * (size_t)t + k] = {key, (uint32_t)(3 * t + k)};
        }
    std::sort(rec.begin(), rec.end());
    // groups: first appearance -> edge; order edges by first appearance
    std::vector<std::pair<uint32_t, size_t>> first; // (side index, group start)
    for (size_t i = 0; i < n3;)
    {
        size_t j = i + 1;
        while (j < n3 && rec[j].first == rec[i].first)
            j++;
        if (j - i > 2)
            throw std::runtime_error("config: mesh edge shared by more than two triangles");
        first.push_back({rec[i].second, i});
        i = j;
    }
    std::sort(first.begin(), first.end());
    m.ne = (int)first.size();
    m.edges.resize(m.ne);
    for (int e = 0; e < m.ne; e++)
    {
        const size_t i = first[e].second;
        const uint64_t key = rec[i].first;
        Edge ed;
        ed.v0 = (int)(key >> 32);
        ed.v1 = (int)(key & 0xffffffffu);
        ed.t0 = (int)(rec[i].second / 3);
        const bool two = i + 1 < n3 && rec[i + 1].first == key;
        ed.t1 = two ? (int)(rec[i + 1].second / 3) : -1;
        ed.boundary = !two;
        m.edges[e] = ed;
    }
    triangle_edge_ids(m);
}

// geometry_type "delaunay" (cTriangulator::DelaunayTriangulation,
// TriangulatorDelaunay.cpp:45-160): subdivision[0] * subdivision[1] seeded
// points on the cloth rectangle, Delaunay-triangulated.  Like the reference,
// every triangle's first and last vertex are swapped when its normal is not
// +y (always, for the planar start), and UV = the point.
void make_delaunay(const ClothConfig &c, ClothModel &m)
{
    const float height = (float)c.cloth_size[0], width = (float)c.cloth_size[1];
    const long long want = (long long)c.subdivision[0] * c.subdivision[1];
    if (want > (1LL << 28))
        throw std::runtime_error("config: delaunay cloth too large");
    delaunay_cloth(width, height, (int)want, c.delaunay_seed, (float)c.delaunay_min_spacing,
                   m.xrest, m.tri);
    m.nv = (int)(m.xrest.size() / 3);
    m.nt = (int)m.tri.size();
    edges_from_triangles(m);
    apply_init_transform(c, m);
}

// Edges as cObjUtil::BuildEdge files them (ObjUtil.cpp:141-224): every
// side {min, max} of every triangle, tri0 = the first triangle holding it,
// tri1 = the second; edges numbered in ascending (min, max) order (the order
// of the std::map the reference keys them by).
void edges_sorted_pairs(ClothModel &m)
{
    const size_t n3 = 3 * (size_t)m.nt;
    std::vector<std::pair<uint64_t, uint32_t>> rec(n3);
    for (int t = 0; t < m.nt; t++)
        for (int k = 0; k < 3; k++)
        {
            const int a = m.tri[t][k], b = m.tri[t][(k + 1) % 3];
            if (a == b)
                throw std::runtime_error("config: degenerate triangle in the mesh");
            const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            rec[3 * (size_t)t + k] = {key, (uint32_t)(3 * t + k)};
        }
    std::sort(rec.begin(), rec.end());
    m.edges.clear();
    for (size_t i = 0; i < n3;)
    {
        size_t j = i + 1;
        while (j < n3 && rec[j].first == rec[i].first)
            j++;
        if (j - i > 2)
            throw std::runtime_error("config: mesh edge shared by more than two triangles");
        Edge ed;
        ed.v0 = (int)(rec[i].first >> 32);
        ed.v1 = (int)(rec[i].first & 0xffffffffu);
        ed.t0 = (int)(rec[i].second / 3);
        ed.t1 = j - i == 2 ? (int)(rec[i + 1].second / 3) : -1;
        ed.boundary = j - i == 1;
        m.edges.push_back(ed);
        i = j;
    }
    m.ne = (int)m.edges.size();
    triangle_edge_ids(m);
}

// enable_cloth_from_obj (BaseCloth.cpp:299-357): the OBJ's triangles; per
// vertex y += 0.2, uv *= 10/3 (in float), position += cloth_obj_translation;
// with enable_relocate_obj the top moves to y = 0.27 and x, z are centred
// on the running mean of the vertices within 1 cm of the top (MoveObjPos,
// BaseCloth.cpp:237-279).  No init transform or UV rotation, as in the
// reference's OBJ branch.
void make_obj(const ClothConfig &c, ClothModel &m)
{
    ObjTriangles obj = read_obj_triangles(c.obj_path);
    m.nv = (int)(obj.pos.size() / 3);
    m.nt = (int)obj.tri.size();
    m.tri = obj.tri;
    m.xrest = obj.pos;
    m.uv = obj.uv;
    const float uv_scale = (float)(10 / 3.0);
    for (int v = 0; v < m.nv; v++)
    {
        double *p = &m.xrest[3 * (size_t)v];
        p[1] += 0.2;
        m.uv[2 * (size_t)v] *= uv_scale;
        m.uv[2 * (size_t)v + 1] *= uv_scale;
        for (int a = 0; a < 3; a++)
            p[a] += c.obj_translation[a];
    }
    if (c.relocate_obj)
    {
        float highest = -1e3f;
        for (int v = 0; v < m.nv; v++)
            if (m.xrest[3 * (size_t)v + 1] > highest)
                highest = (float)m.xrest[3 * (size_t)v + 1];
        const double given = 0.27, shift = given - highest;
        for (int v = 0; v < m.nv; v++)
            m.xrest[3 * (size_t)v + 1] += shift;
        double o[4] = {0, 0, 0, 0};
        int count = 0;
        for (int v = 0; v < m.nv; v++)
        {
            const double *p = &m.xrest[3 * (size_t)v];
            if ((given - p[1]) < 0.01)
            {
                const double q[4] = {p[0], p[1], p[2], 1.0};
                for (int a = 0; a < 4; a++)
                    o[a] = (o[a] * count + q[a]) / (count + 1);
                count += 1;
            }
        }
        for (int v = 0; v < m.nv; v++)
        {
            m.xrest[3 * (size_t)v] -= o[0];
            m.xrest[3 * (size_t)v + 2] -= o[2];
        }
    }
    m.flat_xy.resize(2 * (size_t)m.nv);
    for (int v = 0; v < m.nv; v++)
    {
        m.flat_xy[2 * v] = (float)m.xrest[3 * (size_t)v];
        m.flat_xy[2 * v + 1] = (float)m.xrest[3 * (size_t)v + 1];
    }
    edges_sorted_pairs(m);
}

void make_grid(const ClothConfig &c, ClothModel &m)
{
    const double height = c.cloth_size[0], width = c.cloth_size[1];
    const int hd = c.subdivision[0], wd = c.subdivision[1];
    const double uh = height / hd, uw = width / wd;
    const int hl = hd + 1, wl = wd + 1;
    m.nv = wl * hl;
    m.nt = 2 * hd * wd;
    m.ne = wl * hd + hl * wd + hd * wd;
    m.xrest.assign(3 * (size_t)m.nv, 0.0);
    for (int i = 0; i < wl; i++)
        for (int j = 0; j < hl; j++)
        {
            double *p = &m.xrest[3 * (size_t)(i * hl + j)];
            p[0] = i * uw - width / 2;
            p[1] = j * uh - height / 2;
            p[2] = 0.0;
        }
    m.tri.resize(m.nt);
    m.edges.assign(m.ne, Edge{-1, -1, -1, -1, false});
    std::vector<char> set(m.ne, 0);
    auto put = [&](int id, int a, int b, int t0, int t1, bool boundary) {
        m.edges[id] = Edge{a, b, t0, t1, boundary};
        set[id] = 1;
    };
    const int stride = 3 * hd + 1; // edge slots per row of cells
    for (int row = 0; row < wd; row++)
        for (int col = 0; col < hd; col++)
        {
            const int lu = row * hl + col, ru = lu + 1, ld = lu + hl, rd = ld + 1;
            const bool even = ((row + col) & 1) == 0;
            const int t = 2 * (row * hd + col); // first triangle of the cell
            if (even)
            {
                m.tri[t] = {lu, ld, rd};
                m.tri[t + 1] = {lu, rd, ru};
            }
            else
            {
                m.tri[t] = {lu, ld, ru};
                m.tri[t + 1] = {ld, rd, ru};
            }
            const int top = row * stride + col;
            const int left = row * stride + hd + 2 * col;
            // diagonal: owned by the two triangles of this cell
            if (even)
                put(left + 1, lu, rd, t, t + 1, false);
            else
                put(left + 1, ru, ld, t, t + 1, false);
            // bottom: in the first triangle for even cells, the second for odd
            {
                const int own = even ? t : t + 1;
                const bool last = row == wd - 1;
                put(top + stride, ld, rd, own, last ? -1 : own + 2 * hd, last);
            }
            // right: second triangle of this cell and first of the next cell
            {
                const bool last = col == hd - 1;
                put(left + 2, ru, rd, t + 1, last ? -1 : t + 2, last);
            }
            if (row == 0)
                put(top, lu, ru, even ? t + 1 : t, -1, true);
            if (col == 0)
                put(left, lu, ld, t, -1, true);
        }
    for (int e = 0; e < m.ne; e++)
        if (!set[e])
            throw std::runtime_error("internal: grid edge slot not filled");

    triangle_edge_ids(m);
    apply_init_transform(c, m);
}

// init transform (cRotUtil::TransformMat, XYZ Euler: Rz*Ry*Rx, then
// translation) and UV = rot(uv_rotation) * float(untransformed xy)
// (Triangulator.cpp:116-131, 621-658); also keeps the untransformed xy
void apply_init_transform(const ClothConfig &c, ClothModel &m)
{
    m.flat_xy.resize(2 * (size_t)m.nv);
    for (int v = 0; v < m.nv; v++)
    {
        m.flat_xy[2 * v] = (float)m.xrest[3 * (size_t)v];
        m.flat_xy[2 * v + 1] = (float)m.xrest[3 * (size_t)v + 1];
    }
    const bool identity = c.init_pos[0] == 0 && c.init_pos[1] == 0 && c.init_pos[2] == 0 &&
                          c.init_euler[0] == 0 && c.init_euler[1] == 0 && c.init_euler[2] == 0;
    double R[3][3];
    {
        const double cx = std::cos(c.init_euler[0]), sx = std::sin(c.init_euler[0]);
        const double cy = std::cos(c.init_euler[1]), sy = std::sin(c.init_euler[1]);
        const double cz = std::cos(c.init_euler[2]), sz = std::sin(c.init_euler[2]);
        const double Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
        const double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
        const double Rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
        double T[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                T[a][b] = (Rz[a][0] * Ry[0][b] + Rz[a][1] * Ry[1][b]) + Rz[a][2] * Ry[2][b];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                R[a][b] = (T[a][0] * Rx[0][b] + T[a][1] * Rx[1][b]) + T[a][2] * Rx[2][b];
    }
    const float ang = (float)((double)c.uv_rotation_deg / 180.0 * M_PI);
    const float cr = (float)std::cos((double)ang), sr = (float)std::sin((double)ang);
    const float msr = -sr;
    m.uv.resize(2 * (size_t)m.nv);
    for (int v = 0; v < m.nv; v++)
    {
        double *p = &m.xrest[3 * (size_t)v];
        const float px = (float)p[0], py = (float)p[1];
        m.uv[2 * v] = cr * px + msr * py;
        m.uv[2 * v + 1] = sr * px + cr * py;
        if (!identity)
        {
            double o[3];
            for (int a = 0; a < 3; a++)
                o[a] = ((R[a][0] * p[0] + R[a][1] * p[1]) + R[a][2] * p[2]) + c.init_pos[a];
            p[0] = o[0];
            p[1] = o[1];
            p[2] = o[2];
        }
    }
}

// Eigen-order helpers used where the reference's rounding must be kept
inline double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// cMathUtil::CalcTriangleArea returns float (MathUtil2.cpp:344-357)
float triangle_area(const double *p0, const double *p1, const double *p2)
{
    const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                         a[0] * b[1] - a[1] * b[0]};
    return (float)(0.5 * std::sqrt(dot3(c, c)));
}

// CalcBaraffStretchCoef (BaraffMaterial.cpp:18-32): the inverse's entries
// are rounded to float before forming (-a-c, a, c) and (-b-d, b, d)
void stretch_coef(const double inv[2][2], float *cu, float *cv)
{
    const float a = (float)inv[0][0], b = (float)inv[0][1];
    const float c = (float)inv[1][0], d = (float)inv[1][1];
    cu[0] = -a - c;
    cu[1] = a;
    cu[2] = c;
    cv[0] = -b - d;
    cv[1] = b;
    cv[2] = d;
}

void rest_data(ClothModel &m)
{
    const ClothConfig &c = m.cfg;
    // lumped mass (cBaraffCloth::InitMass, BaraffCloth.cpp:209-236)
    m.area.resize(m.nt);
    m.mass.assign(m.nv, 0.0);
    for (int t = 0; t < m.nt; t++)
    {
        const auto &v = m.tri[t];
        const float A = triangle_area(&m.xrest[3 * (size_t)v[0]], &m.xrest[3 * (size_t)v[1]],
                                      &m.xrest[3 * (size_t)v[2]]);
        m.area[t] = A;
        const double add = (double)A / 3 * (double)c.density;
        for (int k = 0; k < 3; k++)
            m.mass[v[k]] = m.mass[v[k]] + add;
    }
    // stretch coefficients (cBaraffMaterial::Init, BaraffMaterial.cpp:33-76)
    m.coef.resize(12 * (size_t)m.nt);
    const double cs = std::cos(M_PI / 4), sn = std::sin(M_PI / 4);
    for (int t = 0; t < m.nt; t++)
    {
        const auto &v = m.tri[t];
        const float *u0 = &m.uv[2 * v[0]], *u1 = &m.uv[2 * v[1]], *u2 = &m.uv[2 * v[2]];
        const double D[2][2] = {{(double)(u1[0] - u0[0]), (double)(u2[0] - u0[0])},
                                {(double)(u1[1] - u0[1]), (double)(u2[1] - u0[1])}};
        const double invdet = 1.0 / (D[0][0] * D[1][1] - D[1][0] * D[0][1]);
        const double inv[2][2] = {{D[1][1] * invdet, -D[0][1] * invdet},
                                  {-D[1][0] * invdet, D[0][0] * invdet}};
        float *cf = &m.coef[12 * (size_t)t];
        stretch_coef(inv, cf, cf + 3);
        // Dm^-1 * R45^T with R45 = rot(pi/4)
        const double rt[2][2] = {{cs, sn}, {-sn, cs}};
        double r[2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                r[a][b] = inv[a][0] * rt[0][b] + inv[a][1] * rt[1][b];
        stretch_coef(r, cf + 6, cf + 9);
    }
    for (int k = 0; k < 3; k++)
        m.k[k] = (double)c.stretch_k[k];

    // dihedral rest data (cDihedralMaterial::Init, DihedralBending.cpp:23-89)
    std::vector<float> elen(m.ne), heron(m.nt);
    for (int e = 0; e < m.ne; e++)
    {
        const double *a = &m.xrest[3 * (size_t)m.edges[e].v0];
        const double *b = &m.xrest[3 * (size_t)m.edges[e].v1];
        const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        // 4-vector norm (w difference 0), 2-wide packet order
        elen[e] = (float)std::sqrt((dx * dx + dz * dz) + (dy * dy + 0.0 * 0.0));
    }
    for (int t = 0; t < m.nt; t++)
    {
        const double a = elen[m.tri_e[t][0]], b = elen[m.tri_e[t][1]], cc = elen[m.tri_e[t][2]];
        const double p = (a + b + cc) / 2;
        heron[t] = (float)std::sqrt(p * (p - a) * (p - b) * (p - cc));
    }
    const double kb = (double)c.bending_k;
    const double edge_k = ((kb + kb) + kb) / 3; // mean of a 3-vector of float(k)
    m.bend_factor.resize(m.interior.size());
    for (size_t ie = 0; ie < m.interior.size(); ie++)
    {
        const Edge &e = m.edges[m.interior[ie]];
        const float el = elen[m.interior[ie]];
        const float h0 = (2.0f * heron[e.t0]) / el, h1 = (2.0f * heron[e.t1]) / el;
        float hs = h0 + h1;
        if (hs < 1e-7f)
            hs = 1e-7f;
        m.bend_factor[ie] = -edge_k * (double)el / (double)hs;
    }
    // W = (1 + h a) M + h (b - h) K (BaraffCloth.cpp:274-275)
    m.sm = 1 + c.dt * (double)c.rayleigh_a;
    m.sk = c.dt * ((double)c.rayleigh_b - c.dt);
}

int other_vertex(const std::array<int, 3> &t, int a, int b)
{
    for (int v : t)
        if (v != a && v != b)
            return v;
    throw std::runtime_error("internal: degenerate triangle");
}

} // namespace

void ClothModel::rebuild_pins_from_list()
{
    pinned.assign(nv, 0);
    for (int p : pins)
        if (p >= 0 && p < nv)
            pinned[p] = 1;
}

void ClothModel::reference_ell(std::vector<int> &rp, std::vector<int> &cl) const
{
    std::vector<std::vector<int>> rows(nv);
    for (int v = 0; v < nv; v++)
        rows[v].push_back(v);
    for (int e : interior)
        for (int a : stencil[e])
            for (int b : stencil[e])
                rows[a].push_back(b);
    rp.assign(nv + 1, 0);
    cl.clear();
    for (int v = 0; v < nv; v++)
    {
        auto &r = rows[v];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        cl.insert(cl.end(), r.begin(), r.end());
        rp[v + 1] = (int)cl.size();
    }
}

// Per-vertex element lists, the block pattern (the reference ELL rule plus
// every triangle pair; the latter is inside the rule whenever each triangle
// has an interior edge) and the per-block contribution lists in ascending
// element order, built row by row with a column -> slot scatter array.
void ClothModel::build_blocks()
{
    const int nie = (int)interior.size();
    vt_ptr.assign(nv + 1, 0);
    ve_ptr.assign(nv + 1, 0);
    for (int t = 0; t < nt; t++)
        for (int i = 0; i < 3; i++)
            vt_ptr[tri[t][i] + 1]++;
    for (int ie = 0; ie < nie; ie++)
        for (int i = 0; i < 4; i++)
            ve_ptr[stencil[interior[ie]][i] + 1]++;
    for (int v = 0; v < nv; v++)
    {
        vt_ptr[v + 1] += vt_ptr[v];
        ve_ptr[v + 1] += ve_ptr[v];
    }
    vt_list.assign(vt_ptr[nv], 0);
    ve_list.assign(ve_ptr[nv], 0);
    {
        std::vector<int> tf(vt_ptr.begin(), vt_ptr.end() - 1);
        std::vector<int> ef(ve_ptr.begin(), ve_ptr.end() - 1);
        for (int t = 0; t < nt; t++)
            for (int i = 0; i < 3; i++)
                vt_list[tf[tri[t][i]]++] = (t << 2) | i;
        for (int ie = 0; ie < nie; ie++)
            for (int i = 0; i < 4; i++)
                ve_list[ef[stencil[interior[ie]][i]]++] = (ie << 2) | i;
    }
    // rows: {v} U incident interior-edge stencils U incident triangles
    std::vector<int> mark(nv, -1), row;
    row_ptr.assign(nv + 1, 0);
    cols.clear();
    cols.reserve(14 * (size_t)nv);
    diag_slot.resize(nv);
    auto add = [&](int v, int c) {
        if (mark[c] != v)
        {
            mark[c] = v;
            row.push_back(c);
        }
    };
    for (int v = 0; v < nv; v++)
    {
        row.clear();
        add(v, v);
        for (int q = ve_ptr[v]; q < ve_ptr[v + 1]; q++)
            for (int c : stencil[interior[ve_list[q] >> 2]])
                add(v, c);
        for (int q = vt_ptr[v]; q < vt_ptr[v + 1]; q++)
            for (int c : tri[vt_list[q] >> 2])
                add(v, c);
        std::sort(row.begin(), row.end());
        diag_slot[v] = (int)(std::lower_bound(row.begin(), row.end(), v) - row.begin());
        cols.insert(cols.end(), row.begin(), row.end());
        row_ptr[v + 1] = (int)cols.size();
    }
    // contribution lists: count, then fill, row by row; slot_of[c] is the
    // block (v, c) of the current row v (only columns of row v are read)
    const int nb = nnzb();
    std::vector<int> &slot_of = mark;
    s_ptr.assign(nb + 1, 0);
    b_ptr.assign(nb + 1, 0);
    std::vector<int> sfill, bfill;
    for (int pass = 0; pass < 2; pass++)
    {
        if (pass == 1)
        {
            for (int b = 0; b < nb; b++)
            {
                s_ptr[b + 1] += s_ptr[b];
                b_ptr[b + 1] += b_ptr[b];
            }
            s_list.assign(s_ptr[nb], 0);
            b_list.assign(b_ptr[nb], 0);
            sfill.assign(s_ptr.begin(), s_ptr.end() - 1);
            bfill.assign(b_ptr.begin(), b_ptr.end() - 1);
        }
        for (int v = 0; v < nv; v++)
        {
            for (int k = row_ptr[v]; k < row_ptr[v + 1]; k++)
                slot_of[cols[k]] = k;
            for (int q = vt_ptr[v]; q < vt_ptr[v + 1]; q++)
            {
                const int t = vt_list[q] >> 2, i = vt_list[q] & 3;
                for (int j = 0; j < 3; j++)
                {
                    const int b = slot_of[tri[t][j]];
                    if (pass == 0)
                        s_ptr[b + 1]++;
                    else
                        s_list[sfill[b]++] = (t << 4) | (i << 2) | j;
                }
            }
            for (int q = ve_ptr[v]; q < ve_ptr[v + 1]; q++)
            {
                const int ie = ve_list[q] >> 2, i = ve_list[q] & 3;
                const auto &st = stencil[interior[ie]];
                for (int j = 0; j < 4; j++)
                {
                    const int b = slot_of[st[j]];
                    if (pass == 0)
                        b_ptr[b + 1]++;
                    else
                        b_list[bfill[b]++] = (ie << 4) | (i << 2) | j;
                }
            }
        }
    }
}

ClothModel ClothModel::from_json_text(const std::string &text)
{
    Json j = Json::parse(text);
    if (j.kind != Json::Object)
        throw std::runtime_error("config: top level must be an object");
    if (j.has("cloth_type"))
    {
        const std::string t = j.at("cloth_type").as_string();
        if (t != "fem" && t != "fem_hip" && t != "fem_gpu")
            throw std::runtime_error("config: cloth_type \"" + t +
                                     "\" is not the implicit method (\"fem\", or \"fem_gpu\")");
        if (t == "fem_gpu")
        {
            ClothConfig c = ClothConfig::from_json(j);
            c.gpu_form = true;
            return build(c);
        }
    }
    return build(ClothConfig::from_json(j));
}

ClothModel ClothModel::build(const ClothConfig &cfg)
{
    ClothModel m;
    m.cfg = cfg;
    if (cfg.from_obj)
        make_obj(cfg, m);
    else if (cfg.geometry_type == "delaunay")
        make_delaunay(cfg, m);
    else if (cfg.geometry_type == "regular_triangle")
        make_grid(cfg, m);
    else
        throw std::runtime_error("config: geometry_type \"mesh\" needs the mesh arrays (bw_create_mesh)");
    m.finish();
    return m;
}

ClothModel ClothModel::build_mesh(const ClothConfig &cfg, const double *pos3, int nv,
                                  const int *tri3, int nt)
{
    if (nv < 3 || nt < 1 || !pos3 || !tri3)
        throw std::runtime_error("config: a mesh needs >= 3 vertices and >= 1 triangle");
    ClothModel m;
    m.cfg = cfg;
    m.nv = nv;
    m.nt = nt;
    m.xrest.assign(pos3, pos3 + 3 * (size_t)nv);
    m.tri.resize(nt);
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
        {
            const int v = tri3[3 * (size_t)t + k];
            if (v < 0 || v >= nv)
                throw std::runtime_error("config: mesh triangle vertex out of range");
            m.tri[t][k] = v;
        }
    edges_from_triangles(m);
    apply_init_transform(cfg, m);
    m.finish();
    return m;
}

void ClothModel::finish()
{
    ClothModel &m = *this;
    const ClothConfig &cfg = m.cfg;
    // bending stencils (BaseCloth.cpp:383-404): v2, v3 = the third vertices
    // of tri0 and tri1
    m.stencil.resize(m.ne);
    for (int e = 0; e < m.ne; e++)
    {
        const Edge &ed = m.edges[e];
        if (ed.boundary)
        {
            m.stencil[e] = {-1, -1, -1, -1};
            continue;
        }
        m.stencil[e] = {ed.v0, ed.v1, other_vertex(m.tri[ed.t0], ed.v0, ed.v1),
                        other_vertex(m.tri[ed.t1], ed.v0, ed.v1)};
        m.interior.push_back(e);
    }
    rest_data(m);

    // pins: nearest vertex in UV, first minimum wins (BaseCloth.cpp:151-189)
    for (auto &target : cfg.pin_uv)
    {
        double best = std::numeric_limits<double>::quiet_NaN();
        int id = -1;
        for (int v = 0; v < m.nv; v++)
        {
            const float du = m.uv[2 * v] - target[0], dv = m.uv[2 * v + 1] - target[1];
            const double dist = (double)std::sqrt(du * du + dv * dv);
            if (std::isnan(best) || dist < best)
            {
                best = dist;
                id = v;
            }
        }
        m.pins.push_back(id);
    }
    // pin_top_edge: every vertex on the top side (largest untransformed y)
    if (cfg.pin_top_edge)
    {
        float top = -3.4e38f;
        for (int v = 0; v < m.nv; v++)
            top = std::max(top, m.flat_xy[2 * v + 1]);
        for (int v = 0; v < m.nv; v++)
            if (m.flat_xy[2 * v + 1] == top)
                m.pins.push_back(v);
    }
    m.rebuild_pins_from_list();

    m.build_blocks();
    if (cfg.gpu_form)
        m.build_gpu_form();
}

namespace
{
// CalcCotangent (QBendingMaterial.cpp:11-19): dot products rounded to float,
// cot = AB.AC / sqrt(|AB|^2 |AC|^2 - (AB.AC)^2) in float
double cotangent(const double *A, const double *B, const double *C)
{
    double ab[3], ac[3];
    for (int k = 0; k < 3; k++)
    {
        ab[k] = B[k] - A[k];
        ac[k] = C[k] - A[k];
    }
    const float abac = (float)(ab[0] * ac[0] + ab[1] * ac[1] + ab[2] * ac[2]);
    const float abab = (float)(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
    const float acac = (float)(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
    return abac / std::sqrt(abab * acac - abac * abac);
}
} // namespace

// The GPU form's rest data (cBaraffClothGpu::Init, BaraffClothGpu.cpp:22-45,
// InitBendingMatrix :566-650, UpdateBendingMatrix :652-679; cBaseCloth::
// InitMass, BaseCloth.cpp:418-447).  Quadratic bending: per interior edge
// with stencil (v0, v1, v2, v3), coef = (c03 + c04, c01 + c02, -c01 - c03,
// -c02 - c04) from the clamped (1 degree) cotangents of the rest positions,
// base_ij = float(coef_i coef_j * float(3 / (A0 + A1))), edge stiffness
// K_e = Ku cos^2(theta) + Kv sin^2(theta) with theta in [0, pi/2] the edge's
// angle to the u axis, and block (v_i, v_j) += base_ij * K_e * I3, summed
// here in edge order (the reference adds with float atomics).
void ClothModel::build_gpu_form()
{
    const ClothConfig &c = cfg;
    std::vector<double> shared(nv, 0.0);
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++)
            shared[tri[t][k]] += (double)area[t] / 3;
    mass32.resize(nv);
    grav32.resize(nv);
    for (int v = 0; v < nv; v++)
    {
        const double mv = shared[v] * (double)c.density; // float mClothDensity
        mass32[v] = (float)mv;
        grav32[v] = (float)(-9.8 * mv); // InitGravity, BaraffClothGpu.cpp:556-567
    }
    qb.assign(cols.size(), 0.0f);
    const float threshold = (float)(1.0 / std::sin(1.0 / 180 * M_PI));
    const float Ku = c.bending_uv[0], Kv = c.bending_uv[1];
    for (int e = 0; e < ne; e++)
    {
        const Edge &ed = edges[e];
        if (ed.boundary)
            continue;
        const auto &sv = stencil[e];
        const double *x[4];
        for (int k = 0; k < 4; k++)
            x[k] = &xrest[3 * (size_t)sv[k]];
        double c01 = cotangent(x[0], x[1], x[2]), c02 = cotangent(x[0], x[1], x[3]);
        double c03 = cotangent(x[1], x[0], x[2]), c04 = cotangent(x[1], x[0], x[3]);
        for (double *cc : {&c01, &c02, &c03, &c04})
            *cc = std::min<double>(std::max<double>(*cc, -threshold), threshold);
        const double coef[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
        if (!(std::isfinite(coef[0]) && std::isfinite(coef[1]) && std::isfinite(coef[2]) &&
              std::isfinite(coef[3])))
            throw std::runtime_error("config: degenerate bending stencil (fem_gpu)");
        const float prefix = (float)(3.0 / (area[ed.t0] + area[ed.t1]));
        // edge direction in UV (tVector2f::normalized), angle to the u axis
        float du = uv[2 * (size_t)ed.v1] - uv[2 * (size_t)ed.v0];
        float dv = uv[2 * (size_t)ed.v1 + 1] - uv[2 * (size_t)ed.v0 + 1];
        const float n2 = du * du + dv * dv;
        if (n2 > 0)
        {
            const float n = std::sqrt(n2);
            du = du / n;
            dv = dv / n;
        }
        float theta = std::acos(du * 1.0f + dv * 0.0f);
        if (theta > (float)(M_PI / 2.0f))
            theta = (float)(M_PI - theta);
        const float ct = std::cos(theta), st = std::sin(theta);
        const float Ke = Ku * ct * ct + Kv * st * st;
        for (int i = 0; i < 4; i++)
        {
            const int r = sv[i];
            for (int j = 0; j < 4; j++)
            {
                const float base = (float)(coef[i] * coef[j] * prefix);
                const int col = sv[j];
                const int *b = cols.data() + row_ptr[r], *e2 = cols.data() + row_ptr[r + 1];
                const int *it = std::lower_bound(b, e2, col);
                if (it == e2 || *it != col)
                    throw std::runtime_error("config: bending block outside the ELL pattern");
                qb[it - cols.data()] += base * Ke;
            }
        }
    }
}

} // namespace bw
