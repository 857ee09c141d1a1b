#include "mesh_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <stdexcept>

namespace bw
{

namespace
{

// splitmix64 (Steele, Lea, Flood 2014): the k-th value of the stream of
// `seed` is mix(seed + k * golden)
struct SplitMix
{
    uint64_t state;
    double next01()
    {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

} // namespace

std::vector<std::array<float, 2>> rectangle_points(float width, float height, int n,
                                                   uint64_t seed, float min_spacing)
{
    if (!(width > 0) || !(height > 0) || n < 4)
        throw std::runtime_error("config: delaunay cloth needs a positive size and >= 4 points");
    std::vector<std::array<float, 2>> pts;
    // boundary (BuildRectangleBoundary): corners and sides, counter-clockwise
    // from the bottom-left corner's right neighbour
    const float density = std::sqrt(height * width / (float)n);
    const int nx = std::max((int)std::round(width / density), 2);
    const int ny = std::max((int)std::round(height / density), 2);
    const double left = -width / 2.0, right = width / 2.0, down = -height / 2.0, up = height / 2.0;
    auto px = [&](int i) -> float {
        return i == 0 ? (float)left : i == nx - 1 ? (float)right
                                                 : (float)(left + i * ((double)width / (nx - 1)));
    };
    auto py = [&](int i) -> float {
        return i == 0 ? (float)down : i == ny - 1 ? (float)up
                                                 : (float)(down + i * ((double)height / (ny - 1)));
    };
    for (int i = 1; i < nx; i++)
        pts.push_back({px(i), (float)down});
    for (int i = 1; i < ny; i++)
        pts.push_back({(float)right, py(i)});
    for (int i = nx - 2; i >= 0; i--)
        pts.push_back({px(i), (float)up});
    for (int i = ny - 2; i >= 0; i--)
        pts.push_back({(float)left, py(i)});
    if ((int)pts.size() >= n)
        return pts; // the boundary alone reaches n: the reference adds no darts either

    // darts, with a background grid of cell r/sqrt(2) (at most one kept
    // point per cell; a point within r of a dart is at most 2 cells away)
    const float thr = min_spacing > 0 ? min_spacing
                                      : std::sqrt((width * height / (float)n) / (float)M_PI);
    const double cell = thr / std::sqrt(2.0);
    const int gx = (int)std::ceil(width / cell) + 1, gy = (int)std::ceil(height / cell) + 1;
    std::vector<int> grid((size_t)gx * gy, -1);
    auto cell_of = [&](float x, float y, int &cx, int &cy) {
        cx = std::min(gx - 1, std::max(0, (int)((x - left) / cell)));
        cy = std::min(gy - 1, std::max(0, (int)((y - down) / cell)));
    };
    auto too_close = [&](float x, float y) {
        int cx, cy;
        cell_of(x, y, cx, cy);
        float best = 3.4e38f;
        for (int j = std::max(0, cy - 2); j <= std::min(gy - 1, cy + 2); j++)
            for (int i = std::max(0, cx - 2); i <= std::min(gx - 1, cx + 2); i++)
            {
                const int q = grid[(size_t)j * gx + i];
                if (q < 0)
                    continue;
                const float dx = pts[q][0] - x, dy = pts[q][1] - y;
                best = std::min(best, std::sqrt(dx * dx + dy * dy));
            }
        return !(best > thr);
    };
    auto put = [&](int q) {
        int cx, cy;
        cell_of(pts[q][0], pts[q][1], cx, cy);
        grid[(size_t)cy * gx + cx] = q;
    };
    for (int q = 0; q < (int)pts.size(); q++)
        put(q);
    SplitMix rng{seed};
    const long long max_tries = 2000LL * n + 100000;
    long long tries = 0;
    while ((int)pts.size() < n)
    {
        if (++tries > max_tries)
            throw std::runtime_error("config: could not place the delaunay cloth's points");
        const float x = (float)(left + rng.next01() * width);
        const float y = (float)(down + rng.next01() * height);
        if (too_close(x, y))
            continue;
        pts.push_back({x, y});
        put((int)pts.size() - 1);
    }
    return pts;
}

namespace
{

struct Tri
{
    int v[3]; // counter-clockwise
    int n[3]; // neighbour across the edge opposite v[i] (-1: none)
};

inline double orient(const double *a, const double *b, const double *c)
{
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
}

// > 0 when d lies inside the circumcircle of counter-clockwise (a, b, c)
inline double incircle(const double *a, const double *b, const double *c, const double *d)
{
    const double adx = a[0] - d[0], ady = a[1] - d[1];
    const double bdx = b[0] - d[0], bdy = b[1] - d[1];
    const double cdx = c[0] - d[0], cdy = c[1] - d[1];
    const double ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
    return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
}

// position along the Hilbert curve of (x, y) in [0, 2^16)^2
uint64_t hilbert(uint32_t x, uint32_t y)
{
    uint64_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1)
    {
        const uint32_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0)
        {
            if (rx == 1)
            {
                x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
                y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
            }
            std::swap(x, y);
        }
    }
    return d;
}

} // namespace

std::vector<int> hilbert_order(const std::vector<float> &xy2)
{
    const int n = (int)(xy2.size() / 2);
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 2; a++)
        {
            lo[a] = std::min(lo[a], (double)xy2[2 * i + a]);
            hi[a] = std::max(hi[a], (double)xy2[2 * i + a]);
        }
    const double D = std::max(hi[0] - lo[0], hi[1] - lo[1]) + 1e-30;
    std::vector<std::pair<uint64_t, int>> key(n);
    for (int i = 0; i < n; i++)
    {
        const uint32_t qx = (uint32_t)std::min(65535.0, (xy2[2 * i] - lo[0]) / D * 65535.0);
        const uint32_t qy = (uint32_t)std::min(65535.0, (xy2[2 * i + 1] - lo[1]) / D * 65535.0);
        key[i] = {hilbert(qx, qy), i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(n);
    for (int i = 0; i < n; i++)
        order[i] = key[i].second;
    return order;
}

std::vector<std::array<int, 3>> delaunay(const std::vector<std::array<float, 2>> &pts)
{
    const int n = (int)pts.size();
    if (n < 3)
        throw std::runtime_error("config: a triangulation needs 3 points");
    std::vector<double> P(2 * ((size_t)n + 3));
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 2; a++)
        {
            P[2 * i + a] = pts[i][a];
            lo[a] = std::min(lo[a], P[2 * i + a]);
            hi[a] = std::max(hi[a], P[2 * i + a]);
        }
    const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]);
    const double D = std::max(hi[0] - lo[0], hi[1] - lo[1]) + 1e-9;
    // super triangle, well outside the points
    const double sup[3][2] = {{cx - 8.3 * D, cy - 6.1 * D}, {cx + 8.7 * D, cy - 6.3 * D},
                              {cx + 0.1 * D, cy + 9.2 * D}};
    for (int k = 0; k < 3; k++)
    {
        P[2 * ((size_t)n + k)] = sup[k][0];
        P[2 * ((size_t)n + k) + 1] = sup[k][1];
    }
    auto pt = [&](int i) { return &P[2 * (size_t)i]; };

    std::vector<Tri> T;
    T.reserve(2 * (size_t)n + 8);
    T.push_back({{n, n + 1, n + 2}, {-1, -1, -1}});

    // insertion order along a Hilbert curve
    std::vector<std::pair<uint64_t, int>> order(n);
    for (int i = 0; i < n; i++)
    {
        const uint32_t qx = (uint32_t)std::min(65535.0, (P[2 * i] - lo[0]) / D * 65535.0);
        const uint32_t qy = (uint32_t)std::min(65535.0, (P[2 * i + 1] - lo[1]) / D * 65535.0);
        order[i] = {hilbert(qx, qy), i};
    }
    std::sort(order.begin(), order.end());

    // Lawson's incremental insertion: split the triangle (or the edge) that
    // holds the point, then flip the edges facing it while their opposite
    // vertex lies inside the new triangle's circumcircle.  Every step keeps
    // a valid triangulation, so rounding in the circle test can at worst
    // leave an edge unflipped; it cannot break the mesh.
    auto set_back = [&](int nb, int from, int to) {
        if (nb >= 0)
            for (int j = 0; j < 3; j++)
                if (T[nb].n[j] == from)
                    T[nb].n[j] = to;
    };
    std::vector<int> stack;
    // t = (a, b, p) with p at index 2; flip edge (a, b) if needed
    auto legalize = [&](int t) {
        const int u = T[t].n[2];
        if (u < 0)
            return;
        const int a = T[t].v[0], b = T[t].v[1], pv = T[t].v[2];
        int j = 0;
        while (j < 3 && (T[u].v[j] == a || T[u].v[j] == b))
            j++;
        const int q = T[u].v[j];
        if (!(incircle(pt(a), pt(b), pt(pv), pt(q)) > 0))
            return;
        const int n_qb = T[u].n[(j + 2) % 3], n_aq = T[u].n[(j + 1) % 3];
        const int n_bp = T[t].n[0], n_pa = T[t].n[1];
        T[t] = {{a, q, pv}, {u, n_pa, n_aq}};
        T[u] = {{q, b, pv}, {n_bp, t, n_qb}};
        set_back(n_bp, t, u);
        set_back(n_aq, u, t);
        stack.push_back(t);
        stack.push_back(u);
    };
    int last = 0;
    for (auto &o : order)
    {
        const int pi = o.second;
        const double *p = pt(pi);
        // walk to a triangle containing p
        int t = last;
        long long steps = 0;
        for (;;)
        {
            int next = -1;
            for (int i = 0; i < 3; i++)
            {
                const int a = T[t].v[(i + 1) % 3], b = T[t].v[(i + 2) % 3];
                if (orient(pt(a), pt(b), p) < 0)
                {
                    next = T[t].n[i];
                    break;
                }
            }
            if (next < 0)
                break;
            t = next;
            if (++steps > 4 * (long long)T.size() + 64)
                throw std::runtime_error("internal: delaunay point location did not end");
        }
        int on_edge = -1;
        for (int i = 0; i < 3; i++)
            if (orient(pt(T[t].v[(i + 1) % 3]), pt(T[t].v[(i + 2) % 3]), p) == 0)
                on_edge = i;
        stack.clear();
        if (on_edge < 0)
        {
            // split t = (a, b, c) into (a, b, p), (b, c, p), (c, a, p)
            const int a = T[t].v[0], b = T[t].v[1], c = T[t].v[2];
            const int n_bc = T[t].n[0], n_ca = T[t].n[1], n_ab = T[t].n[2];
            const int t1 = (int)T.size(), t2 = t1 + 1;
            T.push_back({{b, c, pi}, {t2, t, n_bc}});
            T.push_back({{c, a, pi}, {t, t1, n_ca}});
            T[t] = {{a, b, pi}, {t1, t2, n_ab}};
            set_back(n_bc, t, t1);
            set_back(n_ca, t, t2);
            stack.push_back(t);
            stack.push_back(t1);
            stack.push_back(t2);
        }
        else
        {
            // p on edge (a, b) of t = (c, a, b); u = (d, b, a) beyond it
            const int i = on_edge;
            const int c = T[t].v[i], a = T[t].v[(i + 1) % 3], b = T[t].v[(i + 2) % 3];
            const int n_ca = T[t].n[(i + 2) % 3], n_bc = T[t].n[(i + 1) % 3];
            const int u = T[t].n[i];
            const int t2 = (int)T.size();
            T.push_back({{0, 0, 0}, {-1, -1, -1}});
            int u2 = -1, n_db = -1, n_ad = -1, d = -1;
            if (u >= 0)
            {
                int j = 0;
                while (j < 3 && (T[u].v[j] == a || T[u].v[j] == b))
                    j++;
                d = T[u].v[j];
                n_db = T[u].n[(j + 2) % 3]; // edge (d, b) is opposite a
                n_ad = T[u].n[(j + 1) % 3]; // edge (a, d) is opposite b
                u2 = (int)T.size();
                T.push_back({{0, 0, 0}, {-1, -1, -1}});
            }
            T[t] = {{c, a, pi}, {u2, t2, n_ca}};
            T[t2] = {{b, c, pi}, {t, u >= 0 ? u : -1, n_bc}};
            set_back(n_bc, t, t2);
            stack.push_back(t);
            stack.push_back(t2);
            if (u >= 0)
            {
                T[u] = {{d, b, pi}, {t2, u2, n_db}};
                T[u2] = {{a, d, pi}, {u, t, n_ad}};
                set_back(n_ad, u, u2);
                stack.push_back(u);
                stack.push_back(u2);
            }
        }
        long long flips = 0;
        while (!stack.empty())
        {
            const int x = stack.back();
            stack.pop_back();
            if (T[x].v[2] != pi)
                continue;
            legalize(x);
            if (++flips > 100000)
                throw std::runtime_error("internal: delaunay edge flips did not end");
        }
        last = t;
    }
    std::vector<std::array<int, 3>> out;
    out.reserve(2 * (size_t)n);
    for (const Tri &t : T)
        if (t.v[0] < n && t.v[1] < n && t.v[2] < n)
            out.push_back({t.v[0], t.v[1], t.v[2]});
    return out;
}

void delaunay_cloth(float width, float height, int n, uint64_t seed, float min_spacing,
                    std::vector<double> &xyz, std::vector<std::array<int, 3>> &tri)
{
    const auto pts = rectangle_points(width, height, n, seed, min_spacing);
    const auto tris = delaunay(pts);
    xyz.assign(3 * pts.size(), 0.0);
    for (size_t v = 0; v < pts.size(); v++)
    {
        xyz[3 * v] = pts[v][0];
        xyz[3 * v + 1] = pts[v][1];
    }
    tri.resize(tris.size());
    for (size_t t = 0; t < tris.size(); t++)
        tri[t] = {tris[t][2], tris[t][1], tris[t][0]};
}

namespace
{
// next whitespace-separated token of line from position i
bool next_token(const std::string &line, size_t &i, std::string &tok)
{
    while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r'))
        i++;
    const size_t b = i;
    while (i < line.size() && line[i] != ' ' && line[i] != '\t' && line[i] != '\r')
        i++;
    tok = line.substr(b, i - b);
    return !tok.empty();
}

float obj_real(const std::string &tok, const std::string &path)
{
    char *end = nullptr;
    const double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str())
        throw std::runtime_error("config: bad number \"" + tok + "\" in " + path);
    return (float)v;
}

// tinyobjloader's fixIndex: 1-based, negative = relative to the count so far
int obj_index(const std::string &tok, int count, const std::string &path)
{
    char *end = nullptr;
    const long v = std::strtol(tok.c_str(), &end, 10);
    if (end == tok.c_str() || v == 0)
        throw std::runtime_error("config: bad face index \"" + tok + "\" in " + path);
    const long r = v > 0 ? v - 1 : count + v;
    if (r < 0 || r >= count)
        throw std::runtime_error("config: face index out of range in " + path);
    return (int)r;
}
} // namespace

ObjTriangles read_obj_triangles(const std::string &path)
{
    std::ifstream in(path);
    if (!in)
        throw std::runtime_error("config: cannot open cloth_obj_path " + path);
    std::vector<float> v, vt;
    struct Corner
    {
        int v, t;
    };
    std::vector<std::array<Corner, 3>> faces;
    int shapes = 0;
    size_t faces_at_group = 0;
    std::string line, key, tok;
    while (std::getline(in, line))
    {
        size_t i = 0;
        if (!next_token(line, i, key) || key[0] == '#')
            continue;
        if (key == "v")
        {
            for (int a = 0; a < 3; a++)
            {
                if (!next_token(line, i, tok))
                    throw std::runtime_error("config: short vertex line in " + path);
                v.push_back(obj_real(tok, path));
            }
        }
        else if (key == "vt")
        {
            for (int a = 0; a < 2; a++)
                vt.push_back(next_token(line, i, tok) ? obj_real(tok, path) : 0.0f);
        }
        else if (key == "o" || key == "g")
        {
            if (faces.size() > faces_at_group)
                shapes++;
            faces_at_group = faces.size();
        }
        else if (key == "f")
        {
            std::vector<Corner> cs;
            while (next_token(line, i, tok))
            {
                const size_t s1 = tok.find('/');
                Corner c{obj_index(tok.substr(0, s1), (int)(v.size() / 3), path), -1};
                if (s1 != std::string::npos)
                {
                    const size_t s2 = tok.find('/', s1 + 1);
                    const std::string t = tok.substr(s1 + 1, s2 == std::string::npos
                                                                 ? std::string::npos
                                                                 : s2 - s1 - 1);
                    if (!t.empty())
                        c.t = obj_index(t, (int)(vt.size() / 2), path);
                }
                cs.push_back(c);
            }
            if (cs.size() < 3)
                continue; // degenerate face, skipped as tinyobjloader does
            if (cs.size() > 3)
                throw std::runtime_error("config: " + path +
                                         " has a face with more than 3 vertices (triangles only)");
            faces.push_back({cs[0], cs[1], cs[2]});
        }
    }
    if (faces.size() > faces_at_group)
        shapes++;
    if (shapes != 1)
        throw std::runtime_error("config: " + path + " must hold exactly one shape of faces");
    ObjTriangles out;
    int nv = 0;
    for (auto &f : faces)
        for (auto &c : f)
            nv = std::max(nv, c.v + 1);
    std::vector<char> used(nv, 0);
    out.pos.assign(3 * (size_t)nv, 0.0);
    out.uv.assign(2 * (size_t)nv, 0.0f);
    out.tri.reserve(faces.size());
    for (auto &f : faces)
    {
        for (auto &c : f)
        {
            if (!used[c.v])
                for (int a = 0; a < 3; a++)
                    out.pos[3 * (size_t)c.v + a] = v[3 * (size_t)c.v + a];
            used[c.v] = 1;
            if (c.t >= 0)
            {
                out.uv[2 * (size_t)c.v] = vt[2 * (size_t)c.t];
                out.uv[2 * (size_t)c.v + 1] = vt[2 * (size_t)c.t + 1];
            }
        }
        out.tri.push_back({f[0].v, f[1].v, f[2].v});
    }
    for (int k = 0; k < nv; k++)
        if (!used[k])
            throw std::runtime_error("config: vertex " + std::to_string(k) + " of " + path +
                                     " is in no face");
    return out;
}

} // namespace bw
