// Kinematic bodies and the cloth-body collision detector of the plain method
// (cBVHCollisionDetecter, BVHCollisionDetecter.cpp; cKinematicBody,
// KinematicBody.cpp).
#pragma once
#include <vector>

#include <hip/hip_runtime.h>

namespace bw
{

// body primitives per LDS tile of the detection kernels
constexpr int DETECT_TILE = 256;

// A kinematic body: a triangle mesh whose edges are numbered as
// cObjUtil::BuildEdge numbers them (ascending {min, max}, tri0 the first
// triangle with the side, ObjUtil.cpp:141-224) and whose triangle and vertex
// normals are fixed at creation (cKinematicBody::Init, KinematicBody.cpp:
// 114-116: CalcTriangleInitArea, UpdateTriangleNormal,
// UpdateVertexNormalFromTriangleNormal).  Positions may move afterwards.
class Body
{
public:
    Body(const double *xyz, int nv, const int *tri, int nt);
    ~Body();
    Body(const Body &) = delete;
    Body &operator=(const Body &) = delete;
    void set_pos(const double *xyz);
    int tiles(int n) const { return (n + DETECT_TILE - 1) / DETECT_TILE; }

    int nv = 0, nt = 0, ne = 0;
    std::vector<double> x;      // 3 nv
    std::vector<int> tri;       // 3 nt
    std::vector<int> edge;      // 4 ne: id0, id1, tri0, tri1 (-1)
    std::vector<double> tn, vn; // 3 nt, 3 nv
    std::vector<double> edge_n; // 3 ne: normalised sum of the edge's triangle normals
    std::vector<unsigned char> in_tri; // per vertex: used by a triangle
    // boxes (lo xyz, hi xyz) of each tile of DETECT_TILE consecutive
    // triangles, vertices and edges, refreshed with the positions
    std::vector<double> tile_box_tri, tile_box_vert, tile_box_edge;
    double *tile_box_dev = nullptr; // the three lists back to back
    // device copies
    double *x_dev = nullptr, *tn_dev = nullptr, *vn_dev = nullptr, *edge_n_dev = nullptr;
    int *tri_dev = nullptr, *edge_dev = nullptr;
    unsigned char *in_tri_dev = nullptr;

private:
    void refresh_tiles();
};

// one detected pair as the kernels write it
struct Hit
{
    double n[3];
    float bary[3];
    float depth;
    int flag; // obj0_is_cloth_obj1_is_rb
    int key0, key1;
    int pad;
};

// Detection kernels (collide.hip).  Each runs in two passes: counts (out ==
// nullptr) then writes at the given offsets, one thread per cloth vertex,
// cloth triangle or cloth edge, looping over the body's triangles, vertices
// or edges in tiles.
struct ClothView
{
    int nv, nt, ne;
    int vbase, tbase;      // first global vertex / triangle
    const int *vrow;       // global vertex -> row
    const int *vt_ptr;     // global vertex -> triangle list (non-empty: used)
    const int *tri_rows;   // 3 rows per global triangle
    const int *edge_rows;  // 2 rows per cloth edge of this cloth
    const double *X;       // positions by row
    const double *tn;      // triangle normals by global triangle
};
void launch_detect_points(const ClothView &c, const Body &b, int *count, const int *offset,
                          Hit *out, hipStream_t s);
void launch_detect_body_points(const ClothView &c, const Body &b, int *count, const int *offset,
                               Hit *out, hipStream_t s);
void launch_detect_edges(const ClothView &c, const Body &b, int cloth_first, int *count,
                         const int *offset, Hit *out, hipStream_t s);

} // namespace bw
