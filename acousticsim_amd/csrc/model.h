// Host-side cloth model: configuration, mesh topology, rest data, the block
// (ELL) column structure and the per-block contribution lists that let the
// device assemble K and f by gathering in a fixed order (no atomics).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "json_lite.h"

namespace bw
{

enum class Precision
{
    Mixed,
    FP64
};
enum class SolveMode
{
    Tolerance,
    Reference
};

// Everything cBaraffCloth::Init reads (BaraffCloth.cpp:22-83,
// BaseCloth.cpp:44-55, Triangulator.cpp:32-139), plus solver keys.
struct ClothConfig
{
    std::string geometry_type = "regular_triangle";
    double cloth_size[2] = {0.5, 0.5}; // [height (y), width (x)]
    int subdivision[2] = {20, 20};     // [height_div, width_div]
    double init_pos[3] = {0, 0, 0};
    double init_euler[3] = {0, 0, 0};
    float uv_rotation_deg = 0;
    float density = 0.2f;    // float member (BaseCloth.h:53)
    float stretch_k[3] = {}; // tVector3f (BaraffCloth.h:25)
    float bending_k = 0;     // [0] of the bending vector (BaraffCloth.cpp:38-41)
    float bending_uv[2] = {0, 0}; // warp, weft bending of the GPU form (BaraffClothGpu.cpp:25-28)
    // cloth_type "fem_gpu": the reference's GPU form (cBaraffClothGpu) as an
    // A/B comparison mode
    bool gpu_form = false;
    float rayleigh_a = 0, rayleigh_b = 0; // float members (BaraffCloth.h:27)
    double dt = 0.01;
    bool gravity = true;
    int max_iters = 500;
    std::vector<std::array<float, 2>> pin_uv; // constraints.static_point
    uint64_t delaunay_seed = 0; // "delaunay_seed" (this library's key)
    double delaunay_min_spacing = 0; // "delaunay_min_spacing"; 0: the reference's rule
    bool pin_top_edge = false;  // "pin_top_edge" (this library's key)
    // OBJ cloth (BaseCloth.cpp:299-357): "enable_cloth_from_obj",
    // "cloth_obj_path", "enable_relocate_obj", "cloth_obj_translation"
    bool from_obj = false;
    std::string obj_path;
    bool relocate_obj = false;
    double obj_translation[3] = {0, 0, 0};
    Precision precision = Precision::Mixed;
    SolveMode solve_mode = SolveMode::Tolerance;
    double tolerance = 1e-6;
    int solver_max_iters = 20000;

    static ClothConfig from_json(const Json &j);
};

struct Edge
{
    int v0, v1, t0, t1;
    bool boundary;
};

struct ClothModel
{
    ClothConfig cfg;
    int nv = 0, nt = 0, ne = 0;
    std::vector<double> xrest;              // 3 nv
    std::vector<float> uv;                  // 2 nv
    std::vector<float> flat_xy;             // 2 nv, untransformed rest xy
    std::vector<std::array<int, 3>> tri;    // vertex ids
    std::vector<std::array<int, 3>> tri_e;  // edge ids (id0,id1),(id1,id2),(id2,id0)
    std::vector<Edge> edges;
    std::vector<std::array<int, 4>> stencil; // -1 on boundary edges
    std::vector<int> interior;               // interior edge ids, ascending

    // rest data
    std::vector<double> mass;  // nv
    std::vector<float> coef;   // 12 nt: cu, cv (warp/weft), cu, cv (bias)
    std::vector<float> area;   // nt
    std::vector<double> bend_factor; // per interior edge: -k_e * e / (h0 + h1)
    double k[3] = {};          // stiffness as double(float)
    double sk = 0, sm = 0;     // W = sm M + sk K

    std::vector<int> pins; // static vertices (may repeat)
    std::vector<uint8_t> pinned;

    // GPU form (cloth_type "fem_gpu"): constant quadratic-bending matrix as
    // one scalar per block (the block is scalar * I3), FP32 masses and the
    // pinned vertices' FP32 target positions
    std::vector<float> qb;     // nnzb, block order of cols
    std::vector<float> mass32; // nv
    std::vector<float> grav32; // nv: y gravity force float(-9.8 m)
    void build_gpu_form();

    // block structure, reference numbering
    std::vector<int> row_ptr, cols, diag_slot;
    // per-block contribution lists in ascending element order.  Stretch:
    // (triangle << 4) | (i << 2) | j ; bending: (interior edge << 4) | (i << 2) | j
    std::vector<int> s_ptr, s_list, b_ptr, b_list;
    // per-vertex force lists: (triangle << 2) | i ; (interior edge << 2) | i
    std::vector<int> vt_ptr, vt_list, ve_ptr, ve_list;

    static ClothModel build(const ClothConfig &cfg);
    // a cloth on a given triangle mesh (rest positions 3 nv, triangles
    // 3 nt); edges, UV (= rest xy) and rest data as for geometry "delaunay"
    static ClothModel build_mesh(const ClothConfig &cfg, const double *pos3, int nv,
                                 const int *tri3, int nt);
    // stencils, rest data, pins and blocks, once the topology is in place
    void finish();
    // parse the cloth JSON (cloth_type, when present, must be "fem" or
    // "fem_hip") and build
    static ClothModel from_json_text(const std::string &text);
    void rebuild_pins_from_list();
    // element lists, block pattern and contribution lists from the topology
    void build_blocks();
    int nnzb() const { return (int)cols.size(); }
    // the reference GPU ELL rule (BaraffClothGpu.cpp:438-500)
    void reference_ell(std::vector<int> &rp, std::vector<int> &cl) const;
};

} // namespace bw
