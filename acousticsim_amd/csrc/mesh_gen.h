// Irregular cloth meshes: seeded point sets on a rectangle and their
// Delaunay triangulation (geometry_type "delaunay"; the reference's
// cTriangulator::DelaunayTriangulation, TriangulatorDelaunay.cpp:45-160,
// builds the same kind of mesh with an unseeded RNG and an O(V^2) dart test).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace bw
{

// n points in the width x height rectangle centred at the origin: the
// rectangle boundary points of cTriangulator::BuildRectangleBoundary
// (TriangulatorDelaunay.cpp:162-216: corners included, spacing
// sqrt(w h / n) along each side), then uniform random points (splitmix64
// stream of `seed`) kept when farther than sqrt(w h / (n pi)) from every
// point kept so far (TriangulatorDelaunay.cpp:62-80), or than min_spacing
// when that is positive; when the boundary alone holds n points or more,
// only the boundary (as the reference's loop adds nothing).  Coordinates are
// float values, as the reference stores them; side points lie exactly on
// the sides.  Throws if the rule cannot place n points.
std::vector<std::array<float, 2>> rectangle_points(float width, float height, int n,
                                                   uint64_t seed, float min_spacing = 0);

// Delaunay triangulation (Lawson's incremental edge flips, points inserted
// in a space-filling-curve order with a walking point location).
// Triangles are counter-clockwise.  Collinear points on the hull are kept.
std::vector<std::array<int, 3>> delaunay(const std::vector<std::array<float, 2>> &pts);

// The cloth mesh of geometry_type "delaunay": rectangle_points, their
// Delaunay triangulation, rest positions (x, y, 0) and triangles with the
// first and last vertex swapped, as the reference swaps every triangle whose
// normal is not +y (TriangulatorDelaunay.cpp:110-125; the planar start has
// normal -z after the swap).
void delaunay_cloth(float width, float height, int n, uint64_t seed, float min_spacing,
                    std::vector<double> &xyz, std::vector<std::array<int, 3>> &tri);

// Triangles of a Wavefront OBJ file as cObjUtil::LoadObj reads them through
// tinyobjloader 2.0.0 (ObjUtil.cpp:10-130): positions "v x y z" and texture
// coordinates "vt u v" stored as float values, faces "f" with 1-based or
// negative (relative) indices in the forms i, i/t, i//n, i/t/n.  Vertex ids
// are the OBJ position indices of the faces; every id up to the largest one
// used must appear in some face, and a vertex's texture coordinate is the
// one of its last face corner that names one (0, 0 if none).  Throws on
// faces with more than three corners (the reference accepts triangles
// only), on more than one shape ("o"/"g" groups holding faces), and on bad
// indices or files.
struct ObjTriangles
{
    std::vector<double> pos;               // 3 nv, float values
    std::vector<float> uv;                 // 2 nv
    std::vector<std::array<int, 3>> tri;   // nt
};
ObjTriangles read_obj_triangles(const std::string &path);

// Vertex order along a Hilbert curve over the points' bounding box
// (xy2: 2 floats per point); returns the permutation (order[k] = point).
std::vector<int> hilbert_order(const std::vector<float> &xy2);

} // namespace bw
