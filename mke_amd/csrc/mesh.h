// Host-side sphere mesh: generator (mke_mesh_sphere restatement) and loader
// (Mesh::load restatement).  Plain C++; no device code.
#pragma once
#include <string>
#include <vector>

namespace phg {

// Arrays as Mesh::load leaves them (src/mesh.h:354-427, src/mesh.cpp:323-519).
struct SphereMesh {
    int n_points = 0, n_triangles = 0, n_zones = 0;
    std::vector<double> pr;      // [V][3] global coordinates
    std::vector<double> coords;  // [V][n_zones][2] chart coordinates (x=lat, y=lon)
    std::vector<int> tri;        // [T][3] 0-based point ids, file order
    std::vector<int> tri_zone;   // [T] 0-based zone (file zone - 1)
    std::vector<int> adj_ptr;    // [V+1] point -> triangles, triangle-id order
    std::vector<int> adj;
    std::vector<int> is_boundary;  // [V]
    std::vector<int> inner, outer, p2io;
};

// Text written by `mke_mesh_sphere --type full --coord local --iter level`
// (utils/icosahedron.cpp:481-624).
std::string sphere_mesh_text(int level);

// Mesh::load (src/mesh.cpp:323-519) of P1 meshes.  Returns false and sets err
// ("bad file format, line = N") on malformed input, like the reference.
bool load_mesh_text(const char* text, size_t len, SphereMesh& m, std::string& err);

// Binary image of a loaded mesh (every array above, little-endian, with a
// checksum): written once, read back in a fraction of the text path's time.
bool save_mesh_binary(const SphereMesh& m, const char* path, std::string& err);
bool load_mesh_binary(const char* path, SphereMesh& m, std::string& err);

// print_function (src/phelm.cpp:248-309): "# points", one line per point
// (x y z f u v, %.16f) with u, v the zone-0 chart coordinates and x, y, z
// either u, v, 0 (no coordinate maps) or the unit-sphere map cos u cos v,
// cos u sin v, sin u (sphere_xyz); "# triangles" (1-based); "# end".
std::string print_function_text(const SphereMesh& m, const double* u, bool sphere_xyz);

}  // namespace phg
