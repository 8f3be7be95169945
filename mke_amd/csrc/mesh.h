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

}  // namespace phg
