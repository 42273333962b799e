// Mesh of the reference driver (include/mesh.hpp:20-129) backed by a gfx950
// context.  The AoS Vertex / Edge / Cell arrays keep the reference's member
// names so code written against them compiles unchanged; the unordered_set
// adjacency becomes sorted vectors (ascending ids).  The authoritative copy of
// the state lives on the device: operators push the host cell state first and
// write their side effects (R, gradients, face states) back into the arrays.
#ifndef O2F_HOST_MESH_HPP
#define O2F_HOST_MESH_HPP

#include <cstddef>
#include <string>
#include <vector>

#include "o2f.h"
#include "params.hpp"

namespace navier_stokes {

struct vector_3D {
    double x, y, z;
};

struct Vertex {
    vector_3D p;
    std::vector<std::size_t> edgeIDs;
    std::vector<std::size_t> cellIDs;
    char boundaryCondition = '\0';
};

struct Edge {
    vector_3D centre;
    std::size_t leftVertexID, rightVertexID;
    std::size_t leftCellID = static_cast<std::size_t>(-1), rightCellID = static_cast<std::size_t>(-1);
    char boundaryCondition = '\0';
    double edgeLength;
    vector_3D edgeNormal;  // unit normal pointing into the left cell
    double rho_L, rho_u_L, rho_v_L, E_L;
    double rho_R, rho_u_R, rho_v_R, E_R;
    double tau_xx = 0.0, tau_xy_yx = 0.0, tau_yy = 0.0;
};

struct Cell {
    vector_3D centre;
    std::vector<std::size_t> verticeIDs;
    std::vector<std::size_t> edgeIDs;
    std::vector<std::size_t> neighbourIDs;
    bool border = false;
    double cellVolume;
    double rho, rho_u, rho_v, E;
    double rho_x, rho_u_x, rho_v_x, E_x, rho_y, rho_u_y, rho_v_y, E_y;
    double R_rho, R_rho_u, R_rho_v, R_E;
    double rho_old, rho_u_old, rho_v_old, E_old;
    double scalar;
};

struct Mesh {
    std::vector<Vertex> vertices;
    std::vector<Edge> edges;
    std::vector<Cell> cells;
    double x_min, x_max, y_min, y_max;

    // Reads params.mesh_name (MSH 2.2) and creates the device context
    // (Mesh::Mesh, src/mesh.cpp:7-283).  Throws std::runtime_error.
    explicit Mesh(const SimulationParams &params);
    ~Mesh();
    Mesh(const Mesh &) = delete;
    Mesh &operator=(const Mesh &) = delete;

    o2f_ctx *ctx() const { return ctx_; }
    // mesh_reorder != "none": this mesh's cell i is the mesh file's cell
    // cell_perm[i] (empty: the file's numbering)
    const std::vector<int> &cell_perm() const { return cell_perm_; }
    // a state [4][n_cells] from the file's cell order to this mesh's, and back
    void state_from_file_order(std::vector<double> &U) const;
    void state_to_file_order(std::vector<double> &U) const;
    void set_params(const SimulationParams &params);  // hot reload
    // host cell state -> device, device -> host cell state
    void push_state();
    void pull_state();
    // device side effects -> host arrays
    void store_residual(const double *R);  // 4*n_cells SoA into Cell::R_*
    void pull_gradients();
    void pull_faces();
    // throw std::runtime_error / std::invalid_argument for a failed C ABI call
    void check(int status) const;

   private:
    o2f_ctx *ctx_ = nullptr;
    int tile_cells_ = -1;  // the tile setting last given to the context
    std::vector<int> cell_perm_;
};

}  // namespace navier_stokes

#endif
