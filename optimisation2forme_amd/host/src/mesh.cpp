#include "mesh.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>

namespace navier_stokes {

void Mesh::check(int status) const {
    if (status == O2F_OK) return;
    const std::string msg = ctx_ ? o2f_last_error(ctx_) : o2f_global_error();
    if (status == O2F_ERR_BC) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

Mesh::Mesh(const SimulationParams &params) {
    o2f_mesh_desc d{};
    check(o2f_msh2_read(params.mesh_name.c_str(), &d));
    struct Free {
        o2f_mesh_desc *d;
        ~Free() { o2f_msh2_free(d); }
    } guard{&d};
    if (params.mesh_reorder != "none") {
        int method;
        if (params.mesh_reorder == "rcm")
            method = O2F_REORDER_RCM;
        else if (params.mesh_reorder == "rcm_xcd")
            method = O2F_REORDER_RCM_XCD;
        else
            throw std::invalid_argument("Unknown mesh_reorder: " + params.mesh_reorder);
        o2f_mesh_desc r{};
        cell_perm_.resize(d.n_cells);
        check(o2f_mesh_reorder(&d, method, &r, cell_perm_.data(), nullptr));
        o2f_msh2_free(&d);
        d = r;  // freed by the guard
    }
    if (params.verbosity > 0) std::printf("\nNombre de Noeuds : %d\n", d.n_vertices);
    int ne = 0;
    check(o2f_mesh_topology(&d, &ne, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    const int nc = d.n_cells, nv = d.n_vertices;
    std::vector<int> ev(2 * ne), ec(2 * ne), ce(4 * nc);
    std::vector<char> eb(ne);
    std::vector<double> eg(5 * ne), cg(3 * nc);
    check(o2f_mesh_topology(&d, &ne, ev.data(), ec.data(), eb.data(), eg.data(), cg.data(), ce.data()));
    const o2f_params op = to_o2f_params(params);
    check(o2f_ctx_create(&d, &op, &ctx_));
    check(o2f_set_tile_cells(ctx_, params.tile_cells));
    tile_cells_ = params.tile_cells;

    vertices.resize(nv);
    x_min = y_min = std::numeric_limits<double>::infinity();
    x_max = y_max = -std::numeric_limits<double>::infinity();
    for (int i = 0; i < nv; ++i) {
        vertices[i].p = {d.xy[2 * i], d.xy[2 * i + 1], 0.0};
        x_min = std::min(x_min, d.xy[2 * i]);
        x_max = std::max(x_max, d.xy[2 * i]);
        y_min = std::min(y_min, d.xy[2 * i + 1]);
        y_max = std::max(y_max, d.xy[2 * i + 1]);
    }
    for (int j = 0; j < d.n_bnd; ++j) {
        vertices[d.bnd_vtx[2 * j]].boundaryCondition = d.bnd_tag[j];
        vertices[d.bnd_vtx[2 * j + 1]].boundaryCondition = d.bnd_tag[j];
    }
    cells.resize(nc);
    for (int c = 0; c < nc; ++c) {
        Cell &cell = cells[c];
        cell.cellVolume = cg[c];
        cell.centre = {cg[nc + c], cg[2 * nc + c], 0.0};
        for (int k = 0; k < d.cell_nv[c]; ++k) {
            const int v = d.cell_vtx[4 * c + k];
            cell.verticeIDs.push_back(v);
            vertices[v].cellIDs.push_back(c);
        }
        for (int k = 0; k < 4 && ce[4 * c + k] >= 0; ++k) cell.edgeIDs.push_back(ce[4 * c + k]);
    }
    edges.resize(ne);
    for (int e = 0; e < ne; ++e) {
        Edge &edge = edges[e];
        edge.leftVertexID = ev[2 * e];
        edge.rightVertexID = ev[2 * e + 1];
        edge.leftCellID = ec[2 * e];
        edge.boundaryCondition = eb[e];
        edge.edgeLength = eg[e];
        edge.edgeNormal = {eg[ne + e], eg[2 * ne + e], 0.0};
        edge.centre = {eg[3 * ne + e], eg[4 * ne + e], 0.0};
        vertices[edge.leftVertexID].edgeIDs.push_back(e);
        vertices[edge.rightVertexID].edgeIDs.push_back(e);
        if (ec[2 * e + 1] >= 0) {
            edge.rightCellID = ec[2 * e + 1];
            cells[edge.leftCellID].neighbourIDs.push_back(edge.rightCellID);
            cells[edge.rightCellID].neighbourIDs.push_back(edge.leftCellID);
        } else {
            cells[edge.leftCellID].border = true;
        }
    }
    for (Vertex &v : vertices) std::sort(v.cellIDs.begin(), v.cellIDs.end());
    for (Cell &c : cells) std::sort(c.neighbourIDs.begin(), c.neighbourIDs.end());
}

Mesh::~Mesh() { o2f_ctx_destroy(ctx_); }

void Mesh::state_from_file_order(std::vector<double> &U) const {
    if (cell_perm_.empty()) return;
    const size_t nc = cell_perm_.size();
    std::vector<double> out(U.size());
    for (size_t q = 0; q < 4; ++q)
        for (size_t i = 0; i < nc; ++i) out[q * nc + i] = U[q * nc + cell_perm_[i]];
    U.swap(out);
}

void Mesh::state_to_file_order(std::vector<double> &U) const {
    if (cell_perm_.empty()) return;
    const size_t nc = cell_perm_.size();
    std::vector<double> out(U.size());
    for (size_t q = 0; q < 4; ++q)
        for (size_t i = 0; i < nc; ++i) out[q * nc + cell_perm_[i]] = U[q * nc + i];
    U.swap(out);
}

void Mesh::set_params(const SimulationParams &params) {
    const o2f_params op = to_o2f_params(params);
    check(o2f_set_params(ctx_, &op));
    if (params.tile_cells != tile_cells_) {  // the tiles are rebuilt only on a change
        check(o2f_set_tile_cells(ctx_, params.tile_cells));
        tile_cells_ = params.tile_cells;
    }
}

void Mesh::push_state() {
    const size_t nc = cells.size();
    std::vector<double> U(4 * nc);
    for (size_t i = 0; i < nc; ++i) {
        U[i] = cells[i].rho;
        U[nc + i] = cells[i].rho_u;
        U[2 * nc + i] = cells[i].rho_v;
        U[3 * nc + i] = cells[i].E;
    }
    check(o2f_set_state(ctx_, U.data()));
}

void Mesh::pull_state() {
    const size_t nc = cells.size();
    std::vector<double> U(4 * nc);
    check(o2f_get_state(ctx_, U.data()));
    for (size_t i = 0; i < nc; ++i) {
        cells[i].rho = U[i];
        cells[i].rho_u = U[nc + i];
        cells[i].rho_v = U[2 * nc + i];
        cells[i].E = U[3 * nc + i];
    }
}

void Mesh::store_residual(const double *R) {
    const size_t nc = cells.size();
    for (size_t i = 0; i < nc; ++i) {
        cells[i].R_rho = R[i];
        cells[i].R_rho_u = R[nc + i];
        cells[i].R_rho_v = R[2 * nc + i];
        cells[i].R_E = R[3 * nc + i];
    }
}

void Mesh::pull_gradients() {
    const size_t nc = cells.size();
    std::vector<double> G(8 * nc);
    check(o2f_get_gradients(ctx_, G.data()));
    for (size_t i = 0; i < nc; ++i) {
        Cell &c = cells[i];
        c.rho_x = G[i];
        c.rho_u_x = G[nc + i];
        c.rho_v_x = G[2 * nc + i];
        c.E_x = G[3 * nc + i];
        c.rho_y = G[4 * nc + i];
        c.rho_u_y = G[5 * nc + i];
        c.rho_v_y = G[6 * nc + i];
        c.E_y = G[7 * nc + i];
    }
}

void Mesh::pull_faces() {
    const size_t ne = edges.size();
    std::vector<double> L(4 * ne), R(4 * ne);
    check(o2f_get_face_states(ctx_, L.data(), R.data()));
    for (size_t e = 0; e < ne; ++e) {
        Edge &ed = edges[e];
        ed.rho_L = L[e];
        ed.rho_u_L = L[ne + e];
        ed.rho_v_L = L[2 * ne + e];
        ed.E_L = L[3 * ne + e];
        ed.rho_R = R[e];
        ed.rho_u_R = R[ne + e];
        ed.rho_v_R = R[2 * ne + e];
        ed.E_R = R[3 * ne + e];
    }
}

}  // namespace navier_stokes
