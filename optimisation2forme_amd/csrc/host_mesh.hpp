// Host-side mesh preparation for the gfx950 residual path.
//
// Builds, once per mesh, everything the device kernels treat as read-only:
// topology with the reference's ordering semantics, FP64 geometry, and the
// cell->vertex interpolation stencils (weights precomputed here, so the
// transcendental acos / tan / hypot / atan2 of src/advection.cpp:52-104 never
// run on the device).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "o2f.h"

namespace o2f {

// Stencil kinds of a vertex (how its value is formed from its cells).
enum : uint8_t {
    kStencilZero = 0,  // no branch of src/advection.cpp:29-128 writes it: value 0
    kStencilCopy = 1,  // one cell: copy (src/advection.cpp:120-128)
    kStencilAvg2 = 2,  // two cells: 0.5*(a+b) (src/advection.cpp:106-119)
    kStencilWsum = 3,  // w0*a0 + w1*a1 + ... left to right (:29-105)
};

struct HostMesh {
    int nv = 0, nc = 0, ne = 0;
    // vertices
    std::vector<double> vx, vy;
    std::vector<char> vbc;
    std::vector<int> vptr;        // CSR over incident cells
    std::vector<int> vcell_asc;   // ascending cell id
    std::vector<int> vcell_ang;   // ascending atan2 angle (Vertex::cellIDsOrdered)
    // cells
    std::vector<int> cnv, cvtx;   // cvtx: 4 per cell, -1 pad
    std::vector<double> cvol, ccx, ccy;
    std::vector<int> cedge;       // 4 per cell, ascending edge id, -1 pad
    std::vector<int> cnbr;        // 4 per cell, distinct face neighbours, -1 pad
    std::vector<uint8_t> cinfo;   // bits 0-3: cell is left cell of cedge[k]; bits 4-6: edge count
    std::vector<int> cevx;        // 8 per cell: (min, max) vertex of each edge slot, -1 pad
    // edges (std::map order of the sorted vertex pair)
    std::vector<int> eva, evb, ecl, ecr;
    std::vector<char> ebc;
    std::vector<double> elen, enx, eny, emx, emy;
    // interpolation stencils (same CSR layout as vptr)
    std::vector<uint8_t> skind;
    std::vector<int> scell;
    std::vector<double> sw;
    // wall edges in ascending id (C_L/C_D, adjoint RHS)
    std::vector<int> wall_edges;
    double min_h = 0.0;
};

// Mesh::Mesh semantics (src/mesh.cpp:7-248).  Returns false with `err` set.
bool build_host_mesh(const o2f_mesh_desc &d, int interp_mode, HostMesh &hm, std::string &err);
// Recomputes the stencil weights from current positions and cell centres
// (the reference recomputes them on every call, src/advection.cpp:26-131).
void compute_stencils(HostMesh &hm, int interp_mode);
// min over cells of 2 V / perimeter (src/simulation.cpp:98-104).
double compute_min_h(const HostMesh &hm);
// applyShapePerturbationNACA12 + updateDerivedGeometry
// (src/shape_optimization.cpp:748-890).  which: 0 camber m, 1 camber
// position p, anything else: no change (returns false).
bool perturb_naca(HostMesh &hm, int which, double pos_x, double c, double m0, double p0,
                  double t, double eps);

// Derivative of the derived geometry along a vertex displacement dxy
// (2 per vertex), in the layout of the HostMesh geometry arrays
// (host_tangent.cpp).
struct GeomTangent {
    std::vector<double> cvol, ccx, ccy, elen, enx, eny, emx, emy, sw;
};
// geometry: O2F_GEOM_REFERENCE, only the edges touching a wall vertex and
// their left cells follow the vertices (updateDerivedGeometry,
// src/shape_optimization.cpp:748-805); O2F_GEOM_FULL, every cell and edge
// does.  The interpolation weights follow the cell centres and vertices.
void geometry_tangent(const HostMesh &hm, const double *dxy, int geometry, GeomTangent &t);
// Entries of a geometry set given by index: cells with (V, c_x, c_y) in
// cval, edges with (L, n_x, n_y, mid_x, mid_y) in eval, stencil weights by
// their index into sw.  geometry_tangent_sparse lists the derivative's
// entries that can be non-zero (every other one is exactly zero).
struct GeomSparse {
    std::vector<int> cell, edge, swi;
    std::vector<double> cval, eval, swv;
};
void geometry_tangent_sparse(const HostMesh &hm, const double *dxy, int geometry, GeomSparse &t);
// d(vertex position)/d(parameter) of applyShapePerturbationNACA12
// (src/shape_optimization.cpp:807-890): which 0 camber m, 1 camber
// position p, 2 thickness t (y = yc + side * yt, x fixed, x_r >= 0.9
// frozen).  Returns false for any other `which`.
bool naca_displacement(const HostMesh &hm, int which, double pos_x, double c, double m0, double p0,
                       std::vector<double> &dxy);

// perturb_naca followed by compute_stencils, without copying the mesh: the
// entries of the perturbed geometry that can differ from hm's (the cells
// and edges updateDerivedGeometry refreshes, the stencil weights of the
// vertices that moved or whose cells did), each computed by the same code
// as the full update, so hm's geometry with these entries replaced is the
// full update's, bit for bit.
bool perturb_naca_sparse(const HostMesh &hm, int interp_mode, int which, double pos_x, double c, double m0,
                         double p0, double t, double eps, GeomSparse &out);

// MSH 2.2 ASCII reader (Gmsh API element order, see o2f_msh2_read).
bool read_msh2(const std::string &path, std::vector<double> &xy, std::vector<int> &cell_nv,
               std::vector<int> &cell_vtx, std::vector<int> &bnd_vtx, std::vector<char> &bnd_tag,
               std::string &err);

}  // namespace o2f
