// Mesh renumbering for locality (host only).
//
// The residual's stage kernels gather by cell, edge and vertex id, so their
// speed depends on how the mesh is numbered: on the 200k-triangle transonic
// workload a random numbering runs a step 3.3x slower than the layer-by-layer
// numbering of the generated O-mesh, and reverse Cuthill-McKee brings the
// random one back to within 5 %.  Meshes read from a mesher keep its numbering
// (Mesh::Mesh, src/mesh.cpp:9-111, uses the Gmsh order as it comes); this
// file gives callers a renumbered copy and the two permutations, so that
// states can be mapped in and out.
//
// Cells: reverse Cuthill-McKee over the graph of cells sharing a vertex,
// each connected component started from a pseudo-peripheral cell (George and
// Liu: repeat the level structure from a minimum-degree cell of the deepest
// level while the depth grows), neighbours visited by ascending degree, then
// ascending id.  Vertices: in the order the renumbered cells first use them.
// Optionally the cells are then laid out for the eight XCDs of an MI355X: the
// order is cut into eight contiguous chunks and their blocks of 64 cells (one
// workgroup of the four-lane cell kernels) interleaved, block 8 j + x holding
// block j of chunk x, so that the workgroups the hardware sends to one XCD
// (every eighth) cover one contiguous part of the mesh and share its L2.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "o2f.h"

namespace o2f {
extern thread_local std::string g_error;
}
using o2f::g_error;

namespace {

struct CellGraph {
    std::vector<int> ptr, adj;  // CSR, neighbours sorted by id
    int degree(int c) const { return ptr[c + 1] - ptr[c]; }
};

CellGraph cells_sharing_a_vertex(const o2f_mesh_desc &m) {
    const int nc = m.n_cells, nv = m.n_vertices;
    std::vector<int> vptr(nv + 1, 0), vcell;
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < m.cell_nv[c]; ++k) ++vptr[m.cell_vtx[4 * c + k] + 1];
    std::partial_sum(vptr.begin(), vptr.end(), vptr.begin());
    vcell.resize(vptr[nv]);
    {
        std::vector<int> pos(vptr.begin(), vptr.end() - 1);
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < m.cell_nv[c]; ++k) vcell[pos[m.cell_vtx[4 * c + k]]++] = c;
    }
    CellGraph g;
    g.ptr.assign(nc + 1, 0);
    std::vector<int> nb;
    for (int c = 0; c < nc; ++c) {
        nb.clear();
        for (int k = 0; k < m.cell_nv[c]; ++k) {
            const int v = m.cell_vtx[4 * c + k];
            for (int i = vptr[v]; i < vptr[v + 1]; ++i)
                if (vcell[i] != c) nb.push_back(vcell[i]);
        }
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        g.adj.insert(g.adj.end(), nb.begin(), nb.end());
        g.ptr[c + 1] = static_cast<int>(g.adj.size());
    }
    return g;
}

// Breadth-first level structure from `root` over the unvisited part of the
// root's component; returns the depth and leaves the last level in `last`.
int levels(const CellGraph &g, int root, std::vector<int> &stamp, int mark, std::vector<int> &last) {
    std::vector<int> cur{root}, next;
    stamp[root] = mark;
    int depth = 0;
    for (;;) {
        next.clear();
        for (int c : cur)
            for (int i = g.ptr[c]; i < g.ptr[c + 1]; ++i) {
                const int n = g.adj[i];
                if (stamp[n] != mark) {
                    stamp[n] = mark;
                    next.push_back(n);
                }
            }
        if (next.empty()) break;
        ++depth;
        cur.swap(next);
    }
    last = cur;
    return depth;
}

std::vector<int> rcm_order(const CellGraph &g, int nc) {
    std::vector<int> order;
    order.reserve(nc);
    std::vector<char> done(nc, 0);
    std::vector<int> stamp(nc, -1), last;
    int mark = 0;
    // components in ascending id of their minimum-degree cell
    std::vector<int> by_degree(nc);
    std::iota(by_degree.begin(), by_degree.end(), 0);
    std::stable_sort(by_degree.begin(), by_degree.end(),
                     [&](int a, int b) { return g.degree(a) < g.degree(b); });
    for (int seed : by_degree) {
        if (done[seed]) continue;
        // pseudo-peripheral cell of this component
        int root = seed;
        int depth = levels(g, root, stamp, mark++, last);
        for (;;) {
            const int cand = *std::min_element(last.begin(), last.end(), [&](int a, int b) {
                return g.degree(a) != g.degree(b) ? g.degree(a) < g.degree(b) : a < b;
            });
            std::vector<int> last2;
            const int d2 = levels(g, cand, stamp, mark++, last2);
            if (d2 <= depth) break;
            root = cand;
            depth = d2;
            last.swap(last2);
        }
        // Cuthill-McKee from the root
        const size_t begin = order.size();
        order.push_back(root);
        done[root] = 1;
        std::vector<int> nb;
        for (size_t h = begin; h < order.size(); ++h) {
            const int c = order[h];
            nb.clear();
            for (int i = g.ptr[c]; i < g.ptr[c + 1]; ++i)
                if (!done[g.adj[i]]) nb.push_back(g.adj[i]);
            std::sort(nb.begin(), nb.end(), [&](int a, int b) {
                return g.degree(a) != g.degree(b) ? g.degree(a) < g.degree(b) : a < b;
            });
            for (int n : nb) {
                done[n] = 1;
                order.push_back(n);
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

std::vector<int> xcd_interleave(const std::vector<int> &order, int block, int nxcd) {
    const int nc = static_cast<int>(order.size());
    const int nb = (nc + block - 1) / block, per = (nb + nxcd - 1) / nxcd;
    std::vector<int> out;
    out.reserve(nc);
    for (int j = 0; j < per; ++j)
        for (int x = 0; x < nxcd; ++x) {
            const int b = x * per + j;
            if (b >= nb) continue;
            for (int i = b * block; i < std::min(nc, (b + 1) * block); ++i) out.push_back(order[i]);
        }
    return out;
}

}  // namespace

extern "C" int o2f_mesh_reorder(const o2f_mesh_desc *mesh, int method, o2f_mesh_desc *out, int *cell_perm,
                                int *vertex_perm) {
    if (!mesh || !out) {
        g_error = "o2f_mesh_reorder: null argument";
        return O2F_ERR_ARG;
    }
    if (method != O2F_REORDER_RCM && method != O2F_REORDER_RCM_XCD) {
        g_error = "o2f_mesh_reorder: unknown method";
        return O2F_ERR_ARG;
    }
    const int nc = mesh->n_cells, nv = mesh->n_vertices, nb = mesh->n_bnd;
    if (nc < 0 || nv < 0 || nb < 0 || (nc > 0 && (!mesh->cell_nv || !mesh->cell_vtx)) || (nv > 0 && !mesh->xy) ||
        (nb > 0 && (!mesh->bnd_vtx || !mesh->bnd_tag))) {
        g_error = "o2f_mesh_reorder: incomplete mesh description";
        return O2F_ERR_ARG;
    }
    for (int c = 0; c < nc; ++c) {
        const int n = mesh->cell_nv[c];
        if (n != 3 && n != 4) {
            g_error = "o2f_mesh_reorder: cells must be triangles or quadrilaterals";
            return O2F_ERR_ARG;
        }
        for (int k = 0; k < n; ++k) {
            const int v = mesh->cell_vtx[4 * c + k];
            if (v < 0 || v >= nv) {
                g_error = "o2f_mesh_reorder: cell references a missing vertex";
                return O2F_ERR_ARG;
            }
        }
    }
    for (int j = 0; j < 2 * nb; ++j)
        if (mesh->bnd_vtx[j] < 0 || mesh->bnd_vtx[j] >= nv) {
            g_error = "o2f_mesh_reorder: boundary element references a missing vertex";
            return O2F_ERR_ARG;
        }
    std::vector<int> corder = rcm_order(cells_sharing_a_vertex(*mesh), nc);
    if (method == O2F_REORDER_RCM_XCD) corder = xcd_interleave(corder, 64, 8);
    // vertices by first use; vertices no cell uses keep their relative order
    std::vector<int> vnew(nv, -1), vorder;
    vorder.reserve(nv);
    for (int c : corder)
        for (int k = 0; k < mesh->cell_nv[c]; ++k) {
            const int v = mesh->cell_vtx[4 * c + k];
            if (vnew[v] < 0) {
                vnew[v] = static_cast<int>(vorder.size());
                vorder.push_back(v);
            }
        }
    for (int v = 0; v < nv; ++v)
        if (vnew[v] < 0) {
            vnew[v] = static_cast<int>(vorder.size());
            vorder.push_back(v);
        }
    auto *pxy = new double[2 * static_cast<size_t>(nv) + 1];
    auto *pcnv = new int[nc + 1];
    auto *pcv = new int[4 * static_cast<size_t>(nc) + 1];
    auto *pbv = new int[2 * static_cast<size_t>(nb) + 1];
    auto *pbt = new char[nb + 1];
    for (int i = 0; i < nv; ++i) {
        pxy[2 * i] = mesh->xy[2 * vorder[i]];
        pxy[2 * i + 1] = mesh->xy[2 * vorder[i] + 1];
    }
    for (int i = 0; i < nc; ++i) {
        const int c = corder[i];
        pcnv[i] = mesh->cell_nv[c];
        for (int k = 0; k < 4; ++k) pcv[4 * i + k] = (k < pcnv[i]) ? vnew[mesh->cell_vtx[4 * c + k]] : -1;
    }
    for (int j = 0; j < 2 * nb; ++j) pbv[j] = vnew[mesh->bnd_vtx[j]];
    if (nb > 0) std::memcpy(pbt, mesh->bnd_tag, nb);
    out->n_vertices = nv;
    out->xy = pxy;
    out->n_cells = nc;
    out->cell_nv = pcnv;
    out->cell_vtx = pcv;
    out->n_bnd = nb;
    out->bnd_vtx = pbv;
    out->bnd_tag = pbt;
    if (cell_perm) std::copy(corder.begin(), corder.end(), cell_perm);
    if (vertex_perm) std::copy(vorder.begin(), vorder.end(), vertex_perm);
    return O2F_OK;
}
