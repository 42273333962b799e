// Host-side cell tiles of the fused stage kernel (o2f_tiles.hip).
//
// One workgroup evaluates one residual stage for a compact tile of cells it
// owns, with every intermediate of the stage (vertex states, face states,
// edge fluxes, cell and vertex stresses) kept in LDS.  The tables below tell
// it, per tile, which entities it must form:
//
//   owned cells O       updated and written by this tile;
//   face cells F ⊇ O    O plus the face neighbours of O: their face states
//                       on the tile's edges enter the tile's HLLC fluxes;
//   gradient cells Gd   F plus every cell in the interpolation stencils of
//                       the vertices of O (their stresses enter the vertex
//                       stresses the update reads); each needs its
//                       Green-Gauss gradient, so its vertices' states;
//   state vertices Vg   the edge end points of every cell of Gd;
//   stress vertices Vo  the edge end points of the cells of O;
//   edges E             the edges of the cells of O.
//
// Gd is stored as [O | F \ O | Gd \ F], so a cell's local index says which
// of the three it is.  Every entity is formed with the same arithmetic as in
// the four-kernel stage (o2f_kernels.hip), so a tile's results are those of
// the separate kernels bit for bit; entities on a tile's rim are formed by
// every tile that needs them.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "host_mesh.hpp"

namespace o2f {

// One gradient cell of a tile (64 bytes, read by one lane).
struct TileCell {
    int c;           // global cell id
    int edge[4];     // its edges (cedge), -1 pad
    int nbr[4];      // its face neighbours (cnbr), -1 pad
    uint16_t lv[8];  // local state-vertex index of each edge slot's end points (cevx order), 0xffff pad
    int16_t le[4];   // local edge index of each edge slot, -1 where the edge is not among the tile's
    uint8_t info;    // cinfo
    uint8_t pad[3];
};
static_assert(sizeof(TileCell) == 64, "TileCell is one 64-byte record");

// Local stress-vertex index of each edge slot's end points of an owned cell.
struct TileOwn {
    uint16_t lo[8];
};

// A state or stress vertex of a tile: its stencil (offset b into the
// stencil arrays, length | kind << 16 | boundary tag << 24) and, for a
// stress vertex, the offset of its local stencil in lsten (-1 otherwise).
struct TileVertex {
    int v, b, nkb, ls;
};
static_assert(sizeof(TileVertex) == 16, "TileVertex is one 16-byte record");

// Stencil entries a vertex record carries (longer stencils read the rest
// from the mesh's stencil arrays).
constexpr int kTileSlots = 8;

// A state vertex as the stage reads it: its first kTileSlots stencil cells
// (global ids, -1 past the stencil), so the cells' states load right after
// the record; the stencil offset b (weights, longer stencils) and nkb as in
// TileVertex.
struct TileGather {
    int cid[kTileSlots];
    int b, nkb;
    int pad[2];
};
static_assert(sizeof(TileGather) == 48, "TileGather is three 16-byte words");

// A stress vertex as the stage reads it: its first kTileSlots stencil
// weights and local gradient-cell indices (0 / 0xffff past the stencil's
// used entries), nkb, the stencil offset b and its local stencil offset ls
// (entries past kTileSlots).
struct TileStress {
    double w[kTileSlots];
    uint16_t lc[kTileSlots];
    int nkb, b, ls, pad;
};
static_assert(sizeof(TileStress) == 96, "TileStress is six 16-byte words");

// Offsets and counts of one tile in the concatenated tables.
struct TileHead {
    int cell0, nO, nF, nGd;  // TileCell records [cell0, cell0 + nGd); TileOwn records [cell0, cell0 + nO)
    int vg0, nVg;            // state vertices
    int e0, nE;              // edges (global ids)
    int vo0, nVo;            // stress vertices
    int member, slot;        // the mesh of a batch it belongs to, and its index among that mesh's tiles
    int pad[4];
};
static_assert(sizeof(TileHead) == 64, "TileHead is one 64-byte record");

struct TileSet {
    int target = 0;  // cells per tile asked for
    std::vector<TileHead> head;
    std::vector<TileCell> cell;
    std::vector<TileOwn> own;  // indexed like cell (only the first nO of each tile are used)
    std::vector<TileVertex> vg;   // state vertices
    std::vector<int> edge;        // global edge ids
    std::vector<TileVertex> vo;   // stress vertices
    std::vector<uint16_t> lsten;  // local Gd index of each stencil cell, in stencil order
    // what the stage reads in place of vg / vo / edge (one entry per entry of
    // those): state-vertex gather records, stress-vertex records, and each
    // tile edge's unit normal (n_x, n_y)
    std::vector<TileGather> vgf;
    std::vector<TileStress> vof;
    std::vector<double> en;
    // largest per-tile counts (LDS is sized for them)
    int max_vg = 0, max_vo = 0, max_gd = 0, max_e = 0, max_o = 0;
    // sums over the tiles (work with the rim counted as often as it is formed)
    long sum_gd = 0, sum_f = 0, sum_vg = 0, sum_vo = 0, sum_e = 0;
    // LDS layout of one tile sized for the largest counts, in doubles: the
    // state / stress vertices from 0, the cell stresses from *tq, the face
    // states from *fs (each region a multiple of four doubles, so rows stay
    // 16-byte aligned); returns the bytes
    size_t lds_layout(int *tq, int *fs) const;
    size_t lds_bytes() const { return lds_layout(nullptr, nullptr); }
};

// Cells grouped into compact tiles of at most `target` cells: a tile grows
// from the lowest-numbered unassigned cell, always adding the candidate cell
// that shares the most vertices with the tile (then the one reached first,
// then the lowest id), which keeps its rim short.  Returns false (err set)
// when the mesh has a boundary edge without a boundary tag (its right face
// state would be undefined) or a table index would not fit its field.
bool build_tiles(const HostMesh &hm, int target, TileSet &ts, std::string &err);

// Every table entry checked against the mesh: the owned cells partition the
// mesh; each tile lists exactly the cells, vertices and edges its stage
// reads, with local indices that point at the right global entity.
// Returns false with a description of the first violation.
bool check_tiles(const HostMesh &hm, const TileSet &ts, std::string &err);

// The tiles of several meshes (a batch) as one set: member m's tiles follow
// member m - 1's, each head naming its member and its slot among that
// member's tiles, its offsets moved past the earlier members' records and
// its stress vertices' local stencils moved with them; the maxima are those
// over every member.  Global ids inside the records stay the member's own.
void concat_tiles(const std::vector<TileSet> &sets, TileSet &all);
// Every tile of the concatenation against its member's own tile: same
// counts, same records, same local stencils, member and slot in the head.
bool check_concat(const std::vector<TileSet> &sets, const TileSet &all, std::string &err);

}  // namespace o2f
