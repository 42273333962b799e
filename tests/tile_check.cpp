// CPU check of the fused stage's tile tables (host_tiles.cpp).  Built and
// run by tests/test_tiles_cpu.py.
//
//   tile_check <dir> <cells per tile>...
//
// <dir> holds the mesh as raw arrays (xy.bin float64, cnv.bin / cvx.bin /
// bv.bin int32, bt.bin int8).  For each tile size prints
//
//   tiles <target> <ok|FAIL: why> <tiles> <owned cells summed> <largest tile> <gradient cells / nc>
//
// and then, for the first size, one line per planted mistake
//
//   planted <kind> <detected|missed>
//
// where each mistake changes one table entry the way a plausible slip in
// the builder would (a wrong local vertex, a missing edge, a stencil cell
// pointing elsewhere, a vertex record with another vertex's stencil, a
// cell owned twice, a face neighbour left out, a tile's slot, the read
// records of a state or stress vertex, an edge normal) and
// check_tiles must reject it.  A last line
//
//   untagged <refused|accepted>
//
// builds tiles for the same mesh with one boundary edge's tag removed.
// The first two sizes also form a two-member batch (concat_tiles):
//
//   concat <ok|FAIL>
//
// followed by planted slips in the concatenation (a head naming the wrong
// member, a record offset or a local stencil offset one off).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "host_mesh.hpp"
#include "host_tiles.hpp"

using namespace o2f;

namespace {

template <class T>
std::vector<T> read_all(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    const size_t n = static_cast<size_t>(f.tellg()) / sizeof(T);
    f.seekg(0);
    std::vector<T> v(n);
    f.read(reinterpret_cast<char *>(v.data()), n * sizeof(T));
    return v;
}

void planted(const HostMesh &hm, const TileSet &good) {
    auto report = [&](const char *kind, TileSet ts) {
        std::string err;
        std::printf("planted %s %s\n", kind, check_tiles(hm, ts, err) ? "missed" : "detected");
    };
    // the tile with the most face cells beyond its own
    int t = 0;
    for (int i = 0; i < (int)good.head.size(); ++i)
        if (good.head[i].nF - good.head[i].nO > good.head[t].nF - good.head[t].nO) t = i;
    const TileHead h = good.head[t];
    {
        TileSet ts = good;  // an edge end point mapped to the wrong local vertex
        TileCell &r = ts.cell[h.cell0];
        r.lv[1] = (uint16_t)((r.lv[1] + 1) % h.nVg);
        report("vertex_slot", ts);
    }
    {
        TileSet ts = good;  // a slot's local edge index off by one
        TileCell &r = ts.cell[h.cell0];
        r.le[0] = (int16_t)((r.le[0] + 1) % h.nE);
        report("edge_slot", ts);
    }
    {
        TileSet ts = good;  // a stencil cell pointing at the next local cell
        const int off = ts.vo[h.vo0].ls;
        ts.lsten[off] = (uint16_t)((ts.lsten[off] + 1) % h.nGd);
        report("stencil_cell", ts);
    }
    {
        TileSet ts = good;  // the owned cells of the next tile begin with one of this tile's
        const TileHead &h2 = ts.head[(t + 1) % ts.head.size()];
        ts.cell[h2.cell0].c = ts.cell[h.cell0].c;
        report("owned_twice", ts);
    }
    {
        TileSet ts = good;  // one face neighbour counted as a gradient-only cell
        ts.head[t].nF -= 1;
        report("face_cell_dropped", ts);
    }
    {
        TileSet ts = good;  // a state vertex record with its neighbour's stencil offset
        ts.vg[h.vg0].b = ts.vg[h.vg0 + 1].b;
        report("vertex_stencil", ts);
    }
    {
        TileSet ts = good;  // a stress vertex's owned-cell slot pointing at another vertex
        TileOwn &o = ts.own[h.cell0];
        o.lo[0] = (uint16_t)((o.lo[0] + 1) % h.nVo);
        report("stress_slot", ts);
    }
    {
        TileSet ts = good;  // a tile's slot (its speed partial) that of the next tile
        ts.head[t].slot += 1;
        report("tile_slot", ts);
    }
    {
        TileSet ts = good;  // a state vertex's read record with its next stencil cell first
        TileGather &g = ts.vgf[h.vg0];
        std::swap(g.cid[0], g.cid[1]);
        report("gather_cell", ts);
    }
    {
        TileSet ts = good;  // a stress vertex's first weight rounded one unit up
        double &w = ts.vof[h.vo0].w[0];
        w = std::nextafter(w, 2.0);
        report("stress_weight", ts);
    }
    {
        TileSet ts = good;  // a tile edge's normal with its components swapped
        std::swap(ts.en[2 * h.e0], ts.en[2 * h.e0 + 1]);
        report("edge_normal", ts);
    }
}

// Two tile sets of the mesh as the members of a batch: the concatenation
// must check, and each planted slip in it must be caught.
void concat(const TileSet &a, const TileSet &b) {
    const std::vector<TileSet> sets = {a, b};
    TileSet all;
    concat_tiles(sets, all);
    std::string err;
    std::printf("concat %s\n", check_concat(sets, all, err) ? "ok" : "FAIL");
    auto report = [&](const char *kind, const TileSet &bad) {
        std::printf("planted %s %s\n", kind, check_concat(sets, bad, err) ? "missed" : "detected");
    };
    const int k = (int)a.head.size();  // the second member's first tile
    {
        TileSet bad = all;  // the second member's tile named as the first member's
        bad.head[k + 1].member = 0;
        report("concat_member", bad);
    }
    {
        TileSet bad = all;  // a record offset one short
        bad.head[k + 1].cell0 -= 1;
        report("concat_cell_offset", bad);
    }
    {
        TileSet bad = all;  // a local stencil offset off by one (a stencil of two or more cells)
        for (size_t i = b.vo.size() / 2; i < b.vo.size(); ++i)
            if ((b.vo[i].nkb & 0xffff) >= 2 && ((b.vo[i].nkb >> 16) & 0xff) == kStencilWsum) {
                bad.vo[a.vo.size() + i].ls += 1;
                break;
            }
        report("concat_stencil_offset", bad);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string dir = argv[1];
    std::vector<double> xy = read_all<double>(dir + "/xy.bin");
    std::vector<int> cnv = read_all<int>(dir + "/cnv.bin"), cvx = read_all<int>(dir + "/cvx.bin"),
                     bv = read_all<int>(dir + "/bv.bin");
    std::vector<char> bt = read_all<char>(dir + "/bt.bin");
    o2f_mesh_desc d{};
    d.n_vertices = (int)xy.size() / 2;
    d.xy = xy.data();
    d.n_cells = (int)cnv.size();
    d.cell_nv = cnv.data();
    d.cell_vtx = cvx.data();
    d.n_bnd = (int)bt.size();
    d.bnd_vtx = bv.data();
    d.bnd_tag = bt.data();
    HostMesh hm;
    std::string err;
    if (!build_host_mesh(d, O2F_INTERP_MVC, hm, err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    std::vector<TileSet> built;
    for (int a = 2; a < argc; ++a) {
        const int target = std::atoi(argv[a]);
        TileSet ts;
        if (!build_tiles(hm, target, ts, err)) {
            std::printf("tiles %d FAIL: %s\n", target, err.c_str());
            continue;
        }
        long owned = 0;
        for (const TileHead &h : ts.head) owned += h.nO;
        std::printf("tiles %d ok %d %ld %d %.4f\n", target, (int)ts.head.size(), owned, ts.max_o,
                    (double)ts.sum_gd / hm.nc);
        if (a == 2) planted(hm, ts);
        if (built.size() < 2) built.push_back(ts);
    }
    if (built.size() == 2) concat(built[0], built[1]);
    // one boundary edge without a tag
    for (int e = 0; e < hm.ne; ++e)
        if (hm.ecr[e] < 0) {
            hm.ebc[e] = '\0';
            break;
        }
    TileSet ts;
    std::printf("untagged %s\n", build_tiles(hm, std::atoi(argv[2]), ts, err) ? "accepted" : "refused");
    return 0;
}
