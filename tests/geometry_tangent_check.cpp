// CPU check of the host geometry tangent (host_tangent.cpp) against central
// differences of the geometry builders it differentiates (host_mesh.cpp).
// Built and run by tests/test_geometry_tangent_cpu.py.
//
//   geometry_tangent_check <dir>
//
// <dir> holds the mesh as raw arrays (xy.bin float64, cnv.bin / cvx.bin /
// bv.bin int32, bt.bin int8).  For a seeded random displacement of every
// vertex (O2F_GEOM_FULL: the meshes at X +- h dX rebuilt from scratch) and
// for the NACA camber and camber-position displacements (O2F_GEOM_REFERENCE:
// the mesh perturbed in place by perturb_naca, as
// applyShapePerturbationNACA12 + updateDerivedGeometry do), prints one line
// per geometry array:
//
//   <case> <array> <max |tangent - central difference| / max |difference|>
//
// over the entries whose central differences at h and h / 10 agree (no
// orientation or clamp switch inside the step), and the number of entries
// left out; and per case one line
//
//   <case> sparse <entries differing from the dense form> <entries listed>
//
// for geometry_tangent_sparse expanded into dense arrays, and lines
//
//   <case> perturb <entries differing from the full update> <entries listed>
//
// for perturb_naca_sparse against a copy updated by perturb_naca and
// compute_stencils (eps = 1e-6, -1e-6, 1e-2).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "host_mesh.hpp"

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

struct Field {
    const char *name;
    std::vector<double> HostMesh::*mesh;
    std::vector<double> GeomTangent::*tangent;
};
const Field kFields[] = {
    {"cvol", &HostMesh::cvol, &GeomTangent::cvol}, {"ccx", &HostMesh::ccx, &GeomTangent::ccx},
    {"ccy", &HostMesh::ccy, &GeomTangent::ccy},    {"elen", &HostMesh::elen, &GeomTangent::elen},
    {"enx", &HostMesh::enx, &GeomTangent::enx},    {"eny", &HostMesh::eny, &GeomTangent::eny},
    {"emx", &HostMesh::emx, &GeomTangent::emx},    {"emy", &HostMesh::emy, &GeomTangent::emy},
    {"sw", &HostMesh::sw, &GeomTangent::sw},
};

void report(const char *label, const HostMesh &p1, const HostMesh &m1, const HostMesh &p2, const HostMesh &m2,
            double h, const GeomTangent &t) {
    for (const Field &f : kFields) {
        const std::vector<double> &a1 = p1.*f.mesh, &b1 = m1.*f.mesh, &a2 = p2.*f.mesh, &b2 = m2.*f.mesh;
        const std::vector<double> &tv = t.*f.tangent;
        double scale = 0.0;
        for (size_t i = 0; i < tv.size(); ++i) scale = std::max(scale, std::abs((a1[i] - b1[i]) / (2 * h)));
        double err = 0.0;
        int skipped = 0;
        for (size_t i = 0; i < tv.size(); ++i) {
            const double d1 = (a1[i] - b1[i]) / (2 * h), d2 = (a2[i] - b2[i]) / (2 * h / 10);
            if (scale > 0 && std::abs(d1 - d2) > 1e-5 * scale) {
                ++skipped;
                continue;
            }
            err = std::max(err, std::abs(tv[i] - d1));
        }
        std::printf("%s %s %.6e %d %.6e\n", label, f.name, scale > 0 ? err / scale : err, skipped, scale);
    }
}

// geometry_tangent_sparse expanded into dense arrays against
// geometry_tangent: prints "<case> sparse <entries that differ bitwise>".
void compare_sparse(const char *label, const HostMesh &hm, const double *dxy, int geometry, const GeomTangent &t) {
    GeomSparse sp;
    geometry_tangent_sparse(hm, dxy, geometry, sp);
    GeomTangent d;
    d.cvol.assign(hm.nc, 0.0);
    d.ccx.assign(hm.nc, 0.0);
    d.ccy.assign(hm.nc, 0.0);
    for (auto *v : {&d.elen, &d.enx, &d.eny, &d.emx, &d.emy}) v->assign(hm.ne, 0.0);
    d.sw.assign(hm.sw.size(), 0.0);
    for (size_t i = 0; i < sp.cell.size(); ++i) {
        d.cvol[sp.cell[i]] = sp.cval[3 * i];
        d.ccx[sp.cell[i]] = sp.cval[3 * i + 1];
        d.ccy[sp.cell[i]] = sp.cval[3 * i + 2];
    }
    for (size_t i = 0; i < sp.edge.size(); ++i) {
        const int e = sp.edge[i];
        d.elen[e] = sp.eval[5 * i];
        d.enx[e] = sp.eval[5 * i + 1];
        d.eny[e] = sp.eval[5 * i + 2];
        d.emx[e] = sp.eval[5 * i + 3];
        d.emy[e] = sp.eval[5 * i + 4];
    }
    for (size_t i = 0; i < sp.swi.size(); ++i) d.sw[sp.swi[i]] = sp.swv[i];
    long bad = 0;
    for (const Field &f : kFields) {
        const std::vector<double> &a = t.*f.tangent, &b = d.*f.tangent;
        for (size_t i = 0; i < a.size(); ++i) bad += std::memcmp(&a[i], &b[i], sizeof(double)) != 0;
    }
    std::printf("%s sparse %ld %zu\n", label, bad, sp.cell.size() + sp.edge.size() + sp.swi.size());
}

// perturb_naca_sparse applied to hm's geometry against the full update
// (a copy, perturb_naca, compute_stencils): prints
// "<case> perturb <entries that differ bitwise> <entries listed>".
void compare_perturbation(const char *label, const HostMesh &hm, int which, double pos_x, double chord, double m0,
                          double p0, double t0, double eps) {
    HostMesh q = hm;
    perturb_naca(q, which, pos_x, chord, m0, p0, t0, eps);
    compute_stencils(q, O2F_INTERP_MVC);
    GeomSparse sp;
    perturb_naca_sparse(hm, O2F_INTERP_MVC, which, pos_x, chord, m0, p0, t0, eps, sp);
    HostMesh r = hm;
    for (size_t i = 0; i < sp.cell.size(); ++i) {
        r.cvol[sp.cell[i]] = sp.cval[3 * i];
        r.ccx[sp.cell[i]] = sp.cval[3 * i + 1];
        r.ccy[sp.cell[i]] = sp.cval[3 * i + 2];
    }
    for (size_t i = 0; i < sp.edge.size(); ++i) {
        const int e = sp.edge[i];
        r.elen[e] = sp.eval[5 * i];
        r.enx[e] = sp.eval[5 * i + 1];
        r.eny[e] = sp.eval[5 * i + 2];
        r.emx[e] = sp.eval[5 * i + 3];
        r.emy[e] = sp.eval[5 * i + 4];
    }
    for (size_t i = 0; i < sp.swi.size(); ++i) r.sw[sp.swi[i]] = sp.swv[i];
    long bad = 0;
    for (const Field &f : kFields) {
        const std::vector<double> &a = q.*f.mesh, &b = r.*f.mesh;
        for (size_t i = 0; i < a.size(); ++i) bad += std::memcmp(&a[i], &b[i], sizeof(double)) != 0;
    }
    std::printf("%s perturb %ld %zu\n", label, bad, sp.cell.size() + sp.edge.size() + sp.swi.size());
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const auto xy = read_all<double>(dir + "/xy.bin");
    const auto cnv = read_all<int>(dir + "/cnv.bin");
    const auto cvx = read_all<int>(dir + "/cvx.bin");
    const auto bv = read_all<int>(dir + "/bv.bin");
    const auto bt = read_all<char>(dir + "/bt.bin");
    auto build = [&](const std::vector<double> &X) {
        o2f_mesh_desc d{};
        d.n_vertices = static_cast<int>(X.size() / 2);
        d.xy = X.data();
        d.n_cells = static_cast<int>(cnv.size());
        d.cell_nv = cnv.data();
        d.cell_vtx = cvx.data();
        d.n_bnd = static_cast<int>(bt.size());
        d.bnd_vtx = bv.data();
        d.bnd_tag = bt.data();
        HostMesh hm;
        std::string err;
        if (!build_host_mesh(d, O2F_INTERP_MVC, hm, err)) std::fprintf(stderr, "%s\n", err.c_str());
        return hm;
    };
    const HostMesh hm = build(xy);
    const double h = 1e-6;
    // every vertex moves; the meshes at X +- h dX are rebuilt
    std::mt19937_64 rng(1);
    std::normal_distribution<double> normal;
    std::vector<double> dxy(xy.size());
    for (double &v : dxy) v = 1e-3 * normal(rng);
    GeomTangent t;
    geometry_tangent(hm, dxy.data(), O2F_GEOM_FULL, t);
    auto moved = [&](double s) {
        std::vector<double> X = xy;
        for (size_t i = 0; i < X.size(); ++i) X[i] += s * dxy[i];
        return build(X);
    };
    report("full", moved(h), moved(-h), moved(h / 10), moved(-h / 10), h, t);
    compare_sparse("full", hm, dxy.data(), O2F_GEOM_FULL, t);
    // the NACA camber / camber position through the reference's in-place update
    const double pos_x = 0.355, chord = 1.0, m0 = 0.03, p0 = 0.45, t0 = 0.11;
    for (int which = 0; which < 2; ++which) {
        std::vector<double> d;
        naca_displacement(hm, which, pos_x, chord, m0, p0, d);
        GeomTangent tr;
        geometry_tangent(hm, d.data(), O2F_GEOM_REFERENCE, tr);
        auto perturbed = [&](double e) {
            HostMesh q = hm;
            perturb_naca(q, which, pos_x, chord, m0, p0, t0, e);
            compute_stencils(q, O2F_INTERP_MVC);
            return q;
        };
        report(which ? "naca_p" : "naca_m", perturbed(h), perturbed(-h), perturbed(h / 10), perturbed(-h / 10), h,
               tr);
        compare_sparse(which ? "naca_p" : "naca_m", hm, d.data(), O2F_GEOM_REFERENCE, tr);
        for (double e : {1e-6, -1e-6, 1e-2})
            compare_perturbation(which ? "naca_p" : "naca_m", hm, which, pos_x, chord, m0, p0, t0, e);
    }
    return 0;
}
