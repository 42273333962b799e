// Adjoint-based shape gradient of the reference
// (src/shape_optimization.cpp:19-145, :171-189, :1022-1067) on the gfx950 path.
#include "shape_optimization.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace navier_stokes {

void computeResidual(Mesh &mesh, const SimulationParams &params, std::vector<double> &R_rho,
                     std::vector<double> &R_rho_u, std::vector<double> &R_rho_v, std::vector<double> &R_E) {
    mesh.set_params(params);
    mesh.push_state();
    const size_t nc = mesh.cells.size();
    std::vector<double> R(4 * nc);
    mesh.check(o2f_residual(mesh.ctx(), R.data()));
    mesh.store_residual(R.data());
    R_rho.assign(R.begin(), R.begin() + nc);
    R_rho_u.assign(R.begin() + nc, R.begin() + 2 * nc);
    R_rho_v.assign(R.begin() + 2 * nc, R.begin() + 3 * nc);
    R_E.assign(R.begin() + 3 * nc, R.end());
}

// ---- the shape-parameter file -------------------------------------------
//
// Grammar (the reference's, src/shape_optimization.cpp:19-98, which
// run_opti.sh:91-129 rewrites between runs): text after '#' is a comment;
// each remaining non-blank line assigns one value, as `key *= value` (an
// optimised parameter), `key = value` or `key value`; "*=" is looked for
// before '='; the value is read by std::stod.  A key assigned again keeps
// its first position, takes the new value, and is optimised iff its last
// assignment used "*=".
namespace {

struct ShapeParam {
    std::string name;
    double value;
    bool optimised;
};

bool is_blank(char ch) { return std::isspace(static_cast<unsigned char>(ch)) != 0; }

std::string strip_blanks(const std::string &s, size_t from = 0, size_t to = std::string::npos) {
    to = std::min(to, s.size());
    while (from < to && is_blank(s[from])) ++from;
    while (to > from && is_blank(s[to - 1])) --to;
    return s.substr(from, to - from);
}

// key and value of one line; false when the line assigns nothing
bool parse_param_line(std::string line, std::string &key, std::string &value, bool &optimised) {
    line = line.substr(0, line.find('#'));
    optimised = false;
    size_t at = line.find("*="), width = 2;
    if (at != std::string::npos) {
        optimised = true;
    } else {
        at = line.find('=');
        width = 1;
    }
    if (at != std::string::npos) {
        key = strip_blanks(line, 0, at);
        value = strip_blanks(line, at + width);
    } else {  // `key value`: the first two blank-separated words
        std::istringstream words(line);
        key.clear();
        value.clear();
        words >> key >> value;
    }
    return !key.empty() && !value.empty();
}

std::vector<ShapeParam> read_shape_params(const std::string &path) {
    std::vector<ShapeParam> params;
    if (path.empty() || path == "none") return params;
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Impossible d'ouvrir le fichier de paramètres: " + path);
    std::string key, value;
    bool optimised;
    for (std::string line; std::getline(in, line);) {
        if (!parse_param_line(line, key, value, optimised)) continue;
        const double v = std::stod(value);
        auto known = std::find_if(params.begin(), params.end(), [&](const ShapeParam &q) { return q.name == key; });
        if (known == params.end())
            params.push_back({key, v, optimised});
        else
            *known = {key, v, optimised};
    }
    return params;
}

double value_or(const std::vector<ShapeParam> &params, const std::string &name, double fallback) {
    for (const ShapeParam &q : params)
        if (q.name == name) return q.value;
    return fallback;
}

std::string format(const char *fmt, double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, fmt, v);
    return buf;
}

// The shape log that run_opti.sh parses (parse_log_J_and_grads): one line,
// " J=" in fixed notation with 8 decimals, then " name=gradient" per
// optimised parameter in scientific notation with 6 (the reference's
// appendShapeOptLog, src/shape_optimization.cpp:112-145, writes the same
// bytes).  The file is replaced, not appended to.
void write_shape_log(const std::string &path, double J, const std::vector<std::string> &names,
                     const std::vector<double> &grad) {
    if (path.empty() || path == "none") return;
    const std::filesystem::path dir = std::filesystem::path(path).parent_path();
    if (!dir.empty()) {
        std::error_code ec;
        std::filesystem::create_directories(dir, ec);
        if (ec) throw std::runtime_error("Impossible de créer le dossier: " + dir.string());
    }
    std::string text = " J=" + format("%.8f", J);
    for (size_t k = 0; k < names.size(); ++k) text += " " + names[k] + "=" + format("%.6e", grad[k]);
    text += "\n";
    std::ofstream out(path);
    if (!out) throw std::runtime_error("Impossible d'ouvrir " + path + " en écriture.");
    out << text;
}

// The console summary of optimizeShape (:1046-1063): J, then per parameter
// its gradient and the direction that lowers J.
void print_preview(double J, const std::vector<std::string> &names, const std::vector<double> &grad) {
    std::string text = "[preview] J courant = " + format("%.8f", J) + "\n" +
                       "[preview] Gradient de forme (dJ/dθ) et sens pour diminuer J :\n";
    for (size_t k = 0; k < names.size(); ++k) {
        const double g = grad[k];
        const char *hint = g > 0.0   ? "-> diminuer ce paramètre (pas négatif)"
                             : g < 0.0 ? "-> augmenter ce paramètre (pas positif)"
                                       : "-> gradient ~ 0 (aucune direction claire)";
        text += "  - " + names[k] + " : dJ/d" + names[k] + " = " + format("%.8f", g) + "    " + hint + "\n";
    }
    std::cout << text << std::flush;
}

}  // namespace

void loadParamsGeneric(const std::string &filepath, std::vector<std::string> &names_all,
                       std::vector<double> &values_all, std::vector<std::string> &names) {
    names_all.clear();
    values_all.clear();
    names.clear();
    for (const ShapeParam &q : read_shape_params(filepath)) {
        names_all.push_back(q.name);
        values_all.push_back(q.value);
        if (q.optimised) names.push_back(q.name);
    }
}

void optimizeShape(Mesh &mesh, const SimulationParams &params, double eps) {
    const std::vector<ShapeParam> file = read_shape_params(params.shape_param_file);
    std::vector<double> values_all;
    std::vector<std::string> names, names_all;
    for (const ShapeParam &q : file) {
        names_all.push_back(q.name);
        values_all.push_back(q.value);
        if (q.optimised) names.push_back(q.name);
    }
    const double beta = value_or(file, "beta", 1.0);
    double J_current = 0.0;
    mesh.check(o2f_objective(mesh.ctx(), beta, &J_current));
    const size_t nc = mesh.cells.size();
    std::vector<double> lambda(4 * nc);
    int iters = 0;
    double relres = 0.0;
    // solveAdjointSystem (src/shape_optimization.cpp:658-740), or the
    // extensions the config selects
    const bool block_jacobi = params.adjoint_preconditioner == "block_jacobi";
    int st;
    if (params.adjoint_operator == "exact") {
        const int opts = (params.adjoint_rhs == "exact" ? O2F_ADJ_EXACT_RHS : 0) |
                         (block_jacobi ? O2F_ADJ_BLOCK_JACOBI : 0);
        st = o2f_bicgstab_exact(mesh.ctx(), beta, params.adjoint_tol, params.adjoint_chunk,
                                params.adjoint_max_iters, lambda.data(), &iters, &relres, opts);
    } else if (block_jacobi) {
        st = o2f_bicgstab_options(mesh.ctx(), beta, eps, params.adjoint_tol, params.adjoint_chunk,
                                  params.adjoint_max_iters, O2F_ADJ_BLOCK_JACOBI, lambda.data(), &iters, &relres);
    } else {
        st = o2f_bicgstab(mesh.ctx(), beta, eps, params.adjoint_tol, params.adjoint_chunk,
                          params.adjoint_max_iters, lambda.data(), &iters, &relres);
    }
    std::cout << "[adjoint] iters=" << iters << "  relres=" << relres << std::endl;
    mesh.check(st);
    std::vector<const char *> na, no;
    for (auto &s : names_all) na.push_back(s.c_str());
    for (auto &s : names) no.push_back(s.c_str());
    std::vector<double> grad(names.size());
    if (params.shape_gradient == "tangent" || params.shape_gradient == "tangent_full")
        mesh.check(o2f_shape_gradient_tangent(
            mesh.ctx(), na.data(), values_all.data(), (int)na.size(), no.data(), (int)no.size(), lambda.data(), beta,
            params.shape_gradient == "tangent_full" ? O2F_GEOM_FULL : O2F_GEOM_REFERENCE, grad.data()));
    else
        mesh.check(o2f_shape_gradient(mesh.ctx(), na.data(), values_all.data(), (int)na.size(), no.data(),
                                      (int)no.size(), lambda.data(), beta, eps, grad.data()));
    print_preview(J_current, names, grad);
    write_shape_log(params.shape_log_file, J_current, names, grad);
}

}  // namespace navier_stokes
