// Adjoint-based shape gradient of the reference
// (src/shape_optimization.cpp:19-145, :171-189, :1022-1067) on the gfx950 path.
#include "shape_optimization.hpp"

#include <algorithm>
#include <cctype>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

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

namespace {
void trim(std::string &s) {
    s.erase(s.begin(), std::find_if(s.begin(), s.end(), [](unsigned char c) { return !std::isspace(c); }));
    s.erase(std::find_if(s.rbegin(), s.rend(), [](unsigned char c) { return !std::isspace(c); }).base(), s.end());
}

double param_or(const std::vector<std::string> &names, const std::vector<double> &values, const std::string &key,
                double def) {
    auto it = std::find(names.begin(), names.end(), key);
    return it == names.end() ? def : values[it - names.begin()];
}
}  // namespace

void loadParamsGeneric(const std::string &filepath, std::vector<std::string> &names_all,
                       std::vector<double> &values_all, std::vector<std::string> &names) {
    names_all.clear();
    values_all.clear();
    names.clear();
    if (filepath.empty() || filepath == "none") return;
    std::ifstream in(filepath);
    if (!in) throw std::runtime_error("Impossible d'ouvrir le fichier de paramètres: " + filepath);
    std::unordered_map<std::string, size_t> pos;
    std::unordered_set<std::string> opt;
    std::string line;
    while (std::getline(in, line)) {
        if (auto h = line.find('#'); h != std::string::npos) line.erase(h);
        trim(line);
        if (line.empty()) continue;
        bool marked = false;
        std::string key, sval;
        if (auto k = line.find("*="); k != std::string::npos) {
            marked = true;
            key = line.substr(0, k);
            sval = line.substr(k + 2);
        } else if (auto e = line.find('='); e != std::string::npos) {
            key = line.substr(0, e);
            sval = line.substr(e + 1);
        } else {
            std::istringstream ss(line);
            ss >> key >> sval;
        }
        trim(key);
        trim(sval);
        if (key.empty() || sval.empty()) continue;
        const double v = std::stod(sval);
        auto it = pos.find(key);
        if (it == pos.end()) {
            pos[key] = names_all.size();
            names_all.push_back(key);
            values_all.push_back(v);
        } else {
            values_all[it->second] = v;
        }
        if (marked)
            opt.insert(key);
        else
            opt.erase(key);
    }
    for (const auto &k : names_all)
        if (opt.count(k)) names.push_back(k);
}

void optimizeShape(Mesh &mesh, const SimulationParams &params, double eps) {
    std::vector<double> values_all;
    std::vector<std::string> names, names_all;
    loadParamsGeneric(params.shape_param_file, names_all, values_all, names);
    const double beta = param_or(names_all, values_all, "beta", 1);
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
    std::cout << std::fixed << std::setprecision(8);
    std::cout << "[preview] J courant = " << J_current << "\n";
    std::cout << "[preview] Gradient de forme (dJ/dθ) et sens pour diminuer J :\n";
    for (size_t k = 0; k < names.size(); ++k) {
        const double g = grad[k];
        const char *sens = (g > 0.0)   ? "-> diminuer ce paramètre (pas négatif)"
                           : (g < 0.0) ? "-> augmenter ce paramètre (pas positif)"
                                       : "-> gradient ~ 0 (aucune direction claire)";
        std::cout << "  - " << names[k] << " : dJ/d" << names[k] << " = " << g << "    " << sens << "\n";
    }
    // appendShapeOptLog (:112-145): " J=%.8f name=%e ..." parsed by run_opti.sh
    const std::string &path = params.shape_log_file;
    if (!path.empty() && path != "none") {
        std::filesystem::path p(path);
        if (!p.parent_path().empty()) {
            std::error_code ec;
            if (!std::filesystem::create_directories(p.parent_path(), ec) && ec)
                throw std::runtime_error("Impossible de créer le dossier: " + p.parent_path().string());
        }
        std::ofstream out(path);
        if (!out) throw std::runtime_error("Impossible d'ouvrir " + path + " en écriture.");
        out << std::fixed << std::setprecision(8) << " J=" << J_current;
        out << std::scientific << std::setprecision(6);
        for (size_t k = 0; k < names.size(); ++k) out << ' ' << names[k] << '=' << grad[k];
        out << '\n';
    }
}

}  // namespace navier_stokes
