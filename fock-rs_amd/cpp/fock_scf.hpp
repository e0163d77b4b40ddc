// fock_scf.hpp -- C++ mirror of the fock-rs `scf` crate API over the C ABI of libfock.so.
//
// Same names, argument meaning and outcome rules as the Rust crates:
//   GaussianPrimitive::new_normalized   basis/src/primitives.rs:19-48 (norm, new)
//   ContractedGaussian::evaluate        basis/src/contracted_gaussian.rs:320-331
//   BasisSet::get                       basis_set/src/lib.rs:79-119 (shell expansion, z,y,x order)
//   Atom                                basis_set/src/atom.rs:7-43
//   read_xyz_file                       chemfiles/src/xyz.rs:13-31 (Angstrom x 1.89)
//   HartreeFockResult                   scf/src/lib.rs:27-37
//   MolecularOrbitals                   scf/src/molecular_orbitals.rs:46-99 (|c| > 0.05)
//   try_scf                             scf/src/lib.rs:48-78 -> std::optional, empty on divergence
// The numerical work runs in libfock.so (fock_scf_run); there is no CPU path here.
#pragma once

#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <optional>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fock.h"

namespace fock_rs {

struct Vector3 {
    double x = 0, y = 0, z = 0;
};

// Rust f64::powi (square-and-multiply)
inline double powi(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

struct GaussianPrimitive {
    std::array<int, 3> angular{0, 0, 0};
    double exponent = 0;
    double coefficient = 0;

    static double norm(const std::array<int, 3> &ang, double alpha) {
        auto prod = [](int lo, int hi) {
            long p = 1;
            for (int v = lo; v <= hi; ++v) p *= v;
            return p;
        };
        const double a = std::sqrt(std::sqrt(powi(0.6366197723675814 * alpha, 3)));
        const double num = powi(8.0 * alpha, ang[0] + ang[1] + ang[2]);
        const double den = (double)(prod(ang[0] + 1, 2 * ang[0]) * prod(ang[1] + 1, 2 * ang[1]) * prod(ang[2] + 1, 2 * ang[2]));
        return a * std::sqrt(num / den);
    }
    static GaussianPrimitive new_normalized(std::array<int, 3> ang, double exponent, double coefficient) {
        return GaussianPrimitive{ang, exponent, coefficient * norm(ang, exponent)};
    }
};

struct ContractedGaussian {
    Vector3 position;
    std::vector<GaussianPrimitive> primitives;

    double evaluate(const Vector3 &at) const {
        const double dx = position.x - at.x, dy = position.y - at.y, dz = position.z - at.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        double acc = 0.0;
        for (const auto &g : primitives)
            acc += g.coefficient * powi(dx, g.angular[0]) * powi(dy, g.angular[1]) * powi(dz, g.angular[2]) *
                   std::exp(-g.exponent * d2);
        return acc;
    }
};

struct PointCharge {
    Vector3 position;
    double charge;
};

struct Atom {
    Vector3 position;
    int atom_type = 0;  // atomic number
    std::vector<ContractedGaussian> basis;
    size_t electron_count() const { return (size_t)atom_type; }
    PointCharge point_charge() const { return PointCharge{position, (double)atom_type}; }
};

// ------------------------------------------------------------------ minimal JSON (BSE subset)
namespace json {
struct Value {
    enum Kind { Null, Number, String, Array, Object } kind = Null;
    double num = 0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;
    const Value &operator[](const std::string &k) const {
        auto it = obj.find(k);
        if (it == obj.end()) throw std::runtime_error("json: missing key " + k);
        return it->second;
    }
    bool has(const std::string &k) const { return obj.count(k) != 0; }
};
inline void ws(const std::string &s, size_t &i) {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
}
inline Value parse(const std::string &s, size_t &i) {
    ws(s, i);
    Value v;
    if (i >= s.size()) throw std::runtime_error("json: unexpected end");
    char c = s[i];
    if (c == '{') {
        v.kind = Value::Object;
        ++i;
        ws(s, i);
        if (s[i] == '}') { ++i; return v; }
        for (;;) {
            ws(s, i);
            Value key = parse(s, i);
            ws(s, i);
            if (s[i] != ':') throw std::runtime_error("json: expected ':'");
            ++i;
            v.obj[key.str] = parse(s, i);
            ws(s, i);
            if (s[i] == ',') { ++i; continue; }
            if (s[i] == '}') { ++i; return v; }
            throw std::runtime_error("json: expected ',' or '}'");
        }
    }
    if (c == '[') {
        v.kind = Value::Array;
        ++i;
        ws(s, i);
        if (s[i] == ']') { ++i; return v; }
        for (;;) {
            v.arr.push_back(parse(s, i));
            ws(s, i);
            if (s[i] == ',') { ++i; continue; }
            if (s[i] == ']') { ++i; return v; }
            throw std::runtime_error("json: expected ',' or ']'");
        }
    }
    if (c == '"') {
        v.kind = Value::String;
        ++i;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) ++i;
            v.str.push_back(s[i++]);
        }
        ++i;
        return v;
    }
    if (c == 'n') { i += 4; return v; }
    if (c == 't' || c == 'f') { i += (c == 't') ? 4 : 5; v.kind = Value::Number; v.num = (c == 't'); return v; }
    size_t start = i;
    while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '-' || s[i] == '+' || s[i] == '.' ||
                            s[i] == 'e' || s[i] == 'E'))
        ++i;
    v.kind = Value::Number;
    v.num = std::strtod(s.substr(start, i - start).c_str(), nullptr);
    return v;
}
}  // namespace json

// Cartesian components of a shell in the reference's order (basis_set/src/lib.rs:87-99)
inline std::vector<std::array<int, 3>> cartesian_components(int L) {
    std::vector<std::array<int, 3>> out;
    for (int x = 0; x <= L; ++x)
        for (int y = 0; y <= L; ++y)
            for (int z = 0; z <= L; ++z)
                if (x + y + z == L) out.push_back({x, y, z});
    return out;
}

class BasisSet {
  public:
    static BasisSet load(const std::string &path) {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("cannot open basis set " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        size_t i = 0;
        BasisSet b;
        b.root_ = std::make_shared<json::Value>(json::parse(text, i));
        return b;
    }
    // basis_set/src/lib.rs:79-119 (panics there when the element is missing; throws here)
    Atom get(const Vector3 &position, int atom_type) const {
        const auto &els = (*root_)["elements"];
        const std::string key = std::to_string(atom_type);
        if (!els.has(key)) throw std::runtime_error("Failed to get atom from basis set");
        Atom atom{position, atom_type, {}};
        for (const auto &shell : els[key]["electron_shells"].arr) {
            const auto &am = shell["angular_momentum"].arr;
            for (size_t idx = 0; idx < am.size(); ++idx) {
                const int L = (int)am[idx].num;
                for (const auto &ang : cartesian_components(L)) {
                    ContractedGaussian cg{position, {}};
                    const auto &ex = shell["exponents"].arr;
                    const auto &co = shell["coefficients"].arr[idx].arr;
                    for (size_t k = 0; k < ex.size(); ++k)
                        cg.primitives.push_back(GaussianPrimitive::new_normalized(
                            ang, std::strtod(ex[k].str.c_str(), nullptr), std::strtod(co[k].str.c_str(), nullptr)));
                    atom.basis.push_back(cg);
                }
            }
        }
        return atom;
    }

  private:
    std::shared_ptr<json::Value> root_;
};

// chemfiles/src/xyz.rs:13-31
inline std::vector<Atom> read_xyz_file(const std::string &path, const BasisSet &basis) {
    static const std::map<std::string, int> sym = {{"H", 1}, {"He", 2}, {"C", 6}, {"N", 7},
                                                   {"O", 8}, {"F", 9},  {"Fe", 26}, {"B", 5}};
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string content = ss.str();
    const std::regex pat(R"(\s*(\w+)\s+([+\-0-9.]+)\s+([+\-0-9.]+)\s+([0-9.+\-]+).*\b)");
    std::vector<Atom> atoms;
    for (auto it = std::sregex_iterator(content.begin(), content.end(), pat); it != std::sregex_iterator(); ++it) {
        const auto &m = *it;
        auto s = sym.find(m[1].str());
        if (s == sym.end()) throw std::runtime_error("not implemented yet");
        atoms.push_back(basis.get(Vector3{std::stod(m[2].str()) * 1.89, std::stod(m[3].str()) * 1.89,
                                          std::stod(m[4].str()) * 1.89},
                                  s->second));
    }
    return atoms;
}

struct MolecularOrbital {
    std::vector<ContractedGaussian> basis_functions;
    std::vector<double> coefficients;
    double evaluate(const Vector3 &at) const {
        double s = 0.0;
        for (size_t i = 0; i < coefficients.size(); ++i) s += coefficients[i] * basis_functions[i].evaluate(at);
        return s;
    }
};

struct MolecularOrbitals {
    static constexpr double MIN_COEFFICIENT_MAGNITUDE = 0.05;
    std::vector<MolecularOrbital> molecular_orbitals;
    // coeff: column-major n x n
    MolecularOrbitals(const std::vector<ContractedGaussian> &basis, const std::vector<double> &coeff, size_t n) {
        for (size_t k = 0; k < n; ++k) {
            MolecularOrbital mo;
            for (size_t i = 0; i < n; ++i) {
                const double c = coeff[i + k * n];
                if (std::fabs(c) > MIN_COEFFICIENT_MAGNITUDE) {
                    mo.basis_functions.push_back(basis[i]);
                    mo.coefficients.push_back(c);
                }
            }
            molecular_orbitals.push_back(std::move(mo));
        }
    }
    const MolecularOrbital &operator[](size_t i) const { return molecular_orbitals[i]; }
};

struct HartreeFockResult {
    MolecularOrbitals orbitals;
    std::vector<double> orbital_energies;
    std::vector<double> density_matrix;      // column-major n x n
    std::vector<double> coefficient_matrix;  // column-major n x n
    double electronic_energy;
    double total_energy;
    size_t iterations;
    size_t n_electrons;
    size_t n_basis;
};

class Context {
  public:
    explicit Context(int device = 0) {
        if (fock_ctx_create(device, &ctx_) != FOCK_OK) throw std::runtime_error("fock_ctx_create failed");
    }
    ~Context() { fock_ctx_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    fock_ctx *get() const { return ctx_; }

  private:
    fock_ctx *ctx_ = nullptr;
};

// scf/src/lib.rs:73-258 on the GPU; opts default to the reference's arithmetic.
inline std::optional<HartreeFockResult> try_scf(const std::vector<Atom> &atoms, size_t max_iters, double epsilon,
                                                int molecule_charge, Context &ctx,
                                                const fock_scf_opts *base_opts = nullptr) {
    std::vector<ContractedGaussian> basis;
    for (const auto &a : atoms) basis.insert(basis.end(), a.basis.begin(), a.basis.end());
    const size_t n = basis.size();
    std::vector<double> centers, exps, coefs, pos, q;
    std::vector<int32_t> ang, off{0};
    for (const auto &f : basis) {
        centers.insert(centers.end(), {f.position.x, f.position.y, f.position.z});
        ang.insert(ang.end(), {f.primitives[0].angular[0], f.primitives[0].angular[1], f.primitives[0].angular[2]});
        for (const auto &p : f.primitives) {
            exps.push_back(p.exponent);
            coefs.push_back(p.coefficient);
        }
        off.push_back((int32_t)exps.size());
    }
    for (const auto &a : atoms) {
        const PointCharge pc = a.point_charge();
        pos.insert(pos.end(), {pc.position.x, pc.position.y, pc.position.z});
        q.push_back(pc.charge);
    }
    fock_basis_desc bd{(int32_t)n, centers.data(), ang.data(), off.data(), exps.data(), coefs.data()};
    fock_nuclei_desc nd{(int32_t)atoms.size(), pos.data(), q.data()};
    fock_scf_opts opts;
    if (base_opts) opts = *base_opts;
    else fock_scf_opts_default(&opts);
    opts.max_iters = (int32_t)max_iters;
    opts.epsilon = epsilon;
    opts.molecule_charge = molecule_charge;
    std::vector<double> C(n * n), P(n * n), eps(n);
    fock_scf_out out{};
    out.coefficients = C.data();
    out.density = P.data();
    out.orbital_energies = eps.data();
    const fock_status st = fock_scf_run(ctx.get(), &bd, &nd, &opts, &out);
    if (st == FOCK_DIVERGED) return std::nullopt;
    if (st != FOCK_OK && st != FOCK_NOT_CONVERGED) throw std::runtime_error(fock_ctx_last_error(ctx.get()));
    return HartreeFockResult{MolecularOrbitals(basis, C, n),
                             eps,
                             P,
                             C,
                             out.electronic_energy,
                             out.total_energy,
                             (size_t)out.iterations,
                             (size_t)out.n_electrons,
                             n};
}

}  // namespace fock_rs
