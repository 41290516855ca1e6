// cli.cpp -- rbm-matfree: drop-in replacement for the reference executables
// rbm-matfree-N{N}-M{M}-K{K}[-opt] (code/build.sh:40-41) and, installed under the
// name rbm-N{N}-M{M}-K{K}[-opt], for the dense-S executables (code/build.sh:37-39),
// on top of libnqes.
//
// Same options and semantics as the reference's CmdArgs (rbm.cc:900-953, parsed
// with the rules of code/argparse.h:435-554), same stdout (seed line, one line of
// K ascending E/N values per SR step, convergence line; rbm.cc:1126, 1188, 1190),
// same files (--dump-rbm, --correlation-data; rbm.cc:406-471, 828-889) and exit
// codes (0; 1 for -c 0; 255 for option errors; 127 for a floating-point error with
// --throw-fperror).  N, M, K come from the executable name when it contains
// N<n>-M<m>-K<k> (install it under the reference's names), else from
// --num-spins/--num-hidden/--num-states; J comes from --jmat FILE.mat (variable J)
// or $NQES_JMAT_DIR/Ising2d_N<N>_alpha1.0.mat.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <optional>
#include <random>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nqes.h"
#include "eig.hpp"
#include "matv5.hpp"
#include "srloop.hpp"

using cplx = std::complex<double>;

namespace {

// ------------------------------------------------------------------ options
struct Opt {
    std::vector<std::string> keys;
    std::string help;
    bool is_flag = false;
    bool optional_ = false;  // may stay unset (no default, not required)
    std::optional<std::string> dflt;
    std::optional<std::string> value;
    std::string error;
    std::string keys_str() const
    {
        std::string s;
        for (size_t i = 0; i < keys.size(); i++) s += (i ? "," : "") + std::string(keys[i].size() > 1 ? "--" : "-") + keys[i];
        return s;
    }
};

struct Options {
    std::vector<std::unique_ptr<Opt>> all;
    std::map<std::string, Opt *> by_key;
    Opt *add(const std::string &keys, const std::string &help, std::optional<std::string> dflt, bool flag = false)
    {
        auto o = std::make_unique<Opt>();
        std::stringstream ss(keys);
        std::string k;
        while (std::getline(ss, k, ',')) o->keys.push_back(k);
        o->help = help;
        o->dflt = dflt;
        o->is_flag = flag;
        for (auto &kk : o->keys) by_key[kk] = o.get();
        all.push_back(std::move(o));
        return all.back().get();
    }
    std::string get(const std::string &key) const { return *by_key.at(key)->value; }
    bool has(const std::string &key) const { return by_key.at(key)->value.has_value(); }
};

bool is_value(const std::vector<std::string> &p, size_t i)
{
    return i < p.size() && (p[i].empty() || p[i][0] != '-' || (p[i].size() > 1 && std::isdigit((unsigned char)p[i][1])));
}

void parse_args(Options &o, const std::vector<std::string> &params, bool &help)
{
    auto set = [&](const std::string &key, size_t &i, bool is_short, const std::optional<std::string> &eq) {
        auto it = o.by_key.find(key);
        if (it == o.by_key.end()) {
            std::cerr << "unrecognised commandline argument: " << key << std::endl;
            return;
        }
        Opt *e = it->second;
        if (eq) e->value = *eq;
        else if (e->is_flag) e->value = "true";
        else if (!is_short && is_value(params, i + 1)) e->value = params[++i];
        else e->error = "No value provided for: " + key;
    };
    auto add_param = [&](size_t &i, size_t start) {
        size_t eq = params[i].find('=');
        if (eq != std::string::npos) set(params[i].substr(start, eq - start), i, false, params[i].substr(eq + 1));
        else set(params[i].substr(start), i, false, std::nullopt);
    };
    for (size_t i = 0; i < params.size(); i++) {
        if (is_value(params, i)) continue;  // no positional arguments
        if (params[i].size() > 1 && params[i][1] == '-') {
            if (params[i] == "--help") {
                help = true;
                continue;
            }
            add_param(i, 2);
        } else {
            if (params[i] == "-?") {
                help = true;
                continue;
            }
            size_t j_end = std::min(params[i].size(), params[i].find('=')) - 1;
            for (size_t j = 1; j < j_end; j++) set(std::string(1, params[i][j]), i, true, std::nullopt);
            add_param(i, j_end);
        }
    }
    for (auto &e : o.all)
        if (!e->value && e->error.empty()) {
            if (e->dflt) e->value = *e->dflt;
            else if (!e->is_flag && !e->optional_) e->error = "Argument missing: " + e->keys_str() + " (" + e->help + ")";
        }
}

template <typename T> T conv(Options &o, const std::string &key, T (*f)(const std::string &))
{
    Opt *e = o.by_key.at(key);
    try {
        return f(*e->value);
    } catch (const std::exception &) {
        std::cerr << "Invalid argument, could not convert \"" << *e->value << "\" for " << e->keys_str() << " ("
                  << e->help << ")" << std::endl;
        exit(-1);
    }
}
int to_int(const std::string &s) { return std::stoi(s); }
double to_dbl(const std::string &s) { return std::stod(s); }
unsigned long long to_u64(const std::string &s) { return std::stoull(s); }
bool to_bool(const std::string &s)
{
    std::string l = s;
    for (auto &ch : l) ch = (char)std::tolower((unsigned char)ch);
    return l == "true" || s == "1";
}

// ------------------------------------------------------------------ pcg64 and the
// libstdc++ normal distribution (weight initialisation, rbm.cc:242-270)
struct Pcg64 {
    unsigned __int128 st;
    static constexpr unsigned __int128 MULT =
        ((unsigned __int128)2549297995355413924ULL << 64) | 4865540595714422341ULL;
    static constexpr unsigned __int128 INC =
        ((unsigned __int128)6364136223846793005ULL << 64) | 1442695040888963407ULL;
    explicit Pcg64(uint64_t seed) { st = ((unsigned __int128)seed + INC) * MULT + INC; }
    uint64_t operator()()
    {
        st = st * MULT + INC;
        unsigned rot = (unsigned)(st >> 122) & 63u;
        uint64_t x = (uint64_t)(st >> 64) ^ (uint64_t)st;
        return (x >> rot) | (x << ((64u - rot) & 63u));
    }
    double canonical()
    {
        double r = (double)(*this)() * 0x1p-64;
        return r >= 1.0 ? 0x1.fffffffffffffp-1 : r;
    }
};

struct Normal {  // Marsaglia polar method with a cached second value
    bool saved_ok = false;
    double saved = 0;
    double operator()(Pcg64 &g)
    {
        if (saved_ok) {
            saved_ok = false;
            return saved;
        }
        double x, y, r2;
        do {
            x = 2.0 * g.canonical() - 1.0;
            y = 2.0 * g.canonical() - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        double mult = std::sqrt(-2 * std::log(r2) / r2);
        saved = x * mult;
        saved_ok = true;
        return y * mult;
    }
};

// ------------------------------------------------------------------ weights <-> .mat
// W layout: RBM k at [k*L, (k+1)*L) = a[N], b[M], w[N][M]
std::vector<cplx> load_rbm(const std::string &path, int N, int M, int K)
{
    auto vars = matv5::read_file(path);
    auto need = [&](const char *name, std::vector<int64_t> dims) -> const matv5::Var & {
        const matv5::Var *v = matv5::find(vars, name);
        if (!v) throw std::runtime_error(path + ": cannot read variable " + name);
        if (v->mclass != 6) throw std::runtime_error(path + ": variable " + name + ": floating point type mismatch");
        if (!v->complex_) throw std::runtime_error(path + ": variable " + name + ": not a complex matrix");
        if (v->dims.size() != dims.size())
            throw std::runtime_error(path + ": variable " + name + ": not a rank " + std::to_string(dims.size()) +
                                     " matrix");
        for (size_t d = 0; d < dims.size(); d++)
            if (v->dims[d] != dims[d])
                throw std::runtime_error(path + ": variable " + name + ": invalid dimension " + std::to_string(d));
        return *v;
    };
    const auto &a = need("a", {K, N});
    const auto &b = need("b", {K, M});
    const auto &w = need("w", {K, N, M});
    const long L = N + M + (long)N * M;
    std::vector<cplx> W((size_t)K * L);
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < N; i++) W[k * L + i] = cplx(a.re[(size_t)i * K + k], a.im[(size_t)i * K + k]);
        for (int j = 0; j < M; j++) W[k * L + N + j] = cplx(b.re[(size_t)j * K + k], b.im[(size_t)j * K + k]);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < M; j++) {
                size_t src = (size_t)j * K * N + (size_t)i * K + k;
                W[k * L + N + M + (long)i * M + j] = cplx(w.re[src], w.im[src]);
            }
    }
    return W;
}

void dump_rbm(const std::string &path, const std::vector<cplx> &W, int N, int M, int K)
{
    const long L = N + M + (long)N * M;
    matv5::Var a{"a", {K, N}, true, 6, {}, {}}, b{"b", {K, M}, true, 6, {}, {}}, w{"w", {K, N, M}, true, 6, {}, {}};
    a.re.resize((size_t)K * N);
    a.im.resize((size_t)K * N);
    b.re.resize((size_t)K * M);
    b.im.resize((size_t)K * M);
    w.re.resize((size_t)K * N * M);
    w.im.resize((size_t)K * N * M);
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < N; i++) {
            a.re[(size_t)i * K + k] = W[k * L + i].real();
            a.im[(size_t)i * K + k] = W[k * L + i].imag();
        }
        for (int j = 0; j < M; j++) {
            b.re[(size_t)j * K + k] = W[k * L + N + j].real();
            b.im[(size_t)j * K + k] = W[k * L + N + j].imag();
        }
        for (int i = 0; i < N; i++)
            for (int j = 0; j < M; j++) {
                size_t dst = (size_t)j * K * N + (size_t)i * K + k;
                w.re[dst] = W[k * L + N + M + (long)i * M + j].real();
                w.im[dst] = W[k * L + N + M + (long)i * M + j].imag();
            }
    }
    matv5::write_file(path, {a, b, w});
}

// ------------------------------------------------------------------ couplings
// Couplings of the models of rbm.cc:965-1031 (diagonal: E_zz(s) = 2 sum_{i>j} J_ij s_i s_j)
// and rbm.cc:1066-1089 (pair-flip: J_ij of sum_{i>j} J_ij (XX + YY + delta ZZ))
std::vector<double> model_J(const std::string &model, int N, const std::string &jmat)
{
    std::vector<double> J((size_t)N * N, 0.0);
    if (model == "trappedion") {
        std::string path = jmat;
        if (path.empty()) {
            const char *dir = getenv("NQES_JMAT_DIR");
            if (!dir) throw std::runtime_error("no coupling matrix: pass --jmat FILE.mat or set NQES_JMAT_DIR");
            path = std::string(dir) + "/Ising2d_N" + std::to_string(N) + "_alpha1.0.mat";
        }
        auto vars = matv5::read_file(path);
        const matv5::Var *v = matv5::find(vars, "J");
        if (!v || v->dims.size() != 2 || v->dims[0] != N || v->dims[1] != N)
            throw std::runtime_error(path + ": no " + std::to_string(N) + "x" + std::to_string(N) + " variable J");
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) J[(size_t)i * N + j] = v->re[(size_t)j * N + i];
    } else if (model == "ising" || model == "periodicising") {
        // -sum_i s_i s_{i-1} (open) [+ s_0 s_{N-1} bond, periodic]
        for (int i = 1; i < N; i++) J[(size_t)i * N + i - 1] = J[(size_t)(i - 1) * N + i] = -0.5;
        if (model == "periodicising" && N > 1) {
            J[(size_t)(N - 1) * N] += -0.5;
            J[N - 1] += -0.5;
        }
    } else if (model == "longrange") {
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++)
                if (i != j) J[(size_t)i * N + j] = std::pow((double)std::abs(i - j), -0.2);
    } else if (model == "hs") {
        // hs_model (rbm.cc:1066-1078): J_ij = 1/d^2, d = sin(pi (i-j)/N) N/pi, i > j
        for (int i = 1; i < N; i++)
            for (int j = 0; j < i; j++) {
                const double d = std::sin(M_PI * (i - j) / N) * N / M_PI;
                J[(size_t)i * N + j] = 1.0 / (d * d);
            }
    } else if (model == "xxz") {
        // xxz_model (rbm.cc:1083-1089): -1 on the bonds (i, i-1) and (N-1, 0)
        for (int i = 1; i < N; i++)
            for (int j = 0; j < i; j++)
                if (i - j == 1 || (i == N - 1 && j == 0)) J[(size_t)i * N + j] = -1.0;
    } else {
        throw std::runtime_error("unknown model " + model);
    }
    return J;
}

// the library's model: ising-like (trappedion and the diagonal models) or hs-like
int model_id(const std::string &model)
{
    if (model == "hs") return NQES_MODEL_HS;
    if (model == "xxz") return NQES_MODEL_XXZ;
    return NQES_MODEL_TRAPPEDION;
}

// Eigen's default row-vector formatting: stream precision 6, every coefficient right
// aligned to the widest one, single spaces between
std::string format_row(const std::vector<double> &v)
{
    std::vector<std::string> s;
    size_t w = 0;
    for (double x : v) {
        std::ostringstream os;
        os << x;
        s.push_back(os.str());
        w = std::max(w, s.back().size());
    }
    std::ostringstream out;
    for (size_t i = 0; i < s.size(); i++) out << (i ? " " : "") << std::setw((int)w) << s[i];
    return out.str();
}

using srloop::sorted_real_eigs;

// dump_states (rbm.cc:474-526): variable s, complex, dims [2^(K N), rows, cols] with
// unit dimensions dropped, column-major; fcn(cfg) is Phi (what 0), H_loc (1) or O (2)
// of the multi-spin configuration cfg (replica k = bits [k N, k N + N)).  Phi^+ is
// the pseudo-inverse, as in rbm.cc:1158, 1167.
void dump_configs(nqes_ctx *ctx, const std::string &prefix, int what, int N, int M, int K,
                  const std::vector<cplx> &W)
{
    const int64_t dim0 = 1ll << (N * K), L = N + M + (int64_t)N * M, P = K * L;
    const int rows = what == 2 ? (int)P : K, cols = what == 2 ? 1 : K;
    std::vector<int64_t> dims{dim0};
    if (rows > 1) dims.push_back(rows);
    if (cols > 1) dims.push_back(cols);
    matv5::Var v{"s", dims, true, 6, {}, {}};
    v.re.assign((size_t)v.numel(), 0.0);
    v.im.assign((size_t)v.numel(), 0.0);
    const int64_t batch = 1 << 16;
    const size_t kk = (size_t)K * K;
    std::vector<cplx> phi(batch * kk), res(what == 1 ? batch * kk : 0), th(what == 2 ? batch * kk * M : 0);
    std::vector<cplx> pinv(kk), G(kk), X(kk), val((size_t)rows * cols), colk(K);
    for (int64_t c0 = 0; c0 < dim0; c0 += batch) {
        const int64_t nb = std::min(batch, dim0 - c0);
        int st = nqes_enumerate(ctx, reinterpret_cast<const double *>(W.data()), c0, nb,
                                reinterpret_cast<double *>(phi.data()),
                                what == 1 ? reinterpret_cast<double *>(res.data()) : nullptr,
                                what == 2 ? reinterpret_cast<double *>(th.data()) : nullptr);
        if (st != NQES_OK) throw std::runtime_error(nqes_last_error(ctx));
        for (int64_t e = 0; e < nb; e++) {
            const cplx *ph = phi.data() + e * kk;
            if (what == 0) {
                for (size_t q = 0; q < kk; q++) val[q] = ph[q];
            } else {
                eigk::pinv(K, ph, pinv.data());
                if (what == 1) {
                    const cplx *r = res.data() + e * kk;
                    for (int i = 0; i < K; i++)
                        for (int j = 0; j < K; j++) {
                            cplx acc(0.0, 0.0);
                            for (int l = 0; l < K; l++) acc += pinv[i * K + l] * r[l * K + j];
                            val[i * K + j] = acc;
                        }
                } else {
                    // deriv (rbm.cc:199-239): det(Phi^+ derPhi); columns != k of Phi^+ derPhi
                    // are those of G = Phi^+ Phi
                    const uint64_t cfg = (uint64_t)(c0 + e);
                    const cplx *t = th.data() + e * kk * M;
                    for (int i = 0; i < K; i++)
                        for (int j = 0; j < K; j++) {
                            cplx acc(0.0, 0.0);
                            for (int l = 0; l < K; l++) acc += pinv[i * K + l] * ph[l * K + j];
                            G[i * K + j] = acc;
                        }
                    auto spin = [&](int kk2, int i) { return (cfg >> (kk2 * N + i)) & 1u; };
                    for (int64_t p = 0; p < P; p++) {
                        const int k = (int)(p / L);
                        const int64_t r = p % L;
                        for (int kk2 = 0; kk2 < K; kk2++) {
                            const cplx f = ph[kk2 * K + k];
                            if (r < N) colk[kk2] = spin(kk2, (int)r) ? -f : f;
                            else if (r < N + M) colk[kk2] = f * t[(kk2 * K + k) * M + (r - N)];
                            else {
                                const int i = (int)((r - N - M) / M), j = (int)((r - N - M) % M);
                                const cplx fv = f * t[(kk2 * K + k) * M + j];
                                colk[kk2] = spin(kk2, i) ? -fv : fv;
                            }
                        }
                        for (int i = 0; i < K; i++) {
                            cplx y(0.0, 0.0);
                            for (int l = 0; l < K; l++) y += pinv[i * K + l] * colk[l];
                            for (int j = 0; j < K; j++) X[i * K + j] = (j == k) ? y : G[i * K + j];
                        }
                        val[p] = eigk::det(K, X.data());
                    }
                }
            }
            const int64_t cfg = c0 + e;
            for (int i = 0; i < rows; i++)
                for (int j = 0; j < cols; j++) {
                    const size_t dst = (size_t)j * rows * dim0 + (size_t)i * dim0 + cfg;
                    v.re[dst] = val[(size_t)i * cols + j].real();
                    v.im[dst] = val[(size_t)i * cols + j].imag();
                }
        }
    }
    matv5::write_file(prefix + "N" + std::to_string(N) + "M" + std::to_string(M) + "K" + std::to_string(K) + ".mat",
                      {v});
}

[[noreturn]] void fp_error()
{
    std::cerr << "Floating point error occured. Exiting." << std::endl;
    exit(127);
}

} // namespace

int main(int argc, char **argv)
{
    Options o;
    o.add("model", "Model to simulate [trappedion, ising, periodicising, longrange, hs, xxz]", std::nullopt);
    o.add("h", "External field h", std::nullopt);
    o.add("n,niters", "Number of SR iterations", std::nullopt);
    o.add("m,mc-steps-percpu", "Monte Carlo samples per chain", std::nullopt);
    o.add("c,ncpus", "Number of Markov chains", std::nullopt);
    o.add("burn-in-steps", "Burn-in sweeps per chain", "64");
    o.add("tol", "Stopping tolerance of the per-site energies", "1e-06");
    o.add("lrate-init", "Initial learning rate, lrate(t) = lrate(0) f^t", "0.1");
    o.add("lrate-decay", "Decay factor f of the learning rate", "0.99");
    o.add("lrate-cutoff", "Minimal learning rate", "0.01");
    o.add("reglam-init", "Initial lambda regulating S_kk", "0.01");
    o.add("reglam-decay", "Decay factor of lambda", "0.9");
    o.add("reglam-cutoff", "Minimal lambda", "0.001");
    o.add("minres-rtol", "MINRES-QLP tolerance", "1e-10");
    o.add("minres-maxit", "MINRES-QLP iteration limit", "18446744073709551615");
    o.add("seed", "Seed of the RNG (default: random)", std::nullopt)->optional_ = true;
    o.add("load-rbm", "Load the initial RBM weights from this .mat file", "");
    o.add("load-single-rbm", "Load one RBM for all K, spin-flipped except the first", "");
    o.add("dump-rbm", "Write the final weights to <prefix>N<N>M<M>K<K>.mat", "");
    o.add("throw-fperror", "Exit with code 127 on a floating point error", "false", true);
    o.add("apply-gauge", "Apply a_i -> a_i - i pi/2 (odd i) to dumped weights", "false", true);
    o.add("inverse-gauge", "With --apply-gauge: a_i -> a_i + i pi/2 (odd i) on loaded weights", "false", true);
    o.add("correlation-data", "File name for the correlation function data", std::nullopt);
    o.add("dump-states", "If N*K < 32: dump Phi of every spin configuration to <prefix>N<N>M<M>K<K>.mat", "");
    o.add("dump-ham", "If N*K < 32: dump H_loc of every spin configuration", "");
    o.add("dump-deriv", "If N*K < 32: dump the derivatives O of every spin configuration", "");
    // extensions of this build
    o.add("num-spins", "N (default: from the executable name)", "0");
    o.add("num-hidden", "M (default: from the executable name)", "0");
    o.add("num-states", "K (default: from the executable name)", "0");
    o.add("jmat", "MAT file holding the N x N couplings J (trappedion)", "");
    o.add("precision", "fp64 or mixed", "fp64");
    o.add("device", "HIP device", "0");
    o.add("init-real-sigma", "If > 0: initial weights real N(0, sigma^2) instead of the complex default", "0");
    o.add("stats", "Print per-step timings to stderr", "false", true);

    std::vector<std::string> params(argv + 1, argv + argc);
    bool help = false;
    parse_args(o, params, help);
    std::string prog = argv[0];
    if (help) {
        std::cout << "Usage: " << prog << "  [options...]" << std::endl << std::endl << "Options:" << std::endl;
        for (auto &e : o.all)
            std::cout << std::setw(17) << e->keys_str() << " : " << e->help << " ["
                      << (e->dflt ? "default: " + *e->dflt : (e->optional_ ? "optional" : "required")) << "]"
                      << std::endl;
        return 0;
    }
    for (auto &e : o.all)
        if (!e->error.empty()) {
            std::cerr << e->error << std::endl;
            return 255;
        }

    // N, M, K
    int N = conv<int>(o, "num-spins", to_int), M = conv<int>(o, "num-hidden", to_int),
        K = conv<int>(o, "num-states", to_int);
    std::smatch mt;
    std::string base = prog.substr(prog.find_last_of('/') == std::string::npos ? 0 : prog.find_last_of('/') + 1);
    if (std::regex_search(base, mt, std::regex("N([0-9]+)-M([0-9]+)-K([0-9]+)"))) {
        if (!N) N = std::stoi(mt[1]);
        if (!M) M = std::stoi(mt[2]);
        if (!K) K = std::stoi(mt[3]);
    }
    const bool dense_name = base.rfind("rbm-N", 0) == 0;
    if (!N || !M || !K) {
        std::cerr << "N, M, K unknown: name the executable rbm-matfree-N<N>-M<M>-K<K> or pass "
                     "--num-spins/--num-hidden/--num-states"
                  << std::endl;
        return 255;
    }
    // the reference defines the dump options only when N*K < 32 (rbm.cc:943-953)
    for (const char *d : {"dump-states", "dump-ham", "dump-deriv"})
        if (N * K >= 32 && !o.get(d).empty()) {
            std::cerr << "Unrecognized argument --" << d << " (only for N*K < 32)" << std::endl;
            return 255;
        }
    std::string model = o.get("model");
    for (auto &ch : model) ch = (char)std::tolower((unsigned char)ch);
    const double h = conv<double>(o, "h", to_dbl);
    const int niters = conv<int>(o, "niters", to_int);
    const int m = conv<int>(o, "m", to_int);
    const int C = conv<int>(o, "c", to_int);
    const int B = conv<int>(o, "burn-in-steps", to_int);
    const double tol = conv<double>(o, "tol", to_dbl);
    const double lrate0 = conv<double>(o, "lrate-init", to_dbl), lratef = conv<double>(o, "lrate-decay", to_dbl),
                 lrate_cut = conv<double>(o, "lrate-cutoff", to_dbl);
    const double reglam0 = conv<double>(o, "reglam-init", to_dbl),
                 reglam_decay = conv<double>(o, "reglam-decay", to_dbl),
                 reglam_cut = conv<double>(o, "reglam-cutoff", to_dbl);
    const double rtol = conv<double>(o, "minres-rtol", to_dbl);
    const unsigned long long maxit = conv<unsigned long long>(o, "minres-maxit", to_u64);
    const bool throw_fp = to_bool(o.get("throw-fperror"));
    const bool apply_gauge = to_bool(o.get("apply-gauge")), inverse_gauge = to_bool(o.get("inverse-gauge"));
    const bool want_stats = to_bool(o.get("stats"));
    const std::string corr_file = o.get("correlation-data");
    const double sigma0 = conv<double>(o, "init-real-sigma", to_dbl);
    int precision;
    if (o.get("precision") == "fp64") precision = NQES_PREC_FP64;
    else if (o.get("precision") == "mixed") precision = NQES_PREC_MIXED;
    else {
        std::cerr << "Invalid argument \"" << o.get("precision") << "\" for --precision (fp64 or mixed)" << std::endl;
        return 255;
    }
    const char *models[] = {"ising", "periodicising", "hs", "xxz", "longrange", "trappedion"};
    bool known = false;
    for (auto mm : models) known |= model == mm;
    if (!known) {
        std::cerr << "Invalid argument \"" << o.get("model") << "\" for --model (Model to simulate). Error: enum is "
                     "only accepting [ising, periodicising, hs, xxz, longrange, trappedion]"
                  << std::endl;
        return 255;
    }
    if (!C) {
        std::cerr << "Number of CPU cores cannot be zero." << std::endl;
        return 1;
    }

    try {
        std::vector<double> J = model_J(model, N, o.get("jmat"));
        std::random_device rd;
        const uint64_t seed = o.has("seed") ? conv<unsigned long long>(o, "seed", to_u64) : (uint64_t)rd();
        std::cout << "Random seed is " << seed << std::endl;

        const long L = N + M + (long)N * M;
        std::vector<cplx> W;
        if (!o.get("load-rbm").empty()) {
            W = load_rbm(o.get("load-rbm"), N, M, K);
            if (apply_gauge)
                for (int k = 0; k < K; k++)
                    for (int i = 1; i < N; i += 2) W[k * L + i] += cplx(0, inverse_gauge ? M_PI_2 : -M_PI_2);
        } else if (!o.get("load-single-rbm").empty()) {
            auto W1 = load_rbm(o.get("load-single-rbm"), N, M, 1);
            W.assign((size_t)K * L, 0.0);
            for (int k = 0; k < K; k++)
                for (long q = 0; q < L; q++) {
                    bool b_block = q >= N && q < N + M;
                    W[k * L + q] = (k == 0 || b_block) ? W1[q] : -W1[q];
                }
        } else {
            Pcg64 rg(seed - 1);
            Normal nd;
            W.assign((size_t)K * L, 0.0);
            for (int k = 0; k < K; k++) {
                if (sigma0 > 0) {
                    for (long q = 0; q < L; q++) W[k * L + q] = cplx(sigma0 * nd(rg), 0.0);
                    continue;
                }
                const double sn = std::sqrt((double)N), sm = std::sqrt((double)M), snm = std::sqrt((double)N * M);
                for (int i = 0; i < N; i++) {
                    double re = nd(rg), im = nd(rg);
                    W[k * L + i] = cplx(re / sn, im / sn);
                }
                for (int j = 0; j < M; j++) {
                    double re = nd(rg), im = nd(rg);
                    W[k * L + N + j] = cplx(re / sm, im / sm);
                }
                for (long q = 0; q < (long)N * M; q++) {
                    double re = nd(rg), im = nd(rg);
                    W[k * L + N + M + q] = cplx(re / snm, im / snm);
                }
            }
        }

        nqes_ctx *ctx = nqes_create(conv<int>(o, "device", to_int));
        if (!ctx) throw std::runtime_error("no HIP device");
        auto check = [&](int st) {
            if (st == NQES_OK) return;
            if (st == NQES_ENONFINITE && throw_fp) fp_error();
            throw std::runtime_error(nqes_last_error(ctx));
        };
        check(nqes_problem_init(ctx, N, M, K, h, J.data(), model_id(model), precision));
        // installed as rbm-N<N>-M<M>-K<K>[-opt] (the reference's dense-S build, code/build.sh:37-39):
        // S as a dense matrix on the FP64 path
        if (dense_name && precision == NQES_PREC_FP64) check(nqes_set_dense_S(ctx, 1));
        // --dump-states/ham/deriv with the initial weights (rbm.cc:1149-1172)
        const std::string dumps[3] = {o.get("dump-states"), o.get("dump-ham"), o.get("dump-deriv")};
        for (int what = 0; what < 3; what++)
            if (!dumps[what].empty()) dump_configs(ctx, dumps[what], what, N, M, K, W);

        srloop::Schedule sched{lrate0, lratef, lrate_cut, reglam0, reglam_decay, reglam_cut};
        srloop::Convergence conv_test(K);
        std::vector<cplx> emat(K * K);
        for (int iter = 0; iter < niters; iter++) {
            nqes_stats st;
            int rc = nqes_sr_step(ctx, reinterpret_cast<double *>(W.data()), seed + (uint64_t)iter * C, C, m, B,
                                  sched.step_lrate(), sched.step_reglam(), rtol, maxit,
                                  reinterpret_cast<double *>(emat.data()), &st);
            if (rc == NQES_ENONFINITE && !throw_fp) {
                // the reference would carry NaNs on; report and stop cleanly
                for (auto &e : emat) e /= (double)N;
                std::cout << format_row(sorted_real_eigs(emat, K)) << std::endl;
                std::cerr << "non-finite local energy or gradient; stopping (use --throw-fperror for exit code 127)"
                          << std::endl;
                break;
            }
            check(rc);
            for (auto &e : emat) e /= (double)N;
            sched.advance();
            std::vector<double> evals = sorted_real_eigs(emat, K);
            std::cout << format_row(evals) << std::endl;
            if (want_stats)
                std::cerr << "step " << iter << ": " << st.t_total_ms << " ms (sample " << st.t_sample_ms
                          << ", solve " << st.t_solve_ms << ", " << st.minres_iters << " MINRES-QLP iterations, "
                          << st.ulp_ties << " near-tie accept decisions)" << std::endl;
            if (conv_test.update(iter, evals, tol)) {
                std::cout << "Converged after " << iter + 1 << " steps." << std::endl;
                break;
            }
        }

        // correlation functions of the final state (rbm.cc:1198-1229)
        std::vector<cplx> ev(K), V(K * K);
        eigk::eig(K, emat.data(), ev.data(), V.data());
        std::vector<int> order(K);
        for (int k = 0; k < K; k++) order[k] = k;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return ev[a].real() < ev[b].real(); });
        std::vector<cplx> Vinv(K * K);
        {
            // inverse of the eigenvector matrix by Gauss-Jordan with partial pivoting
            std::vector<cplx> A(V), I(K * K, 0.0);
            for (int k = 0; k < K; k++) I[k * K + k] = 1.0;
            for (int col = 0; col < K; col++) {
                int p = col;
                for (int r = col + 1; r < K; r++)
                    if (std::abs(A[r * K + col]) > std::abs(A[p * K + col])) p = r;
                for (int c2 = 0; c2 < K; c2++) {
                    std::swap(A[col * K + c2], A[p * K + c2]);
                    std::swap(I[col * K + c2], I[p * K + c2]);
                }
                cplx d = A[col * K + col];
                for (int c2 = 0; c2 < K; c2++) {
                    A[col * K + c2] /= d;
                    I[col * K + c2] /= d;
                }
                for (int r = 0; r < K; r++)
                    if (r != col) {
                        cplx f = A[r * K + col];
                        for (int c2 = 0; c2 < K; c2++) {
                            A[r * K + c2] -= f * A[col * K + c2];
                            I[r * K + c2] -= f * I[col * K + c2];
                        }
                    }
            }
            Vinv = I;
        }
        std::vector<cplx> cx((size_t)N * K * K), cz((size_t)N * K * K), cxx((size_t)N * N * K * K),
            czz((size_t)N * N * K * K);
        check(nqes_correlations(ctx, reinterpret_cast<double *>(W.data()), seed + (uint64_t)niters * C, C, m, B,
                                reinterpret_cast<double *>(cx.data()), reinterpret_cast<double *>(cz.data()),
                                reinterpret_cast<double *>(cxx.data()), reinterpret_cast<double *>(czz.data())));
        // res(idx, k) = Re (V^-1 cor V)_{order[k], order[k]}
        auto project = [&](const std::vector<cplx> &cor, long nidx, std::vector<double> &out) {
            out.assign((size_t)nidx * K, 0.0);
            for (long idx = 0; idx < nidx; idx++)
                for (int k = 0; k < K; k++) {
                    int e = order[k];
                    cplx acc = 0.0;
                    for (int r = 0; r < K; r++)
                        for (int c2 = 0; c2 < K; c2++)
                            acc += Vinv[e * K + r] * cor[(size_t)idx * K * K + r * K + c2] * V[c2 * K + e];
                    out[(size_t)idx * K + k] = acc.real();
                }
        };
        std::vector<double> px, pz, pxx, pzz;
        project(cx, N, px);
        project(cz, N, pz);
        project(cxx, (long)N * N, pxx);
        project(czz, (long)N * N, pzz);
        matv5::Var x{"x", {N, K}, false, 6, {}, {}}, z{"z", {N, K}, false, 6, {}, {}};
        matv5::Var xx{"xx", {N, N, K}, false, 6, {}, {}}, zz{"zz", {N, N, K}, false, 6, {}, {}};
        x.re.resize((size_t)K * N);
        z.re.resize((size_t)K * N);
        xx.re.resize((size_t)K * N * N);
        zz.re.resize((size_t)K * N * N);
        for (int k = 0; k < K; k++) {
            for (int i = 0; i < N; i++) {
                x.re[(size_t)k * N + i] = px[(size_t)i * K + k];
                z.re[(size_t)k * N + i] = pz[(size_t)i * K + k];
            }
            for (int j = 0; j < N; j++)
                for (int i = 0; i < N; i++) {
                    xx.re[(size_t)k * N * N + (size_t)j * N + i] = pxx[((size_t)i * N + j) * K + k];
                    zz.re[(size_t)k * N * N + (size_t)j * N + i] = pzz[((size_t)i * N + j) * K + k];
                }
        }
        matv5::write_file(corr_file, {x, z, xx, zz});

        if (!o.get("dump-rbm").empty()) {
            if (apply_gauge)
                for (int k = 0; k < K; k++)
                    for (int i = 1; i < N; i += 2) W[k * L + i] -= cplx(0, M_PI_2);
            std::ostringstream name;
            name << o.get("dump-rbm") << "N" << N << "M" << M << "K" << K << ".mat";
            dump_rbm(name.str(), W, N, M, K);
        }
        nqes_destroy(ctx);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
