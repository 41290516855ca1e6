// srloop.hpp -- the outer SR loop's rules, shared by the rbm-matfree executable and
// the solve server (host only).
//
// run() of the reference (rbm.cc:1173-1197):
//   - SR step t draws chain c from pcg64(seed + t*C + c) (rbm.cc:1179);
//   - the step is taken with max(lrate, lrate_cutoff) and max(reglam, reglam_cutoff),
//     after which lrate *= lrate_decay and reglam *= reglam_decay (rbm.cc:1180-1185);
//   - the energies are the ascending real parts of the eigenvalues of e_mat / N
//     (rbm.cc:1183-1187);
//   - the run stops after step t > 0 when |evals - best| < tol, where best is the
//     lowest-sum energy vector seen so far, first set at t = 0 (rbm.cc:1188-1196).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "eig.hpp"

namespace srloop {

struct Schedule {
    double lrate, lrate_decay, lrate_cutoff;
    double reglam, reglam_decay, reglam_cutoff;

    double step_lrate() const { return std::max(lrate, lrate_cutoff); }
    double step_reglam() const { return std::max(reglam, reglam_cutoff); }
    void advance()
    {
        reglam *= reglam_decay;
        lrate *= lrate_decay;
    }
};

// ascending real parts of the eigenvalues of the K x K row-major matrix A
inline std::vector<double> sorted_real_eigs(const std::vector<std::complex<double>> &A, int K)
{
    std::vector<std::complex<double>> ev(K);
    eigk::eig(K, A.data(), ev.data(), nullptr);
    std::vector<double> r(K);
    for (int k = 0; k < K; k++) r[k] = ev[k].real();
    std::sort(r.begin(), r.end());
    return r;
}

class Convergence {
public:
    explicit Convergence(int K) : best_(K, 0.0) {}
    // feeds the energies of step `iter`; true when the run should stop
    bool update(int iter, const std::vector<double> &evals, double tol)
    {
        double dn = 0.0, se = 0.0, sb = 0.0;
        for (size_t k = 0; k < best_.size(); k++) {
            dn += (evals[k] - best_[k]) * (evals[k] - best_[k]);
            se += evals[k];
            sb += best_[k];
        }
        if (iter && std::sqrt(dn) < tol) return true;
        if (!iter || se < sb) best_ = evals;
        return false;
    }

private:
    std::vector<double> best_;
};

} // namespace srloop
