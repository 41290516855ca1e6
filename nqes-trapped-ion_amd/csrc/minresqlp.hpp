// minresqlp.hpp -- MINRES-QLP (Choi, Paige & Saunders, SIAM J. Sci. Comput. 33,
// 2011) for Hermitian, possibly singular systems A x = b with a matrix-free
// operator; complex vectors, real scalar recurrences in FP64.
//
// Follows the reference's minresqlp() (code/minresqlp.h:85-390) step for step:
// Lanczos (:158-181), left rotation Q_k (:183-199), right rotations P_{k-2,k} and
// P_{k-1,k} (:200-215), the xnorm recurrence (:217-237), the MINRES/QLP update of
// w and x (:238-300; the MINRES branch is taken only when the flag changed in the
// same iteration, exactly as the reference's condition reads), the next right
// rotation (:302-311), norm estimates (:312-331) and the stopping tests
// (:332-373).  shift = 0, maxxnorm = 1e7, Acondlim = 1e15, TranCond = 1e7, as at
// the reference's call sites (rbm.cc:661, 686).
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "cplx64.hpp"

namespace nqmq {

using nq64::cd;

enum Flag {
    InternalError = -2,
    EigenVector = -1,
    Zero = 0,
    Converged = 1,
    ConvergedLeastSquare = 2,
    MachinePrecisionReached = 3,
    MachinePrecisionReachedLeastSquare = 4,
    ConvergedToEigenVector = 5,
    NormLimitExceeded = 6,
    ConditionLimitExceeded = 7,
    MaxIterationsReached = 8,
    ExactlySingular = 9
};

struct Result {
    int flag;
    uint64_t iters;
    uint64_t qlp_iters;
};

inline double sgn(double a) { return std::copysign(1.0, a); }

// plane rotation [c s; s -c] [a; b] = [r; 0] (minresqlp.h:85-110)
inline void sym_ortho(double a, double b, double &c, double &s, double &r)
{
    if (b == 0.0) {
        c = (a == 0.0) ? 1.0 : sgn(a);
        s = 0.0;
        r = std::fabs(a);
    } else if (a == 0.0) {
        c = 0.0;
        s = sgn(b);
        r = std::fabs(b);
    } else if (std::fabs(b) > std::fabs(a)) {
        double t = a / b;
        s = sgn(b) / std::sqrt(1.0 + t * t);
        c = s * t;
        r = b / s;
    } else {
        double t = b / a;
        c = sgn(a) / std::sqrt(1.0 + t * t);
        s = c * t;
        r = a / c;
    }
}

inline double vnorm(const std::vector<cd> &x)
{
    double s = 0.0;
    for (const cd &z : x) s += z.re * z.re + z.im * z.im;
    return std::sqrt(s);
}

inline double vdot_re(const std::vector<cd> &x, const std::vector<cd> &y)
{
    double s = 0.0;
    for (size_t i = 0; i < x.size(); i++) s += x[i].re * y[i].re + x[i].im * y[i].im;
    return s;
}

// Op: callable (const std::vector<cd>& v, std::vector<cd>& out) -> bool (false aborts)
template <typename Op>
Result solve(Op &&A, const std::vector<cd> &b, std::vector<cd> &x, double rtol, uint64_t maxit, bool &aborted)
{
    using namespace nq64;
    const double maxxnorm = 1e7, Acondlim = 1e15, TranCond = 1e7;
    const size_t n = b.size();
    std::vector<cd> r1(b), r2(b), r3(b), v(n), w(n, zero()), wl(n, zero()), wl2(n, zero()), xl2(n, zero());
    aborted = false;
    double beta1 = vnorm(r2);

    const int flag0 = InternalError;
    int flag = flag0;
    uint64_t iters = 0, QLPiter = 0;
    double beta = 0, tau = 0, taul = 0, phi = beta1, betan = beta1, gmin = 0, gminl = 0;
    double cs = -1, sn = 0, cr1 = -1, sr1 = 0, cr2 = -1, sr2 = 0, dltan = 0, eplnn = 0;
    double gama = 0, gamal = 0, gamal2 = 0, eta = 0, etal = 0, etal2 = 0;
    double vepln = 0, veplnl = 0, veplnl2 = 0, ul3 = 0, ul2 = 0, ul = 0, u = 0;
    double gama_QLP = 0, gamal_QLP = 0, vepln_QLP = 0, u_QLP = 0, ul_QLP = 0;
    double rnorm = betan, xnorm = 0, xl2norm = 0, Axnorm = 0, Anorm = 0, Acond = 1;
    double relres = rnorm / (beta1 + 1e-50);
    x.assign(n, zero());
    if (beta1 == 0.0) flag = Zero;

    while (flag == flag0 && iters < maxit) {
        iters++;
        double betal = beta;
        beta = betan;
        for (size_t i = 0; i < n; i++) v[i] = divr(r3[i], beta);
        if (!A(v, r3)) {
            aborted = true;
            break;
        }
        if (iters > 1) {
            double f = beta / betal;
            for (size_t i = 0; i < n; i++) r3[i] = sub(r3[i], scale(f, r1[i]));
        }
        double alfa = vdot_re(r3, v);
        {
            double f = alfa / beta;
            for (size_t i = 0; i < n; i++) r3[i] = sub(r3[i], scale(f, r2[i]));
        }
        r1 = r2;
        r2 = r3;
        betan = vnorm(r3);
        if (iters == 1 && betan == 0.0) {
            if (alfa == 0.0) {
                flag = Zero;
            } else {
                flag = EigenVector;
                for (size_t i = 0; i < n; i++) x[i] = divr(b[i], alfa);
            }
            break;
        }
        double pnorm = std::sqrt(betal * betal + alfa * alfa + betan * betan);

        double dbar = dltan;
        double dlta = cs * dbar + sn * alfa;
        double epln = eplnn;
        double gbar = sn * dbar - cs * alfa;
        eplnn = sn * betan;
        dltan = -cs * betan;
        double dlta_QLP = dlta;
        double gamal3 = gamal2;
        gamal2 = gamal;
        gamal = gama;
        sym_ortho(gbar, betan, cs, sn, gama);
        double gama_tmp = gama;
        double taul2 = taul;
        taul = tau;
        tau = cs * phi;
        Axnorm = std::sqrt(Axnorm * Axnorm + tau * tau);
        phi = sn * phi;
        if (iters > 2) {
            veplnl2 = veplnl;
            etal2 = etal;
            etal = eta;
            double dlta_tmp = sr2 * vepln - cr2 * dlta;
            veplnl = cr2 * vepln + sr2 * dlta;
            dlta = dlta_tmp;
            eta = sr2 * gama;
            gama = -cr2 * gama;
        }
        if (iters > 1) {
            sym_ortho(gamal, dlta, cr1, sr1, gamal);
            vepln = sr1 * gama;
            gama = -cr1 * gama;
        }

        double ul4 = ul3;
        ul3 = ul2;
        if (iters > 2) ul2 = (taul2 - etal2 * ul4 - veplnl2 * ul3) / gamal2;
        if (iters > 1) ul = (taul - etal * ul3 - veplnl * ul2) / gamal;
        double xnorm_tmp = std::sqrt(xl2norm * xl2norm + ul2 * ul2 + ul * ul);
        if (std::fabs(gama) > DBL_MIN && xnorm_tmp < maxxnorm) {
            u = (tau - eta * ul2 - vepln * ul) / gama;
            if (std::sqrt(xnorm_tmp * xnorm_tmp + u * u) > maxxnorm) {
                u = 0;
                flag = NormLimitExceeded;
            }
        } else {
            u = 0;
            flag = ExactlySingular;
        }
        xl2norm = std::sqrt(xl2norm * xl2norm + ul2 * ul2);
        xnorm = std::sqrt(xl2norm * xl2norm + ul * ul + u * u);

        if (Acond < TranCond && flag != flag0 && QLPiter == 0) {
            wl2 = wl;
            wl = w;
            for (size_t i = 0; i < n; i++)
                w[i] = divr(sub(sub(v[i], scale(epln, wl2[i])), scale(dlta_QLP, wl[i])), gama_tmp);
            if (xnorm < maxxnorm) {
                for (size_t i = 0; i < n; i++) x[i] = add(x[i], scale(tau, w[i]));
            } else {
                flag = NormLimitExceeded;
            }
        } else {
            QLPiter++;
            if (QLPiter == 1 && iters > 1) {
                if (iters > 3)
                    for (size_t i = 0; i < n; i++)
                        wl2[i] = add(add(scale(gamal3, wl2[i]), scale(veplnl2, wl[i])), scale(etal, w[i]));
                if (iters > 2)
                    for (size_t i = 0; i < n; i++) wl[i] = add(scale(gamal_QLP, wl[i]), scale(vepln_QLP, w[i]));
                for (size_t i = 0; i < n; i++) w[i] = scale(gama_QLP, w[i]);
                for (size_t i = 0; i < n; i++) xl2[i] = sub(sub(x[i], scale(ul_QLP, wl[i])), scale(u_QLP, w[i]));
            }
            if (iters == 1) {
                wl2 = wl;
                for (size_t i = 0; i < n; i++) wl[i] = scale(sr1, v[i]);
                for (size_t i = 0; i < n; i++) w[i] = scale(cr1, neg(v[i]));
            } else if (iters == 2) {
                wl2 = wl;
                for (size_t i = 0; i < n; i++) {
                    cd wi = w[i], vi = v[i];
                    wl[i] = add(scale(cr1, wi), scale(sr1, vi));
                    w[i] = sub(scale(sr1, wi), scale(cr1, vi));
                }
            } else {
                for (size_t i = 0; i < n; i++) {
                    cd o_wl = wl[i], o_w = w[i], vi = v[i];
                    cd nw = sub(scale(sr2, o_wl), scale(cr2, vi));
                    cd nwl2 = add(scale(cr2, o_wl), scale(sr2, vi));
                    cd nv = add(scale(cr1, o_w), scale(sr1, nw));
                    cd nw2 = sub(scale(sr1, o_w), scale(cr1, nw));
                    wl2[i] = nwl2;
                    v[i] = nv;
                    w[i] = nw2;
                    wl[i] = nv;
                }
            }
            for (size_t i = 0; i < n; i++) xl2[i] = add(xl2[i], scale(ul2, wl2[i]));
            for (size_t i = 0; i < n; i++) x[i] = add(add(xl2[i], scale(ul, wl[i])), scale(u, w[i]));
        }

        double gamal_tmp = gamal;
        sym_ortho(gamal, eplnn, cr2, sr2, gamal);
        gamal_QLP = gamal_tmp;
        vepln_QLP = vepln;
        gama_QLP = gama;
        ul_QLP = ul;
        u_QLP = u;
        double abs_gama = std::fabs(gama);
        Anorm = std::fmax(std::fmax(Anorm, pnorm), std::fmax(gamal, abs_gama));
        if (iters == 1) {
            gmin = gama;
            gminl = gmin;
        } else {
            double gminl2 = gminl;
            gminl = gmin;
            gmin = std::fmin(std::fmin(gminl2, gamal), abs_gama);
        }
        double Acondl = Acond;
        Acond = Anorm / gmin;
        double rnorml = rnorm;
        double relresl = relres;
        if (flag != ExactlySingular) rnorm = phi;
        relres = rnorm / (Anorm * xnorm + beta1);
        double rootl = std::sqrt(gbar * gbar + dltan * dltan);
        double relAresl = rootl / Anorm;
        double epsx = Anorm * xnorm * DBL_EPSILON;
        if (flag == flag0 || flag == ExactlySingular) {
            double t1 = 1 + relres, t2 = 1 + relAresl;
            if (iters >= maxit) flag = MaxIterationsReached;
            if (Acond >= Acondlim) flag = ConditionLimitExceeded;
            if (xnorm >= maxxnorm) flag = NormLimitExceeded;
            if (epsx >= beta1) flag = ConvergedToEigenVector;
            if (t2 <= 1) flag = MachinePrecisionReachedLeastSquare;
            if (t1 <= 1) flag = MachinePrecisionReached;
            if (relAresl <= rtol) flag = ConvergedLeastSquare;
            if (relres <= rtol) flag = Converged;
        }
        if (flag == ConvergedLeastSquare || flag == MachinePrecisionReachedLeastSquare ||
            flag == NormLimitExceeded || flag == ConditionLimitExceeded) {
            iters--;
            Acond = Acondl;
            rnorm = rnorml;
            relres = relresl;
        }
    }
    Result res;
    res.flag = flag;
    res.iters = iters;
    res.qlp_iters = QLPiter;
    return res;
}

} // namespace nqmq
