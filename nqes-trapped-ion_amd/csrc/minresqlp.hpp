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
#include <type_traits>
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

// The operator as two calls: launch(v) starts A v (v stays untouched until the matching
// finish), finish(out) writes it (false aborts).  The solver launches A v_{k+1} as soon
// as v_{k+1} = r3 / betan and the stopping decision are known, and updates w, x while it
// runs: every element and every reduction is computed as in the one-call order, so the
// results are the same bit for bit.
template <typename AOp>
Result solve_async(AOp &A, const std::vector<cd> &b, std::vector<cd> &x, double rtol, uint64_t maxit, bool &aborted)
{
    using namespace nq64;
    const double maxxnorm = 1e7, Acondlim = 1e15, TranCond = 1e7;
    const size_t n = b.size();
    // r1, r2, r3 and w, wl, wl2 rotate through three buffers each (the reference copies
    // r1 = r2, r2 = r3, wl2 = wl, wl = w); v and the next v alternate between two
    std::vector<cd> rb[3] = {b, b, b}, wb[3] = {std::vector<cd>(n, zero()), std::vector<cd>(n, zero()),
                                                std::vector<cd>(n, zero())};
    std::vector<cd> vb[2] = {std::vector<cd>(n), std::vector<cd>(n)}, xl2(n, zero());
    int ir1 = 0, ir2 = 1, ir3 = 2;
    cd *r1 = rb[ir1].data(), *r2 = rb[ir2].data(), *r3 = rb[ir3].data();
    cd *w = wb[0].data(), *wl = wb[1].data(), *wl2 = wb[2].data();
    int vcur = 0;
    aborted = false;
    double beta1 = 0.0;
    for (size_t i = 0; i < n; i++) beta1 += b[i].re * b[i].re + b[i].im * b[i].im;
    beta1 = std::sqrt(beta1);

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

    bool launched = true;
    if (flag == flag0 && iters < maxit) {  // v_1 = r3 / beta1
        cd *v = vb[vcur].data();
        for (size_t i = 0; i < n; i++) v[i] = divr(r3[i], betan);
        launched = A.launch(vb[vcur]);
    }
    while (flag == flag0 && iters < maxit) {
        iters++;
        double betal = beta;
        beta = betan;
        cd *v = vb[vcur].data();
        if (!launched || !A.finish(rb[ir3])) {
            aborted = true;
            break;
        }
        // r3 -= (beta / betal) r1 with alfa = r3^H v, then r3 -= (alfa / beta) r2 with
        // betan = |r3| (each reduction in index order, as a separate pass would)
        double alfa = 0.0;
        if (iters > 1) {
            const double f = beta / betal;
            for (size_t i = 0; i < n; i++) {
                r3[i] = sub(r3[i], scale(f, r1[i]));
                alfa += r3[i].re * v[i].re + r3[i].im * v[i].im;
            }
        } else {
            for (size_t i = 0; i < n; i++) alfa += r3[i].re * v[i].re + r3[i].im * v[i].im;
        }
        {
            const double f = alfa / beta;
            double s2 = 0.0;
            for (size_t i = 0; i < n; i++) {
                r3[i] = sub(r3[i], scale(f, r2[i]));
                s2 += r3[i].re * r3[i].re + r3[i].im * r3[i].im;
            }
            betan = std::sqrt(s2);
        }
        {
            const int t = ir1;
            ir1 = ir2;
            ir2 = ir3;
            ir3 = t;
            r1 = rb[ir1].data();
            r2 = rb[ir2].data();
            r3 = rb[ir3].data();
        }
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

        // which update of w and x this iteration makes, and the scalars it reads (the
        // tail below moves some of them on)
        const bool minres_step = Acond < TranCond && flag != flag0 && QLPiter == 0;
        bool x_step = false, qlp_first = false;
        if (minres_step) {
            x_step = xnorm < maxxnorm;
            if (!x_step) flag = NormLimitExceeded;
        } else {
            QLPiter++;
            qlp_first = QLPiter == 1 && iters > 1;
        }
        const uint64_t it = iters;
        const double s_epln = epln, s_dlta = dlta_QLP, s_gama = gama_tmp, s_tau = tau;
        const double s_gamal3 = gamal3, s_veplnl2 = veplnl2, s_etal = etal, s_gamalQ = gamal_QLP,
                     s_veplnQ = vepln_QLP, s_gamaQ = gama_QLP, s_ulQ = ul_QLP, s_uQ = u_QLP;
        const double s_cr1 = cr1, s_sr1 = sr1, s_cr2 = cr2, s_sr2 = sr2, s_ul2 = ul2, s_ul = ul, s_u = u;

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

        // the next operator call starts here: v_{k+1} = r3 / betan (r3 is now r2)
        if (flag == flag0 && iters < maxit) {
            cd *vn = vb[1 - vcur].data();
            for (size_t i = 0; i < n; i++) vn[i] = divr(r2[i], betan);
            launched = A.launch(vb[1 - vcur]);
        }

        if (minres_step) {
            cd *t = wl2;
            wl2 = wl;
            wl = w;
            w = t;
            for (size_t i = 0; i < n; i++) w[i] = divr(sub(sub(v[i], scale(s_epln, wl2[i])), scale(s_dlta, wl[i])), s_gama);
            if (x_step)
                for (size_t i = 0; i < n; i++) x[i] = add(x[i], scale(s_tau, w[i]));
        } else {
            if (qlp_first) {
                if (it > 3)
                    for (size_t i = 0; i < n; i++)
                        wl2[i] = add(add(scale(s_gamal3, wl2[i]), scale(s_veplnl2, wl[i])), scale(s_etal, w[i]));
                if (it > 2)
                    for (size_t i = 0; i < n; i++) wl[i] = add(scale(s_gamalQ, wl[i]), scale(s_veplnQ, w[i]));
                for (size_t i = 0; i < n; i++) w[i] = scale(s_gamaQ, w[i]);
                for (size_t i = 0; i < n; i++) xl2[i] = sub(sub(x[i], scale(s_ulQ, wl[i])), scale(s_uQ, w[i]));
            }
            if (it == 1) {
                cd *t = wl2;
                wl2 = wl;
                wl = t;
                for (size_t i = 0; i < n; i++) wl[i] = scale(s_sr1, v[i]);
                for (size_t i = 0; i < n; i++) w[i] = scale(s_cr1, neg(v[i]));
            } else if (it == 2) {
                cd *t = wl2;
                wl2 = wl;
                wl = t;
                for (size_t i = 0; i < n; i++) {
                    cd wi = w[i], vi = v[i];
                    wl[i] = add(scale(s_cr1, wi), scale(s_sr1, vi));
                    w[i] = sub(scale(s_sr1, wi), scale(s_cr1, vi));
                }
            } else {
                for (size_t i = 0; i < n; i++) {
                    cd o_wl = wl[i], o_w = w[i], vi = v[i];
                    cd nw = sub(scale(s_sr2, o_wl), scale(s_cr2, vi));
                    cd nwl2 = add(scale(s_cr2, o_wl), scale(s_sr2, vi));
                    cd nv = add(scale(s_cr1, o_w), scale(s_sr1, nw));
                    cd nw2 = sub(scale(s_sr1, o_w), scale(s_cr1, nw));
                    wl2[i] = nwl2;
                    w[i] = nw2;
                    wl[i] = nv;
                }
            }
            for (size_t i = 0; i < n; i++) xl2[i] = add(xl2[i], scale(s_ul2, wl2[i]));
            for (size_t i = 0; i < n; i++) x[i] = add(add(xl2[i], scale(s_ul, wl[i])), scale(s_u, w[i]));
        }
        vcur = 1 - vcur;
    }
    Result res;
    res.flag = flag;
    res.iters = iters;
    res.qlp_iters = QLPiter;
    return res;
}

// Op: callable (const std::vector<cd>& v, std::vector<cd>& out) -> bool (false aborts),
// run as launch (remember v) and finish (call it)
template <typename Op>
struct SyncOp {
    Op &A;
    const std::vector<cd> *v = nullptr;
    bool launch(const std::vector<cd> &vv)
    {
        v = &vv;
        return true;
    }
    bool finish(std::vector<cd> &out) { return A(*v, out); }
};

template <typename Op>
Result solve(Op &&A, const std::vector<cd> &b, std::vector<cd> &x, double rtol, uint64_t maxit, bool &aborted)
{
    SyncOp<typename std::remove_reference<Op>::type> op{A};
    return solve_async(op, b, x, rtol, maxit, aborted);
}

} // namespace nqmq
