// eig.cpp -- see eig.hpp.
#include "eig.hpp"

#include <cmath>
#include <vector>

namespace eigk {

namespace {

using Mat = std::vector<cplx>;

inline cplx &at(Mat &m, int n, int r, int c) { return m[(size_t)r * n + c]; }

// Householder reduction to upper Hessenberg form, in place
void hessenberg(int n, Mat &H)
{
    for (int k = 0; k < n - 2; k++) {
        double alpha = 0.0;
        for (int i = k + 1; i < n; i++) alpha += std::norm(at(H, n, i, k));
        alpha = std::sqrt(alpha);
        if (alpha == 0.0) continue;
        cplx x0 = at(H, n, k + 1, k);
        cplx ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cplx(1.0, 0.0);
        std::vector<cplx> v(n, 0.0);
        v[k + 1] = x0 + ph * alpha;
        for (int i = k + 2; i < n; i++) v[i] = at(H, n, i, k);
        double vn = 0.0;
        for (int i = k + 1; i < n; i++) vn += std::norm(v[i]);
        if (vn == 0.0) continue;
        // H <- (I - 2 v v^H / vn) H (I - 2 v v^H / vn)
        for (int c = 0; c < n; c++) {
            cplx s = 0.0;
            for (int i = k + 1; i < n; i++) s += std::conj(v[i]) * at(H, n, i, c);
            s *= 2.0 / vn;
            for (int i = k + 1; i < n; i++) at(H, n, i, c) -= v[i] * s;
        }
        for (int r = 0; r < n; r++) {
            cplx s = 0.0;
            for (int i = k + 1; i < n; i++) s += at(H, n, r, i) * v[i];
            s *= 2.0 / vn;
            for (int i = k + 1; i < n; i++) at(H, n, r, i) -= s * std::conj(v[i]);
        }
        for (int i = k + 2; i < n; i++) at(H, n, i, k) = 0.0;
    }
}

// eigenvalue of [[a, b], [c, d]] closest to d
cplx wilkinson(cplx a, cplx b, cplx c, cplx d)
{
    cplx tr = a + d, det = a * d - b * c;
    cplx disc = std::sqrt(tr * tr * 0.25 - det);
    cplx l1 = tr * 0.5 + disc, l2 = tr * 0.5 - disc;
    return std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
}

// solves (A - mu I) x = b by LU with partial pivoting; b overwritten by x
void shifted_solve(int n, const cplx *A, cplx mu, std::vector<cplx> &b)
{
    Mat M(A, A + (size_t)n * n);
    for (int i = 0; i < n; i++) at(M, n, i, i) -= mu;
    double scale = 0.0;
    for (auto &z : M) scale = std::fmax(scale, std::abs(z));
    const double tiny = (scale > 0 ? scale : 1.0) * 1e-14;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (std::abs(at(M, n, i, k)) > std::abs(at(M, n, p, k))) p = i;
        if (p != k) {
            for (int c = 0; c < n; c++) std::swap(at(M, n, k, c), at(M, n, p, c));
            std::swap(b[k], b[p]);
        }
        if (std::abs(at(M, n, k, k)) < tiny) at(M, n, k, k) = tiny;
        for (int i = k + 1; i < n; i++) {
            cplx f = at(M, n, i, k) / at(M, n, k, k);
            for (int c = k; c < n; c++) at(M, n, i, c) -= f * at(M, n, k, c);
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; k--) {
        cplx s = b[k];
        for (int c = k + 1; c < n; c++) s -= at(M, n, k, c) * b[c];
        b[k] = s / at(M, n, k, k);
    }
}

} // namespace

bool eig(int n, const cplx *A, cplx *evals, cplx *V)
{
    if (n <= 0 || n > 16) return false;
    Mat H(A, A + (size_t)n * n);
    hessenberg(n, H);
    const double eps = 2.220446049250313e-16;
    int hi = n - 1, iters = 0;
    bool ok = true;
    while (hi > 0) {
        double sub = std::abs(at(H, n, hi, hi - 1));
        double diag = std::abs(at(H, n, hi, hi)) + std::abs(at(H, n, hi - 1, hi - 1));
        if (sub <= eps * (diag > 0 ? diag : 1.0)) {
            at(H, n, hi, hi - 1) = 0.0;
            hi--;
            iters = 0;
            continue;
        }
        if (++iters > 100 * n) {
            ok = false;
            break;
        }
        // active window [lo, hi]: lowest lo with no negligible subdiagonal below it
        int lo = hi - 1;
        while (lo > 0) {
            double s = std::abs(at(H, n, lo, lo - 1));
            double d = std::abs(at(H, n, lo, lo)) + std::abs(at(H, n, lo - 1, lo - 1));
            if (s <= eps * (d > 0 ? d : 1.0)) break;
            lo--;
        }
        cplx mu = wilkinson(at(H, n, hi - 1, hi - 1), at(H, n, hi - 1, hi), at(H, n, hi, hi - 1), at(H, n, hi, hi));
        if (iters % 11 == 10) mu += cplx(sub, 0.0);  // exceptional shift
        for (int i = lo; i <= hi; i++) at(H, n, i, i) -= mu;
        std::vector<cplx> cs(n), sn(n);
        for (int k = lo; k < hi; k++) {
            cplx a = at(H, n, k, k), b = at(H, n, k + 1, k);
            double r = std::sqrt(std::norm(a) + std::norm(b));
            cplx c = r > 0 ? a / r : cplx(1.0, 0.0), s = r > 0 ? b / r : cplx(0.0, 0.0);
            cs[k] = c;
            sn[k] = s;
            for (int j = k; j < n; j++) {
                cplx x = at(H, n, k, j), y = at(H, n, k + 1, j);
                at(H, n, k, j) = std::conj(c) * x + std::conj(s) * y;
                at(H, n, k + 1, j) = -s * x + c * y;
            }
        }
        for (int k = lo; k < hi; k++) {
            cplx c = cs[k], s = sn[k];
            for (int i = 0; i <= std::min(k + 2, hi); i++) {
                cplx x = at(H, n, i, k), y = at(H, n, i, k + 1);
                at(H, n, i, k) = x * c + y * s;
                at(H, n, i, k + 1) = -x * std::conj(s) + y * std::conj(c);
            }
        }
        for (int i = lo; i <= hi; i++) at(H, n, i, i) += mu;
    }
    for (int i = 0; i < n; i++) evals[i] = at(H, n, i, i);
    if (V) {
        double anorm = 0.0;
        for (int i = 0; i < n * n; i++) anorm = std::fmax(anorm, std::abs(A[i]));
        for (int e = 0; e < n; e++) {
            std::vector<cplx> x(n);
            for (int i = 0; i < n; i++) x[i] = cplx(1.0 + 0.1 * i + 0.01 * e, 0.05 * (i - e));
            cplx mu = evals[e] + cplx(anorm * 1e-10, 0.0);
            for (int it = 0; it < 3; it++) {
                shifted_solve(n, A, mu, x);
                double nr = 0.0;
                for (auto &z : x) nr += std::norm(z);
                nr = std::sqrt(nr);
                for (auto &z : x) z /= nr;
            }
            for (int i = 0; i < n; i++) V[(size_t)i * n + e] = x[i];
        }
    }
    return ok;
}

int pinv(int n, const cplx *A, cplx *X, double rtol)
{
    // B = A (columns b_p), V = I; rotate column pairs until mutually orthogonal
    std::vector<cplx> B(A, A + (size_t)n * n), V((size_t)n * n, cplx(0.0, 0.0));
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    auto col = [&](std::vector<cplx> &Mx, int r, int c) -> cplx & { return Mx[(size_t)r * n + c]; };
    for (int sweep = 0; sweep < 60; sweep++) {
        bool rotated = false;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double alpha = 0.0, beta = 0.0;
                cplx gamma(0.0, 0.0);
                for (int r = 0; r < n; r++) {
                    alpha += std::norm(col(B, r, p));
                    beta += std::norm(col(B, r, q));
                    gamma += std::conj(col(B, r, p)) * col(B, r, q);
                }
                const double ag = std::abs(gamma);
                if (ag == 0.0 || ag <= 1e-15 * std::sqrt(alpha * beta)) continue;
                rotated = true;
                const cplx ph = gamma / ag;  // b_q conj(ph) has a real overlap with b_p
                const double zeta = (beta - alpha) / (2.0 * ag);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < n; r++) {
                    const cplx bp = col(B, r, p), bq = col(B, r, q) * std::conj(ph);
                    col(B, r, p) = c * bp - s * bq;
                    col(B, r, q) = s * bp + c * bq;
                    const cplx vp = col(V, r, p), vq = col(V, r, q) * std::conj(ph);
                    col(V, r, p) = c * vp - s * vq;
                    col(V, r, q) = s * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    // A V = B = U S: s_p = |b_p|, u_p = b_p / s_p; pinv(A) = V S^+ U^H = sum_p v_p b_p^H / s_p^2
    std::vector<double> sv(n);
    double smax = 0.0;
    for (int p = 0; p < n; p++) {
        double a2 = 0.0;
        for (int r = 0; r < n; r++) a2 += std::norm(col(B, r, p));
        sv[p] = std::sqrt(a2);
        smax = std::max(smax, sv[p]);
    }
    for (int i = 0; i < n * n; i++) X[i] = 0.0;
    int rank = 0;
    for (int p = 0; p < n; p++) {
        if (!(sv[p] > rtol * smax)) continue;
        rank++;
        const double inv2 = 1.0 / (sv[p] * sv[p]);
        for (int r = 0; r < n; r++)
            for (int c = 0; c < n; c++) X[(size_t)r * n + c] += col(V, r, p) * std::conj(col(B, c, p)) * inv2;
    }
    return rank;
}

cplx det(int n, const cplx *A)
{
    std::vector<cplx> L(A, A + (size_t)n * n);
    cplx d(1.0, 0.0);
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int r = k + 1; r < n; r++)
            if (std::abs(L[(size_t)r * n + k]) > std::abs(L[(size_t)piv * n + k])) piv = r;
        if (L[(size_t)piv * n + k] == cplx(0.0, 0.0)) return cplx(0.0, 0.0);
        if (piv != k) {
            for (int c = 0; c < n; c++) std::swap(L[(size_t)k * n + c], L[(size_t)piv * n + c]);
            d = -d;
        }
        d *= L[(size_t)k * n + k];
        for (int r = k + 1; r < n; r++) {
            const cplx f = L[(size_t)r * n + k] / L[(size_t)k * n + k];
            for (int c = k; c < n; c++) L[(size_t)r * n + c] -= f * L[(size_t)k * n + c];
        }
    }
    return d;
}

} // namespace eigk
