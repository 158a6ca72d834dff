// poly.cpp -- host side of the GMRES-polynomial preconditioner: harmonic Ritz
// values of a small Arnoldi Hessenberg matrix (Francis double-shift QR) and
// their Leja ordering, grouped into the linear / quadratic factors applied
// on the device.
//
// The preconditioner is p(B) with 1 - a p(a) = prod_k (1 - a / theta_k), the
// theta_k being the harmonic Ritz values of B from d Arnoldi steps (the
// residual polynomial of GMRES(d)).  A conjugate pair, or two consecutive
// real roots, forms the quadratic factor 1 - c1 a + c2 a^2 with c1 = 1/t1 +
// 1/t2, c2 = 1/(t1 t2); a lone real root the linear factor 1 - a/t.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace svfsi {

namespace {

// Eigenvalues of an upper Hessenberg matrix h (n x n, row-major, destroyed)
// by the implicit double-shift QR algorithm with deflation on small
// subdiagonals.  Returns false if an eigenvalue does not converge.
bool hessenberg_eigenvalues(int n, std::vector<double>& h, std::vector<std::complex<double>>& ev)
{
    auto H = [&](int i, int j) -> double& { return h[(size_t)i * n + j]; };
    ev.assign(n, {0.0, 0.0});
    double anorm = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = std::max(i - 1, 0); j < n; ++j) anorm += std::fabs(H(i, j));
    int hi = n - 1;
    int its = 0;
    double t = 0.0; // accumulated exceptional shifts
    while (hi >= 0) {
        // find a negligible subdiagonal element
        int l = hi;
        for (; l > 0; --l) {
            const double s = std::fabs(H(l - 1, l - 1)) + std::fabs(H(l, l));
            const double sc = s == 0.0 ? anorm : s;
            if (std::fabs(H(l, l - 1)) <= 1e-15 * sc) {
                H(l, l - 1) = 0.0;
                break;
            }
        }
        const double x = H(hi, hi);
        if (l == hi) { // one root
            ev[hi] = {x + t, 0.0};
            --hi;
            its = 0;
            continue;
        }
        const double y = H(hi - 1, hi - 1);
        const double w = H(hi, hi - 1) * H(hi - 1, hi);
        if (l == hi - 1) { // two roots
            const double p = 0.5 * (y - x);
            const double q = p * p + w;
            const double z = std::sqrt(std::fabs(q));
            const double xs = x + t;
            if (q >= 0.0) {
                const double zz = p + (p >= 0.0 ? z : -z);
                ev[hi - 1] = {xs + zz, 0.0};
                ev[hi] = {zz != 0.0 ? xs - w / zz : xs + zz, 0.0};
            } else {
                ev[hi - 1] = {xs + p, z};
                ev[hi] = {xs + p, -z};
            }
            hi -= 2;
            its = 0;
            continue;
        }
        if (its == 60) return false;
        double xx = x, yy = y, ww = w;
        if (its == 10 || its == 20) { // exceptional shift
            t += xx;
            for (int i = 0; i <= hi; ++i) H(i, i) -= xx;
            const double s = std::fabs(H(hi, hi - 1)) + std::fabs(H(hi - 1, hi - 2));
            xx = yy = 0.75 * s;
            ww = -0.4375 * s * s;
        }
        ++its;
        // look for two consecutive small subdiagonals; form the first column
        // of (H - s1)(H - s2)
        int m = hi - 2;
        double p = 0.0, q = 0.0, r = 0.0;
        for (; m >= l; --m) {
            const double z = H(m, m);
            const double rr = xx - z;
            const double ss = yy - z;
            p = (rr * ss - ww) / H(m + 1, m) + H(m, m + 1);
            q = H(m + 1, m + 1) - z - rr - ss;
            r = H(m + 2, m + 1);
            const double sc = std::fabs(p) + std::fabs(q) + std::fabs(r);
            p /= sc;
            q /= sc;
            r /= sc;
            if (m == l) break;
            const double u = std::fabs(H(m, m - 1)) * (std::fabs(q) + std::fabs(r));
            const double v = std::fabs(p) * (std::fabs(H(m - 1, m - 1)) + std::fabs(z) + std::fabs(H(m + 1, m + 1)));
            if (u <= 1e-15 * v) break;
        }
        for (int i = m + 2; i <= hi; ++i) {
            H(i, i - 2) = 0.0;
            if (i != m + 2) H(i, i - 3) = 0.0;
        }
        // double-shift QR sweep with 3x3 Householder reflections, restricted
        // to the active block l..hi (only eigenvalues are wanted)
        for (int k = m; k <= hi - 1; ++k) {
            double scale = 1.0;
            if (k != m) {
                p = H(k, k - 1);
                q = H(k + 1, k - 1);
                r = (k != hi - 1) ? H(k + 2, k - 1) : 0.0;
                scale = std::fabs(p) + std::fabs(q) + std::fabs(r);
                if (scale == 0.0) continue;
                p /= scale;
                q /= scale;
                r /= scale;
            }
            const double s0 = std::sqrt(p * p + q * q + r * r);
            const double s = p >= 0.0 ? s0 : -s0;
            if (s == 0.0) continue;
            if (k == m) {
                if (l != m) H(k, k - 1) = -H(k, k - 1);
            } else {
                H(k, k - 1) = -s * scale;
            }
            const double pp = p + s;
            const double xk = pp / s, yk = q / s, zk = r / s;
            const double qq = q / pp, rr = r / pp;
            for (int j = k; j <= hi; ++j) { // rows k, k+1, k+2
                double v = H(k, j) + qq * H(k + 1, j);
                if (k != hi - 1) {
                    v += rr * H(k + 2, j);
                    H(k + 2, j) -= v * zk;
                }
                H(k + 1, j) -= v * yk;
                H(k, j) -= v * xk;
            }
            const int imax = std::min(hi, k + 3);
            for (int i = l; i <= imax; ++i) { // columns k, k+1, k+2
                double v = xk * H(i, k) + yk * H(i, k + 1);
                if (k != hi - 1) {
                    v += zk * H(i, k + 2);
                    H(i, k + 2) -= v * rr;
                }
                H(i, k + 1) -= v * qq;
                H(i, k) -= v;
            }
        }
    }
    return true;
}

} // namespace

// Harmonic Ritz values from the unrotated Arnoldi Hessenberg Hu ((d+1) x d,
// column-major with leading dimension ld: Hu[j*ld + i] = h(i, j)), in Leja
// order with conjugate pairs adjacent, turned into factors:
// coef[2*s] = c1, coef[2*s+1] = c2 (quadratic) or c1 = 1/t, c2 = 0 with
// kind[s] = 1 (linear).  Returns the number of factors, or -1 on failure.
int poly_factors(int d, int ld, const double* Hu, std::vector<double>& coef, std::vector<int>& kind)
{
    if (d < 1) return -1;
    // M = H_d + h(d+1,d)^2 H_d^{-T} e_d e_d^T (still upper Hessenberg)
    std::vector<double> Hd((size_t)d * d), M((size_t)d * d);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) Hd[(size_t)i * d + j] = (i <= j + 1) ? Hu[(size_t)j * ld + i] : 0.0;
    const double hd1 = Hu[(size_t)(d - 1) * ld + d];
    // solve H_d^T f = e_d by Gaussian elimination with partial pivoting
    std::vector<double> T((size_t)d * d), f(d, 0.0);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) T[(size_t)i * d + j] = Hd[(size_t)j * d + i];
    f[d - 1] = 1.0;
    for (int c = 0; c < d; ++c) {
        int piv = c;
        for (int r = c + 1; r < d; ++r)
            if (std::fabs(T[(size_t)r * d + c]) > std::fabs(T[(size_t)piv * d + c])) piv = r;
        if (T[(size_t)piv * d + c] == 0.0) return -1;
        if (piv != c) {
            for (int j = 0; j < d; ++j) std::swap(T[(size_t)c * d + j], T[(size_t)piv * d + j]);
            std::swap(f[c], f[piv]);
        }
        for (int r = c + 1; r < d; ++r) {
            const double g = T[(size_t)r * d + c] / T[(size_t)c * d + c];
            for (int j = c; j < d; ++j) T[(size_t)r * d + j] -= g * T[(size_t)c * d + j];
            f[r] -= g * f[c];
        }
    }
    for (int r = d - 1; r >= 0; --r) {
        double v = f[r];
        for (int j = r + 1; j < d; ++j) v -= T[(size_t)r * d + j] * f[j];
        f[r] = v / T[(size_t)r * d + r];
    }
    M = Hd;
    for (int i = 0; i < d; ++i) M[(size_t)i * d + (d - 1)] += hd1 * hd1 * f[i];
    std::vector<std::complex<double>> th;
    if (!hessenberg_eigenvalues(d, M, th)) return -1;
    for (const auto& z : th)
        if (!(std::abs(z) > 0.0) || !std::isfinite(z.real()) || !std::isfinite(z.imag())) return -1;
    // Leja order (largest modulus first, then maximal product of distances),
    // each complex root followed by its conjugate
    std::vector<std::complex<double>> rem = th, ord;
    auto take = [&](size_t k) {
        const std::complex<double> z = rem[k];
        rem.erase(rem.begin() + (long)k);
        ord.push_back(z);
        if (std::fabs(z.imag()) > 1e-12 * std::abs(z) && !rem.empty()) {
            size_t best = 0;
            for (size_t q = 1; q < rem.size(); ++q)
                if (std::abs(rem[q] - std::conj(z)) < std::abs(rem[best] - std::conj(z))) best = q;
            ord.push_back(std::conj(z)); // the exact conjugate keeps the factor real
            rem.erase(rem.begin() + (long)best);
        }
    };
    {
        size_t k = 0;
        for (size_t q = 1; q < rem.size(); ++q)
            if (std::abs(rem[q]) > std::abs(rem[k])) k = q;
        take(k);
    }
    while (!rem.empty()) {
        size_t k = 0;
        double best = -1.0;
        for (size_t q = 0; q < rem.size(); ++q) {
            double lp = 0.0; // log of the product of distances
            for (const auto& o : ord) lp += std::log(std::abs(rem[q] - o) + 1e-300);
            if (lp > best) { best = lp; k = q; }
        }
        take(k);
    }
    coef.clear();
    kind.clear();
    for (size_t k = 0; k < ord.size();) {
        const std::complex<double> z = ord[k];
        const bool cplx = std::fabs(z.imag()) > 1e-12 * std::abs(z);
        if (cplx) {
            const double m = std::norm(z);
            coef.push_back(2.0 * z.real() / m);
            coef.push_back(1.0 / m);
            kind.push_back(2);
            k += 2;
        } else if (k + 1 < ord.size() && std::fabs(ord[k + 1].imag()) <= 1e-12 * std::abs(ord[k + 1])) {
            const double t1 = z.real(), t2 = ord[k + 1].real();
            coef.push_back(1.0 / t1 + 1.0 / t2);
            coef.push_back(1.0 / (t1 * t2));
            kind.push_back(2);
            k += 2;
        } else {
            coef.push_back(1.0 / z.real());
            coef.push_back(0.0);
            kind.push_back(1);
            k += 1;
        }
    }
    return (int)kind.size();
}

} // namespace svfsi
