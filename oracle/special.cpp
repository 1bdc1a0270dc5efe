// Oracle special functions and the 1-D lattice Green's function.
//
// lgf1d(order, n, c) is the Green's function of the 1-D finite-difference
// stencil of the given order shifted by c >= 0 (the symbol of the two spectral
// directions), as used by the 1-unbounded LGF kernels of the reference
// (src/lgf_1unb2spe.hpp:161-611, src/green_kernels.hpp:579-623):
//     g(n) = 1/(2 pi) int_{-pi}^{pi} cos(n t) / (c + s(t)) dt ,   s = stencil symbol.
// The reference evaluates closed forms from the roots of the stencil
// polynomial; the oracle instead evaluates the integral with the periodic
// trapezoidal rule (exponentially convergent for c > 0, resolution chosen from
// the pole distance), and for c = 0 builds the regularised solution
// -|n|/2 + sum_i K_i (r_i^|n| - 1) from the roots r_i inside the unit circle,
// with K_i fixed by the stencil equation near n = 0. Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

double bessel_k0(double x) { return std::cyl_bessel_k(0.0, x); }
double bessel_k1(double x) { return std::cyl_bessel_k(1.0, x); }
double expint_e1(double x) { return -std::expint(-x); }

double lgf_symbol(int order, double k) {
    const double s1 = std::sin(0.5 * k), s2 = std::sin(k), s3 = std::sin(1.5 * k), s4 = std::sin(2.0 * k);
    switch (order) {
        case 2: return 4.0 * s1 * s1;
        case 4: return 4.0 * (4.0 / 3.0 * s1 * s1 - 1.0 / 12.0 * s2 * s2);
        case 6: return 4.0 * (1.5 * s1 * s1 - 0.15 * s2 * s2 + 1.0 / 90.0 * s3 * s3);
        case 8: return 4.0 * (1.6 * s1 * s1 - 0.2 * s2 * s2 + 8.0 / 315.0 * s3 * s3 - 1.0 / 560.0 * s4 * s4);
        default: throw std::runtime_error("unknown LGF order");
    }
}

namespace {

using cl = std::complex<long double>;

// stencil coefficients s_m (m = 0..w) of the symbol: s(t) = s_0 + 2 sum_m s_m cos(m t)
std::vector<long double> stencil(int order) {
    const int w = order / 2, M = 64;
    std::vector<long double> s(w + 1, 0.0L);
    for (int m = 0; m <= w; ++m) {
        long double acc = 0;
        for (int j = 0; j < M; ++j) {
            const long double t = 2.0L * 3.14159265358979323846264338327950288L * j / M;
            acc += (long double)lgf_symbol(order, (double)t) * cosl(m * t);
        }
        s[m] = acc / M;
    }
    return s;
}

// roots of p (coefficients low->high) by Durand-Kerner
std::vector<cl> poly_roots(const std::vector<long double>& p) {
    const int deg = (int)p.size() - 1;
    std::vector<cl> z(deg);
    for (int i = 0; i < deg; ++i) z[i] = std::pow(cl(0.4L, 0.9L), i);
    auto eval = [&](cl x) { cl a = p[deg]; for (int i = deg - 1; i >= 0; --i) a = a * x + p[i]; return a; };
    for (int it = 0; it < 2000; ++it) {
        long double delta = 0;
        for (int i = 0; i < deg; ++i) {
            cl den = p[deg];
            for (int j = 0; j < deg; ++j) if (j != i) den *= (z[i] - z[j]);
            const cl dz = eval(z[i]) / den;
            z[i] -= dz;
            delta = std::max(delta, std::abs(dz));
        }
        if (delta < 1e-30L) break;
    }
    return z;
}

// regularised c = 0 solution, coefficients (b = slope, K_i, r_i)
struct ZeroSol { long double b; std::vector<cl> K, r; };

ZeroSol zero_solution(int order) {
    const int w = order / 2;
    const auto s = stencil(order);
    // P(z) = sum_{m=-w}^{w} s_|m| z^{m+w}
    std::vector<long double> P(2 * w + 1);
    for (int m = -w; m <= w; ++m) P[m + w] = s[std::abs(m)];
    // divide by (z-1)^2 twice by synthetic division
    for (int rep = 0; rep < 2; ++rep) {
        std::vector<long double> q(P.size() - 1);
        long double carry = 0;
        for (int i = (int)P.size() - 1; i >= 1; --i) { carry = P[i] + carry; q[i - 1] = carry; }
        P = q;
    }
    ZeroSol zs;
    if (P.size() > 1) {
        for (const cl& z : poly_roots(P)) if (std::abs(z) < 1.0L) zs.r.push_back(z);
    }
    const int nk = (int)zs.r.size();   // w-1
    // unknowns: b, K_0..K_{nk-1}; equations (D g)(n) = delta_n for n = 0..w-1, g(n) = b|n| + sum K_i (r_i^|n| - 1)
    const int nu = 1 + nk;
    std::vector<std::vector<cl>> A(nu, std::vector<cl>(nu + 1, cl(0)));
    for (int n = 0; n < nu; ++n) {
        for (int m = -w; m <= w; ++m) {
            const int j = std::abs(n + m);
            const long double sm = s[std::abs(m)];
            A[n][0] += sm * (long double)j;
            for (int i = 0; i < nk; ++i) A[n][1 + i] += sm * (std::pow(zs.r[i], j) - cl(1));
        }
        A[n][nu] = (n == 0) ? cl(1) : cl(0);
    }
    for (int c = 0; c < nu; ++c) {   // Gauss-Jordan with partial pivoting
        int piv = c;
        for (int r = c + 1; r < nu; ++r) if (std::abs(A[r][c]) > std::abs(A[piv][c])) piv = r;
        std::swap(A[c], A[piv]);
        for (int r = 0; r < nu; ++r) {
            if (r == c) continue;
            const cl f = A[r][c] / A[c][c];
            for (int k = c; k <= nu; ++k) A[r][k] -= f * A[c][k];
        }
    }
    zs.b = (A[0][nu] / A[0][0]).real();
    for (int i = 0; i < nk; ++i) zs.K.push_back(A[1 + i][nu] / A[1 + i][1 + i]);
    return zs;
}

std::mutex g_mtx;
std::map<int, ZeroSol> g_zero;
std::map<std::pair<int, double>, std::vector<double>> g_pos;

const std::vector<double>& positive_c(int order, double c, int nmin) {
    const auto key = std::make_pair(order, c);
    {
        std::lock_guard<std::mutex> lk(g_mtx);
        auto it = g_pos.find(key);
        if (it != g_pos.end() && (int)it->second.size() > nmin) return it->second;
    }
    // pole distance ~ sqrt(c) for small c; aliasing error ~ exp(-M * dist)
    const double dist = std::min(1.0, std::sqrt(c) * 0.9);
    int M = 64;
    while (M < 2 * (nmin + 1) + (int)(80.0 / dist)) M *= 2;
    std::vector<cplx> f(M), F(M);
    for (int j = 0; j < M; ++j) {
        const double t = 2.0 * M_PI * j / M;
        f[j] = 1.0 / (c + lgf_symbol(order, t));
    }
    fft_c2c(M, SIGN_BWD, f.data(), F.data());
    std::vector<double> g(M / 2);
    for (int n = 0; n < M / 2; ++n) g[n] = F[n].real() / M;
    std::lock_guard<std::mutex> lk(g_mtx);
    return g_pos[key] = g;
}

}  // namespace

double lgf1d(int order, int n, double c) {
    if (c < 0) throw std::runtime_error("lgf1d: c must be nonnegative");
    n = std::abs(n);
    if (c == 0.0) {
        ZeroSol* zs;
        {
            std::lock_guard<std::mutex> lk(g_mtx);
            auto it = g_zero.find(order);
            if (it == g_zero.end()) it = g_zero.emplace(order, zero_solution(order)).first;
            zs = &it->second;
        }
        cl v = zs->b * (long double)n;
        for (size_t i = 0; i < zs->K.size(); ++i) v += zs->K[i] * (std::pow(zs->r[i], n) - cl(1));
        return (double)v.real();
    }
    const auto& g = positive_c(order, c, n);
    return g[n];
}

// The far-field LGF expansions are evaluated in lgf_expansion.cpp.

}  // namespace oracle
