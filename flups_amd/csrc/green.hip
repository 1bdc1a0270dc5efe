// Green's functions evaluated on the device at setup, and the spectral multiply.
//
// Fill kernels restate the drivers of src/green_functions.cpp
//   3 unbounded directions          :27-167
//   2 unbounded + 1 spectral        :192-391
//   1 unbounded + 2 spectral        :407-538
//   fully spectral                  :560-708
// with the closed-form kernels of src/green_kernels.hpp (HEJ_2..10, CHAT_2,
// LGF_2..8, MEHR_4L/6L/4F/6F) evaluated in FP64. Special functions are computed here from their
// integral or series definitions:
//   K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt, trapezoidal rule with
//             step 1/4 (error ~ exp(-pi^2 / h) for an integrand analytic in
//             |Im t| < pi/2), Abramowitz & Stegun 9.6.24;
//   E_1(x)  = -gamma - ln x - sum (-x)^k/(k k!) for x <= 1 (A&S 5.1.11) and
//             the continued fraction of A&S 5.1.22 (modified Lentz) for x > 1.
// The multiply kernels restate src/dothemagic_std.ipp:56-83 and
// src/dothemagic_rot.ipp:102-349.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "core.hpp"
#include "green.hpp"
#include "group.hpp"

namespace flups {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEuler = 0.57721566490153286061;

__device__ double bessel_k(int nu, double x) {
    // exp(-x) factored out so that large x does not underflow before scaling
    // the step keeps exp(x (1 - cos y) - 2 pi y / h) below 1e-17 for the best strip half-width y
    const double hstep = fmin(0.25, 2.0 * kPi / sqrt(90.0 * fmax(x, 1e-300)));
    double s = 0.5;   // t = 0 term: exp(-x (cosh 0 - 1)) cosh(0) = 1, halved
    for (int i = 1; i < 4000; ++i) {
        const double t = i * hstep;
        const double term = exp(-x * (cosh(t) - 1.0)) * (nu ? cosh(t) : 1.0);
        s += term;
        if (term < 1e-18 * s) break;
    }
    return hstep * s * exp(-x);
}

__device__ double expint_e1(double x) {
    if (x <= 1.0) {
        double sum = 0.0, term = 1.0;
        for (int k = 1; k < 60; ++k) {
            term *= -x / k;
            const double add = -term / k;
            sum += add;
            if (fabs(add) < 1e-18 * fabs(sum)) break;
        }
        return -kEuler - log(x) + sum;
    }
    // E1(x) = exp(-x) / (x + 1/(1 + 1/(x + 2/(1 + 2/(x + ...)))))  evaluated as b_k continued fraction
    double b = x + 1.0, c = 1e300, d = 1.0 / b, h = d;
    for (int i = 1; i < 300; ++i) {
        const double an = -(double)i * i;
        b += 2.0;
        d = 1.0 / (an * d + b);
        c = b + an / c;
        const double del = c * d;
        h *= del;
        if (fabs(del - 1.0) < 1e-17) break;
    }
    return h * exp(-x);
}

__device__ __forceinline__ int symid(int i, double ss) {
    if (ss == 0.0 || i <= ss) return i;
    const int v = (int)fabs(2.0 * ss - i);
    return -(v > 1 ? v : 1);
}

__device__ double hej3_poly(int t, double rho) {
    const double r2 = rho * rho;
    switch (t) {
        case HEJ_4: return rho;
        case HEJ_6: return rho * (7.0 / 4.0 - r2 / 4.0);
        case HEJ_8: return rho * (19.0 / 8.0 - (2.0 / 3.0) * r2 + r2 * r2 / 24.0);
        case HEJ_10: return rho * (187.0 / 64.0 - (233.0 / 192.0) * r2 + (29.0 / 192.0) * r2 * r2 - r2 * r2 * r2 / 192.0);
        default: return 0.0;
    }
}
__device__ double hej2_q(int t, double r2) {
    switch (t) {
        case HEJ_4: return 0.5;
        case HEJ_6: return 0.75 - r2 / 8.0;
        case HEJ_8: return 11.0 / 12.0 - 7.0 * r2 / 24.0 + r2 * r2 / 48.0;
        case HEJ_10: return 25.0 / 24.0 - 23.0 * r2 / 48.0 + 13.0 * r2 * r2 / 192.0 - r2 * r2 * r2 / 384.0;
        default: return 0.0;
    }
}
__device__ int lgf_order(int t) {
    return t == LGF_2 ? 2 : t == LGF_4 ? 4 : t == LGF_6 ? 6 : t == LGF_8 ? 8 : 0;
}
__device__ bool is_mehr(int t) { return t == MEHR_4L || t == MEHR_6L || t == MEHR_4F || t == MEHR_6F; }
// kernels given by a table near the origin and an expansion beyond (LGF, Mehrstellen)
__device__ bool is_table(int t) { return lgf_order(t) != 0 || is_mehr(t); }

// Mehrstellen symbols in y = sin^2(k h / 2) (src/green_kernels.hpp:744-773)
__device__ double mehr_spectral(int t, const double k[3], double h) {
    double y[3];
    for (int d = 0; d < 3; ++d) {
        const double sv = sin(0.5 * k[d] * h);
        y[d] = sv * sv;
    }
    const double s1 = y[0] + y[1] + y[2], s2 = y[0] * y[1] + y[1] * y[2] + y[0] * y[2], s3 = y[0] * y[1] * y[2];
    const double q = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
    const bool six = (t == MEHR_6L || t == MEHR_6F);
    const double left = -4.0 * s1 + 8.0 / 3.0 * s2 - (six ? 32.0 / 15.0 * s3 : 0.0);
    double right = 1.0;
    if (t == MEHR_4F) right = 1.0 - s1 / 3.0;
    if (t == MEHR_6F) right = 1.0 - s1 / 3.0 + 8.0 / 45.0 * s2 - q / 15.0;
    return h * h * right / left;
}

// sine integral Si(x), x >= 0: power series up to 4; beyond, Si = pi/2 + Im E1(ix)
// with E1 from its continued fraction (Abramowitz & Stegun 5.1.22, modified Lentz)
__device__ double sine_integral(double x) {
    if (x <= 4.0) {
        double t = x, s = x;
        const double x2 = x * x;
        for (int k = 1; k < 40; ++k) {
            t *= -x2 / ((2.0 * k) * (2.0 * k + 1));
            const double term = t / (2 * k + 1);
            s += term;
            if (fabs(term) < 1e-18 * fabs(s)) break;
        }
        return s;
    }
    // complex arithmetic on (re, im) pairs, z = i x
    double br = 1.0, bi = x;                   // b = z + 1
    double cr = 1e300, ci = 0.0;               // c = 1 / tiny
    double dr, di;
    {
        const double m = br * br + bi * bi;
        dr = br / m;
        di = -bi / m;
    }
    double hr = dr, hi = di;
    for (int i = 1; i < 2000; ++i) {
        const double an = -(double)i * i;
        br += 2.0;
        // d = 1 / (an d + b)
        double tr = an * dr + br, ti = an * di + bi;
        double m = tr * tr + ti * ti;
        dr = tr / m;
        di = -ti / m;
        // c = b + an / c
        m = cr * cr + ci * ci;
        const double qr = an * cr / m, qi = -an * ci / m;
        cr = br + qr;
        ci = bi + qi;
        const double er = cr * dr - ci * di, ei = cr * di + ci * dr;   // del = c d
        const double nr = hr * er - hi * ei, ni = hr * ei + hi * er;
        hr = nr;
        hi = ni;
        if (fabs(er - 1.0) + fabs(ei) < 1e-17) break;
    }
    // E1 = exp(-i x) h
    const double cx = cos(x), sx = sin(x);
    const double e1i = cx * hi - sx * hr;
    return 0.5 * kPi + e1i;
}
__device__ double lgf_symbol(int order, double k) {
    const double s1 = sin(0.5 * k), s2 = sin(k), s3 = sin(1.5 * k), s4 = sin(2.0 * k);
    switch (order) {
        case 2: return 4.0 * s1 * s1;
        case 4: return 4.0 * (4.0 / 3.0 * s1 * s1 - 1.0 / 12.0 * s2 * s2);
        case 6: return 4.0 * (1.5 * s1 * s1 - 0.15 * s2 * s2 + 1.0 / 90.0 * s3 * s3);
        default: return 4.0 * (1.6 * s1 * s1 - 0.2 * s2 * s2 + 8.0 / 315.0 * s3 * s3 - 1.0 / 560.0 * s4 * s4);
    }
}
__device__ double ipow(double x, int e) {
    double r = 1.0;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}
// far-field expansion: sum coef n1^a n2^b n3^c n^-p log(n)^l
__device__ double expansion(const ExpTerm* t, int nt, double n1, double n2, double n3) {
    const double n = sqrt(n1 * n1 + n2 * n2 + n3 * n3), ln = log(n);
    double s = 0.0;
    for (int i = 0; i < nt; ++i)
        s += t[i].coef * ipow(n1, t[i].a) * ipow(n2, t[i].b) * ipow(n3, t[i].c) / ipow(n, t[i].p) * (t[i].l ? ln : 1.0);
    return s;
}

struct FillGeom {
    int n[3], nf, axis;
};
__device__ __forceinline__ void unflat(long long e, const FillGeom& g, int i[3]) {
    i[0] = (int)(e % g.n[0]);
    i[1] = (int)((e / g.n[0]) % g.n[1]);
    i[2] = (int)(e / ((long long)g.n[0] * g.n[1]));
}

__global__ void fill_3unb(double* g, FillGeom geo, GreenArgs A) {
    const long long tot = (long long)geo.n[0] * geo.n[1] * geo.n[2];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        int i[3];
        unflat(e, geo, i);
        const int is[3] = {symid(i[0], A.symstart[0]), symid(i[1], A.symstart[1]), symid(i[2], A.symstart[2])};
        const double x0 = is[0] * A.hfact[0], x1 = is[1] * A.hfact[1], x2 = is[2] * A.hfact[2];
        const double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
        const bool ord = is_table(A.type);
        double v;
        if (ord) {
            const int ax0 = geo.axis, ax1 = (ax0 + 1) % 3, ax2 = (ax0 + 2) % 3;
            const int a = abs(is[ax0]), b = abs(is[ax1]), c = abs(is[ax2]);
            const double gv = (a < 32 && b < 32 && c < 32) ? A.table[a + 32 * b + 1024 * c] : expansion(A.exp, A.nexp, a, b, c);
            v = -gv / A.hfact[0];
        } else if (A.type == CHAT_2) {
            v = -1.0 / (4.0 * kPi * r);
        } else if (A.type == HEJ_0) {   // -Si(rho) / (2 pi^2 sigma rho), sigma = h / pi
            const double rho = r / A.length;
            v = rho > 0.0 ? -sine_integral(rho) / (rho * 2.0 * kPi * kPi * A.length) : 0.0;
        } else {
            const double rho = r / A.length;
            v = -1.0 / (4.0 * kPi * r) * (erf(rho / sqrt(2.0)) + hej3_poly(A.type, rho) * exp(-0.5 * rho * rho) / sqrt(2.0 * kPi));
        }
        if (e == 0 && !ord) v = A.g0;
        g[e * geo.nf] = v;
    }
}

__global__ void fill_2unb(double* g, FillGeom geo, GreenArgs A) {
    const long long tot = (long long)geo.n[0] * geo.n[1] * geo.n[2];
    const double sh = A.hfact[0] + A.hfact[1] + A.hfact[2];
    const double sk = A.kfact[0] + A.kfact[1] + A.kfact[2];
    const double req = sqrt(A.hfact[0] * A.hfact[1] + A.hfact[1] * A.hfact[2] + A.hfact[2] * A.hfact[0]) / sqrt(kPi);
    const bool ord = is_table(A.type);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        int i[3];
        unflat(e, geo, i);
        const int is[3] = {symid(i[0], A.symstart[0]), symid(i[1], A.symstart[1]), symid(i[2], A.symstart[2])};
        const double k = (is[0] + A.koffset[0]) * A.kfact[0] + (is[1] + A.koffset[1]) * A.kfact[1] + (is[2] + A.koffset[2]) * A.kfact[2];
        const double x0 = is[0] * A.hfact[0], x1 = is[1] * A.hfact[1], x2 = is[2] * A.hfact[2];
        const double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
        const bool r0 = r <= sh * 0.2;
        const bool k0 = !r0 && k <= sk * 0.2;
        double v = 0.0;
        if (ord) {
            if (r0 || k0) {
                const int ax0 = geo.axis, ax1 = (ax0 + 1) % 3, ax2 = (ax0 + 2) % 3;
                const int a[3] = {abs(is[ax0]), abs(is[ax1]), abs(is[ax2])};
                const int j1 = (a[0] == 0) ? a[1] : a[0];
                const int j2 = (a[0] == 0) ? a[2] : (a[1] == 0 ? a[2] : a[1]);
                const double gv = (j1 < 32 && j2 < 32) ? A.table[j1 + 32 * j2] : expansion(A.exp, A.nexp, j1, j2, 0.0);
                v = -gv;
            }
        } else if (A.type == CHAT_2) {
            if (r0) {
                const double kr = k * req;
                v = -(1.0 - kr * bessel_k(1, kr)) / (kPi * kr * kr);
            } else if (k0) {
                v = log(r) / (2.0 * kPi);
            } else {
                v = -bessel_k(0, fabs(k) * r) / (2.0 * kPi);
            }
        } else if (A.type == HEJ_0) {
            // src/green_kernels.hpp:416-428, the same expression at r = 0 and on k = 0, with
            // Ji0c(rho) from the table of the integer squared distances s = |i|^2 (isotropic
            // grid: rho = pi sqrt(s) < 30 for s <= 91); elsewhere zero_ = 0
            // (src/green_kernels.hpp:430-432), which no spectral rewrite replaces for HEJ_0
            // (src/Solver.cpp:1089-1101)
            if (r0 || k0) {
                const double rho = r / A.length;
                const int s = (A.hfact[0] != 0.0 ? is[0] * is[0] : 0) + (A.hfact[1] != 0.0 ? is[1] * is[1] : 0) +
                              (A.hfact[2] != 0.0 ? is[2] * is[2] : 0);
                v = (rho < 30.0 && s < A.ntable) ? (A.table[s] + log(2.0 * A.length) - kEuler) / (2.0 * kPi)
                                                 : log(r) / (2.0 * kPi);
            } else {
                v = 0.0;
            }
        } else {
            if (r0) {
                v = -(0.5 * kEuler - log(sqrt(2.0) * A.length) + hej2_q(A.type, 0.0)) / (2.0 * kPi);
            } else if (k0) {
                const double rho2 = (r / A.length) * (r / A.length);
                v = (log(r) - hej2_q(A.type, rho2) * exp(-0.5 * rho2) + 0.5 * expint_e1(0.5 * rho2)) / (2.0 * kPi);
            } else {
                v = 0.0;   // zero_ (src/green_kernels.hpp:430-432); the spectral rewrite replaces these planes
            }
        }
        if (e == 0 && A.type == CHAT_2) v = 0.25 / (2.0 * kPi) * (kPi - 6.0 + 2.0 * log(0.5 * kPi * req * req));
        g[e * geo.nf] = v;
    }
}

// ---- 1-D lattice Green's function of the order-2w stencil shifted by c ----
// g(n) = 1/(2 pi) int cos(n t)/(c + s(t)) dt, s(t) = s_0 + 2 sum_m s_m cos(m t).
// With z = exp(-L) the roots of c + s = 0 inside the unit circle, c + s_0 + 2 sum s_m cosh(m L) = 0;
// the residues give g(n) = sum_i A_i exp(-|n| L_i), A_i = -1 / (2 sum_m m s_m sinh(m L_i)).
// The roots are found in mu = cosh(L) - 1 (a degree-w polynomial c + sum_j p_j mu^j), so that
// L = 2 asinh(sqrt(mu/2)) stays accurate for small c. For c = 0 the double root z = 1 is
// regularised into -|n|/2 and the constant that makes g(0) = 0.
struct cz {
    double x, y;
};
__device__ __forceinline__ cz operator+(cz a, cz b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cz operator-(cz a, cz b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cz operator*(cz a, cz b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cz operator*(double s, cz a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ cz cdiv(cz a, cz b) {
    const double den = b.x * b.x + b.y * b.y;
    return {(a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den};
}
__device__ __forceinline__ double cabs_(cz a) { return hypot(a.x, a.y); }
__device__ cz cexp_(cz a) {
    const double e = exp(a.x);
    double s, c;
    sincos(a.y, &s, &c);
    return {e * c, e * s};
}
__device__ cz clog_(cz a) { return {log(cabs_(a)), atan2(a.y, a.x)}; }
__device__ cz csqrt_(cz a) {
    const double r = cabs_(a);
    if (r == 0.0) return {0.0, 0.0};
    const double t = sqrt(0.5 * (r + fabs(a.x)));
    if (a.x >= 0.0) return {t, a.y / (2.0 * t)};
    return {fabs(a.y) / (2.0 * t), a.y >= 0 ? t : -t};
}
__device__ cz csinh_(cz z) {
    if (cabs_(z) < 0.1) {   // odd series, error below z^13/13!
        const cz z2 = z * z;
        cz t = z, s = z;
        for (int k = 1; k <= 5; ++k) {
            t = (1.0 / ((2.0 * k) * (2.0 * k + 1.0))) * (t * z2);
            s = s + t;
        }
        return s;
    }
    return 0.5 * (cexp_(z) - cexp_({-z.x, -z.y}));
}
__device__ cz casinh_(cz w) {
    if (cabs_(w) < 0.01) {   // w - w^3/6 + 3 w^5/40 - 5 w^7/112
        const cz w2 = w * w;
        return w * (cz{1.0, 0.0} + w2 * (cz{-1.0 / 6.0, 0.0} + w2 * (cz{3.0 / 40.0, 0.0} + w2 * cz{-5.0 / 112.0, 0.0})));
    }
    return clog_(w + csqrt_(w * w + cz{1.0, 0.0}));
}
// roots of sum_{j=0}^{deg} a_j mu^j (a_deg != 0), deg <= 4, Durand-Kerner then Newton polish
__device__ int poly_roots(const double* a, int deg, cz* r) {
    for (int i = 0; i < deg; ++i) {
        const double ang = 0.4 + 2.0 * kPi * i / deg;
        r[i] = {cos(ang), sin(ang)};
    }
    auto eval = [&](cz x) {
        cz v = {a[deg], 0.0};
        for (int j = deg - 1; j >= 0; --j) v = v * x + cz{a[j], 0.0};
        return v;
    };
    auto deriv = [&](cz x) {
        cz v = {deg * a[deg], 0.0};
        for (int j = deg - 1; j >= 1; --j) v = v * x + cz{j * a[j], 0.0};
        return v;
    };
    for (int it = 0; it < 500; ++it) {
        double delta = 0.0;
        for (int i = 0; i < deg; ++i) {
            cz den = {a[deg], 0.0};
            for (int j = 0; j < deg; ++j)
                if (j != i) den = den * (r[i] - r[j]);
            const cz dz = cdiv(eval(r[i]), den);
            r[i] = r[i] - dz;
            delta = fmax(delta, cabs_(dz) / fmax(cabs_(r[i]), 1e-300));
        }
        if (delta < 1e-16) break;
    }
    for (int i = 0; i < deg; ++i)
        for (int it = 0; it < 3; ++it) {
            const cz dp = deriv(r[i]);
            if (cabs_(dp) == 0.0) break;
            r[i] = r[i] - cdiv(eval(r[i]), dp);
        }
    return deg;
}
// roots and residues of the 1-D LGF for one shift c (they do not depend on n)
struct Lgf1d {
    int nr;
    bool c0;
    cz L[4], Ai[4];
};
__device__ Lgf1d lgf1d_setup(const GreenArgs& A, int order, double c) {
    const int w = order / 2;
    const double* s = A.lgf_s + 5 * (w - 1);   // s_0..s_w
    const double* p = A.lgf_p + 5 * (w - 1);   // p_0 (=0)..p_w
    double a[5];
    cz mu[4];
    Lgf1d R;
    R.c0 = c == 0.0;
    if (R.c0) {
        for (int j = 1; j <= w; ++j) a[j - 1] = p[j];
        R.nr = poly_roots(a, w - 1, mu);
    } else {
        a[0] = c;
        for (int j = 1; j <= w; ++j) a[j] = p[j];
        R.nr = poly_roots(a, w, mu);
    }
    for (int i = 0; i < R.nr; ++i) {
        cz L = 2.0 * casinh_(csqrt_(0.5 * mu[i]));
        if (L.x < 0.0) L = {-L.x, -L.y};
        cz den = {0.0, 0.0};
        for (int m = 1; m <= w; ++m) den = den + (m * s[m]) * csinh_((double)m * L);
        R.L[i] = L;
        R.Ai[i] = cdiv({-0.5, 0.0}, den);
    }
    return R;
}
__device__ double lgf1d_at(const Lgf1d& R, int n) {
    const double an = fabs((double)n);
    double g = R.c0 ? -0.5 * an : 0.0;
    for (int i = 0; i < R.nr; ++i) {
        cz term = R.Ai[i] * cexp_((-an) * R.L[i]);
        if (R.c0) term = term - R.Ai[i];
        g += term.x;
    }
    return g;
}
__device__ double lgf1d_eval(const GreenArgs& A, int order, int n, double c) { return lgf1d_at(lgf1d_setup(A, order, c), n); }

// 1-unbounded LGF kernels: one thread per column along the unbounded direction ud (the
// roots depend on the transverse wave numbers only), then the column's values for every n
__global__ void fill_1unb_lgf(double* g, FillGeom geo, GreenArgs A, int ud) {
    const int da = (ud + 1) % 3, db = (ud + 2) % 3;
    const long long ncol = (long long)geo.n[da] * geo.n[db];
    const long long sd[3] = {1, geo.n[0], (long long)geo.n[0] * geo.n[1]};
    const double h = A.hfact[0] + A.hfact[1] + A.hfact[2];
    const int ord = lgf_order(A.type);
    for (long long col = blockIdx.x * (long long)blockDim.x + threadIdx.x; col < ncol; col += (long long)gridDim.x * blockDim.x) {
        int i[3];
        i[da] = (int)(col % geo.n[da]);
        i[db] = (int)(col / geo.n[da]);
        i[ud] = 0;
        double c = 0.0;
        for (int dd = 0; dd < 3; ++dd) {
            const int is = symid(i[dd], A.symstart[dd]);
            c += lgf_symbol(ord, (is + A.koffset[dd]) * A.kfact[dd] * h);
        }
        const Lgf1d R = lgf1d_setup(A, ord, c);
        const long long e0 = i[da] * sd[da] + i[db] * sd[db];
        for (int iu = 0; iu < geo.n[ud]; ++iu) {
            const double x = symid(iu, A.symstart[ud]) * A.hfact[ud];
            const double r = sqrt(x * x);
            const int nn = (int)round(fabs(r) / h);
            g[(e0 + iu * sd[ud]) * geo.nf] = -h * lgf1d_at(R, nn);
        }
    }
}

// 1-D lattice Green's functions of the Mehrstellen stencils along the unbounded direction
// (src/lgf_1unb2spe.hpp:500-611): transverse wave numbers k1 >= k2 (times h), the root r of
// a1 r^2 + a0 r + a1 = 0 inside the unit circle, r = lambda - sqrt(lambda - 1) sqrt(lambda + 1)
// with lambda = -a0 / (2 a1) (src/lgf_1unb2spe.hpp:80-98); "left": a1 G(n-1) + a0 G(n) +
// a1 G(n+1) = delta_n, "full": the right-hand stencil b applied to it.
__device__ double r_from_lambda(double lambda) {
    if (isinf(lambda)) return 0.0;
    if (fabs(lambda) > 1.0e2) {   // 1/(2l) + 1/(8l^3) + 1/(16l^5) + 5/(128l^7) + 7/(256l^9)
        const double y = 1.0 / lambda;
        const double e[10] = {0, 1. / 2., 0, 1. / 8., 0, 1. / 16., 0, 5. / 128., 0, 7. / 256.};
        double out = e[9];
        for (int i = 8; i >= 0; --i) out = y * out + e[i];
        return out;
    }
    if (lambda >= 1.0) return lambda - sqrt(lambda - 1.0) * sqrt(lambda + 1.0);
    if (lambda > -1.0) return lambda;   // both roots on the unit circle: the real part
    return lambda + sqrt(1.0 - lambda) * sqrt(-1.0 - lambda);
}
__device__ double mehr_1d(int type, int n, double k1, double k2) {
    const bool six = type == MEHR_6L || type == MEHR_6F, full = type == MEHR_4F || type == MEHR_6F;
    const int an = abs(n);
    if (k1 == 0 && k2 == 0) {
        double v = -0.5 * an;
        if (full && !six) v -= (n == 0) * (1. / 12.);
        if (full && six) v -= (an == 1) * (-1. / 240.) + (n == 0) * (11. / 120.);
        return v;
    }
    const double s0 = sin(k1 / 2), s1 = sin(k2 / 2);
    const double y0 = s0 * s0, y1 = s1 * s1;
    const double a0 = six ? -8. / 5. * y0 * y1 + 8. / 3. * (y0 + y1) + 2.0 : -8. / 3. * (y0 * y1 - y0 - y1) + 2.0;
    const double a1 = six ? -8. / 15. * y0 * y1 + 2. / 3. * (y0 + y1) - 1.0 : 2. / 3. * (y0 + y1) - 1.0;
    const double lambda = -a0 / (2. * a1);
    const double r = r_from_lambda(lambda);
    if (!full) {
        if (!isfinite(lambda)) return six ? -(double)(n == 0) / a0 : (double)(n == 0) / a0;
        return (r / a1) * pow(r, (double)an) / (r * r - 1.0);
    }
    if (!six) {
        const double b0 = 5. / 6. - 1. / 3. * (y0 + y1), b1 = 1. / 12.;
        if (!isfinite(lambda)) return (n == 0) * (b0 / a0) + (n == 1) * (b1 / a0);
        if (n == 0) return (r / a1) * (b0 + 2. * b1 * r) / (r * r - 1.);
        return (b1 + b0 * r + b1 * r * r) * pow(r, (double)an) / (a1 * (r * r - 1));
    }
    const double b0 = -1. / 15. * (y0 * y0 + y1 * y1) + 8. / 45. * y0 * y1 - 11. / 45. * (y0 + y1) + 97. / 120.;
    const double b1 = -2. / 45. * (y0 + y1) + 1. / 10.;
    const double b2 = -1. / 240.;
    const double pr = b0 * r * r + b1 * (r + r * r * r) + b2 * (1.0 + r * r * r * r);
    const double base = pr * pow(r, (double)(an - 1)) / (a1 * (r * r - 1.0));
    return base + (an == 1) * (b2 / a1) + (n == 0) * (b1 * a1 - b2 * a0) / (a1 * a1);
}

__global__ void fill_1unb(double* g, FillGeom geo, GreenArgs A) {
    const long long tot = (long long)geo.n[0] * geo.n[1] * geo.n[2];
    const double sk = A.kfact[0] + A.kfact[1] + A.kfact[2];
    const double h = A.hfact[0] + A.hfact[1] + A.hfact[2];
    const int ord = lgf_order(A.type);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        int i[3];
        unflat(e, geo, i);
        const int is[3] = {symid(i[0], A.symstart[0]), symid(i[1], A.symstart[1]), symid(i[2], A.symstart[2])};
        const double kk[3] = {(is[0] + A.koffset[0]) * A.kfact[0], (is[1] + A.koffset[1]) * A.kfact[1], (is[2] + A.koffset[2]) * A.kfact[2]};
        const double k = sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2]);
        const double x0 = is[0] * A.hfact[0], x1 = is[1] * A.hfact[1], x2 = is[2] * A.hfact[2];
        const double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
        const bool k0 = k <= sk * 0.2;
        double v;
        if (ord) {
            double c = 0.0;
            for (int dd = 0; dd < 3; ++dd) c += lgf_symbol(ord, kk[dd] * h);
            const int nn = (int)round(fabs(r) / h);
            v = -h * lgf1d_eval(A, ord, nn, c);
        } else if (is_mehr(A.type)) {
            // src/green_kernels.hpp:624-671: one k is zero; |k1| >= |k2| >= 0
            const double a0 = fabs(kk[0]), a1 = fabs(kk[1]), a2 = fabs(kk[2]);
            const double k1 = fmax(a0, fmax(a1, a2)), k2 = a0 + a1 + a2 - k1;
            const int nn = (int)round(fabs(r) / h);
            v = -h * mehr_1d(A.type, nn, k1 * h, k2 * h);
        } else if (A.type == CHAT_2) {
            v = k0 ? 0.5 * fabs(r) : -0.5 * exp(-k * r) / k;
        } else {
            const double sig = A.length, rho = r / sig, rho2 = rho * rho;
            const double cst = 1.0 / (sqrt(2.0) * sqrt(kPi));
            if (k0) {
                const double er = 0.5 * r * erf(r / (sqrt(2.0) * sig));
                const double ex = exp(-0.5 * rho2);
                double tail;
                switch (A.type) {
                    case HEJ_2: tail = 1.0 - ex; break;
                    case HEJ_4: tail = 0.5 * (1.0 - ex); break;
                    case HEJ_6: tail = 0.125 * (3.0 - ex * (rho2 + 3.0)); break;
                    case HEJ_8: tail = 5.0 / 16.0 - ex * (-rho2 * rho2 / 48.0 + 0.25 * rho2 + 5.0 / 16.0); break;
                    default: tail = 35.0 / 128.0 - ex * (rho2 * rho2 * rho2 / 384.0 - 23.0 / 384.0 * rho2 * rho2 + 47.0 / 128.0 * rho2 + 35.0 / 128.0);
                }
                v = er - tail * sig * cst;
            } else {
                const double s = k * sig, s2 = s * s;
                const double sub = (s * rho > 100.0) ? 0.0
                                                     : ((1.0 - erf((s - rho) / sqrt(2.0))) * exp(-s * rho) +
                                                        (1.0 - erf((s + rho) / sqrt(2.0))) * exp(s * rho));
                double poly;
                switch (A.type) {
                    case HEJ_2: poly = 0.0; break;
                    case HEJ_4: poly = 0.25; break;
                    case HEJ_6: poly = 5.0 / 16.0 + (s2 - rho2) / 16.0; break;
                    case HEJ_8: poly = 11.0 / 32.0 + s2 / 12.0 - 0.125 * rho2 - s2 * rho2 / 48.0 + (s2 * s2 + rho2 * rho2) / 96.0; break;
                    default:
                        poly = 93.0 / 256.0 + 73.0 / 768.0 * s2 - 47.0 / 256.0 * rho2 - 17.0 / 384.0 * s2 * rho2 + 11.0 / 768.0 * s2 * s2 +
                               23.0 / 768.0 * rho2 * rho2 + (s2 * rho2 * rho2 - s2 * s2 * rho2) / 256.0 +
                               (s2 * s2 * s2 - rho2 * rho2 * rho2) / 768.0;
                }
                v = -0.25 * sig / s * sub - sig * sqrt(2.0) / sqrt(kPi) * poly * exp(-0.5 * (s2 + rho2));
            }
        }
        g[e * geo.nf] = v;
    }
}

__global__ void fill_0unb(double* g, FillGeom geo, GreenArgs A, int lo0, int lo1, int lo2) {
    const long long tot = (long long)geo.n[0] * geo.n[1] * geo.n[2];
    const int ord = lgf_order(A.type);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        int i[3];
        unflat(e, geo, i);
        if (e == 0 && A.koffset[0] + A.koffset[1] + A.koffset[2] < 0.2) {
            g[0] = 0.0;
            continue;
        }
        if (i[0] < lo0 || i[1] < lo1 || i[2] < lo2) continue;
        const int il[3] = {symid(i[0], A.symstart[0]), symid(i[1], A.symstart[1]), symid(i[2], A.symstart[2])};
        const double k[3] = {(il[0] + A.koffset[0]) * A.kfact[0], (il[1] + A.koffset[1]) * A.kfact[1], (il[2] + A.koffset[2]) * A.kfact[2]};
        const double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        double v;
        if (ord) {
            v = -A.hgrid * A.hgrid / (lgf_symbol(ord, k[0] * A.hgrid) + lgf_symbol(ord, k[1] * A.hgrid) + lgf_symbol(ord, k[2] * A.hgrid));
        } else if (is_mehr(A.type)) {
            v = mehr_spectral(A.type, k, A.hgrid);
        } else if (A.type == CHAT_2 || A.type == HEJ_0) {
            v = -1.0 / ksq;
        } else {
            const double hh = 0.5 * ksq * A.length * A.length;
            const int jmax = A.type == HEJ_2 ? 0 : A.type == HEJ_4 ? 1 : A.type == HEJ_6 ? 2 : A.type == HEJ_8 ? 3 : 4;
            double p = 1.0, term = 1.0;
            for (int j = 1; j <= jmax; ++j) { term *= hh / j; p += term; }
            v = -p * exp(-hh) / ksq;
        }
        g[e * geo.nf] = v;
    }
}

__global__ void scale_kernel(double* g, size_t n, double f) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) g[i] *= f;
}

// the fully spectral zero mode after the 2-unbounded/1-spectral completion (green_functions.cpp:703-706)
__global__ void zero_first(double* g) { g[0] = 0.0; }

// S: storage of the data and the Green's function (double, or float for FP32 solves); FP64 arithmetic
template <typename S>
struct MultArgs {
    S* d;
    const S* g;
    size_t per;
    int nf, lda;
    double normfact;
};
template <typename S>
struct mult_std {
    using Args = MultArgs<S>;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void run(const Args& a, Blk blk) {
        const size_t per = a.per;
        const int nf = a.nf;
        for (size_t i = blk.x * (size_t)NT + threadIdx.x; i < per; i += (size_t)blk.nx * NT) {
            for (int c = 0; c < a.lda; ++c) {
                S* x = a.d + c * per * nf;
                if (nf == 1) {
                    x[i] = (S)((double)x[i] * a.normfact * (double)a.g[i]);
                } else {
                    const double p = x[2 * i], q = x[2 * i + 1], gr = a.g[2 * i], gi = a.g[2 * i + 1];
                    x[2 * i] = (S)(a.normfact * (p * gr - q * gi));
                    x[2 * i + 1] = (S)(a.normfact * (p * gi + q * gr));
                }
            }
        }
    }
};

__global__ void to_float(const double* a, float* b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

template <typename S>
struct RotMultArgs {
    S* d;
    const S* g;
    RotArgs R;
};
template <typename S>
struct mult_rot {
    using Args = RotMultArgs<S>;
    static constexpr int NT = 256;
    // derivative symbol along one axis at k h: k itself (spectral), the modified wavenumber of
    // the centred differences of order 2, 4, 6 otherwise. The difference symbols are one
    // out-of-line function: inlined, their sines (18 per point) made the kernel tens of thousands
    // of instructions long, and the spectral path paid for it in instruction-cache misses.
    static __device__ __noinline__ double sym_fd(int order, double k, double hh) {
        switch (order) {
            case FD2: return sin(k * hh) / hh;
            case FD4: return (4.0 / 3.0) * sin(k * hh) / hh - (1.0 / 6.0) * sin(2.0 * k * hh) / hh;
            default: return 1.5 * sin(k * hh) / hh - 0.3 * sin(2.0 * k * hh) / hh + (1.0 / 30.0) * sin(3.0 * k * hh) / hh;
        }
    }
    // I: index type, 32-bit below 2^31 points per component (the divisions by the extents are
    // then 32-bit ones), 64-bit above; FD: centred-difference symbols (else spectral, k itself)
    template <typename I, bool FD>
    static __device__ __forceinline__ void body(const Args& a, Blk blk) {
        auto sym = [](int order, double k, double hh) -> double { return FD ? sym_fd(order, k, hh) : k; };
        S* d = a.d;
        const S* g = a.g;
        const RotArgs& R = a.R;
        const I n0 = (I)R.n[0], n1 = (I)R.n[1];
        const I per = n0 * n1 * (I)R.n[2];
        for (I e = (I)blk.x * NT + threadIdx.x; e < per; e += (I)blk.nx * NT) {
            const I r = e / n0;
            const int i0 = (int)(e - r * n0), i1 = (int)(r % n1), i2 = (int)(r / n1);
            const int is[3] = {symid(i0, R.symstart[0]), symid(i1, R.symstart[1]), symid(i2, R.symstart[2])};
            double Kr[3][3], Ki[3][3];
            for (int dd = 0; dd < 3; ++dd)
                for (int c = 0; c < 3; ++c) {
                    const double base = is[dd] + R.koffset[dd];
                    if (R.nf == 1) {
                        Kr[dd][c] = sym(R.order, base * (R.kf[dd][c][0] + R.kf[dd][c][1]), R.h[dd]);
                        Ki[dd][c] = 0.0;
                    } else {
                        Kr[dd][c] = sym(R.order, base * R.kf[dd][c][0], R.h[dd]);
                        Ki[dd][c] = sym(R.order, base * R.kf[dd][c][1], R.h[dd]);
                    }
                }
            // complex values as one 8- or 16-byte load each
            using V = typename std::conditional<std::is_same<S, float>::value, float2, double2>::type;
            double fr[3], fi[3];
            for (int c = 0; c < 3; ++c) {
                if (R.nf == 2) {
                    const V v = reinterpret_cast<const V*>(d)[c * per + e];
                    fr[c] = (double)v.x;
                    fi[c] = (double)v.y;
                } else {
                    fr[c] = (double)d[c * per + e];
                    fi[c] = 0.0;
                }
            }
            auto dr = [&](int c, int dd) { return fr[c] * Kr[dd][c] - fi[c] * Ki[dd][c]; };
            auto di = [&](int c, int dd) { return fr[c] * Ki[dd][c] + fi[c] * Kr[dd][c]; };
            const double rr[3] = {dr(2, 1) - dr(1, 2), dr(0, 2) - dr(2, 0), dr(1, 0) - dr(0, 1)};
            const double ri[3] = {di(2, 1) - di(1, 2), di(0, 2) - di(2, 0), di(1, 0) - di(0, 1)};
            if (R.nf == 1) {
                for (int c = 0; c < 3; ++c) d[c * per + e] = (S)(R.normfact * rr[c] * (double)g[e]);
            } else {
                const V gv = reinterpret_cast<const V*>(g)[e];
                const double gr = gv.x, gi = gv.y;
                for (int c = 0; c < 3; ++c) {
                    V o;
                    o.x = (S)(R.normfact * (rr[c] * gr - ri[c] * gi));
                    o.y = (S)(R.normfact * (rr[c] * gi + ri[c] * gr));
                    reinterpret_cast<V*>(d)[c * per + e] = o;
                }
            }
        }
    }
    static __device__ __forceinline__ void run(const Args& a, Blk blk) {
        const size_t per = (size_t)a.R.n[0] * a.R.n[1] * a.R.n[2];
        const bool fd = a.R.order != SPE;
        if (3 * per * a.R.nf < (1u << 31)) {
            if (fd) body<unsigned, true>(a, blk);
            else body<unsigned, false>(a, blk);
        } else {
            if (fd) body<size_t, true>(a, blk);
            else body<size_t, false>(a, blk);
        }
    }
};

unsigned grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return (unsigned)b;
}

}  // namespace

void launch_green_fill(int nunb, double* g, const int n[3], int nf, int axis, const GreenArgs& A, const int* lo,
                       hipStream_t s) {
    FillGeom geo;
    for (int i = 0; i < 3; ++i) geo.n[i] = n[i];
    geo.nf = nf;
    geo.axis = axis;
    const size_t tot = (size_t)n[0] * n[1] * n[2];
    const unsigned gb = grid_for(tot);
    if (nunb == 3) hipLaunchKernelGGL(fill_3unb, dim3(gb), dim3(256), 0, s, g, geo, A);
    else if (nunb == 2) hipLaunchKernelGGL(fill_2unb, dim3(gb), dim3(256), 0, s, g, geo, A);
    else if (nunb == 1 && (A.type == LGF_2 || A.type == LGF_4 || A.type == LGF_6 || A.type == LGF_8)) {
        const int ud = A.hfact[0] != 0.0 ? 0 : A.hfact[1] != 0.0 ? 1 : 2;   // the unbounded direction
        const size_t ncol = tot / (size_t)n[ud];
        hipLaunchKernelGGL(fill_1unb_lgf, dim3(grid_for(ncol)), dim3(256), 0, s, g, geo, A, ud);
    } else if (nunb == 1) hipLaunchKernelGGL(fill_1unb, dim3(gb), dim3(256), 0, s, g, geo, A);
    else hipLaunchKernelGGL(fill_0unb, dim3(gb), dim3(256), 0, s, g, geo, A, lo ? lo[0] : 0, lo ? lo[1] : 0, lo ? lo[2] : 0);
    HIP_CHECK(hipGetLastError());
}

void launch_scale(double* g, size_t n, double f, hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, s, g, n, f);
    HIP_CHECK(hipGetLastError());
}

// f32: the data and the Green's function are float arrays (FP32 solves), else double
void launch_mult_std(void* d, const void* g, size_t per, int nf, int lda, double normfact, bool f32, hipStream_t s) {
    if (f32) launch_k<mult_std<float>>(dim3(grid_for(per)), dim3(256), 0, s, {(float*)d, (const float*)g, per, nf, lda, normfact});
    else launch_k<mult_std<double>>(dim3(grid_for(per)), dim3(256), 0, s, {(double*)d, (const double*)g, per, nf, lda, normfact});
    HIP_CHECK(hipGetLastError());
}

void launch_mult_rot(void* d, const void* g, const RotArgs& R, bool f32, hipStream_t s) {
    const size_t per = (size_t)R.n[0] * R.n[1] * R.n[2];
    if (f32) launch_k<mult_rot<float>>(dim3(grid_for(per)), dim3(256), 0, s, {(float*)d, (const float*)g, R});
    else launch_k<mult_rot<double>>(dim3(grid_for(per)), dim3(256), 0, s, {(double*)d, (const double*)g, R});
    HIP_CHECK(hipGetLastError());
}

// b[i] = (float)a[i]: FP32 copies of FP64 Green's functions (rounded once)
void launch_to_float(const double* a, float* b, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(to_float, dim3(grid_for(n)), dim3(256), 0, s, a, b, n);
    HIP_CHECK(hipGetLastError());
}

}  // namespace flups
