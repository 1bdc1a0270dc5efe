// Oracle Green's functions.
//   drivers: src/green_functions.cpp:27-167 (3 unbounded), 192-391 (2), 407-538 (1), 560-708 (0)
//   kernels: src/green_kernels.hpp (HEJ_0..10, CHAT_2, LGF_2..8, MEHR_4L/6L/4F/6F)
//   tables:  src/green_functions.hpp:30-92 (.ker files, 32^3 / 32^2 raw float64)
//   1-D Mehrstellen lattice Green's functions: src/lgf_1unb2spe.hpp:80-98, 500-611
//   Ji0c: src/ji0.hpp:23-41 (long double series, the reference's summation order)
// Special functions come from libstdc++ (std::cyl_bessel_k, std::expint) and,
// for the sine integral, a series / complex continued fraction written here:
// implementations independent of both the reference and the product.
// Pins beyond the golden error lines (tests/test_oracle.py): HEJ_0 with two unbounded
// directions by the planes its transformed G must have (zero_ = 0 off r = 0 and k = 0, no
// spectral rewrite: src/green_kernels.hpp:430-432, src/Solver.cpp:1089-1101); HEJ_2..10 with
// one unbounded direction by quadrature of their Fourier-space kernels and by the solve's
// convergence order.
// Test infrastructure only.
#include <cmath>
#include <complex>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

double kernel_expansion(const std::string& kdir, const std::string& name, double n1, double n2, double n3);

// sine integral Si(x), x >= 0: power series below 4, above it Si = pi/2 + Im E1(ix)
// with E1 from its continued fraction (Abramowitz & Stegun 5.1.22), modified Lentz
double sine_integral(double x) {
    if (x < 0) return -sine_integral(-x);
    if (x <= 4.0) {
        long double t = x, s = x, x2 = (long double)x * x;
        for (int k = 1; k < 60; ++k) {
            t *= -x2 / ((2.0L * k) * (2.0L * k + 1));
            const long double term = t / (2 * k + 1);
            s += term;
            if (fabsl(term) < 1e-22L * fabsl(s)) break;
        }
        return (double)s;
    }
    using C = std::complex<long double>;
    const C z(0.0L, (long double)x);
    // E1(z) = exp(-z) / (z + 1 - 1/(z + 3 - 4/(z + 5 - 9/(z + 7 - ...))))
    const long double tiny = 1e-300L;
    C b = z + 1.0L, c = 1.0L / tiny, d = 1.0L / b, h = d;
    for (int i = 1; i < 1000; ++i) {
        const long double an = -(long double)i * i;
        b += 2.0L;
        d = 1.0L / (an * d + b);
        c = b + an / c;
        const C del = c * d;
        h *= del;
        if (std::abs(del - 1.0L) < 1e-21L) break;
    }
    const C e1 = std::exp(-z) * h;
    return (double)(1.57079632679489661923132169163975144L + e1.imag());
}

// Ji0c(x) = int_0^x (1 - J0(u)) / u du = -1/2 sum_{n=1}^{50} (-x^2/4)^n / (n!^2 n), summed in
// long double from n = 50 down, with 1/n!^2 from the running product n! (src/ji0.hpp:23-41).
// Near x = 30 the terms reach 1e9 and the sum keeps about 1e-10 relative accuracy: this is
// the reference's value, kept as it is.
double ji0c(double x) {
    static long double c[51];
    static std::once_flag once;
    std::call_once(once, [] {
        long double f = 1.0L;
        c[0] = 1.0L;
        for (int n = 1; n <= 50; ++n) {
            f *= (long double)n;
            c[n] = 1.0L / (f * f);
        }
    });
    const long double t = (long double)(-(0.25 * x * x));
    long double val = 0.0L;
    for (int n = 50; n > 0; --n) val -= std::pow(t, (long double)n) * c[n] / (long double)n * 0.5L;
    return (double)val;
}

// 1-D lattice Green's functions of the Mehrstellen stencils along the unbounded direction,
// transverse wave numbers k1 >= k2 (times h): the root r of a1 r^2 + a0 r + a1 = 0 inside the
// unit circle gives G(n) = r^{|n|+1} / (a1 (r^2 - 1)) ("left": a1 G(n-1) + a0 G(n) + a1 G(n+1)
// = delta_n); "full" applies the right-hand stencil b to it (src/lgf_1unb2spe.hpp:500-611).
namespace {
double r_from_lambda(double lambda) {   // lambda - sqrt(lambda - 1) sqrt(lambda + 1), |r| <= 1
    if (std::isinf(lambda)) return 0.0;
    if (std::fabs(lambda) > 1.0e2) {   // 1/(2l) + 1/(8l^3) + 1/(16l^5) + 5/(128l^7) + 7/(256l^9)
        static const double e[10] = {0, 1. / 2., 0, 1. / 8., 0, 1. / 16., 0, 5. / 128., 0, 7. / 256.};
        const double y = 1.0 / lambda;
        double out = e[9];
        for (int i = 8; i >= 0; --i) out = y * out + e[i];
        return out;
    }
    const std::complex<double> a = std::sqrt(std::complex<double>(lambda - 1.0)), b = std::sqrt(std::complex<double>(lambda + 1.0));
    return std::real(lambda - a * b);
}
}  // namespace

double mehr_1d(int kind, int n, double k1, double k2) {
    // kind: 0 = 4 left, 1 = 6 left, 2 = 4 full, 3 = 6 full
    const bool six = kind == 1 || kind == 3, full = kind >= 2;
    const int an = std::abs(n);
    if (k1 == 0 && k2 == 0) {
        double v = -0.5 * an;
        if (full && !six) v -= (n == 0) * 1. / 12.;
        if (full && six) v -= (an == 1) * (-1. / 240.) + (n == 0) * (11. / 120.);
        return v;
    }
    const double y0 = std::pow(std::sin(k1 / 2), 2), y1 = std::pow(std::sin(k2 / 2), 2);
    double a0, a1;
    if (!six) {
        a0 = -8. / 3. * (y0 * y1 - y0 - y1) + 2.0;
        a1 = 2. / 3. * (y0 + y1) - 1.0;
    } else {
        a0 = -8. / 5. * y0 * y1 + 8. / 3. * (y0 + y1) + 2.0;
        a1 = -8. / 15. * y0 * y1 + 2. / 3. * (y0 + y1) - 1.0;
    }
    const double lambda = -a0 / (2. * a1);
    const double r = r_from_lambda(lambda);
    if (!full) {
        const double base = (r / a1) * std::pow(r, an) / (r * r - 1.0);
        if (std::isfinite(lambda)) return base;
        return six ? -(double)(n == 0) / a0 : (double)(n == 0) / a0;
    }
    if (!six) {
        const double b0 = 5. / 6. - 1. / 3. * (y0 + y1), b1 = 1. / 12.;
        const double pr = b1 + b0 * r + b1 * r * r;
        if (std::isfinite(lambda)) {
            const double base = pr * std::pow(r, an) / (a1 * (r * r - 1));
            return n == 0 ? (r / a1) * (b0 + 2. * b1 * r) / (r * r - 1.) : base;
        }
        return (n == 0) * (b0 / a0) + (n == 1) * (b1 / a0);
    }
    const double b0 = -1. / 15. * (std::pow(y0, 2) + std::pow(y1, 2)) + 8. / 45. * y0 * y1 - 11. / 45. * (y0 + y1) + 97. / 120.;
    const double b1 = -2. / 45. * (y0 + y1) + 1. / 10.;
    const double b2 = -1. / 240.;
    const double pr = b0 * std::pow(r, 2) + b1 * (r + std::pow(r, 3)) + b2 * (1.0 + std::pow(r, 4));
    const double base = pr * std::pow(r, an - 1) / (a1 * (r * r - 1.0));
    return base + (an == 1) * (b2 / a1) + (n == 0) * (b1 * a1 - b2 * a0) / (a1 * a1);
}

namespace {

const double PI = M_PI;
const double EULER = 0.57721566490153286061;

double e1(double x) { return -std::expint(-x); }

int lgf_order(GreenType t) {
    switch (t) {
        case LGF_2: return 2;
        case LGF_4: return 4;
        case LGF_6: return 6;
        case LGF_8: return 8;
        default: return 0;
    }
}

bool is_mehr(GreenType t) { return t == MEHR_4L || t == MEHR_6L || t == MEHR_4F || t == MEHR_6F; }

// kernels given near the origin by a 32^3 / 32^2 table and beyond it by an expansion
// (green_functions.hpp:30-92 for the file names, lgf_3unb0spe.hpp for the expansions)
bool table_kernel(GreenType t, int dim, std::string* file, std::string* exp) {
    const int ord = lgf_order(t);
    char f[64], e[64];
    if (ord) {
        std::snprintf(f, sizeof(f), "LGF_%d_%dd_32.ker", ord, dim);
        std::snprintf(e, sizeof(e), "lgf_%d_%dunb0spe", ord, dim);
    } else if (is_mehr(t)) {
        const char* nm = t == MEHR_4L ? "4L" : t == MEHR_6L ? "6L" : t == MEHR_4F ? "4F" : "6F";
        const char* lo = t == MEHR_4L ? "4l" : t == MEHR_6L ? "6l" : t == MEHR_4F ? "4f" : "6f";
        const bool lump = (t == MEHR_4L || t == MEHR_6L) && dim == 2;   // 4L and 6L share the 2-D table
        std::snprintf(f, sizeof(f), "MEHR_%s_%dd_32.ker", lump ? "4L6L" : nm, dim);
        std::snprintf(e, sizeof(e), "mehr_%s_%dunb0spe", lump ? "4l6l" : lo, dim);
    } else {
        return false;
    }
    *file = f;
    *exp = e;
    return true;
}

std::mutex g_tab_mtx;
std::map<std::string, std::vector<double>> g_tabs;

const std::vector<double>& kernel_table(const std::string& kdir, const std::string& file, int dim) {
    const std::string name = kdir + "/" + file;
    std::lock_guard<std::mutex> lk(g_tab_mtx);
    auto it = g_tabs.find(name);
    if (it != g_tabs.end()) return it->second;
    const size_t n = (dim == 3) ? 32 * 32 * 32 : 32 * 32;
    std::vector<double> v(n);
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) throw std::runtime_error("unable to read " + name);
    if (std::fread(v.data(), sizeof(double), n, f) != n) { std::fclose(f); throw std::runtime_error("short .ker file"); }
    std::fclose(f);
    return g_tabs[name] = v;
}

// symmetrised index of green_functions (Topology.hpp:415-430): mirrored indices are negative
inline int symid(int i, double ss) {
    if (ss == 0.0 || i <= ss) return i;
    return -std::max((int)std::fabs(2.0 * ss - i), 1);
}

inline double& gat(const Topo& t, double* g, int i0, int i1, int i2) {
    return g[((size_t)i0 + (size_t)t.n[0] * ((size_t)i1 + (size_t)t.n[1] * i2)) * t.nf];
}

// polynomial parts of the HEJ kernels, green_kernels.hpp:26-54, 340-415, 461-566, 681-717
double hej3_poly(GreenType t, double rho) {
    const double r2 = rho * rho;
    switch (t) {
        case HEJ_2: return 0.0;
        case HEJ_4: return rho;
        case HEJ_6: return rho * (7.0 / 4.0 - r2 / 4.0);
        case HEJ_8: return rho * (19.0 / 8.0 - (2.0 / 3.0) * r2 + r2 * r2 / 24.0);
        case HEJ_10: return rho * (187.0 / 64.0 - (233.0 / 192.0) * r2 + (29.0 / 192.0) * r2 * r2 - r2 * r2 * r2 / 192.0);
        default: throw std::runtime_error("not a HEJ kernel");
    }
}
double hej3_g0(GreenType t, double eps) {
    const double base = -M_SQRT2 / (eps * std::sqrt(PI * PI * PI));
    switch (t) {
        case HEJ_2: return base / 4.0;
        case HEJ_4: return base * 3.0 / 8.0;
        case HEJ_6: return base * 15.0 / 32.0;
        case HEJ_8: return base * 35.0 / 64.0;
        case HEJ_10: return base * 315.0 / 512.0;
        default: throw std::runtime_error("not a HEJ kernel");
    }
}
// k=0 (2 unbounded) exponential coefficient Q(rho^2) and the r=0 constant
double hej2_q(GreenType t, double r2) {
    switch (t) {
        case HEJ_2: return 0.0;
        case HEJ_4: return 0.5;
        case HEJ_6: return 0.75 - r2 / 8.0;
        case HEJ_8: return 11.0 / 12.0 - 7.0 * r2 / 24.0 + r2 * r2 / 48.0;
        case HEJ_10: return 25.0 / 24.0 - 23.0 * r2 / 48.0 + 13.0 * r2 * r2 / 192.0 - r2 * r2 * r2 / 384.0;
        default: throw std::runtime_error("not a HEJ kernel");
    }
}
double hej2_q0(GreenType t) { return hej2_q(t, 0.0); }

double hej_spectral(GreenType t, double ksq, double sig) {
    const double s2 = ksq * sig * sig, h = 0.5 * s2;
    double p = 1.0, term = 1.0;
    int jmax = 0;
    switch (t) {
        case HEJ_2: jmax = 0; break;
        case HEJ_4: jmax = 1; break;
        case HEJ_6: jmax = 2; break;
        case HEJ_8: jmax = 3; break;
        case HEJ_10: jmax = 4; break;
        default: throw std::runtime_error("not a HEJ kernel");
    }
    for (int j = 1; j <= jmax; ++j) { term *= h / j; p += term; }
    return -p * std::exp(-h) / ksq;
}

// Mehrstellen symbols in y = sin^2(k h / 2) (green_kernels.hpp:744-773)
double mehr_spectral(GreenType t, const double k[3], double h) {
    double y[3];
    for (int d = 0; d < 3; ++d) {
        const double sv = std::sin(0.5 * k[d] * h);
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

double lgf_spectral(int order, const double k[3], double h) {
    double s = 0;
    for (int d = 0; d < 3; ++d) s += lgf_symbol(order, k[d] * h);
    return -h * h / s;
}

}  // namespace

void green_fill_3unb(const Topo& t, const double hfact[3], const double symstart[3], double* g, GreenType type,
                     double length, const std::string& kdir) {
    for (int d = 0; d < 3; ++d)
        if (hfact[d] == 0.0) throw std::runtime_error("hfact cannot be 0");
    std::string file, expn;
    const bool ord = table_kernel(type, 3, &file, &expn);
    const std::vector<double>* tab = nullptr;
    if (ord) {
        if (!(hfact[0] == hfact[1] && hfact[1] == hfact[2])) throw std::runtime_error("LGF needs an isotropic grid");
        tab = &kernel_table(kdir, file, 3);
    } else if (!(type == CHAT_2 || (type >= HEJ_2 && type <= HEJ_0))) {
        throw std::runtime_error("kernel not restated in the oracle");
    }
    for (int i2 = 0; i2 < t.n[2]; ++i2)
        for (int i1 = 0; i1 < t.n[1]; ++i1)
            for (int i0 = 0; i0 < t.n[0]; ++i0) {
                const int is[3] = {symid(i0, symstart[0]), symid(i1, symstart[1]), symid(i2, symstart[2])};
                const double x0 = is[0] * hfact[0], x1 = is[1] * hfact[1], x2 = is[2] * hfact[2];
                const double r = std::sqrt(x0 * x0 + x1 * x1 + x2 * x2);
                double v;
                if (ord) {
                    const int ax0 = t.axis, ax1 = (ax0 + 1) % 3, ax2 = (ax0 + 2) % 3;
                    const int a = std::abs(is[ax0]), b = std::abs(is[ax1]), c = std::abs(is[ax2]);
                    double gv;
                    if (a < 32 && b < 32 && c < 32) gv = (*tab)[a + 32 * b + 32 * 32 * c];
                    else gv = kernel_expansion(kdir, expn, a, b, c);
                    v = -gv / hfact[0];
                } else if (type == CHAT_2) {
                    v = -1.0 / (4.0 * PI * r);
                } else if (type == HEJ_0) {   // green_kernels.hpp:55-61, length = h / pi
                    const double rho = r / length;
                    v = -sine_integral(rho) / (rho * 2.0 * PI * PI * length);
                } else {
                    const double rho = r / length;
                    v = -1.0 / (4.0 * PI * r) *
                        (std::erf(rho / M_SQRT2) + hej3_poly(type, rho) * std::exp(-0.5 * rho * rho) / std::sqrt(2.0 * PI));
                }
                gat(t, g, i0, i1, i2) = v;
            }
    if (!ord) {
        double g0;
        if (type == CHAT_2) {
            const double vol = hfact[0] * hfact[1] * hfact[2];
            g0 = -0.5 * std::pow(1.5 / (2.0 * PI) * vol, 2.0 / 3.0) / vol;
        } else if (type == HEJ_0) {
            g0 = -1.0 / (2.0 * PI * PI * length);   // green_functions.cpp:66-68
        } else {
            g0 = hej3_g0(type, length);
        }
        g[0] = g0;
    }
}

void green_fill_2unb(const Topo& t, const double hfact[3], const double kfact[3], const double koffset[3],
                     const double symstart[3], double* g, GreenType type, double length, const std::string& kdir) {
    const bool iso = (hfact[0] == 0.0 && hfact[1] == hfact[2]) || (hfact[1] == 0.0 && hfact[0] == hfact[2]) ||
                     (hfact[2] == 0.0 && hfact[0] == hfact[1]);
    std::string file, expn;
    const bool ord = table_kernel(type, 2, &file, &expn);
    const std::vector<double>* tab = nullptr;
    if (ord) {
        if (!iso) throw std::runtime_error("LGF needs an isotropic grid");
        tab = &kernel_table(kdir, file, 2);
    } else if (!(type == CHAT_2 || (type >= HEJ_2 && type <= HEJ_0))) {
        throw std::runtime_error("kernel not restated in the oracle");
    }
    const double sh = hfact[0] + hfact[1] + hfact[2];
    const double sk = kfact[0] + kfact[1] + kfact[2];
    const double req = std::sqrt(hfact[0] * hfact[1] + hfact[1] * hfact[2] + hfact[2] * hfact[0]) / std::sqrt(PI);
    // the 2-D table lookup takes the first two nonzero |indices| in the Green topology's axis order
    // (green_kernels.hpp:134-151)
    for (int i2 = 0; i2 < t.n[2]; ++i2)
        for (int i1 = 0; i1 < t.n[1]; ++i1)
            for (int i0 = 0; i0 < t.n[0]; ++i0) {
                const int is[3] = {symid(i0, symstart[0]), symid(i1, symstart[1]), symid(i2, symstart[2])};
                const double k = (is[0] + koffset[0]) * kfact[0] + (is[1] + koffset[1]) * kfact[1] + (is[2] + koffset[2]) * kfact[2];
                const double x0 = is[0] * hfact[0], x1 = is[1] * hfact[1], x2 = is[2] * hfact[2];
                const double r = std::sqrt(x0 * x0 + x1 * x1 + x2 * x2);
                double v;
                const bool r0 = r <= sh * 0.2;
                const bool k0 = !r0 && k <= sk * 0.2;
                if (ord) {
                    if (r0 || k0) {
                        const int ax0 = t.axis, ax1 = (ax0 + 1) % 3, ax2 = (ax0 + 2) % 3;
                        const int a[3] = {std::abs(is[ax0]), std::abs(is[ax1]), std::abs(is[ax2])};
                        const int j1 = (a[0] == 0) ? a[1] : a[0];
                        const int j2 = (a[0] == 0) ? a[2] : (a[1] == 0 ? a[2] : a[1]);
                        double gv;
                        if (j1 < 32 && j2 < 32) gv = (*tab)[j1 + 32 * j2];
                        else gv = kernel_expansion(kdir, expn, j1, j2, 0.0);
                        v = -gv;
                    } else {
                        v = 0.0;
                    }
                } else if (type == CHAT_2) {
                    if (r0) {
                        const double kr = k * req;
                        v = -(1.0 - kr * std::cyl_bessel_k(1.0, kr)) / (PI * kr * kr);
                    } else if (k0) {
                        v = std::log(r) / (2.0 * PI);
                    } else {
                        v = -std::cyl_bessel_k(0.0, std::fabs(k) * r) / (2.0 * PI);
                    }
                } else if (type == HEJ_0) {
                    // green_kernels.hpp:416-428: the same expression at r = 0 and on k = 0; elsewhere
                    // zero_ = 0 (green_kernels.hpp:430-432), which no spectral rewrite replaces for
                    // HEJ_0 (Solver.cpp:1089-1101)
                    if (r0 || k0) {
                        const double rho = r / length;
                        v = rho < 30.0 ? (ji0c(rho) + std::log(2 * length) - EULER) / (2.0 * PI) : std::log(r) / (2.0 * PI);
                    } else {
                        v = 0.0;
                    }
                } else {
                    if (r0) {
                        v = -(0.5 * EULER - std::log(M_SQRT2 * length) + hej2_q0(type)) / (2.0 * PI);
                    } else if (k0) {
                        const double rho2 = (r / length) * (r / length);
                        v = (std::log(r) - hej2_q(type, rho2) * std::exp(-0.5 * rho2) + 0.5 * e1(0.5 * rho2)) / (2.0 * PI);
                    } else {
                        v = 0.0;   // zero_ (green_kernels.hpp:430-432); the spectral rewrite replaces these planes
                    }
                }
                gat(t, g, i0, i1, i2) = v;
            }
    if (type == CHAT_2) g[0] = 0.25 / (2.0 * PI) * (PI - 6.0 + 2.0 * std::log(0.5 * PI * req * req));
}

// HEJ_2..10 with one unbounded direction (green_kernels.hpp:461-566): G(k, r) for the
// transverse wave number k > 0, or the k = 0 form (k0)
double hej_1unb(GreenType type, double r, double k, double sig, bool k0) {
    const double rho = r / sig, rho2 = rho * rho;
    const double c = 1.0 / (M_SQRT2 * std::sqrt(PI));
    if (k0) {
        const double er = 0.5 * r * std::erf(r / (M_SQRT2 * sig));
        const double ex = std::exp(-0.5 * rho2);
        double tail;
        switch (type) {
            case HEJ_2: tail = (1.0 - ex); break;
            case HEJ_4: tail = 0.5 * (1.0 - ex); break;
            case HEJ_6: tail = 0.125 * (3.0 - ex * (rho2 + 3.0)); break;
            case HEJ_8: tail = 5.0 / 16.0 - ex * (-rho2 * rho2 / 48.0 + 0.25 * rho2 + 5.0 / 16.0); break;
            default:
                tail = 35.0 / 128.0 - ex * (rho2 * rho2 * rho2 / 384.0 - 23.0 / 384.0 * rho2 * rho2 + 47.0 / 128.0 * rho2 +
                                            35.0 / 128.0);
        }
        return er - tail * sig * c;
    }
    const double s = k * sig, s2 = s * s;
    const double sub = (s * rho > 100.0) ? 0.0
                                         : ((1.0 - std::erf((s - rho) / M_SQRT2)) * std::exp(-s * rho) +
                                            (1.0 - std::erf((s + rho) / M_SQRT2)) * std::exp(s * rho));
    double poly;
    switch (type) {
        case HEJ_2: poly = 0.0; break;
        case HEJ_4: poly = 0.25; break;
        case HEJ_6: poly = 5.0 / 16.0 + (s2 - rho2) / 16.0; break;
        case HEJ_8: poly = 11.0 / 32.0 + s2 / 12.0 - 0.125 * rho2 - s2 * rho2 / 48.0 + (s2 * s2 + rho2 * rho2) / 96.0; break;
        default:
            poly = 93.0 / 256.0 + 73.0 / 768.0 * s2 - 47.0 / 256.0 * rho2 - 17.0 / 384.0 * s2 * rho2 + 11.0 / 768.0 * s2 * s2 +
                   23.0 / 768.0 * rho2 * rho2 + (s2 * rho2 * rho2 - s2 * s2 * rho2) / 256.0 +
                   (s2 * s2 * s2 - rho2 * rho2 * rho2) / 768.0;
    }
    return -0.25 * sig / s * sub - sig * M_SQRT2 / std::sqrt(PI) * poly * std::exp(-0.5 * (s2 + rho2));
}

void green_fill_1unb(const Topo& t, const double hfact[3], const double kfact[3], const double koffset[3],
                     const double symstart[3], double* g, GreenType type, double length) {
    const int ord = lgf_order(type);
    const bool mehr = type == MEHR_4L || type == MEHR_6L || type == MEHR_4F || type == MEHR_6F;
    if (!ord && !mehr && !(type == CHAT_2 || (type >= HEJ_2 && type <= HEJ_10)))
        throw std::runtime_error("kernel not restated in the oracle");
    const double sk = kfact[0] + kfact[1] + kfact[2];
    const double h = hfact[0] + hfact[1] + hfact[2];
    for (int i2 = 0; i2 < t.n[2]; ++i2)
        for (int i1 = 0; i1 < t.n[1]; ++i1)
            for (int i0 = 0; i0 < t.n[0]; ++i0) {
                const int is[3] = {symid(i0, symstart[0]), symid(i1, symstart[1]), symid(i2, symstart[2])};
                const double kk[3] = {(is[0] + koffset[0]) * kfact[0], (is[1] + koffset[1]) * kfact[1],
                                      (is[2] + koffset[2]) * kfact[2]};
                const double k = std::sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2]);
                const double x0 = is[0] * hfact[0], x1 = is[1] * hfact[1], x2 = is[2] * hfact[2];
                const double r = std::sqrt(x0 * x0 + x1 * x1 + x2 * x2);
                const bool k0 = k <= sk * 0.2;
                double v;
                if (ord) {
                    double c = 0;
                    for (int d = 0; d < 3; ++d) c += lgf_symbol(ord, kk[d] * h);
                    const int n = (int)std::round(std::fabs(r) / h);
                    v = -h * lgf1d(ord, n, c);
                } else if (mehr) {
                    // green_kernels.hpp:624-671: one k is zero; |k1| >= |k2| >= 0
                    const double a0 = std::fabs(kk[0]), a1 = std::fabs(kk[1]), a2 = std::fabs(kk[2]);
                    const double k1 = std::max(a0, std::max(a1, a2));
                    const double k2 = a0 + a1 + a2 - k1;
                    const int n = (int)std::round(std::fabs(r) / h);
                    const int kind = type == MEHR_4L ? 0 : type == MEHR_6L ? 1 : type == MEHR_4F ? 2 : 3;
                    v = -h * mehr_1d(kind, n, k1 * h, k2 * h);
                } else if (type == CHAT_2) {
                    v = k0 ? 0.5 * std::fabs(r) : -0.5 * std::exp(-k * r) / k;
                } else {
                    v = hej_1unb(type, r, k, length, k0);
                }
                gat(t, g, i0, i1, i2) = v;
            }
}

void green_fill_0unb(const Topo& t, double hgrid, const double kfact[3], const double koffset[3], const double symstart[3],
                     double* g, GreenType type, double length, const int* istart_custom) {
    if (kfact[0] == 0.0 || kfact[1] == 0.0) throw std::runtime_error("dk cannot be 0");
    const int ord = lgf_order(type);
    if (!ord && !is_mehr(type) && !(type == CHAT_2 || type == HEJ_0 || (type >= HEJ_2 && type <= HEJ_10)))
        throw std::runtime_error("kernel not restated in the oracle");
    int lo[3] = {0, 0, 0};
    if (istart_custom)
        for (int d = 0; d < 3; ++d) lo[d] = std::min(std::max(istart_custom[d], 0), t.n[d]);
    for (int i2 = lo[2]; i2 < t.n[2]; ++i2)
        for (int i1 = lo[1]; i1 < t.n[1]; ++i1)
            for (int i0 = lo[0]; i0 < t.n[0]; ++i0) {
                const int il[3] = {symid(i0, symstart[0]), symid(i1, symstart[1]), symid(i2, symstart[2])};
                const double k[3] = {(il[0] + koffset[0]) * kfact[0], (il[1] + koffset[1]) * kfact[1],
                                     (il[2] + koffset[2]) * kfact[2]};
                const double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                double v;
                if (ord) v = lgf_spectral(ord, k, hgrid);
                else if (is_mehr(type)) v = mehr_spectral(type, k, hgrid);
                else if (type == CHAT_2 || type == HEJ_0) v = -1.0 / ksq;
                else v = hej_spectral(type, ksq, length);
                gat(t, g, i0, i1, i2) = v;
            }
    if (koffset[0] + koffset[1] + koffset[2] < 0.2) g[0] = 0.0;
}

}  // namespace oracle
