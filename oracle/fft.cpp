// Oracle 1-D transforms: mixed-radix Stockham complex FFT of any length and the
// FFTW r2r/r2c/c2r kinds built on it from their textbook definitions
// (FFTW manual 4.8 "What FFTW Really Computes"). Test infrastructure only.
#include <cmath>
#include <map>
#include <memory>
#include <mutex>

#include "oracle.hpp"

namespace oracle {

namespace {

// exp(-2 pi i t / m), t reduced modulo m, evaluated in long double
cplx root(long long t, long long m) {
    t %= m;
    if (t < 0) t += m;
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)m;
    return cplx((double)cosl(a), (double)sinl(a));
}

struct Stage {
    int p, span;                      // radix and product of the earlier radices
    std::vector<cplx> tw;             // tw[k*p + q] = root(q*k, span*p)
    std::vector<cplx> w;              // w[t] = root(t, p)
};

struct CPlan {
    int n;
    std::vector<Stage> st;
    explicit CPlan(int n_) : n(n_) {
        int m = n, span = 1;
        std::vector<int> rad;
        for (int p : {4, 2, 3, 5, 7}) {
            while (m % p == 0) { rad.push_back(p); m /= p; }
        }
        for (int p = 11; m > 1; p += 2) {
            while (m % p == 0) { rad.push_back(p); m /= p; }
        }
        for (int p : rad) {
            Stage s;
            s.p = p;
            s.span = span;
            s.tw.resize((size_t)span * p);
            for (int k = 0; k < span; ++k)
                for (int q = 0; q < p; ++q) s.tw[(size_t)k * p + q] = root((long long)q * k, (long long)span * p);
            s.w.resize(p);
            for (int t = 0; t < p; ++t) s.w[t] = root(t, p);
            st.push_back(std::move(s));
            span *= p;
        }
    }
    // out-of-place Stockham autosort; x is destroyed
    void exec(int sign, cplx* x, cplx* y, cplx* out) const {
        cplx* a = x;
        cplx* b = y;
        std::vector<cplx> v(64), r(64);
        for (const Stage& s : st) {
            const int p = s.p, span = s.span, np = n / p;
            if ((int)v.size() < p) { v.resize(p); r.resize(p); }
            for (int j = 0; j < np; ++j) {
                const int k = j % span;
                for (int q = 0; q < p; ++q) {
                    cplx t = s.tw[(size_t)k * p + q];
                    if (sign > 0) t = std::conj(t);
                    v[q] = a[j + q * np] * t;
                }
                for (int u = 0; u < p; ++u) {
                    cplx acc = 0;
                    for (int q = 0; q < p; ++q) {
                        cplx t = s.w[(q * u) % p];
                        if (sign > 0) t = std::conj(t);
                        acc += v[q] * t;
                    }
                    r[u] = acc;
                }
                const int base = (j / span) * span * p + k;
                for (int u = 0; u < p; ++u) b[base + u * span] = r[u];
            }
            std::swap(a, b);
        }
        for (int i = 0; i < n; ++i) out[i] = a[i];
    }
};

std::mutex g_mtx;
std::map<int, std::shared_ptr<CPlan>> g_plans;

std::shared_ptr<CPlan> get_plan(int n) {
    std::lock_guard<std::mutex> lk(g_mtx);
    auto it = g_plans.find(n);
    if (it != g_plans.end()) return it->second;
    auto p = std::make_shared<CPlan>(n);
    g_plans[n] = p;
    return p;
}

}  // namespace

void fft_c2c(int n, int sign, const cplx* in, cplx* out) {
    if (n == 1) { out[0] = in[0]; return; }
    auto p = get_plan(n);
    std::vector<cplx> x(in, in + n), y(n);
    p->exec(sign, x.data(), y.data(), out);
}

void fft_r2c(int n, const double* in, cplx* out) {
    std::vector<cplx> x(n), y(n);
    for (int i = 0; i < n; ++i) x[i] = in[i];
    fft_c2c(n, SIGN_FWD, x.data(), y.data());
    for (int k = 0; k <= n / 2; ++k) out[k] = y[k];
}

void fft_c2r(int n, const cplx* in, double* out) {
    // Hermitian extension; the imaginary parts of the DC and (even n) Nyquist
    // entries are ignored, as FFTW's c2r does.
    std::vector<cplx> x(n), y(n);
    x[0] = cplx(in[0].real(), 0.0);
    for (int k = 1; k < (n + 1) / 2; ++k) {
        x[k] = in[k];
        x[n - k] = std::conj(in[k]);
    }
    if (n % 2 == 0) x[n / 2] = cplx(in[n / 2].real(), 0.0);
    fft_c2c(n, SIGN_BWD, x.data(), y.data());
    for (int i = 0; i < n; ++i) out[i] = y[i].real();
}

namespace {
// Y[k] = Re( e^{-i pi k/(2n)} FFT_{2n}([x, reverse x])[k] ) = 2 sum x_j cos(pi (j+1/2) k / n)
void dct2(int n, const double* x, double* Y) {
    const int M = 2 * n;
    std::vector<cplx> e(M), E(M);
    for (int j = 0; j < n; ++j) { e[j] = x[j]; e[M - 1 - j] = x[j]; }
    fft_c2c(M, SIGN_FWD, e.data(), E.data());
    for (int k = 0; k < n; ++k) Y[k] = (root(k, 4LL * n) * E[k]).real();
}
// Y[k] = x0 + 2 sum_{j>=1} x_j cos(pi j (k+1/2)/n)
void dct3(int n, const double* x, double* Y) {
    const int M = 2 * n;
    std::vector<cplx> u(M, 0.0), U(M);
    for (int j = 0; j < n; ++j) u[j] = (j == 0 ? 1.0 : 2.0) * x[j] * std::conj(root(j, 4LL * n));
    fft_c2c(M, SIGN_BWD, u.data(), U.data());
    for (int k = 0; k < n; ++k) Y[k] = U[k].real();
}
// Y[k] = 2 sum x_j cos(pi (j+1/2)(k+1/2)/n)
void dct4(int n, const double* x, double* Y) {
    const int M = 2 * n;
    std::vector<cplx> u(M, 0.0), U(M);
    for (int j = 0; j < n; ++j) u[j] = x[j] * root(j, 4LL * n);
    fft_c2c(M, SIGN_FWD, u.data(), U.data());
    for (int k = 0; k < n; ++k) Y[k] = 2.0 * (root(2LL * k + 1, 8LL * n) * U[k]).real();
}
}  // namespace

void fft_r2r(int n, R2R kind, const double* in, double* out) {
    std::vector<double> x(in, in + n), y(n);
    switch (kind) {
        case REDFT00: {  // even extension of period 2(n-1)
            if (n == 1) { y[0] = x[0]; break; }
            const int M = 2 * (n - 1);
            std::vector<cplx> e(M), E(M);
            for (int j = 0; j < n; ++j) e[j] = x[j];
            for (int j = 1; j <= n - 2; ++j) e[M - j] = x[j];
            fft_c2c(M, SIGN_FWD, e.data(), E.data());
            for (int k = 0; k < n; ++k) y[k] = E[k].real();
            break;
        }
        case REDFT10: dct2(n, x.data(), y.data()); break;
        case REDFT01: dct3(n, x.data(), y.data()); break;
        case REDFT11: dct4(n, x.data(), y.data()); break;
        case RODFT00: {  // odd extension of period 2(n+1)
            const int M = 2 * (n + 1);
            std::vector<cplx> e(M, 0.0), E(M);
            for (int j = 0; j < n; ++j) { e[j + 1] = x[j]; e[M - 1 - j] = -x[j]; }
            fft_c2c(M, SIGN_FWD, e.data(), E.data());
            for (int k = 0; k < n; ++k) y[k] = -E[k + 1].imag();
            break;
        }
        case RODFT10: {  // DST-II(x)[k] = DCT-II((-1)^j x_j)[n-1-k]
            std::vector<double> t(n), T(n);
            for (int j = 0; j < n; ++j) t[j] = (j % 2 ? -x[j] : x[j]);
            dct2(n, t.data(), T.data());
            for (int k = 0; k < n; ++k) y[k] = T[n - 1 - k];
            break;
        }
        case RODFT01: {  // DST-III(x)[k] = (-1)^k DCT-III(reverse x)[k]
            std::vector<double> t(n), T(n);
            for (int j = 0; j < n; ++j) t[j] = x[n - 1 - j];
            dct3(n, t.data(), T.data());
            for (int k = 0; k < n; ++k) y[k] = (k % 2 ? -T[k] : T[k]);
            break;
        }
        case RODFT11: {  // DST-IV(x)[k] = (-1)^k DCT-IV(reverse x)[k]
            std::vector<double> t(n), T(n);
            for (int j = 0; j < n; ++j) t[j] = x[n - 1 - j];
            dct4(n, t.data(), T.data());
            for (int k = 0; k < n; ++k) y[k] = (k % 2 ? -T[k] : T[k]);
            break;
        }
    }
    for (int i = 0; i < n; ++i) out[i] = y[i];
}

}  // namespace oracle
