// Oracle restatement of the per-direction plan rules.
//   constructor: src/FFTW_plan_dim.cpp:77-161 (type from bc sum, volfact, kfact)
//   cell init:   src/FFTW_plan_dim_cell.cpp:19-391
//   node init:   src/FFTW_plan_dim_node.cpp:19-462
// Test infrastructure only.
#include <cmath>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

static const double TWO_PI = 2.0 * M_PI;

Plan::Plan(int lda_, int dim_, const double h[3], const double L[3], const int* bl, const int* br, int sign_, bool green_,
           bool node_)
    : lda(lda_), dim(dim_), sign(sign_), green(green_), node(node_) {
    bcl.assign(bl, bl + lda);
    bcr.assign(br, br + lda);
    const int t = bcl[0] + bcr[0];
    if (t <= SYMSYM) {
        type = SYMSYM; volfact = 1.0; kfact = TWO_PI / (2.0 * L[dim]); spectral = green;
    } else if (t <= MIXUNB) {
        type = MIXUNB; volfact = h[dim]; kfact = TWO_PI / (4.0 * L[dim]); spectral = false;
    } else if (t == PERPER) {
        type = PERPER; volfact = 1.0; kfact = TWO_PI / L[dim]; spectral = green;
    } else if (t == UNBUNB) {
        type = UNBUNB; volfact = h[dim]; kfact = TWO_PI / (2.0 * L[dim]); spectral = false;
    } else if (t == EMPTY) {
        type = EMPTY; volfact = 1.0; kfact = 0.0; spectral = false;
    } else {
        throw std::runtime_error("invalid combination of boundary conditions");
    }
    n_in.assign(lda, 1);
    start_in.assign(lda, 0);
    start_out.assign(lda, 0);
    postpro.assign(lda, PP_NONE);
    kind.assign(lda, REDFT00);
    imult.assign(lda, false);
    for (int c = 0; c < lda; ++c)
        if (bcl[c] + bcr[c] > type) throw std::runtime_error("component bc not compatible with the plan");
}

void Plan::init(const int size[3], bool isComplex) {
    const int N = size[dim];
    auto set = [&](int c, int nin, int sin, int sout, int pp, int k, bool im) {
        n_in[c] = nin; start_in[c] = sin; start_out[c] = sout; postpro[c] = pp; kind[c] = k; imult[c] = im;
    };
    const bool fwd = (sign == SIGN_FWD);
    if (type == SYMSYM) {
        if (isComplex) throw std::runtime_error("r2r on complex data");
        const bool type4 = bcl[0] != bcr[0];
        for (int c = 1; c < lda; ++c)
            if ((bcl[c] != bcr[c]) != type4) throw std::runtime_error("mixed EVEN-ODD / symmetric components");
        symstart = 0; fieldstart = 0; r2c = false;
        normfact *= node ? 1.0 / (2.0 * (N - 1)) : 1.0 / (2.0 * N);
        for (int c = 0; c < lda; ++c) {
            const int l = bcl[c], r = bcr[c];
            if (green) {
                // spectral Green: only sizes and the half-mode shift matter
                if (!node) {
                    n_in[c] = N; n_out = (l != r) ? N : N + 1; koffset = (l != r) ? 0.5 : 0.0;
                } else {
                    if (l == EVEN && r == EVEN) { n_in[c] = N; koffset = 0.0; }
                    else if (l == ODD && r == ODD) { n_in[c] = N - 2; koffset = 0.0; }
                    else { n_in[c] = N - 1; koffset = 0.5; }
                    n_out = N;
                }
                return;  // the reference leaves the loop after the first component
            }
            if (!node) {
                if (l == EVEN && r == EVEN) {
                    n_out = N + 1; koffset = 0.0;
                    if (fwd) set(c, N, 0, 0, NULL_LAST, REDFT10, false);
                    else set(c, N, 0, 0, PP_NONE, REDFT01, false);
                } else if (l == EVEN && r == ODD) {
                    n_out = N; koffset = 0.5;
                    set(c, N, 0, 0, PP_NONE, REDFT11, false);
                } else if (l == ODD && r == ODD) {
                    n_out = N + 1; koffset = 0.0;
                    if (fwd) set(c, N, 0, 1, NULL_FIRST, RODFT10, true);
                    else set(c, N, 1, 0, NULL_LAST, RODFT01, true);
                } else if (l == ODD && r == EVEN) {
                    n_out = N; koffset = 0.5;
                    set(c, N, 0, 0, PP_NONE, RODFT11, true);
                }
            } else {
                if (l == EVEN && r == EVEN) {
                    n_out = N; koffset = 0.0;
                    set(c, N, 0, 0, PP_NONE, REDFT00, false);
                } else if (l == EVEN && r == ODD) {
                    n_out = N; koffset = 0.5;
                    if (fwd) set(c, N - 1, 0, 0, NULL_LAST, REDFT01, false);
                    else set(c, N - 1, 0, 0, NULL_LAST, REDFT10, false);
                } else if (l == ODD && r == ODD) {
                    n_out = N; koffset = 0.0; fieldstart = 0;
                    set(c, N - 2, 1, 1, NULL_FIRST + NULL_LAST, RODFT00, true);
                } else if (l == ODD && r == EVEN) {
                    n_out = N; koffset = 0.5; fieldstart = 0;
                    if (fwd) set(c, N - 1, 1, 0, NULL_LAST, RODFT01, true);
                    else set(c, N - 1, 0, 1, NULL_FIRST, RODFT10, true);
                }
            }
        }
    } else if (type == MIXUNB) {
        if (isComplex) throw std::runtime_error("r2r on complex data");
        r2c = false;
        if (green) fieldstart = 0;
        else if (bcl[0] == UNB) fieldstart = node ? N - 1 : N;
        else fieldstart = 0;
        symstart = 0;
        normfact *= node ? 1.0 / (2.0 * ((2 * N - 1) - 1.0)) : 1.0 / (4.0 * N);
        for (int c = 0; c < lda; ++c) {
            const int l = bcl[c], r = bcr[c];
            const bool even = (l == EVEN && r == UNB) || (l == UNB && r == EVEN);
            const bool odd = (l == ODD && r == UNB) || (l == UNB && r == ODD);
            koffset = 0.0;
            if (green) {
                const int n = node ? 2 * N - 1 : 2 * N + 1;
                n_out = n;
                set(c, n, 0, 0, PP_NONE, REDFT00, false);
            } else if (!node) {
                n_out = 2 * N + 1;
                if (even) {
                    if (fwd) set(c, 2 * N, 0, 0, NULL_LAST, REDFT10, false);
                    else set(c, 2 * N, 0, 0, PP_NONE, REDFT01, false);
                } else if (odd) {
                    if (fwd) set(c, 2 * N, 0, 1, NULL_FIRST, RODFT10, true);
                    else set(c, 2 * N, 1, 0, NULL_LAST, RODFT01, true);
                } else {
                    throw std::runtime_error("unable to init the mixed plan");
                }
            } else {
                if (even) {
                    n_out = 2 * N - 1;
                    set(c, 2 * N - 1, 0, 0, PP_NONE, REDFT00, false);
                } else if (odd) {
                    n_out = 2 * N - 1;
                    set(c, 2 * N - 3, 1, 1, NULL_FIRST + NULL_LAST, RODFT00, true);
                } else {
                    throw std::runtime_error("unable to init the mixed plan");
                }
            }
        }
    } else if (type == PERPER) {
        if (!node) {
            n_in[0] = N;
            if (isComplex) { n_out = N; r2c = false; }
            else { n_out = N / 2 + 1; r2c = true; }
            symstart = isComplex ? N / 2.0 : 0.0;
            normfact *= 1.0 / N;
            for (int c = 0; c < lda; ++c) set(c, n_in[0], 0, 0, PP_NONE, 0, false);
        } else {
            for (int c = 0; c < lda; ++c) n_in[c] = N - 1;
            if (isComplex) { n_out = N; r2c = false; }
            else { n_out = n_in[0] / 2 + 2; r2c = true; }
            symstart = isComplex ? (N - 1) / 2.0 : 0.0;
            normfact *= 1.0 / (N - 1);
            for (int c = 0; c < lda; ++c) set(c, N - 1, 0, 0, ENFORCE_PERIOD, 0, false);
        }
        fieldstart = 0;
    } else if (type == UNBUNB) {
        const int M = node ? 2 * (N - 1) : 2 * N;
        n_in[0] = M;
        if (isComplex) { n_out = M; r2c = false; }
        else { n_out = M / 2 + 1; r2c = true; }
        fieldstart = 0;
        symstart = node ? N - 1 : N;
        normfact *= 1.0 / M;
        for (int c = 0; c < lda; ++c) set(c, M, 0, 0, PP_NONE, 0, false);
    }
    // EMPTY: nothing to do
}

}  // namespace oracle
