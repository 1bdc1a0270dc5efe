// Per-direction plan rules of the FLUPS solve path, written as rule tables.
//
// What each boundary pair means for the transform (kind, lengths, offsets,
// post-corrections, mode shift and scale factors) follows the reference:
//   type / volfact / kfact       src/FFTW_plan_dim.cpp:77-161
//   cell-centred data             src/FFTW_plan_dim_cell.cpp:19-391
//   node-centred data             src/FFTW_plan_dim_node.cpp:19-462
// Lengths are expressed in terms of N = the extent of the direction when the
// plan is reached (size[dim] of the dry run, src/Solver.cpp:629-717).
#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "core.hpp"

namespace flups {

void fail(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error{buf};
}

PType plan_type(int l, int r) {
    const int t = l + r;
    if (t <= SYMSYM) return SYMSYM;
    if (t <= MIXUNB) return MIXUNB;
    if (t == PERPER) return PERPER;
    if (t == UNBUNB) return UNBUNB;
    if (t == EMPTY) return EMPTY;
    fail("invalid combination of boundary conditions %d %d", l, r);
}

namespace {

// one row of a rule table: n_in = a*N + b, start in/out, post, kind, imult
struct Rule {
    int a, b, si, so, post, kind, imult;
};

// symmetric directions, index [EE, EO, OE, OO][fwd, bwd]
const Rule kCellSym[4][2] = {
    {{1, 0, 0, 0, POST_LAST, REDFT10, 0}, {1, 0, 0, 0, POST_NONE, REDFT01, 0}},                     // EVEN-EVEN
    {{1, 0, 0, 0, POST_NONE, REDFT11, 0}, {1, 0, 0, 0, POST_NONE, REDFT11, 0}},                     // EVEN-ODD
    {{1, 0, 0, 0, POST_NONE, RODFT11, 1}, {1, 0, 0, 0, POST_NONE, RODFT11, 1}},                     // ODD-EVEN
    {{1, 0, 0, 1, POST_FIRST, RODFT10, 1}, {1, 0, 1, 0, POST_LAST, RODFT01, 1}},                    // ODD-ODD
};
const Rule kNodeSym[4][2] = {
    {{1, 0, 0, 0, POST_NONE, REDFT00, 0}, {1, 0, 0, 0, POST_NONE, REDFT00, 0}},
    {{1, -1, 0, 0, POST_LAST, REDFT01, 0}, {1, -1, 0, 0, POST_LAST, REDFT10, 0}},
    {{1, -1, 1, 0, POST_LAST, RODFT01, 1}, {1, -1, 0, 1, POST_FIRST, RODFT10, 1}},
    {{1, -2, 1, 1, POST_FIRST | POST_LAST, RODFT00, 1}, {1, -2, 1, 1, POST_FIRST | POST_LAST, RODFT00, 1}},
};
// semi-unbounded directions, index [even side, odd side][fwd, bwd]
const Rule kCellMix[2][2] = {
    {{2, 0, 0, 0, POST_LAST, REDFT10, 0}, {2, 0, 0, 0, POST_NONE, REDFT01, 0}},
    {{2, 0, 0, 1, POST_FIRST, RODFT10, 1}, {2, 0, 1, 0, POST_LAST, RODFT01, 1}},
};
const Rule kNodeMix[2][2] = {
    {{2, -1, 0, 0, POST_NONE, REDFT00, 0}, {2, -1, 0, 0, POST_NONE, REDFT00, 0}},
    {{2, -3, 1, 1, POST_FIRST | POST_LAST, RODFT00, 1}, {2, -3, 1, 1, POST_FIRST | POST_LAST, RODFT00, 1}},
};

int sym_index(int l, int r) { return (l == EVEN ? 0 : 2) + (r == EVEN ? 0 : 1); }

}  // namespace

DirPlan make_dirplan(int dim, int lda, const int* bcl, const int* bcr, const double h[3], const double L[3], int sign,
                     bool green, bool node) {
    DirPlan p;
    p.dim = dim;
    p.lda = lda;
    p.sign = sign;
    p.green = green;
    p.node = node;
    for (int c = 0; c < lda; ++c) { p.bcl[c] = bcl[c]; p.bcr[c] = bcr[c]; }
    p.type = plan_type(bcl[0], bcr[0]);
    const double twopi = 2.0 * M_PI;
    switch (p.type) {
        case SYMSYM: p.volfact = 1.0; p.kfact = twopi / (2.0 * L[dim]); p.spectral = green; break;
        case MIXUNB: p.volfact = h[dim]; p.kfact = twopi / (4.0 * L[dim]); break;
        case PERPER: p.volfact = 1.0; p.kfact = twopi / L[dim]; p.spectral = green; break;
        case UNBUNB: p.volfact = h[dim]; p.kfact = twopi / (2.0 * L[dim]); break;
        case EMPTY: p.volfact = 1.0; p.kfact = 0.0; break;
    }
    for (int c = 0; c < lda; ++c)
        FLUPS_REQUIRE(bcl[c] + bcr[c] <= p.type, "component %d boundary %d %d not compatible with plan type %d", c, bcl[c],
                      bcr[c], (int)p.type);
    return p;
}

void init_dirplan(DirPlan& p, const int size[3], bool isComplex) {
    const int N = size[p.dim];
    const bool fwd = (p.sign == FWD);
    auto apply = [&](int c, const Rule& r, int base) {
        p.n_in[c] = r.a * base + r.b;
        p.start_in[c] = r.si;
        p.start_out[c] = r.so;
        p.post[c] = r.post;
        p.kind[c] = r.kind;
        p.imult[c] = r.imult != 0;
    };
    switch (p.type) {
        case SYMSYM: {
            FLUPS_REQUIRE(!isComplex, "a symmetric direction cannot act on complex data");
            const bool type4 = p.bcl[0] != p.bcr[0];
            for (int c = 1; c < p.lda; ++c)
                FLUPS_REQUIRE((p.bcl[c] != p.bcr[c]) == type4,
                              "one component has an EVEN-ODD condition while another uses EVEN-EVEN or ODD-ODD");
            p.normfact *= p.node ? 1.0 / (2.0 * (N - 1)) : 1.0 / (2.0 * N);
            p.koffset = type4 ? 0.5 : 0.0;
            if (p.green) {   // spectral Green's function: extents only
                if (p.node) {
                    p.n_out = N;
                    p.n_in[0] = (p.bcl[0] == EVEN && p.bcr[0] == EVEN) ? N : (p.bcl[0] == ODD && p.bcr[0] == ODD) ? N - 2 : N - 1;
                } else {
                    p.n_out = type4 ? N : N + 1;
                    p.n_in[0] = N;
                }
                return;
            }
            p.n_out = p.node ? N : (type4 ? N : N + 1);
            for (int c = 0; c < p.lda; ++c)
                apply(c, (p.node ? kNodeSym : kCellSym)[sym_index(p.bcl[c], p.bcr[c])][fwd ? 0 : 1], N);
            return;
        }
        case MIXUNB: {
            FLUPS_REQUIRE(!isComplex, "a semi-unbounded direction cannot act on complex data");
            p.fieldstart = p.green ? 0 : (p.bcl[0] == UNB ? (p.node ? N - 1 : N) : 0);
            p.normfact *= p.node ? 1.0 / (2.0 * ((2 * N - 1) - 1.0)) : 1.0 / (4.0 * N);
            p.koffset = 0.0;
            if (p.green) {   // Green's function is always even: DCT-I over the doubled domain
                const int n = p.node ? 2 * N - 1 : 2 * N + 1;
                p.n_out = n;
                p.n_in[0] = n;
                p.kind[0] = REDFT00;
                return;
            }
            p.n_out = p.node ? 2 * N - 1 : 2 * N + 1;
            for (int c = 0; c < p.lda; ++c) {
                const int l = p.bcl[c], r = p.bcr[c];
                const bool even = (l == EVEN && r == UNB) || (l == UNB && r == EVEN);
                const bool odd = (l == ODD && r == UNB) || (l == UNB && r == ODD);
                FLUPS_REQUIRE(even || odd, "unsupported semi-unbounded boundary %d %d", l, r);
                apply(c, (p.node ? kNodeMix : kCellMix)[even ? 0 : 1][fwd ? 0 : 1], N);
            }
            return;
        }
        case PERPER: {
            const int n = p.node ? N - 1 : N;
            p.r2c = !isComplex;
            p.n_out = isComplex ? (p.node ? N : N) : (p.node ? n / 2 + 2 : n / 2 + 1);
            p.symstart = isComplex ? n / 2.0 : 0.0;
            p.normfact *= 1.0 / n;
            for (int c = 0; c < p.lda; ++c) {
                p.n_in[c] = n;
                p.post[c] = p.node ? POST_PERIOD : POST_NONE;
            }
            return;
        }
        case UNBUNB: {
            const int M = p.node ? 2 * (N - 1) : 2 * N;
            p.r2c = !isComplex;
            p.n_out = isComplex ? M : M / 2 + 1;
            p.symstart = p.node ? N - 1 : N;
            p.normfact *= 1.0 / M;
            for (int c = 0; c < p.lda; ++c) p.n_in[c] = M;
            return;
        }
        case EMPTY: return;
    }
}

}  // namespace flups
