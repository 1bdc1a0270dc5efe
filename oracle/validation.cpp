// Oracle restatement of the reference's validation harnesses, the producers of
// samples/validation/data_ref (samples/validation/src/validation_3d.cpp:28-486
// with MANUFACTURED_SOLUTION, validation_3d.hpp:64-165, main.cpp:70-135) and
// samples/solve_vtube/data_ref (solve_vtube/src/vtube.cpp: vortex tubes with
// image tubes, ROT solve; main_vtube.cpp:110-200 for the boundary conditions).
// Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "oracle.hpp"

namespace oracle {

namespace {

struct Manu {
    double freq = 1, sign[2] = {0, 0}, center = 0.5, sigma = 0.5;
};
const double C_BUMP = 10.0;

double bump(double y) { return std::fabs(y) >= 1.0 ? 0.0 : std::exp(C_BUMP * (1.0 - 1.0 / (1.0 - y * y))); }
double bump_d2(double ysq, double sl) {
    if (std::fabs(ysq) >= 1.0) return 0.0;
    return std::exp(C_BUMP * (1.0 - 1.0 / (1.0 - ysq))) * (2.0 * C_BUMP * (2.0 * (C_BUMP - 1.0) * ysq + 3.0 * ysq * ysq - 1.0)) /
           std::pow(ysq - 1.0, 4) / (sl * sl);
}

enum FKind { ODDODD, EVENEVEN, ODDEVEN, EVENODD, UNBK, CST };

double fsol(FKind k, double x, double L, const Manu& p) {
    switch (k) {
        case ODDODD: return std::sin((2.0 * M_PI / L * p.freq) * x);
        case EVENEVEN: return std::cos((M_PI / L * p.freq) * x);
        case ODDEVEN: return std::sin((M_PI / L * (p.freq + 0.5)) * x);
        case EVENODD: return std::cos((M_PI / L * (p.freq + 0.5)) * x);
        case UNBK: {
            const double sl = p.sigma * L;
            return bump((x - p.center * L) / sl) + p.sign[0] * bump((x + p.center * L) / sl) +
                   p.sign[1] * bump((x - (2.0 - p.center) * L) / sl);
        }
        default: return 1.0;
    }
}
double frhs(FKind k, double x, double L, const Manu& p) {
    switch (k) {
        case ODDODD: { const double a = 2.0 * M_PI / L * p.freq; return -a * a * std::sin(a * x); }
        case EVENEVEN: { const double a = M_PI / L * p.freq; return -a * a * std::cos(a * x); }
        case ODDEVEN: { const double a = M_PI / L * (p.freq + 0.5); return -a * a * std::sin(a * x); }
        case EVENODD: { const double a = M_PI / L * (p.freq + 0.5); return -a * a * std::cos(a * x); }
        case UNBK: {
            const double sl = p.sigma * L;
            const double y0 = (x - p.center * L) / sl, y1 = (x + p.center * L) / sl, y2 = (x - (2.0 - p.center) * L) / sl;
            return bump_d2(y0 * y0, sl) + p.sign[0] * bump_d2(y1 * y1, sl) + p.sign[1] * bump_d2(y2 * y2, sl);
        }
        default: return 0.0;
    }
}

}  // namespace

void validation_geometry(int center, const int bc6[6], int N, int nglob[3], double h[3]) {
    const bool cell = (center == 1);
    const int fact = cell ? 0 : 1;
    const bool twod = (bc6[4] == NONE);
    const int res = N - fact;
    nglob[0] = res + fact; nglob[1] = res + fact; nglob[2] = twod ? 1 : res + fact;
    h[0] = 1.0 / (nglob[0] - fact);
    h[1] = 1.0 / (nglob[1] - fact);
    h[2] = nglob[2] == 1 ? 1.0 : 1.0 / (nglob[2] - fact);
}

void validation_fields(int center, const int bc6[6], int N, int lda, double* rhs_out, double* sol_out) {
    const bool cell = (center == 1);
    int nglob[3];
    double h[3];
    validation_geometry(center, bc6, N, nglob, h);
    const double L[3] = {1.0, 1.0, 1.0};
    const size_t per = (size_t)nglob[0] * nglob[1] * nglob[2];
    std::vector<double> rhs(per * lda, 0.0), sol(per * lda, 1.0);
    for (int lia = 0; lia < lda; ++lia) {
        Manu p[3];
        p[0].freq = 1; p[1].freq = 2; p[2].freq = 4;
        FKind fk[3];
        for (int d = 0; d < 3; ++d) {
            const int l = bc6[2 * d], r = bc6[2 * d + 1];
            if ((l == PER && r == PER) || (l == ODD && r == ODD)) fk[d] = ODDODD;
            else if (l == EVEN && r == EVEN) fk[d] = EVENEVEN;
            else if (l == ODD && r == EVEN) fk[d] = ODDEVEN;
            else if (l == EVEN && r == ODD) fk[d] = EVENODD;
            else if (l == UNB) {
                if (r == ODD) { p[d].center = 0.7; p[d].sign[1] = -1.0; }
                else if (r == EVEN) { p[d].center = 0.7; p[d].sign[1] = 1.0; }
                fk[d] = UNBK;
            } else if (r == UNB) {
                if (l == ODD) { p[d].center = 0.3; p[d].sign[0] = -1.0; }
                else if (l == EVEN) { p[d].center = 0.3; p[d].sign[0] = 1.0; }
                fk[d] = UNBK;
            } else {
                fk[d] = CST;
            }
        }
        const double shift = cell ? 0.5 : 0.0;
        for (int i2 = 0; i2 < nglob[2]; ++i2)
            for (int i1 = 0; i1 < nglob[1]; ++i1)
                for (int i0 = 0; i0 < nglob[0]; ++i0) {
                    const size_t id = lia * per + (size_t)i0 + (size_t)nglob[0] * ((size_t)i1 + (size_t)nglob[1] * i2);
                    const double x[3] = {(i0 + shift) * h[0], (i1 + shift) * h[1], (i2 + shift) * h[2]};
                    for (int d = 0; d < 3; ++d) {
                        const int d2 = (d + 1) % 3, d3 = (d + 2) % 3;
                        sol[id] *= fsol(fk[d], x[d], L[d], p[d]);
                        rhs[id] += frhs(fk[d], x[d], L[d], p[d]) * fsol(fk[d2], x[d2], L[d2], p[d2]) * fsol(fk[d3], x[d3], L[d3], p[d3]);
                    }
                }
    }
    std::copy(rhs.begin(), rhs.end(), rhs_out);
    std::copy(sol.begin(), sol.end(), sol_out);
}

void validation_case(int center, const int bc6[6], int kernel, int N, int lda, const std::string& kdir, double* err) {
    const bool cell = (center == 1);
    int nglob[3];
    double h[3];
    validation_geometry(center, bc6, N, nglob, h);
    const double L[3] = {1.0, 1.0, 1.0};
    std::vector<int> bc(6 * lda);
    for (int f = 0; f < 6; ++f)
        for (int c = 0; c < lda; ++c) bc[f * lda + c] = bc6[f];
    Solver s(nglob, bc.data(), lda, h, L, !cell, NOD);
    s.kernel_dir = kdir;
    s.set_green((GreenType)kernel);
    s.setup();
    const size_t per = (size_t)nglob[0] * nglob[1] * nglob[2];
    std::vector<double> rhs(per * lda), sol(per * lda), field(per * lda, 0.0);
    validation_fields(center, bc6, N, lda, rhs.data(), sol.data());
    s.solve(field.data(), rhs.data(), STD);
    double vol = 1.0;
    for (int d = 0; d < 3; ++d)
        if (bc6[2 * d] != NONE && bc6[2 * d + 1] != NONE) vol *= h[d];
    for (int lia = 0; lia < lda; ++lia) {
        double e2 = 0, ei = 0;
        for (size_t i = 0; i < per; ++i) {
            const double e = sol[lia * per + i] - field[lia * per + i];
            ei = std::max(ei, std::fabs(e));
            e2 += e * e * vol;
        }
        err[2 * lia] = std::sqrt(e2);
        err[2 * lia + 1] = ei;
    }
}


// vortex-tube Biot-Savart case: tube along vdir, image tubes by symx/symy in {-1,0,1},
// Gaussian core sigma, cell-centred N^3 on [0,1]^3, CHAT_2 kernel unless given
void vtube_case(int vdir, int symx, int symy, int N, int kernel, int order, double sigma, const std::string& kdir, double* err,
                int* bc_out, double* rhs_out, double* sol_out) {
    const int lda = 3;
    const int nglob[3] = {N, N, N};
    const double L[3] = {1, 1, 1};
    const double h[3] = {1.0 / N, 1.0 / N, 1.0 / N};
    const int dir2 = vdir, dir0 = (vdir + 1) % 3, dir1 = (vdir + 2) % 3;
    int bc[3][2][3];
    double xc = 0.5, yc = 0.5, xs = 0.0, ys = 0.0;
    // main_vtube.cpp:117-190: odd/odd/even along the tube, images by symmetry, unbounded otherwise
    for (int c = 0; c < 3; ++c) {
        bc[dir2][0][c] = bc[dir2][1][c] = (c == dir2) ? EVEN : ODD;
        bc[dir0][0][c] = bc[dir0][1][c] = UNB;
        bc[dir1][0][c] = bc[dir1][1][c] = UNB;
    }
    if (symy == 1) { bc[dir1][0][dir0] = EVEN; bc[dir1][0][dir1] = ODD; bc[dir1][0][dir2] = EVEN; }
    if (symy == -1) { bc[dir1][0][dir0] = ODD; bc[dir1][0][dir1] = EVEN; bc[dir1][0][dir2] = ODD; }
    if (symx == 1) { bc[dir0][0][dir0] = ODD; bc[dir0][0][dir1] = EVEN; bc[dir0][0][dir2] = EVEN; }
    if (symx == -1) { bc[dir0][0][dir0] = EVEN; bc[dir0][0][dir1] = ODD; bc[dir0][0][dir2] = ODD; }
    if (symx) { xc = 0.25; xs = symx; }
    if (symy) { yc = 0.25; ys = symy; }
    std::vector<int> flat(6 * lda);
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < 3; ++c) flat[(d * 2 + s) * lda + c] = bc[d][s][c];
    if (bc_out) std::copy(flat.begin(), flat.end(), bc_out);
    const size_t per = (size_t)N * N * N;
    std::vector<double> rhs(3 * per, 0.0), sol(3 * per, 0.0), field(3 * per, 0.0);
    const double eps = 100 * 2.220446049250313e-16;
    for (int i2 = 0; i2 < N; ++i2)
        for (int i1 = 0; i1 < N; ++i1)
            for (int i0 = 0; i0 < N; ++i0) {
                const size_t id = (size_t)i0 + (size_t)N * ((size_t)i1 + (size_t)N * i2);
                const double pos[3] = {(i0 + 0.5) * h[0], (i1 + 0.5) * h[1], (i2 + 0.5) * h[2]};
                const double sx[4] = {-1, +1, -1, +1}, sy[4] = {-1, -1, +1, +1}, w[4] = {1.0, xs, ys, xs * ys};
                for (int t = 0; t < 4; ++t) {
                    const double x = pos[dir0] + sx[t] * xc * L[dir0];
                    const double y = pos[dir1] + sy[t] * yc * L[dir1];
                    const double th = std::atan2(y, x), r = std::sqrt(x * x + y * y), rho = r / sigma;
                    const double vort = 1.0 / (2.0 * M_PI * sigma * sigma) * std::exp(-rho * rho * 0.5);
                    const double vel = (r < eps) ? 0.0 : 1.0 / (2.0 * M_PI * r) * (1.0 - std::exp(-rho * rho * 0.5));
                    rhs[dir2 * per + id] += -vort * w[t];
                    sol[dir0 * per + id] += -std::sin(th) * vel * w[t];
                    sol[dir1 * per + id] += +std::cos(th) * vel * w[t];
                }
            }
    if (rhs_out) std::copy(rhs.begin(), rhs.end(), rhs_out);
    if (sol_out) std::copy(sol.begin(), sol.end(), sol_out);
    if (!err) return;
    Solver s(nglob, flat.data(), lda, h, L, false, (DiffType)order);
    s.kernel_dir = kdir;
    s.set_green((GreenType)kernel);
    s.setup();
    s.solve(field.data(), rhs.data(), ROT);
    const double vol = h[0] * h[1] * h[2];
    for (int c = 0; c < 3; ++c) {
        double e2 = 0, ei = 0;
        for (size_t i = 0; i < per; ++i) {
            const double e = sol[c * per + i] - field[c * per + i];
            ei = std::max(ei, std::fabs(e));
            e2 += e * e * vol;
        }
        err[2 * c] = std::sqrt(e2);
        err[2 * c + 1] = ei;
    }
}

}  // namespace oracle
