// Oracle restatement of the FLUPS solve pipeline on one rank.
//   constructor / dry run: src/Solver.cpp:23-202, 589-811
//   green:                 src/Solver.cpp:980-1154
//   solve / do_FFT:        src/Solver.cpp:1207-1272, 1402-1462
//   pencil semantics:      src/FFTW_plan_dim.cpp:497-720 (execute + postprocess)
//   pencil switches:       src/SwitchTopoX_a2a.cpp (intersection copy with fieldstart shift)
// Arrays are dense, dimension 0 fastest, complex interleaved, components slowest.
// Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

// green fills, green.cpp
void green_fill_3unb(const Topo& t, const double hfact[3], const double symstart[3], double* g, GreenType type,
                     double length, const std::string& kdir);
void green_fill_2unb(const Topo& t, const double hfact[3], const double kfact[3], const double koffset[3],
                     const double symstart[3], double* g, GreenType type, double length, const std::string& kdir);
void green_fill_1unb(const Topo& t, const double hfact[3], const double kfact[3], const double koffset[3],
                     const double symstart[3], double* g, GreenType type, double length);
void green_fill_0unb(const Topo& t, double hgrid, const double kfact[3], const double koffset[3], const double symstart[3],
                     double* g, GreenType type, double length, const int* istart_custom);

namespace {

inline size_t idx(const Topo& t, int i0, int i1, int i2, int lia) {
    return (size_t)i0 * t.nf + (size_t)t.n[0] * t.nf * ((size_t)i1 + (size_t)t.n[1] * ((size_t)i2 + (size_t)t.n[2] * lia));
}

// copy src (topo a) into a zero array of topo b, element i of a landing at i + shift[d] of b
std::vector<double> switch_topo(const std::vector<double>& src, const Topo& a, const Topo& b, const int shift[3]) {
    if (a.nf != b.nf) throw std::runtime_error("switch between real and complex topologies");
    std::vector<double> dst(b.count(), 0.0);
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = std::max(0, -shift[d]);
        hi[d] = std::min(a.n[d], b.n[d] - shift[d]);
    }
    for (int lia = 0; lia < a.lda; ++lia)
        for (int i2 = lo[2]; i2 < hi[2]; ++i2)
            for (int i1 = lo[1]; i1 < hi[1]; ++i1)
                for (int i0 = lo[0]; i0 < hi[0]; ++i0)
                    for (int c = 0; c < a.nf; ++c)
                        dst[idx(b, i0 + shift[0], i1 + shift[1], i2 + shift[2], lia) + c] = src[idx(a, i0, i1, i2, lia) + c];
    return dst;
}

// one in-place pencil transform with the reference's offsets (FFTW_plan_dim.cpp:670-717)
void exec_pencil(const Plan& p, int lia, bool nf_in_complex, double* buf) {
    const int n = p.n_in[lia], si = p.start_in[lia], so = p.start_out[lia];
    if (p.type == SYMSYM || p.type == MIXUNB) {
        // on a complex topology the r2r acts on the real parts with stride 2
        // (FFTW_plan_dim.cpp:326-330; happens for Green's functions only)
        const int st = nf_in_complex ? 2 : 1;
        std::vector<double> in(n), out(n);
        for (int j = 0; j < n; ++j) in[j] = buf[si + st * j];
        fft_r2r(n, (R2R)p.kind[lia], in.data(), out.data());
        for (int j = 0; j < n; ++j) buf[so + st * j] = out[j];
    } else if (p.r2c) {
        if (p.sign == SIGN_FWD) {
            std::vector<double> in(buf + si, buf + si + n);
            std::vector<cplx> out(n / 2 + 1);
            fft_r2c(n, in.data(), out.data());
            for (int k = 0; k <= n / 2; ++k) { buf[2 * (so + k)] = out[k].real(); buf[2 * (so + k) + 1] = out[k].imag(); }
        } else {
            std::vector<cplx> in(n / 2 + 1);
            for (int k = 0; k <= n / 2; ++k) in[k] = cplx(buf[2 * (si + k)], buf[2 * (si + k) + 1]);
            std::vector<double> out(n);
            fft_c2r(n, in.data(), out.data());
            std::copy(out.begin(), out.end(), buf + so);
        }
    } else {
        std::vector<cplx> in(n), out(n);
        for (int k = 0; k < n; ++k) in[k] = cplx(buf[2 * (si + k)], buf[2 * (si + k) + 1]);
        fft_c2c(n, p.sign, in.data(), out.data());
        for (int k = 0; k < n; ++k) { buf[2 * (so + k)] = out[k].real(); buf[2 * (so + k) + 1] = out[k].imag(); }
    }
}

// FFTW_plan_dim.cpp:497-622; nloc = pencil length in elements after the transform
void postpro_pencil(const Plan& p, int lia, int nloc, bool real_after, double* buf) {
    const int pp = p.postpro[lia];
    const bool f = pp & NULL_FIRST, l = pp & NULL_LAST, per = pp & ENFORCE_PERIOD;
    if (f && !l) buf[0] = 0.0;
    else if (!f && l) buf[nloc - 1] = 0.0;
    else if (f && l) { buf[0] = 0.0; buf[nloc - 1] = 0.0; }
    else if (per) {
        const int nfftw = (p.sign == SIGN_FWD) ? p.n_out - 1 : p.n_in[lia];
        const bool real_dmn = p.r2c && p.sign == SIGN_BWD;
        if (real_dmn) buf[nfftw] = buf[0];
        else { buf[2 * nfftw] = buf[0]; buf[2 * nfftw + 1] = buf[1]; }
    }
    (void)real_after;
}

// apply plan p along its dim: array on topo tin -> array on topo tout (same doubles per pencil)
std::vector<double> apply_plan(const Plan& p, const std::vector<double>& src, const Topo& tin, const Topo& tout) {
    const int d = p.dim;
    const int len = tin.n[d] * tin.nf;
    if (len != tout.n[d] * tout.nf) throw std::runtime_error("pencil size mismatch");
    std::vector<double> dst(tout.count(), 0.0);
    const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
    const int lda = tin.lda;
    const long npen = (long)tin.n[o1] * tin.n[o2] * lda;
#pragma omp parallel
    {
        std::vector<double> buf(len);
#pragma omp for schedule(static)
        for (long q = 0; q < npen; ++q) {
            const int j1 = (int)(q % tin.n[o1]);
            const int j2 = (int)((q / tin.n[o1]) % tin.n[o2]);
            const int lia = (int)(q / ((long)tin.n[o1] * tin.n[o2]));
            int ii[3];
            ii[o1] = j1; ii[o2] = j2;
            for (int k = 0; k < tin.n[d]; ++k) {
                ii[d] = k;
                const size_t s = idx(tin, ii[0], ii[1], ii[2], lia);
                for (int c = 0; c < tin.nf; ++c) buf[(size_t)k * tin.nf + c] = src[s + c];
            }
            exec_pencil(p, lia, tin.nf == 2, buf.data());
            postpro_pencil(p, lia, tout.n[d], tout.nf == 1, buf.data());
            for (int k = 0; k < tout.n[d]; ++k) {
                ii[d] = k;
                const size_t s = idx(tout, ii[0], ii[1], ii[2], lia);
                for (int c = 0; c < tout.nf; ++c) dst[s + c] = buf[(size_t)k * tout.nf + c];
            }
        }
    }
    return dst;
}

void sort_plans(std::vector<Plan>& v) {
    std::stable_sort(v.begin(), v.end(), [](const Plan& a, const Plan& b) { return a.type < b.type; });
}

}  // namespace

Solver::Solver(const int nglob_[3], const int* bc, int lda_, const double h_[3], const double L_[3], bool node_,
               DiffType odiff_)
    : lda(lda_), node(node_), odiff(odiff_) {
    for (int d = 0; d < 3; ++d) { nglob[d] = nglob_[d]; h[d] = h_[d]; L[d] = L_[d]; }
    // bc layout: bc[(d*2+side)*lda + lia]
    auto bcp = [&](int d, int s) { return bc + (d * 2 + s) * lda; };
    std::vector<int> dbc(6 * lda);
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < lda; ++c) {
                int v = bcp(d, s)[c];
                dbc[(d * 2 + s) * lda + c] = (v == EVEN) ? ODD : (v == ODD) ? EVEN : v;
            }
    for (int d = 0; d < 3; ++d) {
        fwd.emplace_back(lda, d, h, L, bcp(d, 0), bcp(d, 1), SIGN_FWD, false, node);
        bwd.emplace_back(lda, d, h, L, bcp(d, 0), bcp(d, 1), SIGN_BWD, false, node);
        gplan.emplace_back(1, d, h, L, bcp(d, 0), bcp(d, 1), SIGN_FWD, true, node);
    }
    sort_plans(fwd);
    sort_plans(bwd);
    sort_plans(gplan);
    if (odiff != NOD) {
        for (int i = 0; i < 3; ++i) {
            const int d = bwd[i].dim;
            bwd_diff.emplace_back(lda, d, h, L, &dbc[(d * 2 + 0) * lda], &dbc[(d * 2 + 1) * lda], SIGN_BWD, false, node);
        }
    }
    ndim = 3;
    for (auto& p : fwd) if (p.type == EMPTY) --ndim;
    if (ndim < 2) throw std::runtime_error("only 2D and 3D fields are handled");

    init_topos(fwd, hat, false);
    init_topos(bwd, nullptr, false);
    init_topos(gplan, gtopo, true);
    if (odiff != NOD) init_topos(bwd_diff, nullptr, false);
    normfact = 1.0; volfact = 1.0;
    for (auto& p : fwd) { normfact *= p.normfact; volfact *= p.volfact; }
}

void Solver::init_topos(std::vector<Plan>& plans, Topo* topos, bool isGreen) {
    int size[3] = {nglob[0], nglob[1], nglob[2]};
    bool cplxd = false;
    for (int ip = 0; ip < ndim; ++ip) {
        plans[ip].init(size, cplxd);
        size[plans[ip].dim] = plans[ip].n_out;
        cplxd = cplxd || plans[ip].r2c;
        if (!isGreen && topos) {
            Topo t;
            for (int d = 0; d < 3; ++d) t.n[d] = size[d];
            t.nf = cplxd ? 2 : 1;
            t.lda = lda;
            if (plans[ip].r2c) t.to_real(plans[ip].dim);   // state before the plan runs
            topos[ip] = t;
        }
    }
    if (isGreen && topos) {
        for (int ip = ndim - 1; ip >= 0; --ip) {
            Topo t;
            for (int d = 0; d < 3; ++d) t.n[d] = size[d];
            t.nf = cplxd ? 2 : 1;
            t.lda = 1;
            t.axis = plans[ip].dim;
            if (plans[ip].r2c_by_fft()) {
                t.to_real(plans[ip].dim);
                size[plans[ip].dim] *= 2;
                cplxd = false;
            }
            topos[ip] = t;
        }
    }
}

void Solver::spectral_info(double kf[3], double ko[3], double ss[3]) const {
    for (int ip = 0; ip < 3; ++ip) {
        const int d = fwd[ip].dim;
        kf[d] = fwd[ip].kfact; ko[d] = fwd[ip].koffset; ss[d] = fwd[ip].symstart;
    }
}

void Solver::setup() {
    compute_green();
    setup_done = true;
}

void Solver::compute_green() {
    bool isSpec[3] = {false, false, false};
    double hfact[3] = {0, 0, 0}, kfact[3] = {0, 0, 0}, koffset[3] = {0, 0, 0}, symstart[3] = {0, 0, 0};
    double length = alpha * h[0];
    if (gtype == HEJ_0) length = h[0] / M_PI;
    const bool iso_needed = (gtype >= HEJ_2 && gtype <= HEJ_0) || gtype == LGF_2;
    if (iso_needed && ((ndim == 3 && (h[0] != h[1] || h[1] != h[2])) || (ndim == 2 && h[0] != h[1])))
        throw std::runtime_error("regularized kernel or LGF requires dx=dy=dz");
    for (int ip = 0; ip < ndim; ++ip) {
        const Plan& p = gplan[ip];
        const int d = p.dim;
        isSpec[d] = p.spectral;
        symstart[d] = p.symstart;
        hfact[d] = h[d];
        if (p.spectral) { hfact[d] = 0.0; kfact[d] = p.kfact; koffset[d] = p.koffset; }
        if (p.type == EMPTY) hfact[d] = 0.0;
    }
    int nspec = 0;
    for (int d = 0; d < ndim; ++d) if (isSpec[d]) ++nspec;
    const int nunb = ndim - nspec;

    Topo t0 = gtopo[0];
    std::vector<double> g(t0.count(), 0.0);
    if (nunb == 3) green_fill_3unb(t0, hfact, symstart, g.data(), gtype, length, kernel_dir);
    else if (nunb == 2) green_fill_2unb(t0, hfact, kfact, koffset, symstart, g.data(), gtype, length, kernel_dir);
    else if (nunb == 1) green_fill_1unb(t0, hfact, kfact, koffset, symstart, g.data(), gtype, length);
    else green_fill_0unb(t0, h[0], kfact, koffset, symstart, g.data(), gtype, length, nullptr);

    Topo cur = t0;
    for (int ip = 0; ip < ndim; ++ip) {
        const Plan& p = gplan[ip];
        Topo tin = gtopo[ip];
        if (ip > 0) {   // pencil switch: identical extents on one rank
            if (tin.nf != cur.nf) throw std::runtime_error("green topo mismatch");
            for (int d = 0; d < 3; ++d) if (tin.n[d] != cur.n[d]) throw std::runtime_error("green topo size mismatch");
        }
        Topo tout = tin;
        if (p.r2c_by_fft()) tout.to_complex(p.dim);
        if (!isSpec[p.dim]) g = apply_plan(p, g, tin, tout);
        gtopo[ip] = tout;
        cur = tout;
    }
    for (double& v : g) v *= volfact;

    const bool hej = gtype >= HEJ_2 && gtype <= HEJ_10;
    const bool lgf = gtype == LGF_2 || gtype == LGF_4 || gtype == LGF_6 || gtype == LGF_8;
    const bool meh = gtype >= MEHR_4L && gtype <= MEHR_6F;
    if (ndim == 3 && nspec == 1 && (hej || lgf || meh)) {
        int istart_c[3];
        for (int ip = 0; ip < 3; ++ip) {
            istart_c[ip] = isSpec[ip] ? 1 : 0;
            kfact[gplan[ip].dim] = gplan[ip].kfact;
        }
        green_fill_0unb(cur, h[0], kfact, koffset, symstart, g.data(), gtype, length, istart_c);
    }
    // the field's spectral topology must match
    Topo fh = hat[ndim - 1];
    if (fwd[ndim - 1].r2c) fh.to_complex(fwd[ndim - 1].dim);
    if (fh.nf != cur.nf || fh.n[0] != cur.n[0] || fh.n[1] != cur.n[1] || fh.n[2] != cur.n[2])
        throw std::runtime_error("Green and field spectral topologies differ");
    green.swap(g);
}

std::vector<double> Solver::forward(const double* rhs) const {
    Topo phys;
    for (int d = 0; d < 3; ++d) phys.n[d] = nglob[d];
    phys.nf = 1; phys.lda = lda;
    std::vector<double> cur(rhs, rhs + phys.count());
    Topo tc = phys;
    for (int ip = 0; ip < ndim; ++ip) {
        const Plan& p = fwd[ip];
        Topo tin = hat[ip];
        int shift[3] = {0, 0, 0};
        shift[p.dim] = p.fieldstart;
        cur = switch_topo(cur, tc, tin, shift);
        Topo tout = tin;
        if (p.r2c) tout.to_complex(p.dim);
        cur = apply_plan(p, cur, tin, tout);
        tc = tout;
    }
    return cur;
}

void Solver::multiply(std::vector<double>& data, SolverType type) const {
    Topo t = hat[ndim - 1];
    if (fwd[ndim - 1].r2c) t.to_complex(fwd[ndim - 1].dim);
    const size_t per = (size_t)t.n[0] * t.n[1] * t.n[2];
    if (type == STD) {
        for (int lia = 0; lia < lda; ++lia) {
            double* dd = data.data() + lia * per * t.nf;
            if (t.nf == 1) {
                for (size_t i = 0; i < per; ++i) dd[i] *= normfact * green[i];
            } else {
                for (size_t i = 0; i < per; ++i) {
                    const double a = dd[2 * i], b = dd[2 * i + 1], c = green[2 * i], e = green[2 * i + 1];
                    dd[2 * i] = normfact * (a * c - b * e);
                    dd[2 * i + 1] = normfact * (a * e + b * c);
                }
            }
        }
        return;
    }
    // ROT: Solver.cpp:1484-1537 and dothemagic_rot.ipp
    if (lda != 3) throw std::runtime_error("ROT requires lda = 3");
    if (odiff == NOD) throw std::runtime_error("ROT requires a differentiation order");
    double kabs[3], koff[3], ss[3];
    spectral_info(kabs, koff, ss);
    double kf[3][3][2];
    for (int ip = 0; ip < 3; ++ip) {
        const int d = fwd[ip].dim;
        for (int c = 0; c < 3; ++c) {
            int ph = 0;
            if (fwd[ip].imult[c]) --ph;
            if (bwd_diff[ip].imult[c]) ++ph;
            if (ph == 0) { kf[d][c][0] = 0.0; kf[d][c][1] = kabs[d]; }
            else if (ph == 1) { kf[d][c][0] = -kabs[d]; kf[d][c][1] = 0.0; }
            else { kf[d][c][0] = kabs[d]; kf[d][c][1] = 0.0; }
        }
    }
    auto sym = [&](int i, int d) -> int {
        if (ss[d] == 0.0 || i <= ss[d]) return i;
        return -std::max((int)std::fabs(2.0 * ss[d] - i), 1);
    };
    // derivative symbol applied to a wavenumber factor kf (times index)
    auto dsym = [&](double k, double hh) -> double {
        switch (odiff) {
            case SPE: return k;
            case FD2: return std::sin(k * hh) / hh;
            case FD4: return (4.0 / 3.0) * std::sin(k * hh) / hh - (1.0 / 6.0) * std::sin(2.0 * k * hh) / hh;
            case FD6:
                return 1.5 * std::sin(k * hh) / hh - 0.3 * std::sin(2.0 * k * hh) / hh + (1.0 / 30.0) * std::sin(3.0 * k * hh) / hh;
            default: return 0.0;
        }
    };
    for (int i2 = 0; i2 < t.n[2]; ++i2)
        for (int i1 = 0; i1 < t.n[1]; ++i1)
            for (int i0 = 0; i0 < t.n[0]; ++i0) {
                const int is[3] = {sym(i0, 0), sym(i1, 1), sym(i2, 2)};
                const size_t e = (size_t)i0 + (size_t)t.n[0] * ((size_t)i1 + (size_t)t.n[1] * i2);
                // K[d][c] = derivative along d of component c, complex
                double Kr[3][3], Ki[3][3];
                for (int d = 0; d < 3; ++d)
                    for (int c = 0; c < 3; ++c) {
                        const double base = is[d] + koff[d];
                        if (t.nf == 1) {
                            Kr[d][c] = dsym(base * (kf[d][c][0] + kf[d][c][1]), h[d]);
                            Ki[d][c] = 0.0;
                        } else {
                            Kr[d][c] = dsym(base * kf[d][c][0], h[d]);
                            Ki[d][c] = dsym(base * kf[d][c][1], h[d]);
                        }
                    }
                double fr[3], fi[3];
                for (int c = 0; c < 3; ++c) {
                    fr[c] = data[(c * per + e) * t.nf];
                    fi[c] = t.nf == 2 ? data[(c * per + e) * 2 + 1] : 0.0;
                }
                auto dr = [&](int c, int d) { return fr[c] * Kr[d][c] - fi[c] * Ki[d][c]; };
                auto di = [&](int c, int d) { return fr[c] * Ki[d][c] + fi[c] * Kr[d][c]; };
                const double r0r = dr(2, 1) - dr(1, 2), r0i = di(2, 1) - di(1, 2);
                const double r1r = dr(0, 2) - dr(2, 0), r1i = di(0, 2) - di(2, 0);
                const double r2r = dr(1, 0) - dr(0, 1), r2i = di(1, 0) - di(0, 1);
                const double rr[3] = {r0r, r1r, r2r}, ri[3] = {r0i, r1i, r2i};
                if (t.nf == 1) {
                    const double g = green[e];
                    for (int c = 0; c < 3; ++c) data[c * per + e] = normfact * rr[c] * g;
                } else {
                    const double gr = green[2 * e], gi = green[2 * e + 1];
                    for (int c = 0; c < 3; ++c) {
                        data[(c * per + e) * 2] = normfact * (rr[c] * gr - ri[c] * gi);
                        data[(c * per + e) * 2 + 1] = normfact * (rr[c] * gi + ri[c] * gr);
                    }
                }
            }
}

void Solver::backward(std::vector<double>& data, double* field, SolverType type) const {
    const std::vector<Plan>& plans = (type == STD) ? bwd : bwd_diff;
    Topo tc = hat[ndim - 1];
    if (fwd[ndim - 1].r2c) tc.to_complex(fwd[ndim - 1].dim);
    std::vector<double> cur = std::move(data);
    for (int ip = ndim - 1; ip >= 0; --ip) {
        const Plan& p = plans[ip];
        Topo tout = tc;
        if (fwd[ip].r2c) tout.to_real(fwd[ip].dim);
        cur = apply_plan(p, cur, tc, tout);
        // switch back to the previous topology
        Topo prev;
        if (ip == 0) {
            for (int d = 0; d < 3; ++d) prev.n[d] = nglob[d];
            prev.nf = 1; prev.lda = lda;
        } else {
            prev = hat[ip - 1];
            if (fwd[ip - 1].r2c) prev.to_complex(fwd[ip - 1].dim);
        }
        int shift[3] = {0, 0, 0};
        shift[fwd[ip].dim] = -fwd[ip].fieldstart;
        cur = switch_topo(cur, tout, prev, shift);
        tc = prev;
    }
    std::copy(cur.begin(), cur.end(), field);
}

void Solver::solve(double* field, const double* rhs, SolverType type) {
    if (!setup_done) setup();
    if (type == ROT && lda != 3) throw std::runtime_error("ROT requires lda = 3");
    if (type == ROT && odiff == NOD) throw std::runtime_error("ROT requires a differentiation order");
    std::vector<double> d = forward(rhs);
    multiply(d, type);
    backward(d, field, type);
}

}  // namespace oracle
