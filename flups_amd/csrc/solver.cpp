// Solver set-up and solve orchestration for one GPU.
//
//   constructor: plan creation, sorting and dry run (src/Solver.cpp:23-202, 589-811)
//   setup:       Green's function on the device (src/Solver.cpp:219-469, 980-1193)
//   solve:       copy-in, forward passes, multiply, backward passes, copy-out
//                (src/Solver.cpp:1207-1272, 1402-1542)
#include "solver.hpp"

#include "group.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace flups {

size_t generic_lds_bytes(const PassDesc& P, int ppb);
bool generic_alias(const PassDesc& P);
void launch_generic(const PassDesc& P, hipStream_t s);
bool row_r2r_fits(const PassDesc& P);
bool col_real_fits(const PassDesc& P);
void launch_col_real(const PassDesc& P, hipStream_t s);
bool row_rc_fits(const PassDesc& P);
void launch_row_rc(const PassDesc& P, hipStream_t s);
std::vector<double> row_r2r_table(int kind, int n);
void launch_row_r2r(const PassDesc& P, hipStream_t s);
bool col_r2r_fits(const PassDesc& P);
bool col_c2c_fits(const PassDesc& P);
bool split_pq(int n, int& pf, int& q);
int r2r_h(int kind, int n);
bool row_c2c_fits(const PassDesc& P);
void launch_row_c2c(const PassDesc& P, hipStream_t s);
void launch_col_c2c(const PassDesc& P, hipStream_t s);
void launch_col_r2r(const PassDesc& P, hipStream_t s);
FastPlan* make_fast_plan(SolverImpl& s);   // nullptr if the fused pipeline does not apply
void destroy_fast_plan(FastPlan* f);
void fast_solve(FastPlan* f, double* field, const double* rhs, const View& uin, const View& uout, hipStream_t s,
                int type);
void fast_collect(FastPlan* f);
bool fast_profiling(const FastPlan* f);
void fast_profile(FastPlan* f, bool on);
int fast_profile_get(FastPlan* f, int i, const char** name, double* ms, long* calls, double* bytes);
size_t fast_device_bytes(const FastPlan* f);
void fast_stage_bytes(const FastPlan* f, size_t* b0, size_t* b1);
void fast_solve_on(FastPlan* f, void* s0, void* s1, double* field, const double* rhs, const View& uin, const View& uout,
                   hipStream_t s, int type);
void scratch_init();                           // scratch.hip
double* scratch_device_ptr(size_t* bytes);
HybridZ* make_hybrid_z(SolverImpl& s);
void destroy_hybrid_z(HybridZ* h);
size_t hybrid_device_bytes(const HybridZ* h);
double hybrid_bytes(const HybridZ* h);
void hybrid_z(const HybridZ* h, void* data, hipStream_t st);
void fast_schedule(const FastPlan* f, int out[3]);
double fast_bytes(const FastPlan* f);
void fast_autotune(FastPlan* f, double* out, const double* in, const View& uin, const View& uout, const double* dgreen,
                   double normfact, hipStream_t st);

// ---------------------------------------------------------------------------
// topology sizes (src/Topology.cpp:81-98): the fastest direction is padded so
// that a row is a whole number of `alignment` bytes
void Topo::cmpt_sizes() {
    for (int d = 0; d < 3; ++d) {
        nloc[d] = nglob[d];
        nmem[d] = nloc[d];
        if (d == axis) {
            const int modulo = (int)((nloc[d] * nf * sizeof(double)) % alignment);
            const int delta = (int)((alignment - modulo) / sizeof(double));
            nmem[d] += (modulo == 0) ? 0 : delta / nf;
        }
    }
}

std::string kernel_dir() {
    if (const char* e = std::getenv("FLUPS_KERNEL_PATH")) return e;
    Dl_info info;
    if (dladdr((void*)&kernel_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t s = p.find_last_of('/');
        return (s == std::string::npos ? std::string(".") : p.substr(0, s)) + "/kernel";
    }
    return "kernel";
}

void Timers::add(const std::string& n, double host_s, double dev_ms, long calls) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& x : e)
        if (x.name == n) {
            x.host_s += host_s;
            x.dev_ms += dev_ms;
            x.calls += calls;
            return;
        }
    Entry x;
    x.name = n;
    x.host_s = host_s;
    x.dev_ms = dev_ms;
    x.calls = calls;
    e.push_back(x);
}

// one line per timer, in first-recorded order, indented by nesting depth; the share of each
// stage is taken against the device time of its parent
void Timers::display(FILE* f) {
    std::lock_guard<std::mutex> lk(mu);
    std::fprintf(f, "===== FLUPS profiler \"%s\" =====\n", name.c_str());
    std::fprintf(f, "%-44s %8s %12s %12s %12s %8s\n", "timer", "calls", "host [s]", "device [ms]", "mean [ms]", "% parent");
    for (const auto& x : e) {
        const size_t cut = x.name.find_last_of('/');
        const int depth = (int)std::count(x.name.begin(), x.name.end(), '/');
        double parent = 0;
        if (cut != std::string::npos)
            for (const auto& y : e)
                if (y.name == x.name.substr(0, cut)) parent = y.dev_ms;
        const std::string label = std::string(2 * depth, ' ') + (cut == std::string::npos ? x.name : x.name.substr(cut + 1));
        const double mean = x.calls ? (x.dev_ms > 0 ? x.dev_ms : 1e3 * x.host_s) / x.calls : 0.0;
        if (parent > 0)
            std::fprintf(f, "%-44s %8ld %12.6f %12.4f %12.5f %8.1f\n", label.c_str(), x.calls, x.host_s, x.dev_ms, mean,
                         100.0 * x.dev_ms / parent);
        else
            std::fprintf(f, "%-44s %8ld %12.6f %12.4f %12.5f %8s\n", label.c_str(), x.calls, x.host_s, x.dev_ms, mean, "-");
    }
    std::fflush(f);
}

View dense_view(void* p, const Shape& sh, int es) {
    View v;
    v.p = p;
    v.es = es;
    for (int d = 0; d < 3; ++d) v.n[d] = sh.n[d];
    v.nf = sh.nf;
    v.s[0] = 1;
    v.s[1] = sh.n[0];
    v.s[2] = (long long)sh.n[0] * sh.n[1];
    v.sc = (long long)sh.per();
    return v;
}

View SolverImpl::user_view(double* p) const {
    View v;
    v.p = p;
    for (int d = 0; d < 3; ++d) v.n[d] = phys_.nglob[d];
    v.nf = 1;
    const int a = phys_.axis, a1 = (a + 1) % 3, a2 = (a + 2) % 3;
    v.s[a] = 1;
    v.s[a1] = phys_.nmem[a];
    v.s[a2] = (long long)phys_.nmem[a] * phys_.nmem[a1];
    v.sc = (long long)phys_.memdim();
    return v;
}

// ---------------------------------------------------------------------------
SolverImpl::SolverImpl(const Topo* t, const int* bc, const double h[3], const double L[3], int odiff, const int center[3])
    : phys_(*t), lda_(t->lda), odiff_(odiff) {
    FLUPS_REQUIRE(center[0] == center[1] && center[0] == center[2], "all directions must share the same data location");
    FLUPS_REQUIRE(t->nproc[0] * t->nproc[1] * t->nproc[2] == 1, "this library runs on one GPU: nproc must be 1,1,1");
    FLUPS_REQUIRE(lda_ >= 1 && lda_ <= 3, "lda must be 1..3");
    node_ = (center[0] == 0);
    for (int d = 0; d < 3; ++d) { h_[d] = h[d]; L_[d] = L[d]; }
    auto bcp = [&](int d, int s) { return bc + (d * 2 + s) * lda_; };
    int dbc[3][2][3];
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < lda_; ++c) {
                const int v = bcp(d, s)[c];
                dbc[d][s][c] = (v == EVEN) ? ODD : (v == ODD) ? EVEN : v;
            }
    std::vector<DirPlan> f, b, g;
    for (int d = 0; d < 3; ++d) {
        f.push_back(make_dirplan(d, lda_, bcp(d, 0), bcp(d, 1), h_, L_, FWD, false, node_));
        b.push_back(make_dirplan(d, lda_, bcp(d, 0), bcp(d, 1), h_, L_, BWD, false, node_));
        g.push_back(make_dirplan(d, 1, bcp(d, 0), bcp(d, 1), h_, L_, FWD, true, node_));
    }
    auto by_type = [](const DirPlan& a, const DirPlan& c) { return a.type < c.type; };   // src/FFTW_plan_dim.cpp:35-51
    std::stable_sort(f.begin(), f.end(), by_type);
    std::stable_sort(b.begin(), b.end(), by_type);
    std::stable_sort(g.begin(), g.end(), by_type);
    for (int i = 0; i < 3; ++i) { fwd_[i] = f[i]; bwd_[i] = b[i]; gpl_[i] = g[i]; }
    if (odiff_ != NOD)
        for (int i = 0; i < 3; ++i) {
            const int d = bwd_[i].dim;
            bwdd_[i] = make_dirplan(d, lda_, dbc[d][0], dbc[d][1], h_, L_, BWD, false, node_);
        }
    ndim_ = 3;
    for (int i = 0; i < 3; ++i)
        if (fwd_[i].type == EMPTY) --ndim_;
    FLUPS_REQUIRE(ndim_ > 1, "only 2-D and 3-D fields are handled");

    // dry runs (src/Solver.cpp:629-717)
    auto dry = [&](DirPlan* plans, bool record) {
        int size[3] = {phys_.nglob[0], phys_.nglob[1], phys_.nglob[2]};
        bool cplx = false;
        for (int ip = 0; ip < ndim_; ++ip) {
            init_dirplan(plans[ip], size, cplx);
            const int before = size[plans[ip].dim];
            (void)before;
            size[plans[ip].dim] = plans[ip].n_out;
            cplx = cplx || plans[ip].r2c;
            if (record) {
                Shape s;
                for (int d = 0; d < 3; ++d) s.n[d] = size[d];
                s.nf = cplx ? 2 : 1;
                s.lda = lda_;
                stage_[ip] = s;
                if (plans[ip].r2c) s.to_real(plans[ip].dim);
                hat_[ip] = s;
            }
        }
        return std::make_pair(std::vector<int>(size, size + 3), cplx);
    };
    dry(fwd_, true);
    dry(bwd_, false);
    if (odiff_ != NOD) dry(bwdd_, false);
    auto gfinal = dry(gpl_, false);
    // Green's topologies, built backwards (src/Solver.cpp:731-789)
    {
        std::vector<int> size = gfinal.first;
        bool cplx = gfinal.second;
        for (int ip = ndim_ - 1; ip >= 0; --ip) {
            Shape s;
            for (int d = 0; d < 3; ++d) s.n[d] = size[d];
            s.nf = cplx ? 2 : 1;
            s.lda = 1;
            Shape post = s;
            if (gpl_[ip].r2c_by_fft()) {
                s.to_real(gpl_[ip].dim);
                size[gpl_[ip].dim] *= 2;
                cplx = false;
            }
            gpre_[ip] = s;
            gpost_[ip] = post;
        }
    }
    normfact_ = 1.0;
    volfact_ = 1.0;
    for (int i = 0; i < 3; ++i) { normfact_ *= fwd_[i].normfact; volfact_ *= fwd_[i].volfact; }
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    int lds = 0;
    HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    lds_budget_ = (size_t)lds;
    if (const char* e = std::getenv("FLUPS_GENERIC_LDS")) generic_lds_cap_ = (size_t)std::atoll(e);
    kdir_ = kernel_dir();
}

SolverImpl::~SolverImpl() {
    if (fast_) destroy_fast_plan(fast_);
    if (hybrid_) destroy_hybrid_z(hybrid_);
    for (auto& kv : tw_) (void)hipFree(kv.second);
    for (auto& kv : r2r_tab_) (void)hipFree(kv.second);
    for (void* p : dstage_) if (p) (void)hipFree(p);
    if (dgreen_) (void)hipFree(dgreen_);
    if (dgreen32_) (void)hipFree(dgreen32_);
    if (dstaging_) (void)hipFree(dstaging_);
    if (dinner_) (void)hipFree(dinner_);
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    if (cap_) (void)hipStreamDestroy(cap_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    for (hipEvent_t e : pev_) (void)hipEventDestroy(e);
}

size_t SolverImpl::alloc_size() const {
    size_t m = 1;
    for (int ip = 0; ip < ndim_; ++ip) m = std::max(m, std::max(hat_[ip].count(), stage_[ip].count()));
    return std::max(m, phys_.memsize());
}

void SolverImpl::spectral_info(double kf[3], double ko[3], double ss[3]) const {
    for (int ip = 0; ip < 3; ++ip) {
        const int d = fwd_[ip].dim;
        kf[d] = fwd_[ip].kfact;
        ko[d] = fwd_[ip].koffset;
        ss[d] = fwd_[ip].symstart;
    }
}

// ---------------------------------------------------------------------------
const double* SolverImpl::twiddles(int M) {
    auto it = tw_.find(M);
    if (it != tw_.end()) return it->second;
    FLUPS_REQUIRE(!frozen_, "twiddle table of length %d missing in a solve that may not allocate", M);
    std::vector<double> t(2 * (size_t)M);
    for (int i = 0; i < M; ++i) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)i / (long double)M;
        t[2 * i] = (double)cosl(a);
        t[2 * i + 1] = (double)sinl(a);
    }
    double* d = nullptr;
    HIP_CHECK(hipMalloc(&d, t.size() * sizeof(double)));
    HIP_CHECK(hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    tw_[M] = d;
    return d;
}

const double* SolverImpl::r2r_table(int kind, int n) {
    const auto key = std::make_pair(kind, n);
    auto it = r2r_tab_.find(key);
    if (it != r2r_tab_.end()) return it->second;
    FLUPS_REQUIRE(!frozen_, "real-to-real table (kind %d, n %d) missing in a solve that may not allocate", kind, n);
    const auto t = row_r2r_table(kind, n);
    double* d = nullptr;
    HIP_CHECK(hipMalloc(&d, t.size() * sizeof(double)));
    HIP_CHECK(hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    r2r_tab_[key] = d;
    return d;
}

void SolverImpl::plan_fft(PassDesc& P, int M) {
    P.M = M;
    P.nst = 0;
    int m = M;
    for (int p : {4, 2, 3, 5, 7}) {
        while (m % p == 0) { P.rad[P.nst++] = p; m /= p; }
    }
    for (int p = 11; m > 1; p += 2) {
        while (m % p == 0) { P.rad[P.nst++] = p; m /= p; }
    }
    FLUPS_REQUIRE(P.nst <= 16, "transform length %d has too many factors", M);
    P.tw = twiddles(M);
}

// Arrays larger than half the 256 MB Infinity Cache are read or written with non-temporal
// accesses where that was measured to pay: the row r2r pass between the caller's rows and a
// stage array (256^3, x EVEN-ODD / y PER / z UNB: 2828 -> 2849 solves/s). Non-temporal stores of
// the forward paired-column pass into the spectral array the fused z pass reads next measured
// slower (2800 -> 2694-2728) and are not used.
static constexpr double kNtBytes = 134217728.0;
static double view_bytes(const View& v) { return (double)v.es * v.nf * v.n[0] * v.n[1] * v.n[2]; }

PassDesc SolverImpl::pass_desc(const DirPlan& p, int comp, bool forward, const View& src, const View& dst, const Shape& tin,
                               const Shape& tout, int gshift, int sshift) {
    PassDesc P;
    P.d = p.dim;
    P.sign = p.sign;
    P.n_out_plan = p.n_out;
    P.post = p.post[comp];
    P.n = p.n_in[comp];
    P.si = p.start_in[comp];
    P.so = p.start_out[comp];
    P.kind = p.kind[comp];
    const int d = p.dim;
    P.pen_in = tin.n[d];
    P.nf_in = tin.nf;
    P.pen_out = tout.n[d];
    P.nf_out = tout.nf;
    FLUPS_REQUIRE(P.pen_in * P.nf_in == P.pen_out * P.nf_out, "pencil size mismatch");
    P.gshift = gshift;
    P.sshift = sshift;
    int M = 1;
    if (p.type == SYMSYM || p.type == MIXUNB) {
        P.op = OP_R2R;
        P.r2r_stride = tin.nf;
        switch (P.kind) {
            case REDFT00: M = 2 * (P.n - 1); break;
            case RODFT00: M = 2 * (P.n + 1); break;
            default: M = 2 * P.n;
        }
        FLUPS_REQUIRE(P.n >= 1 && M >= 1, "degenerate real-to-real transform of length %d", P.n);
    } else if (p.r2c) {
        P.op = forward ? OP_R2C : OP_C2R;
        M = (P.n % 2 == 0) ? P.n / 2 : P.n;
    } else {
        P.op = OP_C2C;
        M = P.n;
    }
    plan_fft(P, M);
    P.src = static_cast<const char*>(src.p) + (size_t)src.es * comp * src.sc;
    P.src_f32 = src.es == 4;
    P.src_nd = src.n[d];
    P.src_nf = src.nf;
    P.dst = static_cast<char*>(dst.p) + (size_t)dst.es * comp * dst.sc;
    P.dst_f32 = dst.es == 4;
    P.dst_nd = dst.n[d];
    P.dst_nf = dst.nf;
    for (int i = 0; i < 3; ++i) { P.src_s[i] = src.s[i]; P.dst_s[i] = dst.s[i]; }
    P.n1 = tin.n[(d + 1) % 3];
    P.n2 = tin.n[(d + 2) % 3];
    // pencil grid: the other dim with the smaller source stride varies fastest between pencils,
    // so that neighbouring pencils of a block are neighbours in memory
    {
        const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        const bool swap = src.s[d2] < src.s[d1];
        P.fa = swap ? d2 : d1;
        P.fb = swap ? d1 : d2;
        P.na = swap ? P.n2 : P.n1;
    }
    // contiguous real-to-real pencils of a power-of-two length: the row r2r kernel
    if (!std::getenv("FLUPS_GENERIC_PASS") && row_r2r_fits(P)) {
        P.fastk = 1;
        int pf = 1, q = r2r_h(P.kind, P.n);
        split_pq(r2r_h(P.kind, P.n), pf, q);
        P.tw = twiddles(q);
        P.aux = r2r_table(P.kind, P.n);
        const int fa = P.fa, fb = P.fb;
        const bool plain = P.gshift == 0 && P.sshift == 0 && P.src_nd == P.n && P.dst_nd == P.n && P.pen_in == P.n &&
                           P.si == 0 && P.so == 0 && P.post == 0 && P.kind != REDFT00 && P.kind != RODFT00;
        // pairs of reals as one vector access: 16-byte (double) or 8-byte (float) aligned
        P.vec = plain && ((uintptr_t)P.src % (2 * src.es) == 0) && ((uintptr_t)P.dst % (2 * dst.es) == 0) && P.src_s[fa] % 2 == 0 &&
                P.src_s[fb] % 2 == 0 && P.dst_s[fa] % 2 == 0 && P.dst_s[fb] % 2 == 0;
        P.nt = P.vec && view_bytes(src) >= kNtBytes && view_bytes(dst) >= kNtBytes;
        return P;
    }
    // strided real-to-real pencils, neighbours along a unit-stride dimension: the same transforms
    // on tiles of neighbouring pencils
    if (!std::getenv("FLUPS_GENERIC_PASS") && col_r2r_fits(P)) {
        P.fastk = 4;
        int pf = 1, q = r2r_h(P.kind, P.n);
        split_pq(r2r_h(P.kind, P.n), pf, q);
        P.tw = twiddles(q);
        P.aux = r2r_table(P.kind, P.n);
        return P;
    }
    // contiguous complex pencils (x after an r2c direction): staged rows, thread teams
    if (!std::getenv("FLUPS_GENERIC_PASS") && row_c2c_fits(P)) {
        P.fastk = 6;
        int pf = 1, q = P.n;
        split_pq(P.n, pf, q);
        P.tw = twiddles(q);
        P.aux = twiddles(P.n);
        return P;
    }
    // strided complex pencils: tiles of neighbouring pencils through the register FFT
    // (lengths 3, 5, 7 times a power of two: Q-point column FFTs after the small DFTs)
    if (!std::getenv("FLUPS_GENERIC_PASS") && col_c2c_fits(P)) {
        P.fastk = 5;
        int pf = 1, q = P.n;
        split_pq(P.n, pf, q);
        P.tw = twiddles(q);
        P.aux = twiddles(P.n);
        return P;
    }
    // real-to-complex pencils along the contiguous direction: packed half-length FFT in registers
    if (!std::getenv("FLUPS_GENERIC_PASS") && row_rc_fits(P)) {
        P.fastk = 3;
        int pf = 1, q = P.n / 2;
        split_pq(P.n / 2, pf, q);
        P.tw = twiddles(q);
        P.aux = twiddles(P.n);
        return P;
    }
    // real-to-complex pencils along a strided direction, neighbours along a unit-stride one:
    // pairs of real pencils as one complex column
    if (!std::getenv("FLUPS_GENERIC_PASS") && col_real_fits(P)) {
        P.fastk = 2;
        int pf = 1, q = P.n;
        split_pq(P.n, pf, q);
        P.tw = twiddles(q);
        P.aux = twiddles(P.n);
        const bool fwd = P.op == OP_R2C;
        const void* real = fwd ? P.src : P.dst;
        const long long* rs = fwd ? P.src_s : P.dst_s;
        const int res = fwd ? src.es : dst.es;
        P.vec = ((uintptr_t)real % (2 * res) == 0) && rs[P.d] % 2 == 0 && rs[P.fb] % 2 == 0 && P.na % 2 == 0;
        return P;
    }
    P.alias = generic_alias(P) ? 1 : 0;
    const size_t per_pen = generic_lds_bytes(P, 1);
    FLUPS_REQUIRE(per_pen <= lds_budget_, "pencil of %d elements (transform length %d) does not fit the general path's LDS",
                  P.pen_in, M);
    // several blocks per CU: cap the LDS of one block
    const size_t cap = std::max(per_pen, std::min(lds_budget_, generic_lds_cap_));
    int ppb = (int)std::min<size_t>(16, cap / std::max<size_t>(per_pen, 1));
    while (ppb > 1 && generic_lds_bytes(P, ppb) > cap) --ppb;
    // small problems: fewer pencils per block, so that about 1024 blocks share the 256 CUs
    ppb = std::min(ppb, std::max(1, (P.n1 * P.n2 + 1023) / 1024));
    P.ppb = std::max(1, std::min(ppb, P.n1 * P.n2));
    return P;
}

void SolverImpl::run_pass(PassDesc P) {
    if (P.fastk == 1) launch_row_r2r(P, stream_);
    else if (P.fastk == 2) launch_col_real(P, stream_);
    else if (P.fastk == 3) launch_row_rc(P, stream_);
    else if (P.fastk == 4) launch_col_r2r(P, stream_);
    else if (P.fastk == 5) launch_col_c2c(P, stream_);
    else if (P.fastk == 6) launch_row_c2c(P, stream_);
    else launch_generic(P, stream_);
}

// ---------------------------------------------------------------------------
void SolverImpl::fill_green_args(GreenArgs& A, const double hfact[3], const double kfact[3], const double koffset[3],
                                 const double symstart[3], double length, int nunb) {
    A.type = gtype_;
    A.length = length;
    A.hgrid = h_[0];
    for (int d = 0; d < 3; ++d) {
        A.hfact[d] = hfact[d];
        A.kfact[d] = kfact[d];
        A.koffset[d] = koffset[d];
        A.symstart[d] = symstart[d];
    }
    if (nunb == 3) {
        if (gtype_ == CHAT_2) {
            const double vol = hfact[0] * hfact[1] * hfact[2];
            A.g0 = -0.5 * std::pow(1.5 / (2.0 * M_PI) * vol, 2.0 / 3.0) / vol;
        } else if (gtype_ >= HEJ_2 && gtype_ <= HEJ_10) {
            static const double mult[5] = {1.0 / 4.0, 3.0 / 8.0, 15.0 / 32.0, 35.0 / 64.0, 315.0 / 512.0};
            A.g0 = -M_SQRT2 * mult[gtype_ - HEJ_2] / (length * std::sqrt(M_PI * M_PI * M_PI));
        } else if (gtype_ == HEJ_0) {
            A.g0 = -1.0 / (2.0 * M_PI * M_PI * length);   // src/green_functions.cpp:66-68
        }
    }
    // stencil coefficients of the 1-D lattice Green's functions: s_m (symbol s_0 + 2 sum s_m cos(m t))
    // and p_j, the coefficients of s(t) in powers of mu = cos(t) - 1 (Chebyshev T_m(1 + mu))
    static const double S[4][5] = {{2.0, -1.0, 0, 0, 0},
                                   {5.0 / 2.0, -4.0 / 3.0, 1.0 / 12.0, 0, 0},
                                   {49.0 / 18.0, -3.0 / 2.0, 3.0 / 20.0, -1.0 / 90.0, 0},
                                   {205.0 / 72.0, -8.0 / 5.0, 1.0 / 5.0, -8.0 / 315.0, 1.0 / 560.0}};
    for (int o = 0; o < 4; ++o) {
        const int w = o + 1;
        std::vector<std::vector<double>> T(w + 1, std::vector<double>(w + 1, 0.0));
        T[0][0] = 1.0;
        if (w >= 1) { T[1][0] = 1.0; T[1][1] = 1.0; }
        for (int m = 1; m < w; ++m)   // T_{m+1} = 2 (1 + mu) T_m - T_{m-1}
            for (int j = 0; j <= w; ++j) {
                double v = 2.0 * T[m][j] - T[m - 1][j];
                if (j > 0) v += 2.0 * T[m][j - 1];
                T[m + 1][j] = v;
            }
        for (int j = 0; j <= w; ++j) {
            double pj = (j == 0) ? S[o][0] : 0.0;
            for (int m = 1; m <= w; ++m) pj += 2.0 * S[o][m] * T[m][j];
            A.lgf_p[5 * o + j] = (j == 0) ? 0.0 : pj;   // p_0 = s(0) = 0 exactly
            A.lgf_s[5 * o + j] = S[o][j];
        }
    }
}

namespace {
// Ji0c(x) = int_0^x (1 - J0(u)) / u du = -1/2 sum_{n=1}^{50} (-x^2/4)^n / (n!^2 n): long double,
// summed from n = 50 down, 1/n!^2 from the running product (src/ji0.hpp:23-41)
double ji0c_series(double x) {
    long double c[51];
    long double f = 1.0L;
    c[0] = 1.0L;
    for (int n = 1; n <= 50; ++n) {
        f *= (long double)n;
        c[n] = 1.0L / (f * f);
    }
    const long double t = (long double)(-(0.25 * x * x));
    long double val = 0.0L;
    for (int n = 50; n > 0; --n) val -= std::pow(t, (long double)n) * c[n] / (long double)n * 0.5L;
    return (double)val;
}

std::vector<double> read_table(const std::string& path, size_t n) {
    std::vector<double> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    FLUPS_REQUIRE(f != nullptr, "unable to read Green's table %s", path.c_str());
    const size_t got = std::fread(v.data(), sizeof(double), n, f);
    std::fclose(f);
    FLUPS_REQUIRE(got == n, "short Green's table %s", path.c_str());
    return v;
}
std::vector<ExpTerm> read_expansion(const std::string& path, const std::string& name) {
    std::ifstream f(path);
    FLUPS_REQUIRE((bool)f, "unable to read %s", path.c_str());
    std::vector<ExpTerm> out;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string nm;
        ExpTerm t;
        is >> nm >> t.a >> t.b >> t.c >> t.p >> t.l >> t.coef;
        if (nm == name) out.push_back(t);
    }
    FLUPS_REQUIRE(!out.empty(), "no expansion %s in %s", name.c_str(), path.c_str());
    return out;
}
}  // namespace

void SolverImpl::compute_green() {
    bool spec[3] = {false, false, false};
    double hfact[3] = {0, 0, 0}, kfact[3] = {0, 0, 0}, koffset[3] = {0, 0, 0}, symstart[3] = {0, 0, 0};
    double length = alpha_ * h_[0];
    if (gtype_ == HEJ_0) length = h_[0] / M_PI;
    const bool iso = (gtype_ >= HEJ_2 && gtype_ <= MEHR_6F) || gtype_ == LGF_2;   // regularised, LGF and Mehrstellen
    FLUPS_REQUIRE(!(iso && ((ndim_ == 3 && (h_[0] != h_[1] || h_[1] != h_[2])) || (ndim_ == 2 && h_[0] != h_[1]))),
                  "a regularized kernel or an LGF requires dx = dy = dz");
    FLUPS_REQUIRE(gtype_ >= CHAT_2 && gtype_ <= MEHR_6F, "unknown Green's kernel %d", gtype_);
    for (int ip = 0; ip < ndim_; ++ip) {
        const DirPlan& p = gpl_[ip];
        const int d = p.dim;
        spec[d] = p.spectral;
        symstart[d] = p.symstart;
        hfact[d] = h_[d];
        if (p.spectral) { hfact[d] = 0.0; kfact[d] = p.kfact; koffset[d] = p.koffset; }
        if (p.type == EMPTY) hfact[d] = 0.0;
    }
    int nspec = 0;
    for (int d = 0; d < ndim_; ++d) if (spec[d]) ++nspec;
    const int nunb = ndim_ - nspec;
    const bool mehr = gtype_ >= MEHR_4L && gtype_ <= MEHR_6F;
    // src/green_functions.cpp:444-446 (HEJ_0, one unbounded direction: unavailable in the reference too)
    FLUPS_REQUIRE(!(gtype_ == HEJ_0 && nunb == 1), "HEJ_0 kernel not available for 1D unbounded problems");

    GreenArgs A;
    fill_green_args(A, hfact, kfact, koffset, symstart, length, nunb);
    const int ord = (gtype_ == LGF_2) ? 2 : (gtype_ == LGF_4) ? 4 : (gtype_ == LGF_6) ? 6 : (gtype_ == LGF_8) ? 8 : 0;
    double* dtab = nullptr;
    ExpTerm* dexp = nullptr;
    if ((ord || mehr) && (nunb == 3 || nunb == 2)) {
        // table near the origin + far-field expansion (src/green_functions.hpp:30-92 for the file names)
        const int gd = nunb;
        char name[256], ename[64];
        if (ord) {
            std::snprintf(name, sizeof(name), "%s/LGF_%d_%dd_32.ker", kdir_.c_str(), ord, gd);
            std::snprintf(ename, sizeof(ename), "lgf_%d_%dunb0spe", ord, gd);
        } else {
            const char* up = gtype_ == MEHR_4L ? "4L" : gtype_ == MEHR_6L ? "6L" : gtype_ == MEHR_4F ? "4F" : "6F";
            const char* lo = gtype_ == MEHR_4L ? "4l" : gtype_ == MEHR_6L ? "6l" : gtype_ == MEHR_4F ? "4f" : "6f";
            const bool lump = (gtype_ == MEHR_4L || gtype_ == MEHR_6L) && gd == 2;   // 4L and 6L share the 2-D table
            std::snprintf(name, sizeof(name), "%s/MEHR_%s_%dd_32.ker", kdir_.c_str(), lump ? "4L6L" : up, gd);
            std::snprintf(ename, sizeof(ename), "mehr_%s_%dunb0spe", lump ? "4l6l" : lo, gd);
        }
        const auto tab = read_table(name, gd == 3 ? 32 * 32 * 32 : 32 * 32);
        const auto ex = read_expansion(kdir_ + "/lgf_expansion.txt", ename);
        HIP_CHECK(hipMalloc(&dtab, tab.size() * sizeof(double)));
        HIP_CHECK(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
        HIP_CHECK(hipMalloc(&dexp, ex.size() * sizeof(ExpTerm)));
        HIP_CHECK(hipMemcpyAsync(dexp, ex.data(), ex.size() * sizeof(ExpTerm), hipMemcpyHostToDevice, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        A.table = dtab;
        A.exp = dexp;
        A.nexp = (int)ex.size();
    }
    if (gtype_ == HEJ_0 && nunb == 2) {
        // Ji0c at rho = pi sqrt(s) for the integer squared distances s of the isotropic grid
        // with rho < 30 (s <= 91), summed on the host in long double as src/ji0.hpp:23-41 does
        // (its cancellation near rho = 30 leaves about 1e-10 relative error; the same summation
        // keeps the Green's function equal to the reference's)
        std::vector<double> tab(92);
        for (int q = 0; q < 92; ++q) tab[q] = ji0c_series(std::sqrt((double)q) * h_[0] / length);
        HIP_CHECK(hipMalloc(&dtab, tab.size() * sizeof(double)));
        HIP_CHECK(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        A.table = dtab;
        A.ntable = (int)tab.size();
    }
    // fill on the first Green's topology, then forward transforms of the non-spectral directions
    double* a = nullptr;
    size_t cap = 0;
    for (int ip = 0; ip < ndim_; ++ip) cap = std::max(cap, std::max(gpre_[ip].count(), gpost_[ip].count()));
    HIP_CHECK(hipMalloc(&a, cap * sizeof(double)));
    HIP_CHECK(hipMemsetAsync(a, 0, cap * sizeof(double), stream_));
    launch_green_fill(nunb, a, gpre_[0].n, gpre_[0].nf, gpl_[0].dim, A, nullptr, stream_);
    double* b = nullptr;
    HIP_CHECK(hipMalloc(&b, cap * sizeof(double)));
    for (int ip = 0; ip < ndim_; ++ip) {
        const DirPlan& p = gpl_[ip];
        if (spec[p.dim]) continue;
        View vs = dense_view(a, gpre_[ip]);
        View vd = dense_view(b, gpost_[ip]);
        PassDesc P = pass_desc(p, 0, true, vs, vd, gpre_[ip], gpost_[ip], 0, 0);
        run_pass(P);
        std::swap(a, b);
    }
    const Shape& gfin = gpost_[ndim_ - 1];
    launch_scale(a, gfin.count(), volfact_, stream_);
    const bool hej = gtype_ >= HEJ_2 && gtype_ <= HEJ_10;
    const bool lgf = ord != 0;
    if (ndim_ == 3 && nspec == 1 && (hej || lgf || mehr)) {
        // 2 unbounded + 1 spectral: the spectral kernel away from the k = 0 plane (src/Solver.cpp:1092-1101)
        int lo[3];
        for (int ip = 0; ip < 3; ++ip) {
            lo[ip] = spec[ip] ? 1 : 0;
            kfact[gpl_[ip].dim] = gpl_[ip].kfact;
        }
        GreenArgs B = A;
        for (int d = 0; d < 3; ++d) B.kfact[d] = kfact[d];
        launch_green_fill(0, a, gfin.n, gfin.nf, gpl_[ndim_ - 1].dim, B, lo, stream_);
    }
    // the field's spectral stage must have the same extents
    const Shape& fs = stage_[ndim_ - 1];
    FLUPS_REQUIRE(fs.nf == gfin.nf && fs.n[0] == gfin.n[0] && fs.n[1] == gfin.n[1] && fs.n[2] == gfin.n[2],
                  "Green's function and field spectral layouts differ");
    HIP_CHECK(hipStreamSynchronize(stream_));
    HIP_CHECK(hipFree(b));
    if (dtab) HIP_CHECK(hipFree(dtab));
    if (dexp) HIP_CHECK(hipFree(dexp));
    dgreen_ = a;
}

void SolverImpl::setup() {
    FLUPS_REQUIRE(!setup_done_, "setup was already done");
    const auto t_setup = std::chrono::steady_clock::now();
    if (const char* e = std::getenv("FLUPS_GRAPH")) use_graph_ = std::atoi(e) != 0;
    HIP_CHECK(hipEventCreate(&ev0_));
    HIP_CHECK(hipEventCreate(&ev1_));
    compute_green();
    if (prof_) {
        HIP_CHECK(hipStreamSynchronize(stream_));
        prof_->add("setup/green", std::chrono::duration<double>(std::chrono::steady_clock::now() - t_setup).count(), 0.0);
    }
    fast_ = make_fast_plan(*this);
    if (fast_ && prof_) fast_profile(fast_, true);   // per-kernel events of the fused pipeline
    const char* at_env = std::getenv("FLUPS_AUTOTUNE");
    if (fast_ && (autotune_ || (at_env && std::atoi(at_env) != 0))) {
        // time the candidate schedules on scratch arrays of the caller's layout
        const size_t n = phys_.memsize();
        double *din = nullptr, *dout = nullptr;
        HIP_CHECK(hipMalloc(&din, n * sizeof(double)));
        HIP_CHECK(hipMalloc(&dout, n * sizeof(double)));
        HIP_CHECK(hipMemsetAsync(din, 0, n * sizeof(double), stream_));
        try {
            fast_autotune(fast_, dout, din, user_view(din), user_view(dout), dgreen_, normfact_, stream_);
        } catch (...) {
            (void)hipFree(din);
            (void)hipFree(dout);
            throw;
        }
        HIP_CHECK(hipFree(din));
        HIP_CHECK(hipFree(dout));
    }
    if (fast_ && lda_ != 3) {   // the fused pipeline keeps its own compressed copy; ROT needs lda = 3
        HIP_CHECK(hipFree(dgreen_));
        dgreen_ = nullptr;
    }
    // FP32 requests the fused pipeline cannot run: the general path with FP32 stage arrays and an
    // FP32 copy of the Green's function (FP64 arithmetic in the kernels); the FP64 function is
    // released once the fused last direction has taken its compressed copy
    gf32_ = fp32_ && !fast_;
    ses_ = gf32_ ? 4 : 8;
    if (gf32_) {
        const size_t gn = gpost_[ndim_ - 1].count();
        HIP_CHECK(hipMalloc(&dgreen32_, gn * sizeof(float)));
        launch_to_float(dgreen_, dgreen32_, gn, stream_);
    }
    if (!fast_) hybrid_ = make_hybrid_z(*this);
    if (gf32_) {
        HIP_CHECK(hipStreamSynchronize(stream_));
        HIP_CHECK(hipFree(dgreen_));
        dgreen_ = nullptr;
    }
    if (!fast_) {
        for (int ip = 0; ip < ndim_; ++ip) {
            const size_t n = std::max(hat_[ip].count(), stage_[ip].count());
            HIP_CHECK(hipMalloc(&dstage_[ip], n * ses_));
        }
        warm_tables();
        warmed_ = true;
    }
    scratch_init();   // this device's static scratch address, for flups_solve_noalloc
    HIP_CHECK(hipStreamSynchronize(stream_));
    setup_done_ = true;
    if (prof_) prof_->add("setup", std::chrono::duration<double>(std::chrono::steady_clock::now() - t_setup).count(), 0.0);
}

// ---------------------------------------------------------------------------
// forward transforms: user layout (uin) -> dense spectral array `spec`
void SolverImpl::forward_passes(const View& uin, void* spec) {
    for (int ip = 0; ip < ndim_; ++ip) {
        const DirPlan& p = fwd_[ip];
        const View src = (ip == 0) ? uin : dense_view(dstage_[ip - 1], stage_[ip - 1], ses_);
        const View dst = dense_view(ip == ndim_ - 1 ? spec : dstage_[ip], stage_[ip], ses_);
        SameStage same;   // the components' launches are independent
        for (int c = 0; c < lda_; ++c) run_pass(pass_desc(p, c, true, src, dst, hat_[ip], stage_[ip], p.fieldstart, 0));
    }
}

// multiply the dense spectral array by normfact * G (STD) or take the curl of G * data (ROT)
void SolverImpl::multiply(void* spec, int type) {
    const Shape& sp = stage_[ndim_ - 1];
    const void* g = gf32_ ? (const void*)dgreen32_ : (const void*)dgreen_;
    if (type == 0) {
        launch_mult_std(spec, g, (size_t)sp.n[0] * sp.n[1] * sp.n[2], sp.nf, lda_, normfact_, gf32_, stream_);
        return;
    }
    RotArgs R;
    for (int d = 0; d < 3; ++d) { R.n[d] = sp.n[d]; R.h[d] = h_[d]; }
    R.nf = sp.nf;
    R.order = odiff_;
    R.normfact = normfact_;
    double kabs[3];
    spectral_info(kabs, R.koffset, R.symstart);
    for (int ip = 0; ip < 3; ++ip) {
        const int d = fwd_[ip].dim;
        for (int c = 0; c < 3; ++c) {
            const int ph = (bwdd_[ip].imult[c] ? 1 : 0) - (fwd_[ip].imult[c] ? 1 : 0);
            R.kf[d][c][0] = ph == 0 ? 0.0 : (ph == 1 ? -kabs[d] : kabs[d]);
            R.kf[d][c][1] = ph == 0 ? kabs[d] : 0.0;
        }
    }
    launch_mult_rot(spec, g, R, gf32_, stream_);
}

// backward transforms: dense spectral array `spec` -> user layout (uout); `diff`
// selects the plans of the differentiated field (after a ROT multiply)
void SolverImpl::backward_passes(const void* spec, const View& uout, bool diff) {
    const DirPlan* back = diff ? bwdd_ : bwd_;
    for (int ip = ndim_ - 1; ip >= 0; --ip) {
        const DirPlan& p = back[ip];
        const View src = dense_view(ip == ndim_ - 1 ? const_cast<void*>(spec) : dstage_[ip], stage_[ip], ses_);
        const View dst = (ip == 0) ? uout : dense_view(dstage_[ip - 1], stage_[ip - 1], ses_);
        SameStage same;
        for (int c = 0; c < lda_; ++c)
            run_pass(pass_desc(p, c, false, src, dst, stage_[ip], hat_[ip], 0, fwd_[ip].fieldstart));
    }
}

void SolverImpl::solve_generic(double* field, const double* rhs, int type) {
    if (hybrid_ && type == 0) {
        // general passes for the first directions, the last one fused with the multiply
        const View uin = user_view(const_cast<double*>(rhs)), uout = user_view(field);
        for (int ip = 0; ip < ndim_ - 1; ++ip) {
            const DirPlan& p = fwd_[ip];
            const View src = (ip == 0) ? uin : dense_view(dstage_[ip - 1], stage_[ip - 1], ses_);
            const View dst = dense_view(dstage_[ip], stage_[ip], ses_);
            SameStage same;
            for (int c = 0; c < lda_; ++c) run_pass(pass_desc(p, c, true, src, dst, hat_[ip], stage_[ip], p.fieldstart, 0));
        }
        pmark("fftw forward");
        hybrid_z(hybrid_, dstage_[ndim_ - 2], stream_);
        pmark("fftw + domagic (last direction fused)");
        for (int ip = ndim_ - 2; ip >= 0; --ip) {
            const DirPlan& p = bwd_[ip];
            const View src = dense_view(dstage_[ip], stage_[ip], ses_);
            const View dst = (ip == 0) ? uout : dense_view(dstage_[ip - 1], stage_[ip - 1], ses_);
            SameStage same;
            for (int c = 0; c < lda_; ++c)
                run_pass(pass_desc(p, c, false, src, dst, stage_[ip], hat_[ip], 0, fwd_[ip].fieldstart));
        }
        pmark("fftw backward");
        return;
    }
    forward_passes(user_view(const_cast<double*>(rhs)), dstage_[ndim_ - 1]);
    pmark("fftw forward");
    multiply(dstage_[ndim_ - 1], type);
    pmark("domagic");
    backward_passes(dstage_[ndim_ - 1], user_view(field), type != 0);
    pmark("fftw backward");
}

// profiler stages of flups_solve: an event on the solve stream closes each stage; pmark is a
// no-op without a profiler handle (and so under the graph captures of batches and servers)
void SolverImpl::pmark(const char* label) {
    if (!prof_) return;
    if (pn_ == (int)pev_.size()) {
        hipEvent_t e = nullptr;
        HIP_CHECK(hipEventCreate(&e));
        pev_.push_back(e);
        plab_.emplace_back();
    }
    plab_[pn_] = label;
    HIP_CHECK(hipEventRecord(pev_[pn_], stream_));
    ++pn_;
}

void SolverImpl::pflush(const std::string& prefix) {
    for (int i = 1; i < pn_; ++i) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, pev_[i - 1], pev_[i]));
        prof_->add(prefix + plab_[i], 0.0, ms);
    }
    pn_ = 0;
}

// ---------------------------------------------------------------------------
// advanced interface (src/Solver.cpp:1281-1542): the solve's steps one by one on
// device buffers of flups_get_allocSize doubles. The physical layout is the
// user topology; the spectral layout is the dense array of inner_topo_spectral.
void SolverImpl::prepare_advanced() {
    FLUPS_REQUIRE(setup_done_, "flups_setup must be called first");
    FLUPS_REQUIRE(!gf32_, "the advanced interface works on FP64 buffers: this solver was set up for FP32 "
                          "(flups_set_precision(s, 1)) and keeps FP32 stage arrays");
    if (!dstage_[0]) {
        for (int ip = 0; ip < ndim_; ++ip) {
            const size_t m = std::max(hat_[ip].count(), stage_[ip].count());
            HIP_CHECK(hipMalloc(&dstage_[ip], m * ses_));
        }
    }
    if (!dgreen_) compute_green();   // the fused pipeline released the full Green's function
}

static bool device_ptr(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

double* SolverImpl::inner_buffer() {
    FLUPS_REQUIRE(setup_done_, "flups_setup must be called first");
    FLUPS_REQUIRE(!gf32_, "the advanced interface works on FP64 buffers: this solver was set up for FP32 "
                          "(flups_set_precision(s, 1)) and keeps FP32 stage arrays");
    if (!dinner_) {
        HIP_CHECK(hipMalloc(&dinner_, alloc_size() * sizeof(double)));
        HIP_CHECK(hipMemset(dinner_, 0, alloc_size() * sizeof(double)));
    }
    return dinner_;
}

void SolverImpl::do_copy(const Topo* t, double* data, int sign) {
    FLUPS_REQUIRE(data, "data is NULL");
    FLUPS_REQUIRE(t && t->lda == lda_, "the topology's lda must match the solver's (%d)", lda_);
    FLUPS_REQUIRE(t->memsize() == phys_.memsize() && t->axis == phys_.axis,
                  "do_copy works on the solver's physical topology");
    FLUPS_REQUIRE(sign == SIGN_FORWARD || sign == SIGN_BACKWARD, "sign must be SIGN_FORWARD or SIGN_BACKWARD");
    double* own = inner_buffer();
    const size_t n = phys_.memsize() * sizeof(double);
    if (sign == SIGN_FORWARD) HIP_CHECK(hipMemcpyAsync(own, data, n, hipMemcpyDefault, stream_));
    else HIP_CHECK(hipMemcpyAsync(data, own, n, hipMemcpyDefault, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
}

void SolverImpl::do_fft(double* data, int sign) {
    FLUPS_REQUIRE(data, "data is NULL");
    FLUPS_REQUIRE(device_ptr(data), "do_FFT works on device buffers (flups_get_innerBuffer)");
    FLUPS_REQUIRE(sign == SIGN_FORWARD || sign == SIGN_BACKWARD || sign == SIGN_BACKWARD_DIFF, "unknown sign %d", sign);
    prepare_advanced();
    // in place: the first pass reads `data` entirely before the last one writes it (ndim >= 2)
    if (sign == SIGN_FORWARD) forward_passes(user_view(data), data);
    else backward_passes(data, user_view(data), sign == SIGN_BACKWARD_DIFF);
    HIP_CHECK(hipStreamSynchronize(stream_));
}

void SolverImpl::do_mult(double* data, int type) {
    FLUPS_REQUIRE(data, "data is NULL");
    FLUPS_REQUIRE(device_ptr(data), "do_mult works on device buffers (flups_get_innerBuffer)");
    check_type(type);
    prepare_advanced();
    multiply(data, type);
    HIP_CHECK(hipStreamSynchronize(stream_));
}

Topo SolverImpl::inner_topo_spectral() const {
    const Shape& sp = stage_[ndim_ - 1];
    Topo t;
    t.axis = 0;
    t.lda = lda_;
    t.nf = sp.nf;
    t.alignment = sizeof(double);
    for (int d = 0; d < 3; ++d) t.nglob[d] = sp.n[d];
    t.cmpt_sizes();
    return t;
}

void SolverImpl::check_type(int type) const {
    FLUPS_REQUIRE(setup_done_, "flups_setup must be called before flups_solve");
    FLUPS_REQUIRE(type == 0 || type == 1, "unknown solver type %d", type);
    FLUPS_REQUIRE(type == 0 || lda_ == 3, "the ROT solver needs lda = 3 (vectors)");
    FLUPS_REQUIRE(type == 0 || odiff_ != NOD, "the ROT solver needs a differentiation order (SPE, FD2, FD4, FD6)");
}

// buffers a solve of this type needs beyond those of setup (the general path's stage
// arrays when the fused pipeline cannot run it)
void SolverImpl::prepare(int type) {
    check_type(type);
    if (fast_ && (type == 0 || lda_ == 3)) return;
    if (!dstage_[0]) {
        for (int ip = 0; ip < ndim_; ++ip) {
            const size_t m = std::max(hat_[ip].count(), stage_[ip].count());
            HIP_CHECK(hipMalloc(&dstage_[ip], m * ses_));
        }
    }
    if (!warmed_) {
        warm_tables();
        warmed_ = true;
    }
}

// build every pass descriptor once, with the views the solve uses (stage arrays; the
// user layout for the first and last pass), so that the tables of the kernels they
// select exist before a solve is queued (a solve under graph capture may not allocate)
void SolverImpl::warm_tables() {
    const View u = user_view(static_cast<double*>(dstage_[0]));
    std::vector<int> fk(2 * ndim_, 0);
    for (int ip = 0; ip < ndim_; ++ip) {
        const View prev = (ip == 0) ? u : dense_view(dstage_[ip - 1], stage_[ip - 1], ses_);
        const View cur = dense_view(dstage_[ip], stage_[ip], ses_);
        for (int c = 0; c < lda_; ++c) {
            const PassDesc f = pass_desc(fwd_[ip], c, true, prev, cur, hat_[ip], stage_[ip], fwd_[ip].fieldstart, 0);
            const PassDesc b = pass_desc(bwd_[ip], c, false, cur, prev, stage_[ip], hat_[ip], 0, fwd_[ip].fieldstart);
            if (odiff_ != NOD) (void)pass_desc(bwdd_[ip], c, false, cur, prev, stage_[ip], hat_[ip], 0, fwd_[ip].fieldstart);
            if (c == 0) {
                const bool fused = hybrid_ && ip == ndim_ - 1;   // the last direction runs with the multiply
                fk[ip] = fused ? -2 : f.op == OP_NONE ? -1 : f.fastk;
                fk[ndim_ + ip] = fused ? -2 : b.op == OP_NONE ? -1 : b.fastk;
            }
        }
    }
    pass_kernels_ = fk;
}

// queue the solve's kernels on the solver's stream; field and rhs are device
// pointers, nothing is allocated or synchronised (usable under graph capture
// once prepare(type) has run)
void SolverImpl::enqueue(double* field, const double* rhs, int type) {
    if (fast_ && (type == 0 || lda_ == 3))
        fast_solve(fast_, field, rhs, user_view(const_cast<double*>(rhs)), user_view(field), stream_, type);
    else {
        FLUPS_REQUIRE(dstage_[0] && (dgreen_ || dgreen32_), "solver not prepared for solve type %d", type);
        solve_generic(field, rhs, type);
    }
}

// flups_solve_noalloc: the solve of `enqueue` with its stage arrays carved from the device's
// static scratch instead of the solver's own; field and rhs are device pointers; nothing is
// allocated, freed or synchronised (the tables were built at setup; a missing one is an
// error, not an allocation)
void SolverImpl::solve_noalloc(double* field, const double* rhs, int type) {
    check_type(type);
    FLUPS_REQUIRE(field && rhs, "field and rhs must not be NULL");
    hipPointerAttribute_t at;
    for (const void* p : {(const void*)field, (const void*)rhs}) {
        const bool dev = hipPointerGetAttributes(&at, p) == hipSuccess && at.type == hipMemoryTypeDevice;
        if (!dev) (void)hipGetLastError();
        FLUPS_REQUIRE(dev, "flups_solve_noalloc takes device pointers only (host arrays would need staging buffers)");
    }
    size_t cap = 0;
    double* base = scratch_device_ptr(&cap);
    FLUPS_REQUIRE(base, "no static scratch on this device: flups_setup of this solver ran on another device");
    const bool fused = fast_ && (type == 0 || lda_ == 3);
    size_t need = 0;
    size_t off[3] = {0, 0, 0};
    if (fused) {
        size_t b0 = 0, b1 = 0;
        fast_stage_bytes(fast_, &b0, &b1);
        off[1] = b0;
        need = b0 + b1;
    } else {
        for (int ip = 0; ip < ndim_; ++ip) {
            off[ip] = need;
            need += ((size_t)ses_ * std::max(hat_[ip].count(), stage_[ip].count()) + 255) / 256 * 256;
        }
    }
    FLUPS_REQUIRE(need <= cap, "this solve needs %zu bytes of stage arrays; the static scratch of flups_solve_noalloc holds %zu",
                  need, cap);
    char* sb = reinterpret_cast<char*>(base);
    if (fused) {
        fast_solve_on(fast_, sb + off[0], sb + off[1], field, rhs, user_view(const_cast<double*>(rhs)), user_view(field),
                      stream_, type);
        return;
    }
    FLUPS_REQUIRE(dgreen_ || dgreen32_, "solver not set up for solve type %d", type);
    void* keep[3] = {dstage_[0], dstage_[1], dstage_[2]};
    Timers* keep_prof = prof_;
    for (int ip = 0; ip < 3; ++ip) dstage_[ip] = ip < ndim_ ? static_cast<void*>(sb + off[ip]) : nullptr;
    prof_ = nullptr;   // profiler stages create events: not in a call that may not allocate
    frozen_ = true;
    auto restore = [&]() {
        for (int ip = 0; ip < 3; ++ip) dstage_[ip] = keep[ip];
        prof_ = keep_prof;
        frozen_ = false;
    };
    try {
        solve_generic(field, rhs, type);
    } catch (...) {
        restore();
        throw;
    }
    restore();
}

double SolverImpl::last_solve_ns() {
    if (last_pending_) {   // an asynchronous solve: its end event, once the stream reached it
        HIP_CHECK(hipEventSynchronize(ev1_));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
        last_ns_ = ms * 1e6;
        last_pending_ = false;
    }
    return last_ns_;
}

size_t SolverImpl::device_bytes() const {
    size_t b = 0;
    for (const auto& kv : tw_) b += 2 * sizeof(double) * (size_t)kv.first;
    for (const auto& kv : r2r_tab_) b += 4 * sizeof(double) * (size_t)(kv.first.second / 2 + 1);
    for (int ip = 0; ip < 3; ++ip)
        if (dstage_[ip]) b += (size_t)ses_ * std::max(hat_[ip].count(), stage_[ip].count());
    if (dgreen_) b += sizeof(double) * gpost_[ndim_ - 1].count();
    if (dgreen32_) b += sizeof(float) * gpost_[ndim_ - 1].count();
    b += sizeof(double) * staging_n_;
    if (fast_) b += fast_device_bytes(fast_);
    if (hybrid_) b += hybrid_device_bytes(hybrid_);
    return b;
}

// upper bound of the device memory setup + a solve of this type will hold (setup's
// transient Green's buffers included), for admission decisions before setup
size_t SolverImpl::predicted_device_bytes(int type) const {
    size_t cap = 0, stages = 0;
    for (int ip = 0; ip < ndim_; ++ip) {
        cap = std::max(cap, std::max(gpre_[ip].count(), gpost_[ip].count()));
        stages += std::max(hat_[ip].count(), stage_[ip].count());
    }
    const size_t setup_peak = 2 * cap * sizeof(double);
    const size_t steady = (gpost_[ndim_ - 1].count() + stages) * sizeof(double);
    (void)type;
    return std::max(setup_peak, steady) + (size_t)(4u << 20);   // + twiddle tables, LGF tables, slack
}

void SolverImpl::solve(double* field, const double* rhs, int type) {
    FLUPS_REQUIRE(field && rhs, "field and rhs must not be NULL");
    prepare(type);
    hipPointerAttribute_t at_f, at_r;
    auto is_dev = [](const void* p, hipPointerAttribute_t* a) {
        if (hipPointerGetAttributes(a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
        return a->type == hipMemoryTypeDevice;
    };
    const bool fdev = is_dev(field, &at_f), rdev = is_dev(rhs, &at_r);
    const size_t n = phys_.memsize();
    double* din = const_cast<double*>(rhs);
    double* dout = field;
    const auto t0 = std::chrono::steady_clock::now();
    pn_ = 0;
    pmark("start");
    if (!fdev || !rdev) {
        if (staging_n_ < 2 * n) {
            if (dstaging_) HIP_CHECK(hipFree(dstaging_));
            HIP_CHECK(hipMalloc(&dstaging_, 2 * n * sizeof(double)));
            staging_n_ = 2 * n;
        }
        if (!rdev) {
            din = dstaging_;
            HIP_CHECK(hipMemcpyAsync(din, rhs, n * sizeof(double), hipMemcpyHostToDevice, stream_));
        }
        if (!fdev) dout = dstaging_ + n;
    }
    if (din != rhs) pmark("copy rhs");
    HIP_CHECK(hipEventRecord(ev0_, stream_));
    if (use_graph_ && !profiling_ && !prof_) launch_graph(dout, din, type);
    else enqueue(dout, din, type);
    HIP_CHECK(hipEventRecord(ev1_, stream_));
    const bool fused = fast_ && (type == 0 || lda_ == 3);
    if (fused) pmark("fused pipeline (fftw + domagic)");
    if (!fdev) HIP_CHECK(hipMemcpyAsync(field, dout, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (!fdev) pmark("copy field");
    if (async_ && fdev && rdev && !prof_ && !(fused && fast_profiling(fast_))) {
        // the caller gave a stream (flups_set_stream): return with the solve queued on it
        last_pending_ = true;
        return;
    }
    HIP_CHECK(hipStreamSynchronize(stream_));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
    last_ns_ = ms * 1e6;
    last_pending_ = false;
    if (prof_) {
        // the whole solve: device time from the first to the last mark, host wall-clock
        float all = 0;
        if (pn_ > 1) HIP_CHECK(hipEventElapsedTime(&all, pev_[0], pev_[pn_ - 1]));
        prof_->add("solve", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), all);
        pflush("solve/");
    }
    if (fused) {
        double before[5] = {0, 0, 0, 0, 0};
        const char* names[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        if (prof_)
            for (int i = 0; i < 5; ++i) {
                long c = 0;
                double b = 0;
                fast_profile_get(fast_, i, &names[i], &before[i], &c, &b);
            }
        fast_collect(fast_);
        if (prof_)
            for (int i = 0; i < 5; ++i) {
                long c = 0;
                double b = 0, after = 0;
                const char* nm = nullptr;
                if (fast_profile_get(fast_, i, &nm, &after, &c, &b) && after > before[i])
                    prof_->add(std::string("solve/fused pipeline (fftw + domagic)/") + nm, 0.0, after - before[i]);
            }
    }
}

// The solve's kernels are recorded once per (field, rhs, type) into a hipGraph and
// replayed with one launch: the same kernels with the same arguments, fewer
// host launches. The capture runs on a private stream (the default stream
// cannot be captured); the replay goes on the solver's stream.
void SolverImpl::launch_graph(double* field, const double* rhs, int type) {
    if (!(gexec_ && gfield_ == field && grhs_ == rhs && gtype_key_ == type)) {
        if (gexec_) {
            HIP_CHECK(hipGraphExecDestroy(gexec_));
            gexec_ = nullptr;
        }
        if (!cap_) HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
        hipStream_t keep = stream_;
        stream_ = cap_;
        try {
            enqueue(field, rhs, type);
        } catch (...) {
            stream_ = keep;
            (void)hipStreamEndCapture(cap_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        stream_ = keep;
        HIP_CHECK(hipStreamEndCapture(cap_, &g));
        const hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_CHECK(e);
        gfield_ = field;
        grhs_ = rhs;
        gtype_key_ = type;
    }
    HIP_CHECK(hipGraphLaunch(gexec_, stream_));
}

// minimum HBM traffic of one solve: every pass reads its input and writes its
// output once (zero padding not read, truncated halves not written), the
// multiply reads the field and the Green's function and writes the field
double SolverImpl::algorithmic_bytes(int type) const {
    if (fast_ && (type == 0 || lda_ == 3)) return fast_bytes(fast_);
    const double w = sizeof(double);
    const double user = (double)phys_.nglob[0] * phys_.nglob[1] * phys_.nglob[2] * lda_;
    double b = 0;
    const int nfull = (hybrid_ && type == 0) ? ndim_ - 1 : ndim_;
    for (int ip = 0; ip < nfull; ++ip) {
        const double in = ip == 0 ? user : (double)stage_[ip - 1].count();
        b += 2.0 * w * (in + (double)stage_[ip].count());   // forward pass and its backward mirror
    }
    if (hybrid_ && type == 0) return b + hybrid_bytes(hybrid_);
    const Shape& sp = stage_[ndim_ - 1];
    b += w * (2.0 * (double)sp.count() + (double)gpost_[ndim_ - 1].count());
    (void)type;
    return b;
}

void SolverImpl::profile(bool on) {
    profiling_ = on;
    if (fast_) fast_profile(fast_, on);
}
void SolverImpl::schedule(int out[3]) const {
    out[0] = out[1] = out[2] = -1;
    if (fast_) fast_schedule(fast_, out);
}
int SolverImpl::profile_get(int i, const char** name, double* ms, long* calls, double* bytes) const {
    return fast_ ? fast_profile_get(fast_, i, name, ms, calls, bytes) : 0;
}

}  // namespace flups
