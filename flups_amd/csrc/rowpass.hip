// Fast kernels of the general path: host side. The kernels and their launch templates are in
// rowpass_kernels.hpp; this unit holds the fit tests and tables, the FP64 instantiations and the
// dispatch on the storage types of a pass (P.src_f32, P.dst_f32).
#include "rowpass_kernels.hpp"

namespace flups {

FLUPS_ROWPASS_IO(rowpass_dd, double, double)
FLUPS_RZ(rz_d, double)

namespace {
void rowpass_io(int which, const PassDesc& P, hipStream_t s) {
    if (!P.src_f32 && !P.dst_f32) rowpass_dd(which, P, s);
    else if (!P.src_f32) rowpass_df(which, P, s);
    else if (P.dst_f32) rowpass_ff(which, P, s);
    else rowpass_fd(which, P, s);
}
}  // namespace

// length H of the complex FFT behind a real-to-real transform of n points: n/2 for the
// DCT/DST-II, III, IV, n - 1 for DCT-I (extension of 2(n-1)), n + 1 for DST-I (2(n+1))
int r2r_h(int kind, int n) { return kind == REDFT00 ? n - 1 : kind == RODFT00 ? n + 1 : n / 2; }

// n = pf q with pf in {1, 3, 5, 7} and q a power of two (false otherwise)
bool split_pq(int n, int& pf, int& q) {
    for (int f : {1, 3, 5, 7}) {
        if (n % f) continue;
        const int r = n / f;
        if (r > 0 && (r & (r - 1)) == 0) {
            pf = f;
            q = r;
            return true;
        }
    }
    return false;
}

namespace {
// lengths of the register column FFT: powers of two in [lo, hi], or PF times a power of two in
// [8, 128] with 3 <= PF <= maxpf. Holding 8 PF values per thread (Q/8 threads per pencil) suits
// the strided kernels that load from HBM at PF = 3 (half the generic kernel's time); at PF = 7
// it needs 300 registers and gains nothing, and for contiguous r2c rows it left too few waves
// (x of 96 in 96 x 128 x 64: 24 us against 14 us of the generic kernel). Thread teams (PF
// teams of Q/8 threads, 8 values each, inputs from LDS) serve the rest.
bool fft_len_ok(int n, int lo, int hi, int maxpf) {
    int pf = 0, q = 0;
    if (!split_pq(n, pf, q)) return false;
    if (pf == 1) return q >= lo && q <= hi;
    return pf <= maxpf && q >= 8 && q <= 128;
}
}  // namespace

namespace {
// c2c pencils: exactly the transform's n points, or n + 1 with the periodic copy of a node-
// centred periodic direction (position n gets X[0], forward and backward)
bool c2c_pencil_ok(const PassDesc& P) {
    const int n = P.n;
    if (P.pen_in == n && P.pen_out == n) return P.post == 0;
    const int q = P.sign == FWD ? P.n_out_plan - 1 : n;
    return P.pen_in == n + 1 && P.pen_out == n + 1 && P.post == POST_PERIOD && q == n;
}
}  // namespace

namespace {
// the plan semantics the r2r pencil kernels reproduce (kinds, length, offsets, corrections)
bool r2r_plan_fits(const PassDesc& P) {
    const int k = P.kind;
    const int n = P.n, h = r2r_h(k, n);
    if (P.op != OP_R2R) return false;
    const bool e00 = k == REDFT00 || k == RODFT00;
    if (!(e00 || k == REDFT10 || k == REDFT01 || k == REDFT11 || k == RODFT10 || k == RODFT01 || k == RODFT11)) return false;
    if ((!e00 && n % 2) || !fft_len_ok(h, 16, 1024, 7)) return false;
    if (!(P.src_nf == 1 && P.dst_nf == 1 && P.nf_in == 1 && P.nf_out == 1 && P.r2r_stride == 1 && P.pen_in == P.pen_out &&
          P.pen_in >= n && P.pen_in <= n + 2 && P.gshift >= 0 && P.sshift >= 0))
        return false;
    if (P.si < 0 || P.si > 1 || P.so < 0 || P.so > 1 || P.si + n > P.pen_in || (P.post & POST_PERIOD)) return false;
    // every stored output position (k + sshift for k < dst_nd) is a transform value or a point
    // the post-correction zeroes (the backward pass of an EVEN-EVEN direction stores n of its
    // n + 1 positions: the last one, left over from the forward pass, is not stored)
    const int last = std::min(P.pen_out, P.dst_nd + P.sshift);
    const bool head = P.so == 0 || (P.post & POST_FIRST) || P.sshift >= P.so;
    const bool tail = P.so + n >= last || (P.so + n == P.pen_out - 1 && (P.post & POST_LAST));
    return head && tail;
}
}  // namespace

bool row_r2r_fits(const PassDesc& P) { return r2r_plan_fits(P) && P.src_s[P.d] == 1 && P.dst_s[P.d] == 1; }

bool col_r2r_fits(const PassDesc& P) {
    return r2r_plan_fits(P) && P.fa != P.d && P.src_s[P.fa] == 1 && P.dst_s[P.fa] == 1 && (P.n1 * P.n2) % P.na == 0;
}

// pre / post twiddles of the row r2r kernels, (H + 1) complex each, exact to rounding
std::vector<double> row_r2r_table(int kind, int n) {
    const int dct = kind >= RODFT10 ? kind - 4 : kind;
    const int h = r2r_h(kind, n);
    const bool e00 = kind == REDFT00 || kind == RODFT00;
    const long double pi = 3.14159265358979323846264338327950288L, N = n;
    std::vector<double> t(4 * (size_t)(h + 1) + 2 * (size_t)h);   // A, B, then exp(-2 pi i e / h)
    auto put = [&](size_t i, long double angle) {
        t[2 * i] = (double)cosl(angle);
        t[2 * i + 1] = (double)sinl(angle);
    };
    for (int k = 0; k <= h; ++k) {
        if (e00) {   // A[k] = e^{-i pi k / h} (the 2h-point twiddle of the packed real FFT); B unused
            put(k, -pi * k / h);
            put(h + 1 + k, 0.0L);
        } else if (dct == REDFT11) {
            put(k, -pi * (k + 0.25L) / N);
            put(h + 1 + k, -pi * k / N);
        } else if (dct == REDFT10) {
            put(k, -2.0L * pi * k / N);
            put(h + 1 + k, -pi * k / (2.0L * N));
        } else {
            put(k, pi * k / (2.0L * N));
            put(h + 1 + k, 2.0L * pi * k / N);
        }
    }
    for (int e = 0; e < h; ++e) put(2 * (size_t)(h + 1) + e, -2.0L * pi * e / h);
    return t;
}

bool row_rc_fits(const PassDesc& P) {
    const int n = P.n, h = n / 2;
    if (!(P.op == OP_R2C || P.op == OP_C2R) || n % 2 || !fft_len_ok(h, 16, 1024, 7)) return false;
    const bool fwd = P.op == OP_R2C;
    const int d = P.d;
    if (P.src_s[d] != 1 || P.dst_s[d] != 1 || P.si != 0 || P.so != 0 || (P.post & ~POST_PERIOD) || P.gshift < 0 || P.sshift < 0)
        return false;
    // the periodic copy lands inside the pencil (r2c: complex value n_out - 1, c2r: real value n)
    if ((P.post & POST_PERIOD) && (fwd ? P.n_out_plan - 1 >= P.pen_out || P.n_out_plan < 1 : n >= P.pen_out)) return false;
    // buffer of the pencil in LDS: at most 2H + H/4 doubles
    if (P.pen_in * P.nf_in > 2 * h + h / 4 || P.pen_out * P.nf_out > 2 * h + h / 4) return false;
    if (fwd) return P.nf_in == 1 && P.src_nf == 1 && P.nf_out == 2 && P.dst_nf == 2 && P.pen_in >= n && P.pen_out >= h + 1;
    return P.nf_in == 2 && P.src_nf == 2 && P.nf_out == 1 && P.dst_nf == 1 && P.pen_in >= h + 1 && P.pen_out >= n;
}

bool col_real_fits(const PassDesc& P) {
    const int n = P.n;
    if (!(P.op == OP_R2C || P.op == OP_C2R) || !fft_len_ok(n, 32, 1024, 5)) return false;   // 8 PF values per thread
    const bool fwd = P.op == OP_R2C;
    const int d = P.d, fa = P.fa;
    if (fa == d || P.src_s[fa] != 1 || P.dst_s[fa] != 1) return false;
    // the real side may carry the two padding reals of the in-place r2c layout (n + 2 per
    // pencil): only rows below n are data, the rest is neither read nor written here. A
    // zero-padded (unbounded) direction reads src_nd < n rows and stores dst_nd < n back.
    // a node-centred periodic direction carries one more complex value (forward: X[0] again at
    // n/2 + 1) or real value (backward: x[n] = x[0])
    const bool per = P.post == POST_PERIOD;
    if (P.post != 0 && !(per && (fwd ? P.n_out_plan - 1 == n / 2 + 1 : true))) return false;
    const int nc = n / 2 + 1 + (per ? 1 : 0);   // complex values of the spectral pencil
    if (fwd && !(P.nf_in == 1 && P.nf_out == 2 && P.src_nf == 1 && P.dst_nf == 2 && P.pen_in >= n && P.pen_out == nc &&
                 P.src_nd >= 1 && P.src_nd <= P.pen_in && P.dst_nd == nc))
        return false;
    if (!fwd && !(P.nf_in == 2 && P.nf_out == 1 && P.src_nf == 2 && P.dst_nf == 1 && P.pen_in == nc && P.pen_out >= n + (per ? 1 : 0) &&
                  P.src_nd == nc && P.dst_nd >= 1 && P.dst_nd <= P.pen_out))
        return false;
    return P.si == 0 && P.so == 0 && P.gshift == 0 && P.sshift == 0 && (P.n1 * P.n2) % P.na == 0;
}

bool col_c2c_fits(const PassDesc& P) {
    const int n = P.n;
    if (P.op != OP_C2C || !fft_len_ok(n, 32, 2048, 7) || P.M != n) return false;
    const int d = P.d, fa = P.fa;
    if (fa == d || P.src_s[fa] != 1 || P.dst_s[fa] != 1 || (P.n1 * P.n2) % P.na != 0) return false;
    if (!(P.nf_in == 2 && P.nf_out == 2 && P.src_nf == 2 && P.dst_nf == 2 && c2c_pencil_ok(P))) return false;
    return P.si == 0 && P.so == 0 && P.gshift >= 0 && P.sshift >= 0 && P.src_nd >= 1 && P.dst_nd >= 1;
}

bool row_c2c_fits(const PassDesc& P) {
    const int n = P.n, d = P.d;
    if (P.op != OP_C2C || !fft_len_ok(n, 16, 1024, 7) || P.M != n) return false;
    if (P.src_s[d] != 1 || P.dst_s[d] != 1) return false;
    if (!(P.nf_in == 2 && P.nf_out == 2 && P.src_nf == 2 && P.dst_nf == 2 && c2c_pencil_ok(P))) return false;
    return P.si == 0 && P.so == 0 && P.gshift >= 0 && P.sshift >= 0 && P.src_nd >= 1 && P.dst_nd >= 1;
}

// H = M / 2: a power of two in [16, 1024] or 3, 5, 7 times one in [8, 128]
bool row_rz_len_ok(int M) { return M % 2 == 0 && fft_len_ok(M / 2, 16, 1024, 7); }

void launch_row_r2r(const PassDesc& P, hipStream_t s) { rowpass_io(RP_ROW_R2R, P, s); }
void launch_col_r2r(const PassDesc& P, hipStream_t s) { rowpass_io(RP_COL_R2R, P, s); }
void launch_row_rc(const PassDesc& P, hipStream_t s) { rowpass_io(RP_ROW_RC, P, s); }
void launch_col_real(const PassDesc& P, hipStream_t s) { rowpass_io(RP_COL_REAL, P, s); }
void launch_col_c2c(const PassDesc& P, hipStream_t s) { rowpass_io(RP_COL_C2C, P, s); }
void launch_row_c2c(const PassDesc& P, hipStream_t s) { rowpass_io(RP_ROW_C2C, P, s); }
// in place on one storage type: the data and the Green's function are both A.f32 ? float : double
void launch_col_rz(const RzArgs& A, hipStream_t s) { A.f32 ? rz_f(0, A, s) : rz_d(0, A, s); }
void launch_row_rz(const RzArgs& A, hipStream_t s) { A.f32 ? rz_f(1, A, s) : rz_d(1, A, s); }

}  // namespace flups
