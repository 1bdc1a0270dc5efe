// Fused pipeline for problems whose every direction is unbounded (UNB-UNB) or,
// cell-centred, periodic (PER-PER), transformed in the order x, y, z.
//
// The reference runs, per solve, 3 x (pencil switch + batched 1-D FFT), the
// multiply, then 3 x (FFT + switch back) (src/Solver.cpp:1207-1272). For the
// unbounded case the transforms are: r2c of the zero-padded x rows (length
// M0 = 2 N0 cell / 2 (N0-1) node), c2c along y (M1), c2c along z (M2),
// multiply by normfact * G, and the inverse chain with truncation. Here that is
// five kernels, each one HBM read and one HBM write of the field:
//   x_r2c     user rows (real)        -> S0 (nx = M0/2+1 complex, N1, N2)
//   col_fft   S0 along y, zero-padded -> S1 (nx, M1, N2)
//   z_fused   S1 along z: forward FFT, x normfact*G, inverse FFT, keep N2 (in place)
//   col_fft   S1 along y, inverse, keep N1 -> S0
//   x_c2r     S0 rows -> user rows (real), keep N0
// Periodic directions are transformed without padding (M = N, nothing truncated).
// Stage arrays keep x fastest with the row padded to a multiple of 8 complex
// values, so a tile of 8 consecutive x columns is one aligned 128-byte segment
// for FP64 (16 columns for FP32). The pencil switches of the reference become
// these strided tile loads; nothing is transposed in memory.
//
// Column FFTs: a team of M/8 threads owns one column, 8 values per thread at
// rows j + i*M/8. Radix-8 Stockham stages (a final radix-4 or radix-2 stage
// when M is not a power of 8) exchange through LDS; the first stage reads its
// inputs straight from HBM and the last stage keeps its outputs in registers,
// in the same row pattern. The forward z transform therefore ends exactly where
// the inverse one starts, so the Green's multiply sits in registers between
// them. The x rows use the half-length packed complex FFT for r2c / c2r.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "fft_core.hpp"
#include "group.hpp"
#include "solver.hpp"

namespace flups {

namespace {

// Non-temporal accesses (the nt bit: the lines are not kept in the Infinity Cache) for data
// that no later pass of the solve reads from the cache, when the arrays are too large for the
// 256 MB cache anyway (the launch's `nt` flag, set for stage arrays of 128 MiB or more; smaller
// problems keep every access cached, the next pass then reads what fits from the cache):
//  * the caller's rows, read once by x_r2c and written once by x_c2r;
//  * the stores of x_r2c into S0 and of y_fwd into S1: kept lines are written back as dirty
//    victims while the next pass streams; stored non-temporally, y_fwd takes 143 instead of
//    160 us and z_fused 328 instead of 343 us (FP32: 169 instead of 182 us) at 256^3.
// The z pass's and y_inv's stores stay cached: y_inv and x_c2r read them back first (no gain
// measured from non-temporal stores there).
typedef double d2nt __attribute__((ext_vector_type(2)));
typedef float f2nt __attribute__((ext_vector_type(2)));
__device__ __forceinline__ d2nt ld_nt(const double* p) { return __builtin_nontemporal_load(reinterpret_cast<const d2nt*>(p)); }
__device__ __forceinline__ void st_nt(double* p, double a, double b) {
    __builtin_nontemporal_store(d2nt{a, b}, reinterpret_cast<d2nt*>(p));
}
__device__ __forceinline__ void st_nt(C2<double>* p, C2<double> v) {
    __builtin_nontemporal_store(d2nt{v.x, v.y}, reinterpret_cast<d2nt*>(p));
}
__device__ __forceinline__ void st_nt(C2<float>* p, C2<float> v) {
    __builtin_nontemporal_store(f2nt{v.x, v.y}, reinterpret_cast<f2nt*>(p));
}

// ---------------------------------------------------------------------------
// column passes along y or z of the stage arrays
struct ColArgs {
    const void* src;
    void* dst;
    const void* green;   // z_fused only: normfact * G, [kz'][ky'][kx], kx padded to ldx
    const void* tw;
    int nx;              // valid x columns
    int ldx;             // x row length in complex elements (multiple of 8/16)
    long long s_col;     // element stride between rows of a column (src and dst share it for z; y differs below)
    long long s_col_dst;
    long long s_other_src, s_other_dst;   // stride of the other (batch) dimension
    int nother;
    int nin, nout;       // valid input rows, rows to store
    int gy, gmy, gmz;    // z_fused: ky' extent, M1 (for ky' = min(y, M1-y)), M2
    int cstride_src, cstride_dst;   // component strides in complex elements (low 32 bits)
    long long comp_src, comp_dst;
    int rev;             // 1: blocks walk the batch from its last entry
    int xcd;             // z_fused: 1 = XCD-grouped block numbering
    int nopair;          // z_fused: 1 = the other dim has no mirror rows (one row per block, G stored per row)
    int x0;              // z_fused / z_rot: spectral x index of column 0 (slab offset, a multiple of TX)
    int gntx;            // z_fused / z_rot: x tiles of one ky' row of the compressed G
    int gtx;             // z_rot: tile width of the compressed G (a multiple of the kernel's TX)
    const void* twm;     // z_fused with PF > 1: M-point twiddle table (tw is then the M/PF-point one)
    int crot;            // col_fft: 1 = component c is read from slot (c + 1) mod 3 (after z_rot)
    int nt;              // col_fft forward: 1 = non-temporal stores
};

// TX columns per block, M/8 threads per column; block = TX * M/8 threads.
// HALF: zero-padded input of M/2 rows (forward) or output truncated to M/2 rows
// (inverse); nin / nout are then not consulted. Lanes past nx load a valid
// column (the last one) and store nothing.
template <typename R, int M, int TX, bool FWD, bool HALF>
struct col_fft {
    using Args = ColArgs;
    static constexpr int NT = TX * M / 8;
    static __device__ __forceinline__ void run(const ColArgs& a, Blk bk) {
        __shared__ C2<R> lds[M * TX];
        constexpr int T = M / 8;
        const int c = threadIdx.x % TX;
        const int j = threadIdx.x / TX;
        const int ntx = (a.nx + TX - 1) / TX;
        int blk = a.xcd ? xcd_group(bk.x, bk.nx) : bk.x;
        if (a.rev) blk = bk.nx - 1 - blk;
        const int tile = blk % ntx;
        const int other = blk / ntx;
        const int x = tile * TX + c;
        const bool live = x < a.nx;
        const int cin = a.crot ? (bk.y + 1) % 3 : bk.y;   // source slot of this component
        const C2<R>* src = reinterpret_cast<const C2<R>*>(a.src) + cin * a.comp_src + other * a.s_other_src + min(x, a.nx - 1);
        C2<R>* dst = reinterpret_cast<C2<R>*>(a.dst) + bk.y * a.comp_dst + other * a.s_other_dst + x;
        const C2<R>* tw = reinterpret_cast<const C2<R>*>(a.tw);
        constexpr bool HIN = HALF && FWD, HOUT = HALF && !FWD;
        C2<R> v[8];
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = j + i * T;
            if (HIN) v[i] = i < 4 ? src[row * a.s_col] : C2<R>{0, 0};
            else if (HOUT) v[i] = src[row * a.s_col];
            else v[i] = row < a.nin ? src[row * a.s_col] : C2<R>{0, 0};
        }
        column_fft<R, M, FWD, TX, HIN>(v, lds + c, j, tw);
        if (live) {
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int row = j + i * T;
                if (HOUT ? i < 4 : (HIN || row < a.nout)) {
                    if (FWD && a.nt) st_nt(dst + row * a.s_col_dst, v[i]);   // y_fwd into a large S1
                    else dst[row * a.s_col_dst] = v[i];
                }
            }
        }
    }
};

// Fused z pass, in place on S1: forward FFT along z, multiply by normfact * G,
// inverse FFT, keep the first nin rows. G is real and even in ky and kz, so one
// block serves the y rows ky and M1 - ky of one x tile with the same G values:
// its slice of the compressed G, [ky'][tile][kz'][TX], is contiguous and read
// once. PREFETCH issues the loads of both column sets before the first FFT
// (more bytes in flight, more registers); otherwise the second set is loaded
// after the first one is stored.
// HALF: nin = nout = M/2 (unbounded direction): only four rows per thread are
// loaded and stored, the zero half is pruned from the first forward stage.
// PF = 3: M = 3 Q (the general path's fused last direction), 8 PF values per thread through
// column_fft_pq, no HALF pruning.
// S: storage of the field and of G (R, or float with R = double: the general path's FP32 solves,
// FP32 stage arrays with FP64 arithmetic)
template <typename R, int M, int TX, int SCHED, bool HALF, int PF = 1, typename S = R>
struct z_fused {
    using Args = ColArgs;
    static constexpr int NT = TX * M / (8 * PF);
    // the mixed-radix columns: at most 256 registers per lane; the zero-padded FP32 columns up to
    // 512 with the LDS twiddles (68 registers uncapped): 64, eight waves per SIMD, at most 16 bytes
    // spilled (256^3: 2278 against 2254 solves/s uncapped); the other FP32 variants would spill
    // up to 80 bytes under the cap and are left uncapped
    static constexpr int kWaves = PF == 3 ? 2 : (sizeof(R) == 4 && PF == 1 && HALF && SCHED != 1 && M <= 512 ? 8 : 0);
    static __device__ __forceinline__ void run(const ColArgs& a, Blk bk) {
        // SCHED 0: G loaded with the first set, second set loaded after the first is stored;
        //       1: both sets prefetched up front; 2: G loaded just before each multiply.
        static_assert(PF == 1 || !HALF, "the zero-padded half is pruned for powers of two only");
        __shared__ C2<R> lds[M * TX];
        // FP32 (bound by its VALU instructions): the twiddle table copied to LDS, seven LDS reads
        // per radix-8 stage instead of three loads and four complex products
        constexpr bool LTW = sizeof(R) == 4 && PF == 1;
        __shared__ C2<R> ltw[LTW ? M : 1];
        constexpr int Q = M / PF, T = Q / 8, NV = 8 * PF;
        const int c = threadIdx.x % TX;
        const int j = threadIdx.x / TX;
        const int ntx = (a.nx + TX - 1) / TX;
        const int blk = a.xcd ? xcd_group(bk.x, bk.nx) : bk.x;
        const int tile = blk % ntx;
        int y0 = blk / ntx;                 // 0 .. M1/2 (every row with nopair = 1)
        int yrow = y0;                      // the row of the first (or only) column set
        if (a.nopair == 2) {                // the two rows of a pair in neighbouring blocks (same G slice)
            const int r = y0;
            y0 = r >> 1;
            yrow = (r & 1) ? a.gmy - y0 : y0;
            if ((r & 1) && (y0 == 0 || 2 * y0 == a.gmy)) return;   // no mirror row: the whole block leaves
        }
        const int y1 = a.gmy - y0;          // mirror row; none for y0 = 0 (M1) and y0 = M1/2 (itself)
        const bool pair = a.nopair == 0 && y0 != 0 && y1 != y0;
        const bool live = tile * TX + c < a.nx;
        // uniform bases (scalar registers) + one 32-bit per-lane byte offset for every access;
        // lanes past nx load the last valid column and store nothing
        char* base = reinterpret_cast<char*>(a.dst) +
                     (size_t)sizeof(C2<S>) * ((long long)bk.y * a.comp_src + (long long)tile * TX);
        char* b0 = base + (size_t)sizeof(C2<S>) * ((long long)yrow * a.s_other_src);
        char* b1 = base + (size_t)sizeof(C2<S>) * ((long long)y1 * a.s_other_src);
        const size_t rstride = (size_t)sizeof(C2<S>) * T * a.s_col;                      // i -> i + 1
        const unsigned voff = (unsigned)(sizeof(C2<S>) * ((size_t)j * a.s_col + c));     // < 4 GiB: j < M/8
        const unsigned loff = (unsigned)(sizeof(C2<S>) * ((size_t)j * a.s_col + min(c, a.nx - 1 - tile * TX)));
        constexpr int NL = HALF ? 4 : NV;   // rows loaded / stored per thread
        const char* g = reinterpret_cast<const char*>(a.green) +
                        sizeof(S) * ((long long)y0 * a.gntx + a.x0 / TX + tile) * (M / 2 + 1) * TX;
        const C2<R>* tw = reinterpret_cast<const C2<R>*>(a.tw);
        const C2<R>* twm = reinterpret_cast<const C2<R>*>(a.twm);
        R gv[NV];
        C2<R> v[NV], w[NV];
        auto load_g = [&]() {
    #pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int kz = j + i * T;
                gv[i] = (R)*reinterpret_cast<const S*>(g + (unsigned)(sizeof(S) * (min(kz, M - kz) * TX + c)));
            }
        };
        auto load = [&](C2<R>* u, const char* b) {
    #pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int row = j + i * T;
                if (HALF) u[i] = i < 4 ? ld2<R, S>(b + i * rstride + loff) : C2<R>{0, 0};
                else u[i] = row < a.nin ? ld2<R, S>(b + i * rstride + loff) : C2<R>{0, 0};
            }
        };
        if (SCHED != 2) load_g();
        load(v, b0);
        if (SCHED == 1 && pair) load(w, b1);
        if constexpr (LTW) {   // after the data loads are issued: the first FFT stage waits for them anyway
            for (int i = threadIdx.x; i < M; i += NT) ltw[i] = tw[i];
            __syncthreads();
            tw = ltw;
        }
        for (int s = 0; s < 2; ++s) {
            column_fft_pq<R, PF, Q, true, TX, HALF, LTW>(v, lds + c, j, tw, twm, 1);
            if (SCHED == 2) load_g();
    #pragma unroll
            for (int i = 0; i < NV; ++i) {
                v[i].x *= gv[i];
                v[i].y *= gv[i];
            }
            column_fft_pq<R, PF, Q, false, TX, false, LTW>(v, lds + c, j, tw, twm, 1);   // syncs before its first LDS write
            char* b = s == 0 ? b0 : b1;
            if (live) {
    #pragma unroll
                for (int i = 0; i < NL; ++i) {
                    const int row = j + i * T;
                    if (HALF || row < a.nout) st2<R, S>(b + i * rstride + voff, v[i]);
                }
            }
            if (!pair) break;   // uniform over the block
            if (s == 0) {
                if (SCHED == 1) {
    #pragma unroll
                    for (int i = 0; i < NV; ++i) v[i] = w[i];
                } else {
                    load(v, b1);
                }
            }
        }
    }
};

// The general path's fused last direction for M = PF Q with PF = 5 or 7: one row per block (no
// mirror pairing), TX columns staged in LDS, PF thread teams of Q/8 threads (fft_pq_team) for
// the forward FFT, the multiply by normfact G and the inverse FFT; in place.
template <int PF, int Q, int TX, typename S = double>
struct z_team {
    using Args = ColArgs;
    static constexpr int NT = TX * PF * Q / 8;
    static __device__ __forceinline__ void run(const ColArgs& a, Blk bk) {
        constexpr int M = PF * Q, T = Q / 8, TT = M / 8;
        __shared__ C2<double> lds[M * TX];
        const int c = threadIdx.x % TX, tt = threadIdx.x / TX;
        const int k1 = tt / T, j = tt % T;
        const int ntx = (a.nx + TX - 1) / TX;
        const int blk = a.xcd ? xcd_group(bk.x, bk.nx) : bk.x;
        const int tile = blk % ntx, y = blk / ntx;
        const bool live = tile * TX + c < a.nx;
        const int cl = min(c, a.nx - 1 - tile * TX);   // lanes past nx load the last valid column
        C2<S>* col = reinterpret_cast<C2<S>*>(a.dst) + (long long)bk.y * a.comp_src + (long long)y * a.s_other_src +
                     (long long)tile * TX;
        // all eight loads (clamped rows, zero selected afterwards) before the LDS writes
        C2<double> u[8];
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = tt + i * TT;
            const C2<double> x = c2d(col[(long long)min(r, a.nin - 1) * a.s_col + cl]);
            u[i] = r < a.nin ? x : C2<double>{0.0, 0.0};
        }
    #pragma unroll
        for (int i = 0; i < 8; ++i) lds[(tt + i * TT) * TX + c] = u[i];
        __syncthreads();
        const C2<double>* tw = reinterpret_cast<const C2<double>*>(a.tw);
        const C2<double>* twm = reinterpret_cast<const C2<double>*>(a.twm);
        auto load = [&](int m) -> C2<double> { return lds[m * TX + c]; };
        C2<double> v[8];
        fft_pq_team<double, PF, Q, true, TX>(v, lds + c, k1, j, load, tw, twm, 1);
        const S* g = reinterpret_cast<const S*>(a.green) + ((long long)y * a.gntx + tile) * (M / 2 + 1) * TX;
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int kz = k1 + PF * (j + i * T);
            const double gv = (double)g[min(kz, M - kz) * TX + c];
            v[i] = {v[i].x * gv, v[i].y * gv};
        }
        __syncthreads();   // the forward FFT's last LDS reads are done
    #pragma unroll
        for (int i = 0; i < 8; ++i) lds[(k1 + PF * (j + i * T)) * TX + c] = v[i];
        __syncthreads();
        fft_pq_team<double, PF, Q, false, TX>(v, lds + c, k1, j, load, tw, twm, 1);
        if (!live) return;
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = k1 + PF * (j + i * T);
            if (r < a.nout) col[(long long)r * a.s_col + c] = c2t<S>(v[i]);
        }
    }
};

// Rotational (Biot-Savart) z pass, in place on the three components of S1:
// forward FFT of omega_x, omega_y, omega_z along z, psi = normfact G omega, then
// u = curl psi = i K x psi with the derivative symbol K of each direction (the
// wave number itself, or the symbol of the 2nd/4th/6th-order centred
// difference), inverse FFTs, keep the first nin rows. The same y-row pairing
// and Green's slice as z_fused. src/dothemagic_rot.ipp:102-349.
struct RotK {
    double kfact[3], h[3], symstart[3];
    int order;          // 1 spectral, 2 / 4 / 6 centred differences
    const void* sym;    // derivative symbols by mode index: x (nx), then y (M1), then z (M2), type R
    int off[3];         // start of each direction's table
};

// derivative symbol of mode idx along d: symmetrised index (mirrored half negative),
// k = is * kfact, then k itself (spectral) or the symbol of the 2nd/4th/6th-order
// centred difference (src/dothemagic_rot.ipp:14-101); tabulated on the host
double rot_symbol(const RotK& k, int d, int idx) {
    const double ss = k.symstart[d];
    int is = idx;
    if (ss != 0.0 && idx > ss) {
        const int v = (int)std::fabs(2.0 * ss - idx);
        is = -(v > 1 ? v : 1);
    }
    const double kk = is * k.kfact[d], hh = k.h[d];
    switch (k.order) {
        case 2: return std::sin(kk * hh) / hh;
        case 4: return (4.0 / 3.0) * std::sin(kk * hh) / hh - (1.0 / 6.0) * std::sin(2.0 * kk * hh) / hh;
        case 6: return 1.5 * std::sin(kk * hh) / hh - 0.3 * std::sin(2.0 * kk * hh) / hh + (1.0 / 30.0) * std::sin(3.0 * kk * hh) / hh;
        default: return kk;
    }
}

// FP64: the three components' sets of NV values live together (all loads issued before the
// first transform); in place, u_c back in slot c
template <typename R, int M, int TX>
__device__ __forceinline__ void z_rot_body3(const ColArgs& a, const RotK& rk, C2<R>* lds, Blk bk) {
    constexpr int NV = 8, T = M / NV;
    const int c = threadIdx.x % TX;
    const int j = threadIdx.x / TX;
    const int ntx = (a.nx + TX - 1) / TX;
    const int blk = a.xcd ? xcd_group(bk.x, bk.nx) : bk.x;
    const int tile = blk % ntx;
    const int y0 = blk / ntx;
    const int y1 = a.gmy - y0;
    const bool pair = y0 != 0 && y1 != y0;
    const int x = tile * TX + c;
    const bool live = x < a.nx;
    char* base = reinterpret_cast<char*>(a.dst) + (size_t)sizeof(C2<R>) * ((long long)tile * TX);
    const size_t cstride = (size_t)sizeof(C2<R>) * a.comp_src;
    const size_t rstride = (size_t)sizeof(C2<R>) * T * a.s_col;
    const unsigned voff = (unsigned)(sizeof(C2<R>) * ((size_t)j * a.s_col + c));
    // lanes past nx load the last valid column and store nothing
    const unsigned loff = (unsigned)(sizeof(C2<R>) * ((size_t)j * a.s_col + min(c, a.nx - 1 - tile * TX)));
    // the block's TX columns lie in one tile of the G layout (gtx a multiple of TX)
    const int xg = a.x0 + tile * TX;
    const int cg = xg % a.gtx + c;
    const char* g = reinterpret_cast<const char*>(a.green) +
                    sizeof(R) * ((long long)y0 * a.gntx + xg / a.gtx) * (M / 2 + 1) * a.gtx;
    const C2<R>* tw = reinterpret_cast<const C2<R>*>(a.tw);
    const R* sym = reinterpret_cast<const R*>(rk.sym);
    const R kx = sym[rk.off[0] + a.x0 + min(x, a.nx - 1)];
    for (int s = 0; s < 2; ++s) {
        const int y = s == 0 ? y0 : y1;
        char* b = base + (size_t)sizeof(C2<R>) * ((long long)y * a.s_other_src);
        const R ky = sym[rk.off[1] + y];
        C2<R> v[3][NV];
        // the three components' loads all issued before the first FFT (its barriers would
        // otherwise hold each component's loads back until the previous transform is done)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int row = j + i * T;
                v[cc][i] = row < a.nin ? *reinterpret_cast<const C2<R>*>(b + cc * cstride + i * rstride + loff) : C2<R>{0, 0};
            }
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            column_fft<R, M, true, TX>(v[cc], lds + c, j, tw);
        }
        // G of this tile, loaded after the forward transforms (fewer registers live during them)
        R gv[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int kz = j + i * T;
            gv[i] = *reinterpret_cast<const R*>(g + (unsigned)(sizeof(R) * (min(kz, M - kz) * a.gtx + cg)));
        }
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const R kz = sym[rk.off[2] + j + i * T];
            const C2<R> wx = {v[0][i].x * gv[i], v[0][i].y * gv[i]};
            const C2<R> wy = {v[1][i].x * gv[i], v[1][i].y * gv[i]};
            const C2<R> wz = {v[2][i].x * gv[i], v[2][i].y * gv[i]};
            // u = i K x w: (K x w) then multiply by i: (p + i q) i = -q + i p
            const C2<R> cx = {ky * wz.x - kz * wy.x, ky * wz.y - kz * wy.y};
            const C2<R> cy = {kz * wx.x - kx * wz.x, kz * wx.y - kx * wz.y};
            const C2<R> cz = {kx * wy.x - ky * wx.x, kx * wy.y - ky * wx.y};
            v[0][i] = {-cx.y, cx.x};
            v[1][i] = {-cy.y, cy.x};
            v[2][i] = {-cz.y, cz.x};
        }
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            column_fft<R, M, false, TX>(v[cc], lds + c, j, tw);
            if (live) {
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    const int row = j + i * T;
                    if (row < a.nout) *reinterpret_cast<C2<R>*>(b + cc * cstride + i * rstride + voff) = v[cc][i];
                }
            }
        }
        if (!pair) break;   // uniform over the block
    }
}

// FP32: two sets of NV values live at a time (see the comment inside); u_c lands in slot
// (c + 1) mod 3, and the inverse y pass reads it there (ColArgs::crot). For FP64 this order
// (one component read and transformed twice) is slower than the three-set one: 32.5 against
// 27.6 ms at 513^3 x 3, the fewer registers still leave one block of 512 threads per CU; FP32:
// 17.6 against 18.7 ms (no spills at 128 registers, where the three-set order spilled 20)
template <typename R, int M, int TX>
__device__ __forceinline__ void z_rot_body2(const ColArgs& a, const RotK& rk, C2<R>* lds, C2<R>* ltw, Blk bk) {
    constexpr int NV = 8, T = M / NV;
    const int c = threadIdx.x % TX;
    const int j = threadIdx.x / TX;
    const int ntx = (a.nx + TX - 1) / TX;
    const int blk = a.xcd ? xcd_group(bk.x, bk.nx) : bk.x;
    const int tile = blk % ntx;
    const int y0 = blk / ntx;
    const int y1 = a.gmy - y0;
    const bool pair = y0 != 0 && y1 != y0;
    const int x = tile * TX + c;
    const bool live = x < a.nx;
    char* base = reinterpret_cast<char*>(a.dst) + (size_t)sizeof(C2<R>) * ((long long)tile * TX);
    const size_t cstride = (size_t)sizeof(C2<R>) * a.comp_src;
    const size_t rstride = (size_t)sizeof(C2<R>) * T * a.s_col;
    const unsigned voff = (unsigned)(sizeof(C2<R>) * ((size_t)j * a.s_col + c));
    // lanes past nx load the last valid column and store nothing
    const unsigned loff = (unsigned)(sizeof(C2<R>) * ((size_t)j * a.s_col + min(c, a.nx - 1 - tile * TX)));
    // the block's TX columns lie in one tile of the G layout (gtx a multiple of TX)
    const int xg = a.x0 + tile * TX;
    const int cg = xg % a.gtx + c;
    const char* g = reinterpret_cast<const char*>(a.green) +
                    sizeof(R) * ((long long)y0 * a.gntx + xg / a.gtx) * (M / 2 + 1) * a.gtx;
    const R* sym = reinterpret_cast<const R*>(rk.sym);
    const R kx = sym[rk.off[0] + a.x0 + min(x, a.nx - 1)];
    // the twiddle table copied to LDS: seven LDS reads per radix-8 stage instead of three loads and
    // four complex products (the FP32 pass runs seven column FFTs per column set)
    for (int i = threadIdx.x; i < M; i += TX * M / 8) ltw[i] = reinterpret_cast<const C2<R>*>(a.tw)[i];
    __syncthreads();
    const C2<R>* tw = ltw;
    for (int s = 0; s < 2; ++s) {
        const int y = s == 0 ? y0 : y1;
        char* b = base + (size_t)sizeof(C2<R>) * ((long long)y * a.s_other_src);
        const R ky = sym[rk.off[1] + y];
        // Two sets of NV values live at a time (u = i K x psi, psi = G w). The pass is in place and
        // every w_c enters two components of u, so the components are written to the slot of an
        // input this block already holds as psi: u_z to slot x, u_y to slot z, u_x to slot y (w_y is
        // read and transformed twice); the inverse y pass reads component c from slot (c + 1) mod 3.
        C2<R> p[NV], q[NV];
        auto load = [&](C2<R>* u, int cc) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int row = j + i * T;
                u[i] = row < a.nin ? *reinterpret_cast<const C2<R>*>(b + cc * cstride + i * rstride + loff) : C2<R>{0, 0};
            }
        };
        auto store = [&](int cc, const C2<R>* u) {
            if (!live) return;
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int row = j + i * T;
                if (row < a.nout) *reinterpret_cast<C2<R>*>(b + cc * cstride + i * rstride + voff) = u[i];
            }
        };
        auto psi = [&](C2<R>* u) {   // forward transform, times normfact G (re-read for each use: fewer live registers)
            column_fft<R, M, true, TX, false, true>(u, lds + c, j, tw);
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int kz = j + i * T;
                const R gk = *reinterpret_cast<const R*>(g + (unsigned)(sizeof(R) * (min(kz, M - kz) * a.gtx + cg)));
                u[i] = {u[i].x * gk, u[i].y * gk};
            }
        };
        auto kz_of = [&](int i) { return sym[rk.off[2] + j + i * T]; };
        load(p, 0);
        load(q, 1);
        psi(p);   // psi_x
        psi(q);   // psi_y
#pragma unroll
        for (int i = 0; i < NV; ++i) {   // u_z = i (kx psi_y - ky psi_x)
            const C2<R> cz = {kx * q[i].x - ky * p[i].x, kx * q[i].y - ky * p[i].y};
            q[i] = {-cz.y, cz.x};
        }
        column_fft<R, M, false, TX, false, true>(q, lds + c, j, tw);
        store(0, q);
        load(q, 2);
        psi(q);   // psi_z
#pragma unroll
        for (int i = 0; i < NV; ++i) {   // u_y = i (kz psi_x - kx psi_z)
            const R kz = kz_of(i);
            const C2<R> cy = {kz * p[i].x - kx * q[i].x, kz * p[i].y - kx * q[i].y};
            p[i] = {-cy.y, cy.x};
        }
        column_fft<R, M, false, TX, false, true>(p, lds + c, j, tw);
        store(2, p);
        load(p, 1);
        psi(p);   // psi_y again
#pragma unroll
        for (int i = 0; i < NV; ++i) {   // u_x = i (ky psi_z - kz psi_y)
            const R kz = kz_of(i);
            const C2<R> cx = {ky * q[i].x - kz * p[i].x, ky * q[i].y - kz * p[i].y};
            p[i] = {-cx.y, cx.x};
        }
        column_fft<R, M, false, TX, false, true>(p, lds + c, j, tw);
        store(1, p);
        if (!pair) break;   // uniform over the block
    }
}

template <typename R, int M, int TX>
__device__ __forceinline__ void z_rot_body(const ColArgs& a, const RotK& rk, C2<R>* lds, C2<R>* ltw, Blk bk) {
    if constexpr (sizeof(R) == 4) z_rot_body2<R, M, TX>(a, rk, lds, ltw, bk);
    else z_rot_body3<R, M, TX>(a, rk, lds, bk);
}

// at most 256 registers per lane (two waves per SIMD): the three components need 250-285
// (FP64), and the few that spill cost less than halving the waves (513^3 x 3 solve: 34.0 ->
// 29.8 ms)
struct ZRotArgs {
    ColArgs a;
    RotK rk;
};
template <typename R, int M, int TX>
struct z_rot {
    using Args = ZRotArgs;
    static constexpr int NT = TX * M / 8;
    static constexpr int kWaves = 2;
    static __device__ __forceinline__ void run(const ZRotArgs& z, Blk bk) {
        __shared__ C2<R> lds[M * TX];
        __shared__ C2<R> ltw[sizeof(R) == 4 ? M : 1];   // FP32: the twiddle table
        z_rot_body<R, M, TX>(z.a, z.rk, lds, ltw, bk);
    }
};

// ---------------------------------------------------------------------------
// x rows: r2c of the zero-padded real row (length M = 2*H) through the packed
// complex FFT of length H: z[n] = x[2n] + i x[2n+1], Z = FFT_H(z),
// X[k] = (Z[k] + conj Z[H-k])/2 + exp(-2 pi i k/M) (Z[k] - conj Z[H-k])/(2i).
struct RowArgs {
    const void* src;
    void* dst;
    const void* tw;      // H-point table exp(-2 pi i t/H)
    const void* tw2;     // M-point table exp(-2 pi i t/M)
    long long s_row_src; // stride between rows (y) in source elements (real for r2c / complex for c2r)
    long long s_z_src;   // stride between planes (z)
    long long s_row_dst, s_z_dst;
    int n1;              // rows per plane
    int nrows;           // n1 * n2
    int nvalid;          // valid reals per input row (r2c) / reals to store per output row (c2r)
    int nx;              // complex outputs per row (H + 1)
    long long comp_src, comp_dst;
    int rev;             // 1: blocks walk the rows from the last one (reads what the previous pass wrote last)
    int nt;              // 1: the caller's rows and the stores into S0 (x_r2c) are non-temporal
};

// MODE 0: any nvalid, scalar loads; 1: nvalid = H (zero-padded row: the upper half of the
// packed input is zero and pruned from the first stage), 16-byte loads; 2: nvalid = 2H,
// 16-byte loads. Modes 1 and 2 need 16-byte aligned user rows.
template <typename R, int H, int ROWS, int MODE>
struct x_r2c {
    using Args = RowArgs;
    static constexpr int NT = ROWS * H / 8;
    static __device__ __forceinline__ void run(const RowArgs& a, Blk bk) {
        constexpr int HP = H + H / 8;     // padded row buffer
        __shared__ C2<R> lds[HP * ROWS];
        constexpr int T = H / 8;
        const int rl = threadIdx.x / T;   // row within the block
        const int j = threadIdx.x % T;
        const int blk = a.rev ? bk.nx - 1 - bk.x : bk.x;
        const int row = blk * ROWS + rl;
        const bool live = row < a.nrows;
        const int y = live ? row % a.n1 : 0, z = live ? row / a.n1 : 0;
        const double* src = reinterpret_cast<const double*>(a.src) + bk.y * a.comp_src + y * a.s_row_src + z * a.s_z_src;
        C2<R>* dst = reinterpret_cast<C2<R>*>(a.dst) + bk.y * a.comp_dst + y * a.s_row_dst + z * a.s_z_dst;
        C2<R>* col = lds + rl * HP;   // row's buffer
        C2<R> v[8];
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int n = j + i * T;
            const int e0 = 2 * n, e1 = 2 * n + 1;
            if (MODE == 0) {
                R re = 0, im = 0;
                if (live && e0 < a.nvalid) re = (R)src[e0];
                if (live && e1 < a.nvalid) im = (R)src[e1];
                v[i] = {re, im};
            } else if (MODE == 2 || i < 4) {
                // dead rows (past nrows) read row 0 of the plane and store nothing
                if (a.nt) {
                    const d2nt p = ld_nt(src + e0);
                    v[i] = {(R)p.x, (R)p.y};
                } else {
                    const double2 p = *reinterpret_cast<const double2*>(src + e0);
                    v[i] = {(R)p.x, (R)p.y};
                }
            } else {
                v[i] = {0, 0};
            }
        }
        column_fft<R, H, true, 0, MODE == 1>(v, col, j, reinterpret_cast<const C2<R>*>(a.tw));
        __syncthreads();
    #pragma unroll
        for (int i = 0; i < 8; ++i) col[lslot<0>(j + i * T)] = v[i];
        __syncthreads();
        const C2<R>* tw2 = reinterpret_cast<const C2<R>*>(a.tw2);
        if (live) {
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = j + i * T;
                const C2<R> zk = col[lslot<0>(k)], zc = cconj(col[lslot<0>((H - k) & (H - 1))]);
                const C2<R> e = {(zk.x + zc.x) * (R)0.5, (zk.y + zc.y) * (R)0.5};
                const C2<R> df = {(zk.x - zc.x) * (R)0.5, (zk.y - zc.y) * (R)0.5};
                const C2<R> o = {df.y, -df.x};
                const C2<R> r = cadd(e, cmul(tw2[k], o));
                if (a.nt) st_nt(dst + k, r);
                else dst[k] = r;
            }
            if (j == 0) {   // Nyquist: Re Z0 - Im Z0
                const C2<R> z0 = col[0];
                dst[H] = {z0.x - z0.y, (R)0};
            }
        }
    }
};

// x rows inverse: Z[k] = (X[k] + conj X[H-k]) + i exp(+2 pi i k/M) (X[k] - conj X[H-k]), z = IFFT_H(Z),
// x[2n] = Re z[n], x[2n+1] = Im z[n]; the imaginary parts of X[0] and X[H] are ignored.
// MODE as for x_r2c: 1 stores the first H reals of the row (the upper half of the
// inverse is not computed), 2 all 2H; both with 16-byte stores to aligned user rows.
// Each thread loads its X[k] once; X[H - k] comes from the thread holding it through LDS
// (loading both from memory doubled the load instructions: 256^3 FP64 73 -> 69 us).
template <typename R, int H, int ROWS, int MODE>
struct x_c2r {
    using Args = RowArgs;
    static constexpr int NT = ROWS * H / 8;
    static __device__ __forceinline__ void run(const RowArgs& a, Blk bk) {
        constexpr int HP = H + H / 8 + 8;   // slots lslot(0 .. H): X[H] beside the FFT's column
        __shared__ C2<R> lds[HP * ROWS];
        constexpr int T = H / 8;
        const int rl = threadIdx.x / T;
        const int j = threadIdx.x % T;
        const int blk = a.rev ? bk.nx - 1 - bk.x : bk.x;
        const int row = blk * ROWS + rl;
        const bool live = row < a.nrows;
        const int y = live ? row % a.n1 : 0, z = live ? row / a.n1 : 0;
        const C2<R>* src = reinterpret_cast<const C2<R>*>(a.src) + bk.y * a.comp_src + y * a.s_row_src + z * a.s_z_src;
        double* dst = reinterpret_cast<double*>(a.dst) + bk.y * a.comp_dst + y * a.s_row_dst + z * a.s_z_dst;
        const C2<R>* tw2 = reinterpret_cast<const C2<R>*>(a.tw2);
        C2<R>* col = lds + rl * HP;
        C2<R> v[8], xs[8];
    #pragma unroll
        for (int i = 0; i < 8; ++i) xs[i] = live ? src[j + i * T] : C2<R>{0, 0};
        const C2<R> xh = live ? src[H] : C2<R>{0, 0};
    #pragma unroll
        for (int i = 0; i < 8; ++i) col[lslot<0>(j + i * T)] = xs[i];
        if (j == 0) col[lslot<0>(H)] = xh;
        __syncthreads();
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = j + i * T;
            C2<R> xa = xs[i], xb = col[lslot<0>(H - k)];
            if (k == 0) { xa.y = 0; xb.y = 0; }
            const C2<R> bc = cconj(xb);
            const C2<R> e = cadd(xa, bc), f = csub(xa, bc);
            const C2<R> t = cmul(cconj(tw2[k]), f);
            v[i] = {e.x - t.y, e.y + t.x};
        }
        __syncthreads();   // the pairs are read before the FFT writes the column
        column_fft<R, H, false, 0>(v, col, j, reinterpret_cast<const C2<R>*>(a.tw));
        if (live) {
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int n = j + i * T;
                if (MODE == 0) {
                    if (2 * n < a.nvalid) dst[2 * n] = (double)v[i].x;
                    if (2 * n + 1 < a.nvalid) dst[2 * n + 1] = (double)v[i].y;
                } else if (MODE == 2 || i < 4) {
                    if (a.nt) st_nt(dst + 2 * n, (double)v[i].x, (double)v[i].y);
                    else *reinterpret_cast<double2*>(dst + 2 * n) = double2{(double)v[i].x, (double)v[i].y};
                }
            }
        }
    }
};

// normfact * Re G of the general path's complex (nx, n1, n2) Green's function,
// one octant, laid out [ky'][tile][kz'][tx] with tiles of tx x columns (zero past nx)
__global__ void compress_green(const double* g, double* out, float* out32, int nx, int n1, int n2, int gy, int gz, int tx,
                               int ntx, double normfact) {
    const long long tot = (long long)gy * ntx * gz * tx;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e % tx);
        const int kz = (int)((e / tx) % gz);
        const int tile = (int)((e / ((long long)tx * gz)) % ntx);
        const int ky = (int)(e / ((long long)tx * gz * ntx));
        const int x = tile * tx + c;
        double v = 0.0;
        if (x < nx) v = normfact * g[2 * (x + (long long)nx * (ky + (long long)n1 * kz))];
        if (out) out[e] = v;
        if (out32) out32[e] = (float)v;
    }
}

std::vector<double> twiddle_table(int M) {
    std::vector<double> t(2 * (size_t)M);
    for (int i = 0; i < M; ++i) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)i / (long double)M;
        t[2 * i] = (double)cosl(a);
        t[2 * i + 1] = (double)sinl(a);
    }
    return t;
}

bool pow2_in(int m, int lo, int hi) { return m >= lo && m <= hi && (m & (m - 1)) == 0; }

}  // namespace

// ---------------------------------------------------------------------------
struct FastPlan {
    int n[3];          // physical extents
    int M[3];          // transform lengths
    int nx, ldx;       // complex x extent and padded row length
    int lda;
    bool fp32;
    int zsched = 20;         // z pass schedule (see launch_col): half-width tiles, XCD-grouped
    int ysched = 0;          // y passes: bit 2 half-width tiles, bit 4 XCD-grouped
    int order = 12;          // traversal: bit 0 y_fwd, 1 y_inv, 2 x_c2r, 3 x_r2c walk their batch backwards
    size_t nt_bytes = (size_t)128 << 20;   // stage arrays from this size on: non-temporal accesses (FLUPS_NT_MB)
    int gtx = 0;             // x tile width of the compressed Green's layout
    RotK rot{};              // wave-number data of the rotational pass
    void* s0 = nullptr;     // (ldx, N1, N2) complex, FP64 or FP32
    void* s1 = nullptr;     // one x slab of the y-transformed field: (ldxs, M1, N2) complex, z stride zs
    long long ldxs = 0;     // S1 row length: the widest slab
    long long zs = 0;       // z stride of S1 in complex elements: ldxs * M1 + zpad
    std::vector<int> slab;  // first x column of each slab, then nx
    void* green = nullptr;  // normfact * G, [M1/2+1][x tiles][M2/2+1][tile width of the z kernel]
    void* tw[3] = {nullptr, nullptr, nullptr};   // M1, M2 tables and the M0/2 table
    void* twx = nullptr;                          // M0 table (x post-twiddles)
    std::vector<void*> owned;
    // per-kernel timing (HIP events on the solve stream)
    bool profile = false;
    std::vector<hipEvent_t> ev;   // x_r2c, then y_fwd / z / y_inv of every slab, then x_c2r
    double ms[5] = {0, 0, 0, 0, 0};
    long calls = 0;
    double bytes[5] = {0, 0, 0, 0, 0};   // algorithmic HBM bytes per launch
    size_t dev_bytes = 0;                 // device memory held by the plan
};

void build_fast_plan(SolverImpl& s, FastPlan* f, size_t es, int align);
void destroy_fast_plan(FastPlan* f);

FastPlan* make_fast_plan(SolverImpl& s) {
    if (s.ndim_ != 3 || s.phys_.axis != 0) return nullptr;
    // every direction unbounded or (cell-centred) periodic, transformed in the order x, y, z,
    // with no offsets or post-corrections (node-centred periodic directions need one)
    for (int ip = 0; ip < 3; ++ip) {
        const DirPlan* ps[3] = {&s.fwd_[ip], &s.bwd_[ip], &s.gpl_[ip]};
        const PType t = s.fwd_[ip].type;
        if (!(t == UNBUNB || (t == PERPER && !s.node_)) || s.fwd_[ip].dim != ip) return nullptr;
        for (const DirPlan* p : ps)
            for (int c = 0; c < p->lda; ++c)
                if (p->post[c] != 0 || p->start_in[c] != 0 || p->start_out[c] != 0 || p->fieldstart != 0 || p->koffset != 0.0)
                    return nullptr;
    }
    if (getenv("FLUPS_FORCE_GENERIC")) return nullptr;
    auto* f = new FastPlan;
    for (int d = 0; d < 3; ++d) {
        f->n[d] = s.phys_.nglob[d];
        f->M[d] = s.fwd_[d].n_in[0];
    }
    // supported lengths: powers of two (x: packed length M0/2 in [16, 1024]; y, z: [32, 1024])
    if (!pow2_in(f->M[0] / 2, 16, 1024) || !pow2_in(f->M[1], 32, 1024) || !pow2_in(f->M[2], 32, 1024)) {
        delete f;
        return nullptr;
    }
    f->lda = s.lda_;
    f->fp32 = s.fp32_;
    for (int ip = 0; ip < 3; ++ip) {
        f->rot.kfact[ip] = s.fwd_[ip].kfact;
        f->rot.symstart[ip] = s.fwd_[ip].symstart;
        f->rot.h[ip] = s.h_[ip];
    }
    f->rot.order = s.odiff_;
    f->nx = f->M[0] / 2 + 1;
    const int align = f->fp32 ? 16 : 8;
    f->ldx = (f->nx + align - 1) / align * align;
    const size_t es = f->fp32 ? 2 * sizeof(float) : 2 * sizeof(double);
    // x slabs: with FLUPS_SLAB_MB = S > 0, y_fwd, z and y_inv run slab after slab on one S1
    // buffer of about S MiB instead of the whole y-transformed spectrum (S1 of a 256^3 FP64
    // solve: 553 MB), for callers short of device memory. The slabs split the tiles of
    // align columns evenly. Default: one slab. Narrow slabs cost time: each launch is only
    // a few rounds of blocks deep (64 MiB slabs: about 10% slower at 256^3).
    {
        const int tiles = f->ldx / align;
        const double col_bytes = (double)f->M[1] * f->n[2] * es * f->lda;   // one x column of S1
        double target = 0;
        if (const char* e = getenv("FLUPS_SLAB_MB")) target = atof(e) * (1 << 20);
        const int per = target > 0 ? std::max(1, (int)(target / (col_bytes * align))) : tiles;
        const int nslab = (tiles + per - 1) / per;
        int maxw = 0;
        for (int k = 0; k < nslab; ++k) {
            const int t0 = (int)((long long)k * tiles / nslab), t1 = (int)((long long)(k + 1) * tiles / nslab);
            f->slab.push_back(t0 * align);
            maxw = std::max(maxw, (t1 - t0) * align);
        }
        f->slab.push_back(f->nx);
        f->ldxs = maxw;
    }
    // A pad of 1 KiB after each z plane of S1: with ldxs * M1 a large power of two, the rows
    // of a z column would otherwise fall on the same few HBM channels (a plain copy along z
    // then runs at 3.8 instead of 5 TB/s).
    long long zpad = 1024 / (long long)es;
    if (const char* e = getenv("FLUPS_ZPAD")) zpad = atoll(e);
    if (zpad < 0 || zpad % align) {
        delete f;
        fail("FLUPS_ZPAD: a non-negative multiple of %d (complex elements)", align);
    }
    f->zs = f->ldxs * f->M[1] + zpad;
    // per-lane byte offsets of the column kernels are 32-bit: (M/8) rows of the largest stride
    if ((double)(f->M[2] / 8) * f->zs * es >= 4294967296.0) {
        delete f;
        return nullptr;
    }
    if (const char* e = getenv("FLUPS_ZSCHED")) f->zsched = atoi(e);
    if (const char* e = getenv("FLUPS_ORDER")) f->order = atoi(e) & 15;
    if (const char* e = getenv("FLUPS_YSCHED")) f->ysched = atoi(e) & 20;
    if (const char* e = getenv("FLUPS_NT_MB")) f->nt_bytes = (size_t)(atof(e) * (1 << 20));
    if (f->zsched < 0 || f->zsched > 63 || (f->zsched & 3) == 3 || ((f->zsched & 8) && (f->zsched & 7))) {
        delete f;
        fail("FLUPS_ZSCHED: bits 0-1 schedule (0, 1, 2), bit 2 half-width tiles, bit 3 quarter-width tiles (schedule 0), "
             "bit 4 XCD-grouped blocks, bit 5 mirror rows in neighbouring blocks");
    }
    try {
        build_fast_plan(s, f, es, align);
    } catch (...) {
        destroy_fast_plan(f);
        throw;
    }
    return f;
}

// x tile width of the z kernel for the plan's schedule: half the stage tile for
// M2 = 1024 (LDS) or for half-width schedules
static int z_tile(const FastPlan* f) {
    const int align = f->fp32 ? 16 : 8;
    if (f->M[2] == 1024) return align / 2;
    return (f->zsched & 8) ? align / 4 : (f->zsched & 4) ? align / 2 : align;
}

// (re)write the compressed Green's function in the layout of the current z schedule
static void layout_green(FastPlan* f, const double* dgreen, double normfact, hipStream_t st) {
    const int gtx = z_tile(f);
    if (gtx == f->gtx) return;
    FLUPS_REQUIRE(dgreen, "the full Green's function is needed to change the z tile width");
    const int gy = f->M[1] / 2 + 1, gz = f->M[2] / 2 + 1;
    hipLaunchKernelGGL(compress_green, dim3(4096), dim3(256), 0, st, dgreen, f->fp32 ? nullptr : (double*)f->green,
                       f->fp32 ? (float*)f->green : nullptr, f->nx, f->M[1], f->M[2], gy, gz, gtx, (f->nx + gtx - 1) / gtx,
                       normfact);
    HIP_CHECK(hipGetLastError());
    f->gtx = gtx;
}

void build_fast_plan(SolverImpl& s, FastPlan* f, size_t es, int align) {
    const size_t n0 = (size_t)f->ldx * f->n[1] * f->n[2] * f->lda;
    const size_t n1 = (size_t)f->zs * f->n[2] * f->lda;
    HIP_CHECK(hipMalloc(&f->s0, n0 * es));
    HIP_CHECK(hipMalloc(&f->s1, n1 * es));
    f->dev_bytes += (n0 + n1) * es;
    HIP_CHECK(hipMemset(f->s0, 0, n0 * es));
    HIP_CHECK(hipMemset(f->s1, 0, n1 * es));
    const int gy = f->M[1] / 2 + 1, gz = f->M[2] / 2 + 1;
    // room for the widest Green's layout (full-width tiles: ldx columns)
    const size_t gn = (size_t)gz * gy * f->ldx;
    HIP_CHECK(hipMalloc(&f->green, gn * (f->fp32 ? sizeof(float) : sizeof(double))));
    f->dev_bytes += gn * (f->fp32 ? sizeof(float) : sizeof(double));
    // the Green's function of the general path: complex (nx, M1, M2), FP64
    const Shape& gs = s.stage_[2];
    FLUPS_REQUIRE(gs.nf == 2 && gs.n[0] == f->nx && gs.n[1] == f->M[1] && gs.n[2] == f->M[2], "unexpected Green's layout");
    f->gtx = 0;
    layout_green(f, s.dgreen_, s.normfact_, s.stream_);
    auto upload = [&](int M) -> void* {
        const auto t = twiddle_table(M);
        void* d = nullptr;
        if (f->fp32) {
            std::vector<float> tf(t.begin(), t.end());
            HIP_CHECK(hipMalloc(&d, tf.size() * sizeof(float)));
            HIP_CHECK(hipMemcpy(d, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
        } else {
            HIP_CHECK(hipMalloc(&d, t.size() * sizeof(double)));
            HIP_CHECK(hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        f->owned.push_back(d);
        f->dev_bytes += t.size() * (f->fp32 ? sizeof(float) : sizeof(double));
        return d;
    };
    f->tw[0] = upload(f->M[0] / 2);
    f->tw[1] = upload(f->M[1]);
    f->tw[2] = upload(f->M[2]);
    f->twx = upload(f->M[0]);
    if (f->lda == 3) {   // derivative symbols of the rotational pass
        const int nsym[3] = {f->nx, f->M[1], f->M[2]};
        std::vector<double> t;
        for (int d = 0; d < 3; ++d) {
            f->rot.off[d] = (int)t.size();
            for (int i = 0; i < nsym[d]; ++i) t.push_back(rot_symbol(f->rot, d, i));
        }
        void* d = nullptr;
        if (f->fp32) {
            std::vector<float> tf(t.begin(), t.end());
            HIP_CHECK(hipMalloc(&d, tf.size() * sizeof(float)));
            HIP_CHECK(hipMemcpy(d, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
        } else {
            HIP_CHECK(hipMalloc(&d, t.size() * sizeof(double)));
            HIP_CHECK(hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        f->owned.push_back(d);
        f->dev_bytes += t.size() * (f->fp32 ? sizeof(float) : sizeof(double));
        f->rot.sym = d;
    }
    f->ev.assign(3 + 3 * (f->slab.size() - 1), nullptr);
    for (auto& e : f->ev) HIP_CHECK(hipEventCreate(&e));
    {
        // minimum traffic of each pass: valid elements read + written (plus G for the fused pass)
        const double us = sizeof(double);                          // user arrays are always FP64
        const double rs = f->fp32 ? sizeof(float) : sizeof(double), cs = 2 * rs;
        const double N = (double)f->n[0] * f->n[1] * f->n[2];
        const double A = (double)f->nx * f->n[1] * f->n[2];        // S0 complex elements
        const double B = (double)f->nx * f->M[1] * f->n[2];        // S1 complex elements
        const double G = (double)gz * gy * f->nx * rs;
        const double L = f->lda;
        f->bytes[0] = L * (N * us + A * cs);
        f->bytes[1] = L * (A * cs + B * cs);
        f->bytes[2] = L * (2 * B * cs) + G;
        f->bytes[3] = L * (B * cs + A * cs);
        f->bytes[4] = L * (A * cs + N * us);
    }
    HIP_CHECK(hipStreamSynchronize(s.stream_));
}

void destroy_fast_plan(FastPlan* f) {
    if (!f) return;
    (void)hipFree(f->s0);
    (void)hipFree(f->s1);
    (void)hipFree(f->green);
    for (void* p : f->owned) (void)hipFree(p);
    for (hipEvent_t e : f->ev)
        if (e) (void)hipEventDestroy(e);
    delete f;
}

namespace {

// S: storage of the z_fused field and G (S = R but for the general path's FP32 solves)
template <typename R, int M, int TX, bool FWD, bool FUSED, bool HALF, typename S = R>
void launch_col_h(const ColArgs& a, int comps, int zsched, hipStream_t st) {
    if constexpr (FUSED && !std::is_same<R, S>::value) {
        // the general path's FP32 fused last direction: full- or half-width tiles, schedule 0
        const int tx = (zsched & 4) ? TX / 2 : TX;
        dim3 grid(((a.nx + tx - 1) / tx) * (a.nopair == 1 ? a.nother : a.gmy / 2 + 1), comps);
        if (zsched & 4) launch_k<z_fused<R, M, TX / 2, 0, HALF, 1, S>>(grid, dim3(TX / 2 * M / 8), 0, st, a);
        else launch_k<z_fused<R, M, TX, 0, HALF, 1, S>>(grid, dim3(TX * M / 8), 0, st, a);
    } else if (FUSED) {
        // z schedules: bit 0-1 = SCHED of z_fused, 4 = half-width tiles (TX/2 columns)
        const int tx = (zsched & 8) ? TX / 4 : (zsched & 4) ? TX / 2 : TX;
        dim3 grid(((a.nx + tx - 1) / tx) * (a.nopair == 1 ? a.nother : a.nopair == 2 ? 2 * (a.gmy / 2 + 1) : a.gmy / 2 + 1), comps);
        switch (zsched & 15) {
            case 0: launch_k<z_fused<R, M, TX, 0, HALF, 1, S>>(grid, dim3(TX * M / 8), 0, st, a); break;
            case 1: launch_k<z_fused<R, M, TX, 1, HALF, 1, S>>(grid, dim3(TX * M / 8), 0, st, a); break;
            case 2: launch_k<z_fused<R, M, TX, 2, HALF, 1, S>>(grid, dim3(TX * M / 8), 0, st, a); break;
            case 4: launch_k<z_fused<R, M, TX / 2, 0, HALF, 1, S>>(grid, dim3(TX / 2 * M / 8), 0, st, a); break;
            case 5: launch_k<z_fused<R, M, TX / 2, 1, HALF, 1, S>>(grid, dim3(TX / 2 * M / 8), 0, st, a); break;
            case 6: launch_k<z_fused<R, M, TX / 2, 2, HALF, 1, S>>(grid, dim3(TX / 2 * M / 8), 0, st, a); break;
            case 8:
                launch_k<z_fused<R, M, (TX >= 8 ? TX / 4 : 1), 0, HALF, 1, S>>(grid, dim3((TX >= 8 ? TX / 4 : 1) * M / 8), 0, st, a);
                break;
            default: fail("unknown z schedule %d", zsched);
        }
    } else if (zsched & 4) {   // y schedule: half-width tiles
        dim3 grid(((a.nx + TX / 2 - 1) / (TX / 2)) * a.nother, comps);
        launch_k<col_fft<R, M, TX / 2, FWD, HALF>>(grid, dim3(TX / 2 * M / 8), 0, st, a);
    } else {
        dim3 grid(((a.nx + TX - 1) / TX) * a.nother, comps);
        launch_k<col_fft<R, M, TX, FWD, HALF>>(grid, dim3(TX * M / 8), 0, st, a);
    }
    HIP_CHECK(hipGetLastError());
}

// HALF kernels when the column is zero-padded to twice its length (unbounded
// directions): fused pass nin = nout = M/2, forward nin = M/2, inverse nout = M/2
template <typename R, int M, int TX, bool FWD, bool FUSED, typename S = R>
void launch_col(const ColArgs& a, int comps, int zsched, hipStream_t st) {
    const int h = M / 2;
    const bool half = FUSED ? (a.nin == h && a.nout == h) : FWD ? (a.nin == h && a.nout == M) : (a.nin == M && a.nout == h);
    if (half) launch_col_h<R, M, TX, FWD, FUSED, true, S>(a, comps, zsched, st);
    else launch_col_h<R, M, TX, FWD, FUSED, false, S>(a, comps, zsched, st);
}

template <typename R, int TX, bool FWD, bool FUSED, typename S = R>
void col_dispatch(int M, const ColArgs& a, int comps, int zsched, hipStream_t st) {
    switch (M) {
        case 32: launch_col<R, 32, TX, FWD, FUSED, S>(a, comps, zsched, st); break;
        case 64: launch_col<R, 64, TX, FWD, FUSED, S>(a, comps, zsched, st); break;
        case 128: launch_col<R, 128, TX, FWD, FUSED, S>(a, comps, zsched, st); break;
        case 256: launch_col<R, 256, TX, FWD, FUSED, S>(a, comps, zsched, st); break;
        case 512: launch_col<R, 512, TX, FWD, FUSED, S>(a, comps, zsched, st); break;
        case 1024: launch_col<R, 1024, TX / 2, FWD, FUSED, S>(a, comps, zsched & ~12, st); break;   // already half width
        default: fail("unsupported column length %d", M);
    }
}

template <typename R, int TX>
void launch_rot(int M, const ColArgs& a, const RotK& rk, hipStream_t st) {
    dim3 grid(((a.nx + TX - 1) / TX) * (a.gmy / 2 + 1), 1);
    switch (M) {
        case 32: launch_k<z_rot<R, 32, TX>>(grid, dim3(TX * 4), 0, st, {a, rk}); break;
        case 64: launch_k<z_rot<R, 64, TX>>(grid, dim3(TX * 8), 0, st, {a, rk}); break;
        case 128: launch_k<z_rot<R, 128, TX>>(grid, dim3(TX * 16), 0, st, {a, rk}); break;
        case 256: launch_k<z_rot<R, 256, TX>>(grid, dim3(TX * 32), 0, st, {a, rk}); break;
        case 512: launch_k<z_rot<R, 512, TX>>(grid, dim3(TX * 64), 0, st, {a, rk}); break;
        case 1024: launch_k<z_rot<R, 1024, TX>>(grid, dim3(TX * 128), 0, st, {a, rk}); break;
        default: fail("unsupported column length %d", M);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename R, int MODE>
void row_launch(bool fwd, int H, const RowArgs& a, int comps, hipStream_t st) {
#define FLUPS_ROW(HH, RR)                                                                                   \
    case HH: {                                                                                              \
        dim3 grid((a.nrows + RR - 1) / RR, comps);                                                          \
        if (fwd) launch_k<x_r2c<R, HH, RR, MODE>>(grid, dim3(RR * HH / 8), 0, st, a);                       \
        else launch_k<x_c2r<R, HH, RR, MODE>>(grid, dim3(RR * HH / 8), 0, st, a);                           \
        break;                                                                                              \
    }
    switch (H) {
        FLUPS_ROW(16, 128)
        FLUPS_ROW(32, 64)
        FLUPS_ROW(64, 32)
        FLUPS_ROW(128, 16)
        FLUPS_ROW(256, 8)
        FLUPS_ROW(512, 4)
        FLUPS_ROW(1024, 2)
        default: fail("unsupported row length %d", 2 * H);
    }
#undef FLUPS_ROW
    HIP_CHECK(hipGetLastError());
}

// user rows of 16-byte aligned pairs and a zero-padded (nvalid = H) or full (2H) row:
// the vectorised kernels; anything else the scalar one
template <typename R>
void row_dispatch(bool fwd, int H, const RowArgs& a, int comps, hipStream_t st) {
    const void* user = fwd ? a.src : a.dst;
    const long long s_row = fwd ? a.s_row_src : a.s_row_dst, s_z = fwd ? a.s_z_src : a.s_z_dst;
    const long long sc = fwd ? a.comp_src : a.comp_dst;
    const bool aligned = ((uintptr_t)user % 16) == 0 && s_row % 2 == 0 && s_z % 2 == 0 && (comps == 1 || sc % 2 == 0);
    if (aligned && a.nvalid == H) row_launch<R, 1>(fwd, H, a, comps, st);
    else if (aligned && a.nvalid == 2 * H) row_launch<R, 2>(fwd, H, a, comps, st);
    else row_launch<R, 0>(fwd, H, a, comps, st);
}

// s0 / s1: the stage arrays (the plan's own, or caller-provided ones of fast_stage_bytes)
template <typename R>
void run_fast(FastPlan* f, void* s0, void* s1, const double* rhs, double* field, const View& uin, const View& uout,
              hipStream_t st, int type = 0) {
    const int N0 = f->n[0], N1 = f->n[1], N2 = f->n[2];
    const int M0 = f->M[0], M1 = f->M[1], M2 = f->M[2];
    const int H = M0 / 2;
    const long long ldx = f->ldx;
    const long long zs = f->zs;                                // z stride of S1
    const long long c0 = ldx * N1 * N2, c1 = zs * N2;          // component strides (complex)
    constexpr int TX = sizeof(R) == 8 ? 8 : 16;
    // non-temporal accesses where the stage arrays exceed half the 256 MB Infinity Cache
    const bool nt0 = (size_t)c0 * f->lda * sizeof(C2<R>) >= f->nt_bytes;
    const bool nt1 = (size_t)zs * N2 * f->lda * sizeof(C2<R>) >= f->nt_bytes;
    // 1. x r2c: user (real) -> S0
    RowArgs r{};
    r.rev = (f->order >> 3) & 1;
    r.nt = nt0;
    r.src = rhs;
    r.dst = s0;
    r.tw = f->tw[0];
    r.tw2 = f->twx;
    r.s_row_src = uin.s[1];
    r.s_z_src = uin.s[2];
    r.s_row_dst = ldx;
    r.s_z_dst = ldx * N1;
    r.n1 = N1;
    r.nrows = N1 * N2;
    r.nvalid = N0;
    r.nx = f->nx;
    r.comp_src = uin.sc;
    r.comp_dst = c0;
    if (f->profile) HIP_CHECK(hipEventRecord(f->ev[0], st));
    row_dispatch<R>(true, H, r, f->lda, st);
    if (f->profile) HIP_CHECK(hipEventRecord(f->ev[1], st));
    const size_t es = sizeof(C2<R>);
    const int nslab = (int)f->slab.size() - 1;
    for (int k = 0; k < nslab; ++k) {
        const int x0 = f->slab[k], w = f->slab[k + 1] - x0;
        // 2. y forward: columns [x0, x0 + w) of S0 (ldx, N1, N2) -> S1 (ldxs, M1, N2)
        ColArgs a{};
        a.rev = f->order & 1;
        a.nt = nt1;
        a.xcd = (f->ysched >> 4) & 1;
        a.src = reinterpret_cast<const char*>(s0) + es * x0;
        a.dst = s1;
        a.tw = f->tw[1];
        a.nx = w;
        a.ldx = (int)f->ldxs;
        a.s_col = ldx;
        a.s_col_dst = f->ldxs;
        a.s_other_src = ldx * N1;
        a.s_other_dst = zs;
        a.nother = N2;
        a.nin = N1;
        a.nout = M1;
        a.comp_src = c0;
        a.comp_dst = c1;
        col_dispatch<R, TX, true, false>(M1, a, f->lda, f->ysched, st);
        if (f->profile) HIP_CHECK(hipEventRecord(f->ev[2 + 3 * k], st));
        // 3. z fused: S1 in place, forward, * normfact G, inverse, keep N2
        ColArgs z{};
        z.src = s1;
        z.dst = s1;
        z.green = f->green;
        z.tw = f->tw[2];
        z.nx = w;
        z.ldx = (int)f->ldxs;
        z.s_col = zs;
        z.s_col_dst = zs;
        z.s_other_src = f->ldxs;
        z.s_other_dst = f->ldxs;
        z.nother = M1;
        z.nin = N2;
        z.nout = N2;
        z.gy = M1 / 2 + 1;
        z.gmy = M1;
        z.gmz = M2;
        z.comp_src = c1;
        z.comp_dst = c1;
        z.xcd = (f->zsched >> 4) & 1;
        z.nopair = (f->zsched & 32) ? 2 : 0;
        z.x0 = x0;
        z.gntx = (f->nx + f->gtx - 1) / f->gtx;
        if (type == 1) {
            // rotational pass: all three components per block, tile width of the Green's layout
            FLUPS_REQUIRE(f->lda == 3, "the rotational solve needs lda = 3");
            z.gtx = f->gtx;
            if (M2 == 1024 && f->gtx == TX / 2 && sizeof(R) == 8) {
                // M2 = 1024, FP64: the Green's layout width, four columns of 64-byte rows per block
                // of 512 threads (two columns of 32-byte rows: 28.5 instead of 27.8 ms at 513^3 x 3);
                // FP32 keeps a quarter of the stage tile (eight columns would need 1024 threads of
                // 157 registers)
                launch_rot<R, TX / 2>(M2, z, f->rot, st);
            } else if (M2 == 1024 && f->gtx == TX / 2) launch_rot<R, TX / 4>(M2, z, f->rot, st);
            else if (f->gtx == TX) launch_rot<R, TX>(M2, z, f->rot, st);
            else if (f->gtx == TX / 2) launch_rot<R, TX / 2>(M2, z, f->rot, st);
            else launch_rot<R, (TX >= 8 ? TX / 4 : 1)>(M2, z, f->rot, st);
        } else {
            col_dispatch<R, TX, true, true>(M2, z, f->lda, f->zsched, st);
        }
        if (f->profile) HIP_CHECK(hipEventRecord(f->ev[3 + 3 * k], st));
        // 4. y inverse: S1 -> columns [x0, x0 + w) of S0, keep N1
        ColArgs b = a;
        b.rev = (f->order >> 1) & 1;
        b.crot = (type == 1 && sizeof(R) == 4) ? 1 : 0;   // FP32 z_rot leaves u_c in slot (c + 1) mod 3
        b.src = s1;
        b.dst = reinterpret_cast<char*>(s0) + es * x0;
        b.s_col = f->ldxs;
        b.s_col_dst = ldx;
        b.s_other_src = zs;
        b.s_other_dst = ldx * N1;
        b.nin = M1;
        b.nout = N1;
        b.comp_src = c1;
        b.comp_dst = c0;
        col_dispatch<R, TX, false, false>(M1, b, f->lda, f->ysched, st);
        if (f->profile) HIP_CHECK(hipEventRecord(f->ev[4 + 3 * k], st));
    }
    // 5. x c2r: S0 -> user (real), keep N0
    RowArgs q{};
    q.rev = (f->order >> 2) & 1;
    q.nt = nt0;
    q.src = s0;
    q.dst = field;
    q.tw = f->tw[0];
    q.tw2 = f->twx;
    q.s_row_src = ldx;
    q.s_z_src = ldx * N1;
    q.s_row_dst = uout.s[1];
    q.s_z_dst = uout.s[2];
    q.n1 = N1;
    q.nrows = N1 * N2;
    q.nvalid = N0;
    q.nx = f->nx;
    q.comp_src = c0;
    q.comp_dst = uout.sc;
    row_dispatch<R>(false, H, q, f->lda, st);
    if (f->profile) HIP_CHECK(hipEventRecord(f->ev.back(), st));
}

}  // namespace

void fast_solve(FastPlan* f, double* field, const double* rhs, const View& uin, const View& uout, hipStream_t s, int type) {
    if (f->fp32) run_fast<float>(f, f->s0, f->s1, rhs, field, uin, uout, s, type);
    else run_fast<double>(f, f->s0, f->s1, rhs, field, uin, uout, s, type);
}

// bytes of the two stage arrays S0 and S1 (each a whole number of 256-byte blocks)
void fast_stage_bytes(const FastPlan* f, size_t* b0, size_t* b1) {
    const size_t es = f->fp32 ? 2 * sizeof(float) : 2 * sizeof(double);
    const size_t n0 = (size_t)f->ldx * f->n[1] * f->n[2] * f->lda;
    const size_t n1 = (size_t)f->zs * f->n[2] * f->lda;
    *b0 = (n0 * es + 255) / 256 * 256;
    *b1 = (n1 * es + 255) / 256 * 256;
}

// the same solve on caller-provided stage arrays of fast_stage_bytes (no zero fill needed:
// a pass reads only what the pass before it wrote, or padding columns whose values never
// reach a stored element)
void fast_solve_on(FastPlan* f, void* s0, void* s1, double* field, const double* rhs, const View& uin, const View& uout,
                   hipStream_t s, int type) {
    if (f->fp32) run_fast<float>(f, s0, s1, rhs, field, uin, uout, s, type);
    else run_fast<double>(f, s0, s1, rhs, field, uin, uout, s, type);
}

size_t fast_device_bytes(const FastPlan* f) { return f->dev_bytes; }

// ---------------------------------------------------------------------------
// Last direction of the general path fused with the multiply: when the final
// forward transform is a c2c along z (unbounded or cell-centred periodic) of a
// power-of-two length, the forward z FFT, the Green's multiply and the inverse
// z FFT run as one z_fused launch in place on the stage array before it (the
// general passes do the other directions). The other dim (y) may be a
// half-spectrum (r2c) or real-to-real direction: rows are not paired then.
// 2-D problems fuse their last direction (y) the same way, as a 3-D field with
// one row (columns up to 2048 long, two columns per block).
void launch_col_rz(const RzArgs& A, hipStream_t s);   // rowpass.hip
void launch_row_rz(const RzArgs& A, hipStream_t s);
bool row_rz_len_ok(int M);
bool split_pq(int n, int& pf, int& q);

struct HybridZ {
    int M = 0, nx = 0, ny = 0, nz = 0, lda = 1, gtx = 8;
    bool real = false;       // r2c last direction (col_rz): the general path's Green's function, real columns
    bool xrow = false;       // r2c last direction along x (row_rz): real rows
    void* twn = nullptr;     // row_rz: n-point table
    double normfact = 1;
    const void* g = nullptr;     // real: the solver's spectral Green's function (not owned)
    bool f32 = false;            // FP32 stage arrays and Green's function (FP64 arithmetic)
    int pf = 1;              // M = pf 2^k: 1, or 3 (mixed-radix column FFT, 8 pf values per thread)
    bool pair = false, twod = false;
    void* green = nullptr;   // normfact * Re G, [y'][tile][kz'][gtx]
    void* tw = nullptr;      // (M / pf)-point table
    void* twm = nullptr;     // M-point table (pf 3)
    size_t dev_bytes = 0;
    double bytes = 0;        // algorithmic bytes of one launch
};

HybridZ* make_hybrid_z(SolverImpl& s) {
    const int nd = s.ndim_;
    if (nd != 3 && nd != 2) return nullptr;
    const bool f32 = s.gf32_;           // the general path's FP32 solves: float data and Green's function
    const double es = f32 ? 4.0 : 8.0;   // bytes of one stored real
    if (getenv("FLUPS_FORCE_GENERIC")) return nullptr;
    const int ip = nd - 1;
    const DirPlan &f = s.fwd_[ip], &b = s.bwd_[ip];
    if (!(f.type == UNBUNB || (f.type == PERPER && !s.node_))) return nullptr;
    if (f.dim != nd - 1 && !(f.dim == 0 && f.r2c)) return nullptr;
    for (const DirPlan* p : {&f, &b})
        for (int c = 0; c < p->lda; ++c)
            if (p->post[c] != 0 || p->start_in[c] != 0 || p->start_out[c] != 0 || p->fieldstart != 0 || p->koffset != 0.0)
                return nullptr;
    const Shape& in = s.stage_[nd - 2];   // field before the last transform
    const Shape& sp = s.stage_[nd - 1];   // spectral field (and Green's function)
    if (f.r2c && f.dim == 0) {
        // x last (every other direction real-to-real): rows of the real field, the Green's
        // function's rows of the general path's spectral layout
        const int M = f.n_in[0];
        if (in.nf != 1 || sp.nf != 2 || !row_rz_len_ok(M) || sp.n[0] != M / 2 + 1 || in.n[0] > M || !s.dgreen_) return nullptr;
        for (int d = 1; d < nd; ++d)
            if (in.n[d] != sp.n[d]) return nullptr;
        auto* h = new HybridZ;
        h->f32 = f32;
        h->real = true;
        h->xrow = true;
        h->M = M;
        h->nx = in.n[0];
        h->ny = nd == 3 ? in.n[1] * in.n[2] : in.n[1];   // rows
        h->nz = 1;
        h->lda = s.lda_;
        h->normfact = s.normfact_;
        h->g = f32 ? (const void*)s.dgreen32_ : (const void*)s.dgreen_;
        int pf = 1, q = 0;
        split_pq(M / 2, pf, q);
        const auto t = twiddle_table(q), tn = twiddle_table(M);
        try {
            HIP_CHECK(hipMalloc(&h->tw, t.size() * sizeof(double)));
            HIP_CHECK(hipMemcpy(h->tw, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
            HIP_CHECK(hipMalloc(&h->twn, tn.size() * sizeof(double)));
            HIP_CHECK(hipMemcpy(h->twn, tn.data(), tn.size() * sizeof(double), hipMemcpyHostToDevice));
        } catch (...) {
            (void)hipFree(h->tw);
            delete h;
            throw;
        }
        h->dev_bytes = (t.size() + tn.size()) * sizeof(double);
        h->bytes = h->lda * 2.0 * (double)h->nx * h->ny * es + (double)h->ny * (M / 2 + 1) * 2 * es;
        return h;
    }
    if (f.r2c) {
        // every other direction real-to-real: the field before the last transform is real and
        // the Green's function keeps the general path's spectral layout
        const int M = f.n_in[0];
        if (in.nf != 1 || sp.nf != 2 || in.n[0] != sp.n[0] || (nd == 3 && in.n[1] != sp.n[1])) return nullptr;
        if (!pow2_in(M, 32, 1024) || sp.n[nd - 1] != M / 2 + 1 || in.n[nd - 1] > M || !s.dgreen_) return nullptr;
        auto* h = new HybridZ;
        h->f32 = f32;
        h->real = true;
        h->M = M;
        h->nx = in.n[0];
        h->ny = nd == 3 ? in.n[1] : 1;
        h->nz = in.n[nd - 1];
        h->lda = s.lda_;
        h->normfact = s.normfact_;
        h->g = f32 ? (const void*)s.dgreen32_ : (const void*)s.dgreen_;
        const auto t = twiddle_table(M);
        try {
            HIP_CHECK(hipMalloc(&h->tw, t.size() * sizeof(double)));
            HIP_CHECK(hipMemcpy(h->tw, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        } catch (...) {
            delete h;
            throw;
        }
        h->dev_bytes = t.size() * sizeof(double);
        // the field in and out, the Green's function's (M/2 + 1) rows of complex values
        h->bytes = h->lda * 2.0 * (double)h->nx * h->ny * h->nz * es + (double)h->nx * h->ny * (M / 2 + 1) * 2 * es;
        return h;
    }
    if (in.nf != 2 || sp.nf != 2 || in.n[0] != sp.n[0] || (nd == 3 && in.n[1] != sp.n[1])) return nullptr;
    const int M = f.n_in[0];
    const bool p2 = pow2_in(M, 32, nd == 2 ? 2048 : 1024), p3 = M % 3 == 0 && pow2_in(M / 3, 8, 128),
               p5 = M % 5 == 0 && pow2_in(M / 5, 8, 128), p7 = M % 7 == 0 && pow2_in(M / 7, 8, 128);
    if (!(p2 || p3 || p5 || p7) || sp.n[nd - 1] != M) return nullptr;
    const int ny = nd == 3 ? in.n[1] : 1;
    if ((double)(M / 8) * in.n[0] * ny * 16.0 >= 4294967296.0) return nullptr;
    auto* h = new HybridZ;
    h->f32 = f32;
    h->M = M;
    h->nx = in.n[0];
    h->ny = ny;
    h->nz = in.n[nd - 1];
    h->lda = s.lda_;
    // y rows pair as (y, M1 - y) when y is a full unbounded / periodic c2c direction (G even in ky)
    const DirPlan& fy = s.fwd_[1];
    h->pair = nd == 3 && fy.dim == 1 && !fy.r2c && (fy.type == UNBUNB || (fy.type == PERPER && !s.node_)) &&
              fy.koffset == 0.0 && fy.n_in[0] == h->ny;
    // half-width x tiles in 3-D; two columns per block in 2-D (few columns, cache-resident field)
    h->gtx = nd == 2 ? 2 : 4;
    h->twod = nd == 2;
    h->pf = p2 ? 1 : p3 ? 3 : p5 ? 5 : 7;
    if (h->pf >= 5) {   // thread teams, one row per block: no mirror pairing, tiles by length (LDS M gtx complex:
                        // 48 KB up to M = 384, then 4 columns: 28 / 40 / 56 KB at M = 448 / 640 / 896)
        h->pair = false;
        h->gtx = M <= 192 ? 16 : M <= 384 ? 8 : 4;
    }
    // 3 Q points: Q/8 threads per column, so wider tiles keep the blocks larger (LDS M gtx
    // complex values, 48 KB at most)
    if (h->pf == 3) h->gtx = nd == 2 ? (M <= 96 ? 16 : 8) : (M <= 48 ? 32 : M <= 192 ? 16 : 8);
    const int ntx = (h->nx + h->gtx - 1) / h->gtx;
    const int gy = h->pair ? h->ny / 2 + 1 : h->ny, gz = M / 2 + 1;
    const size_t gn = (size_t)gy * ntx * gz * h->gtx;
    try {
        HIP_CHECK(hipMalloc(&h->green, gn * (size_t)es));
        h->dev_bytes += gn * (size_t)es;
        hipLaunchKernelGGL(compress_green, dim3(4096), dim3(256), 0, s.stream_, s.dgreen_, f32 ? nullptr : (double*)h->green,
                           f32 ? (float*)h->green : nullptr, h->nx, h->ny, M, gy, gz, h->gtx, ntx, s.normfact_);
        HIP_CHECK(hipGetLastError());
        const auto t = twiddle_table(M / h->pf);
        HIP_CHECK(hipMalloc(&h->tw, t.size() * sizeof(double)));
        HIP_CHECK(hipMemcpy(h->tw, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        h->dev_bytes += t.size() * sizeof(double);
        if (h->pf > 1) {
            const auto tm = twiddle_table(M);
            HIP_CHECK(hipMalloc(&h->twm, tm.size() * sizeof(double)));
            HIP_CHECK(hipMemcpy(h->twm, tm.data(), tm.size() * sizeof(double), hipMemcpyHostToDevice));
            h->dev_bytes += tm.size() * sizeof(double);
        }
        HIP_CHECK(hipStreamSynchronize(s.stream_));
    } catch (...) {
        (void)hipFree(h->green);
        (void)hipFree(h->tw);
        (void)hipFree(h->twm);
        delete h;
        throw;
    }
    const double cnt = (double)h->nx * h->ny * h->nz * 2 * es;
    h->bytes = h->lda * 2.0 * cnt + (double)gy * gz * h->nx * es;
    return h;
}

void destroy_hybrid_z(HybridZ* h) {
    if (!h) return;
    (void)hipFree(h->green);
    (void)hipFree(h->tw);
    (void)hipFree(h->twm);
    (void)hipFree(h->twn);
    delete h;
}
size_t hybrid_device_bytes(const HybridZ* h) { return h->dev_bytes; }
double hybrid_bytes(const HybridZ* h) { return h->bytes; }

// M = 3 Q: TX columns per block (the compressed G's tile width), Q/8 threads per column
template <int M, int TX, typename S>
void launch_z_pf3(const ColArgs& a, int comps, hipStream_t st) {
    dim3 grid(((a.nx + TX - 1) / TX) * (a.nopair == 1 ? a.nother : a.gmy / 2 + 1), comps);
    launch_k<z_fused<double, M, TX, 0, false, 3, S>>(grid, dim3(TX * M / 24), 0, st, a);
    HIP_CHECK(hipGetLastError());
}
template <int TX, typename S>
void dispatch_z_pf3(int M, const ColArgs& a, int comps, hipStream_t st) {
    switch (M) {
        case 24: launch_z_pf3<24, TX, S>(a, comps, st); break;
        case 48: launch_z_pf3<48, TX, S>(a, comps, st); break;
        case 96:
            if constexpr (TX <= 16) launch_z_pf3<96, TX, S>(a, comps, st);
            else fail("column length %d with tiles of %d", M, TX);
            break;
        case 192:
            if constexpr (TX <= 16) launch_z_pf3<192, TX, S>(a, comps, st);
            else fail("column length %d with tiles of %d", M, TX);
            break;
        case 384:
            if constexpr (TX <= 8) launch_z_pf3<384, TX, S>(a, comps, st);
            else fail("column length %d with tiles of %d", M, TX);
            break;
        default: fail("unsupported column length %d", M);
    }
}

template <int PF, int Q, typename S>
void launch_z_team(int tx, const ColArgs& a, int comps, hipStream_t st) {
    const int ntx = (a.nx + tx - 1) / tx;
    dim3 grid(ntx * a.nother, comps);
    switch (tx) {
        case 16:
            if constexpr (PF * Q <= 192) launch_k<z_team<PF, Q, 16, S>>(grid, dim3(16 * PF * Q / 8), 0, st, a);
            else fail("column length %d with tiles of 16", PF * Q);
            break;
        case 8:
            if constexpr (PF * Q <= 384) launch_k<z_team<PF, Q, 8, S>>(grid, dim3(8 * PF * Q / 8), 0, st, a);
            else fail("column length %d with tiles of 8", PF * Q);
            break;
        default: launch_k<z_team<PF, Q, 4, S>>(grid, dim3(4 * PF * Q / 8), 0, st, a); break;
    }
    HIP_CHECK(hipGetLastError());
}
template <int PF, typename S>
void dispatch_z_team(int M, int tx, const ColArgs& a, int comps, hipStream_t st) {
    switch (M / PF) {
        case 8: launch_z_team<PF, 8, S>(tx, a, comps, st); break;
        case 16: launch_z_team<PF, 16, S>(tx, a, comps, st); break;
        case 32: launch_z_team<PF, 32, S>(tx, a, comps, st); break;
        case 64: launch_z_team<PF, 64, S>(tx, a, comps, st); break;
        case 128: launch_z_team<PF, 128, S>(tx, a, comps, st); break;
        default: fail("unsupported column length %d", M);
    }
}

// in place on the stage array before the z transform: (nx, ny, nz) complex (real for an r2c
// last direction), lda components; S = the storage of the stage array (float for FP32 solves)
template <typename S>
void hybrid_z_t(const HybridZ* h, void* data, hipStream_t st) {
    if (h->xrow) {
        RzArgs A;
        A.f32 = std::is_same<S, float>::value;
        A.d = data;
        A.g = h->g;
        A.tw = static_cast<const double*>(h->tw);
        A.twn = static_cast<const double*>(h->twn);
        A.nin = h->nx;
        A.M = h->M;
        A.lda = h->lda;
        A.rows = h->ny;
        A.comp = (long long)h->nx * h->ny;
        A.normfact = h->normfact;
        launch_row_rz(A, st);
        return;
    }
    if (h->real) {
        RzArgs A;
        A.f32 = std::is_same<S, float>::value;
        A.d = data;
        A.g = h->g;
        A.tw = static_cast<const double*>(h->tw);
        A.nx = h->nx;
        A.ny = h->ny;
        A.nin = h->nz;
        A.M = h->M;
        A.lda = h->lda;
        A.s_col = (long long)h->nx * h->ny;
        A.comp = (long long)h->nx * h->ny * h->nz;
        A.normfact = h->normfact;
        launch_col_rz(A, st);
        return;
    }
    ColArgs z{};
    z.src = data;
    z.dst = data;
    z.green = h->green;
    z.tw = h->tw;
    z.nx = h->nx;
    z.ldx = h->nx;
    z.s_col = (long long)h->nx * h->ny;
    z.s_col_dst = z.s_col;
    z.s_other_src = h->nx;
    z.s_other_dst = h->nx;
    z.nother = h->ny;
    z.nin = h->nz;
    z.nout = h->nz;
    z.gy = h->pair ? h->ny / 2 + 1 : h->ny;
    z.gmy = h->pair ? h->ny : 0;
    z.gmz = h->M;
    z.comp_src = (long long)h->nx * h->ny * h->nz;
    z.comp_dst = z.comp_src;
    z.xcd = 1;
    z.nopair = h->pair ? 0 : 1;
    z.x0 = 0;
    z.gntx = (h->nx + h->gtx - 1) / h->gtx;
    z.twm = h->twm;
    if (h->pf >= 5) {
        if (h->pf == 5) dispatch_z_team<5, S>(h->M, h->gtx, z, h->lda, st);
        else dispatch_z_team<7, S>(h->M, h->gtx, z, h->lda, st);
        return;
    }
    if (h->pf == 3) {
        switch (h->gtx) {
            case 32: dispatch_z_pf3<32, S>(h->M, z, h->lda, st); break;
            case 16: dispatch_z_pf3<16, S>(h->M, z, h->lda, st); break;
            default: dispatch_z_pf3<8, S>(h->M, z, h->lda, st); break;
        }
        return;
    }
    if (h->twod) {   // two columns per block (the tile width of the compressed G), XCD-grouped
        switch (h->M) {
            case 32: launch_col<double, 32, 2, true, true, S>(z, h->lda, 16, st); break;
            case 64: launch_col<double, 64, 2, true, true, S>(z, h->lda, 16, st); break;
            case 128: launch_col<double, 128, 2, true, true, S>(z, h->lda, 16, st); break;
            case 256: launch_col<double, 256, 2, true, true, S>(z, h->lda, 16, st); break;
            case 512: launch_col<double, 512, 2, true, true, S>(z, h->lda, 16, st); break;
            case 1024: launch_col<double, 1024, 2, true, true, S>(z, h->lda, 16, st); break;
            case 2048: launch_col<double, 2048, 2, true, true, S>(z, h->lda, 16, st); break;
            default: fail("unsupported column length %d", h->M);
        }
        return;
    }
    // half-width tiles (schedule bit 2), XCD-grouped; M = 1024 is already half width
    col_dispatch<double, 8, true, true, S>(h->M, z, h->lda, 4 | 16, st);
}
void hybrid_z(const HybridZ* h, void* data, hipStream_t st) {
    if (h->f32) hybrid_z_t<float>(h, data, st);
    else hybrid_z_t<double>(h, data, st);
}

double fast_bytes(const FastPlan* f) { return f->bytes[0] + f->bytes[1] + f->bytes[2] + f->bytes[3] + f->bytes[4]; }

void fast_schedule(const FastPlan* f, int out[3]) {
    out[0] = f->zsched;
    out[1] = f->ysched;
    out[2] = f->order;
}

// Schedule autotuner: times the candidate schedules of the column passes (tile
// width and XCD grouping of the z pass, XCD grouping of the y passes, traversal
// order of the x passes) on the plan's own shapes and keeps the fastest. Every
// candidate computes the same arithmetic per column, so results do not depend
// on the choice. Choices are remembered per shape for the life of the process.
namespace {
std::mutex g_tune_mu;
std::map<std::string, std::array<int, 3>> g_tune;
}  // namespace

void fast_autotune(FastPlan* f, double* out, const double* in, const View& uin, const View& uout, const double* dgreen,
                   double normfact, hipStream_t st) {
    char key[160];
    std::snprintf(key, sizeof(key), "%d,%d,%d|%d,%d,%d|%d|%d", f->n[0], f->n[1], f->n[2], f->M[0], f->M[1], f->M[2], f->lda,
                  (int)f->fp32);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            f->zsched = it->second[0];
            f->ysched = it->second[1];
            f->order = it->second[2];
            layout_green(f, dgreen, normfact, st);
            return;
        }
    }
    auto run = [&]() {
        if (f->fp32) run_fast<float>(f, f->s0, f->s1, in, out, uin, uout, st);
        else run_fast<double>(f, f->s0, f->s1, in, out, uin, uout, st);
    };
    const bool prof = f->profile;
    f->profile = false;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    std::array<int, 3> pick = {f->zsched, f->ysched, f->order};
    for (int z : {20, 16, 4, 0})
        for (int y : {0, 16})
            for (int o : {12, 4}) {
                f->zsched = z;
                f->ysched = y;
                f->order = o;
                layout_green(f, dgreen, normfact, st);
                run();
                HIP_CHECK(hipEventRecord(e0, st));
                for (int r = 0; r < 3; ++r) run();
                HIP_CHECK(hipEventRecord(e1, st));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    pick = {z, y, o};
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    f->profile = prof;
    f->zsched = pick[0];
    f->ysched = pick[1];
    f->order = pick[2];
    layout_green(f, dgreen, normfact, st);
    HIP_CHECK(hipStreamSynchronize(st));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = pick;
}

// after the stream has completed: accumulate the per-kernel times of the last solve
// kernel kinds of one solve: x_r2c, y_fwd / z / y_inv of each slab (summed), x_c2r
void fast_collect(FastPlan* f) {
    if (!f->profile) return;
    const int ne = (int)f->ev.size();
    for (int i = 0; i + 1 < ne; ++i) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, f->ev[i], f->ev[i + 1]));
        const int kind = i == 0 ? 0 : i == ne - 2 ? 4 : 1 + (i - 1) % 3;
        f->ms[kind] += ms;
    }
    ++f->calls;
}
bool fast_profiling(const FastPlan* f) { return f->profile; }
void fast_profile(FastPlan* f, bool on) {
    f->profile = on;
    for (double& m : f->ms) m = 0;
    f->calls = 0;
}
int fast_profile_get(FastPlan* f, int i, const char** name, double* ms, long* calls, double* bytes) {
    static const char* names[5] = {"x_r2c", "y_fwd", "z_fused", "y_inv", "x_c2r"};
    if (i < 0 || i >= 5) return 0;
    *name = names[i];
    *ms = f->ms[i];
    *calls = f->calls;
    *bytes = f->bytes[i];
    return 5;
}

}  // namespace flups
