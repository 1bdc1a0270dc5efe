// Solver and Topology objects behind the FLUPS C interface (one GPU, one rank).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "core.hpp"
#include "green.hpp"

namespace flups {

// user-visible topology (reference: src/Topology.hpp:27-430, single rank)
struct Topo {
    int axis = 0, lda = 1, nf = 1, alignment = 8;
    int nglob[3] = {1, 1, 1}, nloc[3] = {1, 1, 1}, nmem[3] = {1, 1, 1}, nproc[3] = {1, 1, 1}, axproc[3] = {0, 1, 2};
    void* comm = nullptr;
    size_t memdim() const { return (size_t)nmem[0] * nmem[1] * nmem[2] * nf; }
    size_t memsize() const { return memdim() * lda; }
    size_t locsize() const { return (size_t)nloc[0] * nloc[1] * nloc[2] * nf; }
    void cmpt_sizes();
};

// Named timers of a FLUPS_Profiler handle (the reference times its solve through H3LPR's
// profiler: src/flups.cpp:110-114, 219-237, src/Solver.cpp:1218, 1286, 1411-1419, 1475).
// Host wall-clock seconds and device milliseconds (HIP events on the solver's stream) are
// accumulated per name; names nest with '/' ("solve/fftw fwd x").
struct Timers {
    struct Entry {
        std::string name;
        double host_s = 0, dev_ms = 0;
        long calls = 0;
    };
    std::string name;
    std::vector<Entry> e;
    std::mutex mu;
    void add(const std::string& n, double host_s, double dev_ms, long calls = 1);
    void display(FILE* f);
};

struct FastPlan;   // fused unbounded pipeline, fast.hip
struct HybridZ;    // general path with its last direction fused with the multiply, fast.hip

class SolverImpl {
   public:
    SolverImpl(const Topo* t, const int* bc /*[(d*2+side)*lda + c]*/, const double h[3], const double L[3], int odiff,
               const int center[3]);
    ~SolverImpl();
    void set_green(int t) { gtype_ = t; }
    void set_alpha(double a) { alpha_ = a; }
    // async: flups_solve on device arrays returns once its work is queued on the stream (a
    // caller-given stream); otherwise it returns with the result complete
    void set_stream(hipStream_t s, bool async = false) {
        stream_ = s;
        async_ = async;
    }
    void set_fp32(bool f) {
        FLUPS_REQUIRE(!setup_done_, "flups_set_precision must be called before flups_setup");
        fp32_ = f;
    }
    void setup();
    void solve(double* field, const double* rhs, int type);
    void prepare(int type);                                          // allocate what a solve of this type needs
    void enqueue(double* field, const double* rhs, int type);        // device pointers, no sync, no allocation
    void solve_noalloc(double* field, const double* rhs, int type);  // enqueue on the static scratch's stage arrays
    size_t device_bytes() const;                                     // device memory held now
    size_t predicted_device_bytes(int type) const;                   // bound for setup + solve, before setup
    bool is_setup() const { return setup_done_; }
    void set_autotune(bool on) {
        FLUPS_REQUIRE(!setup_done_, "flups_set_autotune must be called before flups_setup");
        autotune_ = on;
    }
    void schedule(int out[3]) const;   // fused pipeline schedule (z, y, order), -1 on the general path
    double algorithmic_bytes(int type) const;   // minimum HBM bytes of one solve
    // advanced interface: the solve's steps on device buffers
    double* inner_buffer();
    void do_copy(const Topo* t, double* data, int sign);
    void do_fft(double* data, int sign);
    void do_mult(double* data, int type);
    const Topo& inner_topo_physical() const { return phys_; }
    Topo inner_topo_spectral() const;
    size_t alloc_size() const;
    void spectral_info(double kf[3], double ko[3], double ss[3]) const;
    // 0 general FP64, 1 fused FP64, 2 fused FP32, 3 general with the last direction fused with the
    // multiply (FP64), 4 general with FP32 stage arrays, 5 the same with the fused last direction
    int path() const { return fast_ ? (fp32_ ? 2 : 1) : hybrid_ ? (gf32_ ? 5 : 3) : (gf32_ ? 4 : 0); }
    // kernel of each general-path pass (component 0): forward passes in order, then the backward ones
    const std::vector<int>& pass_kernels() const { return pass_kernels_; }
    double last_solve_ns();   // waits for the last solve's end when it was queued asynchronously
    int ndim() const { return ndim_; }
    // per-kernel timing of the fused pipeline (HIP events on the solve stream)
    void profile(bool on);
    int profile_get(int i, const char** name, double* ms, long* calls, double* bytes) const;

    // FLUPS_Profiler handle given to flups_init_timed (not owned): setup and every flups_solve
    // record their stages into it
    void set_profiler(Timers* t) { prof_ = t; }
    Timers* profiler() const { return prof_; }

    // building blocks (also used by the fused pipeline)
    hipStream_t stream() const { return stream_; }

   private:
    void compute_green();
    const double* twiddles(int M);
    void plan_fft(PassDesc& P, int M);
    PassDesc pass_desc(const DirPlan& p, int comp, bool forward, const View& src, const View& dst, const Shape& tin,
                       const Shape& tout, int gshift, int sshift);
    void run_pass(PassDesc P);
    void fill_green_args(GreenArgs& A, const double hfact[3], const double kfact[3], const double koffset[3],
                         const double symstart[3], double length, int nunb);
    View user_view(double* p) const;
    void solve_generic(double* field, const double* rhs, int type);
    void forward_passes(const View& uin, void* spec);
    void multiply(void* spec, int type);
    void backward_passes(const void* spec, const View& uout, bool diff);
    void prepare_advanced();
    void warm_tables();
    void launch_graph(double* field, const double* rhs, int type);
    void check_type(int type) const;
    void pmark(const char* label);          // profiler: event on the stream closing the stage `label`
    void pflush(const std::string& prefix);  // after a sync: stage times into the profiler

    Topo phys_;
    int lda_ = 1, ndim_ = 3, odiff_ = NOD, gtype_ = CHAT_2;
    bool node_ = false, fp32_ = false, setup_done_ = false, autotune_ = false;
    double h_[3], L_[3], alpha_ = 2.0, normfact_ = 1, volfact_ = 1;
    DirPlan fwd_[3], bwd_[3], bwdd_[3], gpl_[3];
    Shape hat_[3], stage_[3], gpre_[3], gpost_[3];
    hipStream_t stream_ = nullptr;
    std::map<int, double*> tw_;
    std::map<std::pair<int, int>, double*> r2r_tab_;   // (kind, n) -> pre / post twiddles of the row r2r kernel
    const double* r2r_table(int kind, int n);
    void* dstage_[3] = {nullptr, nullptr, nullptr};   // general path stage arrays: double, or float (gf32_)
    double* dgreen_ = nullptr;
    float* dgreen32_ = nullptr;   // FP32 copy of the Green's function (gf32_)
    bool gf32_ = false;           // FP32 request on the general path: FP32 stage arrays, FP64 arithmetic
    int ses_ = 8;                 // bytes of one real in the stage arrays
    double* dstaging_ = nullptr;   // rhs/field copies for host pointers
    double* dinner_ = nullptr;     // inner buffer of the advanced interface
    std::vector<int> pass_kernels_;
    bool frozen_ = false;   // a solve that may not allocate is running: missing tables are errors
    bool use_graph_ = false, profiling_ = false, warmed_ = false;   // FLUPS_GRAPH=1: replay solves as graphs
    hipStream_t cap_ = nullptr;    // capture stream of the solve graph
    hipGraphExec_t gexec_ = nullptr;
    double* gfield_ = nullptr;
    const double* grhs_ = nullptr;
    int gtype_key_ = -1;
    size_t staging_n_ = 0;
    size_t lds_budget_ = 65536;
    size_t generic_lds_cap_ = 32768;   // LDS of one general-path block (FLUPS_GENERIC_LDS)
    FastPlan* fast_ = nullptr;
    HybridZ* hybrid_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    Timers* prof_ = nullptr;
    std::vector<hipEvent_t> pev_;       // profiler stage events, reused
    std::vector<std::string> plab_;
    int pn_ = 0;
    double last_ns_ = 0;
    bool async_ = false, last_pending_ = false;
    std::string kdir_;
    friend struct FastPlan;
    friend FastPlan* make_fast_plan(SolverImpl&);
    friend void build_fast_plan(SolverImpl&, FastPlan*, size_t, int);
    friend HybridZ* make_hybrid_z(SolverImpl&);
};

std::string kernel_dir();

}  // namespace flups
