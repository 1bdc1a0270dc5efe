// C entry points of libflups (declared in include/flups.h).
// Each forwards to the Topo / SolverImpl objects; failed checks are caught here,
// recorded for flups_get_last_error and, in abort mode, abort the process as the
// reference's FLUPS_CHECK does (src/defines.hpp:301-304).
#define FLUPS_BUILDING_LIBRARY 1
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/flups.h"
#include "solver.hpp"

struct Topology {
    flups::Topo t;
};
struct Solver {
    flups::SolverImpl* impl;
    Topology* phys = nullptr;   // inner topologies handed out by flups_get_innerTopo_*: owned by
    Topology* spec = nullptr;   // the solver, as in the reference (freed with it)
};
struct Profiler {
    flups::Timers t;
};

namespace {
thread_local std::string g_last_error;
int g_abort = 0;

void report(const char* where, const std::string& msg) {
    g_last_error = std::string(where) + ": " + msg;
    std::fprintf(stderr, "[flups] %s\n", g_last_error.c_str());
    if (g_abort) std::abort();
}
}  // namespace

namespace flups {
void report_error(const char* where, const std::string& msg) { report(where, msg); }
}  // namespace flups

namespace {
template <typename F>
auto guard(const char* where, F&& f) -> decltype(f()) {
    try {
        return f();
    } catch (const flups::Error& e) {
        report(where, e.msg);
    } catch (const std::exception& e) {
        report(where, e.what());
    }
    return decltype(f())();
}
}  // namespace

extern "C" {

const char* flups_get_last_error(void) { return g_last_error.c_str(); }
void flups_clear_last_error(void) { g_last_error.clear(); }
void flups_set_abort_on_error(int on) { g_abort = on; }

void* flups_malloc(size_t size) {
    void* p = nullptr;
    if (posix_memalign(&p, 64, size ? size : 1) != 0) return nullptr;
    std::memset(p, 0, size);
    return p;
}
void flups_free(void* data) { std::free(data); }

void flups_info(int argc, char** argv) {
    (void)argc;
    std::fprintf(stderr, "[flups] MI355X FLUPS solve path, single GPU, program %s\n", argv && argv[0] ? argv[0] : "?");
}

// ---- topologies ----
FLUPS_Topology* flups_topo_new(const int axis, const int lda, const int nglob[3], const int nproc[3], const bool isComplex,
                               const int axproc[3], const int alignment, MPI_Comm comm) {
    return guard("flups_topo_new", [&]() -> FLUPS_Topology* {
        FLUPS_REQUIRE(axis >= 0 && axis < 3, "axis must be 0, 1 or 2");
        FLUPS_REQUIRE(lda >= 1, "lda must be >= 1");
        FLUPS_REQUIRE(alignment > 0 && alignment % 8 == 0, "alignment must be a positive multiple of 8 bytes");
        FLUPS_REQUIRE(nproc[0] * nproc[1] * nproc[2] == 1, "the total number of procs (%d) has to be 1 on one GPU",
                      nproc[0] * nproc[1] * nproc[2]);
        auto* T = new Topology;
        T->t.axis = axis;
        T->t.lda = lda;
        T->t.nf = isComplex ? 2 : 1;
        T->t.alignment = alignment;
        T->t.comm = comm;
        for (int d = 0; d < 3; ++d) {
            FLUPS_REQUIRE(nglob[d] >= 1, "nglob must be >= 1");
            T->t.nglob[d] = nglob[d];
            T->t.nproc[d] = nproc[d];
            T->t.axproc[d] = axproc ? axproc[d] : d;
        }
        T->t.cmpt_sizes();
        return T;
    });
}
FLUPS_Topology* flups_init_topology(const int axis, const int lda, const int nglob[3], const int nproc[3], const bool isComplex,
                                    const int axproc[3], const int alignment, MPI_Comm comm) {
    return flups_topo_new(axis, lda, nglob, nproc, isComplex, axproc, alignment, comm);
}
void flups_topo_free(const FLUPS_Topology* t) { delete t; }
bool flups_topo_get_isComplex(const FLUPS_Topology* t) { return t->t.nf == 2; }
int flups_topo_get_lda(const FLUPS_Topology* t) { return t->t.lda; }
int flups_topo_get_axis(const FLUPS_Topology* t) { return t->t.axis; }
int flups_topo_get_nglob(const FLUPS_Topology* t, const int dim) { return t->t.nglob[dim]; }
int flups_topo_get_nloc(const FLUPS_Topology* t, const int dim) { return t->t.nloc[dim]; }
int flups_topo_get_nmem(const FLUPS_Topology* t, const int dim) { return t->t.nmem[dim]; }
int flups_topo_get_nproc(const FLUPS_Topology* t, const int dim) { return t->t.nproc[dim]; }
void flups_topo_get_istartGlob(const FLUPS_Topology* t, int istart[3]) {
    (void)t;
    istart[0] = istart[1] = istart[2] = 0;
}
size_t flups_topo_get_locsize(const FLUPS_Topology* t) { return t->t.locsize(); }
size_t flups_topo_get_memsize(const FLUPS_Topology* t) { return t->t.memsize(); }
int flups_topo_cmpt_rank_fromid(const FLUPS_Topology* t, const int global_id, const int id) {
    (void)t; (void)global_id; (void)id;
    return 0;
}
int flups_topo_cmpt_start_id_from_rank(const FLUPS_Topology* t, const int rank_id, const int id) {
    return rank_id * t->t.nglob[id];
}
MPI_Comm flups_topo_get_comm(FLUPS_Topology* t) { return (MPI_Comm)t->t.comm; }
void flups_topo_ranksplit(const FLUPS_Topology* t, const int rank, int rankd[3]) {
    (void)t; (void)rank;
    rankd[0] = rankd[1] = rankd[2] = 0;
}
int flups_topo_rankindex(const FLUPS_Topology* topo, const int rankd[3]) {
    (void)topo; (void)rankd;
    return 0;
}

// ---- solver ----
FLUPS_Solver* flups_init_timed(FLUPS_Topology* t, FLUPS_BoundaryType* bc[3][2], const double h[3], const double L[3],
                               const FLUPS_DiffType orderDiff, const FLUPS_CenterType center_type[3], FLUPS_Profiler* prof) {
    const auto t0 = std::chrono::steady_clock::now();
    return guard("flups_init", [&]() -> FLUPS_Solver* {
        FLUPS_REQUIRE(t != nullptr, "topology is NULL");
        FLUPS_REQUIRE(t->t.nf == 1, "the physical topology must be real");
        const int lda = t->t.lda;
        std::vector<int> flat(6 * lda);
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s) {
                FLUPS_REQUIRE(bc[d][s] != nullptr, "boundary condition array [%d][%d] is NULL", d, s);
                for (int c = 0; c < lda; ++c) flat[(d * 2 + s) * lda + c] = (int)bc[d][s][c];
            }
        const int center[3] = {(int)center_type[0], (int)center_type[1], (int)center_type[2]};
        auto* S = new Solver;
        S->impl = new flups::SolverImpl(&t->t, flat.data(), h, L, (int)orderDiff, center);
        if (prof) {
            S->impl->set_profiler(&prof->t);
            prof->t.add("init", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0.0);
        }
        return S;
    });
}
FLUPS_Solver* flups_init(FLUPS_Topology* t, FLUPS_BoundaryType* bc[3][2], const double h[3], const double L[3],
                         FLUPS_DiffType orderDiff, const FLUPS_CenterType center_type[3]) {
    return flups_init_timed(t, bc, h, L, orderDiff, center_type, nullptr);
}
void flups_cleanup_solver(FLUPS_Solver* s) {
    if (!s) return;
    guard("flups_cleanup", [&]() { delete s->impl; return 0; });
    delete s->phys;
    delete s->spec;
    delete s;
}
void flups_cleanup(FLUPS_Solver* s) { flups_cleanup_solver(s); }
void flups_cleanup_all(FLUPS_Solver* s) { flups_cleanup_solver(s); }
void flups_cleanup_backend(void) {}

void flups_set_greenType(FLUPS_Solver* s, const FLUPS_GreenType type) {
    guard("flups_set_greenType", [&]() {
        FLUPS_REQUIRE(s != nullptr, "solver is NULL");
        FLUPS_REQUIRE((int)type >= 0 && (int)type <= 14, "unknown Green's function type %d", (int)type);
        s->impl->set_green((int)type);
        return 0;
    });
}
void flups_set_stream(FLUPS_Solver* s, void* stream) {
    // a caller-given stream opts into asynchronous solves on device arrays; NULL restores the
    // default stream and the synchronous return
    if (s) s->impl->set_stream((hipStream_t)stream, stream != nullptr);
}
void flups_set_alpha(FLUPS_Solver* s, const double alpha) {
    if (s) s->impl->set_alpha(alpha);
}
void flups_set_precision(FLUPS_Solver* s, int fp32) {
    guard("flups_set_precision", [&]() {
        FLUPS_REQUIRE(s, "NULL solver");
        s->impl->set_fp32(fp32 != 0);
        return 0;
    });
}
void flups_setup(FLUPS_Solver* s, const bool changeComm) {
    (void)changeComm;
    guard("flups_setup", [&]() {
        FLUPS_REQUIRE(s != nullptr, "solver is NULL");
        s->impl->setup();
        return 0;
    });
}
void flups_solve(FLUPS_Solver* s, double* field, double* rhs, const FLUPS_SolverType type) {
    guard("flups_solve", [&]() {
        FLUPS_REQUIRE(s != nullptr, "solver is NULL");
        s->impl->solve(field, rhs, (int)type);
        return 0;
    });
}
void flups_solve_noalloc(FLUPS_Solver* s, double* field, double* rhs, const FLUPS_SolverType type) {
    guard("flups_solve_noalloc", [&]() {
        FLUPS_REQUIRE(s != nullptr, "solver is NULL");
        s->impl->solve_noalloc(field, rhs, (int)type);
        return 0;
    });
}
size_t flups_get_allocSize(FLUPS_Solver* s) { return s ? s->impl->alloc_size() : 0; }
void flups_get_spectralInfo(FLUPS_Solver* s, double kfact[3], double koffset[3], double symstart[3]) {
    if (s) s->impl->spectral_info(kfact, koffset, symstart);
}
int flups_get_path(FLUPS_Solver* s) { return s ? s->impl->path() : -1; }
int flups_get_pass_kernels(FLUPS_Solver* s, int* kinds, int max) {
    if (!s) return -1;
    const std::vector<int>& k = s->impl->pass_kernels();
    for (int i = 0; i < (int)k.size() && i < max; ++i) kinds[i] = k[i];
    return (int)k.size();
}
// ---- advanced interface (src/flups.h:467-578) ----
double* flups_get_innerBuffer(FLUPS_Solver* s) {
    return guard("flups_get_innerBuffer", [&]() -> double* {
        FLUPS_REQUIRE(s, "NULL solver");
        return s->impl->inner_buffer();
    });
}
// the returned topologies belong to the solver (src/flups.h:517-525): callers do not free them
FLUPS_Topology* flups_get_innerTopo_physical(FLUPS_Solver* s) {
    return guard("flups_get_innerTopo_physical", [&]() -> FLUPS_Topology* {
        FLUPS_REQUIRE(s, "NULL solver");
        if (!s->phys) s->phys = new Topology;
        s->phys->t = s->impl->inner_topo_physical();
        return s->phys;
    });
}
FLUPS_Topology* flups_get_innerTopo_spectral(FLUPS_Solver* s) {
    return guard("flups_get_innerTopo_spectral", [&]() -> FLUPS_Topology* {
        FLUPS_REQUIRE(s, "NULL solver");
        if (!s->spec) s->spec = new Topology;
        s->spec->t = s->impl->inner_topo_spectral();
        return s->spec;
    });
}
void flups_skip_firstSwitchtopo(FLUPS_Solver* s) {
    (void)s;   // the physical layout is always the user topology: there is no first switch to skip
}
void flups_do_copy(FLUPS_Solver* s, const FLUPS_Topology* topo, double* data, const int sign) {
    guard("flups_do_copy", [&]() {
        FLUPS_REQUIRE(s && topo, "NULL solver or topology");
        s->impl->do_copy(&topo->t, data, sign);
        return 0;
    });
}
void flups_do_FFT(FLUPS_Solver* s, double* data, const int sign) {
    guard("flups_do_FFT", [&]() {
        FLUPS_REQUIRE(s, "NULL solver");
        s->impl->do_fft(data, sign);
        return 0;
    });
}
void flups_do_mult(FLUPS_Solver* s, double* data, const FLUPS_SolverType type) {
    guard("flups_do_mult", [&]() {
        FLUPS_REQUIRE(s, "NULL solver");
        s->impl->do_mult(data, (int)type);
        return 0;
    });
}
void flups_switchtopo_info(FLUPS_Solver* s) {
    if (!s) return;
    int sched[3];
    s->impl->schedule(sched);
    std::fprintf(stderr, "[flups] one GPU: pencil switches are fused into the transform passes (path %d, schedule %d %d %d)\n",
                 s->impl->path(), sched[0], sched[1], sched[2]);
}

void flups_set_autotune(FLUPS_Solver* s, int on) {
    guard("flups_set_autotune", [&]() {
        FLUPS_REQUIRE(s, "NULL solver");
        s->impl->set_autotune(on != 0);
        return 0;
    });
}
double flups_get_solve_bytes(FLUPS_Solver* s, const FLUPS_SolverType type) {
    return guard("flups_get_solve_bytes", [&]() -> double {
        FLUPS_REQUIRE(s, "NULL solver");
        return s->impl->algorithmic_bytes((int)type);
    });
}
void flups_get_schedule(FLUPS_Solver* s, int sched[3]) {
    if (s && sched) s->impl->schedule(sched);
}
void flups_profile_enable(FLUPS_Solver* s, int on) {
    if (s) s->impl->profile(on != 0);
}
int flups_profile_get(FLUPS_Solver* s, int idx, char* name, int name_len, double* total_ms, long* calls,
                      double* bytes_per_call) {
    if (!s) return 0;
    const char* nm = "";
    const int n = s->impl->profile_get(idx, &nm, total_ms, calls, bytes_per_call);
    if (n && name && name_len > 0) {
        std::strncpy(name, nm, name_len - 1);
        name[name_len - 1] = 0;
    }
    return n;
}
double flups_get_last_solve_ns(FLUPS_Solver* s) { return s ? s->impl->last_solve_ns() : 0.0; }

void flups_pencilDirs(const FLUPS_BoundaryType* bc[3][2], int dirs[3]) {
    // directions in execution order: stable sort by plan type (src/flups.cpp:200-210)
    int type[3];
    for (int d = 0; d < 3; ++d) type[d] = (int)bc[d][0][0] + (int)bc[d][1][0];
    int order[3] = {0, 1, 2};
    for (int i = 1; i < 3; ++i)
        for (int j = i; j > 0 && type[order[j]] < type[order[j - 1]]; --j) std::swap(order[j], order[j - 1]);
    for (int d = 0; d < 3; ++d) dirs[d] = order[d];
}

// ---- device memory helpers (HBM-resident buffers for callers without a HIP allocator) ----
void* flups_device_malloc(size_t bytes) {
    return guard("flups_device_malloc", [&]() -> void* {
        void* p = nullptr;
        HIP_CHECK(hipMalloc(&p, bytes ? bytes : 1));
        return p;
    });
}
void flups_device_free(void* p) {
    if (p) (void)hipFree(p);
}
int flups_memcpy(void* dst, const void* src, size_t bytes) {
    return guard("flups_memcpy", [&]() {
        HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
        return 1;
    });
}
void* flups_stream_create(void) {
    return guard("flups_stream_create", [&]() -> void* {
        hipStream_t st = nullptr;
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        return (void*)st;
    });
}
void flups_stream_destroy(void* stream) {
    if (stream) (void)hipStreamDestroy((hipStream_t)stream);
}
int flups_stream_synchronize(void* stream) {
    return guard("flups_stream_synchronize", [&]() -> int {
        HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
        return 1;
    });
}
int flups_pointer_is_device(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return a.type == hipMemoryTypeDevice ? 1 : 0;
}

// ---- debug output (src/flups.h:634-638) ----
// data in the topology's layout (host or device memory): one line per pencil of the fastest
// direction, components one after the other (src/defines.hpp:204-240)
void flups_print_data(const FLUPS_Topology* topo, double* data) {
    guard("flups_print_data", [&]() {
        FLUPS_REQUIRE(topo && data, "NULL topology or data");
        const flups::Topo& t = topo->t;
        std::vector<double> host(t.memsize());
        HIP_CHECK(hipMemcpy(host.data(), data, host.size() * sizeof(double), hipMemcpyDefault));
        const int a0 = t.axis, a1 = (a0 + 1) % 3, a2 = (a0 + 2) % 3;
        std::printf("ax0 = %d -- lda = %d -- nf = %d - nmem = %d %d %d -- end = %d %d %d \n", a0, t.lda, t.nf, t.nmem[0],
                    t.nmem[1], t.nmem[2], t.nloc[0], t.nloc[1], t.nloc[2]);
        for (int c = 0; c < t.lda; ++c) {
            std::printf("lia == %d \n", c);
            for (int i2 = 0; i2 < t.nloc[a2]; ++i2)
                for (int i1 = 0; i1 < t.nloc[a1]; ++i1) {
                    if (i1 == 0) std::printf("\n");
                    const double* row = host.data() + c * t.memdim() +
                                        (size_t)t.nf * t.nmem[a0] * ((size_t)i1 + (size_t)t.nmem[a1] * i2);
                    for (int i0 = 0; i0 < t.nloc[a0]; ++i0) {
                        if (t.nf == 1) std::printf("%e \t ", row[i0]);
                        else std::printf("(%e, %e) \t", row[2 * i0], row[2 * i0 + 1]);
                    }
                    std::printf("\n");
                }
        }
        std::fflush(stdout);
        return 0;
    });
}
// HDF5 / XDMF dumps are not part of this library (no HDF5 dependency): the entry point exists
// so that callers link, and it reports that nothing was written
void flups_hdf5_dump(const FLUPS_Topology* topo, const char filename[], const double* data) {
    (void)topo;
    (void)data;
    guard("flups_hdf5_dump", [&]() {
        flups::fail("HDF5 output is not built into this library; %s was not written", filename ? filename : "(null)");
        return 0;
    });
}

// ---- profiler handles (src/flups.cpp:219-237): timers of the solvers made by flups_init_timed ----
FLUPS_Profiler* flups_profiler_new(void) {
    auto* p = new Profiler;
    p->t.name = "default";
    return p;
}
FLUPS_Profiler* flups_profiler_new_n(const char name[]) {
    auto* p = new Profiler;
    p->t.name = name ? name : "default";
    return p;
}
void flups_profiler_free(FLUPS_Profiler* p) { delete p; }
void flups_profiler_disp(FLUPS_Profiler* p) {
    if (p) p->t.display(stdout);
}
int flups_profiler_get(FLUPS_Profiler* p, int idx, char* name, int name_len, double* host_s, double* device_ms,
                       long* calls) {
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->t.mu);
    const int n = (int)p->t.e.size();
    if (idx >= 0 && idx < n) {
        const auto& x = p->t.e[idx];
        if (name && name_len > 0) {
            std::strncpy(name, x.name.c_str(), name_len - 1);
            name[name_len - 1] = 0;
        }
        if (host_s) *host_s = x.host_s;
        if (device_ms) *device_ms = x.dev_ms;
        if (calls) *calls = x.calls;
    }
    return n;
}

}  // extern "C"
