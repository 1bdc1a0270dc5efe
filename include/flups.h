/*
 * flups.h -- C interface of the MI355X FLUPS solve path (libflups.so).
 *
 * Drop-in for the reference's src/flups.h: every entry point keeps the
 * reference's name, signature and argument meaning. Each declaration cites the
 * reference interface it replaces. The library runs on one GPU and one
 * process: topologies must have nproc = {1,1,1}; the MPI communicator is kept
 * and returned, never used.
 *
 * field/rhs pointers passed to flups_solve may be device pointers (HBM
 * resident, the fast path) or host pointers (copied over PCIe around the
 * solve). Work runs on the stream given by flups_set_stream (default stream
 * otherwise). flups_solve returns with the result complete, except when the
 * caller gave a stream with flups_set_stream and both arrays are device
 * arrays (and no profiler records the solve): the solve is then queued on
 * that stream and the call returns at once, as GPU libraries do on a
 * caller's stream; the result is complete once the stream is synchronised.
 *
 * Errors: entry points keep the reference's void/pointer returns. On a failed
 * check the library records a message (flups_get_last_error), prints it to
 * stderr and returns (NULL for constructors); flups_set_abort_on_error(1)
 * restores the reference's abort-on-failure behaviour.
 */
#ifndef FLUPS_H
#define FLUPS_H

#include <stdbool.h>
#include <stddef.h>

#include "flups_interface.h"

#if defined(FLUPS_BUILDING_LIBRARY)
/* pointer-sized slot: holds an MPICH int handle or an Open MPI pointer handle */
typedef void* MPI_Comm;
#elif defined(FLUPS_WITH_MPI) || defined(MPI_VERSION)
#include <mpi.h>
#else
typedef int MPI_Comm;
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Solver   FLUPS_Solver;    /* src/flups.h:26 */
typedef struct Topology FLUPS_Topology;  /* src/flups.h:27 */
typedef struct Profiler FLUPS_Profiler;  /* src/flups.h:28 */

typedef enum BoundaryType FLUPS_BoundaryType;
typedef enum GreenType    FLUPS_GreenType;
typedef enum SolverType   FLUPS_SolverType;
typedef enum DiffType     FLUPS_DiffType;
typedef enum CenterType   FLUPS_CenterType;

/* ---- memory (src/flups.h:51,61) ---- */
void* flups_malloc(size_t size);   /* zeroed, FLUPS_ALIGNMENT-aligned host memory */
void  flups_free(void* data);
void  flups_info(int argc, char** argv);   /* src/flups.h:69 */

/* index of (i0,i1,i2,lia) given in the axsrc frame, in an array whose fastest
   index is axtrg (src/flups.h:109-119) */
static inline size_t flups_locID(const int axsrc, const int i0, const int i1, const int i2, const int lia, const int axtrg,
                                 const int size[3], const int nf) {
    const int i[3] = {i0, i1, i2};
    const int d0 = (3 + axtrg - axsrc) % 3, d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
    const int a0 = axtrg, a1 = (a0 + 1) % 3, a2 = (a0 + 2) % 3;
    return (size_t)i[d0] * nf + (size_t)size[a0] * nf * ((size_t)i[d1] + (size_t)size[a1] * ((size_t)i[d2] + (size_t)lia * size[a2]));
}

/* symmetrised (wave number) index used with flups_get_spectralInfo (src/flups.h:170-184) */
static inline void flups_symID(const int axsrc, const int i0, const int i1, const int i2, const int istart[3],
                               const double symstart[3], const int axtrg, int is[3]) {
    const int ie[3] = {istart[axsrc] + i0, istart[(axsrc + 1) % 3] + i1, istart[(axsrc + 2) % 3] + i2};
    const int d0 = (3 + axtrg - axsrc) % 3;
    int k;
    for (k = 0; k < 3; ++k) {
        const int ax = (axtrg + k) % 3;
        const int e  = ie[(d0 + k) % 3];
        if (symstart[ax] == 0.0 || e <= symstart[ax]) {
            is[k] = e;
        } else {
            double m = 2.0 * symstart[ax] - e;
            int    v = (int)(m < 0 ? -m : m);
            is[k]    = v > 1 ? v : 1;
        }
    }
}

/* ---- topologies (src/flups.h:210-349) ---- */
FLUPS_Topology* flups_topo_new(const int axis, const int lda, const int nglob[3], const int nproc[3], const bool isComplex,
                               const int axproc[3], const int alignment, MPI_Comm comm);          /* src/flups.h:210 */
/* alias with the name used in some FLUPS documentation; identical to flups_topo_new */
FLUPS_Topology* flups_init_topology(const int axis, const int lda, const int nglob[3], const int nproc[3],
                                    const bool isComplex, const int axproc[3], const int alignment, MPI_Comm comm);
void   flups_topo_free(const FLUPS_Topology* t);                                                  /* src/flups.h:217 */
bool   flups_topo_get_isComplex(const FLUPS_Topology* t);                                         /* src/flups.h:226 */
int    flups_topo_get_lda(const FLUPS_Topology* t);                                               /* src/flups.h:234 */
int    flups_topo_get_axis(const FLUPS_Topology* t);                                              /* src/flups.h:242 */
int    flups_topo_get_nglob(const FLUPS_Topology* t, const int dim);                              /* src/flups.h:250 */
int    flups_topo_get_nloc(const FLUPS_Topology* t, const int dim);                               /* src/flups.h:261 */
int    flups_topo_get_nmem(const FLUPS_Topology* t, const int dim);                               /* src/flups.h:272 */
int    flups_topo_get_nproc(const FLUPS_Topology* t, const int dim);                              /* src/flups.h:280 */
void   flups_topo_get_istartGlob(const FLUPS_Topology* t, int istart[3]);                         /* src/flups.h:287 */
size_t flups_topo_get_locsize(const FLUPS_Topology* t);                                           /* src/flups.h:294 */
size_t flups_topo_get_memsize(const FLUPS_Topology* t);                                           /* src/flups.h:301 */
int    flups_topo_cmpt_rank_fromid(const FLUPS_Topology* t, const int global_id, const int id);    /* src/flups.h:314 */
int    flups_topo_cmpt_start_id_from_rank(const FLUPS_Topology* t, const int rank_id, const int id); /* src/flups.h:321 */
MPI_Comm flups_topo_get_comm(FLUPS_Topology* t);                                                  /* src/flups.h:329 */
void   flups_topo_ranksplit(const FLUPS_Topology* t, const int rank, int rankd[3]);               /* src/flups.h:340 */
int    flups_topo_rankindex(const FLUPS_Topology* topo, const int rankd[3]);                      /* src/flups.h:349 */

/* ---- solver (src/flups.h:369-452) ---- */
FLUPS_Solver* flups_init(FLUPS_Topology* t, FLUPS_BoundaryType* bc[3][2], const double h[3], const double L[3],
                         FLUPS_DiffType orderDiff, const FLUPS_CenterType center_type[3]);        /* src/flups.h:369 */
FLUPS_Solver* flups_init_timed(FLUPS_Topology* t, FLUPS_BoundaryType* bc[3][2], const double h[3], const double L[3],
                               const FLUPS_DiffType orderDiff, const FLUPS_CenterType center_type[3],
                               FLUPS_Profiler* prof);                                             /* src/flups.h:375 */
void flups_cleanup(FLUPS_Solver* s);                                                              /* src/flups.h:384 */
void flups_cleanup_solver(FLUPS_Solver* s);                                                       /* src/flups.h:392 */
void flups_cleanup_all(FLUPS_Solver* s);                                                          /* src/flups.h:401 */
void flups_cleanup_backend(void);                                                                 /* src/flups.h:410 */
void flups_set_greenType(FLUPS_Solver* s, const FLUPS_GreenType type);                            /* src/flups.h:421 */
/* stream = a hipStream_t; the reference's hook is a no-op (src/flups.cpp:147). A non-NULL stream
   makes flups_solve on device arrays asynchronous on it; NULL restores the default stream and the
   synchronous return */
void flups_set_stream(FLUPS_Solver* s, void* stream);                                             /* src/flups.h:429 */
void flups_setup(FLUPS_Solver* s, const bool changeComm);                                         /* src/flups.h:441 */
void flups_solve(FLUPS_Solver* s, double* field, double* rhs, const FLUPS_SolverType type);        /* src/flups.h:452 */
/* No counterpart in the reference: for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops). The solve of flups_solve for DEVICE field / rhs
   pointers, with the solve's stage arrays taken from a static device array of 128 MiB per GPU
   (held from library load to process end) instead of the solver's own: no hipMalloc, hipFree
   or other allocation, no synchronisation. The work is queued on the solver's stream
   (flups_set_stream); the result is complete once that stream is synchronised. A solve whose
   stage arrays exceed 128 MiB (e.g. an all-unbounded FP64 problem above about 128^3) fails
   with an error. All calls on one GPU share the array: calls that overlap on one GPU race. */
void flups_solve_noalloc(FLUPS_Solver* s, double* field, double* rhs, const FLUPS_SolverType type);

/* ---- advanced (src/flups.h:467-578) ---- */
size_t flups_get_allocSize(FLUPS_Solver* s);                                                      /* src/flups.h:473 */
void   flups_get_spectralInfo(FLUPS_Solver* s, double kfact[3], double koffset[3], double symstart[3]); /* src/flups.h:485 */
void   flups_set_alpha(FLUPS_Solver* s, const double alpha);                                      /* src/flups.h:494 */
void   flups_pencilDirs(const FLUPS_BoundaryType* bc[3][2], int dirs[3]);                          /* src/flups.h:570 */

/* ---- advanced interface (src/flups.h:467-578): the steps of flups_solve one by one.
   Buffers are device memory of flups_get_allocSize doubles (flups_get_innerBuffer returns
   one); flups_do_copy accepts host or device user arrays. The physical layout is the user
   topology; the spectral layout is described by flups_get_innerTopo_spectral (complex
   interleaved, dimension 0 fastest, dense); mode indices follow flups_get_spectralInfo and
   flups_symID as in the reference. ---- */
double*         flups_get_innerBuffer(FLUPS_Solver* s);                                          /* src/flups.h:509 */
FLUPS_Topology* flups_get_innerTopo_physical(FLUPS_Solver* s);                                   /* src/flups.h:517 */
FLUPS_Topology* flups_get_innerTopo_spectral(FLUPS_Solver* s);                                   /* src/flups.h:525 */
void            flups_skip_firstSwitchtopo(FLUPS_Solver* s);                                     /* src/flups.h:534 */
void            flups_do_copy(FLUPS_Solver* s, const FLUPS_Topology* topo, double* data, const int sign); /* src/flups.h:544 */
void            flups_do_FFT(FLUPS_Solver* s, double* data, const int sign);                    /* src/flups.h:552 */
void            flups_do_mult(FLUPS_Solver* s, double* data, const FLUPS_SolverType type);      /* src/flups.h:560 */
void            flups_switchtopo_info(FLUPS_Solver* s);                                          /* src/flups.h:578 */

/* ---- profiler handles (src/flups.h:592-611). A solver made by flups_init_timed records into
   its handle: "init", "setup" and "setup/green" (host wall-clock), and for every flups_solve
   "solve" (host wall-clock and device time) with its stages timed by HIP events on the solve
   stream: "solve/copy rhs" and "solve/copy field" (host arrays), "solve/fftw forward",
   "solve/domagic", "solve/fftw backward" (general path), "solve/fftw + domagic (last direction
   fused)", or "solve/fused pipeline (fftw + domagic)" with one entry per kernel. The
   pencil switches are part of the transform passes. flups_profiler_disp prints the table to
   stdout. ---- */
FLUPS_Profiler* flups_profiler_new(void);                                                         /* src/flups.h:592 */
FLUPS_Profiler* flups_profiler_new_n(const char name[]);                                          /* src/flups.h:599 */
void            flups_profiler_free(FLUPS_Profiler* p);                                           /* src/flups.h:605 */
void            flups_profiler_disp(FLUPS_Profiler* p);                                           /* src/flups.h:611 */
/* addition: timer idx of the handle (name, host seconds, device milliseconds, calls); returns the
   number of timers, -1 for a NULL handle */
int             flups_profiler_get(FLUPS_Profiler* p, int idx, char* name, int name_len, double* host_s,
                                   double* device_ms, long* calls);

/* ---- debug output (src/flups.h:634-638) ---- */
/* prints data (host or device memory) laid out by topo, one pencil of the fastest direction per line */
void flups_print_data(const FLUPS_Topology* topo, double* data);                                 /* src/flups.h:638 */
/* HDF5 is not a dependency of this library: records an error and writes nothing */
void flups_hdf5_dump(const FLUPS_Topology* topo, const char filename[], const double* data);     /* src/flups.h:634 */

/* ---- additions of this library (no counterpart in the reference) ---- */
const char* flups_get_last_error(void);        /* message of the last failed check, "" if none */
void        flups_clear_last_error(void);
void        flups_set_abort_on_error(int on);  /* 1: abort() on a failed check, as the reference */
/* 0 = FP64 throughout (default), 1 = FP32: the fused unbounded pipeline runs its transforms
   and multiply in FP32 arithmetic with FP64-computed Green's function and twiddles, rounded
   once; every other boundary set runs the general path with FP32 stage arrays and an FP32
   copy of the Green's function (half the HBM bytes) and FP64 arithmetic in the kernels. The
   caller's arrays stay double. Call before flups_setup. The advanced interface
   (flups_do_copy / do_FFT / do_mult) works on FP64 solvers only. */
void        flups_set_precision(FLUPS_Solver* s, int fp32);
/* 0 = general pencil path (FP64), 1 = fused pipeline in FP64, 2 = fused in FP32,
   3 = general passes with the last direction fused with the multiply (FP64),
   4 = general path with FP32 stage arrays, 5 = the same with the fused last direction */
int         flups_get_path(FLUPS_Solver* s);
/* kernel of each pass of the general path after flups_setup (component 0; the forward
   passes in order, then the backward ones): 0 generic pencil kernel, 1 contiguous
   real-to-real, 2 paired real columns (r2c / c2r), 3 contiguous r2c / c2r, 4 strided
   real-to-real, 5 strided c2c, 6 contiguous c2c, -1 no transform, -2 fused with the multiply (the last
   direction of flups_get_path 3). Fills at most max entries, returns the
   number of passes (0 when the fused pipeline runs everything, -1 for a NULL solver) */
int         flups_get_pass_kernels(FLUPS_Solver* s, int* kinds, int max);
/* 1: flups_setup times the fused pipeline's candidate schedules on this problem and keeps
   the fastest (remembered per shape for the process; FLUPS_AUTOTUNE=1 turns it on for all) */
void        flups_set_autotune(FLUPS_Solver* s, int on);
/* minimum HBM bytes of one solve of this type (each pass reads its input and writes its
   output once; the multiply reads the field and the Green's function): the roofline numerator */
double      flups_get_solve_bytes(FLUPS_Solver* s, const FLUPS_SolverType type);
/* schedule of the fused pipeline: z pass, y passes, x traversal order (-1 on the general path) */
void        flups_get_schedule(FLUPS_Solver* s, int sched[3]);
/* nanoseconds of device time of the last flups_solve (HIP events on the solver's stream) */
double      flups_get_last_solve_ns(FLUPS_Solver* s);
/* per-kernel timing of the fused pipeline: enable (resets the counters), then read entry idx:
   name, accumulated milliseconds, number of solves and algorithmic HBM bytes per launch.
   Returns the number of entries (0 if the solver runs the general path). */
void        flups_profile_enable(FLUPS_Solver* s, int on);
int         flups_profile_get(FLUPS_Solver* s, int idx, char* name, int name_len, double* total_ms, long* calls,
                              double* bytes_per_call);

/* device memory through the library's HIP runtime (zero-copy HBM buffers for flups_solve) */
void*       flups_device_malloc(size_t bytes);
void        flups_device_free(void* p);
int         flups_memcpy(void* dst, const void* src, size_t bytes);   /* any direction; 1 on success */
int         flups_pointer_is_device(const void* p);
/* HIP streams of the library's runtime (for flups_set_stream from callers without their own):
   create (default priority, non-blocking), destroy, wait for (1 on success) */
void*       flups_stream_create(void);
void        flups_stream_destroy(void* stream);
int         flups_stream_synchronize(void* stream);

/* ---- serving and batching (no counterpart in the reference: its solver serves one
   problem per process). A request describes a whole problem; caller arrays are dense
   doubles shaped (lda, n2, n1, n0), device or host memory. ---- */
typedef struct {
    int    nglob[3];
    int    center;        /* FLUPS_CenterType, all directions */
    int    bc[3][2];      /* FLUPS_BoundaryType per direction and side, same for every component */
    int    lda;           /* 1 (scalar) or 3 (vector) */
    double h[3], L[3];
    int    green;         /* FLUPS_GreenType */
    double alpha;         /* regularisation factor; <= 0 keeps the default */
    int    fp32;          /* 1: FP32 transforms where the fused pipeline runs */
    int    type;          /* FLUPS_SolverType: STD or ROT */
    int    order_diff;    /* FLUPS_DiffType (ROT) */
} FLUPS_ServeRequest;

typedef struct FLUPS_ServerS FLUPS_Server;
typedef struct FLUPS_BatchS  FLUPS_Batch;

/* serving loop: `workers` threads; solvers cached by request description and evicted
   least-recently-used within `hbm_budget_bytes`; at most `solvers_per_key` solvers per
   description. Requests arriving while others run are admitted as workers and memory
   free up; their solves overlap on the device. */
FLUPS_Server* flups_server_new(size_t hbm_budget_bytes, int workers, int solvers_per_key);
void          flups_server_free(FLUPS_Server* s);
long          flups_server_submit(FLUPS_Server* s, const FLUPS_ServeRequest* req, const double* rhs, double* field);
int           flups_server_wait(FLUPS_Server* s, long ticket, double* latency_ms);   /* 1 done, 0 failed */
void          flups_server_drain(FLUPS_Server* s);
/* out: completed, failed, p50 ms, p99 ms, mean ms, solves/s (first submit to last completion),
   cache hits, cache misses, evictions, peak device bytes, device bytes now, cached solvers */
void          flups_server_stats(FLUPS_Server* s, double out[12]);
void          flups_server_reset_stats(FLUPS_Server* s);
/* device bytes the server reserves to admit a request whose solver is not cached */
size_t        flups_server_request_bytes(const FLUPS_ServeRequest* req);

/* batch: n independent problems (device buffers), set up once, solved by one hipGraph
   launch. flups_batch_run: launches of the same kernel, step and grid in several problems
   run as one grouped launch (blocks mapped to problems through an offset table; up to 32
   launches per group); every problem's result is bit for bit its result solved alone.
   flups_batch_run_branches: the same graph with one branch per problem and its own launches. */
FLUPS_Batch*  flups_batch_new(int n, const FLUPS_ServeRequest* reqs, const double* const* rhs, double* const* field);
void          flups_batch_free(FLUPS_Batch* b);
int           flups_batch_run(FLUPS_Batch* b);              /* 1 on success; waits for completion */
int           flups_batch_run_branches(FLUPS_Batch* b);     /* 1 on success; waits for completion */
int           flups_batch_run_sequential(FLUPS_Batch* b);   /* same problems one by one, no graph */
double        flups_batch_last_ms(FLUPS_Batch* b);          /* device time of the last run / run_branches */
int           flups_batch_launches(FLUPS_Batch* b, int grouped);   /* kernel launches of one run (grouped 1) or run_branches (0) */
size_t        flups_batch_device_bytes(FLUPS_Batch* b);

#ifdef __cplusplus
}
#endif

#endif
