/*
 * A C caller of libflups built against include/flups.h with the real <mpi.h>, in the style of the
 * reference's samples/solve_advanced_C/main_solve_advanced.c:67-214: a topology on
 * MPI_COMM_WORLD, a timed solver, flups_solve on host arrays, then the advanced interface
 * (do_copy, do_FFT, do_mult, a Gaussian filter indexed with flups_symID and
 * flups_get_spectralInfo, the backward do_FFT and do_copy), checked against the analytic
 * solution of a periodic sine. The inner buffer is device memory here, so the filter reads and
 * writes it through flups_memcpy. Exit status 0 when both errors are below 1e-10.
 */
#include <math.h>
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "flups.h"

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    const int    nglob[3]  = {32, 24, 16};
    const int    nproc[3]  = {1, 1, 1};
    const double L[3]      = {1.0, 1.5, 0.5};
    const double h[3]      = {L[0] / nglob[0], L[1] / nglob[1], L[2] / nglob[2]};
    const double c_2pi     = 2.0 * 3.141592653589793;
    const double kw[3]     = {c_2pi / L[0], c_2pi / L[1], c_2pi / L[2]};
    const double ksq_mode  = kw[0] * kw[0] + kw[1] * kw[1] + kw[2] * kw[2];
    const double eps       = 1e-3;   /* filter exp(-eps |k|^2) */
    FLUPS_CenterType center[3] = {CELL_CENTER, CELL_CENTER, CELL_CENTER};
    FLUPS_BoundaryType per[1] = {PER};
    FLUPS_BoundaryType* bc[3][2] = {{per, per}, {per, per}, {per, per}};
    int rank = 0, failed = 0;

    FLUPS_Topology* topo = flups_topo_new(0, 1, nglob, nproc, false, NULL, FLUPS_ALIGNMENT, MPI_COMM_WORLD);
    if (!topo) {
        fprintf(stderr, "flups_topo_new: %s\n", flups_get_last_error());
        return 2;
    }
    MPI_Comm comm = flups_topo_get_comm(topo);
    MPI_Comm_rank(comm, &rank);

    FLUPS_Profiler* prof = flups_profiler_new_n("c driver");
    FLUPS_Solver*   s    = flups_init_timed(topo, bc, h, L, NOD, center, prof);
    if (!s) {
        fprintf(stderr, "flups_init_timed: %s\n", flups_get_last_error());
        return 2;
    }
    flups_set_greenType(s, CHAT_2);
    flups_setup(s, true);

    int nmem[3];
    for (int d = 0; d < 3; ++d) nmem[d] = flups_topo_get_nmem(topo, d);
    const size_t memsize = flups_topo_get_memsize(topo);
    double* rhs   = (double*)flups_malloc(sizeof(double) * memsize);
    double* field = (double*)flups_malloc(sizeof(double) * memsize);
    double* exact = (double*)flups_malloc(sizeof(double) * memsize);
    for (int i2 = 0; i2 < nglob[2]; ++i2)
        for (int i1 = 0; i1 < nglob[1]; ++i1)
            for (int i0 = 0; i0 < nglob[0]; ++i0) {
                const double x = (i0 + 0.5) * h[0], y = (i1 + 0.5) * h[1], z = (i2 + 0.5) * h[2];
                const size_t id = flups_locID(0, i0, i1, i2, 0, 0, nmem, 1);
                rhs[id]   = sin(kw[0] * x) * sin(kw[1] * y) * sin(kw[2] * z);
                exact[id] = -rhs[id] / ksq_mode;
            }

    /* 1. the plain solve on host arrays */
    flups_solve(s, field, rhs, STD);
    double emax = 0.0, smax = 0.0;
    for (size_t i = 0; i < memsize; ++i) {
        emax = fmax(emax, fabs(field[i] - exact[i]));
        smax = fmax(smax, fabs(exact[i]));
    }
    const double err_solve = emax / smax;

    /* 2. the advanced interface with a spectral Gaussian filter */
    double* inner = flups_get_innerBuffer(s);
    const FLUPS_Topology* spec = flups_get_innerTopo_spectral(s);
    flups_do_copy(s, topo, rhs, FLUPS_FORWARD);
    flups_do_FFT(s, inner, FLUPS_FORWARD);
    flups_do_mult(s, inner, STD);
    double kfact[3], koffset[3], symstart[3];
    flups_get_spectralInfo(s, kfact, koffset, symstart);
    const int    ax0 = flups_topo_get_axis(spec), ax1 = (ax0 + 1) % 3, ax2 = (ax0 + 2) % 3;
    const int    nf  = flups_topo_get_isComplex(spec) ? 2 : 1;
    int          nmem_s[3], istart_s[3];
    for (int d = 0; d < 3; ++d) nmem_s[d] = flups_topo_get_nmem(spec, d);
    flups_topo_get_istartGlob(spec, istart_s);
    const size_t nspec = flups_topo_get_memsize(spec);
    double*      host  = (double*)flups_malloc(sizeof(double) * nspec);
    flups_memcpy(host, inner, sizeof(double) * nspec);
    for (int i2 = 0; i2 < flups_topo_get_nloc(spec, ax2); ++i2)
        for (int i1 = 0; i1 < flups_topo_get_nloc(spec, ax1); ++i1) {
            const size_t id = flups_locID(ax0, 0, i1, i2, 0, ax0, nmem_s, nf);
            for (int i0 = 0; i0 < flups_topo_get_nloc(spec, ax0); ++i0) {
                int is[3];
                flups_symID(ax0, i0, i1, i2, istart_s, symstart, 0, is);
                const double k0 = (is[ax0] + koffset[ax0]) * kfact[ax0];
                const double k1 = (is[ax1] + koffset[ax1]) * kfact[ax1];
                const double k2 = (is[ax2] + koffset[ax2]) * kfact[ax2];
                const double f  = exp(-eps * (k0 * k0 + k1 * k1 + k2 * k2));
                for (int c = 0; c < nf; ++c) host[id + (size_t)i0 * nf + c] *= f;
            }
        }
    flups_memcpy(inner, host, sizeof(double) * nspec);
    flups_do_FFT(s, inner, FLUPS_BACKWARD);
    memset(field, 0, sizeof(double) * memsize);
    flups_do_copy(s, topo, field, FLUPS_BACKWARD);
    emax = 0.0;
    const double damp = exp(-eps * ksq_mode);
    for (size_t i = 0; i < memsize; ++i) emax = fmax(emax, fabs(field[i] - damp * exact[i]));
    const double err_filter = emax / (damp * smax);

    if (rank == 0) {
        printf("c driver: %d x %d x %d periodic, CHAT_2: solve rel Linf %.3e, filtered advanced steps rel Linf %.3e\n",
               nglob[0], nglob[1], nglob[2], err_solve, err_filter);
        flups_profiler_disp(prof);
    }
    failed = !(err_solve < 1e-10 && err_filter < 1e-10) || flups_get_last_error()[0] != 0;
    if (flups_get_last_error()[0]) fprintf(stderr, "last error: %s\n", flups_get_last_error());

    flups_free(host);
    flups_free(rhs);
    flups_free(field);
    flups_free(exact);
    flups_cleanup_solver(s);
    flups_profiler_free(prof);
    flups_topo_free(topo);
    MPI_Finalize();
    return failed;
}
