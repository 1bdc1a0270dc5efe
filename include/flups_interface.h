/*
 * flups_interface.h -- enumerations and constants of the FLUPS C interface.
 *
 * Same names and values as the reference's src/flups_interface.h:14-110 so
 * that code written against FLUPS compiles and links unchanged against the
 * MI355X library.
 */
#ifndef FLUPS_INTERFACE_H_
#define FLUPS_INTERFACE_H_

/* transform signs (src/flups_interface.h:14-28) */
#define FLUPS_FORWARD -1
#define FLUPS_BACKWARD 1
#define FLUPS_BACKWARD_DIFF 2

/* memory alignment in bytes used by flups_malloc (src/flups_interface.h:34) */
#define FLUPS_ALIGNMENT 8

/* boundary conditions (src/flups_interface.h:47-53) */
enum BoundaryType {
    EVEN = 0, /* zero flux */
    ODD  = 1, /* zero value */
    PER  = 3, /* periodic */
    UNB  = 4, /* unbounded */
    NONE = 9  /* direction not used (2-D problems) */
};

/* Green's function kernels (src/flups_interface.h:59-75) */
enum GreenType {
    CHAT_2  = 0,
    LGF_2   = 1,
    HEJ_2   = 2,
    HEJ_4   = 3,
    HEJ_6   = 4,
    HEJ_8   = 5,
    HEJ_10  = 6,
    HEJ_0   = 7,
    LGF_4   = 8,
    LGF_6   = 9,
    LGF_8   = 10,
    MEHR_4L = 11,
    MEHR_6L = 12,
    MEHR_4F = 13,
    MEHR_6F = 14
};

/* solver kinds (src/flups_interface.h:86-89) */
enum SolverType {
    STD = 0, /* laplacian(phi) = rhs */
    ROT = 1  /* laplacian(phi) = curl(rhs), lda = 3 */
};

/* spectral derivative used by ROT (src/flups_interface.h:95-101) */
enum DiffType {
    NOD = 0,
    SPE = 1,
    FD2 = 2,
    FD4 = 4,
    FD6 = 6
};

/* data location (src/flups_interface.h:107-110) */
enum CenterType {
    NODE_CENTER = 0,
    CELL_CENTER = 1
};

#endif
