/* itlin.h of the liblinalg_hip drop-in: the option and result blocks of
 * LinA's gmres (LinA/itlin.h:58-84, LinA/gmres.c:308-310), declared here
 * with the same members in the same order so that a caller compiled against
 * LinA (FEM2D/FEM.c:531-541, 737-744) passes them unchanged.  Only the
 * members gmres reads or writes have a meaning here. */
#ifndef LINALG_HIP_ITLIN_H
#define LINALG_HIP_ITLIN_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { None = 0, Minimum = 1, Verbose = 2, Debug = 3 } PRINT_LEVEL;
typedef enum { False = 0, True = 1 } LOGICAL;
/* CheckOnRestart: the residual is tested once per restart cycle;
 * CheckEachIter: the iterate is formed and its residual tested every step */
typedef enum { CheckOnRestart = 0, CheckEachIter = 1 } TERM_CHECK;
typedef enum { Absolute = 0, Relative = 1 } CONV_CHECK;

/* y = A x, and a preconditioner z = M^-1 x: (n, x, y) */
typedef void MATVEC(int, double*, double*);
typedef void PRECON(int, double*, double*);

struct ITLIN_OPT {
    double tol, rho;          /* tol: eta = |r| / |r_0| to reach (clamped to [1e-15, 0.1]) */
    int i_max, maxiter;       /* restart length (<= 0 or > n: min(10, n)); iteration cap (<= 0: 100) */
    TERM_CHECK termcheck;
    CONV_CHECK convcheck;     /* not used by gmres */
    LOGICAL rescale;          /* not used by gmres */
    PRINT_LEVEL errorlevel, monitorlevel, datalevel;
    FILE *errorfile, *monitorfile, *datafile, *iterfile, *resfile, *miscfile;
    double* scale;            /* not used by gmres */
};

struct ITLIN_INFO {
    double precision, normdx, residuum; /* precision: the final eta */
    int iter, rcode, subcode, nomatvec, noprecon, noprecl, noprecr;
    /* rcode: 0 converged, 2 iteration cap reached, 1 singular Hessenberg
     * matrix, 20 n <= 0, 21 tol <= 0, -99 no matvec, -999 out of memory */
};

/* GMRES(i_max) with right preconditioner preconr and left preconditioner
 * preconl (NULL: none) for A x = b.  y enters as the initial iterate and
 * leaves as the solution of the right-preconditioned system A M_r^-1 y = b
 * (the caller recovers x = M_r^-1 y), as in LinA (gmres.c:308-501). */
void gmres(int n, double* y, MATVEC* matvec, PRECON* preconr, PRECON* preconl, double* b, struct ITLIN_OPT* opt,
           struct ITLIN_INFO* info);

#ifdef __cplusplus
}
#endif
#endif
