/* C ABI of libsbz_hip.so: the MI355X perturbation-stage solver behind SymBoltz's
 * CosmologyProblem / solve(prob, ks) API.
 *
 * Conventions: every entry point returns an int status (SBZ_OK on success); the message of the
 * last failure on the calling thread is available from sbz_last_error().  The caller allocates
 * and owns every host buffer; the library owns device memory and the opaque handles.  Per-mode
 * integration failures never fail a call: they are reported in the per-mode retcode array, as the
 * reference warns instead of throwing (src/solve.jl:407-419,557-564).  Calls on one model handle
 * must not run concurrently (one HIP stream per handle). */
#ifndef SBZ_H
#define SBZ_H

#include "sbz_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SBZ_OK 0
#define SBZ_ERR_ARG 1      /* invalid argument (the analogue of check_solve_args, solve.jl:356-361) */
#define SBZ_ERR_HIP 2      /* HIP runtime error */
#define SBZ_ERR_COMPILE 3  /* hiprtc compile or code-object load failure */
#define SBZ_ERR_SOLVE 4    /* a host-side solve failed (background) */
#define SBZ_ERR_NODEV 5    /* no GPU visible */

/* per-mode return codes, mirroring SciMLBase.ReturnCode */
#define SBZ_RET_SUCCESS 0
#define SBZ_RET_MAXITERS 1
#define SBZ_RET_DTLESSTHANMIN 2
#define SBZ_RET_UNSTABLE 3
#define SBZ_RET_INITIALFAILURE 4
#define SBZ_RET_NOTRUN 5

/* Thread-local message of the last failure on this thread ("" if none). */
const char* sbz_last_error(void);

/* Library version (major*10000 + minor*100 + patch). */
int sbz_version(void);

/* Number of visible GPUs (0 if none); never fails on a machine without a GPU. */
int sbz_device_count(int* count);

/* Background + thermodynamics solve for one parameter set.
 * Replaces: solvebg (src/solve.jl:427-435) with the a = 1 callback (:158-203) and the spline
 * knot extraction spline(sol) (src/utils.jl:118-127).
 * params: SBZ_NPAR doubles laid out as in sbz_params.h; on return SBZ_P_TAU0, SBZ_P_KAPPA0 and
 * SBZ_P_TAUREC are filled in (the reference writes them into the problem's tunables).
 * Outputs: knot times t[n], unknowns u[n][5] = (a, X_H+, X_He+, ΔT, _κ) and u'[n][5]. */
int sbz_background_solve(double* params, double rtol, double atol, int max_knots, double* t,
                         double* u, double* du, int* n_knots, int* n_rejected);

#ifdef __cplusplus
}
#endif
#endif
