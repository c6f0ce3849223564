// Host-side background + thermodynamics solve (the input stage of the perturbation hot path).
//
// Mirrors the reference's `solvebg` (src/solve.jl:427-435) with the a = 1 termination callback
// (:158-203): integrate the five background unknowns (a, X_H+, X_He+, ΔT, _κ) with Rodas5P at
// rtol = atol = 1e-7 from τ_bg0 = 1e-6, stop when a = 1, record τ0, κ0 = _κ(τ0) and τrec
// (argmax of the visibility function over the stored steps), and return the step times with
// u and u' = f(τ, u) as the knots of the cubic Hermite background spline (src/utils.jl:118-127).
#include <hip/hip_runtime.h>

#include <cmath>

#include "sbz_bgsolve.h"
#include "sbz_internal.h"

// Background right-hand side at one state (host helper for tests and the Python mirror).
extern "C" int sbz_background_rhs(const double* params, const double* u, double* du) {
    if (!params || !u || !du) return SBZ_ERR_ARG;
    sbz_bg_rhs<double>(params, u, du);
    return SBZ_OK;
}

extern "C" int sbz_background_solve(double* params, double rtol, double atol, int max_knots,
                                    double* t_out, double* u_out, double* du_out, int* n_knots,
                                    int* n_rejected) {
    sbz_clear_error();
    if (!params || !t_out || !u_out || !du_out || !n_knots || max_knots < 5) {
        sbz_set_error("sbz_background_solve: null argument or max_knots < 5");
        return SBZ_ERR_ARG;
    }
    int nrej = 0;
    double tf = 0.0;
    const int st = sbz_bg_solve(params, rtol, atol, max_knots, t_out, u_out, du_out, n_knots, &nrej, &tf);
    if (n_rejected) *n_rejected = nrej;
    if (st != SBZ_BG_OK) {
        sbz_set_error("background: %s at tau = %g", sbz_bg_status_text(st), tf);
        return SBZ_ERR_SOLVE;
    }
    return SBZ_OK;
}
