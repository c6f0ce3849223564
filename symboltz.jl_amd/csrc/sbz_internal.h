// Library-internal helpers: status codes and the thread-local last-error message.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/sbz.h"

void sbz_set_error(const char* fmt, ...);
void sbz_clear_error();

// ---- perturbation solver state shared by the runtime's source files ----------------------------

struct SbzSetTable {
    std::vector<double> knots;  // t[nk], u[nk][5], du[nk][5]
    int nk = 0;
    double params[SBZ_NPAR];
    bool valid = false;
};

struct sbz_model {
    int n = 0, nnz_lu = 0;
    hipModule_t module = nullptr;
    hipFunction_t k_pk = nullptr, k_wave = nullptr, k_wave2 = nullptr, k_pk_modes = nullptr;
    hipFunction_t k_srcbg = nullptr, k_fill = nullptr, k_qreset = nullptr;
    hipFunction_t k_wave4 = nullptr;   // null when the model's border or chains exceed 16 lanes
    hipFunction_t k_tan = nullptr, k_pk_tan = nullptr;   // null when the tangent kernel does not fit LDS
    hipStream_t stream = nullptr;
    int device = 0;
    std::vector<SbzSetTable> sets;
};

// Mirrors struct SbzSolveArgs of sbz_kernel.hip (passed by value as the kernel argument).
struct SbzSolveArgsHost {
    const double* params;
    const double* knots;
    const long long* knot_off;
    const int* knot_n;
    const int* inst_set;
    const double* inst_k;
    const double* inst_t0;
    const int* order;
    int n_inst;
    double rtol, atol;
    long long maxiters;
    const double* saveat;
    int n_save;
    int saveat_stride;
    double* out_u;
    double* out_save;
    double* out_src;
    int* retcode;
    long long* stats;
    unsigned int* queue;
    int S;
    const double* srcbg;
    const int* xcd_off;
    int n_xcd;
    const int* inst_setp;
    const int* inst_setm;
    const double* inst_tscale;
    double* out_s;
    double* out_fend;
};

// Kernel variant of tangent plans (not a caller-selectable algo: sbz_plan_create_tangent sets it).
constexpr int SBZ_ALGO_TANGENT_INTERNAL = 100;

// Work queues of a launch: one per XCD (instances grouped by set mod SBZ_N_XCD) when a plan has
// several parameter sets, else one.
constexpr int SBZ_N_XCD = 8;

// Launch the persistent solve kernel of `algo` with `args` on the model's stream: grid `units`
// wavefronts (four modes each for SBZ_ALGO_WAVE4).
hipError_t sbz_launch_solve(sbz_model* m, int algo, SbzSolveArgsHost* args, unsigned units);
// Resident wavefronts of a wave variant on the model's device (the persistent grid size).
int sbz_resident_waves(sbz_model* m, int algo);
