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
    hipFunction_t k_wave2h = nullptr;  // two modes per wavefront; null unless the model fits 32 lanes only
    hipFunction_t k_wave_alt = nullptr;   // one mode per wavefront, the other border-factor layout
    hipFunction_t k_tan = nullptr, k_pk_tan = nullptr;   // null when the tangent kernel does not fit LDS
    hipStream_t stream = nullptr;
    int device = 0;
    std::vector<SbzSetTable> sets;
};

#include <cstddef>

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
// wavefronts (four modes each for SBZ_ALGO_WAVE4, two for SBZ_ALGO_WAVE2H).
hipError_t sbz_launch_solve(sbz_model* m, int algo, SbzSolveArgsHost* args, unsigned units);
// Resident wavefronts of a wave variant on the model's device (the persistent grid size).
int sbz_resident_waves(sbz_model* m, int algo);

// A solve plan: instances, parameter tables and outputs resident on the device.
struct sbz_plan {
    sbz_model* m = nullptr;
    int n_inst = 0, n_sets = 0, n_save = 0, saveat_stride = 0, lanes = 0;
    double rtol = 1e-5, atol = 1e-5;
    long long maxiters = 100000;
    // device buffers
    double *d_params = nullptr, *d_knots = nullptr, *d_k = nullptr, *d_t0 = nullptr, *d_saveat = nullptr;
    double *d_out = nullptr, *d_save = nullptr, *d_src = nullptr, *d_pk = nullptr;
    double* d_srcbg = nullptr;   // [n_sets][n_save] background rows for the wave kernels' sources
    int save_flags = 0;
    int algo = SBZ_ALGO_WAVE;
    long long *d_koff = nullptr, *d_stats = nullptr;
    int *d_knotn = nullptr, *d_set = nullptr, *d_order = nullptr, *d_ret = nullptr;
    int* d_xcd_off = nullptr;   // per-XCD segments of d_order (null: one queue)
    // tangent plans: the sets of log θ ± h, 1/(2h), tangents and f(τ0) per instance, d log P
    int *d_setp = nullptr, *d_setm = nullptr;
    double *d_tscale = nullptr, *d_s = nullptr, *d_fend = nullptr, *d_dlogpk = nullptr;
    int n_xcd = 1;
    unsigned int* d_queue = nullptr;   // one counter per queue
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    // sbz_plan_run_graph: the run's launches (output fill, source background, queue reset, solve,
    // P(k)) captured once into a graph and replayed
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool pk_ready = false;   // d_pk holds the P(k) of the last run (graph runs compute it)
    // host copies of the processing order and the wavenumbers (mixed-model launches merge the
    // orders of several plans by estimated cost)
    std::vector<int> h_order;
    std::vector<double> h_k;
};

// The kernel arguments of a plan's solve launch.
void sbz_plan_args(const sbz_plan* p, SbzSolveArgsHost* out);
// Read a whole file (false when missing or empty).
bool sbz_read_file(const char* path, std::vector<char>& out);
// Compile HIP source with hiprtc for the current device's architecture into a code object.
int sbz_compile_hiprtc(const char* src, std::vector<char>& code);
