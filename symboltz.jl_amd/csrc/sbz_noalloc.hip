// sbz_solve_batch_noalloc: the one-shot batch solve with no device allocation.
//
// Every device buffer of the solve (background tables, instances, outputs, work counter) lives in
// one file-scope device array, `scratch`, of SBZ_NOALLOC_SCRATCH_BYTES per GPU, held by the
// process until it ends.  Host code never uses the array's name as an address: on each device, at
// the first call with that device current, hipGetSymbolAddress gives the device address of that
// device's copy, which is kept per device and used for every copy and kernel argument.  The
// solve kernels come from the model's code object and take these addresses as plain pointers, as
// they take the plan's buffers in sbz_solve_batch.  All calls on one GPU share its array: calls
// that overlap on one GPU race (the array keeps the last call's contents).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "sbz_internal.h"

static __device__ __align__(256) unsigned char scratch[SBZ_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};   // device address of `scratch` on each device (null: not resolved yet)

// The device address of this device's scratch array; the caller has made `dev` current.
int scratch_of(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) {
        sbz_set_error("sbz_solve_batch_noalloc: device %d beyond the supported %d", dev, kMaxDevices);
        return SBZ_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            sbz_set_error("sbz_solve_batch_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return SBZ_ERR_HIP;
        }
        g_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_ptr[dev]);
    return SBZ_OK;
}

// Bump layout inside the scratch array: 256-byte aligned regions.
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    }
};
}  // namespace

#define NCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            sbz_set_error("sbz_solve_batch_noalloc: %s failed: %s", #x, hipGetErrorString(e_));  \
            return SBZ_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

extern "C" int sbz_solve_batch_noalloc(sbz_model* m, const int* inst_set, const double* inst_k,
                                       const double* inst_t0, int n_inst, const sbz_opts* opts, const double* saveat,
                                       int n_save, double* out_u, double* out_save, int* retcode, long long* stats) {
    sbz_clear_error();
    if (!m || !inst_set || !inst_k || !inst_t0 || !opts || n_inst < 0 || n_save < 0 || (n_save > 0 && !saveat)) {
        sbz_set_error("sbz_solve_batch_noalloc: invalid argument");
        return SBZ_ERR_ARG;
    }
    if (!(opts->rtol > 0) || !(opts->atol > 0)) {
        sbz_set_error("sbz_solve_batch_noalloc: tolerances must be positive");
        return SBZ_ERR_ARG;
    }
    const int algo = opts->algo ? opts->algo : SBZ_ALGO_WAVE;
    if (algo != SBZ_ALGO_WAVE && algo != SBZ_ALGO_WAVE2) {
        sbz_set_error("sbz_solve_batch_noalloc: only the one-mode wavefront kernels (SBZ_ALGO_WAVE, SBZ_ALGO_WAVE2) "
                      "run without allocation");
        return SBZ_ERR_ARG;
    }
    if (opts->save & ~SBZ_SAVE_STATES) {
        sbz_set_error("sbz_solve_batch_noalloc: saves states only (no sources, no per-set saveat)");
        return SBZ_ERR_ARG;
    }
    const int n_sets = (int)m->sets.size();
    for (int i = 0; i < n_inst; ++i)
        if (inst_set[i] < 0 || inst_set[i] >= n_sets || !m->sets[inst_set[i]].valid) {
            sbz_set_error("sbz_solve_batch_noalloc: instance %d refers to parameter set %d without a background table",
                          i, inst_set[i]);
            return SBZ_ERR_ARG;
        }
    for (int i = 1; i < n_save; ++i)
        if (!(saveat[i] >= saveat[i - 1])) {
            sbz_set_error("sbz_solve_batch_noalloc: saveat must be ascending (entry %d)", i);
            return SBZ_ERR_ARG;
        }
    if (n_inst == 0) return SBZ_OK;
    // layout of this call inside the scratch array
    const int n = m->n;
    Carve c;
    const size_t o_params = c.take(sizeof(double) * SBZ_NPAR * (size_t)n_sets);
    size_t knot_doubles = 0;
    for (int s = 0; s < n_sets; ++s) knot_doubles += m->sets[s].knots.size();
    const size_t o_knots = c.take(sizeof(double) * std::max<size_t>(knot_doubles, 1));
    const size_t o_koff = c.take(sizeof(long long) * n_sets);
    const size_t o_kn = c.take(sizeof(int) * n_sets);
    const size_t o_set = c.take(sizeof(int) * (size_t)n_inst);
    const size_t o_k = c.take(sizeof(double) * (size_t)n_inst);
    const size_t o_t0 = c.take(sizeof(double) * (size_t)n_inst);
    const size_t o_saveat = c.take(sizeof(double) * std::max(n_save, 1));
    const size_t o_out = c.take(sizeof(double) * (size_t)n_inst * n);
    const size_t o_save = (n_save > 0) ? c.take(sizeof(double) * (size_t)n_inst * n_save * n) : 0;
    const size_t o_ret = c.take(sizeof(int) * (size_t)n_inst);
    const size_t o_stats = c.take(sizeof(long long) * 6 * (size_t)n_inst);
    const size_t o_queue = c.take(sizeof(unsigned int));
    if (c.off > (size_t)SBZ_NOALLOC_SCRATCH_BYTES) {
        sbz_set_error("sbz_solve_batch_noalloc: the batch needs %zu bytes of device scratch, more than the %d "
                      "available (fewer instances, saves or parameter sets per call, or sbz_solve_batch)",
                      c.off, SBZ_NOALLOC_SCRATCH_BYTES);
        return SBZ_ERR_ARG;
    }
    NCHK(hipSetDevice(m->device));
    unsigned char* base = nullptr;
    const int st = scratch_of(m->device, &base);
    if (st != SBZ_OK) return st;
    hipStream_t s0 = m->stream;
    // inputs: each set's table is copied straight from the model's host copy (no staging buffer)
    size_t kpos = 0;
    for (int s = 0; s < n_sets; ++s) {
        const SbzSetTable& T = m->sets[s];
        const long long off = (long long)kpos;
        const int nk = T.nk;
        NCHK(hipMemcpyAsync(base + o_params + sizeof(double) * SBZ_NPAR * (size_t)s, T.params,
                            sizeof(double) * SBZ_NPAR, hipMemcpyHostToDevice, s0));
        if (!T.knots.empty())
            NCHK(hipMemcpyAsync(base + o_knots + sizeof(double) * kpos, T.knots.data(), sizeof(double) * T.knots.size(),
                                hipMemcpyHostToDevice, s0));
        NCHK(hipMemcpyAsync(base + o_koff + sizeof(long long) * s, &off, sizeof(long long), hipMemcpyHostToDevice, s0));
        NCHK(hipMemcpyAsync(base + o_kn + sizeof(int) * s, &nk, sizeof(int), hipMemcpyHostToDevice, s0));
        NCHK(hipStreamSynchronize(s0));   // `off` and `nk` are locals of this iteration
        kpos += T.knots.size();
    }
    NCHK(hipMemcpyAsync(base + o_set, inst_set, sizeof(int) * (size_t)n_inst, hipMemcpyHostToDevice, s0));
    NCHK(hipMemcpyAsync(base + o_k, inst_k, sizeof(double) * (size_t)n_inst, hipMemcpyHostToDevice, s0));
    NCHK(hipMemcpyAsync(base + o_t0, inst_t0, sizeof(double) * (size_t)n_inst, hipMemcpyHostToDevice, s0));
    if (n_save > 0)
        NCHK(hipMemcpyAsync(base + o_saveat, saveat, sizeof(double) * n_save, hipMemcpyHostToDevice, s0));
    // saves a failed mode never reaches read NaN, as in the plan path; the queue starts at zero
    {
        double* p = reinterpret_cast<double*>(base + o_save);
        long long cnt = (n_save > 0) ? (long long)n_inst * n_save * n : 0;
        double v = __builtin_nan("");
        void* fargs[] = {&p, &cnt, &v};
        if (cnt > 0)
            NCHK(hipModuleLaunchKernel(m->k_fill, (unsigned)std::min<long long>((cnt + 255) / 256, 4096), 1, 1, 256, 1,
                                       1, 0, s0, fargs, nullptr));
        unsigned int* q = reinterpret_cast<unsigned int*>(base + o_queue);
        int nq = 1;
        void* qargs[] = {&q, &nq};
        NCHK(hipModuleLaunchKernel(m->k_qreset, 1, 1, 1, 64, 1, 1, 0, s0, qargs, nullptr));
    }
    SbzSolveArgsHost a;
    a.params = reinterpret_cast<const double*>(base + o_params);
    a.knots = reinterpret_cast<const double*>(base + o_knots);
    a.knot_off = reinterpret_cast<const long long*>(base + o_koff);
    a.knot_n = reinterpret_cast<const int*>(base + o_kn);
    a.inst_set = reinterpret_cast<const int*>(base + o_set);
    a.inst_k = reinterpret_cast<const double*>(base + o_k);
    a.inst_t0 = reinterpret_cast<const double*>(base + o_t0);
    a.order = nullptr;   // instances run in the caller's order
    a.n_inst = n_inst;
    a.rtol = opts->rtol;
    a.atol = opts->atol;
    a.maxiters = opts->maxiters > 0 ? opts->maxiters : 100000;
    a.saveat = (n_save > 0) ? reinterpret_cast<const double*>(base + o_saveat) : nullptr;
    a.n_save = n_save;
    a.saveat_stride = 0;
    a.out_u = reinterpret_cast<double*>(base + o_out);
    a.out_save = (n_save > 0) ? reinterpret_cast<double*>(base + o_save) : nullptr;
    a.out_src = nullptr;
    a.retcode = reinterpret_cast<int*>(base + o_ret);
    a.stats = reinterpret_cast<long long*>(base + o_stats);
    a.queue = reinterpret_cast<unsigned int*>(base + o_queue);
    const int waves = std::max(1, std::min(n_inst, opts->lanes > 0 ? opts->lanes : sbz_resident_waves(m, algo)));
    a.S = waves;
    a.srcbg = nullptr;
    a.xcd_off = nullptr;   // one queue
    a.n_xcd = 1;
    NCHK(sbz_launch_solve(m, algo, &a, (unsigned)waves));
    if (out_u) NCHK(hipMemcpyAsync(out_u, a.out_u, sizeof(double) * (size_t)n_inst * n, hipMemcpyDeviceToHost, s0));
    if (out_save && n_save > 0)
        NCHK(hipMemcpyAsync(out_save, a.out_save, sizeof(double) * (size_t)n_inst * n_save * n, hipMemcpyDeviceToHost,
                            s0));
    if (retcode) NCHK(hipMemcpyAsync(retcode, a.retcode, sizeof(int) * (size_t)n_inst, hipMemcpyDeviceToHost, s0));
    if (stats)
        NCHK(hipMemcpyAsync(stats, a.stats, sizeof(long long) * 6 * (size_t)n_inst, hipMemcpyDeviceToHost, s0));
    NCHK(hipStreamSynchronize(s0));
    return SBZ_OK;
}
