// Host emulation of the wavefront-per-mode kernel (sbz_wave_kernel) — TEST INFRASTRUCTURE ONLY.
//
// The emitted kernel source is compiled unchanged for the CPU.  One 64-lane workgroup is
// emulated with 64 cooperative contexts (ucontext) on one OS thread: every __syncthreads, lane
// exchange (readlane, shfl_xor) and vote (__all) is a real barrier at which all 64 lanes must
// meet, and the emulator aborts if lanes arrive at different barrier sites or some lanes exit
// while others wait (non-uniform control flow around a barrier).  This checks the kernel's
// arithmetic, its cross-lane communication and, built with AddressSanitizer, its indexing of the
// global buffers without a GPU.
#include <ucontext.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define __HIPCC_RTC__ 1  // the kernel then skips <hip/hip_runtime.h>
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __restrict__ __restrict
#define __launch_bounds__(...)
#define __constant__ static
#define __shared__ static

struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
static EmuDim3 threadIdx, blockIdx, blockDim{64, 1, 1}, gridDim{1, 1, 1};

namespace emu {
constexpr int W = 64;
constexpr size_t STACK = 1 << 20;
static ucontext_t sched_ctx, lane_ctx[W];
static std::vector<char> stacks;
static int cur = 0;
static bool done[W];
static int site[W];   // source line of the barrier each lane waits at
static double xd[2][W];   // exchange buffers, alternated per exchange (one barrier suffices)
static long long xi[2][W];
static int gen[W];
static void (*body)(void*) = nullptr;
static void* body_arg = nullptr;
static char error[256];

static void barrier_at(int line) {
    site[cur] = line;
    swapcontext(&lane_ctx[cur], &sched_ctx);
}
static void syncthreads(int line) { barrier_at(line); }
static int readlane_i(int v, int l, int s) {
    const int g = gen[cur] ^= 1;
    xi[g][cur] = v;
    barrier_at(s);
    return (int)xi[g][l];
}
static double shfl_xor_d(double v, int off, int s) {
    const int g = gen[cur] ^= 1;
    xd[g][cur] = v;
    barrier_at(s);
    return xd[g][cur ^ off];
}
// DPP moves used by the kernel: quad_perm (ctrl < 0x100), row_mirror (0x140), row_half_mirror (0x141)
static int mov_dpp(int v, int ctrl, int s) {
    const int g = gen[cur] ^= 1;
    xi[g][cur] = v;
    barrier_at(s);
    int src;
    if (ctrl < 0x100) src = (cur & ~3) | ((ctrl >> (2 * (cur & 3))) & 3);
    else if (ctrl == 0x140) src = (cur & ~15) | (15 - (cur & 15));
    else if (ctrl == 0x141) src = (cur & ~7) | (7 - (cur & 7));
    else if (ctrl >= 0x150 && ctrl <= 0x15F) src = (cur & ~15) | (ctrl - 0x150);   // row_newbcast
    else {
        std::fprintf(stderr, "wave_emu: unsupported DPP control 0x%x\n", ctrl);
        std::abort();
    }
    return (int)xi[g][src];
}
static int all(int p, int s) {
    const int g = gen[cur] ^= 1;
    xi[g][cur] = p ? 1 : 0;
    barrier_at(s);
    int r = 1;
    for (int l = 0; l < W; ++l) r &= (int)xi[g][l];
    return r;
}
static int any(int p, int s) {
    const int g = gen[cur] ^= 1;
    xi[g][cur] = p ? 1 : 0;
    barrier_at(s);
    int r = 0;
    for (int l = 0; l < W; ++l) r |= (int)xi[g][l];
    return r;
}
static unsigned long long ballot(int p, int s) {
    const int g = gen[cur] ^= 1;
    xi[g][cur] = p ? 1 : 0;
    barrier_at(s);
    unsigned long long m = 0;
    for (int l = 0; l < W; ++l)
        if (xi[g][l]) m |= 1ull << l;
    return m;
}
static void lane_main() {
    body(body_arg);
    done[cur] = true;
    swapcontext(&lane_ctx[cur], &sched_ctx);
}
// run body(arg) on 64 lanes; returns 0 or -1 (message in `error`)
static int run(void (*fn)(void*), void* arg) {
    body = fn;
    body_arg = arg;
    stacks.assign(STACK * W, 0);
    for (int l = 0; l < W; ++l) {
        done[l] = false;
        gen[l] = 0;
        getcontext(&lane_ctx[l]);
        lane_ctx[l].uc_stack.ss_sp = stacks.data() + STACK * l;
        lane_ctx[l].uc_stack.ss_size = STACK;
        lane_ctx[l].uc_link = nullptr;
        makecontext(&lane_ctx[l], lane_main, 0);
    }
    for (;;) {
        for (int l = 0; l < W; ++l) {
            cur = l;
            threadIdx.x = l;
            site[l] = -1;
            swapcontext(&sched_ctx, &lane_ctx[l]);
        }
        int nd = 0;
        for (int l = 0; l < W; ++l) nd += done[l];
        if (nd == W) return 0;
        if (nd != 0) {
            snprintf(error, sizeof error, "%d lanes exited while %d wait at a barrier", nd, W - nd);
            return -1;
        }
        for (int l = 1; l < W; ++l)
            if (site[l] != site[0]) {
                snprintf(error, sizeof error, "lanes 0 and %d wait at different barriers (source lines %d, %d)",
                         l, site[0], site[l]);
                return -1;
            }
    }
}
}  // namespace emu

#define __syncthreads() emu::syncthreads(__LINE__)
#define __builtin_amdgcn_readlane(v, l) emu::readlane_i((v), (l), __LINE__)
#define __shfl_xor(v, off, w) emu::shfl_xor_d((v), (off), __LINE__)
#define __all(p) emu::all((p), __LINE__)
#define __any(p) emu::any((p), __LINE__)
#define __ballot(p) emu::ballot((p), __LINE__)
#define __builtin_amdgcn_mov_dpp(v, ctrl, rm, bm, bc) emu::mov_dpp((v), (ctrl), __LINE__)
static inline unsigned atomicAdd(unsigned* p, unsigned v) {
    const unsigned o = *p;
    *p = o + v;
    return o;
}
static inline long long __double_as_longlong(double d) {
    long long r;
    std::memcpy(&r, &d, 8);
    return r;
}
static inline double __longlong_as_double(long long x) {
    double r;
    std::memcpy(&r, &x, 8);
    return r;
}

#include "emu_kernel.inc"

// 0: sbz_wave_kernel (fused stage combinations), 1: sbz_wave_kernel_o2 (two-pass combinations),
// 2: sbz_wave4_kernel (four modes per wavefront; models whose border and chains fit 16 lanes),
// 3: sbz_wave_kernel_altlu (the other LDS layout of the border factors),
// 4: sbz_wave2h_kernel (two modes per wavefront; models whose border and chains fit 32 lanes only)
static int g_o2 = 0;
extern "C" void emu_set_o2(int on) { g_o2 = on; }
extern "C" int emu_has_wave4(void) { return SBZ_WAVE4_OK; }
extern "C" int emu_has_wave2h(void) { return SBZ_WAVE2H_OK; }
static void wave_body(void* a) {
#if SBZ_WAVE4_OK
    if (g_o2 == 2) {
        sbz_wave4_kernel(*(SbzSolveArgs*)a);
        return;
    }
#endif
#if SBZ_WAVE2H_OK
    if (g_o2 == 4) {
        sbz_wave2h_kernel(*(SbzSolveArgs*)a);
        return;
    }
#endif
    if (g_o2 == 3) sbz_wave_kernel_altlu(*(SbzSolveArgs*)a);
    else if (g_o2) sbz_wave_kernel_o2(*(SbzSolveArgs*)a);
    else sbz_wave_kernel(*(SbzSolveArgs*)a);
}

extern "C" const char* emu_error(void) { return emu::error; }
extern "C" int emu_n(void) { return SBZ_N; }

// With the flag set, emu_wave_solve hands the kernel a source background table built by
// sbz_srcbg_row (the body of sbz_srcbg_kernel) instead of letting it evaluate the background per mode.
static int g_srcbg = 0;
extern "C" void emu_set_srcbg(int on) { g_srcbg = on; }

// Solve n_inst instances with the wave kernel on one emulated workgroup.  The knot table of every
// set is laid out as the runtime uploads it: t[nk], u[nk][5], du[nk][5] at knot_off[set].
extern "C" int emu_wave_solve(const double* params, const double* knots, const long long* knot_off,
                              const int* knot_n, const int* inst_set, const double* inst_k,
                              const double* inst_t0, int n_inst, double rtol, double atol, long long maxiters,
                              const double* saveat, int n_save, double* out_u, double* out_save, double* out_src,
                              int* retcode, long long* stats) {
    unsigned queue = 0;
    SbzSolveArgs a;
    std::memset(&a, 0, sizeof a);
    a.params = params;
    a.knots = knots;
    a.knot_off = knot_off;
    a.knot_n = knot_n;
    a.inst_set = inst_set;
    a.inst_k = inst_k;
    a.inst_t0 = inst_t0;
    a.order = nullptr;
    a.n_inst = n_inst;
    a.rtol = rtol;
    a.atol = atol;
    a.maxiters = maxiters;
    a.saveat = n_save ? saveat : nullptr;
    a.n_save = n_save;
    a.out_u = out_u;
    a.out_save = out_save;
    a.out_src = out_src;
    a.retcode = retcode;
    a.stats = stats;
    a.queue = &queue;
    a.S = 1;
    a.xcd_off = nullptr;
    a.n_xcd = 1;
    std::vector<double> tab;
    if (g_srcbg && out_src && n_save > 0) {
        int n_sets = 1;
        for (int i = 0; i < n_inst; ++i) n_sets = inst_set[i] + 1 > n_sets ? inst_set[i] + 1 : n_sets;
        const long long n_rows = (long long)n_sets * n_save;
        tab.assign((size_t)n_rows * SBZ_SRCBG_STRIDE, std::nan(""));
        for (long long g = 0; g < n_rows; ++g)
            sbz_srcbg_row(params, knots, knot_off, knot_n, saveat, 0, n_save, g, tab.data());
        a.srcbg = tab.data();
    }
    return emu::run(wave_body, &a);
}

#if SBZ_TAN_OK
// Tangent solve (sbz_wave_tan_kernel) of n_inst instances of set 0 with the perturbed sets 1 (log θ + h)
// and 2 (log θ - h): params [3][SBZ_NPAR], knots of the three sets at knot_off[0..2].
static void tan_body(void* a) { sbz_wave_tan_kernel(*(SbzSolveArgs*)a); }
extern "C" int emu_tan_solve(const double* params, const double* knots, const long long* knot_off, const int* knot_n,
                             const double* inst_k, const double* inst_t0, int n_inst, double rtol, double atol,
                             long long maxiters, double tscale, double* out_u, double* out_s, double* out_fend,
                             int* retcode, long long* stats) {
    unsigned queue = 0;
    std::vector<int> set0(n_inst, 0), setp(n_inst, 1), setm(n_inst, 2);
    std::vector<double> tsc(n_inst, tscale);
    SbzSolveArgs a;
    std::memset(&a, 0, sizeof a);
    a.params = params;
    a.knots = knots;
    a.knot_off = knot_off;
    a.knot_n = knot_n;
    a.inst_set = set0.data();
    a.inst_setp = setp.data();
    a.inst_setm = setm.data();
    a.inst_tscale = tsc.data();
    a.inst_k = inst_k;
    a.inst_t0 = inst_t0;
    a.n_inst = n_inst;
    a.rtol = rtol;
    a.atol = atol;
    a.maxiters = maxiters;
    a.out_u = out_u;
    a.out_s = out_s;
    a.out_fend = out_fend;
    a.retcode = retcode;
    a.stats = stats;
    a.queue = &queue;
    a.S = 1;
    a.n_xcd = 1;
    return emu::run(tan_body, &a);
}
extern "C" int emu_has_tan(void) { return 1; }
#else
extern "C" int emu_has_tan(void) { return 0; }
#endif
