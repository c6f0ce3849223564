// Host runtime of the perturbation solver: model handles (code objects compiled by hiprtc or
// ahead of time by hipcc, cached on disk), background tables, and solve plans whose inputs stay
// resident in device memory between launches.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "sbz_internal.h"
#include "sbz_thermo.h"

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            sbz_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return SBZ_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

// Row stride of the source background table (SBZ_SRCBG_STRIDE in sbz_kernel.hip).
static constexpr int kSrcBgStride = 80;


bool sbz_read_file(const char* path, std::vector<char>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !out.empty();
}

int sbz_compile_hiprtc(const char* src, std::vector<char>& code) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src, "sbz_model.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        sbz_set_error("hiprtcCreateProgram failed");
        return SBZ_ERR_COMPILE;
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        sbz_set_error("hiprtc compile failed: %.900s", log.c_str());
        hiprtcDestroyProgram(&prog);
        return SBZ_ERR_COMPILE;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return SBZ_OK;
}

extern "C" int sbz_compile(const char* hip_src, const char* out_path) {
    sbz_clear_error();
    if (!hip_src || !out_path) {
        sbz_set_error("sbz_compile: null argument");
        return SBZ_ERR_ARG;
    }
    std::vector<char> code;
    int st = sbz_compile_hiprtc(hip_src, code);
    if (st != SBZ_OK) return st;
    std::ofstream f(out_path, std::ios::binary);
    f.write(code.data(), (std::streamsize)code.size());
    return f ? SBZ_OK : SBZ_ERR_ARG;
}

extern "C" int sbz_model_create(const char* hip_src, const sbz_model_desc* desc, sbz_model** out) {
    sbz_clear_error();
    if (!desc || !out || desc->n <= 0) {
        sbz_set_error("sbz_model_create: invalid descriptor");
        return SBZ_ERR_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        sbz_set_error("sbz_model_create: no GPU visible");
        return SBZ_ERR_NODEV;
    }
    std::vector<char> code;
    bool have = desc->cache_path && sbz_read_file(desc->cache_path, code);
    if (!have) {
        if (!hip_src) {
            sbz_set_error("sbz_model_create: no cached code object at %s and no source",
                          desc->cache_path ? desc->cache_path : "(null)");
            return SBZ_ERR_COMPILE;
        }
        int st = sbz_compile_hiprtc(hip_src, code);
        if (st != SBZ_OK) return st;
        if (desc->cache_path) {
            std::ofstream f(desc->cache_path, std::ios::binary);
            f.write(code.data(), (std::streamsize)code.size());
        }
    }
    sbz_model* m = new sbz_model();
    m->n = desc->n;
    m->nnz_lu = desc->nnz_lu;
    (void)hipGetDevice(&m->device);
    hipError_t e = hipModuleLoadData(&m->module, code.data());
    if (e != hipSuccess) {
        sbz_set_error("hipModuleLoadData: %s", hipGetErrorString(e));
        delete m;
        return SBZ_ERR_COMPILE;
    }
    if (hipModuleGetFunction(&m->k_wave, m->module, "sbz_wave_kernel") != hipSuccess ||
        hipModuleGetFunction(&m->k_wave2, m->module, "sbz_wave_kernel_o2") != hipSuccess ||
        hipModuleGetFunction(&m->k_pk_modes, m->module, "sbz_pk_modes_kernel") != hipSuccess ||
        hipModuleGetFunction(&m->k_srcbg, m->module, "sbz_srcbg_kernel") != hipSuccess ||
        hipModuleGetFunction(&m->k_fill, m->module, "sbz_fill_kernel") != hipSuccess ||
        hipModuleGetFunction(&m->k_qreset, m->module, "sbz_queue_reset_kernel") != hipSuccess ||
        hipModuleGetFunction(&m->k_pk, m->module, "sbz_pk_kernel") != hipSuccess) {
        sbz_set_error("code object lacks the solver kernels");
        (void)hipModuleUnload(m->module);
        delete m;
        return SBZ_ERR_COMPILE;
    }
    if (hipModuleGetFunction(&m->k_wave4, m->module, "sbz_wave4_kernel") != hipSuccess) m->k_wave4 = nullptr;
    if (hipModuleGetFunction(&m->k_wave2h, m->module, "sbz_wave2h_kernel") != hipSuccess) m->k_wave2h = nullptr;
    if (hipModuleGetFunction(&m->k_wave_alt, m->module, "sbz_wave_kernel_altlu") != hipSuccess) m->k_wave_alt = nullptr;
    if (hipModuleGetFunction(&m->k_tan, m->module, "sbz_wave_tan_kernel") != hipSuccess ||
        hipModuleGetFunction(&m->k_pk_tan, m->module, "sbz_pk_tan_kernel") != hipSuccess)
        m->k_tan = m->k_pk_tan = nullptr;
    (void)hipGetLastError();
    if (hipStreamCreate(&m->stream) != hipSuccess) {
        sbz_set_error("hipStreamCreate failed");
        (void)hipModuleUnload(m->module);
        delete m;
        return SBZ_ERR_HIP;
    }
    *out = m;
    return SBZ_OK;
}

extern "C" int sbz_model_destroy(sbz_model* m) {
    if (!m) return SBZ_OK;
    if (m->stream) (void)hipStreamDestroy(m->stream);
    if (m->module) (void)hipModuleUnload(m->module);
    delete m;
    return SBZ_OK;
}

extern "C" int sbz_bg_upload(sbz_model* m, int set, const double* t, const double* u, const double* du,
                             int n_knots, const double* params) {
    sbz_clear_error();
    if (!m || set < 0 || !t || !u || !du || !params || n_knots < 2) {
        sbz_set_error("sbz_bg_upload: invalid argument");
        return SBZ_ERR_ARG;
    }
    for (int i = 1; i < n_knots; ++i)
        if (!(t[i] > t[i - 1])) {
            sbz_set_error("sbz_bg_upload: knot times are not strictly increasing at %d", i);
            return SBZ_ERR_ARG;
        }
    if ((int)m->sets.size() <= set) m->sets.resize(set + 1);
    SbzSetTable& s = m->sets[set];
    s.nk = n_knots;
    s.knots.assign(11 * (size_t)n_knots, 0.0);
    std::memcpy(s.knots.data(), t, sizeof(double) * n_knots);
    std::memcpy(s.knots.data() + n_knots, u, sizeof(double) * 5 * n_knots);
    std::memcpy(s.knots.data() + 6 * (size_t)n_knots, du, sizeof(double) * 5 * n_knots);
    std::memcpy(s.params, params, sizeof(s.params));
    sbz_params_derive(s.params);
    s.valid = true;
    return SBZ_OK;
}

static void plan_free(sbz_plan* p) {
    void* bufs[] = {p->d_params, p->d_knots, p->d_k, p->d_t0, p->d_saveat, p->d_out, p->d_save, p->d_src,
                    p->d_pk, p->d_koff, p->d_stats, p->d_knotn, p->d_set, p->d_order, p->d_ret, p->d_queue,
                    p->d_xcd_off, p->d_setp, p->d_setm, p->d_tscale, p->d_s, p->d_fend, p->d_dlogpk,
                    p->d_srcbg};
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
}

template <class T>
static hipError_t upload(T** dst, const T* src, size_t n) {
    hipError_t e = hipMalloc((void**)dst, sizeof(T) * std::max<size_t>(n, 1));
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice);
}

hipError_t sbz_launch_solve(sbz_model* m, int algo, SbzSolveArgsHost* args, unsigned units) {
    void* kargs[] = {args};
    hipFunction_t kf = (algo == SBZ_ALGO_WAVE2)             ? m->k_wave2
                       : (algo == SBZ_ALGO_WAVE4)           ? m->k_wave4
                       : (algo == SBZ_ALGO_WAVE2H)          ? m->k_wave2h
                       : (algo == SBZ_ALGO_WAVE_ALTLU)      ? m->k_wave_alt
                       : (algo == SBZ_ALGO_TANGENT_INTERNAL) ? m->k_tan
                                                            : m->k_wave;
    return hipModuleLaunchKernel(kf, units, 1, 1, 64, 1, 1, 0, m->stream, kargs, nullptr);
}

int sbz_resident_waves(sbz_model* m, int algo) {
    int per_cu = 0, ncu = 0;
    hipFunction_t kw = (algo == SBZ_ALGO_WAVE2)             ? m->k_wave2
                       : (algo == SBZ_ALGO_WAVE4)           ? m->k_wave4
                       : (algo == SBZ_ALGO_WAVE2H)          ? m->k_wave2h
                       : (algo == SBZ_ALGO_WAVE_ALTLU)      ? m->k_wave_alt
                       : (algo == SBZ_ALGO_TANGENT_INTERNAL) ? m->k_tan
                                                            : m->k_wave;
    if (!kw) return 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kw, 64, 0) != hipSuccess || per_cu < 1)
        per_cu = (algo == SBZ_ALGO_WAVE2) ? 8 : (algo == SBZ_ALGO_WAVE4 || algo == SBZ_ALGO_WAVE2H) ? 2 : 4;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device) != hipSuccess || ncu < 1)
        ncu = 256;
    return per_cu * ncu;
}

extern "C" int sbz_model_capacity(sbz_model* m, int algo, int* capacity) {
    sbz_clear_error();
    if (!m || !capacity ||
        (algo != SBZ_ALGO_WAVE && algo != SBZ_ALGO_WAVE2 && algo != SBZ_ALGO_WAVE4 && algo != SBZ_ALGO_WAVE_ALTLU &&
         algo != SBZ_ALGO_WAVE2H)) {
        sbz_set_error("sbz_model_capacity: invalid argument");
        return SBZ_ERR_ARG;
    }
    hipFunction_t kf = (algo == SBZ_ALGO_WAVE2)        ? m->k_wave2
                       : (algo == SBZ_ALGO_WAVE4)      ? m->k_wave4
                       : (algo == SBZ_ALGO_WAVE2H)     ? m->k_wave2h
                       : (algo == SBZ_ALGO_WAVE_ALTLU) ? m->k_wave_alt
                                                       : m->k_wave;
    if (!kf) {
        *capacity = 0;
        return SBZ_OK;
    }
    int per_cu = 0, ncu = 0;
    HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, 64, 0));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device));
    *capacity = per_cu * ncu * (algo == SBZ_ALGO_WAVE4 ? 4 : algo == SBZ_ALGO_WAVE2H ? 2 : 1);
    return SBZ_OK;
}

extern "C" int sbz_plan_create(sbz_model* m, const int* inst_set, const double* inst_k, const double* inst_t0,
                               int n_inst, const sbz_opts* opts, const double* saveat, int n_save, sbz_plan** out) {
    sbz_clear_error();
    if (!m || !inst_set || !inst_k || !inst_t0 || !opts || !out || n_inst < 0 || n_save < 0 ||
        (n_save > 0 && !saveat)) {
        sbz_set_error("sbz_plan_create: invalid argument");
        return SBZ_ERR_ARG;
    }
    if (!(opts->rtol > 0) || !(opts->atol > 0)) {
        sbz_set_error("sbz_plan_create: tolerances must be positive");
        return SBZ_ERR_ARG;
    }
    int n_sets = (int)m->sets.size();
    for (int i = 0; i < n_inst; ++i) {
        if (inst_set[i] < 0 || inst_set[i] >= n_sets || !m->sets[inst_set[i]].valid) {
            sbz_set_error("sbz_plan_create: instance %d refers to parameter set %d without a background table", i,
                          inst_set[i]);
            return SBZ_ERR_ARG;
        }
    }
    // SBZ_SAVEAT_PER_SET: saveat holds one ascending row of n_save times per parameter set
    const bool per_set = n_save > 0 && (opts->save & SBZ_SAVEAT_PER_SET);
    const int n_rows = per_set ? n_sets : 1;
    for (int r = 0; r < n_rows; ++r)
        for (int i = 1; i < n_save; ++i)
            if (!(saveat[(size_t)r * n_save + i] >= saveat[(size_t)r * n_save + i - 1])) {
                sbz_set_error("sbz_plan_create: saveat must be ascending (row %d, entry %d)", r, i);
                return SBZ_ERR_ARG;
            }
    (void)hipSetDevice(m->device);
    sbz_plan* p = new sbz_plan();
    p->m = m;
    p->n_inst = n_inst;
    p->n_sets = n_sets;
    p->n_save = n_save;
    p->rtol = opts->rtol;
    p->atol = opts->atol;
    p->maxiters = opts->maxiters > 0 ? opts->maxiters : 100000;
    p->save_flags = n_save > 0 ? opts->save : 0;
    if (n_save > 0 && !(p->save_flags & (SBZ_SAVE_STATES | SBZ_SAVE_SOURCES))) p->save_flags |= SBZ_SAVE_STATES;
    p->saveat_stride = per_set ? n_save : 0;
    if (opts->algo != 0 && opts->algo != SBZ_ALGO_WAVE && opts->algo != SBZ_ALGO_WAVE2 && opts->algo != SBZ_ALGO_WAVE4 &&
        opts->algo != SBZ_ALGO_WAVE_ALTLU && opts->algo != SBZ_ALGO_WAVE2H) {
        sbz_set_error("sbz_plan_create: unknown algo %d", opts->algo);
        delete p;
        return SBZ_ERR_ARG;
    }
    if (opts->algo == SBZ_ALGO_WAVE_ALTLU && !m->k_wave_alt) {
        sbz_set_error("sbz_plan_create: the model's code object has no SBZ_ALGO_WAVE_ALTLU kernel");
        delete p;
        return SBZ_ERR_ARG;
    }
    if (opts->algo == SBZ_ALGO_WAVE4 && (!m->k_wave4 || p->save_flags != 0)) {
        sbz_set_error(!m->k_wave4 ? "sbz_plan_create: SBZ_ALGO_WAVE4 needs a model whose border and chains fit 16 lanes"
                                  : "sbz_plan_create: SBZ_ALGO_WAVE4 saves no states or sources (final states only)");
        delete p;
        return SBZ_ERR_ARG;
    }
    if (opts->algo == SBZ_ALGO_WAVE2H && (!m->k_wave2h || p->save_flags != 0)) {
        sbz_set_error(!m->k_wave2h
                          ? "sbz_plan_create: SBZ_ALGO_WAVE2H needs a model without the four-mode kernel whose border "
                            "and chains fit 32 lanes"
                          : "sbz_plan_create: SBZ_ALGO_WAVE2H saves no states or sources (final states only)");
        delete p;
        return SBZ_ERR_ARG;
    }
    p->algo = opts->algo ? opts->algo : SBZ_ALGO_WAVE;
    // `lanes`: concurrently resident wavefronts (the persistent grid)
    int lanes = opts->lanes > 0 ? opts->lanes : n_inst;
    if (opts->lanes <= 0) {
        // persistent grid: as many wavefronts as can be resident at once, each pulling instances
        // (four at a time for SBZ_ALGO_WAVE4, two for SBZ_ALGO_WAVE2H)
        const int need = (p->algo == SBZ_ALGO_WAVE4) ? (n_inst + 3) / 4 : (p->algo == SBZ_ALGO_WAVE2H) ? (n_inst + 1) / 2 : n_inst;
        lanes = std::min(need, sbz_resident_waves(m, p->algo));
    }
    lanes = std::max(1, std::min(lanes, std::max(n_inst, 1)));
    p->lanes = lanes;
    // concatenated background tables
    std::vector<double> params((size_t)n_sets * SBZ_NPAR), knots;
    std::vector<long long> koff(n_sets);
    std::vector<int> kn(n_sets);
    for (int s = 0; s < n_sets; ++s) {
        std::memcpy(params.data() + (size_t)s * SBZ_NPAR, m->sets[s].params, sizeof(double) * SBZ_NPAR);
        koff[s] = (long long)knots.size();
        kn[s] = m->sets[s].nk;
        knots.insert(knots.end(), m->sets[s].knots.begin(), m->sets[s].knots.end());
    }
    // longest-first order: estimated cost grows with k (step count ∝ k at high k).  With several
    // parameter sets and a wavefront kernel, the order is cut into one segment per XCD (the sets
    // with set mod 8 = x), each longest first, so every XCD's L2 holds only its sets' knot tables.
    std::vector<int> order(n_inst);
    std::iota(order.begin(), order.end(), 0);
    int used_sets = 0;
    {
        std::vector<char> seen(n_sets, 0);
        for (int i = 0; i < n_inst; ++i)
            if (!seen[inst_set[i]]) { seen[inst_set[i]] = 1; ++used_sets; }
    }
    p->n_xcd = (used_sets >= SBZ_N_XCD) ? SBZ_N_XCD : 1;
    std::vector<int> xcd_off(SBZ_N_XCD + 1, 0);
    if (p->n_xcd > 1) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            const int xa = inst_set[a] % SBZ_N_XCD, xb = inst_set[b] % SBZ_N_XCD;
            return xa != xb ? xa < xb : inst_k[a] > inst_k[b];
        });
        for (int i = 0; i < n_inst; ++i) ++xcd_off[inst_set[i] % SBZ_N_XCD + 1];
        for (int x = 0; x < SBZ_N_XCD; ++x) xcd_off[x + 1] += xcd_off[x];
    } else {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst_k[a] > inst_k[b]; });
    }
    p->h_order = order;
    p->h_k.assign(inst_k, inst_k + n_inst);
    hipError_t e = hipSuccess;
#define UP(dst, src, n) \
    if (e == hipSuccess) e = upload(dst, src, n)
    UP(&p->d_params, params.data(), params.size());
    UP(&p->d_knots, knots.data(), knots.size());
    UP(&p->d_koff, koff.data(), koff.size());
    UP(&p->d_knotn, kn.data(), kn.size());
    UP(&p->d_set, inst_set, (size_t)n_inst);
    UP(&p->d_k, inst_k, (size_t)n_inst);
    UP(&p->d_t0, inst_t0, (size_t)n_inst);
    UP(&p->d_order, order.data(), (size_t)n_inst);
    if (p->n_xcd > 1) UP(&p->d_xcd_off, xcd_off.data(), xcd_off.size());
    UP(&p->d_saveat, saveat, (size_t)n_save * n_rows);
#undef UP
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_out, sizeof(double) * std::max<size_t>((size_t)n_inst * m->n, 1));
    if (e == hipSuccess && (p->save_flags & SBZ_SAVE_STATES))
        e = hipMalloc((void**)&p->d_save, sizeof(double) * (size_t)n_inst * n_save * m->n);
    if (e == hipSuccess && (p->save_flags & SBZ_SAVE_SOURCES))
        e = hipMalloc((void**)&p->d_src, sizeof(double) * (size_t)n_inst * n_save * 3);
    if (e == hipSuccess && (p->save_flags & SBZ_SAVE_SOURCES))
        e = hipMalloc((void**)&p->d_srcbg, sizeof(double) * kSrcBgStride * (size_t)n_sets * n_save);
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_ret, sizeof(int) * std::max(n_inst, 1));
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_stats, sizeof(long long) * 6 * std::max(n_inst, 1));
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_pk, sizeof(double) * std::max(n_inst, 1));
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_queue, sizeof(unsigned int) * SBZ_N_XCD);
    if (e == hipSuccess) e = hipEventCreate(&p->ev0);
    if (e == hipSuccess) e = hipEventCreate(&p->ev1);
    if (e != hipSuccess) {
        sbz_set_error("sbz_plan_create: %s", hipGetErrorString(e));
        plan_free(p);
        delete p;
        return SBZ_ERR_HIP;
    }
    *out = p;
    return SBZ_OK;
}

static hipError_t launch_fill(sbz_model* m, double* ptr, long long n, double v) {
    if (!ptr || n <= 0) return hipSuccess;
    void* args[] = {&ptr, &n, &v};
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    return hipModuleLaunchKernel(m->k_fill, (unsigned)blocks, 1, 1, 256, 1, 1, 0, m->stream, args, nullptr);
}

static hipError_t launch_pk(sbz_plan* p) {
    sbz_model* m = p->m;
    const double* params = p->d_params;
    const double* knots = p->d_knots;
    const long long* koff = p->d_koff;
    const int* kn = p->d_knotn;
    const int* set = p->d_set;
    const double* kk = p->d_k;
    const double* uf = p->d_out;
    int n = p->n_inst;
    double* dpk = p->d_pk;
    void* args[] = {&params, &knots, &koff, &kn, &set, &kk, &uf, &n, &dpk};
    return hipModuleLaunchKernel(m->k_pk, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0, m->stream, args, nullptr);
}

// The kernel arguments of a plan's solve launch.
void sbz_plan_args(const sbz_plan* p, SbzSolveArgsHost* out) {
    SbzSolveArgsHost& a = *out;
    a.params = p->d_params;
    a.knots = p->d_knots;
    a.knot_off = p->d_koff;
    a.knot_n = p->d_knotn;
    a.inst_set = p->d_set;
    a.inst_k = p->d_k;
    a.inst_t0 = p->d_t0;
    a.order = p->d_order;
    a.n_inst = p->n_inst;
    a.rtol = p->rtol;
    a.atol = p->atol;
    a.maxiters = p->maxiters;
    a.saveat = p->n_save ? p->d_saveat : nullptr;
    a.n_save = p->n_save;
    a.saveat_stride = p->saveat_stride;
    a.out_u = p->d_out;
    a.out_save = p->d_save;
    a.out_src = p->d_src;
    a.retcode = p->d_ret;
    a.stats = p->d_stats;
    a.queue = p->d_queue;
    a.S = p->lanes;
    a.srcbg = p->d_srcbg;
    a.xcd_off = p->d_xcd_off;
    a.n_xcd = p->n_xcd;
    a.inst_setp = p->d_setp;
    a.inst_setm = p->d_setm;
    a.inst_tscale = p->d_tscale;
    a.out_s = p->d_s;
    a.out_fend = p->d_fend;
}

// The launches of one run on the model's stream: NaN into the save/source outputs, the source
// background table, the queue reset and the persistent solve kernel (bracketed by ev0/ev1 when
// `timed`), then the P(k) kernel when `with_pk`.  Used eagerly and under stream capture.
static hipError_t plan_enqueue(sbz_plan* p, bool timed, bool with_pk) {
    sbz_model* m = p->m;
    hipError_t e = hipSuccess;
    const double nan = __builtin_nan("");
    if (p->d_save) e = launch_fill(m, p->d_save, (long long)p->n_inst * p->n_save * m->n, nan);
    if (e == hipSuccess && p->d_src) e = launch_fill(m, p->d_src, (long long)p->n_inst * p->n_save * 3, nan);
    if (e == hipSuccess && p->d_srcbg) {
        // the background at every save time of every set, shared by all modes of the set
        const double* params = p->d_params;
        const double* knots = p->d_knots;
        const long long* koff = p->d_koff;
        const int* kn = p->d_knotn;
        const double* sv = p->d_saveat;
        int stride = p->saveat_stride, n_save = p->n_save;
        long long n_rows = (long long)p->n_sets * p->n_save;
        double* out = p->d_srcbg;
        void* bargs[] = {&params, &knots, &koff, &kn, &sv, &stride, &n_save, &n_rows, &out};
        e = hipModuleLaunchKernel(m->k_srcbg, (unsigned)((n_rows + 63) / 64), 1, 1, 64, 1, 1, 0, m->stream, bargs,
                                  nullptr);
    }
    unsigned int* q = p->d_queue;
    int nq = p->n_xcd;
    void* qargs[] = {&q, &nq};
    if (e == hipSuccess) e = hipModuleLaunchKernel(m->k_qreset, 1, 1, 1, 64, 1, 1, 0, m->stream, qargs, nullptr);
    SbzSolveArgsHost a;
    sbz_plan_args(p, &a);
    if (e == hipSuccess && timed) e = hipEventRecord(p->ev0, m->stream);
    if (e == hipSuccess) e = sbz_launch_solve(m, p->algo, &a, (unsigned)p->lanes);
    if (e == hipSuccess && timed) e = hipEventRecord(p->ev1, m->stream);
    if (e == hipSuccess && with_pk) e = launch_pk(p);
    return e;
}

extern "C" int sbz_plan_run(sbz_plan* p, float* kernel_ms) {
    sbz_clear_error();
    if (!p) return SBZ_ERR_ARG;
    p->pk_ready = false;
    if (p->n_inst == 0) {
        if (kernel_ms) *kernel_ms = 0.0f;
        return SBZ_OK;
    }
    HIPCHK(plan_enqueue(p, true, false));
    HIPCHK(hipEventSynchronize(p->ev1));
    HIPCHK(hipEventElapsedTime(&p->last_ms, p->ev0, p->ev1));
    if (kernel_ms) *kernel_ms = p->last_ms;
    return SBZ_OK;
}

extern "C" int sbz_plan_launch(sbz_plan* p) {
    sbz_clear_error();
    if (!p) return SBZ_ERR_ARG;
    p->pk_ready = false;
    if (p->n_inst == 0) return SBZ_OK;
    HIPCHK(plan_enqueue(p, true, false));
    return SBZ_OK;
}

extern "C" int sbz_plan_wait(sbz_plan* p, float* kernel_ms) {
    sbz_clear_error();
    if (!p) return SBZ_ERR_ARG;
    if (p->n_inst == 0) {
        if (kernel_ms) *kernel_ms = 0.0f;
        return SBZ_OK;
    }
    HIPCHK(hipEventSynchronize(p->ev1));
    HIPCHK(hipEventElapsedTime(&p->last_ms, p->ev0, p->ev1));
    if (kernel_ms) *kernel_ms = p->last_ms;
    return SBZ_OK;
}

extern "C" int sbz_plan_run_graph(sbz_plan* p, float* graph_ms) {
    sbz_clear_error();
    if (!p) return SBZ_ERR_ARG;
    p->pk_ready = false;
    if (p->n_inst == 0) {
        if (graph_ms) *graph_ms = 0.0f;
        return SBZ_OK;
    }
    sbz_model* m = p->m;
    if (!p->gexec) {
        HIPCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = plan_enqueue(p, false, true);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(m->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            sbz_set_error("sbz_plan_run_graph: capture failed: %s",
                          hipGetErrorString(e != hipSuccess ? e : e2));
            return SBZ_ERR_HIP;
        }
        p->graph = g;
        e = hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            p->gexec = nullptr;
            sbz_set_error("sbz_plan_run_graph: hipGraphInstantiate: %s", hipGetErrorString(e));
            return SBZ_ERR_HIP;
        }
    }
    HIPCHK(hipEventRecord(p->ev0, m->stream));
    HIPCHK(hipGraphLaunch(p->gexec, m->stream));
    HIPCHK(hipEventRecord(p->ev1, m->stream));
    HIPCHK(hipEventSynchronize(p->ev1));
    HIPCHK(hipEventElapsedTime(&p->last_ms, p->ev0, p->ev1));
    if (graph_ms) *graph_ms = p->last_ms;
    p->pk_ready = true;
    return SBZ_OK;
}

extern "C" int sbz_plan_pk(sbz_plan* p, double* pk) {
    sbz_clear_error();
    if (!p || !pk) return SBZ_ERR_ARG;
    if (p->n_inst == 0) return SBZ_OK;
    sbz_model* m = p->m;
    if (!p->pk_ready) HIPCHK(launch_pk(p));
    HIPCHK(hipMemcpyAsync(pk, p->d_pk, sizeof(double) * p->n_inst, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return SBZ_OK;
}

extern "C" int sbz_plan_pk_modes(sbz_plan* p, const int* masks, int n_masks, int at_saves, double* out) {
    sbz_clear_error();
    if (!p || !masks || !out || n_masks < 1) {
        sbz_set_error("sbz_plan_pk_modes: invalid argument");
        return SBZ_ERR_ARG;
    }
    for (int j = 0; j < n_masks; ++j)
        if (masks[j] < 1 || masks[j] > 7) {
            sbz_set_error("sbz_plan_pk_modes: species mask %d is not a non-empty subset of c (1), b (2), h (4)",
                          masks[j]);
            return SBZ_ERR_ARG;
        }
    if (at_saves && (!(p->save_flags & SBZ_SAVE_STATES) || p->n_save < 1 || !p->d_save)) {
        sbz_set_error("sbz_plan_pk_modes: the plan saved no states at saveat times");
        return SBZ_ERR_ARG;
    }
    if (p->n_inst == 0) return SBZ_OK;
    sbz_model* m = p->m;
    const double* params = p->d_params;
    const double* knots = p->d_knots;
    const long long* koff = p->d_koff;
    const int* kn = p->d_knotn;
    const int* set = p->d_set;
    const double* kk = p->d_k;
    const double* states = at_saves ? p->d_save : p->d_out;
    int n = p->n_inst;
    int n_t = at_saves ? p->n_save : 1;
    const double* taus = at_saves ? p->d_saveat : nullptr;
    int stride = p->saveat_stride;
    int nm = n_masks;
    int* d_masks = nullptr;
    double* d_out = nullptr;
    const size_t n_out = (size_t)n_masks * n_t * n;
    HIPCHK(hipMalloc((void**)&d_masks, sizeof(int) * n_masks));
    if (hipMalloc((void**)&d_out, sizeof(double) * n_out) != hipSuccess) {
        (void)hipFree(d_masks);
        sbz_set_error("sbz_plan_pk_modes: hipMalloc of %zu doubles failed", n_out);
        return SBZ_ERR_HIP;
    }
    int st = SBZ_OK;
    void* args[] = {&params, &knots, &koff, &kn, &set, &kk, &states, &n, &n_t, &taus, &stride, &d_masks, &nm, &d_out};
    const long long nth = (long long)n * n_t;
    hipError_t e = hipMemcpyAsync(d_masks, masks, sizeof(int) * n_masks, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess)
        e = hipModuleLaunchKernel(m->k_pk_modes, (unsigned)((nth + 255) / 256), 1, 1, 256, 1, 1, 0, m->stream, args,
                                  nullptr);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * n_out, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        sbz_set_error("sbz_plan_pk_modes: %s", hipGetErrorString(e));
        st = SBZ_ERR_HIP;
    }
    (void)hipFree(d_masks);
    (void)hipFree(d_out);
    return st;
}

extern "C" int sbz_plan_fetch(sbz_plan* p, double* out_u, double* out_save, int* retcode, long long* stats) {
    sbz_clear_error();
    if (!p) return SBZ_ERR_ARG;
    sbz_model* m = p->m;
    size_t n = (size_t)p->n_inst;
    if (n == 0) return SBZ_OK;
    if (out_u) HIPCHK(hipMemcpyAsync(out_u, p->d_out, sizeof(double) * n * m->n, hipMemcpyDeviceToHost, m->stream));
    if (out_save && p->d_save)
        HIPCHK(hipMemcpyAsync(out_save, p->d_save, sizeof(double) * n * p->n_save * m->n, hipMemcpyDeviceToHost,
                              m->stream));
    if (retcode) HIPCHK(hipMemcpyAsync(retcode, p->d_ret, sizeof(int) * n, hipMemcpyDeviceToHost, m->stream));
    if (stats) HIPCHK(hipMemcpyAsync(stats, p->d_stats, sizeof(long long) * 6 * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return SBZ_OK;
}

extern "C" int sbz_plan_fetch_sources(sbz_plan* p, double* out_src) {
    sbz_clear_error();
    if (!p || !out_src) return SBZ_ERR_ARG;
    if (!p->d_src) {
        sbz_set_error("sbz_plan_fetch_sources: plan was created without SBZ_SAVE_SOURCES");
        return SBZ_ERR_ARG;
    }
    HIPCHK(hipMemcpyAsync(out_src, p->d_src, sizeof(double) * (size_t)p->n_inst * p->n_save * 3,
                          hipMemcpyDeviceToHost, p->m->stream));
    HIPCHK(hipStreamSynchronize(p->m->stream));
    return SBZ_OK;
}

extern "C" int sbz_plan_destroy(sbz_plan* p) {
    if (!p) return SBZ_OK;
    plan_free(p);
    delete p;
    return SBZ_OK;
}

extern "C" int sbz_plan_create_tangent(sbz_model* m, const int* inst_set, const int* inst_setp, const int* inst_setm,
                                       const double* inst_tscale, const double* inst_k, const double* inst_t0,
                                       int n_inst, const sbz_opts* opts, sbz_plan** out) {
    sbz_clear_error();
    if (!m || !inst_setp || !inst_setm || !inst_tscale || !opts || !out || n_inst < 0) {
        sbz_set_error("sbz_plan_create_tangent: invalid argument");
        return SBZ_ERR_ARG;
    }
    if (!m->k_tan) {
        sbz_set_error("sbz_plan_create_tangent: the model's tangent kernel does not fit a workgroup's LDS");
        return SBZ_ERR_ARG;
    }
    const int n_sets = (int)m->sets.size();
    for (int i = 0; i < n_inst; ++i) {
        const int sp = inst_setp[i], sm = inst_setm[i];
        if (sp < 0 || sp >= n_sets || !m->sets[sp].valid || sm < 0 || sm >= n_sets || !m->sets[sm].valid) {
            sbz_set_error("sbz_plan_create_tangent: instance %d refers to a perturbed parameter set without a "
                          "background table", i);
            return SBZ_ERR_ARG;
        }
        if (!(inst_tscale[i] > 0.0) || !std::isfinite(inst_tscale[i])) {
            sbz_set_error("sbz_plan_create_tangent: instance %d has a non-positive 1/(2h)", i);
            return SBZ_ERR_ARG;
        }
    }
    sbz_opts o = *opts;
    o.algo = SBZ_ALGO_WAVE;
    o.save = 0;
    sbz_plan* p = nullptr;
    int st = sbz_plan_create(m, inst_set, inst_k, inst_t0, n_inst, &o, nullptr, 0, &p);
    if (st != SBZ_OK) return st;
    p->algo = SBZ_ALGO_TANGENT_INTERNAL;
    if (opts->lanes <= 0) p->lanes = std::max(1, std::min(n_inst, sbz_resident_waves(m, p->algo)));
    hipError_t e = hipSuccess;
    const size_t nn = (size_t)std::max(n_inst, 1);
    e = upload(&p->d_setp, inst_setp, (size_t)n_inst);
    if (e == hipSuccess) e = upload(&p->d_setm, inst_setm, (size_t)n_inst);
    if (e == hipSuccess) e = upload(&p->d_tscale, inst_tscale, (size_t)n_inst);
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_s, sizeof(double) * nn * m->n);
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_fend, sizeof(double) * nn * m->n);
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_dlogpk, sizeof(double) * nn);
    if (e != hipSuccess) {
        sbz_set_error("sbz_plan_create_tangent: %s", hipGetErrorString(e));
        sbz_plan_destroy(p);
        return SBZ_ERR_HIP;
    }
    *out = p;
    return SBZ_OK;
}

extern "C" int sbz_plan_fetch_tangent(sbz_plan* p, double* out_s, double* out_fend) {
    sbz_clear_error();
    if (!p || !p->d_s) {
        sbz_set_error("sbz_plan_fetch_tangent: not a tangent plan");
        return SBZ_ERR_ARG;
    }
    const size_t b = sizeof(double) * (size_t)p->n_inst * p->m->n;
    if (p->n_inst == 0) return SBZ_OK;
    if (out_s) HIPCHK(hipMemcpyAsync(out_s, p->d_s, b, hipMemcpyDeviceToHost, p->m->stream));
    if (out_fend) HIPCHK(hipMemcpyAsync(out_fend, p->d_fend, b, hipMemcpyDeviceToHost, p->m->stream));
    HIPCHK(hipStreamSynchronize(p->m->stream));
    return SBZ_OK;
}

extern "C" int sbz_plan_pk_tangent(sbz_plan* p, double* pk, double* dlogpk) {
    sbz_clear_error();
    if (!p || !p->d_s || !pk || !dlogpk) {
        sbz_set_error("sbz_plan_pk_tangent: not a tangent plan, or a null output");
        return SBZ_ERR_ARG;
    }
    if (p->n_inst == 0) return SBZ_OK;
    sbz_model* m = p->m;
    const double* params = p->d_params;
    const double* knots = p->d_knots;
    const long long* koff = p->d_koff;
    const int* kn = p->d_knotn;
    const int* set = p->d_set;
    const int* setp = p->d_setp;
    const int* setm = p->d_setm;
    const double* tsc = p->d_tscale;
    const double* kk = p->d_k;
    const double* uf = p->d_out;
    const double* sf = p->d_s;
    const double* ff = p->d_fend;
    int n = p->n_inst;
    double* dpk = p->d_pk;
    double* ddl = p->d_dlogpk;
    void* args[] = {&params, &knots, &koff, &kn, &set, &setp, &setm, &tsc, &kk, &uf, &sf, &ff, &n, &dpk, &ddl};
    HIPCHK(hipModuleLaunchKernel(m->k_pk_tan, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0, m->stream, args,
                                 nullptr));
    HIPCHK(hipMemcpyAsync(pk, p->d_pk, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(dlogpk, p->d_dlogpk, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return SBZ_OK;
}

extern "C" int sbz_solve_batch(sbz_model* m, const int* inst_set, const double* inst_k, const double* inst_t0,
                               int n_inst, const sbz_opts* opts, const double* saveat, int n_save, double* out_u,
                               double* out_save, int* retcode, long long* stats) {
    sbz_plan* p = nullptr;
    int st = sbz_plan_create(m, inst_set, inst_k, inst_t0, n_inst, opts, saveat, n_save, &p);
    if (st != SBZ_OK) return st;
    st = sbz_plan_run(p, nullptr);
    if (st == SBZ_OK) st = sbz_plan_fetch(p, out_u, out_save, retcode, stats);
    sbz_plan_destroy(p);
    return st;
}
