// Mixed-model launches: the plans of several models (different state sizes, hierarchies and
// species) integrated by one persistent kernel.  The code object holds every model's device code in
// its own namespace and one kernel, sbz_multi_kernel, whose wavefronts pull (model, instance)
// entries from one queue ordered by estimated cost across all models, and run the entry's model on
// the entry's plan buffers.  Each model's plan keeps its own inputs and outputs, so sbz_plan_fetch
// and sbz_plan_pk read the results as after sbz_plan_run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <vector>

#include "sbz_internal.h"

// Families per code object (the dimension of SbzMultiArgs::fam in the generated source).
static constexpr int kMultiMaxF = 8;

struct SbzMultiArgsHost {
    SbzSolveArgsHost fam[kMultiMaxF];
    const int* mfam;
    const int* minst;
    int n;
    unsigned int* queue;
};

struct sbz_multi {
    int device = 0;
    hipModule_t module = nullptr;
    hipFunction_t k_multi = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<sbz_model*> models;
    int* d_fam = nullptr;
    int* d_inst = nullptr;
    size_t cap = 0;
    unsigned int* d_queue = nullptr;
    int resident = 0;
    std::vector<int> h_fam, h_inst;   // host side of the entry tables: alive while their copies are in flight
};

#define MCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            sbz_set_error("%s failed: %s", #x, hipGetErrorString(e_));                            \
            return SBZ_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

static void multi_free(sbz_multi* mm) {
    if (mm->d_fam) (void)hipFree(mm->d_fam);
    if (mm->d_inst) (void)hipFree(mm->d_inst);
    if (mm->d_queue) (void)hipFree(mm->d_queue);
    if (mm->ev0) (void)hipEventDestroy(mm->ev0);
    if (mm->ev1) (void)hipEventDestroy(mm->ev1);
    if (mm->stream) (void)hipStreamDestroy(mm->stream);
    if (mm->module) (void)hipModuleUnload(mm->module);
}

extern "C" int sbz_multi_create(const char* hip_src, const char* cache_path, sbz_model* const* models, int n_models,
                                sbz_multi** out) {
    sbz_clear_error();
    if (!models || !out || n_models < 1 || n_models > kMultiMaxF) {
        sbz_set_error("sbz_multi_create: invalid argument (1 to %d models)", kMultiMaxF);
        return SBZ_ERR_ARG;
    }
    *out = nullptr;
    for (int i = 0; i < n_models; ++i) {
        if (!models[i] || models[i]->device != models[0]->device) {
            sbz_set_error("sbz_multi_create: model %d is missing or on another device", i);
            return SBZ_ERR_ARG;
        }
        for (int j = 0; j < i; ++j)
            if (models[j] == models[i]) {   // one plan per family: a repeated model would run its plan twice
                sbz_set_error("sbz_multi_create: models %d and %d are the same handle", j, i);
                return SBZ_ERR_ARG;
            }
    }
    std::vector<char> code;
    bool have = cache_path && sbz_read_file(cache_path, code);
    if (!have) {
        if (!hip_src) {
            sbz_set_error("sbz_multi_create: no cached code object at %s and no source", cache_path ? cache_path : "(null)");
            return SBZ_ERR_COMPILE;
        }
        const int st = sbz_compile_hiprtc(hip_src, code);
        if (st != SBZ_OK) return st;
        if (cache_path) {
            std::ofstream f(cache_path, std::ios::binary);
            f.write(code.data(), (std::streamsize)code.size());
        }
    }
    MCHK(hipSetDevice(models[0]->device));
    sbz_multi* mm = new sbz_multi();
    mm->device = models[0]->device;
    mm->models.assign(models, models + n_models);
    hipError_t e = hipModuleLoadData(&mm->module, code.data());
    if (e != hipSuccess) {
        sbz_set_error("sbz_multi_create: hipModuleLoadData: %s", hipGetErrorString(e));
        delete mm;
        return SBZ_ERR_COMPILE;
    }
    if (hipModuleGetFunction(&mm->k_multi, mm->module, "sbz_multi_kernel") != hipSuccess) {
        sbz_set_error("sbz_multi_create: code object lacks sbz_multi_kernel");
        multi_free(mm);
        delete mm;
        return SBZ_ERR_COMPILE;
    }
    // the state size of every family compiled into the code object must be its model's
    hipDeviceptr_t dn = nullptr;
    size_t bytes = 0;
    int nf[kMultiMaxF] = {};
    if (hipModuleGetGlobal(&dn, &bytes, mm->module, "sbz_multi_n") != hipSuccess || bytes != sizeof(nf) ||
        hipMemcpyDtoH(nf, dn, sizeof(nf)) != hipSuccess) {
        sbz_set_error("sbz_multi_create: code object lacks its family table");
        multi_free(mm);
        delete mm;
        return SBZ_ERR_COMPILE;
    }
    for (int i = 0; i < kMultiMaxF; ++i) {
        const int want = i < n_models ? models[i]->n : 0;
        if (nf[i] != want) {
            sbz_set_error("sbz_multi_create: family %d of the code object has %d unknowns, model %d has %d", i, nf[i],
                          i, want);
            multi_free(mm);
            delete mm;
            return SBZ_ERR_ARG;
        }
    }
    e = hipStreamCreate(&mm->stream);
    if (e == hipSuccess) e = hipEventCreate(&mm->ev0);
    if (e == hipSuccess) e = hipEventCreate(&mm->ev1);
    if (e == hipSuccess) e = hipMalloc((void**)&mm->d_queue, sizeof(unsigned int));
    int per_cu = 0, ncu = 0;
    if (e == hipSuccess) e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mm->k_multi, 64, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, mm->device);
    if (e != hipSuccess || per_cu < 1) {
        sbz_set_error("sbz_multi_create: %s", e != hipSuccess ? hipGetErrorString(e) : "kernel does not fit a CU");
        multi_free(mm);
        delete mm;
        return e != hipSuccess ? SBZ_ERR_HIP : SBZ_ERR_COMPILE;
    }
    mm->resident = per_cu * ncu;
    *out = mm;
    return SBZ_OK;
}

extern "C" int sbz_multi_capacity(sbz_multi* mm, int* capacity) {
    sbz_clear_error();
    if (!mm || !capacity) return SBZ_ERR_ARG;
    *capacity = mm->resident;
    return SBZ_OK;
}

extern "C" int sbz_multi_run(sbz_multi* mm, sbz_plan* const* plans, int n_plans, float* kernel_ms) {
    sbz_clear_error();
    if (!mm || !plans || n_plans != (int)mm->models.size()) {
        sbz_set_error("sbz_multi_run: expected one plan per model of the code object (%d)",
                      mm ? (int)mm->models.size() : 0);
        return SBZ_ERR_ARG;
    }
    for (int f = 0; f < n_plans; ++f) {
        const sbz_plan* p = plans[f];
        if (!p || p->m != mm->models[f]) {
            sbz_set_error("sbz_multi_run: plan %d does not belong to model %d of the code object", f, f);
            return SBZ_ERR_ARG;
        }
        if (p->save_flags != 0 || p->d_setp) {
            sbz_set_error("sbz_multi_run: plan %d saves states or sources, or is a tangent plan (final states only)", f);
            return SBZ_ERR_ARG;
        }
    }
    // one queue over all plans, longest first: a mode's steps grow with k and each step's cost with
    // the state size, so k·n orders the entries across models
    struct Entry {
        double cost;
        int fam, inst;
    };
    std::vector<Entry> ent;
    for (int f = 0; f < n_plans; ++f) {
        const sbz_plan* p = plans[f];
        for (int j = 0; j < p->n_inst; ++j) {
            const int i = p->h_order[j];
            ent.push_back({p->h_k[i] * (double)p->m->n, f, i});
        }
    }
    std::stable_sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) { return a.cost > b.cost; });
    const int n = (int)ent.size();
    for (int f = 0; f < n_plans; ++f) plans[f]->pk_ready = false;
    if (n == 0) {
        if (kernel_ms) *kernel_ms = 0.0f;
        return SBZ_OK;
    }
    MCHK(hipSetDevice(mm->device));
    // work a plan still has on its model's stream (an sbz_plan_launch not yet waited for, a pending
    // fetch) finishes before the launch writes the plan's outputs
    for (int f = 0; f < n_plans; ++f) MCHK(hipStreamSynchronize(plans[f]->m->stream));
    MCHK(hipStreamSynchronize(mm->stream));   // copies of an earlier call that stopped on an error
    if ((size_t)n > mm->cap) {
        if (mm->d_fam) MCHK(hipFree(mm->d_fam));
        if (mm->d_inst) MCHK(hipFree(mm->d_inst));
        mm->d_fam = mm->d_inst = nullptr;
        mm->cap = 0;
        MCHK(hipMalloc((void**)&mm->d_fam, sizeof(int) * (size_t)n));
        MCHK(hipMalloc((void**)&mm->d_inst, sizeof(int) * (size_t)n));
        mm->cap = (size_t)n;
    }
    mm->h_fam.resize(n);
    mm->h_inst.resize(n);
    for (int g = 0; g < n; ++g) {
        mm->h_fam[g] = ent[g].fam;
        mm->h_inst[g] = ent[g].inst;
    }
    MCHK(hipMemcpyAsync(mm->d_fam, mm->h_fam.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, mm->stream));
    MCHK(hipMemcpyAsync(mm->d_inst, mm->h_inst.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, mm->stream));
    MCHK(hipMemsetAsync(mm->d_queue, 0, sizeof(unsigned int), mm->stream));
    SbzMultiArgsHost a;
    std::memset(&a, 0, sizeof(a));
    for (int f = 0; f < n_plans; ++f) {
        sbz_plan_args(plans[f], &a.fam[f]);
        a.fam[f].order = nullptr;     // entries name instances directly
        a.fam[f].xcd_off = nullptr;
        a.fam[f].n_xcd = 1;
    }
    a.mfam = mm->d_fam;
    a.minst = mm->d_inst;
    a.n = n;
    a.queue = mm->d_queue;
    void* kargs[] = {&a};
    const unsigned waves = (unsigned)std::max(1, std::min(n, mm->resident));
    MCHK(hipEventRecord(mm->ev0, mm->stream));
    MCHK(hipModuleLaunchKernel(mm->k_multi, waves, 1, 1, 64, 1, 1, 0, mm->stream, kargs, nullptr));
    MCHK(hipEventRecord(mm->ev1, mm->stream));
    MCHK(hipEventSynchronize(mm->ev1));
    float ms = 0.0f;
    MCHK(hipEventElapsedTime(&ms, mm->ev0, mm->ev1));
    for (int f = 0; f < n_plans; ++f) plans[f]->last_ms = ms;
    if (kernel_ms) *kernel_ms = ms;
    return SBZ_OK;
}

extern "C" int sbz_multi_destroy(sbz_multi* mm) {
    if (!mm) return SBZ_OK;
    (void)hipSetDevice(mm->device);
    multi_free(mm);
    delete mm;
    return SBZ_OK;
}
