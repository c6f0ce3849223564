// Mixed-model launches: the plans of several models (different state sizes, hierarchies and
// species) integrated by one persistent kernel.  The code object holds every model's device code in
// its own namespace and one kernel, sbz_multi_kernel, whose wavefronts pull entries from one queue
// ordered by estimated cost across all models.  An entry names an instance, integrated by its
// model's one-mode body; or, for a model with a several-modes body (two or four modes per
// wavefront) whose plan is large enough to be throughput-bound, an entry is a ticket into that
// body, whose rows then pull the model's instances, longest first, from the model's own queue
// until it is empty.  Each model's plan keeps its own inputs and outputs, so
// sbz_plan_fetch and sbz_plan_pk read the results as after sbz_plan_run.
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
    hipModule_t module = nullptr;        // the one-mode kernel's code object
    hipModule_t module_rows = nullptr;   // the several-modes kernel's (sbz_multi_attach_rows)
    hipFunction_t k_multi = nullptr;   // one-mode entries only
    hipFunction_t k_rows = nullptr;    // entries and tickets (a family with a several-modes body)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<sbz_model*> models;
    int* d_fam = nullptr;
    int* d_inst = nullptr;
    size_t cap = 0;
    unsigned int* d_queue = nullptr;   // [0]: the entry queue; [1 + f]: family f's own queue
    int* d_ford = nullptr;              // the several-modes families' instance orders, longest first
    size_t ford_cap = 0;
    int rows[kMultiMaxF] = {};          // modes per wavefront of each family's body
    int resident = 0, resident_rows = 0;   // resident wavefronts of k_multi, k_rows
    // host side of the entry and order tables: alive while their copies are in flight
    std::vector<int> h_fam, h_inst, h_ford;
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
    if (mm->d_ford) (void)hipFree(mm->d_ford);
    if (mm->ev0) (void)hipEventDestroy(mm->ev0);
    if (mm->ev1) (void)hipEventDestroy(mm->ev1);
    if (mm->stream) (void)hipStreamDestroy(mm->stream);
    if (mm->module) (void)hipModuleUnload(mm->module);
    if (mm->module_rows) (void)hipModuleUnload(mm->module_rows);
}

// Load a mixed-model code object (from cache_path, else compiled from hip_src and cached there),
// take its kernel `name`, and check its family table against the models: state sizes, and the
// modes per wavefront of each family's several-modes body (into rows).
static int load_multi_module(const char* who, const char* hip_src, const char* cache_path, const char* name,
                             sbz_model* const* models, int n_models, hipModule_t* mod, hipFunction_t* fn,
                             int* rows) {
    std::vector<char> code;
    bool have = cache_path && sbz_read_file(cache_path, code);
    if (!have) {
        if (!hip_src) {
            sbz_set_error("%s: no cached code object at %s and no source", who, cache_path ? cache_path : "(null)");
            return SBZ_ERR_COMPILE;
        }
        const int st = sbz_compile_hiprtc(hip_src, code);
        if (st != SBZ_OK) return st;
        if (cache_path) {
            std::ofstream f(cache_path, std::ios::binary);
            f.write(code.data(), (std::streamsize)code.size());
        }
    }
    hipError_t e = hipModuleLoadData(mod, code.data());
    if (e != hipSuccess) {
        *mod = nullptr;
        sbz_set_error("%s: hipModuleLoadData: %s", who, hipGetErrorString(e));
        return SBZ_ERR_COMPILE;
    }
    int st = SBZ_OK;
    hipDeviceptr_t dn = nullptr;
    size_t bytes = 0;
    int nf[kMultiMaxF] = {};
    if (hipModuleGetFunction(fn, *mod, name) != hipSuccess) {
        sbz_set_error("%s: code object lacks %s", who, name);
        st = SBZ_ERR_COMPILE;
    } else if (hipModuleGetGlobal(&dn, &bytes, *mod, "sbz_multi_n") != hipSuccess || bytes != sizeof(nf) ||
               hipMemcpyDtoH(nf, dn, sizeof(nf)) != hipSuccess ||
               hipModuleGetGlobal(&dn, &bytes, *mod, "sbz_multi_modes") != hipSuccess ||
               bytes != sizeof(int) * kMultiMaxF || hipMemcpyDtoH(rows, dn, sizeof(int) * kMultiMaxF) != hipSuccess) {
        sbz_set_error("%s: code object lacks its family tables", who);
        st = SBZ_ERR_COMPILE;
    } else {
        // the state size of every family compiled into the code object must be its model's
        for (int i = 0; i < kMultiMaxF && st == SBZ_OK; ++i) {
            const int want = i < n_models ? models[i]->n : 0;
            if (nf[i] != want) {
                sbz_set_error("%s: family %d of the code object has %d unknowns, model %d has %d", who, i, nf[i], i,
                              want);
                st = SBZ_ERR_ARG;
            } else if (i < n_models && rows[i] != 1 && rows[i] != 2 && rows[i] != 4) {
                sbz_set_error("%s: family %d has %d modes per wavefront", who, i, rows[i]);
                st = SBZ_ERR_COMPILE;
            }
        }
    }
    if (st != SBZ_OK) {
        (void)hipModuleUnload(*mod);
        *mod = nullptr;
    }
    return st;
}

static int resident_of(hipFunction_t fn, int device, int* resident) {
    int per_cu = 0, ncu = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        sbz_set_error("mixed-model kernel occupancy query failed");
        return SBZ_ERR_HIP;
    }
    if (per_cu < 1) {
        sbz_set_error("mixed-model kernel does not fit a CU");
        return SBZ_ERR_COMPILE;
    }
    *resident = per_cu * ncu;
    return SBZ_OK;
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
    MCHK(hipSetDevice(models[0]->device));
    sbz_multi* mm = new sbz_multi();
    mm->device = models[0]->device;
    mm->models.assign(models, models + n_models);
    int st = load_multi_module("sbz_multi_create", hip_src, cache_path, "sbz_multi_kernel", models, n_models,
                               &mm->module, &mm->k_multi, mm->rows);
    if (st == SBZ_OK) st = resident_of(mm->k_multi, mm->device, &mm->resident);
    hipError_t e = hipSuccess;
    if (st == SBZ_OK) {
        e = hipStreamCreate(&mm->stream);
        if (e == hipSuccess) e = hipEventCreate(&mm->ev0);
        if (e == hipSuccess) e = hipEventCreate(&mm->ev1);
        if (e == hipSuccess) e = hipMalloc((void**)&mm->d_queue, sizeof(unsigned int) * (1 + kMultiMaxF));
        if (e != hipSuccess) {
            sbz_set_error("sbz_multi_create: %s", hipGetErrorString(e));
            st = SBZ_ERR_HIP;
        }
    }
    if (st != SBZ_OK) {
        multi_free(mm);
        delete mm;
        return st;
    }
    *out = mm;
    return SBZ_OK;
}

extern "C" int sbz_multi_attach_rows(sbz_multi* mm, const char* hip_src, const char* cache_path) {
    sbz_clear_error();
    if (!mm) return SBZ_ERR_ARG;
    if (mm->module_rows) {
        sbz_set_error("sbz_multi_attach_rows: a kernel is attached already");
        return SBZ_ERR_ARG;
    }
    MCHK(hipSetDevice(mm->device));
    int rows[kMultiMaxF] = {};
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    int st = load_multi_module("sbz_multi_attach_rows", hip_src, cache_path, "sbz_multi_kernel_rows",
                               mm->models.data(), (int)mm->models.size(), &mod, &fn, rows);
    if (st != SBZ_OK) return st;
    for (size_t f = 0; f < mm->models.size(); ++f)
        if (rows[f] != mm->rows[f]) {
            sbz_set_error("sbz_multi_attach_rows: family %d has %d modes per wavefront here, %d in the handle's code object",
                          (int)f, rows[f], mm->rows[f]);
            (void)hipModuleUnload(mod);
            return SBZ_ERR_ARG;
        }
    int resident = 0;
    st = resident_of(fn, mm->device, &resident);
    if (st != SBZ_OK) {
        (void)hipModuleUnload(mod);
        return st;
    }
    mm->module_rows = mod;
    mm->k_rows = fn;
    mm->resident_rows = resident;
    return SBZ_OK;
}

extern "C" int sbz_multi_capacity(sbz_multi* mm, int* capacity) {
    sbz_clear_error();
    if (!mm || !capacity) return SBZ_ERR_ARG;
    *capacity = mm->resident;
    return SBZ_OK;
}

extern "C" int sbz_multi_rows(sbz_multi* mm, int* rows) {
    sbz_clear_error();
    if (!mm || !rows) return SBZ_ERR_ARG;
    for (size_t f = 0; f < mm->models.size(); ++f) rows[f] = mm->k_rows ? mm->rows[f] : 1;
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
    // the state size, so k·n orders the entries across models.  A several-modes family (rows[f] > 1)
    // whose plan holds at least twice the resident wavefronts' worth of instances gets them sorted
    // longest first into its own order, and one ticket per rows[f] of them, placed at the cost of
    // the first instance the ticket's wavefront would take; below that the launch is latency-bound,
    // where the one-mode body, fastest per mode, serves it.
    struct Entry {
        double cost;
        int fam, inst;
    };
    std::vector<Entry> ent;
    std::vector<size_t> ford_off(n_plans, 0);
    std::vector<char> use_rows(n_plans, 0);
    bool rows_run = false;
    for (int f = 0; f < n_plans; ++f) {
        use_rows[f] = mm->k_rows && mm->rows[f] > 1 && plans[f]->n_inst >= 2 * mm->resident_rows;
        rows_run = rows_run || use_rows[f];
    }
    MCHK(hipSetDevice(mm->device));
    // work a plan still has on its model's stream (an sbz_plan_launch not yet waited for, a pending
    // fetch) finishes before the launch writes the plan's outputs; copies of an earlier call that
    // stopped on an error finish before the host tables they read are rewritten
    for (int f = 0; f < n_plans; ++f) MCHK(hipStreamSynchronize(plans[f]->m->stream));
    MCHK(hipStreamSynchronize(mm->stream));
    mm->h_ford.clear();
    for (int f = 0; f < n_plans; ++f) {
        const sbz_plan* p = plans[f];
        const double nn = (double)p->m->n;
        if (!use_rows[f]) {
            for (int j = 0; j < p->n_inst; ++j) {
                const int i = p->h_order[j];
                ent.push_back({p->h_k[i] * nn, f, i});
            }
            continue;
        }
        ford_off[f] = mm->h_ford.size();
        std::vector<int> ord(p->n_inst);
        for (int i = 0; i < p->n_inst; ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [p](int x, int y) { return p->h_k[x] > p->h_k[y]; });
        for (int j = 0; j < p->n_inst; j += mm->rows[f]) ent.push_back({p->h_k[ord[j]] * nn, f, -1});
        mm->h_ford.insert(mm->h_ford.end(), ord.begin(), ord.end());
    }
    std::stable_sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) { return a.cost > b.cost; });
    const int n = (int)ent.size();
    for (int f = 0; f < n_plans; ++f) plans[f]->pk_ready = false;
    if (n == 0) {
        if (kernel_ms) *kernel_ms = 0.0f;
        return SBZ_OK;
    }
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
    if (!mm->h_ford.empty()) {
        if (mm->h_ford.size() > mm->ford_cap) {
            if (mm->d_ford) MCHK(hipFree(mm->d_ford));
            mm->d_ford = nullptr;
            mm->ford_cap = 0;
            MCHK(hipMalloc((void**)&mm->d_ford, sizeof(int) * mm->h_ford.size()));
            mm->ford_cap = mm->h_ford.size();
        }
        MCHK(hipMemcpyAsync(mm->d_ford, mm->h_ford.data(), sizeof(int) * mm->h_ford.size(), hipMemcpyHostToDevice,
                            mm->stream));
    }
    MCHK(hipMemsetAsync(mm->d_queue, 0, sizeof(unsigned int) * (1 + kMultiMaxF), mm->stream));
    SbzMultiArgsHost a;
    std::memset(&a, 0, sizeof(a));
    for (int f = 0; f < n_plans; ++f) {
        sbz_plan_args(plans[f], &a.fam[f]);
        a.fam[f].xcd_off = nullptr;
        a.fam[f].n_xcd = 1;
        if (!use_rows[f]) {
            a.fam[f].order = nullptr;   // entries name instances directly
        } else {                        // the rows pull from the family's own queue and order
            a.fam[f].order = mm->d_ford + ford_off[f];
            a.fam[f].queue = mm->d_queue + 1 + f;
        }
    }
    a.mfam = mm->d_fam;
    a.minst = mm->d_inst;
    a.n = n;
    a.queue = mm->d_queue;
    void* kargs[] = {&a};
    // the kernel with the several-modes bodies only when some family runs in one: it holds more
    // registers, which the one-mode bodies of latency-bound launches would pay for
    const unsigned waves = (unsigned)std::max(1, std::min(n, rows_run ? mm->resident_rows : mm->resident));
    MCHK(hipEventRecord(mm->ev0, mm->stream));
    MCHK(hipModuleLaunchKernel(rows_run ? mm->k_rows : mm->k_multi, waves, 1, 1, 64, 1, 1, 0, mm->stream, kargs,
                               nullptr));
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
