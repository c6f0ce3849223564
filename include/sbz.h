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

/* The same solve for n parameter sets at once on the GPU, one thread per set (solvebg of
 * src/solve.jl:427-435 for each cosmology of a batch; the reference runs them one after the other on
 * the CPU).  params [n][SBZ_NPAR] in/out; t [n][max_knots], u and du [n][max_knots][5];
 * n_knots, n_rejected (may be NULL) and status [n] per set, status 0 on success, else 1 maximum
 * iterations, 2 step size below minimum, 3 more steps than max_knots, 4 a = 1 not reached.
 * Entries past n_knots[i] in a set's rows are left untouched.  Per-set failures do not fail the
 * call (the first is described by sbz_last_error). */
int sbz_background_solve_batch(double* params, int n, double rtol, double atol, int max_knots, double* t,
                               double* u, double* du, int* n_knots, int* n_rejected, int* status, float* kernel_ms);

/* Background right-hand side at one state (host; used by tests of the background stage). */
int sbz_background_rhs(const double* params, const double* u, double* du);

/* ---- perturbation solver ------------------------------------------------------------------ */

typedef struct sbz_model sbz_model;
typedef struct sbz_plan sbz_plan;

/* Model descriptor.  n: number of perturbation unknowns (layout: Φ, photon F_0..F_l, G_0..G_l,
 * massless ν F_0..F_l, massive ν ψ_i,0..l for each momentum bin, δ_c, θ_c, δ_b, θ_b[, δ_X, θ_X]);
 * nnz_lu: slots of the static sparse LU pattern (the algorithmic flop count's); cache_path:
 * code-object file loaded if present and written after a hiprtc compile otherwise. */
typedef struct {
    int n;
    int nnz_lu;
    const char* cache_path;
} sbz_model_desc;

/* What a plan stores at its saveat times. */
#define SBZ_SAVE_STATES 1   /* full perturbation states [n_inst][n_save][n] */
#define SBZ_SAVE_SOURCES 2  /* CMB sources S_T, S_E, S_ψ [n_inst][n_save][3] (cosmologies.jl:99-105) */
#define SBZ_SAVEAT_PER_SET 4 /* saveat is [n_sets][n_save]: one time grid per parameter set (each set's
                                own τ0), as source_grid with per-cosmology τs needs; no reference
                                counterpart (the reference solves one cosmology per call) */

/* Kernel variants. */
#define SBZ_ALGO_WAVE 1  /* one 64-lane wavefront per mode: chains in lanes, row-parallel border LU (default) */
#define SBZ_ALGO_WAVE2 3 /* SBZ_ALGO_WAVE compiled for two resident wavefronts per SIMD (throughput) */
#define SBZ_ALGO_WAVE4 4 /* four modes per wavefront, 16 lanes each (throughput); models whose border and
                            chains fit 16 lanes (sbz_model_capacity reports 0 otherwise), no saves */
#define SBZ_ALGO_WAVE_ALTLU 5 /* SBZ_ALGO_WAVE with the other layout of the border factors in LDS: the
                                 model's default keeps strictly-lower and strictly-upper rows apart
                                 (unmasked solves) unless that costs resident workgroups, in which case
                                 it keeps one combined row per lane (masked solves, half the LDS) */
#define SBZ_ALGO_WAVE2H 6 /* two modes per wavefront, 32 lanes each (throughput); models without the
                             four-mode kernel whose border and chains fit 32 lanes
                             (sbz_model_capacity reports 0 otherwise), no saves */

/* Solver options (solvept keywords: reltol, abstol, maxiters; src/solve.jl:543).  lanes: number of
 * resident wavefronts of the persistent grid (0: as many as fit the device).  save: SBZ_SAVE_* flags
 * (0 with saveat given: states).  algo: SBZ_ALGO_* (0: default). */
typedef struct {
    double rtol;
    double atol;
    long long maxiters;
    int lanes;
    int save;
    int algo;
} sbz_opts;

/* Compile emitted HIP source with hiprtc for gfx950 into a code-object file. */
int sbz_compile(const char* hip_src, const char* out_path);

/* Create a model handle from the emitted source (sbz_model_create in place of MTK's generated
 * f!/jac!, src/solve.jl:215-227).  The code object comes from desc->cache_path when it exists,
 * else from hiprtc.  Fails with SBZ_ERR_NODEV when no GPU is visible. */
int sbz_model_create(const char* hip_src, const sbz_model_desc* desc, sbz_model** out);
int sbz_model_destroy(sbz_model* m);

/* Upload the background spline of parameter set `set`: knot times t[n], unknowns u[n][5] and
 * their derivatives du[n][5] (spline(bgsol), src/utils.jl:118-127) and the SBZ_NPAR parameters
 * including τ0 and κ0 (setuppt copies them into the perturbation problem, src/solve.jl:496-513). */
int sbz_bg_upload(sbz_model* m, int set, const double* t, const double* u, const double* du, int n_knots,
                  const double* params);

/* Solve plan: instances (set, k, τ_start) resident on the device.  Replaces the per-mode task
 * loop of solvept (src/solve.jl:543-569; ptprobgen clamps τ_start, :524-531).  n_save output times
 * shared by all instances (saveat). */
int sbz_plan_create(sbz_model* m, const int* inst_set, const double* inst_k, const double* inst_t0, int n_inst,
                    const sbz_opts* opts, const double* saveat, int n_save, sbz_plan** out);
/* Launch the persistent solver kernel and wait; *kernel_ms receives its duration (HIP events on
 * the model's stream). */
int sbz_plan_run(sbz_plan* p, float* kernel_ms);
/* sbz_plan_run in two halves: sbz_plan_launch enqueues the run on the model's stream and returns at
 * once; sbz_plan_wait waits for it and gives its duration.  Plans of different models (each model
 * has its own stream) launched one after the other run concurrently on the device, the next
 * kernel's wavefronts filling the CUs that the previous kernel's tail frees.  Replaces the same
 * per-mode loop as sbz_plan_run (src/solve.jl:543-569). */
int sbz_plan_launch(sbz_plan* p);
int sbz_plan_wait(sbz_plan* p, float* kernel_ms);
/* The same run as one HIP graph launch: on the first call the run's kernels (NaN fill of the
 * save/source outputs, source background table, queue reset, solve, and the P(k) reduction of
 * sbz_plan_pk) are captured from the model's stream into a graph, which every call then replays;
 * *graph_ms receives the duration of the whole graph.  sbz_plan_pk after it copies the P(k) the
 * graph computed.  Replaces the same per-mode loop of solvept (src/solve.jl:543-569) as
 * sbz_plan_run; the graph form has no reference counterpart (it removes the per-launch host
 * overhead of repeated runs of one plan). */
int sbz_plan_run_graph(sbz_plan* p, float* graph_ms);
/* Copy results to host: final states [n_inst][n], saved states [n_inst][n_save][n] (may be
 * null), per-mode return codes, and stats [n_inst][6] = accepted steps, rejected steps, f evals,
 * J evals, LU factorizations, linear solves (DEStats, src/solve.jl:885-891). */
int sbz_plan_fetch(sbz_plan* p, double* out_u, double* out_save, int* retcode, long long* stats);
/* CMB source functions at the saveat times, [n_inst][n_save][3] = (S_T, S_E, S_ψ): the
 * source_grid of src/observables/fourier.jl:267-281 with Ss = [ST, SE, Sψ]. */
int sbz_plan_fetch_sources(sbz_plan* p, double* out_src);
/* Matter power spectrum P(k, τ0) = P0(k) Δ_m² per instance from the final states
 * (spectrum_matter, src/observables/fourier.jl:79-101). */
int sbz_plan_pk(sbz_plan* p, double* pk);
/* P(k, τ) = P0(k) Δ_S(τ)² for species combinations S given as bit masks (1: CDM, 2: baryons,
 * 4: massive ν; a single species is its own Δ, several are ρ-weighted), at τ0 from the final
 * states (at_saves = 0) or at every saveat time from the saved states (at_saves = 1, needs
 * SBZ_SAVE_STATES): spectrum_matter(modes, sol, k, τ) (src/observables/fourier.jl:39-52,90-97).
 * out [n_masks][n_t][n_inst] with n_t = 1 or n_save.  The massive-ν bit needs a model with them. */
int sbz_plan_pk_modes(sbz_plan* p, const int* masks, int n_masks, int at_saves, double* out);
int sbz_plan_destroy(sbz_plan* p);
/* Resident capacity of a kernel variant on the model's device: modes that run at once (one per
 * wavefront for SBZ_ALGO_WAVE and SBZ_ALGO_WAVE2, four per wavefront for SBZ_ALGO_WAVE4, two for
 * SBZ_ALGO_WAVE2H), and 0 for SBZ_ALGO_WAVE4 / SBZ_ALGO_WAVE2H when the model has no such kernel.  No reference counterpart: the
 * autotuner (symboltz/autotune.py) uses it to decide whether higher occupancy can pay off. */
int sbz_model_capacity(sbz_model* m, int algo, int* capacity);

/* Tangent plans: forward-mode sensitivities of the perturbation solve.  Replaces ForwardDiff
 * through solve/spectrum_matter (the reference's AD tests, test/runtests.jl:363-377, differentiate
 * log P(k) with respect to log θ).  Each instance integrates u and s = ∂u/∂log θ together: the
 * augmented linear system u' = J u, s' = J s + J_θ u on the primal's step sequence, one W
 * factorization per step serving both blocks.  inst_setp / inst_setm name the uploaded parameter
 * sets of log θ + h and log θ - h (backgrounds the caller solved at a tight tolerance) and
 * inst_tscale = 1/(2h): the θ-dependence of the background enters as the central difference of
 * those two sets at equal τ.  States only (no saveat); fails with SBZ_ERR_ARG for models whose
 * tangent kernel does not fit a workgroup's LDS.  sbz_plan_run runs it. */
int sbz_plan_create_tangent(sbz_model* m, const int* inst_set, const int* inst_setp, const int* inst_setm,
                            const double* inst_tscale, const double* inst_k, const double* inst_t0, int n_inst,
                            const sbz_opts* opts, sbz_plan** out);
/* Tangents s = ∂u/∂log θ at τ0 [n_inst][n] and f(τ0, u(τ0)) [n_inst][n] (either may be NULL). */
int sbz_plan_fetch_tangent(sbz_plan* p, double* out_s, double* out_fend);
/* P(k, τ0) [n_inst] and d log P / d log θ [n_inst] of a tangent plan, including the shift of τ0
 * with θ (spectrum_matter differentiated, src/observables/fourier.jl:79-101). */
int sbz_plan_pk_tangent(sbz_plan* p, double* pk, double* dlogpk);

/* One-shot: create a plan, run, fetch, destroy. */
int sbz_solve_batch(sbz_model* m, const int* inst_set, const double* inst_k, const double* inst_t0, int n_inst,
                    const sbz_opts* opts, const double* saveat, int n_save, double* out_u, double* out_save,
                    int* retcode, long long* stats);

/* sbz_solve_batch without any device allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc).  Same arguments and results as sbz_solve_batch, with these limits:
 *  - every device buffer of the call (background tables of all uploaded sets, instances, saveat,
 *    outputs, work counter) must fit SBZ_NOALLOC_SCRATCH_BYTES, a device array the library holds on
 *    each GPU it runs on until the process ends; larger calls fail with SBZ_ERR_ARG;
 *  - wavefront kernels only (opts->algo 0, SBZ_ALGO_WAVE or SBZ_ALGO_WAVE2), states only (no
 *    SBZ_SAVE_SOURCES, no SBZ_SAVEAT_PER_SET);
 *  - instances are processed in the given order (the plan entry points sort them longest first:
 *    pass them by descending k for the same load balance);
 *  - all calls on one GPU share its array: calls that overlap on one GPU race. */
#define SBZ_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)
int sbz_solve_batch_noalloc(sbz_model* m, const int* inst_set, const double* inst_k, const double* inst_t0,
                            int n_inst, const sbz_opts* opts, const double* saveat, int n_save, double* out_u,
                            double* out_save, int* retcode, long long* stats);

/* ---- mixed-model launches ---------------------------------------------------------------- */

/* Several models in one persistent launch.  Replaces the per-model solve loops of a stream of
 * requests over several cosmological models (solve(prob, ks) per CosmologyProblem,
 * src/solve.jl:543-569, called once per model): the solves of every model run in one launch
 * whose wavefronts pull (model, instance) entries from one queue ordered longest first across the
 * models, so the whole batch has one tail.  No reference counterpart beyond that loop.
 * hip_src is the mixed-model source (symboltz.multi: each model's device code in its own namespace
 * and the kernel sbz_multi_kernel), loaded from cache_path when that file exists, else compiled
 * with hiprtc and written there; models[0..n_models) are the model handles of its families, in its
 * order (at most 8, one device; their state sizes are checked against the code object).
 * Calls on one handle must not overlap. */
typedef struct sbz_multi sbz_multi;
int sbz_multi_create(const char* hip_src, const char* cache_path, sbz_model* const* models, int n_models,
                     sbz_multi** out);
/* Run plans[f] of models[f] for every family f together (one plan per model, in order; a plan may
 * hold no instances).  Plans save no states or sources (final states only) and are not tangent
 * plans.  Results land in each plan's buffers: sbz_plan_fetch and sbz_plan_pk read them as after
 * sbz_plan_run.  *kernel_ms receives the launch's duration.  With sbz_multi_attach_rows, a family
 * whose body integrates several modes per wavefront (sbz_multi_rows) and whose plan holds at least
 * twice the resident wavefronts of the attached kernel runs in that body (its wavefronts drain the
 * family's instances longest first), in a launch of the attached kernel; every other instance runs
 * in its model's one-mode body, and a run without such a family launches the one-mode kernel. */
int sbz_multi_run(sbz_multi* mm, sbz_plan* const* plans, int n_plans, float* kernel_ms);
/* Resident wavefronts of the one-mode mixed-model kernel (modes that run at once). */
int sbz_multi_capacity(sbz_multi* mm, int* capacity);
/* Attach the kernel with the several-modes bodies, sbz_multi_kernel_rows, from its own code object
 * (source and cache as sbz_multi_create; the same models in the same order).  It is a code object
 * of its own so that the one-mode kernel, which serves the latency-bound launches, keeps the
 * registers of its bodies alone.  No reference counterpart. */
int sbz_multi_attach_rows(sbz_multi* mm, const char* hip_src, const char* cache_path);
/* Modes per wavefront of each family's several-modes body: rows[f] = 4 (16-lane rows), 2 (32-lane
 * rows) or 1 (none, or no such kernel attached), f < n_models.  No reference counterpart. */
int sbz_multi_rows(sbz_multi* mm, int* rows);
int sbz_multi_destroy(sbz_multi* mm);

/* ---- line-of-sight integration ----------------------------------------------------------- */

/* j_l(x) and j_l'(x) = l/(2l+1) j_{l-1} - (l+1)/(2l+1) j_{l+1} on x_i = i dx (i < n_x - 1) plus one
 * duplicate padding row, computed on the device and returned in y, dy [n_x][n_l]: the table of
 * SphericalBesselCache(ls) (src/observables/angular.jl:9-48).  ls strictly ascending. */
int sbz_bessel_table(const int* ls, int n_l, int n_x, double dx, double* y, double* dy);

/* A C_l context for one ascending l list: the Bessel table of sbz_bessel_table is built on the
 * device at creation and kept there; sbz_cmb_compute then runs the stage below for one cosmology
 * per call on the context's own stream (work buffers grow as needed and are reused).  In
 * sbz_cmb_compute kernel_ms[0] is 0 (the table already exists); sbz_cmb_spectra is the one-shot
 * create + compute + destroy.  Calls on one context must not overlap. */
typedef struct sbz_cmb sbz_cmb;
int sbz_cmb_create(const int* ls, int n_l, int n_x, double dx, sbz_cmb** out);
int sbz_cmb_compute(sbz_cmb* h, const double* S_coarse, int n_c, int n_s, const double* taus, int n_t,
                    const double* y_c, const double* w_c, const double* ks, const double* y_f, const double* wk,
                    int n_f, const int* kind, const int* l_limber, const int* modes, int n_m, double* out_cl,
                    double* out_theta, float* kernel_ms);
int sbz_cmb_destroy(sbz_cmb* h);

/* The C_l stage of spectrum_cmb (src/observables/angular.jl:260-341) after the source solve, on
 * the device from the coarse sources to C_l:
 *  - source_kinterp with a ChebyshevInterpolator (src/observables/fourier.jl:232-242,520-532):
 *    S_coarse [n_c][n_t][n_s] on the nodes y_c (barycentric weights w_c) to the fine wavenumbers
 *    ks [n_f] whose transformed values are y_f;
 *  - the Bessel table of the ascending ls [n_l] on x_i = i dx, i < n_x - 1 (angular.jl:9-48);
 *  - los_integrate of every source over τs [n_t] (angular.jl:109-185), the Limber value for
 *    l ≥ l_limber[s] (angular.jl:150-177), then Θ /= k for kind 0 (temperature) and
 *    Θ *= √((l+2)(l+1)l(l-1))/k² for kind 1 (E polarization); kind 2 (lensing) is left as is
 *    (angular.jl:300-311);
 *  - C_l[m][l] = Σ_k wk[k] Θ_A(k, l) Θ_B(k, l) for each pair (A, B) = modes[2m], modes[2m+1],
 *    where the caller's wk holds the k-quadrature weight times (2/π) k² P0(k) (angular.jl:198-223).
 * The sources are Ss = [k S_T, k² S_E, S_ψ] as spectrum_cmb builds them (angular.jl:296-299).
 * out_theta [n_s][n_f][n_l] may be NULL.  kernel_ms[5]: Bessel table, k interpolation, LOS,
 * Limber and factors, k integral. */
int sbz_cmb_spectra(const double* S_coarse, int n_c, int n_s, const double* taus, int n_t, const double* y_c,
                    const double* w_c, const double* ks, const double* y_f, const double* wk, int n_f, const int* ls,
                    int n_l, int n_x, double dx, const int* kind, const int* l_limber, const int* modes, int n_m,
                    double* out_cl, double* out_theta, float* kernel_ms);

/* ---- surrogate MLP (no reference counterpart) -------------------------------------------- */

/* A multilayer perceptron for the learned surrogate of log P(k) (symboltz/surrogate.py): widths
 * w[0..n_layers], GELU after every hidden layer, BF16 weights and activations in MFMA GEMMs with
 * FP32 accumulation, FP32 master weights and Adam.  max_batch bounds every batch; weights are
 * He-uniform from `seed`, biases zero.  Parameters are laid out as all weight matrices ([w_l][w_l+1]
 * row-major, layer by layer) followed by all biases. */
typedef struct sbz_mlp sbz_mlp;
int sbz_mlp_create(int n_layers, const int* widths, int max_batch, unsigned long long seed, sbz_mlp** out);
int sbz_mlp_destroy(sbz_mlp* m);
int sbz_mlp_n_params(sbz_mlp* m, long long* n);
int sbz_mlp_get_params(sbz_mlp* m, float* params);
int sbz_mlp_set_params(sbz_mlp* m, const float* params);
/* Mean-squared-error loss and its gradient on one batch (host x [B][w0], t [B][wL]); no update. */
int sbz_mlp_grad(sbz_mlp* m, const float* x, const float* t, int batch, float* grad, float* loss);
/* Training data kept on the device (host x [n][w0], t [n][wL]). */
int sbz_mlp_set_data(sbz_mlp* m, const float* x, const float* t, long long n);
/* n_batches Adam steps (lr; β = 0.9, 0.999; ε = 1e-8) over consecutive batches of the row list
 * idx[n_batches*batch]; loss[n_batches] (may be null) receives each batch's loss. */
int sbz_mlp_train(sbz_mlp* m, const int* idx, int n_batches, int batch, float lr, float* loss, float* kernel_ms);
/* Inference y [n][wL] = net(x [n][w0]) on host buffers, in batches of max_batch. */
int sbz_mlp_forward(sbz_mlp* m, const float* x, long long n, float* y, float* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif
