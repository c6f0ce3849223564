"""Mixed-model launches: plans of several models in one persistent kernel (sbz_multi_*).

`multi_source(models)` writes one HIP source that holds every model's device code in its own
namespace (the model's emitted functions spliced into the generic kernel, exactly the source of its
own code object) and one kernel, `sbz_multi_kernel`.  Its wavefronts pull entries from a single
queue ordered by estimated cost across the models, so a stream of requests with mixed state sizes
is balanced over the whole device in one launch with one tail, instead of one launch (and one
tail) per model.  An entry either names one instance, which the wavefront integrates with its
model's one-mode Rodas5P body, or, for a model with a several-modes body (four modes per
wavefront when its border and chains fit 16 lanes, two when they fit 32) and a plan large enough
to be throughput-bound, is a ticket: the wavefront runs that body, whose rows pull the model's
instances from the model's own queue (longest first) until it is empty.  Either way the results
land in the model's plan buffers.

`MultiModel(gpu_models)` loads (or compiles and caches) the code object; `run(plans)` takes one
final-state plan per model, in the same order, and runs them together.
"""
import ctypes
import hashlib
import re

from . import _lib
from .build import INCLUDE, KCACHE
from .solver import model_source

MAX_FAMILIES = 8   # kMultiMaxF of csrc/sbz_multi.cpp


def _macro_names(text):
    return sorted(set(re.findall(r"^[ \t]*#[ \t]*define[ \t]+(\w+)", text, re.M)))


def modes_per_wave(full_src):
    """Modes per wavefront of a model's body in the mixed kernel: 4 (SBZ_WAVE4_OK), 2
    (SBZ_WAVE2H_OK) or 1."""
    if re.search(r"^#define SBZ_WAVE4_OK 1$", full_src, re.M):
        return 4
    if re.search(r"^#define SBZ_WAVE2H_OK 1$", full_src, re.M):
        return 2
    return 1


def multi_source(models, rows_kernel=False):
    """One self-contained HIP source for the given models (see the module docstring): with the
    one-mode kernel sbz_multi_kernel, or (rows_kernel) with sbz_multi_kernel_rows, which also takes
    tickets.  The two are separate code objects, so each kernel's registers are its bodies' own."""
    if not 1 <= len(models) <= MAX_FAMILIES:
        raise ValueError(f"a mixed-model code object holds 1 to {MAX_FAMILIES} models")
    rows = [modes_per_wave(model_source(M).full_src) for M in models]
    out = ["// Mixed-model code object: " + "; ".join(str(M) for M in models),
           "#ifndef __HIPCC_RTC__", "#include <hip/hip_runtime.h>", "#endif"]
    nf = len(models)
    abi = set(_macro_names((INCLUDE / "sbz.h").read_text()) + _macro_names((INCLUDE / "sbz_params.h").read_text()))
    for f, M in enumerate(models):
        block = model_source(M).full_src
        pre = ["#define SBZ_STAGE_PRELOAD 0"] if rows_kernel else []   # the kernel short of registers
        out += [f"namespace sbzf{f} {{", "#define SBZ_MULTI_FAMILY 1"] + pre + [block,
                "constexpr int SBZ_MULTI_N = SBZ_N;", f"}}  // namespace sbzf{f}"]
        # every macro of the block is the block's own, so the next model defines its own values,
        # except the C ABI's constants: the ABI header (return codes, parameter slots, and the
        # entry-point declarations, which may be declared once only) stays included
        out += [f"#undef {nm}" for nm in _macro_names(block) if nm not in abi] + ["#undef SBZ_MULTI_FAMILY"]
    lds1 = "0"
    for f in range(nf):
        lds1 = f"sbz_cmax(sizeof(sbzf{f}::WaveLDS), {lds1})"
    lds_rows = lds1
    for f in range(nf):
        if rows[f] > 1:
            lds_rows = f"sbz_cmax({rows[f]} * sizeof(sbzf{f}::WaveRow), {lds_rows})"
    ns = ", ".join(f"sbzf{f}::SBZ_MULTI_N" for f in range(nf))
    rw = ", ".join(str(r) for r in rows)
    out.append(f"""
// ---- the mixed-model kernels -----------------------------------------------------------------
// The kernel argument: every model's solve arguments (the struct of its plan launch, identical in
// layout in every namespace), then one queue of (model, instance) entries, longest first; the
// instance is -1 for a ticket into a several-modes body (sbz_multi_kernel_rows only).
struct SbzMultiArgs {{
    sbzf0::SbzSolveArgs fam[{MAX_FAMILIES}];
    const int* __restrict__ mfam;
    const int* __restrict__ minst;
    int n;
    unsigned int* queue;
}};
constexpr unsigned long sbz_cmax(unsigned long a, unsigned long b) {{ return a > b ? a : b; }}
// state size of each family (0: unused slot); the host checks it against the models it was given
__constant__ int sbz_multi_n[{MAX_FAMILIES}] = {{{ns}}};
// modes per wavefront of each family's several-modes body (1: none; 2, 4: tickets allowed)
__constant__ int sbz_multi_modes[{MAX_FAMILIES}] = {{{rw}}};""")

    def kernel(name, with_rows):
        k = [f"""
// {"one-mode entries and tickets into the several-modes bodies" if with_rows else
    "one-mode entries only (fewer registers: the latency-bound launches)"}
extern "C" __global__ void __launch_bounds__(64) {name}(SbzMultiArgs a) {{
    // the LDS of whichever model the wavefront runs: each family lays out its own rows here
    __shared__ __align__(16) unsigned char lds[{lds_rows if with_rows else lds1}];
    __shared__ int pulled;
    for (;;) {{
        if (threadIdx.x == 0) {{
            const unsigned int g = atomicAdd(a.queue, 1u);
            pulled = (g < (unsigned int)a.n) ? (int)g : -1;
        }}
        __syncthreads();
        const int g = __builtin_amdgcn_readfirstlane(pulled);
        __syncthreads();
        if (g < 0) break;
        const int f = __builtin_amdgcn_readfirstlane(a.mfam[g]);
        const int inst = __builtin_amdgcn_readfirstlane(a.minst[g]);
        switch (f) {{"""]
        for f in range(nf):
            fa = f"reinterpret_cast<const sbzf{f}::SbzSolveArgs&>(a.fam[{f}])"
            one = (f"sbzf{f}::WaveLDS& L = *reinterpret_cast<sbzf{f}::WaveLDS*>(lds);\n"
                   f"                const sbzf{f}::WaveCtx c = sbzf{f}::sbz_wave_ctx(threadIdx.x);\n"
                   f"                sbzf{f}::sbz_wave_mode<true>({fa}, L, c, inst);")
            if with_rows and rows[f] > 1:   # an instance, or a ticket (-1): rows drain the family's queue
                k.append(f"""        case {f}:
            if (inst >= 0) {{
                {one}
            }} else {{
                sbzf{f}::sbz_waverow_run<true, {64 // rows[f]}>({fa}, reinterpret_cast<sbzf{f}::WaveRow*>(lds));
            }}
            break;""")
            else:
                k.append(f"""        case {f}: {{
                {one}
            break;
        }}""")
        k.append("""        default:
            break;
        }
    }
}""")
        return k

    out += kernel("sbz_multi_kernel_rows", True) if rows_kernel else kernel("sbz_multi_kernel", False)
    return "\n".join(out) + "\n"


def multi_cache_path(src):
    return KCACHE / f"multi_{hashlib.sha1(src.encode()).hexdigest()[:16]}.hsaco"


class MultiModel:
    """The models of `gms` (GpuModel handles, at most 8) in one code object, plus the code object
    of their several-modes kernel when some model has a several-modes body (and `rows`)."""

    def __init__(self, gms, rows=True):
        self.gms = list(gms)
        self.L = _lib.lib()
        L = self.L
        vp = ctypes.c_void_p
        L.sbz_multi_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp), ctypes.c_int,
                                       ctypes.POINTER(vp)]
        L.sbz_multi_run.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.sbz_multi_capacity.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.sbz_multi_rows.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.sbz_multi_attach_rows.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
        L.sbz_multi_destroy.argtypes = [vp]
        models = [g.M for g in self.gms]
        src = multi_source(models)
        KCACHE.mkdir(exist_ok=True)
        self.cache_path = multi_cache_path(src)
        hs = (vp * len(self.gms))(*[g.h.value for g in self.gms])
        h = vp()
        _lib.check(L.sbz_multi_create(src.encode(), str(self.cache_path).encode(), hs, len(self.gms), ctypes.byref(h)),
                   "sbz_multi_create")
        self.h = h
        if rows and any(modes_per_wave(model_source(M).full_src) > 1 for M in models):
            src_r = multi_source(models, rows_kernel=True)
            _lib.check(L.sbz_multi_attach_rows(self.h, src_r.encode(), str(multi_cache_path(src_r)).encode()),
                       "sbz_multi_attach_rows")

    def capacity(self):
        c = ctypes.c_int()
        _lib.check(self.L.sbz_multi_capacity(self.h, ctypes.byref(c)), "sbz_multi_capacity")
        return c.value

    def rows(self):
        """Modes per wavefront of each family's several-modes body (1: none)."""
        r = (ctypes.c_int * len(self.gms))()
        _lib.check(self.L.sbz_multi_rows(self.h, r), "sbz_multi_rows")
        return list(r)

    def run(self, plans):
        """Run one final-state plan per model (in the models' order) in one launch; kernel ms."""
        ps = (ctypes.c_void_p * len(plans))(*[p.h.value for p in plans])
        ms = ctypes.c_float()
        _lib.check(self.L.sbz_multi_run(self.h, ps, len(plans), ctypes.byref(ms)), "sbz_multi_run")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self.L.sbz_multi_destroy(self.h)
            self.h = None
