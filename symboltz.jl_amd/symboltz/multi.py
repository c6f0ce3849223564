"""Mixed-model launches: plans of several models in one persistent kernel (sbz_multi_*).

`multi_source(models)` writes one HIP source that holds every model's device code in its own
namespace (the model's emitted functions spliced into the generic kernel, exactly the source of its
own code object) and one kernel, `sbz_multi_kernel`.  Its wavefronts pull entries (model, instance)
from a single queue ordered by estimated cost across the models, and integrate the entry's
instance with that model's Rodas5P body (one mode per wavefront) on that model's plan buffers, so a
stream of requests with mixed state sizes is balanced over the whole device in one launch with one
tail, instead of one launch (and one tail) per model.

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


def multi_source(models):
    """One self-contained HIP source for the given models (see the module docstring)."""
    if not 1 <= len(models) <= MAX_FAMILIES:
        raise ValueError(f"a mixed-model code object holds 1 to {MAX_FAMILIES} models")
    out = ["// Mixed-model code object: " + "; ".join(str(M) for M in models),
           "#ifndef __HIPCC_RTC__", "#include <hip/hip_runtime.h>", "#endif"]
    nf = len(models)
    abi = set(_macro_names((INCLUDE / "sbz.h").read_text()) + _macro_names((INCLUDE / "sbz_params.h").read_text()))
    for f, M in enumerate(models):
        block = model_source(M).full_src
        out += [f"namespace sbzf{f} {{", "#define SBZ_MULTI_FAMILY 1", block,
                "constexpr int SBZ_MULTI_N = SBZ_N;", f"}}  // namespace sbzf{f}"]
        # every macro of the block is the block's own, so the next model defines its own values,
        # except the C ABI's constants: the ABI header (return codes, parameter slots, and the
        # entry-point declarations, which may be declared once only) stays included
        out += [f"#undef {nm}" for nm in _macro_names(block) if nm not in abi] + ["#undef SBZ_MULTI_FAMILY"]
    lds = "0"
    for f in range(nf):
        lds = f"sbz_cmax(sizeof(sbzf{f}::WaveLDS), {lds})"
    ns = ", ".join(f"sbzf{f}::SBZ_MULTI_N" for f in range(nf))
    out.append(f"""
// ---- the mixed-model kernel ------------------------------------------------------------------
// The kernel argument: every model's solve arguments (the struct of its plan launch, identical in
// layout in every namespace), then one queue of (model, instance) entries, longest first.
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

extern "C" __global__ void __launch_bounds__(64) sbz_multi_kernel(SbzMultiArgs a) {{
    // the LDS of whichever model the wavefront runs: each family lays out its own WaveLDS here
    __shared__ __align__(16) unsigned char lds[{lds}];
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
        switch (f) {{""")
    for f in range(nf):
        out.append(f"""        case {f}: {{
            sbzf{f}::WaveLDS& L = *reinterpret_cast<sbzf{f}::WaveLDS*>(lds);
            const sbzf{f}::WaveCtx c = sbzf{f}::sbz_wave_ctx(threadIdx.x);
            sbzf{f}::sbz_wave_mode<true>(reinterpret_cast<const sbzf{f}::SbzSolveArgs&>(a.fam[{f}]), L, c, inst);
            break;
        }}""")
    out.append("""        default:
            break;
        }
    }
}""")
    return "\n".join(out) + "\n"


def multi_hash(models):
    return hashlib.sha1(multi_source(models).encode()).hexdigest()[:16]


class MultiModel:
    """The models of `gms` (GpuModel handles, at most 8) in one code object."""

    def __init__(self, gms):
        self.gms = list(gms)
        self.L = _lib.lib()
        L = self.L
        vp = ctypes.c_void_p
        L.sbz_multi_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp), ctypes.c_int,
                                       ctypes.POINTER(vp)]
        L.sbz_multi_run.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.sbz_multi_capacity.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.sbz_multi_destroy.argtypes = [vp]
        src = multi_source([g.M for g in self.gms])
        KCACHE.mkdir(exist_ok=True)
        self.cache_path = KCACHE / f"multi_{hashlib.sha1(src.encode()).hexdigest()[:16]}.hsaco"
        hs = (vp * len(self.gms))(*[g.h.value for g in self.gms])
        h = vp()
        _lib.check(L.sbz_multi_create(src.encode(), str(self.cache_path).encode(), hs, len(self.gms), ctypes.byref(h)),
                   "sbz_multi_create")
        self.h = h

    def capacity(self):
        c = ctypes.c_int()
        _lib.check(self.L.sbz_multi_capacity(self.h, ctypes.byref(c)), "sbz_multi_capacity")
        return c.value

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
