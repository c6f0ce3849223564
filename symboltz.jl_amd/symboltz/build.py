"""Build steps: the host+runtime shared library and ahead-of-time code objects of the models.

`build_library()` compiles csrc/*.cpp into lib/libsbz_hip.so with hipcc for gfx950.
`build_models()` emits the HIP source of the standard models and compiles each into a gfx950
code object in _kcache/ (named by the hash of its source), so a GPU host loads them without
invoking hiprtc; any other model is compiled by hiprtc on first use and cached the same way.
"""
import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
KCACHE = PKG / "_kcache"
INCLUDE = PKG.parent / "include"

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops"]


def build_library(verbose=False):
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / "libsbz_hip.so"
    srcs = [str(CSRC / f) for f in ("sbz_api.cpp", "sbz_background.cpp", "sbz_solver.cpp", "sbz_cmb.hip",
                                    "sbz_bgsolve.hip",
                                    "sbz_mlp.hip", "sbz_noalloc.hip", "sbz_multi.cpp")]
    cmd = ["hipcc", "-fPIC", "-shared", "-I", str(INCLUDE)] + HIP_FLAGS + srcs + ["-lhiprtc", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libsbz_hip.so build failed:\n{r.stderr[-4000:]}")
    if verbose:
        print("built", out)
    return out


def compile_code_object(src, path):
    """hipcc --genco of one emitted source into a code object."""
    path = Path(path)
    path.parent.mkdir(exist_ok=True)
    tmp = path.with_suffix(".hip")
    tmp.write_text(src)
    cmd = ["hipcc", "--genco"] + HIP_FLAGS + [str(tmp), "-o", str(path)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    tmp.unlink()
    if r.returncode != 0:
        raise RuntimeError(f"code object build failed for {path.name}:\n{r.stderr[-4000:]}")
    return path


def build_probe(M):
    """Host build of a model's emitted functions (test probe, see SBZ_HOST_PROBE in the kernel)."""
    from .solver import model_source
    em = model_source(M)
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / f"probe_{em.full_hash}.so"
    if out.exists():
        return out
    tmp = LIBDIR / f"probe_{em.full_hash}.hip"
    tmp.write_text(em.full_src)
    cmd = ["hipcc", "-fPIC", "-shared", "-DSBZ_HOST_PROBE", "-DSBZ_DEV=__host__ __device__ inline", "-O2"] + \
        HIP_FLAGS[:1] + ["-std=c++17", str(tmp), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    tmp.unlink()
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed:\n{r.stderr[-4000:]}")
    return out


def standard_models():
    """Models used by the tests, the smoke check and the benchmark."""
    from .models import LCDM, w0waCDM
    return [
        LCDM(lmax=5),                                     # reference test fixture (runtests.jl:14-15)
        LCDM(lmax=10, h=False),                           # ΛCDM with massless ν only, lmax 10
        LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4),     # photons/ν lmax 16, one massive ν (4 bins, lmax 10)
        w0waCDM(lmax=6),                                  # w0wa dark-energy fluid with massive ν
        LCDM(lmax=6),                                     # the surrogate's model (3 massive ν)
        LCDM(lmax=32),                                    # the reference's high-lmax test, n = 236 (runtests.jl:808-812)
    ]


def standard_multis():
    """Model lists compiled ahead of time into mixed-model code objects (symboltz.multi): the mixed
    stream's five families, and the three models of the problem-level test."""
    from .models import LCDM, w0waCDM
    from .workloads import mixed_families
    return [mixed_families(), [LCDM(lmax=10, h=False), w0waCDM(lmax=6), LCDM(lmax=6)]]


def build_models(models=None, jobs=4, verbose=False, with_multi=None):
    from concurrent.futures import ThreadPoolExecutor
    from .solver import model_source
    standard = models is None
    if with_multi is None:
        with_multi = standard
    models = standard_models() if standard else models
    KCACHE.mkdir(exist_ok=True)
    work = []
    for M in models:
        em = model_source(M)
        path = KCACHE / f"{em.full_hash}.hsaco"
        if not path.exists():
            work.append((em.full_src, path))
    multis = []
    if with_multi:   # the mixed-model code objects of the standard model lists (both kernels)
        from .multi import multi_cache_path, multi_source
        for fams in standard_multis():
            for rk in (False, True):
                src = multi_source(fams, rows_kernel=rk)
                path = multi_cache_path(src)
                multis.append(path.name)
                if not path.exists():
                    work.append((src, path))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(work) or 1))) as ex:
        for p in ex.map(lambda w: compile_code_object(*w), work):
            if verbose:
                print("built", p)
    keep = {f"{model_source(M).full_hash}.hsaco" for M in models}
    keep |= {f"src_{getattr(model_source(M), 'tag', '')}.json" for M in models}
    keep |= set(multis)
    for f in (KCACHE.iterdir() if standard else []):   # drop code objects and sources of older kernel versions
        if f.name not in keep and (f.suffix == ".hsaco" or f.name.startswith("src_")):
            f.unlink()
    return [KCACHE / f"{model_source(M).full_hash}.hsaco" for M in models]
