"""The GPU perturbation solver behind the C ABI (include/sbz.h).

`GpuModel(M)` owns a model handle: the emitted HIP source of the model's perturbation system,
compiled (or loaded from the code-object cache) and bound to a HIP stream.  Background tables are
uploaded per parameter set; `Plan` holds a batch of (set, k, τ_start) instances resident on the
device, runs the persistent Rodas5P kernel and fetches final states, saved states, return codes,
statistics and P(k).
"""
import ctypes
import hashlib
import json
from pathlib import Path

import numpy as np

from . import _lib
from .emitter import kernel_source
from .build import KCACHE

_src_cache = {}


def model_source(M):
    """Emitted source of a model, memoised in-process and on disk (emission costs seconds)."""
    key = M.key()
    if key in _src_cache:
        return _src_cache[key]
    kern = (Path(__file__).resolve().parent.parent / "csrc" / "sbz_kernel.hip").read_text()
    emit = (Path(__file__).resolve().parent / "emitter.py").read_text()
    tag = hashlib.sha1((repr(key) + kern + emit + _deps_text()).encode()).hexdigest()[:16]
    meta = KCACHE / f"src_{tag}.json"
    if meta.exists():
        d = json.loads(meta.read_text())
        em = _Stored(d)
        em.tag = tag
    else:
        em = kernel_source(M)
        KCACHE.mkdir(exist_ok=True)
        meta.write_text(json.dumps(dict(full_src=em.full_src, full_hash=em.full_hash, n=em.n, nnz_LU=em.nnz_LU,
                                        nnz_J=em.nnz_J, flops_step=em.flops_step, ops={k: list(v) for k, v in em.ops.items()})))
    em.tag = tag
    _src_cache[key] = em
    return em


def _deps_text():
    """The other files the emitted source is made of (their changes must change the tag)."""
    base = Path(__file__).resolve().parent.parent / "csrc"
    inc = Path(__file__).resolve().parent.parent.parent / "include"
    return ("".join((base / f).read_text() for f in ("sbz_thermo.h", "sbz_dual.h", "sbz_rodas5p.h")) +
            "".join((inc / f).read_text() for f in ("sbz.h", "sbz_params.h")))


class _Stored:
    def __init__(self, d):
        self.__dict__.update(d)


class _Desc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nnz_lu", ctypes.c_int), ("cache_path", ctypes.c_char_p)]


class _Opts(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("maxiters", ctypes.c_longlong),
                ("lanes", ctypes.c_int), ("save", ctypes.c_int), ("algo", ctypes.c_int)]


ALGO_WAVE = 1
ALGO_WAVE2 = 3
ALGO_WAVE4 = 4   # four modes per wavefront (models whose border and chains fit 16 lanes; no saves)
ALGO_WAVE_ALTLU = 5   # ALGO_WAVE with the other layout of the border factors in LDS
ALGO_WAVE2H = 6   # two modes per wavefront (models without ALGO_WAVE4 whose border and chains fit 32 lanes; no saves)
ALGO_AUTO = -1   # host-side: symboltz.autotune picks among the wavefront variants


SAVE_STATES = 1
SAVE_SOURCES = 2
SAVEAT_PER_SET = 4   # saveat given as [n_sets][n_save], one time grid per parameter set


def _declare(L):
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    llp = ctypes.POINTER(ctypes.c_longlong)
    L.sbz_model_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Desc), ctypes.POINTER(vp)]
    L.sbz_model_destroy.argtypes = [vp]
    L.sbz_bg_upload.argtypes = [vp, ctypes.c_int, dp, dp, dp, ctypes.c_int, dp]
    L.sbz_plan_create.argtypes = [vp, ip, dp, dp, ctypes.c_int, ctypes.POINTER(_Opts), dp, ctypes.c_int,
                                  ctypes.POINTER(vp)]
    L.sbz_plan_run.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.sbz_plan_run_graph.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.sbz_plan_launch.argtypes = [vp]
    L.sbz_plan_wait.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.sbz_plan_fetch.argtypes = [vp, dp, dp, ip, llp]
    L.sbz_plan_pk.argtypes = [vp, dp]
    L.sbz_plan_fetch_sources.argtypes = [vp, dp]
    L.sbz_plan_destroy.argtypes = [vp]
    L.sbz_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p]


class GpuModel:
    def __init__(self, M):
        self.M = M
        self.L = _lib.lib()
        _declare(self.L)
        em = model_source(M)
        self.em = em
        self.n = em.n
        KCACHE.mkdir(exist_ok=True)
        self.cache_path = KCACHE / f"{em.full_hash}.hsaco"
        desc = _Desc(em.n, em.nnz_LU, str(self.cache_path).encode())
        h = ctypes.c_void_p()
        _lib.check(self.L.sbz_model_create(em.full_src.encode(), ctypes.byref(desc), ctypes.byref(h)),
                   "sbz_model_create")
        self.h = h
        self.n_sets = 0

    def upload_background(self, set_index, t, u, du, P):
        t = np.ascontiguousarray(t, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.ascontiguousarray(du, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        _lib.check(self.L.sbz_bg_upload(self.h, set_index, _lib.ptr(t), _lib.ptr(u), _lib.ptr(du), len(t),
                                        _lib.ptr(P)), "sbz_bg_upload")
        self.n_sets = max(self.n_sets, set_index + 1)

    def plan(self, inst_set, ks, t0, rtol=1e-5, atol=1e-5, maxiters=100000, saveat=None, lanes=0, save=0, algo=0):
        """algo: 0 (library default, one wavefront per SIMD), ALGO_WAVE, ALGO_WAVE2, ALGO_WAVE4 or
        ALGO_AUTO (timed choice among the variants, symboltz.autotune)."""
        if algo == ALGO_AUTO:
            from .autotune import choose
            algo = choose(self, inst_set, ks, t0, rtol, atol, maxiters, saveat, save)
        p = Plan(self, inst_set, ks, t0, rtol, atol, maxiters, saveat, lanes, save, algo)
        p.algo = algo
        return p

    def plan_tangent(self, inst_set, setp, setm, tscale, ks, t0, rtol=1e-5, atol=1e-5, maxiters=100000):
        """A tangent plan (sbz_plan_create_tangent): u and s = ∂u/∂log θ per instance, with the
        parameter sets of log θ ± h (setp, setm) and 1/(2h) (tscale) per instance."""
        return TangentPlan(self, inst_set, setp, setm, tscale, ks, t0, rtol, atol, maxiters)

    def close(self):
        if getattr(self, "h", None):
            self.L.sbz_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    def __init__(self, gm, inst_set, ks, t0, rtol, atol, maxiters, saveat, lanes, save=0, algo=0):
        self.gm = gm
        L = gm.L
        self.set = np.ascontiguousarray(inst_set, dtype=np.int32)
        self.k = np.ascontiguousarray(ks, dtype=np.float64)
        self.t0 = np.ascontiguousarray(t0, dtype=np.float64)
        self.n_inst = len(self.k)
        self.saveat = np.ascontiguousarray(saveat if saveat is not None else np.zeros(0), dtype=np.float64)
        save = int(save) & (SAVE_STATES | SAVE_SOURCES)
        flags = save
        if self.saveat.ndim == 2:   # one row per uploaded parameter set
            if self.saveat.shape[0] != gm.n_sets:
                raise ValueError(f"per-set saveat has {self.saveat.shape[0]} rows for {gm.n_sets} parameter sets")
            flags |= SAVEAT_PER_SET
            self.n_save = self.saveat.shape[1]
        else:
            self.n_save = len(self.saveat)
        opts = _Opts(rtol, atol, int(maxiters), int(lanes), flags, int(algo))
        self.save = save if (save or self.n_save == 0) else SAVE_STATES
        h = ctypes.c_void_p()
        _lib.check(L.sbz_plan_create(gm.h, _lib.ptr(self.set, ctypes.c_int), _lib.ptr(self.k), _lib.ptr(self.t0),
                                     self.n_inst, ctypes.byref(opts), _lib.ptr(self.saveat), self.n_save,
                                     ctypes.byref(h)), "sbz_plan_create")
        self.h = h

    def run(self):
        ms = ctypes.c_float(0.0)
        _lib.check(self.gm.L.sbz_plan_run(self.h, ctypes.byref(ms)), "sbz_plan_run")
        return ms.value

    def launch(self):
        """Enqueue the run and return at once (sbz_plan_launch); `wait()` gives its kernel ms."""
        _lib.check(self.gm.L.sbz_plan_launch(self.h), "sbz_plan_launch")

    def wait(self):
        ms = ctypes.c_float(0.0)
        _lib.check(self.gm.L.sbz_plan_wait(self.h, ctypes.byref(ms)), "sbz_plan_wait")
        return ms.value

    def run_graph(self):
        """The same run replayed from a captured HIP graph (solve and P(k) in one launch); returns
        the graph's duration in ms.  `pk()` afterwards returns the graph's P(k)."""
        ms = ctypes.c_float(0.0)
        _lib.check(self.gm.L.sbz_plan_run_graph(self.h, ctypes.byref(ms)), "sbz_plan_run_graph")
        return ms.value

    def fetch(self, save=False):
        n, m = self.n_inst, self.gm.n
        u = np.zeros((n, m))
        us = np.zeros((n, self.n_save, m)) if (save and self.n_save and (self.save & SAVE_STATES)) else None
        ret = np.zeros(n, dtype=np.int32)
        st = np.zeros((n, 6), dtype=np.int64)
        _lib.check(self.gm.L.sbz_plan_fetch(self.h, _lib.ptr(u), _lib.ptr(us) if us is not None else None,
                                            _lib.ptr(ret, ctypes.c_int), _lib.ptr(st, ctypes.c_longlong)),
                   "sbz_plan_fetch")
        return dict(u=u, usave=us, retcode=ret, stats=st)

    def pk_modes(self, masks, at_saves=False):
        """P(k, τ) for species masks: [n_masks][n_t][n_inst] (n_t = n_save at the saves, else 1)."""
        masks = np.ascontiguousarray(masks, dtype=np.int32)
        n_t = self.n_save if at_saves else 1
        out = np.zeros((len(masks), n_t, self.n_inst))
        L = self.gm.L
        L.sbz_plan_pk_modes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]
        _lib.check(L.sbz_plan_pk_modes(self.h, _lib.ptr(masks, ctypes.c_int), len(masks), int(bool(at_saves)),
                                       _lib.ptr(out)), "sbz_plan_pk_modes")
        return out

    def sources(self):
        out = np.zeros((self.n_inst, self.n_save, 3))
        _lib.check(self.gm.L.sbz_plan_fetch_sources(self.h, _lib.ptr(out)), "sbz_plan_fetch_sources")
        return out

    def pk(self):
        out = np.zeros(self.n_inst)
        _lib.check(self.gm.L.sbz_plan_pk(self.h, _lib.ptr(out)), "sbz_plan_pk")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.gm.L.sbz_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TangentPlan(Plan):
    """Instances integrated with their forward sensitivity s = ∂u/∂log θ (see include/sbz.h)."""

    def __init__(self, gm, inst_set, setp, setm, tscale, ks, t0, rtol, atol, maxiters):
        self.gm = gm
        L = gm.L
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.sbz_plan_create_tangent.argtypes = [ctypes.c_void_p, ip, ip, ip, dp, dp, dp, ctypes.c_int,
                                              ctypes.POINTER(_Opts), ctypes.POINTER(ctypes.c_void_p)]
        L.sbz_plan_fetch_tangent.argtypes = [ctypes.c_void_p, dp, dp]
        L.sbz_plan_pk_tangent.argtypes = [ctypes.c_void_p, dp, dp]
        self.set = np.ascontiguousarray(inst_set, dtype=np.int32)
        self.setp = np.ascontiguousarray(setp, dtype=np.int32)
        self.setm = np.ascontiguousarray(setm, dtype=np.int32)
        self.tscale = np.ascontiguousarray(tscale, dtype=np.float64)
        self.k = np.ascontiguousarray(ks, dtype=np.float64)
        self.t0 = np.ascontiguousarray(t0, dtype=np.float64)
        self.n_inst = len(self.k)
        self.n_save = 0
        self.save = 0
        self.saveat = np.zeros(0)
        self.algo = None
        opts = _Opts(rtol, atol, int(maxiters), 0, 0, 0)
        h = ctypes.c_void_p()
        _lib.check(L.sbz_plan_create_tangent(gm.h, _lib.ptr(self.set, ctypes.c_int), _lib.ptr(self.setp, ctypes.c_int),
                                             _lib.ptr(self.setm, ctypes.c_int), _lib.ptr(self.tscale), _lib.ptr(self.k),
                                             _lib.ptr(self.t0), self.n_inst, ctypes.byref(opts), ctypes.byref(h)),
                   "sbz_plan_create_tangent")
        self.h = h

    def fetch_tangent(self):
        """(s [n_inst][n], f(τ0, u(τ0)) [n_inst][n])."""
        n, m = self.n_inst, self.gm.n
        s, f = np.zeros((n, m)), np.zeros((n, m))
        _lib.check(self.gm.L.sbz_plan_fetch_tangent(self.h, _lib.ptr(s), _lib.ptr(f)), "sbz_plan_fetch_tangent")
        return s, f

    def pk_tangent(self):
        """(P(k, τ0) [n_inst], d log P / d log θ [n_inst])."""
        pk, d = np.zeros(self.n_inst), np.zeros(self.n_inst)
        _lib.check(self.gm.L.sbz_plan_pk_tangent(self.h, _lib.ptr(pk), _lib.ptr(d)), "sbz_plan_pk_tangent")
        return pk, d
