"""Build and call the host emulation of the wave kernel (wave_emu.cpp) — test infrastructure only.

The emitted source of a model is compiled with g++ around the lane shim; the resulting library
solves instances exactly as sbz_wave_kernel does, one emulated 64-lane workgroup at a time."""
import ctypes
import hashlib
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "build"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)


def build(full_src, sanitize=False):
    shim = (HERE / "wave_emu.cpp").read_text()
    tag = hashlib.sha1((full_src + shim + str(sanitize) + Path(__file__).read_text()).encode()).hexdigest()[:16]
    d = BUILD / tag
    lib = d / "libwave_emu.so"
    if lib.exists():
        return lib
    d.mkdir(parents=True, exist_ok=True)
    (d / "emu_kernel.inc").write_text(full_src)
    # -fno-gnu-unique: the kernel's LDS block is a static local of an inline function; without it
    # the dynamic linker would share one such object between the libraries of different models
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-shared", "-fPIC", "-fno-gnu-unique", "-w", "-I", str(d),
           str(HERE / "wave_emu.cpp"),
           "-o", str(lib)]
    if sanitize:
        cmd[1:1] = ["-fsanitize=address", "-fno-omit-frame-pointer"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("emulator build failed:\n" + r.stderr[-4000:])
    return lib


class WaveEmu:
    def __init__(self, full_src, sanitize=False):
        self.L = ctypes.CDLL(str(build(full_src, sanitize)))
        self.L.emu_error.restype = ctypes.c_char_p
        self.L.emu_wave_solve.argtypes = [_dp, _dp, _llp, _ip, _ip, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_longlong, _dp, ctypes.c_int, _dp, _dp, _dp,
                                          _ip, _llp]
        self.n = self.L.emu_n()

    def solve(self, P, t, u, du, ks, t0, rtol=1e-5, atol=1e-5, maxiters=100000, saveat=None, sources=False,
              srcbg=False, o2=False, wave4=False, altlu=False, wave2h=False):
        """srcbg: give the kernel the per-set source background table (as the runtime does for the wave
        kernels) instead of letting it evaluate the background at every save of every mode.
        o2: run sbz_wave_kernel_o2 (the two-wavefronts-per-SIMD build) instead of sbz_wave_kernel.
        wave4: run sbz_wave4_kernel (four modes per wavefront, one emulated wavefront for all).
        altlu: run sbz_wave_kernel_altlu (the other LDS layout of the border factors).
        wave2h: run sbz_wave2h_kernel (two modes per wavefront, one emulated wavefront for all)."""
        p = lambda a, ty=ctypes.c_double: a.ctypes.data_as(ctypes.POINTER(ty))
        nk = len(t)
        knots = np.concatenate([np.asarray(t, float), np.asarray(u, float).ravel(), np.asarray(du, float).ravel()])
        P = np.ascontiguousarray(P, dtype=np.float64)
        koff = np.zeros(1, dtype=np.int64)
        kn = np.array([nk], dtype=np.int32)
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        t0 = np.ascontiguousarray(np.broadcast_to(t0, ks.shape), dtype=np.float64)
        m = len(ks)
        iset = np.zeros(m, dtype=np.int32)
        saveat = np.zeros(0) if saveat is None else np.ascontiguousarray(saveat, dtype=np.float64)
        ns = len(saveat)
        out_u = np.zeros((m, self.n))
        out_save = np.zeros((m, ns, self.n)) if ns else None
        out_src = np.zeros((m, ns, 3)) if (ns and sources) else None
        ret = np.zeros(m, dtype=np.int32)
        st = np.zeros((m, 6), dtype=np.int64)
        self.L.emu_set_srcbg(1 if srcbg else 0)
        self.L.emu_set_o2(4 if wave2h else 3 if altlu else 2 if wave4 else (1 if o2 else 0))
        rc = self.L.emu_wave_solve(p(P), p(knots), p(koff, ctypes.c_longlong), p(kn, ctypes.c_int),
                                   p(iset, ctypes.c_int), p(ks), p(t0), m, rtol, atol, maxiters,
                                   p(saveat) if ns else None, ns, p(out_u), p(out_save) if ns else None,
                                   p(out_src) if out_src is not None else None, p(ret, ctypes.c_int),
                                   p(st, ctypes.c_longlong))
        if rc != 0:
            raise RuntimeError("wave kernel emulation: " + self.L.emu_error().decode())
        return dict(u=out_u, usave=out_save, src=out_src, retcode=ret, stats=st)

    def solve_tangent(self, sets, ks, t0, tscale, rtol=1e-5, atol=1e-5, maxiters=100000):
        """sbz_wave_tan_kernel on the CPU: sets = [(P, t, u, du)] for θ, θ + h, θ - h; returns u, s and
        f(τ0, u) per instance."""
        p = lambda a, ty=ctypes.c_double: a.ctypes.data_as(ctypes.POINTER(ty))
        params = np.ascontiguousarray(np.stack([np.asarray(S[0], float) for S in sets]))
        parts, koff, kn = [], [], []
        off = 0
        for P, t, u, du in sets:
            kn.append(len(t))
            koff.append(off)
            part = np.concatenate([np.asarray(t, float), np.asarray(u, float).ravel(), np.asarray(du, float).ravel()])
            parts.append(part)
            off += len(part)
        knots = np.ascontiguousarray(np.concatenate(parts))
        koff = np.array(koff, dtype=np.int64)
        kn = np.array(kn, dtype=np.int32)
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        t0 = np.ascontiguousarray(np.broadcast_to(t0, ks.shape), dtype=np.float64)
        m = len(ks)
        out_u, out_s, out_f = np.zeros((m, self.n)), np.zeros((m, self.n)), np.zeros((m, self.n))
        ret = np.zeros(m, dtype=np.int32)
        st = np.zeros((m, 6), dtype=np.int64)
        L = self.L
        L.emu_tan_solve.argtypes = [_dp, _dp, _llp, _ip, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_longlong, ctypes.c_double, _dp, _dp, _dp, _ip, _llp]
        rc = L.emu_tan_solve(p(params), p(knots), p(koff, ctypes.c_longlong), p(kn, ctypes.c_int), p(ks), p(t0), m, rtol,
                             atol, maxiters, tscale, p(out_u), p(out_s), p(out_f), p(ret, ctypes.c_int),
                             p(st, ctypes.c_longlong))
        if rc != 0:
            raise RuntimeError("tangent kernel emulation: " + L.emu_error().decode())
        return dict(u=out_u, s=out_s, fend=out_f, retcode=ret, stats=st)
