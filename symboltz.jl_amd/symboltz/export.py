"""Export model code objects for callers outside Python (the Julia binding).

    python -m symboltz.export OUT_DIR [--model lcdm:10 --model lcdm:16,16,10,4 --model w0wa:6 ...]

writes, per model, `<name>.hip` (the emitted source), `<name>.hsaco` (its gfx950 code object, the
`cache_path` of sbz_model_create) and `<name>.json`: the state size `n`, `nnz_lu`, the unknowns in
the order of the state vector, the index of each parameter slot of include/sbz_params.h, and the
momentum-quadrature nodes and weights the massive-ν slots need.  A caller that holds these files
creates the model with sbz_model_create(src, {n, nnz_lu, path_to_hsaco}) and never needs Python.

Model specs: `lcdm:L` (ΛCDM, every hierarchy truncated at L, 4 massive-ν bins), `lcdm:Lg,Lnu,Lh,nx`,
`lcdm-massless:L` (massless ν only), `w0wa:L` (w0waCDM).  Without --model, the standard models.
"""
import argparse
import json
import shutil
from pathlib import Path

from . import models
from .build import compile_code_object, KCACHE, standard_models
from .emitter import layout
from .solver import model_source


def unknown_names(M):
    """Names of the state vector's entries, in order (the layout of include/sbz.h's model descriptor)."""
    L = layout(M)
    names = ["Phi"]
    names += [f"F{l}" for l in range(M.lmax_g + 1)]
    names += [f"G{l}" for l in range(M.lmax_g + 1)]
    if M.has_nu:
        names += [f"Fnu{l}" for l in range(M.lmax_nu + 1)]
    if M.has_h:
        names += [f"psi{i}_{l}" for i in range(M.nx) for l in range(M.lmax_h + 1)]
    names += ["delta_c", "theta_c", "delta_b", "theta_b"]
    if M.de == "w0wa":
        names += ["delta_X", "theta_X"]
    assert len(names) == L["n"]
    return names


def parse_spec(spec):
    kind, _, arg = spec.partition(":")
    vals = [int(v) for v in arg.split(",")] if arg else [10]
    if kind == "lcdm" and len(vals) == 1:
        return models.LCDM(lmax=vals[0])
    if kind == "lcdm" and len(vals) == 4:
        return models.LCDM(lmax_g=vals[0], lmax_nu=vals[1], lmax_h=vals[2], nx=vals[3])
    if kind == "lcdm-massless" and len(vals) == 1:
        return models.LCDM(lmax=vals[0], h=False)
    if kind == "w0wa" and len(vals) == 1:
        return models.w0waCDM(lmax=vals[0])
    raise ValueError(f"unknown model spec {spec!r}")


def model_name(M):
    de = "w0wa" if M.de == "w0wa" else "lcdm"
    nu = f"nu{M.lmax_nu}" if M.has_nu else "nonu"
    h = f"h{M.lmax_h}x{M.nx}" if M.has_h else "noh"
    return f"{de}_g{M.lmax_g}_{nu}_{h}"


def param_slots():
    return {k[2:]: v for k, v in vars(models).items() if k.startswith("P_") and isinstance(v, int)}


def export(M, out):
    out = Path(out)
    out.mkdir(parents=True, exist_ok=True)
    em = model_source(M)
    name = model_name(M)
    (out / f"{name}.hip").write_text(em.full_src)
    cached = KCACHE / f"{em.full_hash}.hsaco"
    if not cached.exists():
        compile_code_object(em.full_src, cached)
    shutil.copyfile(cached, out / f"{name}.hsaco")
    meta = {"name": name, "n": em.n, "nnz_lu": em.nnz_LU, "unknowns": unknown_names(M),
            "params": param_slots(), "npar": models.NPAR}
    if M.has_h:
        x, w = models.quadrature(M.nx)
        meta["quadrature"] = {"x": [float(v) for v in x], "w": [float(v) for v in w]}
    (out / f"{name}.json").write_text(json.dumps(meta, indent=1, ensure_ascii=False))
    return name


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("out")
    ap.add_argument("--model", action="append", default=[])
    a = ap.parse_args(argv)
    ms = [parse_spec(s) for s in a.model] or standard_models()
    for M in ms:
        print(export(M, a.out))


if __name__ == "__main__":
    main()
