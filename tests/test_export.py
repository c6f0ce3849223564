"""The model export for non-Python callers: the files a Julia binding loads (emitted source, code
object, metadata), with the unknown order and parameter slots equal to the C header's."""
import json
import re
from pathlib import Path

from symboltz import models
from symboltz.emitter import layout
from symboltz.export import export, parse_spec, unknown_names

ROOT = Path(__file__).resolve().parent.parent


def test_unknown_names_follow_the_layout():
    for M in (models.LCDM(lmax=5), models.LCDM(lmax=10, h=False), models.w0waCDM(lmax=6),
              models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4)):
        names = unknown_names(M)
        L = layout(M)
        assert len(names) == M.n_unknowns == L["n"]
        assert names[L["F"]] == "F0" and names[L["G"]] == "G0" and names[L["C"]] == "delta_c"
        assert names[L["B"] + 1] == "theta_b"
        if M.has_h:
            assert names[L["H"]] == "psi0_0"
        if M.de == "w0wa":
            assert names[L["X"]] == "delta_X"


def test_parse_spec():
    assert parse_spec("lcdm:10") == models.LCDM(lmax=10)
    assert parse_spec("lcdm:16,16,10,4") == models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4)
    assert parse_spec("lcdm-massless:10") == models.LCDM(lmax=10, h=False)
    assert parse_spec("w0wa:6") == models.w0waCDM(lmax=6)


def test_export_writes_source_code_object_and_metadata(tmp_path):
    M = models.LCDM(lmax=5)
    name = export(M, tmp_path)
    src = (tmp_path / f"{name}.hip").read_text()
    assert "sbz_wave_kernel" in src
    co = (tmp_path / f"{name}.hsaco").read_bytes()
    assert co[:4] == b"\x7fELF" or co.startswith(b"__CLANG_OFFLOAD_BUNDLE__")   # hipModuleLoad takes either
    meta = json.loads((tmp_path / f"{name}.json").read_text())
    assert meta["n"] == M.n_unknowns and meta["nnz_lu"] > meta["n"]
    assert meta["unknowns"] == unknown_names(M)
    assert len(meta["quadrature"]["x"]) == M.nx
    # every slot of include/sbz_params.h (apart from the library-derived ones) is in the metadata
    hdr = (ROOT / "include" / "sbz_params.h").read_text()
    slots = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define SBZ_P_(\w+)\s+(\d+)", hdr)}
    derived = {"RE1F", "RE1K", "RE2F", "RE2K"}
    assert meta["npar"] == int(re.search(r"#define SBZ_NPAR (\d+)", hdr).group(1))
    for k, v in slots.items():
        if k not in derived:
            assert meta["params"][k] == v, k
