"""The mixed-model sources (symboltz.multi): the one-mode kernel and the kernel with the
several-modes bodies, each family in its own namespace, with the family tables the host checks
(sbz_multi_create / sbz_multi_attach_rows).  CPU only: the sources are inspected (the build
compiles both for gfx950)."""
import re

from symboltz import workloads
from symboltz.multi import modes_per_wave, multi_source
from symboltz.solver import model_source


def test_modes_per_wave_follow_the_emitted_eligibility():
    fams = workloads.mixed_families()
    rows = [modes_per_wave(model_source(M).full_src) for M in fams]
    # n = 38 fits 16 lanes; n = 54, 56, 100 fit 32; n = 236's two rows exceed 64 KB of LDS
    assert rows == [4, 2, 2, 2, 1]


def test_two_kernels_two_sources():
    fams = workloads.mixed_families()
    one = multi_source(fams)
    rows = multi_source(fams, rows_kernel=True)
    assert "sbz_multi_kernel(SbzMultiArgs a)" in one and "sbz_multi_kernel_rows(SbzMultiArgs a)" not in one
    assert "sbz_multi_kernel_rows(SbzMultiArgs a)" in rows and "sbz_multi_kernel(SbzMultiArgs a)" not in rows
    assert "sbz_waverow_run" not in one.split("// ---- the mixed-model kernels")[1]
    body = rows.split("// ---- the mixed-model kernels")[1]
    # tickets into the four-modes body of family 0 and the two-modes bodies of families 1-3
    assert "sbzf0::sbz_waverow_run<true, 16>" in body
    for f in (1, 2, 3):
        assert f"sbzf{f}::sbz_waverow_run<true, 32>" in body
    assert "sbzf4::sbz_waverow_run" not in body
    for src in (one, rows):
        assert re.search(r"__constant__ int sbz_multi_modes\[8\] = \{4, 2, 2, 2, 1\};", src)
        assert "__constant__ int sbz_multi_n[8]" in src
        for f in range(len(fams)):
            assert f"namespace sbzf{f} {{" in src
