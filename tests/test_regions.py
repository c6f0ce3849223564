"""The region-timer copies of the wave kernels (symboltz.regions) still find every anchor in the
current kernel source, so the profiling tool keeps working as the kernels change."""
from symboltz import models
from symboltz.regions import NR, wave4_source, wave_source
from symboltz.solver import model_source


def test_region_sources_build_for_both_kernels():
    src4 = model_source(models.LCDM(lmax=10, h=False)).full_src
    out4 = wave4_source(src4)
    assert out4.count("clock64()") >= NR
    src1 = model_source(models.LCDM(lmax_g=16, lmax_nu=16, lmax_h=10, nx=4)).full_src
    out1 = wave_source(src1)
    assert out1.count("clock64()") >= NR - 1
    # the copies differ from the model's own source only inside the instrumented body
    assert out4.startswith(src4[:src4.index("SBZ_DEV void sbz_waverow_run")])
    assert out1.startswith(src1[:src1.index("SBZ_DEV void sbz_wave_mode(")])
