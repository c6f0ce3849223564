"""Physical constants in SI units (CODATA 2018), as in the reference's src/constants.jl:3-24.

The same values are compiled into the HIP library (csrc/sbz_thermo.h); this module serves the
host-side parameter derivations (Ω_γ0 from T0, neutrino masses, the pivot scale, unit
conversions of k).
"""
import math

c = 299792458.0
h = 6.62607015e-34
hbar = h / (2 * math.pi)
kB = 1.380649e-23
GN = 6.6743e-11
sigmaSB = 5.670374419184429e-8
aR = 4 / c * sigmaSB
sigmaT = 6.6524587321e-29
km = 1000.0
pc = 3.0856775814913673e16
Mpc = 1e6 * pc
H100 = 100 * km / Mpc
k0 = H100 * Mpc / c          # h/Mpc in units of H0/c per unit h (constants.jl:20)
eV = 1.602176634e-19
me = 9.1093837015e-31
amu = 1.6605390666e-27
mH = 1.008 * amu              # PeriodicTable hydrogen atomic mass (constants.jl:22)
mHe = 4.002602 * amu          # PeriodicTable helium atomic mass (constants.jl:23)


def k_dimensionless(k_per_Mpc, h):
    """Convert a wavenumber in 1/Mpc to units of H0/c: k / (h H100 / c)  (constants.jl:47-50)."""
    return k_per_Mpc / (h * H100 / c) / Mpc


def Mpc_per_H0c(h):
    """Length of c/H0 in Mpc; P(k) in (c/H0)^3 times this cubed gives Mpc^3."""
    return c / (h * H100) / Mpc
