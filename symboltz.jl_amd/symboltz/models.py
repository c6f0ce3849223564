"""Model descriptions: the ΛCDM and w0waCDM configurations of the reference, as data.

`LCDM(...)` mirrors the keyword interface of `ΛCDM` (src/models/cosmologies.jl:27-48):
photons with polarization (photons.jl:6), massless neutrinos (neutrinos.jl:6), massive
neutrinos with `nx` momentum bins (neutrinos.jl:67), CDM, baryons with RECFAST recombination and
tanh reionization (baryons.jl:145), a cosmological constant, and the Harrison–Zel'dovich
primordial spectrum.  `w0waCDM(...)` replaces Λ by the w0wa fluid (cosmologies.jl:215-219).
Species can be dropped (`nu=False`, `h=False`) and each hierarchy can get its own lmax, as the
reference allows through its component keywords (`γ = photons(g; lmax)`, ...).

`parameters_Planck18` restates src/parameters.jl:1-21; `make_params` derives every parameter the
solver needs and packs them into the SBZ_NPAR vector of include/sbz_params.h.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import constants as C
from .quadrature import momentum_quadrature

# Index constants of include/sbz_params.h
NPAR = 64
MAX_NX = 8
P_H, P_T0, P_OMG, P_OMNU, P_OMC, P_OMB, P_OMH, P_OMDE = 0, 1, 2, 3, 4, 5, 6, 7
P_YHE, P_FHE, P_NH, P_T0H, P_Y0H, P_IRHO0 = 8, 9, 10, 11, 12, 13
P_ZRE1, P_DZRE1, P_NRE1, P_ZRE2, P_DZRE2, P_NRE2 = 14, 15, 16, 17, 18, 19
P_W0, P_WA, P_CS2X, P_TAU0, P_KAPPA0, P_TAUREC = 20, 21, 22, 23, 24, 25
P_AS, P_NS, P_KPIVOT, P_FH, P_REION, P_TAUBG0 = 26, 27, 28, 29, 30, 31
P_X, P_W, P_NX, P_NEFF, P_MEV, P_DE, P_HESW, P_HSW = 32, 40, 48, 49, 50, 51, 52, 53


@dataclass(frozen=True)
class Model:
    """Species content and hierarchy truncations of a cosmological model."""
    name: str = "ΛCDM"
    lmax_g: int = 10          # photon temperature and polarization hierarchies (photons.jl:6)
    lmax_nu: int = 10         # massless neutrino hierarchy (neutrinos.jl:6)
    lmax_h: int = 10          # massive neutrino hierarchy per momentum bin (neutrinos.jl:67)
    nx: int = 4               # massive neutrino momentum bins
    has_nu: bool = True
    has_h: bool = True
    de: str = "L"             # "L" cosmological constant, "w0wa" CPL fluid
    reionization: bool = True
    Hswitch: int = 1
    Heswitch: int = 6

    def __post_init__(self):
        if self.lmax_g < 3:
            raise ValueError("Need lmax >= 3")  # photons.jl:7
        if self.has_nu and self.lmax_nu < 3:
            raise ValueError("Need lmax >= 3")
        if self.has_h and (self.lmax_h < 3 or not 1 <= self.nx <= MAX_NX):
            raise ValueError("massive neutrinos need lmax >= 3 and 1 <= nx <= 8")
        if self.de not in ("L", "w0wa"):
            raise ValueError("dark energy must be 'L' or 'w0wa'")
        if self.Hswitch not in (0, 1):
            raise ValueError(f"Supported H switches are 0 and 1. Got {self.Hswitch}.")
        if self.Heswitch not in (0, 6):
            raise ValueError(f"Supported He switches are 0 and 6. Got {self.Heswitch}.")

    @property
    def n_unknowns(self):
        n = 1 + 2 * (self.lmax_g + 1) + 4
        if self.has_nu:
            n += self.lmax_nu + 1
        if self.has_h:
            n += self.nx * (self.lmax_h + 1)
        if self.de == "w0wa":
            n += 2
        return n

    def key(self):
        return (self.lmax_g, self.lmax_nu if self.has_nu else -1, self.lmax_h if self.has_h else -1,
                self.nx if self.has_h else 0, self.de)


def LCDM(lmax=10, *, nu=True, h=True, lmax_g=None, lmax_nu=None, lmax_h=None, nx=4,
         reionization=True, Hswitch=1, Heswitch=6, name="ΛCDM", de="L"):
    """Return the ΛCDM model description (keywords as `ΛCDM(; lmax, ...)`, cosmologies.jl:27-48)."""
    return Model(name=name, lmax_g=lmax if lmax_g is None else lmax_g,
                 lmax_nu=lmax if lmax_nu is None else lmax_nu,
                 lmax_h=lmax if lmax_h is None else lmax_h, nx=nx, has_nu=nu, has_h=h, de=de,
                 reionization=reionization, Hswitch=Hswitch, Heswitch=Heswitch)


def w0waCDM(lmax=10, **kw):
    """ΛCDM with the w0wa fluid instead of Λ (cosmologies.jl:215-219)."""
    kw.setdefault("name", "w0waCDM")
    return LCDM(lmax, de="w0wa", **kw)


def parameters_Planck18(M):
    """Planck 2018 parameters (src/parameters.jl:1-21)."""
    h = 0.6736
    p = {"h": h, "T0": 2.7255, "Omega_c0": 0.1200 / h**2, "Omega_b0": 0.0224 / h**2, "YHe": 0.2454,
         "ln_As1e10": math.log(2.099e-9 * 1e10), "ns": 0.965}
    if M.has_nu:
        p["Neff"] = 2.99
    if M.has_h:
        p["m_eV"] = 0.02
        p["N"] = 3
    if M.de == "w0wa":
        p.update({"w0": -0.9, "wa": 0.2, "cs2": 1.0})
    return p


_quad_cache = {}


def quadrature(nx):
    if nx not in _quad_cache:
        _quad_cache[nx] = momentum_quadrature(nx)
    return _quad_cache[nx]


def make_params(M, pars, tau_bg0=1e-6):
    """Derive all model parameters and pack them into the SBZ_NPAR layout.

    Derivations follow the reference's bindings: Ω_γ0 from T0 (photons.jl:44-46), massless
    Ω_ν0 = Neff·7/8·(4/11)^(4/3)·Ω_γ0 and T_ν = (4/11)^(1/3) T0 (cosmologies.jl:78-84), massive
    Ω_h0 from N, m and the momentum integral Iρ0 (neutrinos.jl:80-86), f_He (baryons.jl:152),
    Ω_Λ0 = 1 - Σ Ω (cosmologies.jl:221-227), As = exp(ln_As1e10)/1e10 and the pivot 0.05/Mpc
    (inflation.jl:4-6).

    Extension not in the reference: `z_re` moves the hydrogen reionization redshift
    (default 7.6711, hard-coded at baryons.jl:202).
    """
    P = np.zeros(NPAR)
    h = float(pars["h"])
    T0 = float(pars["T0"])
    P[P_H] = h
    P[P_T0] = T0
    H0 = C.H100 * h
    Omg = math.pi**2 / 15 * (C.kB * T0) ** 4 / (C.hbar**3 * C.c**5) * 8 * math.pi * C.GN / (3 * H0**2)
    P[P_OMG] = Omg
    if M.has_nu:
        Neff = float(pars.get("Neff", 3.046))
        P[P_NEFF] = Neff
        P[P_OMNU] = Neff * 7 / 8 * (4 / 11) ** (4 / 3) * Omg
    P[P_OMC] = float(pars["Omega_c0"])
    P[P_OMB] = float(pars["Omega_b0"])
    YHe = float(pars["YHe"])
    P[P_YHE] = YHe
    P[P_FHE] = YHe / (C.mHe / C.mH * (1 - YHe))
    if M.has_h:
        N = float(pars.get("N", 3))
        m_eV = float(pars["m_eV"])
        T0h = (4 / 11) ** (1 / 3) * T0
        x, W = quadrature(M.nx)
        y0 = m_eV * C.eV / (C.kB * T0h)   # m c²/(kB T0) with m c² = m_eV eV
        Irho0 = float(np.sum(W * np.sqrt(x**2 + y0**2)))
        Omh = N * 8 * math.pi / 3 * 2 / (2 * math.pi**2) * (C.kB * T0h) ** 4 / (C.hbar * C.c) ** 3 * Irho0 \
            / ((H0 * C.c) ** 2 / C.GN)
        P[P_NH] = N
        P[P_MEV] = m_eV
        P[P_T0H] = T0h
        P[P_Y0H] = y0
        P[P_IRHO0] = Irho0
        P[P_OMH] = Omh
        P[P_NX] = M.nx
        P[P_X:P_X + M.nx] = x
        P[P_W:P_W + M.nx] = W
    P[P_OMDE] = 1.0 - (P[P_OMG] + P[P_OMNU] + P[P_OMC] + P[P_OMB] + P[P_OMH])
    if M.de == "w0wa":
        P[P_DE] = 1
        P[P_W0] = float(pars.get("w0", -1.0))
        P[P_WA] = float(pars.get("wa", 0.0))
        P[P_CS2X] = float(pars.get("cs2", 1.0))
    P[P_ZRE1], P[P_DZRE1], P[P_NRE1] = float(pars.get("z_re", 7.6711)), 0.5, 1.5
    P[P_ZRE2], P[P_DZRE2], P[P_NRE2] = 3.5, 0.5, 1.0
    P[P_REION] = 1.0 if M.reionization else 0.0
    P[P_FH] = 1.125 if M.Hswitch == 1 else 1.14
    P[P_HESW] = float(M.Heswitch)
    P[P_HSW] = float(M.Hswitch)
    lnAs = pars.get("ln_As1e10")
    P[P_AS] = math.exp(lnAs) / 1e10 if lnAs is not None else float(pars.get("As", float("nan")))
    P[P_NS] = float(pars.get("ns", float("nan")))
    P[P_KPIVOT] = 0.05 / C.Mpc / (C.H100 / C.c) / h
    P[P_TAUBG0] = tau_bg0
    P[P_TAU0] = P[P_KAPPA0] = P[P_TAUREC] = float("nan")
    return P
