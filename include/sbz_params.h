/* Layout of the per-parameter-set scalar vector shared by the host library, the device kernels
 * and every caller of the C ABI.  All entries are double precision.  The vector is what the
 * reference keeps in the tunable part of its MTKParameters (src/solve.jl:496-513: the
 * background solution's tunables τ0 and κ0 are copied into the perturbation problem), plus the
 * model constants that the reference bakes into its generated code (momentum-quadrature nodes,
 * src/models/neutrinos.jl:55-60,72-75).
 *
 * Units follow the reference: τ in 1/H0, k in H0/c (src/constants.jl:47-50). */
#ifndef SBZ_PARAMS_H
#define SBZ_PARAMS_H

#define SBZ_NPAR 64
#define SBZ_MAX_NX 8

#define SBZ_P_H        0   /* dimensionless Hubble parameter h                       metric.jl:20 */
#define SBZ_P_T0       1   /* photon temperature today [K]                           photons.jl:44 */
#define SBZ_P_OMG      2   /* Ω_γ0, from T0 and h                                    photons.jl:44-46 */
#define SBZ_P_OMNU     3   /* Ω_ν0 massless, Neff·7/8·(4/11)^(4/3)·Ω_γ0              cosmologies.jl:80 */
#define SBZ_P_OMC      4   /* Ω_c0                                                   generic_species.jl:24 */
#define SBZ_P_OMB      5   /* Ω_b0 */
#define SBZ_P_OMH      6   /* Ω_h0, all N degenerate massive species                 neutrinos.jl:86 */
#define SBZ_P_OMDE     7   /* Ω_Λ0 or Ω_X0 = 1 - Σ others (flatness)                 cosmologies.jl:221-227 */
#define SBZ_P_YHE      8   /* primordial helium mass fraction                        baryons.jl:151 */
#define SBZ_P_FHE      9   /* n_He/n_H = YHe/(mHe/mH (1-YHe))                        baryons.jl:152 */
#define SBZ_P_NH       10  /* number of degenerate massive neutrinos N (0: none)     neutrinos.jl:79 */
#define SBZ_P_T0H      11  /* massive-ν temperature today (4/11)^(1/3) T0            cosmologies.jl:82-84 */
#define SBZ_P_Y0H      12  /* y0 = m c²/(kB T0h)                                     neutrinos.jl:83 */
#define SBZ_P_IRHO0    13  /* Iρ0 = Σ W_i sqrt(x_i² + y0²)                           neutrinos.jl:84 */
#define SBZ_P_ZRE1     14  /* H+He reionization redshift (reference hard-codes 7.6711, baryons.jl:202) */
#define SBZ_P_DZRE1    15  /* its width (0.5) */
#define SBZ_P_NRE1     16  /* its exponent (3/2) */
#define SBZ_P_ZRE2     17  /* second He reionization redshift (3.5, baryons.jl:203) */
#define SBZ_P_DZRE2    18  /* its width (0.5) */
#define SBZ_P_NRE2     19  /* its exponent (1) */
#define SBZ_P_W0       20  /* w0 of the w0wa fluid                                   dark_energy.jl:26 */
#define SBZ_P_WA       21  /* wa */
#define SBZ_P_CS2X     22  /* rest-frame sound speed² of the w0wa fluid */
#define SBZ_P_TAU0     23  /* conformal time today (a = 1), set by the background    solve.jl:181-186 */
#define SBZ_P_KAPPA0   24  /* optical depth offset κ0 = _κ(τ0)                        solve.jl:187-189 */
#define SBZ_P_TAUREC   25  /* argmax of the visibility function over background knots solve.jl:190-195 */
#define SBZ_P_AS       26  /* primordial amplitude As                                inflation.jl:4 */
#define SBZ_P_NS       27  /* spectral index ns */
#define SBZ_P_KPIVOT   28  /* pivot wavenumber 0.05/Mpc in H0/c                       inflation.jl:6 */
#define SBZ_P_FH       29  /* RECFAST hydrogen fudge factor (1.125 for Hswitch = 1)   baryons.jl:79-80 */
#define SBZ_P_REION    30  /* 1: tanh reionization on, 0: off                         baryons.jl:201-205 */
#define SBZ_P_TAUBG0   31  /* start of the background integration (1e-6)             solve.jl:131 */
#define SBZ_P_X        32  /* momentum-quadrature nodes x_i, i < SBZ_MAX_NX           neutrinos.jl:72 */
#define SBZ_P_W        40  /* momentum-quadrature weights W_i                         neutrinos.jl:72 */
#define SBZ_P_NX       48  /* number of momentum bins (as a double) */
#define SBZ_P_NEFF     49  /* effective number of massless species */
#define SBZ_P_MEV      50  /* individual massive-ν mass [eV] */
#define SBZ_P_DE       51  /* dark-energy model: 0 cosmological constant, 1 w0wa fluid */
#define SBZ_P_HESW     52  /* RECFAST helium switch: 6 (all corrections) or 0        baryons.jl:91-114 */
#define SBZ_P_HSW      53  /* RECFAST hydrogen switch: 1 (K fudge fit) or 0          baryons.jl:76-84 */

/* Derived constants, filled by the library itself (sbz_background_solve, sbz_bg_upload) from the
 * entries above; callers leave them at any value. */
#define SBZ_P_RE1F     54  /* (1 + z_re1)^n1 */
#define SBZ_P_RE1K     55  /* tanh rate 1/(n1 (1 + z_re1)^(n1-1) Δz1)                baryons.jl:131-134 */
#define SBZ_P_RE2F     56  /* (1 + z_re2)^n2 */
#define SBZ_P_RE2K     57  /* 1/(n2 (1 + z_re2)^(n2-1) Δz2) */

#endif
