"""HIP code emitter for the perturbation system of a model.

The reference builds its perturbation ODE symbolically (ModelingToolkit) and generates `f!`,
the analytic sparse Jacobian `jac!` and the initialization system (src/solve.jl:215-227,
src/utils.jl:184-236).  This module does the MI355X counterpart: it states the model's
perturbation equations symbolically (sympy), in terms of the per-mode unknowns u, the wavenumber
k and a small vector B(τ) of background quantities that the device recomputes from the splined
background unknowns, and emits straight-line device code for

  sbz_f      f(τ, u)                                   (the RHS, with common subexpressions)
  sbz_w      W = I/(γh) - J(τ) on the filled sparse pattern of a static LU
  sbz_lu     the numeric LU of W without pivoting in a min-degree elimination order
  sbz_solve  the forward/back substitution with that LU
  sbz_dfdt   ∂f/∂τ = Σ_b (∂f/∂B_b) dB_b/dτ + explicit 1/τ terms
  sbz_ic     adiabatic initial conditions
  sbz_dm     the gauge-invariant matter overdensity Δ_m used by P(k)

together with exact flop counts of each routine (add, sub, mul count 1, fma 2; divisions and
special functions are counted separately).

Equations (component code of the reference): gravity.jl:38-39; photons.jl:30-42,47-53,56-67;
neutrinos.jl:23-26,33-39 (massless) and :109-139 (massive); generic_species.jl:46-57;
baryons via cosmologies.jl:95; dark_energy.jl:50-64; couplings cosmologies.jl:85-97;
P(k) overdensity fourier.jl:39-52.
"""
import hashlib
import re
from pathlib import Path

import sympy as sp
from sympy.printing.c import C99CodePrinter

CSRC = Path(__file__).resolve().parent.parent / "csrc"

# Background vector layout B[] (device: sbz_pt_background in csrc/sbz_kernel.h)
B_NAMES = ["tau", "a", "Hc", "kd", "cs2", "rg", "rn", "rc", "rb", "rh", "ph", "Irho", "IP", "rx", "wx", "wdx"] + \
          [f"E{i}" for i in range(8)] + ["kappa"]
# reciprocals evaluated once per background evaluation (the wavefront kernel's expressions use them
# instead of divisions): 1/ℋ, 1/(Iρ + I_P/3), 1/Iρ, 1/ρ_b, 1/(1 + w_X), 1/τ, 1/E_i
B_INV = ["iHc", "iIrhoP", "iIrho", "irb", "iwx1", "itau"] + [f"iE{i}" for i in range(8)]
B_NAMES = B_NAMES + B_INV
NB = len(B_NAMES)
BIDX = {n: i for i, n in enumerate(B_NAMES)}

# Per-set constants vector C[] (device: computed once per mode from the parameter vector)
C_NAMES = [f"x{i}" for i in range(8)] + [f"w{i}" for i in range(8)] + [f"dl{i}" for i in range(8)] + ["cs2X"] + \
          [f"ix{i}" for i in range(8)] + ["ik", "ik2"]   # 1/x_i; 1/k and 1/k² (per mode)
CIDX = {n: i for i, n in enumerate(C_NAMES)}


def _inv_map(sysm):
    """base expression -> symbol of its precomputed reciprocal (B_INV, 1/x_i, 1/k)."""
    B, C, k = sysm.B, sysm.C, sysm.k
    m = {B["Hc"]: B["iHc"], B["Irho"] + B["IP"] / 3: B["iIrhoP"], B["Irho"]: B["iIrho"], B["rb"]: B["irb"],
         B["wx"] + 1: B["iwx1"], B["tau"]: B["itau"], k: C["ik"]}
    for i in range(8):
        m[B[f"E{i}"]] = B[f"iE{i}"]
        m[C[f"x{i}"]] = C[f"ix{i}"]
    return m


def with_reciprocals(e, invm):
    """e with every b**(-n) (b a key of invm, n a positive integer) replaced by invm[b]**n."""
    e = sp.sympify(e)
    return e.replace(lambda x: x.is_Pow and x.base in invm and x.exp.is_Integer and x.exp < 0,
                     lambda x: invm[x.base] ** (-x.exp))


def layout(M):
    """Index layout of the unknowns (identical to the oracle's and to include/sbz.h)."""
    L = {}
    i = 0
    L["Phi"] = i; i += 1
    L["F"] = i; i += M.lmax_g + 1
    L["G"] = i; i += M.lmax_g + 1
    L["N"] = i
    if M.has_nu:
        i += M.lmax_nu + 1
    L["H"] = i
    if M.has_h:
        i += M.nx * (M.lmax_h + 1)
    L["C"] = i; i += 2
    L["B"] = i; i += 2
    L["X"] = i
    if M.de == "w0wa":
        i += 2
    L["n"] = i
    return L


class System:
    """The symbolic perturbation system of one model."""

    def __init__(self, M):
        self.M = M
        self.L = layout(M)
        n = self.L["n"]
        self.n = n
        self.u = sp.symbols(f"u0:{n}", real=True)
        self.B = {nm: sp.Symbol(f"B_{nm}", real=True) for nm in B_NAMES}
        self.dB = {nm: sp.Symbol(f"dB_{nm}", real=True) for nm in B_NAMES}
        self.C = {nm: sp.Symbol(f"C_{nm}", real=True) for nm in C_NAMES}
        self.k = sp.Symbol("k", positive=True)
        # Φ' and Ψ as placeholders: f_ph is linear in them, so every row splits into a local part
        # plus α Φ' + β Ψ, with Φ' and Ψ themselves linear functionals of u (Woodbury in the kernel)
        self.DPHI = sp.Symbol("DPHI", real=True)
        self.PSI = sp.Symbol("PSI", real=True)
        self.f_ph = self._rhs(ph=True)
        self.f = self._rhs()

    # -- equations ---------------------------------------------------------------------------
    def _rhs(self, ph=False):
        M, L, u, B, C, k = self.M, self.L, self.u, self.B, self.C, self.k
        R = sp.Rational
        pi = sp.pi
        n = self.n
        f = [None] * n
        tau, a, H, kd = B["tau"], B["a"], B["Hc"], B["kd"]
        F = [u[L["F"] + l] for l in range(M.lmax_g + 1)]
        G = [u[L["G"] + l] for l in range(M.lmax_g + 1)]
        Phi = u[L["Phi"]]
        dc, thc, db, thb = u[L["C"]], u[L["C"] + 1], u[L["B"]], u[L["B"] + 1]
        # δρ and Π = Σ (1 + w) ρ σ                                        cosmologies.jl:92-94
        drho = F[0] * B["rg"] + dc * B["rc"] + db * B["rb"]
        Pi = R(4, 3) * B["rg"] * F[2] / 2
        if M.has_nu:
            N = [u[L["N"] + l] for l in range(M.lmax_nu + 1)]
            drho += N[0] * B["rn"]
            Pi += R(4, 3) * B["rn"] * N[2] / 2
        if M.has_h:
            Lh = M.lmax_h
            psi = [[u[L["H"] + i * (Lh + 1) + l] for l in range(Lh + 1)] for i in range(M.nx)]
            E = [B[f"E{i}"] for i in range(M.nx)]
            xs = [C[f"x{i}"] for i in range(M.nx)]
            ws = [C[f"w{i}"] for i in range(M.nx)]
            dls = [C[f"dl{i}"] for i in range(M.nx)]
            Idr = sum(ws[i] * E[i] * psi[i][0] for i in range(M.nx))         # neutrinos.jl:116
            sig_h = R(2, 3) * sum(ws[i] * xs[i] ** 2 / E[i] * psi[i][2] for i in range(M.nx)) / \
                (B["Irho"] + B["IP"] / 3)                                    # neutrinos.jl:121
            drho += Idr / B["Irho"] * B["rh"]
            Pi += (B["rh"] + B["ph"]) * sig_h
        if M.de == "w0wa":
            dX, thX = u[L["X"]], u[L["X"] + 1]
            drho += dX * B["rx"]
        Psi_e = Phi - 12 * pi * a**2 * Pi / k**2                             # gravity.jl:39
        dPhi_e = -R(4, 3) * pi * a**2 / H * drho - k**2 / (3 * H) * Phi - H * Psi_e  # gravity.jl:38
        Psi, dPhi = (self.PSI, self.DPHI) if ph else (Psi_e, dPhi_e)
        self.Psi_e, self.dPhi_e = Psi_e, dPhi_e
        f[L["Phi"]] = dPhi
        # photons                                                           photons.jl:30-33,56-59
        Lg = M.lmax_g
        thg = 3 * k * F[1] / 4
        Pig = F[2] + G[0] + G[2]
        iF, iG = L["F"], L["G"]
        f[iF + 0] = -k * F[1] + 4 * dPhi
        f[iF + 1] = k / 3 * (F[0] - 2 * F[2] + 4 * Psi) - R(4, 3) * kd / k * (thb - thg)
        for l in range(2, Lg):
            f[iF + l] = k / (2 * l + 1) * (l * F[l - 1] - (l + 1) * F[l + 1]) + kd * (F[l] - (Pig / 10 if l == 2 else 0))
        f[iF + Lg] = k * F[Lg - 1] - (Lg + 1) / tau * F[Lg] + kd * F[Lg]
        f[iG + 0] = k * (-G[1]) + kd * (G[0] - Pig / 2)
        f[iG + 1] = k / 3 * (G[0] - 2 * G[2]) + kd * G[1]
        for l in range(2, Lg):
            f[iG + l] = k / (2 * l + 1) * (l * G[l - 1] - (l + 1) * G[l + 1]) + kd * (G[l] - (Pig / 10 if l == 2 else 0))
        f[iG + Lg] = k * G[Lg - 1] - (Lg + 1) / tau * G[Lg] + kd * G[Lg]
        # massless neutrinos                                                neutrinos.jl:23-26
        if M.has_nu:
            Ln, iN = M.lmax_nu, L["N"]
            f[iN + 0] = -k * N[1] + 4 * dPhi
            f[iN + 1] = k / 3 * (N[0] - 2 * N[2] + 4 * Psi)
            for l in range(2, Ln):
                f[iN + l] = k / (2 * l + 1) * (l * N[l - 1] - (l + 1) * N[l + 1])
            f[iN + Ln] = k * N[Ln - 1] - (Ln + 1) / tau * N[Ln]
        # massive neutrinos                                                 neutrinos.jl:128-132
        if M.has_h:
            for i in range(M.nx):
                base = L["H"] + i * (Lh + 1)
                q = xs[i] / E[i]
                ps = psi[i]
                f[base + 0] = -k * q * ps[1] - dPhi * dls[i]
                f[base + 1] = k / 3 * q * (ps[0] - 2 * ps[2]) - k / 3 * E[i] / xs[i] * Psi * dls[i]
                for l in range(2, Lh):
                    f[base + l] = k / (2 * l + 1) * q * (l * ps[l - 1] - (l + 1) * ps[l + 1])
                f[base + Lh] = k / (2 * Lh + 1) * q * (Lh * ps[Lh - 1] - (Lh + 1) * (
                    (2 * Lh + 1) * E[i] / xs[i] * ps[Lh] / (k * tau) - ps[Lh - 1]))
        # CDM (w = c_s² = 0) and baryons                                    generic_species.jl:46-47
        f[L["C"]] = -(thc - 3 * dPhi)
        f[L["C"] + 1] = -H * thc + k**2 * Psi
        f[L["B"]] = -(thb - 3 * dPhi) - 3 * H * B["cs2"] * db
        thint = -kd * 4 * B["rg"] / (3 * B["rb"]) * (thg - thb)              # cosmologies.jl:95
        f[L["B"] + 1] = -H * thb + B["cs2"] * k**2 * db + k**2 * Psi + thint
        # w0wa fluid                                                        dark_energy.jl:50-58
        if M.de == "w0wa":
            w, cs2 = B["wx"], C["cs2X"]
            ca2 = w - B["wdx"] / (3 * H * (1 + w))
            f[L["X"]] = -(1 + w) * (thX - 3 * dPhi) - 3 * H * (cs2 - w) * dX - 9 * (H / k) ** 2 * (1 + w) * (cs2 - ca2) * thX
            f[L["X"] + 1] = -H * (1 - 3 * cs2) * thX + cs2 / (1 + w) * k**2 * dX + k**2 * Psi
        if not ph:
            self.Psi, self.dPhi = Psi, dPhi
        return f

    # -- derived structure ------------------------------------------------------------------
    def jacobian_entries(self):
        """Nonzero entries {(i, j): expr} of J = ∂f/∂u."""
        J = {}
        for i, fi in enumerate(self.f):
            fi = sp.expand(fi) if False else fi
            for j in range(self.n):
                if fi.has(self.u[j]):
                    d = sp.diff(fi, self.u[j])
                    if d != 0:
                        J[(i, j)] = d
        return J

    def ic(self):
        """Adiabatic initial conditions as {index: expr} (zero elsewhere), Φ from the constraint."""
        M, L, B, C, k = self.M, self.L, self.B, self.C, self.k
        R = sp.Rational
        tau, a, kd = B["tau"], B["a"], B["kd"]
        fnu = (B["rn"] + B["rh"]) / (B["rg"] + B["rn"] + B["rh"])           # cosmologies.jl:89
        Psi = 20 * R(1, 2) / (15 + 4 * fnu)                                  # cosmologies.jl:75
        v = {}
        iF, iG = L["F"], L["G"]
        F1 = R(2, 3) * k * tau * Psi
        F2 = -R(8, 15) * k / kd * F1
        v[iF + 0] = -2 * Psi
        v[iF + 1] = F1
        v[iF + 2] = F2
        v[iF + 3] = -R(3, 7) * k / kd * F2
        G2 = R(1, 16) * F2
        v[iG + 0] = R(5, 16) * F2
        v[iG + 1] = -R(1, 16) * k / kd * F2
        v[iG + 2] = G2
        v[iG + 3] = -R(3, 7) * k / kd * G2
        Pi = R(4, 3) * B["rg"] * F2 / 2
        if M.has_nu:
            iN = L["N"]
            N2 = 2 * (R(1, 15) * (k * tau) ** 2 * Psi)
            v[iN + 0] = -2 * Psi
            v[iN + 1] = R(4, 3) / k * (R(1, 2) * k**2 * tau * Psi)
            v[iN + 2] = N2
            v[iN + 3] = R(3, 7) * k * tau * N2
            Pi += R(4, 3) * B["rn"] * N2 / 2
        if M.has_h:
            num = 0
            for i in range(M.nx):
                base = L["H"] + i * (M.lmax_h + 1)
                x, dl, E = C[f"x{i}"], C[f"dl{i}"], B[f"E{i}"]
                v[base + 0] = -R(1, 4) * (-2 * Psi) * dl
                v[base + 1] = -R(1, 3) * E / x * (R(1, 2) * k * tau * Psi) * dl
                p2 = -R(1, 2) * (R(1, 15) * (k * tau) ** 2 * Psi) * dl
                v[base + 2] = p2
                num += C[f"w{i}"] * x**2 / E * p2
            Pi += (B["rh"] + B["ph"]) * R(2, 3) * num / (B["Irho"] + B["IP"] / 3)
        v[L["C"]] = -R(3, 2) * Psi
        v[L["C"] + 1] = R(1, 2) * k**2 * tau * Psi
        v[L["B"]] = -R(3, 2) * Psi
        v[L["B"] + 1] = R(1, 2) * k**2 * tau * Psi
        if M.de == "w0wa":
            v[L["X"]] = -R(3, 2) * (1 + B["wx"]) * Psi
            v[L["X"] + 1] = R(1, 2) * k**2 * tau * Psi
        v[L["Phi"]] = Psi + 12 * sp.pi * a**2 * Pi / k**2                   # gravity.jl:39
        return v

    def sources(self):
        """CMB source functions (S_T, S_E, S_ψ) of cosmologies.jl:99-105 in terms of u, f = u',
        g = u'', the background B, its derivative dB and the second derivative of κ̇."""
        L, u, B, dB, k = self.L, self.u, self.B, self.dB, self.k
        n = self.n
        fs = sp.symbols(f"f0:{n}", real=True)
        gs = sp.symbols(f"g0:{n}", real=True)
        ddkd = sp.Symbol("ddkd", real=True)
        tau0, taurec = sp.Symbol("tau0", real=True), sp.Symbol("taurec", real=True)

        def D(expr):
            out = sum(sp.diff(expr, u[j]) * fs[j] for j in range(n) if expr.has(u[j]))
            out += sum(sp.diff(expr, B[nm]) * dB[nm] for nm in B_NAMES if expr.has(B[nm]))
            return out

        kd, kdd = B["kd"], dB["kd"]
        eK = sp.exp(-B["kappa"])
        v = -kd * eK                                       # v = D(e^{-κ}), baryons.jl:179
        vd = (-kdd + kd**2) * eK
        vdd = (-ddkd + 3 * kd * kdd - kd**3) * eK
        iF, iG = L["F"], L["G"]
        Pig = u[iF + 2] + u[iG] + u[iG + 2]                 # Π_γ, photons.jl:39
        Pigd = fs[iF + 2] + fs[iG] + fs[iG + 2]
        Pigdd = gs[iF + 2] + gs[iG] + gs[iG + 2]
        Psi = self.Psi
        Phi = u[L["Phi"]]
        thb, thbd = u[L["B"] + 1], fs[L["B"] + 1]
        ST_SW = v * (u[iF] / 4 + Psi + Pig / 16)
        ST_ISW = eK * (D(Psi) + fs[L["Phi"]])
        ST_Dop = (vd * thb + v * thbd) / k**2
        ST_pol = sp.Rational(3, 16) / k**2 * (vdd * Pig + 2 * vd * Pigd + v * Pigdd)
        chi = tau0 - B["tau"]
        ST = ST_SW + ST_ISW + ST_Dop + ST_pol
        SE = sp.Rational(3, 16) * v * Pig / (k * chi) ** 2
        Spsi = -(Psi + Phi) * (B["tau"] - taurec) / (tau0 - taurec) / (tau0 - B["tau"])
        return ST, SE, Spsi

    def species_deltas(self):
        """(Δ_s, ρ_s) for s = c, b and, with massive ν, h: Δ_s = δ_s + 3ℋ(1+w_s)θ_s/k², the species of
        total_symbolic_gauge_invariant_overdensities (fourier.jl:39-52)."""
        M, L, u, B, C, k = self.M, self.L, self.u, self.B, self.C, self.k
        H = B["Hc"]
        out = [(u[L["C"]] + 3 * H * u[L["C"] + 1] / k**2, B["rc"]),
               (u[L["B"]] + 3 * H * u[L["B"] + 1] / k**2, B["rb"])]
        if M.has_h:
            Lh = M.lmax_h
            Idr = sum(C[f"w{i}"] * B[f"E{i}"] * u[L["H"] + i * (Lh + 1)] for i in range(M.nx))
            uh = sum(C[f"w{i}"] * C[f"x{i}"] * u[L["H"] + i * (Lh + 1) + 1] for i in range(M.nx)) / \
                (B["Irho"] + B["IP"] / 3)
            wh = B["ph"] / B["rh"]
            out.append((Idr / B["Irho"] + 3 * H * (1 + wh) * (uh * k) / k**2, B["rh"]))
        return out

    def delta_m(self):
        """Δ_m = Σ ρ_s Δ_s / Σ ρ_s over c, b, h; Δ_s = δ_s + 3ℋ(1+w_s)θ_s/k² (fourier.jl:39-52)."""
        M, L, u, B, C, k = self.M, self.L, self.u, self.B, self.C, self.k
        H = B["Hc"]
        Dc = u[L["C"]] + 3 * H * u[L["C"] + 1] / k**2
        Db = u[L["B"]] + 3 * H * u[L["B"] + 1] / k**2
        num = B["rc"] * Dc + B["rb"] * Db
        den = B["rc"] + B["rb"]
        if M.has_h:
            Lh = M.lmax_h
            Idr = sum(C[f"w{i}"] * B[f"E{i}"] * u[L["H"] + i * (Lh + 1)] for i in range(M.nx))
            uh = sum(C[f"w{i}"] * C[f"x{i}"] * u[L["H"] + i * (Lh + 1) + 1] for i in range(M.nx)) / \
                (B["Irho"] + B["IP"] / 3)
            dh = Idr / B["Irho"]
            wh = B["ph"] / B["rh"]
            Dh = dh + 3 * H * (1 + wh) * (uh * k) / k**2
            num += B["rh"] * Dh
            den += B["rh"]
        return num / den


# ------------------------------------------------------------------------------- LU structure
def min_degree_order(n, pattern):
    """Greedy minimum-degree elimination order on the symmetrised pattern (fill-reducing)."""
    adj = [set() for _ in range(n)]
    for (i, j) in pattern:
        if i != j:
            adj[i].add(j)
            adj[j].add(i)
    alive = set(range(n))
    order = []
    while alive:
        p = min(alive, key=lambda v: (len(adj[v]), v))
        order.append(p)
        nb = adj[p]
        for a in nb:
            adj[a].discard(p)
            adj[a] |= (nb - {a})
        alive.discard(p)
        adj[p] = set()
    return order


def symbolic_lu(n, pattern, order):
    """Fill pattern of the LU of a matrix with `pattern`, eliminated in `order` without pivoting.
    Returns (filled set of (i, j), lower dict p -> [rows], upper dict p -> [cols])."""
    pos = {v: r for r, v in enumerate(order)}
    rows = {i: set() for i in range(n)}
    for (i, j) in pattern:
        rows[i].add(j)
    for i in range(n):
        rows[i].add(i)
    filled = set()
    lower, upper = {}, {}
    cols = {j: set() for j in range(n)}
    for i in range(n):
        for j in rows[i]:
            cols[j].add(i)
    for p in order:
        lo = sorted([i for i in cols[p] if pos[i] > pos[p]], key=lambda i: pos[i])
        up = sorted([j for j in rows[p] if pos[j] > pos[p]], key=lambda j: pos[j])
        lower[p], upper[p] = lo, up
        for i in lo:
            for j in up:
                if j not in rows[i]:
                    rows[i].add(j)
                    cols[j].add(i)
    for i in range(n):
        for j in rows[i]:
            filled.add((i, j))
    return filled, lower, upper


# ------------------------------------------------------------------------------- printing
class _Printer(C99CodePrinter):
    def _print_Pow(self, expr):
        b, e = expr.as_base_exp()
        if e.is_Integer:
            ei = int(e)
            bs = self.parenthesize(b, 100)
            if ei == 1:
                return bs
            if ei == 2:
                return f"({bs}*{bs})"
            if ei == 3:
                return f"({bs}*{bs}*{bs})"
            if ei == -1:
                return f"(1.0/{bs})"
            if ei == -2:
                return f"(1.0/({bs}*{bs}))"
            if ei == -3:
                return f"(1.0/({bs}*{bs}*{bs}))"
        if e == sp.Rational(1, 2):
            return f"sqrt({self._print(b)})"
        return f"pow({self._print(b)}, {self._print(e)})"

    def _print_Rational(self, expr):
        return repr(float(expr.p) / float(expr.q))

    def _print_Integer(self, expr):
        return repr(float(int(expr)))

    def _print_Float(self, expr):
        return repr(float(expr))

    def _print_Pi(self, expr):
        return "3.141592653589793"

    def _print_Symbol(self, expr):
        nm = expr.name
        if nm.startswith("B_"):
            return f"B[{BIDX[nm[2:]]}]"
        if nm.startswith("dB_"):
            return f"dB[{BIDX[nm[3:]]}]"
        if nm.startswith("C_"):
            return f"Cs[{CIDX[nm[2:]]}]"
        if re.fullmatch(r"u\d+", nm):
            return f"u[{nm[1:]}]"
        if re.fullmatch(r"f\d+", nm):
            return f"fv[{nm[1:]}]"
        if re.fullmatch(r"g\d+", nm):
            return f"gv[{nm[1:]}]"
        return nm


_PR = _Printer()


def cprint(e):
    return _PR.doprint(e)


def count_ops(exprs):
    """Flop count of expressions: add/sub/mul = 1, division = 1 (reported also separately)."""
    flops = 0
    divs = 0
    special = 0
    for e in exprs:
        for node in sp.preorder_traversal(e):
            if isinstance(node, sp.Add):
                flops += len(node.args) - 1
            elif isinstance(node, sp.Mul):
                nneg = sum(1 for a in node.args if a == -1)
                flops += len(node.args) - 1 - nneg
            elif isinstance(node, sp.Pow):
                b, ex = node.as_base_exp()
                if ex.is_Integer:
                    ei = abs(int(ex))
                    flops += max(ei - 1, 0)
                    if int(ex) < 0:
                        divs += 1
                        flops += 1
                else:
                    special += 1
            elif isinstance(node, sp.Function):
                special += 1
    return flops, divs, special


def _preload(lines):
    """Hoist every u[i], B[i], dB[i], Cs[i] read of a block into locals loaded at its top, ahead of
    a compiler barrier: the LDS reads are then all in flight at once instead of being issued one
    dependency at a time."""
    pat = re.compile(r"\b(u|B|dB|Cs)\[(\d+)\]")
    seen = {}
    for ln in lines:
        for m in pat.finditer(ln):
            seen.setdefault((m.group(1), int(m.group(2))), None)
    loc = {"u": "pu", "B": "pb", "dB": "pd", "Cs": "pc"}
    head = [f"    const double {loc[a]}{i} = {a}[{i}];" for (a, i) in seen]
    body = [pat.sub(lambda m: f"{loc[m.group(1)]}{m.group(2)}", ln) for ln in lines]
    return head + ['    asm volatile("" ::: "memory");'] + body


def _emit_block(assign, tmp_prefix, opt=True, interleave=False):
    """CSE a list of (lhs_code, expr) and return C lines plus the op count.

    With interleave, each output follows directly the temporaries it is the first to need, so
    temporaries die early (lower register pressure for long blocks that store their outputs)."""
    exprs = [e for _, e in assign]
    if opt:
        repl, red = sp.cse(exprs, symbols=sp.numbered_symbols(tmp_prefix), optimizations="basic")
    else:
        repl, red = [], exprs
    lines = []
    if not interleave:
        for s, e in repl:
            lines.append(f"    const double {s} = {cprint(e)};")
        for (lhs, _), e in zip(assign, red):
            lines.append(f"    {lhs} = {cprint(e)};")
    else:
        defs = {s: e for s, e in repl}
        pos = {s: i for i, (s, _) in enumerate(repl)}
        done = set()

        def need(e, acc):
            for sym in e.free_symbols:
                if sym in defs and sym not in done and sym not in acc:
                    acc.add(sym)
                    need(defs[sym], acc)

        for (lhs, _), e in zip(assign, red):
            acc = set()
            need(e, acc)
            for sym in sorted(acc, key=pos.get):
                lines.append(f"    const double {sym} = {cprint(defs[sym])};")
                done.add(sym)
            lines.append(f"    {lhs} = {cprint(e)};")
    ops = count_ops([e for _, e in repl] + list(red))
    return lines, ops


class Emitted:
    """Emitted device source of one model, with its structure metadata."""

    def __init__(self, M, fence=0):
        self.M = M
        self.fence = fence  # emit a compiler memory fence every `fence` LU pivots / solve rows (0: none)
        sysm = System(M)
        self.sys = sysm
        n = sysm.n
        self.n = n
        J = sysm.jacobian_entries()
        self.J = J
        pattern = set(J.keys())
        order = min_degree_order(n, pattern | {(j, i) for (i, j) in pattern})
        filled, lower, upper = symbolic_lu(n, pattern, order)
        self.order, self.lower, self.upper = order, lower, upper
        slots = sorted(filled, key=lambda ij: (ij[0], ij[1]))
        self.slot = {ij: s for s, ij in enumerate(slots)}
        self.nnz_J = len(J)
        self.nnz_LU = len(slots)
        self.ops = {}
        self.src = self._emit()
        self.hash = hashlib.sha1(self.src.encode()).hexdigest()[:16]

    def _emit(self):
        S, sysm, n = self.slot, self.sys, self.n
        out = []
        out.append(f"// Emitted by symboltz.emitter for model {self.M}")
        out.append(f"#define SBZ_N {n}")
        out.append(f"#define SBZ_NNZ_LU {self.nnz_LU}")
        # background vector B[] and constants Cs[] layouts (printer indices; filled by the kernel)
        out.append(f"#define SBZ_NB {NB}")
        out.append(f"#define SBZ_NC {len(C_NAMES)}")
        for nm in ["iHc", "iIrhoP", "iIrho", "irb", "iwx1", "itau", "iE0"]:
            out.append(f"#define SBZ_B_{nm.upper()} {BIDX[nm]}")
        for nm in ["ix0", "ik", "ik2"]:
            out.append(f"#define SBZ_C_{nm.upper()} {CIDX[nm]}")
        out.append(f"#define SBZ_NX {self.M.nx if self.M.has_h else 0}")
        out.append(f"#define SBZ_HAS_W0WA {1 if self.M.de == 'w0wa' else 0}")
        # f
        lines, ops = _emit_block([(f"du[{i}]", e) for i, e in enumerate(sysm.f)], "tf")
        self.ops["f"] = ops
        out.append("template <class UI, class UO>")
        out.append("SBZ_DEV void sbz_f(const double* __restrict__ B, const double* __restrict__ Cs, double k, UI u, UO du) {")
        out += lines
        out.append("}")
        # W = I/(γh) - J on the filled pattern
        assign = []
        for (i, j), s in sorted(S.items(), key=lambda kv: kv[1]):
            e = -self.J.get((i, j), sp.Integer(0))
            if i == j:
                e = sp.Symbol("ig") + e
            assign.append((f"W[{s}]", e))
        lines, ops = _emit_block(assign, "tw")
        self.ops["w"] = ops
        out.append("template <class WT>")
        out.append("SBZ_DEV void sbz_w(const double* __restrict__ B, const double* __restrict__ Cs, double k, double ig, WT W) {")
        out += lines
        out.append("}")
        # LU (no pivoting): diagonal slots keep the reciprocal of the pivot
        lu = ["template <class WT>", "SBZ_DEV bool sbz_lu(WT W) {", "    bool ok = true;"]
        flops_lu = 0
        for p in self.order:
            dp = S[(p, p)]
            lu.append(f"    {{ const double piv = W[{dp}]; ok = ok && (piv != 0.0) && (piv == piv) && (fabs(piv) < 1e300);")
            lu.append(f"      const double rp = 1.0 / piv; W[{dp}] = rp;")
            flops_lu += 1
            for i in self.lower[p]:
                li = S[(i, p)]
                lu.append(f"      {{ const double l = W[{li}] * rp; W[{li}] = l;")
                flops_lu += 1
                for j in self.upper[p]:
                    lu.append(f"        W[{S[(i, j)]}] -= l * W[{S[(p, j)]}];")
                    flops_lu += 2
                lu.append("      }")
            lu.append("    }")
            if self.fence and (self.order.index(p) + 1) % self.fence == 0:
                lu.append('    asm volatile("" ::: "memory");')
        lu.append("    return ok;")
        lu.append("}")
        self.ops["lu"] = (flops_lu, n, 0)
        out += lu
        # solve
        so = ["template <class WT, class BT>", "SBZ_DEV void sbz_solve(WT W, BT b) {"]
        flops_s = 0
        for r, p in enumerate(self.order):
            for i in self.lower[p]:
                so.append(f"    b[{i}] -= W[{S[(i, p)]}] * b[{p}];")
                flops_s += 2
            if self.fence and (r + 1) % self.fence == 0:
                so.append('    asm volatile("" ::: "memory");')
        for r, p in enumerate(reversed(self.order)):
            terms = " ".join(f"- W[{S[(p, j)]}] * b[{j}]" for j in self.upper[p])
            flops_s += 2 * len(self.upper[p]) + 1
            so.append(f"    b[{p}] = (b[{p}] {terms}) * W[{S[(p, p)]}];")
            if self.fence and (r + 1) % self.fence == 0:
                so.append('    asm volatile("" ::: "memory");')
        so.append("}")
        self.ops["solve"] = (flops_s, 0, 0)
        out += so
        # dfdt = Σ_b ∂f/∂B_b dB_b (B_tau has dB = 1 set by the caller)
        dassign = []
        for i, fi in enumerate(sysm.f):
            e = sp.Integer(0)
            for nm in B_NAMES:
                bs = sysm.B[nm]
                if fi.has(bs):
                    e += sp.diff(fi, bs) * sysm.dB[nm]
            dassign.append((f"out[{i}]", e))
        lines, ops = _emit_block(dassign, "td")
        self.ops["dfdt"] = ops
        out.append("template <class UI, class UO>")
        out.append("SBZ_DEV void sbz_dfdt(const double* __restrict__ B, const double* __restrict__ dB, const double* __restrict__ Cs, double k, UI u, UO out) {")
        out += lines
        out.append("}")
        # initial conditions
        ic = sysm.ic()
        lines, _ = _emit_block([(f"u[{i}]", ic.get(i, sp.Integer(0))) for i in range(n)], "ti")
        out.append("template <class UO>")
        out.append("SBZ_DEV void sbz_ic(const double* __restrict__ B, const double* __restrict__ Cs, double k, UO u) {")
        out += lines
        out.append("}")
        # CMB sources at one time (S_ψ is zero before τrec; S_E and S_ψ are set to zero at χ = 0)
        ST, SE, Spsi = sysm.sources()
        lines, ops = _emit_block([("S[0]", ST), ("S[1]", SE), ("const double spsi", Spsi)], "ts")
        self.ops["src"] = ops
        out.append("template <class UI, class FI, class GI>")
        out.append("SBZ_DEV void sbz_src(const double* __restrict__ B, const double* __restrict__ dB, double ddkd, "
                   "const double* __restrict__ Cs, double k, double tau0, double taurec, UI u, FI fv, GI gv, double* S) {")
        out += [ln.replace("    const double spsi =", "    const double spsi =") for ln in lines]
        out.append("    const double tau = B[0];")
        out.append("    S[2] = (tau >= taurec && tau < tau0) ? spsi : 0.0;")
        out.append("    if (!(tau < tau0)) S[1] = 0.0;")
        out.append("}")
        # Δ_m
        lines, _ = _emit_block([("double r", sysm.delta_m())], "tm")
        out.append("template <class UI>")
        out.append("SBZ_DEV double sbz_dm(const double* __restrict__ B, const double* __restrict__ Cs, double k, UI u) {")
        out += [ln.replace("    double r =", "    const double r =") for ln in lines]
        out.append("    return r;")
        out.append("}")
        # Δ of a species combination (bit 0: c, 1: b, 2: h); one species is returned as is
        spc = sysm.species_deltas()
        items = []
        for j, (dj, rj) in enumerate(spc):
            items += [(f"double d{j}", dj), (f"double r{j}", rj)]
        lines, _ = _emit_block(items, "ts")
        out.append("template <class UI>")
        out.append("SBZ_DEV double sbz_dspecies(const double* __restrict__ B, const double* __restrict__ Cs, double k, "
                   "UI u, int mask) {")
        out += lines
        out.append("    double num = 0.0, den = 0.0, one = 0.0;")
        out.append("    int cnt = 0;")
        for j in range(len(spc)):
            out.append(f"    if (mask & {1 << j}) {{ num += r{j} * d{j}; den += r{j}; one = d{j}; ++cnt; }}")
        out.append("    return cnt == 1 ? one : num / den;")
        out.append("}")
        out.append(f"#define SBZ_SPECIES_MASK {(1 << len(spc)) - 1}")
        out += self._emit_wave()
        f_f, f_w, f_lu, f_s, f_d = (self.ops["f"][0] + self.ops["f"][1], self.ops["w"][0] + self.ops["w"][1],
                                    self.ops["lu"][0], self.ops["solve"][0],
                                    self.ops["dfdt"][0] + self.ops["dfdt"][1])
        # Rodas5P step: 8 f + W + LU + 8 solves + dfdt + stage combinations + error norm
        lin = n * (2 * 28 + 2 * 28 + 2 * 8 + 4) + n * 6
        self.flops_step = 8 * f_f + f_w + f_lu + 8 * f_s + f_d + lin
        out.append(f"#define SBZ_FLOPS_F {f_f}")
        out.append(f"#define SBZ_FLOPS_W {f_w}")
        out.append(f"#define SBZ_FLOPS_LU {f_lu}")
        out.append(f"#define SBZ_FLOPS_SOLVE {f_s}")
        out.append(f"#define SBZ_FLOPS_DFDT {f_d}")
        out.append(f"#define SBZ_FLOPS_STEP {self.flops_step}")
        return "\n".join(out) + "\n"

    def _check_lowrank(self, border, Jh, alpha, beta, pvec, qvec):
        """J = ∂h/∂u + α pᵀ + β qᵀ on the border block, checked at a random point."""
        import random
        exprs = [e for e in self.J.values()] + list(Jh.values()) + list(alpha.values()) + list(beta.values()) + \
            list(pvec.values()) + list(qvec.values())
        syms = set()
        for e in exprs:
            syms |= sp.sympify(e).free_symbols
        rng = random.Random(12345)
        vals = {x: sp.Float(rng.uniform(0.5, 2.0)) for x in syms}
        for i in border:
            for j in border:
                full = float(sp.sympify(self.J.get((i, j), 0)).xreplace(vals))
                rec = float(sp.sympify(Jh.get((i, j), 0) + alpha[i] * pvec[j] + beta[i] * qvec[j]).xreplace(vals))
                if abs(full - rec) > 1e-9 * (1.0 + abs(full)):
                    raise ValueError(f"low-rank split of the border Jacobian fails at ({i}, {j})")

    def _emit_wave(self):
        """Structure and straight-line code of the wavefront-per-mode kernel (see sbz_kernel.hip)."""
        M, sysm, n = self.M, self.sys, self.n
        chains, border = wave_structure(M)
        bset = set(border)
        nb = len(border)
        if nb > 64 or len(chains) > 64:
            raise ValueError("wave kernel supports at most 64 border unknowns and 64 chains")
        # Border rows split as f_i = h_i(u) + α_i Φ' + β_i Ψ with Φ' = p·u, Ψ = q·u, so the border
        # block of W is A - α pᵀ - β qᵀ with A = I/(γh) - ∂h/∂u block diagonal (Woodbury).
        u = sysm.u
        DP, PS = sysm.DPHI, sysm.PSI
        alpha = {c: sp.diff(sysm.f_ph[c], DP) for c in border}
        beta = {c: sp.diff(sysm.f_ph[c], PS) for c in border}
        hrow = {c: sysm.f_ph[c].subs({DP: 0, PS: 0}) for c in border}
        pvec = {c: sp.diff(sysm.dPhi_e, u[c]) for c in range(n)}
        qvec = {c: sp.diff(sysm.Psi_e, u[c]) for c in range(n)}
        if any(pvec[c] != 0 or qvec[c] != 0 for c in range(n) if c not in bset):
            raise ValueError("Φ' and Ψ must depend on border unknowns only")
        Jh = {}
        for i in border:
            for j in border:
                d = sp.diff(hrow[i], u[j])
                if d != 0:
                    Jh[(i, j)] = d
        self._check_lowrank(border, Jh, alpha, beta, pvec, qvec)
        bpat = set(Jh.keys())
        loc = {c: r for r, c in enumerate(border)}
        lpat = {(loc[i], loc[j]) for (i, j) in bpat}
        order = min_degree_order(nb, lpat | {(j, i) for (i, j) in lpat})
        lane_of = {border[r]: ln for ln, r in enumerate(order)}     # comp -> lane
        bcomp = [border[r] for r in order]                         # lane -> comp
        # pattern of A in lane space; eliminate lanes in order 0..nb-1
        Lpat = {(lane_of[i], lane_of[j]) for (i, j) in bpat}
        filled, lower, upper = symbolic_lu(nb, Lpat, list(range(nb)))
        chains = sorted(chains, key=lambda ch: -len(ch[3]))   # equal lengths on adjacent lanes
        self.wave = dict(nb=nb, nc=len(chains), bcomp=bcomp, chains=chains)
        mch = max(len(ch[3]) for ch in chains)
        groups = []   # (first lane, last lane + 1, length): chain code is compiled per group
        for ci, ch in enumerate(chains):
            if groups and groups[-1][2] == len(ch[3]):
                groups[-1][1] = ci + 1
            else:
                groups.append([ci, ci + 1, len(ch[3])])
        out = [f"#define SBZ_NBORD {nb}", f"#define SBZ_NCHAIN {len(chains)}", f"#define SBZ_MCHAIN {mch}",
               # four modes per wavefront (16 lanes each) when the border and the chains fit 16 lanes
               f"#define SBZ_WAVE4_OK {1 if (nb <= 16 and len(chains) <= 16) else 0}"]
        # the tangent kernel when its LDS (primal and tangent vectors, three backgrounds, border
        # staging and factors) fits one workgroup's 64 KB
        tan_lds = 8 * (30 * n + 24 * NB + len(C_NAMES) + 2 * nb * (nb + 6) + nb + len(chains) + 1)
        out.append(f"#define SBZ_TAN_OK {1 if tan_lds <= 65536 else 0}")
        out.append("#define SBZ_CHAIN_GROUPS(X) " + " ".join(f"X({a}, {b}, {n_})" for a, b, n_ in groups))
        out.append("__constant__ int sbz_bord_comp[SBZ_NBORD] = {" + ", ".join(map(str, bcomp)) + "};")
        for ch in chains:   # the kernel addresses chain members as base + (l - 3)
            if ch[3] != list(range(ch[3][0], ch[3][0] + len(ch[3]))):
                raise ValueError("wave kernel needs every chain stored contiguously")
        out.append("__constant__ int sbz_chain_base[SBZ_NCHAIN] = {" + ", ".join(str(ch[3][0]) for ch in chains) + "};")
        out.append("__constant__ int sbz_chain_len[SBZ_NCHAIN] = {" + ", ".join(str(len(ch[3])) for ch in chains) + "};")
        out.append("__constant__ int sbz_chain_att[SBZ_NCHAIN] = {" + ", ".join(str(lane_of[ch[4]]) for ch in chains) + "};")
        out.append("__constant__ int sbz_chain_attc[SBZ_NCHAIN] = {" + ", ".join(str(ch[4]) for ch in chains) + "};")
        out.append("__constant__ int sbz_chain_kind[SBZ_NCHAIN] = {" + ", ".join(str(ch[0]) for ch in chains) + "};")
        out.append("__constant__ int sbz_chain_L[SBZ_NCHAIN] = {" + ", ".join(str(ch[1]) for ch in chains) + "};")
        out.append("__constant__ int sbz_chain_bin[SBZ_NCHAIN] = {" + ", ".join(str(ch[2]) for ch in chains) + "};")
        bch = [-1] * nb
        for ci, ch in enumerate(chains):
            bch[lane_of[ch[4]]] = ci
        out.append("__constant__ int sbz_bord_chain[SBZ_NBORD] = {" + ", ".join(map(str, bch)) + "};")
        # border RHS and ∂f/∂τ: every lane evaluates all border rows, keeps its own
        invm = _inv_map(sysm)
        assign = [(f"const double wf{j}", with_reciprocals(sysm.f[c], invm)) for j, c in enumerate(bcomp)]
        lines, _ = _emit_block(assign, "tbf")
        lines = _preload(lines)
        out.append("template <class UI>")
        out.append("SBZ_DEV void sbz_wb_f(int lane, const double* __restrict__ B, const double* __restrict__ Cs, double k, UI u, double* out) {")
        out += lines
        out.append("    if (lane == 0) {")
        out += [f"        out[{c}] = wf{j};" for j, c in enumerate(bcomp)]
        out.append("    }")
        out.append("}")
        dassign = []
        for j, c in enumerate(bcomp):
            fi = sysm.f[c]
            e = sp.Integer(0)
            for nm in B_NAMES:
                if fi.has(sysm.B[nm]):
                    e += sp.diff(fi, sysm.B[nm]) * sysm.dB[nm]
            dassign.append((f"const double wd{j}", with_reciprocals(e, invm)))
        lines, _ = _emit_block(dassign, "tbd")
        lines = _preload(lines)
        out.append("template <class UI>")
        out.append("SBZ_DEV void sbz_wb_dfdt(int lane, const double* __restrict__ B, const double* __restrict__ dB, const double* __restrict__ Cs, double k, UI u, double* out) {")
        out += lines
        out.append("    if (lane == 0) {")
        out += [f"        out[{c}] = wd{j};" for j, c in enumerate(bcomp)]
        out.append("    }")
        out.append("}")
        # A, α, β, p, q of the border, computed once per step by the whole wavefront (one CSE over
        # all entries) and staged through LDS (only nonzero entries are written: the row loads
        # mask the others to zero), from where every lane loads its own row.  Row r: A[r][0..nb), then its coupling to the
        # bottom of an attached chain (whose Schur complement xc = W[c0][p]/d'_0 is folded into
        # the diagonal), α_r, β_r, p_r, q_r.
        att_bottom = {lane_of[ch[4]]: ch[3][0] for ch in chains}
        ws = nb + 5 + (nb + 5 + 1) % 2       # odd row stride in doubles: conflict-free row loads
        ig = sp.Symbol("ig")
        items = []
        for (i, j) in sorted(filled):
            e = -Jh.get((bcomp[i], bcomp[j]), sp.Integer(0)) + (ig if i == j else 0)
            if e != 0:
                items.append((f"sW[{i * ws + j}]", e))
        for i in range(nb):
            c = bcomp[i]
            if i in att_bottom:
                items.append((f"sW[{i * ws + nb}]", -sp.diff(hrow[c], u[att_bottom[i]])))
            for off, e in ((1, alpha[c]), (2, beta[c]), (3, pvec[c]), (4, qvec[c])):
                if e != 0:
                    items.append((f"sW[{i * ws + nb + off}]", e))
        items = [(lhs, with_reciprocals(e, invm)) for lhs, e in items]
        lines, _ = _emit_block(items, "tw", interleave=True)
        out.append(f"#define SBZ_WSTRIDE {ws}")
        # two modes per wavefront (32 lanes each) for the models that have no four-mode kernel but
        # whose border and chains fit 32 lanes, when two rows' LDS fits 64 KB
        row_bytes = 8 * (14 * n + 7 * NB + len(C_NAMES) + nb * ws + nb + len(chains) + 1)
        w4 = nb <= 16 and len(chains) <= 16
        w2h = (not w4) and nb <= 32 and len(chains) <= 32 and 2 * row_bytes <= 65536
        out.append(f"#define SBZ_WAVE2H_OK {1 if w2h else 0}")
        out.append("struct SbzWb {   // per-lane low-rank data of the border solve")
        out.append("    double p, q, z1, z2;   // p_r, q_r; (A^-1 α)_r, (A^-1 β)_r")
        out.append("    double c11, c12, c21, c22;   // (I - [p q]^T A^-1 [α β])^-1, uniform")
        out.append("};")
        # rows that hold a written entry in each staging column: every other entry loads as zero,
        # so the staging needs no zeros of its own (it may share memory with dead stage vectors)
        colmask = [0] * (nb + 5)
        for lhs, _ in items:
            idx = int(lhs[3:-1])
            colmask[idx % ws] |= 1 << (idx // ws)
        out.append("SBZ_DEV void sbz_wb_w(int lane, const double* __restrict__ B, const double* __restrict__ Cs, double k, double ig, double xc, double* sW, double* row, double& cpl, double& al, double& be, SbzWb& wb) {")
        out.append("    if (lane == 0) {")
        out += ["    " + ln for ln in lines]
        out.append("    }")
        out.append("    __syncthreads();")
        out.append("    const int r = lane < SBZ_NBORD ? lane : SBZ_NBORD - 1;")
        out.append("    const double* w = sW + r * SBZ_WSTRIDE;")
        out.append("    const unsigned long long bit = (lane < SBZ_NBORD) ? (1ull << lane) : 0ull;")
        for j in range(nb):
            out.append(f"    row[{j}] = (0x{colmask[j]:x}ull & bit) ? w[{j}] : 0.0;")
        for off, nm in ((0, "cpl"), (1, "al"), (2, "be"), (3, "wb.p"), (4, "wb.q")):
            out.append(f"    {nm} = (0x{colmask[nb + off]:x}ull & bit) ? w[SBZ_NBORD + {off}] : 0.0;")
        for i in sorted(att_bottom):   # selects on values, never on row positions: row stays in registers
            out.append(f"    row[{i}] -= (lane == {i}) ? cpl * xc : 0.0;")
        out.append("    __syncthreads();   // sW is rewritten by the next step")
        out.append("}")
        # row-parallel LU of A (lane r owns row r, pivots in lane order), level scheduled: pivots
        # whose rows and columns do not depend on each other are eliminated in the same round;
        # pivot rows are broadcast with readlane (no LDS round trip)
        lev = [0] * nb
        for j in range(nb):
            deps = [q for q in range(j) if (j, q) in filled or (q, j) in filled]
            lev[j] = 1 + max((lev[q] for q in deps), default=-1)
        nlev = max(lev) + 1
        ucol = {j: [r for r in range(nb) if j in upper[r]] for j in range(nb)}
        levb = [0] * nb
        for j in reversed(range(nb)):
            levb[j] = 1 + max((levb[c] for c in upper[j]), default=-1)
        nlevb = max(levb) + 1
        self.wave.update(lu_rounds=nlev, bwd_rounds=nlevb)
        # the pivot-row and solution broadcasts go through BC::at<j>: readlane over the wavefront
        # (one mode per wavefront) or a DPP row broadcast within 16 lanes (four modes per wavefront)
        lu = ["template <class BC>", "SBZ_DEV bool sbz_wb_lu(int lane, double* row, double& rpo) {", "    bool ok = true;",
              "    rpo = 0.0;"]
        for L_ in range(nlev):
            piv = [j for j in range(nb) if lev[j] == L_]
            lu.append(f"    // round {L_}: pivots {piv} (independent of each other)")
            for j in piv:   # one pivot at a time keeps its broadcast row short-lived
                lu.append("    {")
                lu.append(f"        const double p{j} = BC::template at<{j}>(row[{j}]);")
                lu.append(f"        const double rp{j} = 1.0 / p{j};")
                lu.append(f"        ok = ok && (p{j} != 0.0) && (p{j} == p{j}) && (fabs(p{j}) < 1e300);")
                lu.append(f"        rpo = (lane == {j}) ? rp{j} : rpo;")
                if lower[j]:
                    m = sum(1 << r for r in lower[j])
                    lu.append(f"        const bool on = (0x{m:x}ull >> lane) & 1ull;")
                    lu.append(f"        const double l = on ? row[{j}] * rp{j} : 0.0;")
                    for c in upper[j]:
                        lu.append(f"        row[{c}] -= l * BC::template at<{j}>(row[{c}]);")
                    lu.append(f"        row[{j}] = on ? l : row[{j}];")
                lu.append("    }")
        lu.append("    return ok;")
        lu.append("}")
        out += lu
        # The factors go to LDS in one of two layouts; both are emitted (template argument SPLIT):
        #   split: strictly-lower rows (sL) and strictly-upper rows plus the reciprocal pivot in
        #     column nb (sU = sL + nb·ws), masked once when stored, so the solves (nine per step)
        #     load them unmasked;
        #   combined: one row per lane, L left of the diagonal, U right of it, 1/pivot in column nb:
        #     half the LDS, the solves mask each entry as they load it.
        # SBZ_LU_SPLIT is the default: split unless it would lower the one-mode kernel's resident
        # workgroups per CU (4, one per SIMD, needs <= 40 KB); the other layout is the kernel
        # variant SBZ_ALGO_WAVE_ALTLU, which the autotuner may choose.
        wl_doubles = 14 * n + 9 * NB + len(C_NAMES) + nb + len(chains) + 1 + (0 if nb * ws <= 8 * n else nb * ws)
        occ = lambda b: min(4, 163840 // b)
        split = occ(8 * (wl_doubles + 2 * nb * ws) + 16) >= occ(8 * (wl_doubles + nb * ws) + 16)
        self.wave["lu_split"] = split
        out.append(f"#define SBZ_LU_SPLIT {1 if split else 0}")
        st = ["template <bool SPLIT>",
              "SBZ_DEV void sbz_wb_store(int lane, const double* row, double rpo, double* sLU) {",
              "    if (lane < SBZ_NBORD) {",
              "        if constexpr (SPLIT) {",
              "            double* lr = sLU + lane * SBZ_WSTRIDE;",
              "            double* ur = sLU + (SBZ_NBORD + lane) * SBZ_WSTRIDE;"]
        for j in range(nb):
            st.append(f"            lr[{j}] = ({j} < lane) ? row[{j}] : 0.0; ur[{j}] = ({j} > lane) ? row[{j}] : 0.0;")
        st += ["            ur[SBZ_NBORD] = rpo;",
               "        } else {",
               "            double* mr = sLU + lane * SBZ_WSTRIDE;"]
        for j in range(nb):
            st.append(f"            mr[{j}] = row[{j}];")
        st += ["            mr[SBZ_NBORD] = rpo;",
               "        }",
               "    }",
               "    __syncthreads();",
               "}"]
        out += st

        # A^-1 applied to NR right-hand sides at once (lane r holds entry r of each), level
        # scheduled: x_j is broadcast with readlane and every lane subtracts its (possibly zero)
        # factor entry times x_j
        def tri_body(split, nr):
            so = []
            if split:
                so += ["    const double* lr = sLU + r * SBZ_WSTRIDE;",
                       "    const double* ur = sLU + (SBZ_NBORD + r) * SBZ_WSTRIDE;",
                       "    const double* sU = sLU + SBZ_NBORD * SBZ_WSTRIDE;"]
                lval = lambda j: f"lr[{j}]"
                uval = lambda j: f"ur[{j}]"
                rpv = lambda j: f"sU[{j} * SBZ_WSTRIDE + SBZ_NBORD]"
                rpr = "ur[SBZ_NBORD]"
            else:
                so += ["    const double* mr = sLU + r * SBZ_WSTRIDE;"]
                lval = lambda j: f"(own && {j} < lane) ? mr[{j}] : 0.0"
                uval = lambda j: f"(own && {j} > lane) ? mr[{j}] : 0.0"
                rpv = lambda j: f"sLU[{j} * SBZ_WSTRIDE + SBZ_NBORD]"
                rpr = "mr[SBZ_NBORD]"
            # every factor entry of a pass is loaded before the pass starts (the empty asm is a
            # compiler barrier for memory accesses), so the rounds never wait on LDS latency
            fwd = [j for j in range(nb) if lower[j]]
            so += [f"    const double l{j} = {lval(j)};" for j in fwd]
            so.append('    asm volatile("" ::: "memory");')
            for L_ in range(nlev):
                piv = [j for j in range(nb) if lev[j] == L_ and lower[j]]
                if not piv:
                    continue
                so.append("    {")
                for q in range(nr):
                    for j in piv:
                        so.append(f"        const double x{q}_{j} = BC::template at<{j}>(b{q});")
                for q in range(nr):
                    so.append(f"        b{q} -= " + " + ".join(f"l{j} * x{q}_{j}" for j in piv) + ";")
                so.append("    }")
            bwd = [j for j in range(nb) if ucol[j]]
            so += [f"    const double u{j} = {uval(j)}, rp{j} = {rpv(j)};" for j in bwd]
            so.append(f"    const double rpr = {rpr};")
            so.append('    asm volatile("" ::: "memory");')
            for L_ in range(nlevb):
                piv = [j for j in range(nb) if levb[j] == L_ and ucol[j]]
                if not piv:
                    continue
                so.append("    {")
                for q in range(nr):
                    for j in piv:
                        so.append(f"        const double x{q}_{j} = BC::template at<{j}>(b{q}) * rp{j};")
                for q in range(nr):
                    so.append(f"        b{q} -= " + " + ".join(f"u{j} * x{q}_{j}" for j in piv) + ";")
                so.append("    }")
            for q in range(nr):
                so.append(f"    b{q} *= rpr;")
            return so

        def tri(name, nr):
            args = ", ".join(f"double& b{q}" for q in range(nr))
            so = ["template <bool SPLIT, class BC>",
                  f"SBZ_DEV void {name}(int lane, const double* __restrict__ sLU, {args}) {{",
                  "    const bool own = lane < SBZ_NBORD;   // lanes past the border use zero factors",
                  "    const int r = own ? lane : SBZ_NBORD - 1;",
                  "    if constexpr (SPLIT) {"]
            so += ["    " + ln for ln in tri_body(True, nr)]
            so.append("    } else {")
            so += ["    " + ln for ln in tri_body(False, nr)]
            so += ["    }", "}"]
            return so
        out += tri("sbz_wb_tri", 1)
        out += tri("sbz_wb_tri2", 2)

        # The same solves with the factors kept in registers (four-modes-per-wavefront kernel, where
        # LDS bounds the resident wavefronts): each lane unpacks its factored row once per step into
        # its strictly-lower and strictly-upper parts (its reciprocal pivot stays in rpo); the
        # reciprocal pivots of the backward pivots, broadcast once per step, for the solves that
        # scale after the broadcast
        bwd_piv = [j for j in range(nb) if ucol[j]]
        up = ["SBZ_DEV void sbz_wb_unpack(int lane, const double* row, double* lrow, double* urow) {"]
        up.append("    const bool own = lane < SBZ_NBORD;   // lanes past the border carry zeros")
        for j in range(nb):
            up.append(f"    lrow[{j}] = (own && {j} < lane) ? row[{j}] : 0.0; urow[{j}] = (own && {j} > lane) ? row[{j}] : 0.0;")
        up.append("}")
        up += ["template <class BC>", "SBZ_DEV void sbz_wb_rpb(double rpo, double* rpb) {"]
        up += [f"    rpb[{j}] = BC::template at<{j}>(rpo);" for j in bwd_piv]
        up.append("}")
        out += up

        def tri_r(name, nr):
            args = ", ".join(f"double& b{q}" for q in range(nr))
            so = ["template <class BC, bool OWNSCALE>",
                  f"SBZ_DEV void {name}(const double* lrow, const double* urow, const double* rpb, double rpo, {args}) {{"]
            for L_ in range(nlev):
                piv = [j for j in range(nb) if lev[j] == L_ and lower[j]]
                if not piv:
                    continue
                so.append("    {")
                for q in range(nr):
                    for j in piv:
                        so.append(f"        const double x{q}_{j} = BC::template at<{j}>(b{q});")
                for q in range(nr):
                    so.append(f"        b{q} -= " + " + ".join(f"lrow[{j}] * x{q}_{j}" for j in piv) + ";")
                so.append("    }")
            for L_ in range(nlevb):
                piv = [j for j in range(nb) if levb[j] == L_ and ucol[j]]
                if not piv:
                    continue
                # x_j = b_j / p_j: OWNSCALE scales every lane's own entry by its reciprocal pivot
                # and broadcasts lane j's product (no registers for the broadcast pivots); else the
                # broadcast b_j is scaled by rpb[j].  Both form the same product.
                so.append("    {")
                so.append("        double " + ", ".join(f"x{q}_{j}" for q in range(nr) for j in piv) + ";")
                so.append("        if constexpr (OWNSCALE) {")
                for q in range(nr):
                    so.append(f"            const double s{q} = b{q} * rpo;")
                for q in range(nr):
                    for j in piv:
                        so.append(f"            x{q}_{j} = BC::template at<{j}>(s{q});")
                so.append("        } else {")
                for q in range(nr):
                    for j in piv:
                        so.append(f"            x{q}_{j} = BC::template at<{j}>(b{q}) * rpb[{j}];")
                so.append("        }")
                for q in range(nr):
                    so.append(f"        b{q} -= " + " + ".join(f"urow[{j}] * x{q}_{j}" for j in piv) + ";")
                so.append("    }")
            for q in range(nr):
                so.append(f"    b{q} *= rpo;")
            so.append("}")
            return so
        out += tri_r("sbz_wb_tri_r", 1)
        out += tri_r("sbz_wb_tri2_r", 2)
        return out


def wave_structure(M):
    """Partition of the unknowns for the wavefront-per-mode kernel.

    Chains: the interiors l = 3..L of every hierarchy (photon F and G, massless ν, each massive-ν
    momentum bin).  Each is tridiagonal and couples to the rest only through its l = 2 member (its
    attachment).  Border: every other unknown (Φ, l ≤ 2 multipoles, fluids).
    Returns (chains, border) with chains = [(kind, L, bin, [comp for l = 3..L], attach_comp)];
    kind 0: photon (damped by κ̇), 1: massless ν, 2: massive ν."""
    Lt = layout(M)
    chains = []
    def add(kind, base, L, b=0):
        if L < 3:      # no interior: the whole hierarchy stays in the border
            return
        chains.append((kind, L, b, [base + l for l in range(3, L + 1)], base + 2))
    add(0, Lt["F"], M.lmax_g)
    add(0, Lt["G"], M.lmax_g)
    if M.has_nu:
        add(1, Lt["N"], M.lmax_nu)
    if M.has_h:
        for i in range(M.nx):
            add(2, Lt["H"] + i * (M.lmax_h + 1), M.lmax_h, i)
    in_chain = {c for ch in chains for c in ch[3]}
    border = [i for i in range(Lt["n"]) if i not in in_chain]
    return chains, border


def expand_includes(path, seen=None):
    """Inline local #include "..." directives (hiprtc gets one self-contained source)."""
    path = Path(path)
    seen = set() if seen is None else seen
    out = []
    for line in path.read_text().splitlines():
        m = re.match(r'\s*#include\s+"([^"]+)"', line)
        if m:
            inc = (path.parent / m.group(1)).resolve()
            if inc not in seen:
                seen.add(inc)
                out.append(expand_includes(inc, seen))
            continue
        if line.strip() == "#pragma once":
            continue
        out.append(line)
    return "\n".join(out)


def kernel_source(M, fence=0):
    """Full self-contained HIP source (emitted model + generic kernel) and its metadata."""
    em = Emitted(M, fence)
    kern = expand_includes(CSRC / "sbz_kernel.hip")
    src = kern.replace("// @@SBZ_MODEL@@", em.src)
    em.full_src = src
    em.full_hash = hashlib.sha1(src.encode()).hexdigest()[:16]
    return em
