"""Mechanism compiler: Cantera-style YAML -> device constants + HIP source.

Replaces the reference's Cantera-based preprocessing
(``src/janc_v2/preprocess/load.py:8-135`` and ``preprocess/nondim.py:14-25``)
for the hot path.  Two products:

* ``Mechanism.const_array()`` - the flat float64 array that fills the device
  ``MechConst`` struct (layout declared in ``csrc/janc_types.hpp``); it carries
  every *value* (A, b, Ea/R, third-body efficiencies, NASA-7 polynomials, molar
  masses, nondimensional scales), so a nondim configuration changes only data.
* ``Mechanism.generate_source()`` - straight-line HIP/C++ for the reaction
  rates and the analytic Jacobian.  Only the *structure* (stoichiometry, which
  reactions are three-body, which exponents are integers) is compiled in, so
  every zero entry of the 19x8 stoichiometric matrix disappears at compile
  time.

The YAML is read directly (no Cantera): Cantera's SI conversion of A
(length^3/quantity -> m^3/kmol per order) and Ea (-> J/kmol), and its element
weights (H 1.008, O 15.999, N 14.007), are applied here.
"""
import hashlib
import math
import os

import numpy as np
import yaml

RU = 8.314463                 # preprocess/load.py:6 (Rg)
P_ATM = 101325.0              # model/reaction_model.py:113
MAX_POLY = (5, 6, 6, 5)       # cp, h, h_chem, s coefficient counts
_ELEMENT_WEIGHT = {"H": 1.008, "O": 15.999, "N": 14.007, "C": 12.011,
                   "Ar": 39.95, "He": 4.002602}
_LENGTH_M = {"cm": 0.01, "m": 1.0, "mm": 0.001}
_QUANTITY_KMOL = {"mol": 0.001, "kmol": 1.0}
_ENERGY_J_PER_KMOL = {"cal/mol": 4184.0, "kcal/mol": 4.184e6, "J/mol": 1000.0,
                      "kJ/mol": 1.0e6, "J/kmol": 1.0}

DEFAULT_MECH = os.path.join(os.path.dirname(__file__), "data", "9sp-19r-H2-Air.yaml")


class NondimScales:
    """Reference scales of ``preprocess/nondim.py:3-25`` (all 1 when unset)."""

    def __init__(self, cfg=None):
        cfg = cfg or {"P0": 1.0, "T0": 1.0, "R0": 1.0, "x0": 1.0}
        self.P0, self.T0, self.R0, self.x0 = (float(cfg[k]) for k in ("P0", "T0", "R0", "x0"))
        self.rho0 = self.P0 / (self.R0 * self.T0)
        self.M0 = RU / self.R0
        self.e0 = self.P0 / self.rho0
        self.u0 = math.sqrt(self.P0 / self.rho0)
        self.t0 = self.x0 / self.u0


class Reaction:
    __slots__ = ("equation", "reactants", "products", "three_body", "A_cgs", "b",
                 "Ea", "efficiencies")

    def __init__(self, entry, species):
        self.equation = entry["equation"]
        sep = "<=>" if "<=>" in self.equation else "="
        lhs, rhs = self.equation.split(sep)
        self.reactants, tb_l = _side(lhs, species)
        self.products, tb_r = _side(rhs, species)
        self.three_body = entry.get("type") == "three-body"
        if self.three_body != (tb_l and tb_r):
            raise ValueError("inconsistent third body in %r" % self.equation)
        if entry.get("type") not in (None, "three-body", "elementary"):
            raise NotImplementedError("reaction type %r" % entry.get("type"))
        rc = entry["rate-constant"]
        self.A_cgs, self.b, self.Ea = float(rc["A"]), float(rc["b"]), float(rc["Ea"])
        self.efficiencies = dict(entry.get("efficiencies", {}))


def _side(text, species):
    out, third = {}, False
    for term in text.split(" + "):
        term = term.strip()
        if term == "M":
            third = True
            continue
        bits = term.split()
        coef, name = (float(bits[0]), bits[1]) if len(bits) == 2 else (1.0, bits[0])
        k = species.index(name)
        out[k] = out.get(k, 0.0) + coef
    return out, third


class Mechanism:
    """Parsed mechanism + nondimensional constants for one solver context."""

    def __init__(self, path=DEFAULT_MECH, nondim_config=None):
        if os.path.splitext(path)[1].lower() != ".yaml":
            raise AssertionError("janc only read mech file with .yaml format")
        if not os.path.isfile(path):
            raise FileNotFoundError("No mechanism file detected in the specified directory.")
        self.path = path
        with open(path) as fh:
            self.doc = yaml.safe_load(fh)
        self.nd = NondimScales(nondim_config)
        units = self.doc.get("units", {})
        self._conc = _LENGTH_M[units.get("length", "m")] ** 3 / \
            _QUANTITY_KMOL[units.get("quantity", "kmol")]
        self._ea = _ENERGY_J_PER_KMOL[units.get("activation-energy", "J/kmol")]
        self.species = list(self.doc["phases"][0]["species"])
        self.ns = len(self.species)
        self.reactions = [Reaction(r, self.species) for r in self.doc["reactions"]]
        self.nrx = len(self.reactions)
        active = [any(k in r.reactants or k in r.products for r in self.reactions)
                  for k in range(self.ns)]
        self.na = sum(active)
        if not all(active[:self.na]) or any(active[self.na:]):
            raise AssertionError("Inert species must be the last elements in species_list")
        self._thermo()
        self._kinetics()

    # ---------------- thermo (load.py:91-135) ----------------
    def _thermo(self):
        nd = self.nd
        by_name = {s["name"]: s for s in self.doc["species"]}
        mw, tcr, lo, hi = [], [], [], []
        for name in self.species:
            sp = by_name[name]
            th = sp["thermo"]
            if th["model"] != "NASA7":
                raise NotImplementedError(th["model"])
            tcr.append(float(th["temperature-ranges"][1]))
            lo.append([float(v) for v in th["data"][0]])
            hi.append([float(v) for v in th["data"][1]])
            mw.append(sum(_ELEMENT_WEIGHT[e] * n for e, n in sp["composition"].items()))
        self.molecular_weight = np.array(mw)
        self.species_M = self.molecular_weight / (1000 * nd.M0)
        self.Tcr_all = np.array(tcr) / nd.T0
        if not np.all(self.Tcr_all == self.Tcr_all[0]):
            raise NotImplementedError("species with different NASA-7 switch temperatures")
        self.Tcr = float(self.Tcr_all[0])
        sc = np.array([1, nd.T0, nd.T0 ** 2, nd.T0 ** 3, nd.T0 ** 4, 1 / nd.T0, 1])
        self.nasa = np.stack([np.array(lo) * sc, np.array(hi) * sc])  # (2, ns, 7)
        a = self.nasa
        M = self.species_M[None, :, None]
        # polynomials, highest power first (jnp.polyval order)
        self.cp = a[:, :, 4::-1] / M                                       # a5..a1
        hc = np.stack([a[..., 4] / 5, a[..., 3] / 4, a[..., 2] / 3, a[..., 1] / 2,
                       a[..., 0], a[..., 5]], axis=-1)                     # h/R
        self.h_chem = hc
        self.h = hc / M
        self.s = np.stack([a[..., 4] / 4, a[..., 3] / 3, a[..., 2] / 2, a[..., 1],
                           a[..., 6]], axis=-1)
        self.logc = a[..., 0]

    def injection_Y(self, ratio=1.0):
        """Injected mass fractions of the rotating-detonation example
        (janc_basic_example2.ipynb): H2 and O2 of H2-air at equivalence
        ``ratio`` (air 0.232 O2 by mass), 1e-20 for the other reacting
        species; the last species (N2) by closure."""
        yH2 = (ratio / 8 * 0.232) / (1 + ratio / 8 * 0.232)
        yO2 = 0.232 / (1 + ratio / 8 * 0.232)
        Y = np.full(self.ns - 1, 1e-20)
        Y[self.species.index("H2")] = yH2
        Y[self.species.index("O2")] = yO2
        return Y

    # ---------------- kinetics (load.py:8-88) ----------------
    def _kinetics(self):
        nd, ns = self.nd, self.ns
        A, B, EaR, eff, vsum = [], [], [], [], []
        for r in self.reactions:
            order = sum(r.reactants.values()) + (1 if r.three_body else 0)
            a_si = r.A_cgs * self._conc ** (order - 1)
            a_mol = a_si * ((1e-3) ** (order - 1))
            vf_sum = sum(r.reactants.values())
            vb_sum = sum(r.products.values())
            A.append((nd.t0 / nd.rho0) * (nd.T0 ** r.b) * ((nd.rho0 / nd.M0) ** vf_sum)
                     * nd.M0 * a_mol)
            B.append(r.b)
            EaR.append((r.Ea * self._ea / 1000) / (nd.e0 * nd.M0))
            row = np.zeros(ns)
            if r.three_body:
                for k, name in enumerate(self.species):
                    row[k] = float(r.efficiencies.get(name, 1.0))
            eff.append(row * (nd.rho0 / nd.M0))
            vsum.append(vb_sum - vf_sum)
        self.A, self.B, self.EaR = np.array(A), np.array(B), np.array(EaR)
        self.eff = np.array(eff)
        self.vsum = np.array(vsum)
        self.effM = self.eff[:, :self.na] / self.species_M[None, :self.na]
        for r in self.reactions:
            for k in set(r.reactants) | set(r.products):
                d = r.products.get(k, 0.0) - r.reactants.get(k, 0.0)
                if abs(d) not in (0.0, 1.0, 2.0):
                    raise NotImplementedError("stoichiometric difference %g" % d)
            for v in list(r.reactants.values()) + list(r.products.values()):
                if v not in (1.0, 2.0):
                    raise NotImplementedError("stoichiometric coefficient %g" % v)
            if r.three_body is False and r.efficiencies:
                raise ValueError("efficiencies on a non three-body reaction")

    # ---------------- device constants ----------------
    def layout(self):
        """(name, shape) of every MechConst field, in declaration order."""
        ns, na, nrx = self.ns, self.na, self.nrx
        return [("R", (ns,)), ("M", (ns,)), ("Tcr", (1,)),
                ("cp", (2, ns, 5)), ("h", (2, ns, 6)), ("hc", (2, na, 6)),
                ("s", (2, na, 5)), ("logc", (2, na)),
                ("A", (nrx,)), ("B", (nrx,)), ("EaR", (nrx,)), ("eff", (nrx, ns)),
                ("effM", (nrx, na)), ("Pref", (1,)), ("T0", (1,))]

    def const_array(self):
        vals = {
            "R": 1.0 / self.species_M, "M": self.species_M, "Tcr": [self.Tcr],
            "cp": self.cp, "h": self.h, "hc": self.h_chem[:, :self.na],
            "s": self.s[:, :self.na], "logc": self.logc[:, :self.na],
            "A": self.A, "B": self.B, "EaR": self.EaR, "eff": self.eff,
            "effM": self.effM, "Pref": [P_ATM / self.nd.P0], "T0": [self.nd.T0],
        }
        parts = []
        for name, shape in self.layout():
            arr = np.asarray(vals[name], dtype=np.float64)
            assert arr.size == int(np.prod(shape)), name
            parts.append(arr.reshape(-1))
        return np.ascontiguousarray(np.concatenate(parts))

    def structure_digest(self):
        """Hash of everything compiled into the generated source."""
        h = hashlib.sha256()
        h.update(repr((self.species, self.na)).encode())
        for r in self.reactions:
            h.update(repr((sorted(r.reactants.items()), sorted(r.products.items()),
                           r.three_body, r.b)).encode())
        return h.hexdigest()[:16]

    # ---------------- code generation ----------------
    def generate_source(self):
        return _generate(self)


def _fmt(x):
    return repr(float(x))


def _power_expr(k, n, base="Xc"):
    return " * ".join(["%s[%d]" % (base, k)] * int(n))


FWD_GROUP = 2      # reactions between scheduling barriers (forward rates)


def _generate(m):
    na, ns = m.na, m.ns
    L = []
    e = L.append
    e("// Generated by janc_amd/mechanism.py from %s -- do not edit." % os.path.basename(m.path))
    e("// Structure digest %s: species %s" % (m.structure_digest(), " ".join(m.species)))
    e("#pragma once")
    e("")
    e("// Net production rates w_k and dw_k/d(rho Y_n) for one cell, following")
    e("// model/reaction_model.py:91-130 (reactionConstant_i summed over i).")
    e("// X: unclipped rho*Y/M (all species); Xc: clip(X, 1e-50) for active species;")
    e("// invMX[n] = 1/(M_n*Xc_n); eg: exp(g) of the Gibbs s/R - h/(RT) of the active")
    e("// species, so exp(-sum_k dnu_k g_k) is a quotient of eg products.")
    e("// JT: the Jacobian's storage type.  Rates, exponentials and every d_n are")
    e("// FP64; JT = float rounds each d_n once and accumulates J in FP32 (the")
    e("// mixed-precision form: half the registers of an FP64 J).")
    e("template <typename JT>")
    e("__host__ __device__ __forceinline__ void mech_rates_jacobian(")
    e("    const MechConst& __restrict__ c, const double T, const double invT, const double logT,")
    e("    const double (&X)[JH_NS], const double (&Xc)[JH_NA], const double (&invMX)[JH_NA],")
    e("    const double (&eg)[JH_NA], double (&w)[JH_NA],")
    e("    JT (&J)[JH_NA][JH_NA])")
    e("{")
    e("  const double Pf = c.Pref * invT;   // 101325/P0/T")
    e("  const double iPf = jh::rcp(Pf);")
    e("#pragma unroll")
    e("  for (int k = 0; k < JH_NA; ++k) { w[k] = 0.0;")
    e("#pragma unroll")
    e("    for (int n = 0; n < JH_NA; ++n) J[k][n] = JT(0); }")
    # third-body concentrations first: X is not needed after them
    for i, r in enumerate(m.reactions):
        if r.three_body:
            terms = " + ".join("c.eff[%d][%d] * X[%d]" % (i, k, k) for k in range(ns))
            e("  const double cM%d = %s;" % (i, terms))
    for i, r in enumerate(m.reactions):
        # reactions in pairs between scheduling barriers: the scheduler may
        # overlap two reactions' exp and reciprocal chains (measured 5 %
        # faster than one at a time; three at a time gained nothing more),
        # while more at once would raise the live register count past what
        # two resident waves per SIMD allow
        if i % FWD_GROUP == 0:
            e("#if defined(__HIP_DEVICE_COMPILE__)")
            e("  __builtin_amdgcn_sched_barrier(0);")
            e("#endif")
        e("  { // R%d: %s" % (i, r.equation))
        # kf_i = A * T^B * exp(-Ea/R / T)
        if r.b == 0.0:
            tb = None
        elif r.b == 1.0:
            tb = "T"
        elif r.b == 2.0:
            tb = "(T * T)"
        elif r.b == -1.0:
            tb = "invT"
        elif r.b == -2.0:
            tb = "(invT * invT)"
        else:
            tb = "exp(c.B[%d] * logT)" % i
        kfi = "c.A[%d]" % i
        if tb:
            kfi = "(%s * %s)" % (kfi, tb)
        if r.Ea != 0.0:
            kfi = "%s * exp(-c.EaR[%d] * invT)" % (kfi, i)
        e("    const double kfi = %s;" % kfi)
        pf = " * ".join(_power_expr(k, n) for k, n in sorted(r.reactants.items()))
        pb = " * ".join(_power_expr(k, n) for k, n in sorted(r.products.items()))
        e("    const double kf = kfi * (%s);" % pf)
        # exp(-sum_k dnu_k g_k) = prod(eg of net reactants) / prod(eg of net
        # products): one reciprocal per reaction instead of a live ieg table
        num, den = [], []
        for k in range(na):
            d = r.products.get(k, 0.0) - r.reactants.get(k, 0.0)
            if d != 0.0:
                (den if d > 0 else num).extend(["eg[%d]" % k] * int(abs(d)))
        if num and den:
            eqf = "(%s) * jh::rcp(%s)" % (" * ".join(num), " * ".join(den))
        elif den:
            eqf = "jh::rcp(%s)" % " * ".join(den)
        else:
            eqf = " * ".join(num) if num else "1.0"
        vs = sum(r.products.values()) - sum(r.reactants.values())
        # kb = kf_i / (exp(sum dnu g) Pf^vsum) * prod X^nu_b (reaction_model.py:113)
        if vs == 0.0:
            eq = "(%s)" % eqf
        elif vs == 1.0:
            eq = "(%s * iPf)" % eqf
        elif vs == -1.0:
            eq = "(%s * Pf)" % eqf
        else:
            eq = "(%s * pow(Pf, %s))" % (eqf, _fmt(-vs))
        e("    const double kb = kfi * %s * (%s);" % (eq, pb))
        if r.three_body:
            e("    const double cM = cM%d;" % i)
        e("    const double q = kf - kb;")
        cq = "cM * q" if r.three_body else "q"
        # w_k += dnu_k * cM * q
        dnu = {}
        for k in range(na):
            d = r.products.get(k, 0.0) - r.reactants.get(k, 0.0)
            if d != 0.0:
                dnu[k] = d
        if r.three_body:
            e("    const double cq = cM * q;")
            cq = "cq"
        for k, d in dnu.items():
            e("    w[%d] %s %s;" % (k, "+=" if d > 0 else "-=",
                                  cq if abs(d) == 1.0 else "2.0 * " + cq))
        # d_n = q*effM_n + (invMX_n*cM)*(vf_n kf - vb_n kb)
        involved = sorted(set(r.reactants) | set(r.products))
        dn = {}
        for n in range(na):
            parts = []
            vf, vb = r.reactants.get(n, 0.0), r.products.get(n, 0.0)
            if r.three_body:
                parts.append("q * c.effM[%d][%d]" % (i, n))
            if n in involved:
                if vf and vb:
                    diff = "(%s - %s)" % (_scaled(vf, "kf"), _scaled(vb, "kb"))
                elif vf:
                    diff = _scaled(vf, "kf")
                else:
                    diff = "(-%s)" % _scaled(vb, "kb")
                pre = "(invMX[%d] * cM)" % n if r.three_body else "invMX[%d]" % n
                parts.append("%s * %s" % (pre, diff))
            if parts:
                dn[n] = " + ".join(parts)
        for n, expr in dn.items():
            e("    const JT d%d = JT(%s);" % (n, expr))
        for k, d in dnu.items():
            for n in dn:
                op = "+=" if d > 0 else "-="
                val = "d%d" % n if abs(d) == 1.0 else "JT(2.0) * d%d" % n
                e("    J[%d][%d] %s %s;" % (k, n, op, val))
        e("  }")
    e("#pragma unroll")
    e("  for (int k = 0; k < JH_NA; ++k) { w[k] = c.M[k] * w[k];")
    e("#pragma unroll")
    e("    for (int n = 0; n < JH_NA; ++n) J[k][n] = JT(c.M[k]) * J[k][n]; }")
    e("}")
    return "\n".join(L) + "\n\n" + _generate_vjp(m)


def _generate_vjp(m):
    """Reverse mode of mech_rates_jacobian for the chemistry adjoint.

    The cotangent of the (I/dt - J) x = w solve is w_bar = lambda and
    J_bar = lambda x^T (rank one), so per reaction both contractions reduce
    to the scalar s_i = sum_k dnu_ik M_k lambda_k.  Each reaction's forward
    quantities are recomputed and its reverse applied at once (reactions add
    independently), accumulating X_bar (all species), Xc_bar, invMX_bar and
    T_bar.  gp[k] = d g_k / dT = (h_k/R)/T^2.
    """
    na, ns = m.na, m.ns
    L = []
    e = L.append
    e("__host__ __device__ __forceinline__ void mech_rates_vjp(")
    e("    const MechConst& __restrict__ c, const double T, const double invT, const double logT,")
    e("    const double (&X)[JH_NS], const double (&Xc)[JH_NA], const double (&invMX)[JH_NA],")
    e("    const double (&eg)[JH_NA], const double (&gp)[JH_NA],")
    e("    const double (&lam)[JH_NA], const double (&x)[JH_NA],")
    e("    double (&Xb)[JH_NS], double (&Xcb)[JH_NA], double (&iMXb)[JH_NA], double& Tb)")
    e("{")
    e("  const double Pf = c.Pref * invT;")
    e("  const double iPf = jh::rcp(Pf);")
    e("  double ml[JH_NA];")
    e("#pragma unroll")
    e("  for (int k = 0; k < JH_NA; ++k) ml[k] = c.M[k] * lam[k];")
    # third-body concentrations first: X is not needed after them
    for i, r in enumerate(m.reactions):
        if r.three_body:
            e("  const double cM%d = %s;" % (i, " + ".join("c.eff[%d][%d] * X[%d]" % (i, k, k)
                                                          for k in range(ns))))
    for i, r in enumerate(m.reactions):
        # one reaction at a time (as in the forward): interleaved reactions
        # hold more live values than two waves per SIMD allow
        e("#if defined(__HIP_DEVICE_COMPILE__)")
        e("  __builtin_amdgcn_sched_barrier(0);")
        e("#endif")
        e("  { // R%d: %s" % (i, r.equation))
        if r.b == 0.0:
            tb, dl = None, "0.0"
        elif r.b == 1.0:
            tb, dl = "T", "invT"
        elif r.b == 2.0:
            tb, dl = "(T * T)", "2.0 * invT"
        elif r.b == -1.0:
            tb, dl = "invT", "-invT"
        elif r.b == -2.0:
            tb, dl = "(invT * invT)", "-2.0 * invT"
        else:
            tb, dl = "exp(c.B[%d] * logT)" % i, "c.B[%d] * invT" % i
        kfi = "c.A[%d]" % i
        if tb:
            kfi = "(%s * %s)" % (kfi, tb)
        if r.Ea != 0.0:
            kfi = "%s * exp(-c.EaR[%d] * invT)" % (kfi, i)
            dl = "(%s + c.EaR[%d] * invT * invT)" % (dl, i)
        e("    const double kfi = %s;" % kfi)
        e("    const double dlk = %s;   // d ln kf_i / dT" % dl)
        pf_terms = []
        for k, n in sorted(r.reactants.items()):
            pf_terms += ["Xc[%d]" % k] * int(n)
        pb_terms = []
        for k, n in sorted(r.products.items()):
            pb_terms += ["Xc[%d]" % k] * int(n)
        e("    const double PF = %s;" % " * ".join(pf_terms))
        e("    const double PB = %s;" % " * ".join(pb_terms))
        num, den, dgsum = [], [], []
        for k in range(na):
            d = r.products.get(k, 0.0) - r.reactants.get(k, 0.0)
            if d != 0.0:
                (den if d > 0 else num).extend(["eg[%d]" % k] * int(abs(d)))
                dgsum.append("%s%s gp[%d]" % ("- " if d > 0 else "+ ", "" if abs(d) == 1 else "2.0 *", k))
        vs = sum(r.products.values()) - sum(r.reactants.values())
        # exp(-sum dnu g) as a quotient of eg products (one reciprocal per
        # reaction instead of a live table of 1/eg)
        if num and den:
            eq = "(%s) * jh::rcp(%s)" % (" * ".join(num), " * ".join(den))
        elif den:
            eq = "jh::rcp(%s)" % " * ".join(den)
        else:
            eq = " * ".join(num) if num else "1.0"
        if vs == 1.0:
            eq += " * iPf"
        elif vs == -1.0:
            eq += " * Pf"
        elif vs != 0.0:
            eq += " * pow(Pf, %s)" % _fmt(-vs)
        e("    const double EQ = %s;" % eq)
        e("    const double kf = kfi * PF;")
        e("    const double kb = kfi * EQ * PB;")
        dls = " ".join(dgsum)
        e("    const double dlb = dlk %s + %s * invT;   // d ln kb / dT at fixed X" %
          (dls, _fmt(vs)))
        if r.three_body:
            e("    const double q = kf - kb;")
        dnu = {}
        for k in range(na):
            d = r.products.get(k, 0.0) - r.reactants.get(k, 0.0)
            if d != 0.0:
                dnu[k] = d
        terms = []
        for k, d in dnu.items():
            terms.append(("+ " if d > 0 else "- ") + ("" if abs(d) == 1 else "2.0 * ") + "ml[%d]" % k)
        sexpr = " ".join(terms).lstrip("+ ")
        if sexpr.startswith("- "):
            sexpr = "-" + sexpr[2:]
        e("    const double s = %s;" % sexpr)
        involved = sorted(set(r.reactants) | set(r.products))
        if r.three_body:
            e("    const double cM = cM%d;" % i)
            e("    double sx = 0.0;   // sum_n x_n effM_in")
            for n in range(na):
                e("    sx += x[%d] * c.effM[%d][%d];" % (n, i, n))
            e("    double qb = s * cM + s * sx;")
            e("    double cMb = s * q;")
            cexpr = "cM"
        else:
            e("    double qb = s;")
            cexpr = None
        e("    double kfb = qb, kbb = -qb;")
        for n in involved:
            vf, vb = r.reactants.get(n, 0.0), r.products.get(n, 0.0)
            diff_terms = []
            if vf:
                diff_terms.append(_scaled(vf, "kf"))
            if vb:
                diff_terms.append("-" + _scaled(vb, "kb"))
            diff = " ".join(diff_terms) if len(diff_terms) == 1 else \
                "%s - %s" % (diff_terms[0], diff_terms[1][1:])
            e("    { const double Db = s * x[%d];" % n)
            if r.three_body:
                e("      const double cf = invMX[%d] * cM;" % n)
                e("      cMb += Db * invMX[%d] * (%s);" % (n, diff))
                e("      iMXb[%d] += Db * cM * (%s);" % (n, diff))
            else:
                e("      const double cf = invMX[%d];" % n)
                e("      iMXb[%d] += Db * (%s);" % (n, diff))
            if vf:
                e("      kfb += Db * cf * %s;" % _fmt(vf))
            if vb:
                e("      kbb -= Db * cf * %s;" % _fmt(vb))
            e("    }")
        if r.three_body:
            for k in range(ns):
                e("    Xb[%d] += cMb * c.eff[%d][%d];" % (k, i, k))
        # kf = kfi * PF, kb = kfi * EQ * PB
        e("    Tb += kfb * kf * dlk + kbb * kb * dlb;")
        def prod_grad(terms, coef):
            # d(prod)/d Xc[k] for each distinct k, times coef
            from collections import Counter
            cnt = Counter(terms)
            for t, mult in cnt.items():
                k = int(t[3:-1])
                others = list(terms)
                others.remove(t)
                rest = " * ".join(others) if others else "1.0"
                fac2 = "" if mult == 1 else "%s * " % _fmt(mult)
                e("    Xcb[%d] += %s%s * %s;" % (k, fac2, coef, rest))
        prod_grad(pf_terms, "kfb * kfi")
        prod_grad(pb_terms, "kbb * kfi * EQ")
        e("  }")
    e("}")
    return "\n".join(L) + "\n"


def generate_dims(m):
    return "\n".join([
        "// Generated by janc_amd/mechanism.py from %s -- do not edit." % os.path.basename(m.path),
        "#pragma once",
        "#define JH_NS %d" % m.ns,
        "#define JH_NA %d" % m.na,
        "#define JH_NRX %d" % m.nrx,
        '#define JH_MECH_DIGEST "%s"' % m.structure_digest(), ""])


def _scaled(v, name):
    return name if v == 1.0 else "%s * %s" % (_fmt(v), name)


def main(argv=None):
    import argparse
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("yaml", nargs="?", default=DEFAULT_MECH)
    p.add_argument("-o", "--out", default=os.path.join(os.path.dirname(__file__), "csrc",
                                                        "mech_gen.inc"))
    a = p.parse_args(argv)
    write_generated(Mechanism(a.yaml), os.path.dirname(a.out))
    print("wrote", a.out)


def write_generated(mech, outdir):
    """Write mech_dims.h and mech_gen.inc; return True if either changed."""
    changed = False
    for name, text in (("mech_dims.h", generate_dims(mech)),
                       ("mech_gen.inc", mech.generate_source())):
        path = os.path.join(outdir, name)
        old = open(path).read() if os.path.exists(path) else None
        if old != text:
            with open(path, "w") as fh:
                fh.write(text)
            changed = True
    return changed


if __name__ == "__main__":
    main()
