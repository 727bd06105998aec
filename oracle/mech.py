"""Oracle restatement of the mechanism preprocessing (test infrastructure only).

Follows ``src/janc_v2/preprocess/load.py:8-135`` and
``src/janc_v2/preprocess/nondim.py:1-25`` of the reference.  The reference
reads the mechanism through Cantera; Cantera is absent here, so this module
reads the same YAML file directly and applies Cantera's unit conversions
(cm/mol -> m/kmol for A, cal/mol -> J/kmol for Ea) and Cantera's element
weights (H 1.008, O 15.999, N 14.007).
"""
import math

import numpy as np
import yaml

RG = 8.314463            # preprocess/load.py:6, preprocess/nondim.py:3
ATOMIC_WEIGHT = {"H": 1.008, "O": 15.999, "N": 14.007, "C": 12.011, "AR": 39.95,
                 "HE": 4.002602}
_LEN_TO_M = {"cm": 1e-2, "m": 1.0, "mm": 1e-3}
_QTY_TO_KMOL = {"mol": 1e-3, "kmol": 1.0, "molec": 1.0 / 6.02214076e26}
_EA_TO_J_PER_KMOL = {"cal/mol": 4184.0, "kcal/mol": 4184.0e3, "J/mol": 1e3,
                     "kJ/mol": 1e6, "J/kmol": 1.0}


class Nondim:
    """Reference scales, ``preprocess/nondim.py:3-25``."""

    def __init__(self, cfg=None):
        if cfg is None:
            self.P0, self.T0, self.R0, self.x0 = 1.0, 1.0, 1.0, 1.0
        else:
            self.P0, self.T0 = cfg["P0"], cfg["T0"]
            self.R0, self.x0 = cfg["R0"], cfg["x0"]
        self.rho0 = self.P0 / (self.R0 * self.T0)
        self.M0 = RG / self.R0
        self.e0 = self.P0 / self.rho0
        self.u0 = math.sqrt(self.P0 / self.rho0)
        self.t0 = self.x0 / self.u0


def _parse_equation(eq):
    """'H + O2 + M <=> HO2 + M' -> (reactants dict, products dict, third_body)."""
    lhs, rhs = eq.split("<=>") if "<=>" in eq else eq.split("=>")

    def side(s):
        out, tb = {}, False
        for tok in s.split(" + "):
            tok = tok.strip()
            if tok == "M":
                tb = True
                continue
            parts = tok.split(" ")
            if len(parts) == 2:
                coef, name = float(parts[0]), parts[1]
            else:
                coef, name = 1.0, parts[0]
            out[name] = out.get(name, 0.0) + coef
        return out, tb

    r, tbr = side(lhs)
    p, _ = side(rhs)
    return r, p, tbr


def load_mechanism(path, nondim_config=None):
    """Return the reference's ``ReactionParams`` plus the thermo coefficient set.

    ReactionParams follows ``read_reaction_mechanism`` (load.py:8-88); thermo
    coefficients follow ``get_cantera_coeffs`` (load.py:91-135).  Arrays carry
    no trailing broadcast axes (the reference adds ``(2,3)`` for 2-D).
    """
    nd = Nondim(nondim_config)
    with open(path) as f:
        doc = yaml.safe_load(f)
    units = doc.get("units", {})
    len_f = _LEN_TO_M[units.get("length", "m")]
    qty_f = _QTY_TO_KMOL[units.get("quantity", "kmol")]
    ea_f = _EA_TO_J_PER_KMOL[units.get("activation-energy", "J/kmol")]
    # Cantera converts concentrations to kmol/m^3: one (length^3/quantity) unit
    conc_f = len_f ** 3 / qty_f

    species = list(doc["phases"][0]["species"])
    sp_by_name = {s["name"]: s for s in doc["species"]}
    ns = len(species)
    rxns = doc["reactions"]
    nrx = len(rxns)

    vf_full = np.zeros((nrx, ns))
    vb_full = np.zeros((nrx, ns))
    A, B, Ea = [], [], []
    third = np.zeros((nrx, ns))
    is_tb = np.zeros(nrx)
    for i, r in enumerate(rxns):
        reac, prod, tb = _parse_equation(r["equation"])
        for k, c in reac.items():
            vf_full[i, species.index(k)] += c
        for k, c in prod.items():
            vb_full[i, species.index(k)] += c
        rate = r["rate-constant"]
        # load.py:40 reaction_order = sum(reactants); +1 for three-body (:41-43)
        order = sum(reac.values())
        if r.get("type") == "three-body":
            eff = r.get("efficiencies", {})
            order += 1
            for j, sname in enumerate(species):
                third[i, j] = eff.get(sname, 1.0)      # load.py:44-45
        is_tb[i] = 1.0 * np.any(third[i, :] > 0)      # load.py:46
        a_si = rate["A"] * conc_f ** (order - 1)       # Cantera's SI value
        A.append(a_si * ((1e-3) ** (order - 1)))       # load.py:47
        B.append(float(rate["b"]))                     # load.py:48
        Ea.append(rate["Ea"] * ea_f / 1000)            # load.py:49
    A = np.array(A)
    B = np.array(B)
    Ea = np.array(Ea)

    nonzero = np.any(vf_full + vb_full != 0, axis=0)   # load.py:22-28
    zero_col = np.all(vf_full + vb_full == 0, axis=0)
    assert (not zero_col.any()) or np.all(zero_col[np.argmax(zero_col):]), \
        "Inert species must be the last elements in species_list"
    n_inert = int(zero_col.sum())
    vf = vf_full[:, nonzero]
    vb = vb_full[:, nonzero]
    vf_sum = vf.sum(axis=1)
    vb_sum = vb.sum(axis=1)

    # load.py:56-58 nondimensionalisation, evaluated in the reference's order
    A = (nd.t0 / nd.rho0) * (nd.T0 ** B) * ((nd.rho0 / nd.M0) ** vf_sum) * nd.M0 * A
    EaOverRu = Ea / (nd.e0 * nd.M0)
    third = third * (nd.rho0 / nd.M0)

    rp = {
        "species": species, "vf": vf, "vb": vb, "A": A, "B": B,
        "Ea/Ru": EaOverRu, "third_body_coeffs": third, "is_third_body": is_tb,
        "num_of_reactions": nrx, "num_of_species": ns,
        "num_of_inert_species": n_inert, "vsum": vb_sum - vf_sum,
    }

    # ---- get_cantera_coeffs, load.py:91-135 ----
    Tcr, lo, hi, M = [], [], [], []
    for sname in species:
        th = sp_by_name[sname]["thermo"]
        assert th["model"] == "NASA7"
        Tcr.append(th["temperature-ranges"][1])
        lo.append(th["data"][0])
        hi.append(th["data"][1])
        comp = sp_by_name[sname]["composition"]
        M.append(sum(ATOMIC_WEIGHT[e.upper()] * n for e, n in comp.items()))
    scale = np.array([[1, nd.T0, nd.T0 ** 2, nd.T0 ** 3, nd.T0 ** 4, 1 / nd.T0, 1]])
    lo = np.array(lo) * scale
    hi = np.array(hi) * scale
    species_M = np.array(M) / (1000 * nd.M0)
    Tcr = np.array(Tcr) / nd.T0
    th = {"species_M": species_M, "Tcr": Tcr}
    for tag, c in (("low", lo), ("high", hi)):
        cp = np.flip(c[:, 0:5], axis=1) / species_M[:, None]
        th["cp_" + tag] = cp
        th["dcp_" + tag] = cp[:, 0:-1] * np.array([[4, 3, 2, 1]])
        hpoly = np.flip(np.roll(c[:, 0:6], 1, axis=1), axis=1) * \
            np.array([[1 / 5, 1 / 4, 1 / 3, 1 / 2, 1, 1]])
        th["h_" + tag] = hpoly / species_M[:, None]
        th["h_chem_" + tag] = hpoly
        th["s_" + tag] = np.flip(np.concatenate([c[:, -1:], c[:, 1:5]], axis=1), axis=1) * \
            np.array([[1 / 4, 1 / 3, 1 / 2, 1, 1]])
        th["logcof_" + tag] = c[:, 0]
    th["raw_low"], th["raw_high"] = lo, hi
    th["molecular_weight"] = np.array(M)
    return rp, th, nd
