"""Generate the syngas gas-phase mechanism table used by the chemistry operator.

    python -m lbm_porouschem_amd.chem.make_mechanism [path/to/gri30.yaml]

Source data: GRI-Mech 3.0 (53 species, 325 reactions) in Cantera YAML form, as
shipped in the reference checkout at data/gri30.yaml.  The operator transports 12
gas species, so the detailed mechanism is reduced by species elimination: keep the
12 species of the H2/CO (syngas) sub-mechanism and every reaction whose reactants
and products all lie in that set.  That leaves 40 reactions (32 elementary, 6
three-body, 2 fall-off) - the complete H2/O2 + CO/HCO chemistry of GRI-Mech 3.0.
Third-body efficiencies of eliminated species are dropped (their concentration is
zero).  N2 takes part only as a collider.

Output (SI units, JSON) next to this file: syngas12_gri30.json
  species:   name, molar mass (kg/mol), element counts (C, H, O, N), NASA7 ranges
             and coefficients (low, high)
  reactions: reactant / product stoichiometry (species index, coefficient),
             kind (0 elementary, 1 three-body, 2 Lindemann fall-off, 3 Troe fall-off),
             A [m, mol, s], b, Ea/R [K], low-pressure rate for fall-off, Troe
             (a, T3, T1, T2; T2 = 0 means absent), efficiencies per kept species.
All reactions of the set are reversible; reverse rates come from equilibrium
constants of the NASA7 data at run time.
"""
import json
import os
import re
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/data/gri30.yaml"
OUT = os.path.join(HERE, "syngas12_gri30.json")

# transported gas species, in the order of the operator's state vector
KEEP = ["H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2", "CO", "CO2", "HCO", "N2"]
ATOMIC = {"C": 12.011e-3, "H": 1.008e-3, "O": 15.999e-3, "N": 14.007e-3, "Ar": 39.95e-3}
CAL = 4.184
R_GAS = 8.314462618


def _side(expr):
    out = []
    for term in expr.split(" + "):
        term = term.strip()
        if term in ("M", "(+M)"):
            continue
        m = re.match(r"^([0-9.]*)\s*(\S+)$", term)
        nu = float(m.group(1)) if m.group(1) else 1.0
        out.append((m.group(2), nu))
    return out


def parse_equation(eq):
    eq = eq.replace("(+M)", "").replace("+ M", "").strip()
    rev = "<=>" in eq
    lhs, rhs = re.split(r"\s*<=>\s*|\s*=>\s*", eq)
    merge = lambda side: {n: sum(v for m, v in side if m == n) for n, _ in side}
    return merge(_side(lhs)), merge(_side(rhs)), rev


def arrhenius(rc, order):
    # A in (cm^3/mol)^(order-1)/s -> (m^3/mol)^(order-1)/s; Ea cal/mol -> Ea/R in K
    return {"A": rc["A"] * (1e-6) ** (order - 1), "b": rc["b"], "Ea_R": rc["Ea"] * CAL / R_GAS}


def build(src):
    d = yaml.safe_load(open(src))
    assert d["units"] == {"length": "cm", "time": "s", "quantity": "mol", "activation-energy": "cal/mol"}
    spec = {s["name"]: s for s in d["species"]}
    species = []
    for name in KEEP:
        s = spec[name]
        th = s["thermo"]
        assert th["model"] == "NASA7"
        comp = s["composition"]
        species.append({
            "name": name,
            "M": sum(ATOMIC[e] * n for e, n in comp.items()),
            "elements": [comp.get(e, 0) for e in ("C", "H", "O", "N")],
            "T_ranges": th["temperature-ranges"],
            "nasa_low": th["data"][0],
            "nasa_high": th["data"][1],
        })
    idx = {n: i for i, n in enumerate(KEEP)}
    reactions = []
    for r in d["reactions"]:
        reac, prod, rev = parse_equation(r["equation"])
        if not (set(reac) | set(prod)) <= set(KEEP):
            continue
        assert rev, r["equation"]
        kind = {"elementary": 0, "three-body": 1, "falloff": 2}[r.get("type", "elementary")]
        order = sum(reac.values())
        rec = {
            "equation": r["equation"],
            "reactants": [[idx[n], v] for n, v in reac.items()],
            "products": [[idx[n], v] for n, v in prod.items()],
            "kind": kind,
            "efficiencies": [1.0] * len(KEEP),
        }
        if kind in (1, 2):
            for n, e in r.get("efficiencies", {}).items():
                if n in idx:
                    rec["efficiencies"][idx[n]] = e
        if kind == 0:
            rec.update(arrhenius(r["rate-constant"], order))
        elif kind == 1:
            rec.update(arrhenius(r["rate-constant"], order + 1))
        else:
            rec.update(arrhenius(r["high-P-rate-constant"], order))
            lo = arrhenius(r["low-P-rate-constant"], order + 1)
            rec["low"] = [lo["A"], lo["b"], lo["Ea_R"]]
            if "Troe" in r:
                t = r["Troe"]
                rec["kind"] = 3
                rec["troe"] = [t["A"], t["T3"], t["T1"], t.get("T2", 0.0)]
        reactions.append(rec)
    return {
        "source": "GRI-Mech 3.0 (Cantera YAML), reduced by species elimination to the H2/CO sub-mechanism",
        "species": species,
        "reactions": reactions,
    }


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    mech = build(src)
    with open(OUT, "w") as f:
        json.dump(mech, f, indent=1)
    print(f"{OUT}: {len(mech['species'])} species, {len(mech['reactions'])} reactions")


if __name__ == "__main__":
    main()
