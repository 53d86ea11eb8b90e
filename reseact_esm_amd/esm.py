"""Mechanism ingestion from ESM `reaction_systems` (the JSON model format the
reference's models are written in; the SuperFast system inlined in
prototypes/g2_depsink/g2_nodep.esm:8-1860, referenced from
reseact.esm:12714-12717).

A reaction system declares species (some `constant`: fixed reservoirs such as
O2, CH4, H2O), parameters with defaults (temperature, pressure, photolysis
rates), expression templates, and reactions with substrates, products and a
rate expression tree.  It becomes a mass-action mechanism for the chemistry
operator:

    rate_r = k_r * prod_{substrates s} u_s^{n_s}
    du_i/dt = sum_r nu_{i,r} rate_r,   nu = product coef - substrate stoich,

where k_r is the rate expression evaluated with the parameters and the
constant species' values, and a constant substrate's value multiplies k_r
(constant species are not integrated).  There is no combinatoric 1/n_s!
factor: 2 HO2 -> ... proceeds at k HO2^2 and removes 2 HO2 per event.  The
system's own 49 assertions fix this: all hold to 1.5e-8 without the factor,
29 fail with it.

The result is the plain-array form of `rs_set_mechanism_stoich` (at most two
substrate slots per reaction, any number of products with real
coefficients) plus the species order, initial state and the system's own
tests (expected values with tolerances), which pin the ingestion.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np


# ---------------------------------------------------------------- expressions
_BINARY = {
    "+": lambda a: sum(a),
    "*": lambda a: math.prod(a),
    "-": lambda a: a[0] - sum(a[1:]) if len(a) > 1 else -a[0],
    "/": lambda a: a[0] / a[1],
    "^": lambda a: a[0] ** a[1],
}
_UNARY = {"exp": math.exp, "log": math.log, "log10": math.log10, "sqrt": math.sqrt,
          "abs": abs, "sin": math.sin, "cos": math.cos, "tanh": math.tanh}


def eval_expr(e, env: dict, templates: dict):
    """Evaluates an ESM expression tree: numbers, names (parameters), and
    {"op": ..., "args": [...]} nodes, including apply_expression_template
    (template params bound by "bindings")."""
    if isinstance(e, bool):
        raise ValueError("boolean in expression")
    if isinstance(e, (int, float)):
        return float(e)
    if isinstance(e, str):
        if e not in env:
            raise KeyError(f"unknown name {e!r} in rate expression")
        return float(env[e])
    if not isinstance(e, dict) or "op" not in e:
        raise ValueError(f"bad expression node {e!r}")
    op = e["op"]
    if op == "apply_expression_template":
        t = templates[e["name"]]
        local = dict(env)
        for k, v in e.get("bindings", {}).items():
            local[k] = eval_expr(v, env, templates)
        for k, v in zip(t.get("params", []), e.get("args", [])):
            local[k] = eval_expr(v, env, templates)
        return eval_expr(t["body"], local, templates)
    args = [eval_expr(a, env, templates) for a in e.get("args", [])]
    if op in _BINARY:
        return float(_BINARY[op](args))
    if op in _UNARY:
        if len(args) != 1:
            raise ValueError(f"{op} takes one argument")
        return float(_UNARY[op](args[0]))
    raise ValueError(f"unsupported op {op!r}")


# ---------------------------------------------------------------- the system
@dataclass
class EsmReaction:
    id: str
    substrates: list[tuple[str, float]]
    products: list[tuple[str, float]]
    rate: object


@dataclass
class EsmSystem:
    name: str
    species: dict             # name -> {"default": .., "constant": bool}
    parameters: dict          # name -> default value
    templates: dict
    reactions: list[EsmReaction]
    tests: list = field(default_factory=list)

    @property
    def active(self) -> list[str]:
        """Integrated species, in declaration order."""
        return [s for s, d in self.species.items() if not d.get("constant", False)]


def load_reaction_system(src, name: str | None = None) -> EsmSystem:
    """`src` is a path to an .esm file or its parsed JSON; `name` selects the
    reaction system (the only one if None)."""
    doc = src if isinstance(src, dict) else json.load(open(src))
    systems = doc["reaction_systems"]
    if name is None:
        if len(systems) != 1:
            raise ValueError(f"several reaction systems: {sorted(systems)}")
        name = next(iter(systems))
    rs = systems[name]
    species = {k: {"default": float(v.get("default", 0.0)), "constant": bool(v.get("constant", False))}
               for k, v in rs["species"].items()}
    params = {k: float(v.get("default", 0.0)) if isinstance(v, dict) else float(v)
              for k, v in rs.get("parameters", {}).items()}
    rx = []
    for r in rs["reactions"]:
        sub = [(s["species"], float(s.get("stoichiometry", 1))) for s in (r.get("substrates") or [])]
        pro = [(s["species"], float(s.get("stoichiometry", 1))) for s in (r.get("products") or [])]
        rx.append(EsmReaction(r.get("id", f"R{len(rx) + 1}"), sub, pro, r["rate"]))
    return EsmSystem(name, species, params, rs.get("expression_templates", {}), rx,
                     rs.get("tests", []))


@dataclass
class CompiledMechanism:
    """Plain arrays for rs_set_mechanism_stoich over the active species."""
    species: list[str]
    reactants: np.ndarray     # (nr, 2) int32, -1 = none (a species twice = second order)
    prod_off: np.ndarray      # (nr+1,) int32
    prod_species: np.ndarray  # int32
    prod_coef: np.ndarray     # float64
    k: np.ndarray             # (nr,) float64, constant substrates folded in
    photo: np.ndarray         # (nr,) int32 (0: the box model's photolysis is a parameter)
    ids: list[str]

    @property
    def ns(self) -> int:
        return len(self.species)

    @property
    def nr(self) -> int:
        return int(self.k.shape[0])

    def net(self) -> np.ndarray:
        """Net stoichiometry matrix (ns, nr)."""
        N = np.zeros((self.ns, self.nr))
        for r in range(self.nr):
            for a in self.reactants[r]:
                if a >= 0:
                    N[a, r] -= 1.0
            for q in range(self.prod_off[r], self.prod_off[r + 1]):
                N[self.prod_species[q], r] += self.prod_coef[q]
        return N

    def rhs(self, u: np.ndarray) -> np.ndarray:
        """du/dt of one cell (host evaluation, for checks)."""
        rate = self.k.copy()
        for r in range(self.nr):
            for a in self.reactants[r]:
                if a >= 0:
                    rate[r] *= u[a]
        return self.net() @ rate


def compile_system(sys: EsmSystem, parameter_overrides: dict | None = None,
                   species_values: dict | None = None,
                   sinks: dict | None = None) -> CompiledMechanism:
    """Evaluates every rate with the parameters (defaults, then overrides)
    and folds constant substrates (their default or `species_values`) into k.
    Substrates beyond two slots, or a stoichiometry that is not 1 or 2, are
    outside the mass-action kernels' reach and raise.

    `sinks` {species: k (1/s)} appends first-order losses D(X) = -k X as
    reactions X -> (nothing) named "sink_X": the form of the deposition sink
    the reference merges onto the chemistry by operator_compose
    (prototypes/g2_depsink/g2_dep.esm, D(X) ~ -(k_dry + k_wet) X; the sink
    model itself, superfast_deposition_sink.esm, is not in the reference
    tree, so which species take which k is the caller's)."""
    values = {s: d["default"] for s, d in sys.species.items()}
    if species_values:
        values.update(species_values)
    # rate expressions see the parameters and the constant species (e.g. H2O
    # in the water-dependent HO2 self-reaction); an active species inside a
    # rate would make it non-mass-action and is reported as an unknown name
    env = {s: values[s] for s, d in sys.species.items() if d["constant"]}
    env.update(sys.parameters)
    if parameter_overrides:
        env.update(parameter_overrides)
    active = sys.active
    index = {s: i for i, s in enumerate(active)}
    re_, off, psp, pco, kk, ids = [], [0], [], [], [], []
    for r in sys.reactions:
        k = eval_expr(r.rate, env, sys.templates)
        slots: list[int] = []
        for s, n in r.substrates:
            if n not in (1.0, 2.0):
                raise ValueError(f"{r.id}: substrate stoichiometry {n} not supported")
            if sys.species[s]["constant"]:
                k *= values[s] ** n
                continue
            slots += [index[s]] * int(n)
        if not slots or len(slots) > 2:
            raise ValueError(f"{r.id}: {len(slots)} active substrate slots (1 or 2 supported)")
        # substrates consumed, products formed; a catalyst appears on both sides
        for s, c in r.products:
            if sys.species[s]["constant"]:
                continue
            psp.append(index[s])
            pco.append(c)
        re_.append(slots + [-1] * (2 - len(slots)))
        off.append(len(psp))
        kk.append(k)
        ids.append(r.id)
    for sp, kd in (sinks or {}).items():
        if sp not in index:
            raise KeyError(f"sink on {sp!r}, which is not an active species")
        if not (kd >= 0.0 and math.isfinite(kd)):
            raise ValueError(f"sink rate of {sp} must be finite and >= 0")
        re_.append([index[sp], -1])
        off.append(len(psp))
        kk.append(float(kd))
        ids.append(f"sink_{sp}")
    return CompiledMechanism(active, np.array(re_, np.int32).reshape(-1, 2),
                             np.array(off, np.int32), np.array(psp, np.int32),
                             np.array(pco, np.float64), np.array(kk, np.float64),
                             np.zeros(len(kk), np.int32), ids)


def initial_state(sys: EsmSystem, overrides: dict | None = None) -> np.ndarray:
    vals = {s: d["default"] for s, d in sys.species.items()}
    if overrides:
        vals.update(overrides)
    return np.array([vals[s] for s in sys.active], np.float64)


@dataclass
class Assertion:
    test: str
    variable: str
    time: float
    expected: float
    rel: float
    abs: float


def assertions(sys: EsmSystem) -> list[Assertion]:
    """The system's own checks, with per-assertion tolerances falling back to
    the test's."""
    out = []
    for t in sys.tests:
        tt = t.get("tolerance", {})
        for a in t.get("assertions", []):
            at = a.get("tolerance", {})
            out.append(Assertion(t["id"], a["variable"], float(a["time"]), float(a["expected"]),
                                 float(at.get("rel", tt.get("rel", 0.0))),
                                 float(at.get("abs", tt.get("abs", 0.0)))))
    return out


def check(expected: float, got: float, rel: float, abs_: float) -> bool:
    """|got - expected| <= max(abs, rel |expected|)."""
    return abs(got - expected) <= max(abs_, rel * abs(expected))
