"""Writes tests/golden/superfast_g2_nodep.json: the SuperFast reaction system
(species, parameters, expression templates, 27 reactions and the system's own
6 tests / 49 assertions) from the reference's
prototypes/g2_depsink/g2_nodep.esm, as data for the mechanism-ingestion tests
(the GPU host has no copy of the reference).  Run from the repository root:

    python tests/golden/make_superfast_fixture.py [/root/reference]
"""
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "prototypes", "g2_depsink", "g2_nodep.esm")
doc = json.load(open(src))
rs = doc["reaction_systems"]["SuperFast"]
keep = {k: rs[k] for k in ("species", "parameters", "expression_templates", "reactions", "tests")}
out = {"esm": doc.get("esm"), "source": "prototypes/g2_depsink/g2_nodep.esm (reaction_systems.SuperFast)",
       "reaction_systems": {"SuperFast": keep}}
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "superfast_g2_nodep.json")
with open(dst, "w") as f:
    json.dump(out, f, indent=1, sort_keys=True)
print(dst, os.path.getsize(dst), "bytes")
