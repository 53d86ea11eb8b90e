"""Writes jit_sources.inc: the texts compiled at run time by hipRTC (the
chemistry prelude, the shared launch structs and the kernel template) as C++
raw-string constants for reseact_hip.cpp."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PARTS = [("kJitPrelude", "chem_prelude.h"), ("kJitLaunch", "rs_launch.h"),
         ("kJitChemKernels", "chem_kernels.hip")]
DELIM = "RSJIT"

out = ["// generated by embed_jit.py -- do not edit"]
for name, fn in PARTS:
    text = open(os.path.join(HERE, fn)).read()
    if ")" + DELIM + '"' in text:
        sys.exit(f"{fn} contains the raw-string delimiter")
    out.append(f'static const char {name}[] = R"{DELIM}(\n{text}\n){DELIM}";')
sys.stdout.write("\n".join(out) + "\n")
