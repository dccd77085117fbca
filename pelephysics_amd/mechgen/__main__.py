"""Regenerate mechanism tables and HIP headers from the reference YAML files.

    python -m pelephysics_amd.mechgen [--ref /root/reference] [mech ...]

Writes ``pelephysics_amd/mechanisms/<mech>.json`` (normalised parameter
tables, read by the oracle and the tests) and
``pelephysics_amd/csrc/mech/<mech>.h`` (unrolled device code, small
mechanisms only) and
``pelephysics_amd/csrc/tab/<mech>.h`` (constant-memory tables).
"""
import argparse
import os

from .model import load_yaml, Mechanism
from .emit_hip import emit
from .emit_tables import emit as emit_tables

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
DEFAULT_MECHS = ["LiDryer", "drm19", "dodecane_lu", "heptane_lu_88sk", "CH4_lean_qss", "LuEthylene_qss",
                 "dodecane_lu_qss", "heptane_lu_qss"]
# mechanisms that also get the fully unrolled device code (mech/<m>.h); the
# larger ones are compiled from the constant-memory tables only
UNROLLED = ("LiDryer", "drm19")


# QSS mechanisms are read from the reduced mechanism the reference's QSS
# converter wrote (qssa.yaml: its species and reactions, qssa_species and
# forward_to_remove_idx), as the reference's generator reads it
# (.github/workflows/ci.yml:296-298)
def source(name: str) -> str:
    return "qssa.yaml" if name.endswith("_qss") else "mechanism.yaml"


def generate(mechs, ref="/root/reference"):
    out = []
    for name in mechs:
        src = os.path.join(ref, "Mechanisms", name, source(name))
        m = load_yaml(src, name)
        with open(os.path.join(PKG, "mechanisms", f"{name}.json"), "w") as f:
            f.write(m.to_json())
        if name in UNROLLED:
            with open(os.path.join(PKG, "csrc", "mech", f"{name}.h"), "w") as f:
                f.write(emit(m, f"Mechanisms/{name}/{source(name)}"))
        with open(os.path.join(PKG, "csrc", "tab", f"{name}.h"), "w") as f:
            f.write(emit_tables(m, f"Mechanisms/{name}/{source(name)}"))
        out.append(name)
    return out


def load(name: str) -> Mechanism:
    with open(os.path.join(PKG, "mechanisms", f"{name}.json")) as f:
        return Mechanism.from_json(f.read())


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("mechs", nargs="*", default=DEFAULT_MECHS)
    a = ap.parse_args()
    print(generate(a.mechs, a.ref))
