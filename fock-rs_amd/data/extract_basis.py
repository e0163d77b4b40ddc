"""Extract the basis-set data this package ships from the reference's BSE JSON files.

The reference embeds Basis Set Exchange (BSE) JSON files
(basis_set/src/data/*.json, loaded by basis_set/src/basis_sets.rs:12-24).
This script keeps the fields the loader reads (basis_set/src/lib.rs:21-28:
function_type, angular_momentum, exponents, coefficients) for elements
H..Ne, with the exponent/coefficient strings untouched so that parsing them
gives the same doubles as the reference's `str::parse::<f64>`.

Run once from a checkout that has the reference mounted:
    python fock-rs_amd/data/extract_basis.py /root/reference
"""
import json
import os
import sys

SETS = ["STO-2G", "STO-3G", "STO-6G", "3-21G", "6-31G", "6-311G", "6-31++G", "6-31G_st"]
ELEMENTS = [str(z) for z in range(1, 11)]


def main(ref_root: str) -> None:
    here = os.path.dirname(os.path.abspath(__file__))
    out_dir = os.path.join(here, "basis")
    os.makedirs(out_dir, exist_ok=True)
    for name in SETS:
        with open(os.path.join(ref_root, "basis_set", "src", "data", name + ".json")) as f:
            data = json.load(f)
        elements = {}
        for z in ELEMENTS:
            if z not in data["elements"]:
                continue
            shells = []
            for sh in data["elements"][z]["electron_shells"]:
                shells.append({
                    "function_type": sh["function_type"],
                    "angular_momentum": sh["angular_momentum"],
                    "exponents": sh["exponents"],
                    "coefficients": sh["coefficients"],
                })
            elements[z] = {"electron_shells": shells}
        out = {"name": data["name"], "source": "Basis Set Exchange via fock-rs basis_set/src/data/" + name + ".json",
               "elements": elements}
        with open(os.path.join(out_dir, name + ".json"), "w") as f:
            json.dump(out, f, indent=None, separators=(",", ":"))
            f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
