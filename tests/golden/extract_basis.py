"""Extract H/Li/C/N/O subsets (and Te/I of def2-SVP, with their ECP core
electron counts) of the basis-set-exchange JSON files (orbital bases, and
the def2-universal-jfit auxiliary basis of density-fitted J) that
qdk-chemistry ships (cpp/src/qdk/chemistry/algorithms/microsoft/scf/resources/
compressed/<name>.tar.gz, member basis/<name>.json) into the package's
qdk_chemistry_amd/data/basis/.

Only `electron_shells` (and `ecp_electrons`) of the kept elements are copied;
the values are data (exponents and contraction coefficients as strings,
untouched).  The ECP potentials themselves are not copied: ECP integrals are
not part of this package, only the reduced nuclear charges they imply.  Run once in a
container that holds the reference checkout; the output JSON is committed.
"""
import io
import json
import os
import sys
import tarfile

REF = "/root/reference/cpp/src/qdk/chemistry/algorithms/microsoft/scf/resources/compressed"
NAMES = ["cc-pvdz", "cc-pvtz", "def2-svp", "def2-tzvp", "6-31g_st_", "sto-3g", "def2-universal-jfit"]
ELEMENTS = ["1", "3", "6", "7", "8"]
EXTRA = {"def2-svp": ["52", "53"]}  # Te, I: the one-electron integral fixture's molecule


def main(out_dir):
    os.makedirs(out_dir, exist_ok=True)
    for name in NAMES:
        with tarfile.open(os.path.join(REF, name + ".tar.gz")) as tf:
            member = tf.getmember(f"basis/{name}.json")
            data = json.load(io.TextIOWrapper(tf.extractfile(member)))
        keep = {"name": data.get("name", name), "elements": {}}
        for z in ELEMENTS + EXTRA.get(name, []):
            if z in data["elements"]:
                el = data["elements"][z]
                keep["elements"][z] = {"electron_shells": el["electron_shells"]}
                if "ecp_electrons" in el:
                    keep["elements"][z]["ecp_electrons"] = el["ecp_electrons"]
        with open(os.path.join(out_dir, name + ".json"), "w") as f:
            json.dump(keep, f, indent=0)


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    default = os.path.join(here, "..", "..", "qdk-chemistry_amd", "qdk_chemistry_amd", "data", "basis")
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.normpath(default))
