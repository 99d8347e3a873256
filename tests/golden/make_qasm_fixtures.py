"""Copies the reference's circuit instance files (circuits/qasm/*.qasm: OpenQASM 2.0 data written by
circuits/generators.py:100-102 through Qiskit's exporter) into tests/golden/qasm/, so the front-end
test runs where the reference checkout is absent.  Run in the container that holds the reference:

    python tests/golden/make_qasm_fixtures.py /root/reference/circuits/qasm
"""
import glob
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(src):
    dst = os.path.join(HERE, "qasm")
    os.makedirs(dst, exist_ok=True)
    files = sorted(glob.glob(os.path.join(src, "*.qasm")))
    for f in files:
        shutil.copyfile(f, os.path.join(dst, os.path.basename(f)))
    print(f"{len(files)} instance files -> {dst}")


if __name__ == "__main__":
    main(sys.argv[1])
