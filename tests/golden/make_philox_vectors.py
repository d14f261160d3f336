"""Regenerate philox_vectors.json from rocRAND's open-source Philox4x32-10.

Compiles philox_vectors.cpp (host code only) with hipcc against
/opt/rocm/include/rocrand/rocrand_philox4x32_10.h, runs it, and stores its output.
The first three entries are the Random123 known-answer inputs (all-zero, all-one,
pi digits); the rest are pseudo-random (counter, key) pairs.

Usage: python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "philox_vectors")
        subprocess.run(["hipcc", "-O1", "-o", exe, os.path.join(HERE, "philox_vectors.cpp")],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    vec = json.loads(out)
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump({"source": "rocrand_philox4x32_10.h ten_rounds (host)", "vectors": vec}, f,
                  separators=(",", ":"))
    print("wrote", len(vec), "vectors")


if __name__ == "__main__":
    main()
