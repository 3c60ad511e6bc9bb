"""Write gemm_colmajor_known_answer.json: the fixed-input GEMM known answer of
the reference's own test (test/linearAlgebra/src/TestBlasLapackDoubleGemmHost.cpp:
36-112): A 10x5, B 5x3 column-major, C = A B, tolerance 1e-12.

Only the numbers are kept (inputs and expected output).  Run once with the
reference tree present; the JSON is committed.
"""
import json
import os
import re
import sys

SRC = "/root/reference/test/linearAlgebra/src/TestBlasLapackDoubleGemmHost.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gemm_colmajor_known_answer.json")


def grab(text, name):
    m = re.search(name + r"\s*=\s*\{([^}]*)\}", text)
    return [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]


def main():
    text = open(SRC).read()
    d = {"source": "test/linearAlgebra/src/TestBlasLapackDoubleGemmHost.cpp:36-112",
         "m": 10, "k": 5, "n": 3, "lda": 10, "ldb": 5, "ldc": 10, "tol": 1e-12,
         "A_colmajor": grab(text, "colMajA"), "B_colmajor": grab(text, "colMajB"),
         "C_colmajor": grab(text, "colMajCRef")}
    assert len(d["A_colmajor"]) == 50 and len(d["B_colmajor"]) == 15 and len(d["C_colmajor"]) == 30
    json.dump(d, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
