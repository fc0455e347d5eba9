"""Writes the golden fixtures under tests/golden from PHIST's own test data.

Run in a container where the PHIST sources are mounted (default
/root/reference); the outputs are committed and the tests never read the
sources at run time.

Outputs
  a2_precalc_sprandn25.json  the A2_precalc_result golden vectors of
                             test/kernels/SparseMatTest_def.hpp:815-926 (real
                             case): y = A*x for A = Dsprandn25 and
                             x(i,j) = i + j*25 (0-based), nvec = 1 and 4,
                             checked there to sqrt(eps).
  matrices/*.mm              test matrices from test/matrices (data files of
                             the reference's tests, copied verbatim).
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
MATRICES = ["Dspeye25", "Dsprandn25", "Dsprandn388", "Dsprandn_nodiag25", "Dsprandsym163", "Dspshift25",
            "Dspzero25", "DsymmMat20", "DjadaTestMat512"]


def extract_precalc(text, nv):
    """the real-valued precalc_result block for _N_ == 25 && _NV_ == nv"""
    key = f"_N_ == 25 && _NV_ == {nv}"
    start = text.index(key)
    # the real case follows the '#else' of the IS_COMPLEX switch
    else_pos = text.index("#else", start)
    body_start = text.index("{", text.index("precalc_result[_N_*_NV_]", else_pos))
    body_end = text.index("}", body_start)
    nums = re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?", text[body_start:body_end])
    vals = [float(v) for v in nums]
    assert len(vals) == 25 * nv, (nv, len(vals))
    return vals


def main():
    src = os.path.join(REF, "test", "kernels", "SparseMatTest_def.hpp")
    text = open(src).read()
    out = {
        "source": "test/kernels/SparseMatTest_def.hpp:815-926 (A2_precalc_result, real case)",
        "matrix": "Dsprandn25",
        "input": "x(i,j) = i + j*25, i = 0..24",
        "alpha": 1.0,
        "beta": 0.0,
        "tolerance": "sqrt(eps) absolute (SparseMatTest_def.hpp:948)",
        "nvec1": extract_precalc(text, 1),
        "nvec4": extract_precalc(text, 4),
    }
    with open(os.path.join(HERE, "a2_precalc_sprandn25.json"), "w") as f:
        json.dump(out, f, indent=1)
    os.makedirs(os.path.join(HERE, "matrices"), exist_ok=True)
    for m in MATRICES:
        shutil.copyfile(os.path.join(REF, "test", "matrices", m + ".mm"), os.path.join(HERE, "matrices", m + ".mm"))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
