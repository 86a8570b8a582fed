"""Writes the committed fixtures of tests/golden from the FP64 oracle.

loop_hashes.json: SHA-256 of the FP32 SoA loop buffers (x then y arrays,
loop-major) for a few (config_id, Nl, N); pins the generator bit for bit.
Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import numpy as np  # noqa: E402
import oracle as orc  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    out = []
    for cid, Nl, N in [(1, 16, 64), (2, 8, 4096), (3, 96, 256), (1, 1024, 1024)]:
        x, y = orc.gen_loops(cid, Nl, N)
        out.append({"config_id": cid, "Nl": Nl, "N": N, "sha256_x": sha(x), "sha256_y": sha(y)})
    with open(os.path.join(HERE, "loop_hashes.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
