"""Write tests/golden/family_hashes.json: SHA-256 of the initial conditions (conserved planes, binary64
little-endian bytes) that npde.families produces for a few seeded instances, so that a faster
construction of the same instances can be checked bit for bit against the one that wrote the
datasets and results in this repository.  Usage:  python tests/golden/make_family_hashes.py"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "neural-pde-surrogate_amd"))
from npde import families  # noqa: E402

CASES = [("riemann", i, n) for i in (0, 1, 2, 5, 17, 100003) for n in (64, 512)] + \
        [("blast", i, None) for i in range(12)] + [("blast", 3, 96)] + [("kh", i, 128) for i in (0, 1)]


def digest(U: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(U, dtype="<f8").tobytes()).hexdigest()


def main():
    out = []
    for fam, i, n in CASES:
        U = getattr(families, fam)(i, n)[0] if n is not None else getattr(families, fam)(i)[0]
        out.append({"family": fam, "index": i, "n": n, "shape": list(U.shape), "sha256": digest(U)})
    (Path(__file__).resolve().parent / "family_hashes.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(out)} instances")


if __name__ == "__main__":
    main()
