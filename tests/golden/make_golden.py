"""Writes the fixtures in tests/golden from the oracle (run from the repo root).

fbm_checksums.json: sha256 of the FP32 heights of the benchmark DEMs (side_seed).
"""
import hashlib
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
import pyoracle as O  # noqa: E402

if __name__ == "__main__":
    out = {}
    for n, seed in [(257, 0), (1025, 2), (4097, 1)]:
        out[f"{n}_{seed}"] = hashlib.sha256(O.fbm_heights(n, n, seed).tobytes()).hexdigest()
    (HERE / "fbm_checksums.json").write_text(json.dumps(out, indent=1) + "\n")
    print(out)
