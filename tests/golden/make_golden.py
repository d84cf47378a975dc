"""Regenerate the golden fixtures in tests/golden/.

The reference ships no test vectors, so the fixtures are produced here by the
oracle (oracle/, a restatement of src/knn_brute.cu:801-885) after it has been
pinned by the checks in tests/test_oracle.py (independent numpy brute force,
hand-checked cases, planted mistakes).  Inputs are not stored: they are
regenerated from seeds and identified by SHA-256.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from matching_sift_amd import datagen  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    q, b = datagen.config0()
    sorted4 = oracle.host_brute(q, b, threads=8)
    np.save(os.path.join(OUT, "c0_sorted.npy"), sorted4)
    np.save(os.path.join(OUT, "c0_matches_r08.npy"), oracle.ratio(sorted4, 0.8))
    meta = {
        "config0": {"n_base": 10000, "n_query": 1000, "seed_base": datagen.SEED_BASE,
                    "seed_query": datagen.SEED_QUERY,
                    "sha256_base_f32": datagen.sha256(b), "sha256_query_f32": datagen.sha256(q)},
        "note": "sorted = host_brute float4 {d1,d2,idx1,idx2}; matches at ratio 0.8",
    }
    # LSH codes of 256 base rows under 16 integer hyperplanes (seeded)
    H = datagen.gauss_int(16 * 128, 32.0, 127.0, seed=99).reshape(16, 128)
    codes = oracle.lsh_codes(b[:256], H)
    np.save(os.path.join(OUT, "lsh_codes_256x16.npy"), codes)
    meta["lsh_codes"] = {"rows": 256, "nbits": 16, "hyperplane_seed": 99, "sd": 32.0, "lim": 127.0}
    with open(os.path.join(OUT, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote fixtures to", OUT)


if __name__ == "__main__":
    main()
