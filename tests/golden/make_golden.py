"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

  c1_pairs.txt    sorted (l, r) row pairs of workload c1 (2,600 x 2,300)
  c2_digest.json  count and sha256 of the sorted (l, r) pairs of workload c2
                  (100,000 x 100,000; about 5 minutes on 8 cores)

Usage: python tests/golden/make_golden.py [c1] [c2]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import hyperblocker_amd as hb  # noqa: E402
from hyperblocker_amd import datagen  # noqa: E402
from tests import helpers as H  # noqa: E402


def main(which):
    for name in which:
        L, R, dup, cfg = datagen.make(name)
        plan = hb.HostPlan.from_yaml(cfg["rules"])
        t = time.time()
        pairs, audit, cnt = H.run_oracle(plan, L, R, audit=(name == "c1"))
        dt = time.time() - t
        if name == "c1":
            np.savetxt(os.path.join(HERE, "c1_pairs.txt"), pairs, fmt="%d")
        if audit is None:
            # audit of every rule evaluated alone, on the pairs the plan matches
            # or rejects: full audit of c2 costs as much as the blocking itself
            n_audit = len(H.run_oracle(plan, L, R)[1]) if name != "c2" else None
        else:
            n_audit = len(audit)
        rec = {"workload": name, "count": int(len(pairs)), "sha256": H.digest_pairs(pairs),
               "counters": cnt, "oracle_seconds": round(dt, 1), "audit": n_audit}
        with open(os.path.join(HERE, "%s_digest.json" % name), "w") as f:
            json.dump(rec, f, indent=1)
        print(name, rec)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c1", "c2"])
