"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

  c1_pairs.txt          sorted (l, r) row pairs of workload c1 (2,600 x 2,300)
  c1_digest.json        count, sha256 and audit list of c1
  c2_digest.json        count and sha256 of the sorted (l, r) pairs of workload
                        c2 (100,000 x 100,000; about 5 minutes on 8 cores)
  c2_sample_pairs.txt   the pairs of c2 whose left row is a multiple of 97
  <fixture>.json        count, sha256, pairs and audit list of the CSV fixtures
                        (dblp_acm: the dblp-acm tables and rule shipped with the
                        reference's data/; dblp_google_self: DBLP-Google
                        tableA against itself with that set's rule, since its
                        tableB is not shipped)

Usage: python tests/golden/make_golden.py [c1] [c2] [c2_sample] [dblp_acm] [dblp_google_self]
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

C2_SAMPLE_STEP = 97


def synthetic(name):
    L, R, dup, cfg = datagen.make(name)
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    t = time.time()
    pairs, audit, cnt = H.run_oracle(plan, L, R, audit=(name == "c1"))
    dt = time.time() - t
    rec = {"workload": name, "count": int(len(pairs)), "sha256": H.digest_pairs(pairs),
           "counters": cnt, "oracle_seconds": round(dt, 1),
           "audit": None if audit is None else len(audit)}
    if name == "c1":
        np.savetxt(os.path.join(HERE, "c1_pairs.txt"), pairs, fmt="%d")
        rec["audit_entries"] = H.audit_to_json(audit)
    else:
        # the full audit of c2 costs as much as the blocking itself; the
        # sampled rows below carry their audit list
        rec["audit"] = 0
    with open(os.path.join(HERE, "%s_digest.json" % name), "w") as f:
        json.dump(rec, f, indent=1)
    print(name, {k: v for k, v in rec.items() if k != "audit_entries"})


def c2_sample():
    L, R, dup, cfg = datagen.make("c2")
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    pairs, audit, cnt = H.run_oracle(plan, L, R, l_step=C2_SAMPLE_STEP)
    np.savetxt(os.path.join(HERE, "c2_sample_pairs.txt"), pairs, fmt="%d",
               header="c2 pairs with left row %% %d == 0; audit entries: %d"
               % (C2_SAMPLE_STEP, len(audit)))
    print("c2_sample", len(pairs), "audit", len(audit))


def fixture(name):
    Lh, Rh, plan = H.load_fixture(name)
    L, R = H.host_columns(Lh), H.host_columns(Rh)
    t = time.time()
    pairs, audit, cnt = H.run_oracle(plan, L, R)
    rec = {"fixture": name, "files": H.FIXTURES[name], "rows": [Lh.n_rows, Rh.n_rows],
           "plan": {"n_rules": plan.n_rules, "pred_index": plan.pred_index,
                    "pred_type": plan.pred_type, "pred_threshold": plan.pred_threshold},
           "count": int(len(pairs)), "sha256": H.digest_pairs(pairs), "counters": cnt,
           "oracle_seconds": round(time.time() - t, 1),
           "pairs": pairs.tolist(), "audit_entries": H.audit_to_json(audit)}
    with open(os.path.join(HERE, "%s.json" % name), "w") as f:
        json.dump(rec, f, indent=0)
    print(name, rec["count"], "audit", len(audit), rec["counters"])


def main(which):
    for name in which:
        if name in ("c1", "c2"):
            synthetic(name)
        elif name == "c2_sample":
            c2_sample()
        else:
            fixture(name)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c1", "c2", "c2_sample", "dblp_acm", "dblp_google_self"])
