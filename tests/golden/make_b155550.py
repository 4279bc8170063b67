"""Extract the numbers the parity tests need from the reference's frozen fixtures.

Source: test/fixtures/b155550/xx_u1_bug_order{1,2}.json and xx_none_bug_order2.json of
madhavmen/BondUpdateGalerkin (frozen runs of BUG-Julia b155550: XX chain, Neel start, dt = 0.05;
U(1) at L = 10 for 20 steps, no symmetry at L = 8 for 10 steps; provenance in that directory's
README.md).  Only data is copied: per step the S^z profile, the error against free fermions, the
bond dimensions and the budget status (U(1)); per step the norm, bond dimensions and status, the
final 2^8 amplitudes and their error against the dense propagator (no symmetry).

Run from the repository root with the reference checkout at /root/reference:
    python tests/golden/make_b155550.py
"""
import json
import os

SRC = "/root/reference/test/fixtures/b155550"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "b155550_xx_u1.json")


def main():
    out = {}
    for name in ("xx_u1_bug_order1", "xx_u1_bug_order2"):
        d = json.load(open(os.path.join(SRC, name + ".json")))
        out[name] = {
            "L": d["L"], "dt": d["dt"], "start": d["start"], "order": d["order"],
            "steps": [{"sz": r["sz"], "err_exact": r["err_exact"], "bond_dims": r["bond_dims"],
                       "status": r["status"]} for r in d["trajectory"]],
        }
    d = json.load(open(os.path.join(SRC, "xx_none_bug_order2.json")))
    out["xx_none_bug_order2"] = {
        "L": d["L"], "dt": d["dt"], "start": d["start"], "order": d["order"],
        "err_dense": d["err_dense"],
        "final_state": {"re": d["final_state"]["re"], "im": d["final_state"]["im"]},
        "steps": [{"norm": r["norm"], "bond_dims": r["bond_dims"], "status": r["status"]}
                  for r in d["trajectory"]],
    }
    json.dump(out, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
