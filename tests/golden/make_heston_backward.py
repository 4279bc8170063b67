"""Extract the frozen Heston backward-pricing run the oracle's dense Heston generator is pinned on.

Source: test/fixtures/b155550/qtt_heston_backward.json of madhavmen/BondUpdateGalerkin (a frozen
run of BUG-Julia b155550; provenance in that directory's README.md).  The run's setup is stated in
test/QTTS/Tooling/test_nonlinear_step.jl:159-176: HestonParams(0.03, 0.0, 1.5, 0.09, 0.5, -0.7),
5 + 5 bits on x in [-3, 3] and v in [0, 1] with the end points on the grid, the payoff
0.02 softplus((e^x - 1) / 0.02) constant in v, ten steps of 0.02 at order 3.  Only data is copied:
the per-step bond dimensions and the 1024 final amplitudes (grid index ix + 32 iy).

Run from the repository root with the reference checkout at /root/reference:
    python tests/golden/make_heston_backward.py
"""
import json
import os

SRC = "/root/reference/test/fixtures/b155550/qtt_heston_backward.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "heston_backward.json")


def main():
    d = json.load(open(SRC))
    out = {
        "bits": d["N"], "h": d["h"], "steps": d["steps"], "order": d["order"],
        "params": {"r": 0.03, "q": 0.0, "kappa": 1.5, "theta": 0.09, "sigma": 0.5, "rho": -0.7},
        "x": [-3.0, 3.0], "v": [0.0, 1.0],
        "payoff": "0.02 softplus((exp(x) - 1) / 0.02), constant in v",
        "bond_dims": [r["bond_dims"] for r in d["trajectory"]],
        "final_state": {"re": d["final_state"]["re"], "im": d["final_state"]["im"]},
    }
    json.dump(out, open(OUT, "w"))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
