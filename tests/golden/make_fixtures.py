"""Writes the committed fixtures of this directory.

* ``configs.json``: the seeded parameters of the benchmark workloads (pomerol_amd.workloads).
* ``reference_goldens.json``: golden values held by the reference's own tests, transcribed as
  data with their file:line (test/AndersonTest.cpp:40-64, GF2siteTest.cpp:40-51,
  GF4siteTest.cpp:36-46, Anderson2PGFTest.cpp:37-57, HamiltonianTest.cpp:70-75,
  3PSusc3siteTest.cpp:72-156).

Run: ``python tests/golden/make_fixtures.py`` (no reference code is imported or run).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pomerol_amd import workloads as W  # noqa: E402

I = 1j

GOLDENS = {
    "AndersonTest": {
        "source": "test/AndersonTest.cpp:32-64",
        "params": {"U": 3.7, "mu_over_U": 0.6, "h": 0.1, "V": 1.0, "epsilon": 2.3, "beta": 20.0},
        "tolerance": 1e-8,
        "G_up": [[-0.7545439, -0.14723373], [-0.59517353, -0.34478922], [-0.42646689, -0.4031622],
                 [-0.30758605, -0.39519013], [-0.23068661, -0.36811109], [-0.18013326, -0.33885065],
                 [-0.14541077, -0.31208195], [-0.12043127, -0.28860297], [-0.10171394, -0.26815136],
                 [-0.08721336, -0.25025992]],
        "G_down": [[0.49196891, -0.07241433], [0.44396903, -0.18681652], [0.37248532, -0.24764566],
                   [0.30425235, -0.26969548], [0.24921656, -0.27135953], [0.20705484, -0.26399883],
                   [0.1748209, -0.25319165], [0.14976473, -0.2414213], [0.12986709, -0.22973732],
                   [0.11373954, -0.21855949]],
    },
    "GF2siteTest": {
        "source": "test/GF2siteTest.cpp:35-51",
        "params": {"U": 1.0, "mu": 0.5, "beta": 10.0, "t": -1.0},
        "tolerance": 1e-8,
        "G_im": [-2.53021005e-01, -4.62090702e-01, -4.32482782e-01, -3.65598615e-01, -3.07785174e-01,
                 -2.62894141e-01, -2.28274316e-01, -2.01170772e-01, -1.79539602e-01, -1.61950993e-01],
    },
    "GF4siteTest": {
        "source": "test/GF4siteTest.cpp:32-59",
        "params": {"beta": 10.0},
        "tolerance": 1e-6,
        "G": [[0.00515461461, -0.191132319], [-0.0129218293, -0.35749415], [-0.0063208255, -0.364571553],
              [-0.00244599255, -0.326995909], [-0.000938220077, -0.285235829], [-0.000360621591, -0.248974505],
              [-0.000129046261, -0.219206946], [-3.20102701e-05, -0.194983212], [9.51503858e-06, -0.175149329],
              [2.68929175e-05, -0.158732731]],
    },
    "Anderson2PGFTest": {
        "source": "test/Anderson2PGFTest.cpp:37-57,96-152",
        "params": {"U": 0.5, "mu": 0.25, "levels": [1.02036910873357, -1.02036910873357],
                   "hoppings": [0.296439333614347, 0.296439333614347], "beta": 26.0,
                   "pole_resolution": 1e-5, "coefficient_tolerance": 1e-8},
        "tolerance": 1e-6,
        "chi_uuuu": [-2.342841271771e+01, 0.0, 6.932231165814e-03, 2.037522082872e-03, -2.150424835716e-03,
                     -4.384848776411e-03, -5.253420668000e-03, -5.370700986029e-03, -5.126175681822e-03,
                     -4.732777836189e-03],
    },
    "HamiltonianTest": {
        "source": "test/HamiltonianTest.cpp:40-42,70-75",
        "E0": -2.8860009,
        "tolerance": 1e-7,
    },
    # chi3(n1, n2) for n1, n2 in -1..1, row n1 + 1, column n2 + 1, as [re, im]
    "3PSusc3siteTest": {
        "source": "test/3PSusc3siteTest.cpp:30-158",
        "params": {"U": 4.0, "mu_over_U": 0.6, "t": 1.0, "beta": 5.0},
        "tolerance": 1e-10,
        "PP": {"indices": [["A", "up"], ["C", "up"], ["A", "dn"], ["C", "dn"]],
               "chi3": [[[0.056123393380680404, -0.027886788783494966], [0.03544036019820439, 0.0],
                         [0.008008339386258077, -0.0022610905291281]],
                        [[0.03544036019820439, 0.0], [0.056123393380680404, 0.027886788783494966],
                         [0.026205160724877802, 0.020851328654297216]],
                        [[0.008008339386258187, -0.002261090529127845],
                         [0.026205160724877885, 0.020851328654297483],
                         [0.013637220898198006, 0.012985397056265282]]]},
        "PH": {"indices": [["A", "up"], ["A", "up"], ["C", "dn"], ["C", "dn"]],
               "chi3": [[[-0.03377687240880357, 0.6799848693219751], [-0.01589165211289571, 0.0],
                         [-0.010498942766103334, -0.0019558762969712905]],
                        [[-0.01589165211289571, 0.0], [-0.03377687240880357, -0.6799848693219751],
                         [0.028277168735633868, 0.006544223759963227]],
                        [[-0.010498942766103213, -0.001955876296971393],
                         [0.02827716873563382, 0.006544223759962816],
                         [0.03857021454863498, -0.6687150248684948]]]},
        "xPH": {"indices": [["A", "up"], ["C", "up"], ["C", "dn"], ["A", "dn"]],
                "chi3": [[[0.27954539611247325, -0.16431057067114357], [0.08787949653955825, 0.0],
                          [0.025342850143558054, -0.003942925235138866]],
                         [[0.08787949653955825, 0.0], [0.27954539611247325, 0.16431057067114357],
                          [0.05067298632153447, 0.02536375725665837]],
                         [[0.02534285014355848, -0.00394292523513924],
                          [0.05067298632153436, 0.0253637572566579],
                          [0.11815163820082586, 0.06747758349542371]]]},
    },
}


def main():
    cfgs = {}
    for name, c in (("cfg1", W.cfg1()), ("cfg2", W.cfg2())):
        cfgs[name] = {"n_bath": c.n_bath, "U": c.U, "beta": c.beta, "mu": c.mu, "seed": c.seed,
                      "eps": list(map(float, c.eps)), "V": list(map(float, c.V))}
    _, beta3, tp = W.cfg3_expr()
    cfgs["cfg3"] = {"t": 1.0, "tp": tp, "U": 4.0, "mu": 2.0, "beta": beta3, "seed": W.SEED}
    batch = W.batch_configs(8)
    cfgs["batch_head"] = [{"n_bath": c.n_bath, "U": c.U, "beta": c.beta, "seed": c.seed,
                           "eps": list(map(float, c.eps)), "V": list(map(float, c.V))} for c in batch]
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump(cfgs, f, indent=1)
    with open(os.path.join(HERE, "reference_goldens.json"), "w") as f:
        json.dump(GOLDENS, f, indent=1)


if __name__ == "__main__":
    main()
