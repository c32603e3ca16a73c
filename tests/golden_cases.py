"""Loads the golden fixtures (tests/golden/*.npz, written by the compiled
reference) into a uniform description used by the oracle and GPU tests."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

INTER_1 = ["feq", "rho", "ux", "uy"]
INTER_2 = INTER_1 + ["ueqx", "ueqy", "frx", "fry", "fsx", "fsy"]


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*_s*.npz")))


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    nc = len(d["tau"])
    case = dict(name=name, mask=d["mask"].astype(np.int32), tau=list(d["tau"]),
                g_fluid=list(d["g_fluid"]), g_solid=list(d["g_solid"]), steps=int(d["steps"]),
                ncomp=nc, f0=[d["f0_%d" % c] for c in range(nc)],
                force=[d["force_%d" % c] if ("force_%d" % c) in d else None for c in range(nc)],
                out={})
    for c in range(nc):
        case["out"][("f", c)] = d["f_%d" % c]
        for nm in (INTER_2 if nc == 2 else INTER_1):
            case["out"][(nm, c)] = d["%s_%d" % (nm, c)]
    return case
