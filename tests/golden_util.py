"""Loads the golden vectors written by tests/golden/make_golden.py."""
import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    d["meta"] = json.loads(str(d["meta"]))
    return d


def oracle_params(d, orc):
    a = d["meta"]["args"]
    flux = {"roe": orc.ROE, "hllc": orc.HLLC}[a["flux"]]
    lim = {"none": orc.LIM_NONE, "venkat": orc.LIM_VENKAT, "bj": orc.LIM_BJ}[a["lim"]]
    return orc.params(d["inf"], flux=flux, order=a["order"], limiter=lim, limiter_order=a["lorder"],
                      entropy_fix=a["efix"], venkat_k=a["k"], cfl_min=1.0, cfl_max=1000.0, cfl_ramp=50)


def oracle_mesh(d, orc):
    return orc.Mesh(d["xyz"], d["edge"], d["edge_area"], d["bedge"], d["bedge_area"], d["bedge_type"], d["vol"],
                    d["n2n_ia"], d["n2n_ja"], d["bedge_tag"])
