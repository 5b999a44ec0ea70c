"""Writes tests/golden/lambs_30k.npz from the reference's gravity accuracy
fixture (tests/gravity/lambs.00200_subsamp_30K + tests/gravity/direct.acc).

The fixture is data only: the 30 000 dark particles (big-endian float32 Tipsy
records, stored here as float32 arrays exactly as read) and the reference's
direct-sum accelerations (ASCII '%.14g', stored as float64).  Run from the repo
root in a container that has /root/reference; the GPU box only reads the .npz.
"""
import os
import sys

import numpy as np

REF = "/root/reference/tests/gravity"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from paratreet_amd.ics import read_tipsy, read_acc  # noqa: E402


def main():
    t = read_tipsy(os.path.join(REF, "lambs.00200_subsamp_30K"))
    acc = read_acc(os.path.join(REF, "direct.acc"))
    assert acc.shape == (len(t["mass"]), 3)
    np.savez_compressed(
        os.path.join(HERE, "lambs_30k.npz"),
        pos=t["pos"].astype(np.float32), vel=t["vel"].astype(np.float32),
        mass=t["mass"].astype(np.float32), soft=t["soft"].astype(np.float32),
        time=np.float64(t["time"]), direct_acc=acc)


if __name__ == "__main__":
    main()
