"""Extract golden vectors from the reference's own test data into small
fixtures under tests/golden/.  Run in a container that has the reference
tree (read in place, never imported):

    python tests/golden/make_fixtures.py [/root/reference]

Writes:
  lsst_trace.json        Zemax ray traces through LSST (tests/testdata/
                         LSST_trace_{0,1,2}.txt, as read by
                         tests/test_zemax.py:412-462)
  lsst_wf.npz            Zemax wavefront maps (LSST_wf_*.txt,
                         tests/test_zemax.py:223-269)
  silicaDispersion.csv, kbk7.csv
                         refractive-index tables (tests/test_Medium.py)
  optics/*.yaml          optic descriptions used by those tests
"""
import json
import os
import shutil
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
TD = os.path.join(REF, "tests", "testdata")


def lsst_trace():
    out = []
    for fn in ["LSST_trace_0.txt", "LSST_trace_1.txt", "LSST_trace_2.txt"]:
        path = os.path.join(TD, fn)
        with open(path, encoding='utf-16-le') as f:
            Hx, Hy, Px, Py = np.genfromtxt(f, skip_header=13, max_rows=4, usecols=(6,))
        with open(path, encoding='utf-16-le') as f:
            arr = np.genfromtxt(f, skip_header=22, max_rows=18, usecols=list(range(1, 12)))
        out.append({"file": fn, "Hx": Hx, "Hy": Hy, "Px": Px, "Py": Py,
                    "rows": arr.tolist()})
    with open(os.path.join(OUT, "lsst_trace.json"), "w") as f:
        json.dump({"zemax_surfaces": [4, 5, 7, 9, 10, 11, 12, 13, 14, 15, 16, 17],
                   "traces": out}, f, indent=1)


def lsst_wf():
    thxs = [0.0, 0.0, 0.0, 1.176]
    thys = [0.0, 1.225, 1.75, 1.176]
    fns = ["LSST_wf_0.0_0.0.txt", "LSST_wf_0.0_1.225.txt",
           "LSST_wf_0.0_1.75.txt", "LSST_wf_1.176_1.176.txt"]
    maps = []
    for fn in fns:
        with open(os.path.join(TD, fn), encoding='utf-16-le') as f:
            maps.append(np.loadtxt(f, skiprows=16))
    np.savez_compressed(os.path.join(OUT, "lsst_wf.npz"), thx_deg=np.array(thxs),
                        thy_deg=np.array(thys), maps=np.array(maps))


def copies():
    for fn in ["silicaDispersion.csv", "kbk7.csv"]:
        shutil.copy(os.path.join(TD, fn), os.path.join(OUT, fn))
    os.makedirs(os.path.join(OUT, "optics"), exist_ok=True)
    for fn in ["LSST_g_500.yaml", "LSST_r.yaml"]:
        shutil.copy(os.path.join(REF, "batoid", "data", "LSST", fn),
                    os.path.join(OUT, "optics", fn))


if __name__ == "__main__":
    lsst_trace()
    lsst_wf()
    copies()
    print("fixtures written to", OUT)
