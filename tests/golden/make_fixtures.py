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
  zemax_traces.json      Zemax ray traces through HSC, DECam, CBP and
                         ComCam (tests/test_zemax.py:15-62, 465-667)
  hsc_zemax.npz          HSC wavefront map, Zernike coefficients and
                         Huygens PSF (tests/test_zemax.py:66-220)
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


def _zemax_utf16(path, nrows):
    """Field/pupil coordinates, wavelength and the per-surface rows of a
    UTF-16 Zemax 'Ray Trace Data' export (as tests/test_zemax.py reads
    them)."""
    with open(path, encoding='utf-16-le') as f:
        wl = np.genfromtxt(f, skip_header=8, max_rows=1, usecols=(2,))
    with open(path, encoding='utf-16-le') as f:
        Hx, Hy, Px, Py = np.genfromtxt(f, skip_header=13, max_rows=4, usecols=(6,))
    with open(path, encoding='utf-16-le') as f:
        rows = np.genfromtxt(f, skip_header=22, max_rows=nrows, usecols=list(range(0, 12)))
    return {"wavelength_um": float(wl), "Hx": float(Hx), "Hy": float(Hy),
            "Px": float(Px), "Py": float(Py), "rows": rows.tolist()}


def other_traces():
    """HSC (tests/test_zemax.py:15-62), DECam (465-512), CBP (529-600) and
    ComCam (603-667) Zemax traces."""
    out = {}
    hsc = []
    for fn in ["HSC_raytrace_1.txt", "HSC_raytrace_2.txt", "HSC_raytrace_3.txt"]:
        with open(os.path.join(TD, fn)) as f:
            arr = np.loadtxt(f, skiprows=22, usecols=list(range(0, 12)))
        hsc.append({"file": fn, "rows": arr.tolist()})
    out["HSC"] = hsc
    out["DECam"] = [dict(file=fn, **_zemax_utf16(os.path.join(TD, fn), 40))
                    for fn in ["DECam_raytrace_0.txt", "DECam_raytrace_1.txt",
                               "DECam_raytrace_2.txt"]]
    out["CBP"] = [dict(file=f"CBP_SRT{i}.txt",
                       **_zemax_utf16(os.path.join(TD, f"CBP_SRT{i}.txt"), 18))
                  for i in range(1, 5)]
    out["ComCam"] = [dict(file=f"ComCam_SRT{i}_{b.upper()}.txt", band=b,
                          **_zemax_utf16(os.path.join(TD, f"ComCam_SRT{i}_{b.upper()}.txt"), 18))
                     for b in "ugrizy" for i in (1, 2)]
    with open(os.path.join(OUT, "zemax_traces.json"), "w") as f:
        json.dump(out, f, indent=0)


def hsc_maps():
    """HSC wavefront map (tests/test_zemax.py:154-185), Zernike coefficients
    (188-220) and Huygens PSF (66-151)."""
    with open(os.path.join(TD, "HSC_wavefront.txt")) as f:
        wf = np.loadtxt(f, skiprows=17)
    with open(os.path.join(TD, "HSC_huygensPSF.txt")) as f:
        psf = np.loadtxt(f, skiprows=21)
    zern = [0.0]
    with open(os.path.join(TD, "HSC_Zernike.txt")) as f:
        for i, line in enumerate(f):
            if i > 38:
                zern.append(float(line[9:20]))
    np.savez_compressed(os.path.join(OUT, "hsc_zemax.npz"), wavefront=wf, huygens_psf=psf,
                        zernike=np.array(zern))


def copies():
    for fn in ["silicaDispersion.csv", "kbk7.csv"]:
        shutil.copy(os.path.join(TD, fn), os.path.join(OUT, fn))
    os.makedirs(os.path.join(OUT, "optics"), exist_ok=True)
    for sub, fn in [("LSST", "LSST_g_500.yaml"), ("LSST", "LSST_r.yaml"),
                    ("HSC", "HSC.yaml"), ("HSC", "HSC_old.yaml"), ("HSC", "HSC_no_obsc.yaml"),
                    ("DECam", "DECam.yaml"), ("LSST", "CBP.yaml")] + [
                        ("LSST", f"ComCam_{b}.yaml") for b in "ugrizy"]:
        shutil.copy(os.path.join(REF, "batoid", "data", sub, fn),
                    os.path.join(OUT, "optics", fn))


if __name__ == "__main__":
    lsst_trace()
    lsst_wf()
    other_traces()
    hsc_maps()
    copies()
    print("fixtures written to", OUT)
