"""The reference's Zemax ray-trace comparisons (tests/test_zemax.py) for
HSC, DECam, CBP and ComCam, written once against a small tracer interface
so the same cases run on the oracle (CPU) and on the GPU library.

A tracer provides:
  rays(x, y, z, vx, vy, vz, t, wavelength)   one ray in the global frame
  from_stop(x, y, optic, stop, wavelength, dirCos)
  propagate(rays, t)
  trace_full(optic, rays) -> [(interface name, r[3], v[3])] in the global
                             frame after each Interface, in trace order.

Fixtures: tests/golden/zemax_traces.json and the optics YAMLs under
tests/golden/optics (written by tests/golden/make_fixtures.py).
"""
import json
import os

import numpy as np

import batoid_amd as bt

GOLD = os.path.join(os.path.dirname(__file__), "golden")
_TRACES = None


def traces():
    global _TRACES
    if _TRACES is None:
        with open(os.path.join(GOLD, "zemax_traces.json")) as f:
            _TRACES = json.load(f)
    return _TRACES


def optic(name):
    return bt.Optic.fromYaml(os.path.join(GOLD, "optics", name))


# tests/test_zemax.py:20-28: Zemax surfaces shared with HSC_old.yaml
HSC_SURFACES = [3, 6, 7, 8, 9, 11, 12, 13, 14, 16, 17, 18, 19, 20, 21, 24, 25, 28, 29, 31]
HSC_NAMES = ['PM', 'G1_entrance', 'G1_exit', 'G2_entrance', 'G2_exit',
             'ADC1_entrance', 'ADC1_exit', 'ADC2_entrance', 'ADC2_exit',
             'G3_entrance', 'G3_exit', 'G4_entrance', 'G4_exit',
             'G5_entrance', 'G5_exit', 'F_entrance', 'F_exit',
             'W_entrance', 'W_exit', 'D']


def hsc(tr):
    """tests/test_zemax.py:15-62: 1e-9 m and 1e-9 in direction cosines."""
    telescope = optic("HSC_old.yaml")
    for case in traces()["HSC"]:
        arr = np.array(case["rows"])
        a0 = arr[0]
        rays = tr.rays(a0[1] / 1000, a0[2] / 1000, 16.0, a0[4], a0[5], -a0[6], 0.0, 750e-9)
        out = {name: (r, v) for name, r, v in tr.trace_full(telescope, rays)}
        for name, isurf in zip(HSC_NAMES, HSC_SURFACES):
            r, v = out[name]
            isec = np.array([r[0], r[1], r[2] - 16.0])
            np.testing.assert_allclose(isec, arr[isurf - 1][1:4] / 1000, rtol=0, atol=1e-9,
                                       err_msg=f"{case['file']} {name}")
            np.testing.assert_allclose(v / np.linalg.norm(v), arr[isurf - 1][4:7], rtol=0,
                                       atol=1e-9, err_msg=f"{case['file']} {name}")


DECAM_SURFACES = [2, 3, 5, 7, 9, 10, 11, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 28,
                  29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40]


def decam(tr):
    """tests/test_zemax.py:465-512: rays from the primary mirror as stop,
    1e-9 m and 1e-9 in |direction cosines|."""
    telescope = optic("DECam.yaml")
    for case in traces()["DECam"]:
        arr = np.array(case["rows"])[:, 1:]
        dirCos = bt.utils.fieldToDirCos(np.deg2rad(case["Hx"] * 1.1),
                                        np.deg2rad(case["Hy"] * 1.1), projection='zemax')
        rays = tr.from_stop(case["Px"] * 2.005, case["Py"] * 2.005, telescope, telescope['PM'],
                            700e-9, dirCos)
        steps = tr.trace_full(telescope, rays)
        assert len(steps) == len(DECAM_SURFACES)
        for (name, r, v), iiz in zip(steps, DECAM_SURFACES):
            row = arr[iiz - 1]
            n = 1. / np.sqrt(np.sum(v**2))
            what = f"{case['file']} {name}"
            np.testing.assert_allclose(r[:2], row[:2] / 1e3, rtol=0, atol=1e-9, err_msg=what)
            np.testing.assert_allclose(abs(r[2]), abs(row[2] / 1e3), rtol=0, atol=1e-9,
                                       err_msg=what)
            np.testing.assert_allclose(np.abs(v * n), np.abs(row[3:6]), rtol=0, atol=1e-9,
                                       err_msg=what)


def _zemax_to_batoid(case):
    """x, y, z -> -x, y, -z; metres (tests/test_zemax.py:551-566)."""
    rows = np.array(case["rows"])
    x, y, z = -rows[:, 1] * 1e-3, rows[:, 2] * 1e-3, -rows[:, 3] * 1e-3
    xc, yc, zc = -rows[:, 4], rows[:, 5], -rows[:, 6]
    return -case["Hx"], case["Hy"], case["wavelength_um"] * 1e-6, x, y, z, xc, yc, zc


CBP_INDEX = {'Schmidt_plate_entrance': 1, 'Schmidt_plate_exit': 2, 'stop': 3,
             'turning_flat': 5, 'PM': 7, 'flat_hole': 9, 'L1_entrance': 11, 'L1_exit': 12,
             'L2_entrance': 13, 'L2_exit': 14, 'L3_entrance': 15, 'L3_exit': 16,
             'Detector': 17}


def cbp(tr):
    """tests/test_zemax.py:529-600: tilted turning flat and primary, two
    Sellmeier glasses; 1e-10 m and 1e-10 in direction."""
    cbp_optic = optic("CBP.yaml")
    for case in traces()["CBP"]:
        Hx, Hy, w, x, y, z, xc, yc, zc = _zemax_to_batoid(case)
        v = bt.utils.zemaxToDirCos(Hx * np.deg2rad(2.05), Hy * np.deg2rad(2.05))
        rays = tr.rays(x[0], y[0], z[0], *v, 0.0, w)
        tr.propagate(rays, -0.1)
        steps = tr.trace_full(cbp_optic, rays)
        assert [name for name, _, _ in steps] == list(CBP_INDEX)
        for name, r, bv in steps:
            if name == 'flat_hole':
                continue   # Zemax's tilt there differs (the reference skips it too)
            i = CBP_INDEX[name]
            what = f"{case['file']} {name}"
            np.testing.assert_allclose(r, [x[i], y[i], z[i]], rtol=0, atol=1e-10, err_msg=what)
            bv = bv / np.linalg.norm(bv)
            if bv[2] > 0:
                bv = -bv
            np.testing.assert_allclose(bv, [xc[i], yc[i], zc[i]], rtol=0, atol=1e-10,
                                       err_msg=what)


COMCAM_INDEX = {'M1': 5, 'M2': 6, 'M3': 8, 'L1_entrance': 10, 'L1_exit': 11,
                'L2_entrance': 12, 'L2_exit': 13, 'Filter_entrance': 14, 'Filter_exit': 15,
                'L3_entrance': 16, 'L3_exit': 17, 'Detector': 18}


def comcam(tr, bands="ugrizy"):
    """tests/test_zemax.py:603-667 (positions, 1e-9 m) for every band."""
    for band in bands:
        com = optic(f"ComCam_{band}.yaml")
        for case in traces()["ComCam"]:
            if case["band"] != band:
                continue
            Hx, Hy, w, x, y, z, xc, yc, zc = _zemax_to_batoid(case)
            d = np.array([x[1] - x[0], y[1] - y[0], z[1] - z[0]])
            v = d / np.linalg.norm(d)
            rays = tr.rays(x[0], y[0], z[0], *v, 0.0, w)
            steps = tr.trace_full(com, rays)
            assert [name for name, _, _ in steps] == list(COMCAM_INDEX)
            for name, r, _ in steps:
                i = COMCAM_INDEX[name] - 1
                np.testing.assert_allclose(r, [x[i], y[i], z[i]], rtol=0, atol=1e-9,
                                           err_msg=f"{case['file']} {name}")
