"""Pin the oracle against the reference's Zemax ray traces through HSC,
DECam, CBP and ComCam (tests/test_zemax.py:15-667): Baffles, tilted and
shifted frames, Table / Sumita / Sellmeier glasses, polygon, rectangle,
ray and union obscurations, a mirror as the stop surface.  CPU only."""
import ctypes

import numpy as np
import pytest

import batoid_amd as bt
import oracle
from oracle import raygen as G
from oracle.host import HostRays, HostPlan, host_records
from batoid_amd.plan import PlanSpec, collect_steps
import zemax_cases as Z


def _dv(a):
    a = np.ascontiguousarray(a, dtype=float).ravel()
    return (ctypes.c_double * len(a))(*a)


def _get_n(medium, w):
    tab = host_records(medium)
    w = np.array([w], dtype=float)
    out = np.empty(1)
    oracle.load().or_get_n_n(tab.ptr(), w.ctypes.data, out.ctypes.data, 1, 1)
    return out[0]


class OracleTracer:
    """The oracle, one Interface at a time (optic.py:1045-1049)."""

    def rays(self, x, y, z, vx, vy, vz, t, w):
        h = HostRays(x=[x], y=[y], z=[z], vx=[vx], vy=[vy], vz=[vz], t=[t], wavelength=[w])
        h.coordSys = bt.globalCoordSys
        return h

    def from_stop(self, x, y, optic, stop, w, dirCos):
        ct = bt.CoordTransform(stop.coordSys, optic.coordSys)
        h = G.finish([x], [y], stop.surface, (ct.dr, ct.drot), dirCos, _get_n(optic.inMedium, w),
                     optic.backDist, w)
        h.coordSys = optic.coordSys
        return h

    def propagate(self, h, t):
        oracle.load().or_propagate_in_place(h.struct, t, 1)

    def trace_full(self, optic, h):
        O = oracle.load()
        out = []
        cs = h.coordSys
        for step in collect_steps(optic):
            HostPlan(PlanSpec([step], cs)).trace(h, niter=5, threads=1)
            cs = step[0].coordSys
            g = h.copy()
            ct = bt.CoordTransform(cs, bt.globalCoordSys)
            O.or_apply_forward_transform(_dv(ct.dr), _dv(ct.drot), g.struct, 1)
            out.append((step[0].name, np.array([g["x"][0], g["y"][0], g["z"][0]]),
                        np.array([g["vx"][0], g["vy"][0], g["vz"][0]])))
        return out


@pytest.mark.parametrize("case", ["hsc", "decam", "cbp", "comcam"])
def test_oracle_vs_zemax(case):
    getattr(Z, case)(OracleTracer())
