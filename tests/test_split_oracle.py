"""The oracle's traceSplit work queue (oracle/split.py, following
batoid/optic.py:1184-1282) against an independent enumeration of the ghost
tree.

On the reference's simple telescope (tests/test_rsplit.py:38-80: a plane
lens, a plane mirror, a plane detector, every refractive surface and the
detector 2 % reflective, 98 % transmissive) no ray is ever vignetted and
every ray of a batch takes the same branches, so the set of paths that
traceSplit must return, and each path's flux, follow from the optic's item
list alone: a branch of flux f splits into f*T (same direction through the
item list) and f*R (reversed) at a refractive interface or the detector, a
mirror passes it on in the same list direction with its travel along z
reversed, and a branch below minFlux is dropped.  A path dropped or added by the work
queue, a wrong flux factor or a wrong end of the optic fails here; so does a
flux budget (kept + dropped = incoming, R + T = 1) that does not close."""
import numpy as np
import pytest

import batoid_amd as bt
from oracle import split as OS
from oracle.host import HostRays

R, T = 0.02, 0.98


def _simple_telescope():
    telescope = bt.CompoundOptic(
        name="simple", stopSurface=bt.Interface(bt.Plane()),
        items=[
            bt.Lens(name="L1", items=[
                bt.RefractiveInterface(bt.Plane(), name="L1_entrance",
                                       coordSys=bt.CoordSys(origin=[0, 0, 0.3]),
                                       inMedium=bt.ConstMedium(1.0),
                                       outMedium=bt.ConstMedium(1.1)),
                bt.RefractiveInterface(bt.Plane(), name="L1_exit",
                                       coordSys=bt.CoordSys(origin=[0, 0, 0.2]),
                                       inMedium=bt.ConstMedium(1.1),
                                       outMedium=bt.ConstMedium(1.0))]),
            bt.Mirror(bt.Plane(), name="Mirror"),
            bt.Detector(bt.Plane(), name="detector",
                        coordSys=bt.CoordSys(origin=[0, 0, 0.1]))],
        pupilSize=1.0, backDist=1.0, inMedium=bt.ConstMedium(1.0))
    for name in ("L1_entrance", "L1_exit"):
        telescope[name].forwardCoating = bt.SimpleCoating(R, T)
        telescope[name].reverseCoating = bt.SimpleCoating(R, T)
    telescope['detector'].forwardCoating = bt.SimpleCoating(R, T)
    return telescope


def _leaves(optic):
    """(name, kind) of the Interfaces in trace order, CompoundOptics flattened."""
    out = []
    for item in optic.items:
        if hasattr(item, 'items'):
            out.extend(_leaves(item))
        else:
            out.append((item.name, type(item).__name__))
    return out


def enumerate_ghosts(leaves, minFlux):
    """Depth-first walk of one ray's branch tree over the flattened item
    list: {path: (flux, sign of vz)} leaving forward (past the last item)
    and in reverse (before the first), and the flux dropped below minFlux.
    Rays enter travelling -z; every reflection flips the sign of vz."""
    out_f, out_r, dropped = {}, {}, 0.0
    stack = [(0, +1, 1.0, -1, ())]
    while stack:
        i, d, f, s, path = stack.pop()
        name, kind = leaves[i]
        path = path + (name,)
        if kind == "Mirror":
            branches = [(d, f, -s)]
        else:
            branches = [(d, f * T, s), (-d, f * R, -s)]
        for nd, nf, ns in branches:
            if nf < minFlux:
                dropped += nf
                continue
            j = i + nd
            if j == len(leaves):
                assert path not in out_f
                out_f[path] = (nf, ns)
            elif j < 0:
                assert path not in out_r
                out_r[path] = (nf, ns)
            else:
                stack.append((j, nd, nf, ns, path))
    return out_f, out_r, dropped


def _rays(n, seed=11):
    rng = np.random.default_rng(seed)
    r = 0.5 * np.sqrt(rng.uniform(size=n))
    th = rng.uniform(0, 2 * np.pi, n)
    vx, vy = 0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n)
    vz = -np.sqrt(1 - vx * vx - vy * vy)
    h = HostRays(x=r * np.cos(th), y=r * np.sin(th), z=np.full(n, 1.0),
                 vx=vx, vy=vy, vz=vz, t=0.0, wavelength=500e-9, flux=1.0)
    return OS.Rays(h, bt.CoordSys())


def test_enumeration_closes_the_flux_budget():
    leaves = _leaves(_simple_telescope())
    assert [k for _, k in leaves] == ["RefractiveInterface", "RefractiveInterface",
                                      "Mirror", "Detector"]
    for minFlux in (1e-2, 1e-4, 1e-6):
        f, r, dropped = enumerate_ghosts(leaves, minFlux)
        total = sum(v[0] for v in f.values()) + sum(v[0] for v in r.values()) + dropped
        assert abs(total - 1.0) < 1e-13
    # minFlux above R: only the straight-through path survives
    f, r, _ = enumerate_ghosts(leaves, 0.05)
    assert list(f) == [("L1_entrance", "L1_exit", "Mirror", "detector")] and not r
    assert f[("L1_entrance", "L1_exit", "Mirror", "detector")] == (pytest.approx(T ** 3), 1)


@pytest.mark.parametrize("minFlux", [5e-2, 1e-3, 1e-4, 1e-6])
def test_oracle_trace_split_matches_enumeration(minFlux):
    optic = _simple_telescope()
    n = 64
    out_f, out_r = OS.trace_split(optic, _rays(n), minFlux)
    want_f, want_r, dropped = enumerate_ghosts(_leaves(optic), minFlux)
    for got, want in ((out_f, want_f), (out_r, want_r)):
        paths = [tuple(g.path) for g in got]
        assert len(paths) == len(set(paths)), "a path returned twice"
        assert sorted(paths) == sorted(want), "paths dropped or added"
        for g in got:
            assert len(g) == n
            assert not g.h["vignetted"].any() and not g.h["failed"].any()
            flux, sign = want[tuple(g.path)]
            np.testing.assert_allclose(g.h["flux"], flux, rtol=1e-14, atol=0)
            assert (np.sign(g.h["vz"]) == sign).all()
    kept = sum(float(np.sum(g.h["flux"])) for g in list(out_f) + list(out_r))
    assert kept + n * dropped == pytest.approx(n, rel=1e-13)
