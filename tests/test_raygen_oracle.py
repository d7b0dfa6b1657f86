"""Pin the ray-generation oracle (oracle/raygen.py) and the host bookkeeping
of GPU draws from a numpy PCG64 stream.  CPU only.

What fixes each piece:
* PCG64 jump-ahead: numpy's own Generator, draw for draw;
* the device's closed forms for the lattice and the hexapolar azimuths
  (fftfreq as (k - n//2) / (n d); linspace(0, 2 pi, m, endpoint=False) as
  j * (2 pi / m)): numpy itself, over every size that matters;
* hexapolar ring counts: the reference's test_asPolar properties
  (tests/test_RayVector.py:530-605) and the C2 pupil's closed-form count;
* finishParallel + frame change: the reference's own test_asGrid checks
  (tests/test_RayVector.py:368-527), i.e. the chief ray at backDist and
  the lattice spacing recovered on the pupil plane.
"""
import numpy as np
import pytest

import batoid_amd as bt
from batoid_amd import _pcg
from oracle import raygen as G


def _numpy_draws(seed, k):
    return np.random.default_rng(seed).random(size=k)


@pytest.mark.parametrize("seed", [57721, 2, 5772156649])
def test_pcg64_jump_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    (sh, sl), (ih, il) = _pcg.stream_words(rng)
    state, inc = (sh << 64) | sl, (ih << 64) | il
    ks = [0, 1, 2, 3, 63, 64, 1000, 4097]
    ref = _numpy_draws(seed, max(ks) + 1)
    for k in ks:
        assert G.pcg64_uniform(state, inc, k) == ref[k], k
    # a far draw: numpy's own advance() then one draw
    far = 10_000_019
    rng2 = np.random.default_rng(seed)
    rng2.bit_generator.advance(far)
    assert G.pcg64_uniform(state, inc, far) == rng2.random()
    # the product's host-side jump agrees with the oracle's table composition
    assert _pcg.jump(state, inc, far) == G.pcg64_state_at(state, inc, far)


def test_pcg64_planted_mistake_is_caught():
    """A rotation taken from the wrong bits (a plausible slip) fails."""
    rng = np.random.default_rng(57721)
    (sh, sl), (ih, il) = _pcg.stream_words(rng)
    state, inc = (sh << 64) | sl, (ih << 64) | il
    s1 = G.pcg64_state_at(state, inc, 1)
    x = ((s1 >> 64) ^ s1) & ((1 << 64) - 1)
    rot = (s1 >> 64) & 63                       # wrong: low bits of the high word
    bad = ((x >> rot) | (x << ((64 - rot) & 63))) & ((1 << 64) - 1)
    good = G.pcg64_output(s1)
    ref = _numpy_draws(57721, 1)[0]
    assert (good >> 11) * 2.0 ** -53 == ref
    assert (bad >> 11) * 2.0 ** -53 != ref


def test_uniform_is_lo_plus_span_times_draw():
    """Generator.uniform(lo, hi) = lo + (hi - lo) * random(): the form the
    GPU evaluates with lo and span handed over by the host."""
    for lo, hi in [(-0.5, 0.5), (2.558**2, 4.18**2), (0.0, 2 * np.pi), (0.1**2, 0.2**2)]:
        a = np.random.default_rng(7).uniform(lo, hi, size=5000)
        u = np.random.default_rng(7).random(size=5000)
        np.testing.assert_array_equal(a, lo + (hi - lo) * u)


def test_consume_leaves_numpy_state():
    """After the GPU took 2n draws the host generator is where numpy's would
    be after drawing them, buffered 32-bit half-word included."""
    for n in (1, 1000, 123457):
        a = np.random.default_rng(99)
        b = np.random.default_rng(99)
        a.integers(0, 2**31, dtype=np.uint32)     # leaves a buffered half-word
        b.integers(0, 2**31, dtype=np.uint32)
        a.uniform(size=2 * n)
        _pcg.consume(b, 2 * n)
        assert a.bit_generator.state == b.bit_generator.state
        np.testing.assert_array_equal(a.integers(0, 2**31, size=8, dtype=np.uint32),
                                      b.integers(0, 2**31, size=8, dtype=np.uint32))
    assert _pcg.stream_words(np.random.Generator(np.random.MT19937(1))) is None


def test_linspace_azimuths_are_index_times_step():
    for m in list(range(1, 4000)) + [11076, 11080, 11082, 6 * 1846]:
        a = np.linspace(0, 2 * np.pi, m, endpoint=False)
        np.testing.assert_array_equal(a, np.arange(m) * (2 * np.pi / m))


def test_fftfreq_lattice_closed_form():
    for nx in list(range(1, 70)) + [510, 511, 512, 640, 1023]:
        ny = nx + 3
        uu, vv = G.grid_uv(nx, ny)
        x_d = 1. if nx <= 2 else (nx - (2 if nx % 2 == 0 else 1)) / nx
        y_d = 1. if ny <= 2 else (ny - (2 if ny % 2 == 0 else 1)) / ny
        i = np.arange(nx * ny)
        iy, ix = i // nx, i % nx
        np.testing.assert_array_equal(uu, (ix - nx // 2) * (1.0 / (nx * x_d)))
        np.testing.assert_array_equal(vv, (iy - ny // 2) * (1.0 / (ny * y_d)))


def test_axis_aligned_lattice_dot_is_one_product():
    """np.dot((L, 0), (u, v)) is exactly L*u: the GPU's lx[0]*u + lx[1]*v."""
    rng = np.random.default_rng(3)
    uu, vv = G.grid_uv(96, 97)
    for L in rng.uniform(0.5, 10.0, size=5):
        x, y = G.lattice_xy((L, 0.0), (0.0, L), uu, vv)
        np.testing.assert_array_equal(x, L * uu + 0.0 * vv)
        np.testing.assert_array_equal(y, 0.0 * uu + L * vv)


def test_hexapolar_counts():
    """tests/test_RayVector.py:546-605: multiples of kfold, one centre ray
    when inner == 0, and the C2 pupil's 9,998,544 rays."""
    rng = np.random.default_rng(5772156)
    for _ in range(20):
        inner = rng.uniform(1.0, 3.0)
        outer = inner + rng.uniform(1.0, 3.0)
        nrad = int(rng.integers(1, 21))
        naz = int(rng.integers(20, 41))
        for kfold in (6, 5):
            rr, th = G.hexapolar(outer, inner, nrad, naz, kfold)
            assert len(rr) % kfold == 0
            rr0, th0 = G.hexapolar(outer, 0.0, nrad, naz, kfold)
            assert len(rr0) % kfold == 1
            assert rr0[-1] == 0.0 and th0[-1] == 0.0
            # closed form of the ring rule, utils.py:514-519
            rhos = np.linspace(outer, inner, nrad)
            counts = [max(int((naz * r / outer) // kfold) * kfold, kfold) for r in rhos]
            assert len(rr) == sum(counts)
    rr, _ = G.hexapolar(4.18, 2.558, 1120, 11080)
    assert len(rr) == 9_998_544
    np.testing.assert_array_equal(bt.utils.hexapolar(4.18, 2.558, 1120, 11080, rth=True)[0], rr)


def _finish_args(dirCos, backDist, n=1.0):
    return dict(stop_surface=bt.Plane(), stop_to_ray=(np.zeros(3), np.eye(3)),
                dirCos=dirCos, n=n, backDist=backDist)


def test_oracle_grid_reference_properties():
    """tests/test_RayVector.py:368-431 on the oracle's rays: the chief ray
    sits at backDist from the origin, and on the pupil plane the lattice
    spacing is dx (even and odd nx)."""
    rng = np.random.default_rng(5772156)
    for parity in (0, 1):
        for _ in range(5):
            backDist = rng.uniform(9.0, 11.0)
            wavelength = rng.uniform(300e-9, 1100e-9)
            nx = int(rng.integers(10, 21))
            while nx % 2 != parity:
                nx = int(rng.integers(10, 21))
            lx = rng.uniform(1.0, 10.0)
            dx = lx / (nx - 2) if parity == 0 else lx / (nx - 1)
            dirCos = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1),
                               rng.uniform(-1.2, -0.8)])
            dirCos /= np.sqrt(np.dot(dirCos, dirCos))
            uu, vv = G.grid_uv(nx, nx)
            x, y = G.lattice_xy((lx, 0.0), (0.0, lx), uu, vv)
            h = G.finish(x, y, wavelength=wavelength, **_finish_args(dirCos, backDist))
            cr = (nx // 2) * nx + nx // 2 if parity == 0 else (nx * nx - 1) // 2
            r = np.array([h["x"][cr], h["y"][cr], h["z"][cr]])
            np.testing.assert_allclose(np.sqrt(np.dot(r, r)), backDist)
            np.testing.assert_allclose(h["vx"], dirCos[0], rtol=1e-15)
            # propagate to the pupil plane z = 0
            t = -h["z"] / h["vz"]
            px, py = h["x"] + h["vx"] * t, h["y"] + h["vy"] * t
            np.testing.assert_allclose(np.diff(px)[0], dx)
            np.testing.assert_allclose(np.diff(py)[0], 0, atol=1e-14)
            np.testing.assert_allclose(np.diff(px)[nx - 1], -dx * (nx - 1))
            np.testing.assert_allclose(np.diff(py)[nx - 1], dx)


def test_oracle_polar_centre_ray():
    """tests/test_RayVector.py:565-578: with inner = 0 the last ray crosses
    the pupil centre."""
    dirCos = np.array([0.03, -0.05, -1.0])
    dirCos /= np.sqrt(np.dot(dirCos, dirCos))
    rr, th = G.hexapolar(3.0, 0.0, 7, 30)
    x, y = G.polar_xy(rr, th)
    h = G.finish(x, y, wavelength=600e-9, **_finish_args(dirCos, 10.0))
    t = -h["z"][-1] / h["vz"][-1]
    assert abs(h["x"][-1] + h["vx"][-1] * t) < 1e-14
    assert abs(h["y"][-1] + h["vy"][-1] * t) < 1e-14


def test_hypot_restatement_is_libm():
    """The device's ref_hypot algorithm (restated in oracle/libm.py) equals
    x86-64 libm's hypot bit for bit -- on pupil-edge points, across the
    rescaling ranges and at the special values -- while the correctly
    naive sqrt(x*x + y*y) does not (so the check has teeth)."""
    import ctypes
    import math
    from oracle import libm
    c = ctypes.CDLL("libm.so.6")
    c.hypot.restype = ctypes.c_double
    c.hypot.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(11)
    th = rng.uniform(0, 2 * np.pi, 20000)
    pts = [(r * math.cos(t), r * math.sin(t)) for r in (4.1, 4.18, 0.41, 1.0)
           for t in th.tolist()]
    e = rng.uniform(-330, 330, (20000, 2))
    pts += list(zip((rng.normal(size=20000) * 10.0 ** np.clip(e[:, 0], -323, 307)).tolist(),
                    (rng.normal(size=20000) * 10.0 ** np.clip(e[:, 1], -323, 307)).tolist()))
    pts += [(0.0, 0.0), (-3.0, 4.0), (math.inf, math.nan), (math.nan, 2.0), (1e308, 1e308),
            (5e-324, 5e-324), (0.0, -5e-324)]
    cr_differs = 0
    for x, y in pts:
        a, b = c.hypot(x, y), libm.hypot(x, y)
        assert a == b or (math.isnan(a) and math.isnan(b)), (x, y, a, b)
        if math.isfinite(x) and math.isfinite(y) and 1e-150 < abs(x) < 1e150 \
                and 1e-150 < abs(y) < 1e150:
            cr_differs += math.sqrt(x * x + y * y) != a
    assert cr_differs > 50


def test_hexapolar_rings_vectorised():
    """The device ring table's radii, counts and azimuth steps equal the
    reference's per-ring loop (the oracle's hexapolar), element for element."""
    from batoid_amd.rayVector import _hexapolar_rings
    rng = np.random.default_rng(11)
    cases = [(4.18, 2.558, 1120, 11080, 6), (1.0, 0.0, 7, 30, 6), (0.5, 0.1, 3, 5, 6),
             (2.0, 0.0, 50, 301, 4)]
    cases += [(rng.uniform(0.1, 5), rng.uniform(0, 0.09), int(rng.integers(1, 200)),
               int(rng.integers(1, 5000)), int(rng.integers(1, 9))) for _ in range(40)]
    for outer, inner, nrad, naz, kfold in cases:
        rhos, nphis, dth = _hexapolar_rings(outer, inner, nrad, naz, kfold)
        rr, th = G.hexapolar(outer, inner, nrad, naz, kfold)
        assert int(nphis.sum()) == len(rr)
        off = np.concatenate([[0], np.cumsum(nphis)])
        for k in range(len(nphis)):
            np.testing.assert_array_equal(rr[off[k]:off[k + 1]], rhos[k])
            assert dth[k] == 2 * np.pi / int(nphis[k])
            if nphis[k] > 1:
                assert th[off[k] + 1] == dth[k]
