"""Mathematical pins of the oracle's conical jet and disc intensity (CPU).

The conical jet of the quasar sweep is this package's own model: the
reference has no counterpart (its jets are src/jet.h), so no reference output
can pin it.  Each piece is pinned here by what the mathematics fixes, with
the metric, the ZAMO frame and the Doppler formula computed independently in
numpy from the textbook Boyer-Lindquist Kerr metric:

  cone velocity   u.u = -1; zero angular momentum u_phi = 0 (radial outflow
                  in the ZAMO frame); Lorentz factor relative to the ZAMO
                  alpha u^t = Gamma with alpha = (-g^tt)^(-1/2) from a numeric
                  inverse of g; u^theta = 0, u^r > 0.  Grid: spins 0 .. 0.998
                  and -0.7, radii from 1.001 r_+ (inside the ergosphere) to
                  1e7, polar angles near both axes and at the equator.
  redshift        g = 1 / (Gamma (1 - beta cos psi)), psi the angle between
                  the photon's direction and the outflow in the ZAMO frame
                  (the special-relativistic Doppler factor; a local Lorentz
                  transformation, so exact at every point, and at a = 0,
                  r = 1e7 the flat-space formula); the jet's boost g^(2+alpha).
  sampling gate   never skips a step whose dense output enters the cone
                  (brute force on 2e5 random quintic segments sampled at 2049
                  points), and is the exact bound for each single term of the
                  Hairer interpolant (max s = 1, s(1-s) = 1/4, s^2(1-s) = 4/27,
                  s^2(1-s)^2 = 1/16): a segment that enters the cone by 1e-6
                  of its reach is sampled, one that misses by 1e-3 is not.
  profile         zero outside the cones and the radial range; (r0/r)^2 in the
                  core (|cos theta| >= cos(0.8 theta_c), 1.5 r0 <= r <=
                  0.85 r1); even in cos theta; continuous with continuous
                  first and second derivatives at all six edges (the one-sided
                  deviation grows like the cube of the distance: doubling the
                  distance multiplies it by 8).
  disc intensity  sigma (g T)^4 / pi equals the reference's
                  spec::blackbody_bolometric(g T) and T its
                  NovikovThorneDisc::temperature, bit for bit, on the golden
                  grid of tests/golden/intensity_reference.json (written by the
                  reference build), and live through oracle/_ref where built.

test_pins_catch_planted_mistakes compiles copies of the oracle with one
mistake planted each (a dropped sqrt(Delta/Sigma), a wrong lapse, the wrong
sign of the frame dragging, beta from 1/Gamma instead of 1/Gamma^2, a gate
coefficient halved, the southern cone forgotten by the gate, a cubic instead
of a quintic step, r^-1 instead of r^-2, the taper's inner edge moved, g^3
instead of g^4) and requires the named pin to fail on each.
"""
import ctypes as C
import json
import math
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from blackhole_amd import scene as S

HERE = os.path.dirname(os.path.abspath(__file__))
P = C.c_void_p
D = C.c_double


def bind(L):
    L.oracle_cone_velocity.argtypes = [D, D, D, D, P]
    L.oracle_cone_shape.argtypes = [P, C.c_int64, P, P, P]
    L.oracle_cone_gate.argtypes = [P, C.c_int64, P, P]
    L.oracle_redshift.restype = D
    L.oracle_redshift.argtypes = [P, P]
    L.oracle_jet_boost.restype = D
    L.oracle_jet_boost.argtypes = [D, D]
    L.oracle_make_photon.argtypes = [D, D, D, D, P, P]
    L.oracle_disc_temperature.restype = D
    L.oracle_disc_temperature.argtypes = [P, D]
    L.oracle_disc_intensity.argtypes = [P, C.c_int64, P, P, P, P]
    return L


def _p(a):
    return a.ctypes.data_as(P)


# ---- independent Kerr geometry (Boyer-Lindquist, G = c = M = 1) ------------
def kerr_metric(a, r, th):
    """Covariant metric g_{mu nu} in (t, r, theta, phi)."""
    s, c = math.sin(th), math.cos(th)
    Sig = r * r + a * a * c * c
    Del = r * r - 2 * r + a * a
    g = np.zeros((4, 4))
    g[0, 0] = -(1 - 2 * r / Sig)
    g[0, 3] = g[3, 0] = -2 * a * r * s * s / Sig
    g[1, 1] = Sig / Del
    g[2, 2] = Sig
    g[3, 3] = (r * r + a * a + 2 * a * a * r * s * s / Sig) * s * s
    return g


def horizon(a):
    return 1 + math.sqrt(1 - a * a)


def grid():
    pts = []
    for a in (0.0, 0.3, 0.9, 0.998, -0.7):
        rh = horizon(a)
        for r in (rh * 1.001, rh * 1.01, rh * 1.2, 1.9, 2.5, 6.0, 50.0, 1e3, 1e7):
            if r <= rh:
                continue
            for th in (1e-3, 0.02, 0.087, 0.5, 1.2, math.pi / 2, math.pi - 0.02, math.pi - 1e-3):
                pts.append((a, r, th))
    return pts


GAMMA = 10.0
BETA = math.sqrt(1 - 1 / GAMMA ** 2)


def pin_velocity(L):
    """u.u = -1, u_phi = 0, alpha u^t = Gamma, u^theta = 0, u^r > 0."""
    bad = []
    ergo = 0
    for a, r, th in grid():
        u = np.zeros(4)
        L.oracle_cone_velocity(a, r, th, GAMMA, _p(u))
        g = kerr_metric(a, r, th)
        ergo += g[0, 0] > 0
        terms = np.abs(np.outer(u, u) * g)
        uu = u @ g @ u
        uphi = g[3] @ u
        alpha = 1.0 / math.sqrt(-np.linalg.inv(g)[0, 0])
        if abs(uu + 1) > 1e-12 * (terms.sum() + 1):
            bad.append(("u.u", a, r, th, uu))
        if abs(uphi) > 1e-12 * (abs(g[3, 0] * u[0]) + abs(g[3, 3] * u[3]) + 1e-300):
            bad.append(("u_phi", a, r, th, uphi))
        if abs(alpha * u[0] / GAMMA - 1) > 1e-9:
            bad.append(("Gamma", a, r, th, alpha * u[0]))
        if u[2] != 0.0 or not u[1] > 0.0:
            bad.append(("direction", a, r, th, u))
    assert ergo > 20, "the grid must reach into the ergosphere"
    assert not bad, bad[:5]


def pin_doppler(L):
    """Redshift of a photon against the outflow = the special-relativistic
    Doppler factor in the ZAMO frame; boost = g^(2+alpha)."""
    rng = np.random.default_rng(7)
    bad = []
    for a, r, th in grid():
        g = kerr_metric(a, r, th)
        ginv = np.linalg.inv(g)
        alpha = 1.0 / math.sqrt(-ginv[0, 0])
        omega = -g[0, 3] / g[3, 3]
        u = np.zeros(4)
        L.oracle_cone_velocity(a, r, th, GAMMA, _p(u))
        for _ in range(6):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            y = np.zeros(8)
            L.oracle_make_photon(a, r, th, 0.0, _p(np.ascontiguousarray(d)), _p(y))
            # photon energy and radial momentum in the ZAMO frame, from the
            # covariant momentum (the traced momentum is past-directed)
            E = (y[4] + omega * y[7]) / alpha
            k_r = -y[5] / math.sqrt(g[1, 1])
            cos_psi = k_r / E
            expect = 1.0 / (GAMMA * (1.0 - BETA * cos_psi))
            got = L.oracle_redshift(_p(y), _p(u))
            if abs(got / expect - 1) > 1e-10:
                bad.append((a, r, th, d, got, expect))
    # flat space: a = 0, r = 1e7, head-on and tail-on along the outflow
    for d_r, cos_psi in ((-1.0, 1.0), (1.0, -1.0), (0.0, 0.0)):
        d = np.array([d_r, math.sqrt(1 - d_r * d_r), 0.0])
        y = np.zeros(8)
        L.oracle_make_photon(0.0, 1e7, 0.03, 0.0, _p(d), _p(y))
        u = np.zeros(4)
        L.oracle_cone_velocity(0.0, 1e7, 0.03, GAMMA, _p(u))
        got = L.oracle_redshift(_p(y), _p(u))
        expect = 1.0 / (GAMMA * (1.0 - BETA * cos_psi))
        if abs(got / expect - 1) > 1e-6:
            bad.append(("flat", d_r, got, expect))
    for gg in (0.05, 0.9, 1.0, 3.7, 19.9):
        if abs(L.oracle_jet_boost(0.7, gg) / gg ** 2.7 - 1) > 1e-14:
            bad.append(("boost", gg))
    assert not bad, bad[:5]


def _scene():
    s = S.quasar_diff_frame(30.0)
    return s


def theta_of(c, s):
    """Hairer dense output of theta: c1 + s (c2 + (1-s) (c3 + s (c4 + (1-s) c5)))."""
    s1 = 1.0 - s
    return c[0] + s * (c[1] + s1 * (c[2] + s * (c[3] + s1 * c[4])))


def gate(L, sc, c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    n = c.shape[1]
    out = np.zeros(n, np.uint8)
    L.oracle_cone_gate(C.byref(sc), n, _p(c), _p(out))
    return out.astype(bool)


def pin_gate(L):
    """No step that enters the cone is skipped; the bound is exact per term."""
    sc = _scene()
    tc = sc.jet_cone_angle
    # single terms: reach |c2|, |c3|/4, 4|c4|/27, |c5|/16 is attained
    reach = {1: 1.0, 2: 0.25, 3: 4.0 / 27.0, 4: 1.0 / 16.0}
    for k, f in reach.items():
        for m in (1e-4, 0.01, 0.2):
            for pole in ("north", "south"):
                for eps, must in ((-1e-6, True), (1e-3, False)):
                    cc = np.zeros((5, 1))
                    psi0 = tc + f * m * (1 + eps)
                    cc[0] = psi0 if pole == "north" else math.pi - psi0
                    cc[k] = -m if pole == "north" else m
                    got = gate(L, sc, cc)[0]
                    assert got == must, (k, m, pole, eps, got)
    rng = np.random.default_rng(11)
    n = 200_000
    c = np.empty((5, n))
    north = rng.random(n) < 0.5
    c[0] = rng.uniform(-0.02, 0.45, n)
    c[0] = np.where(north, c[0], math.pi - c[0])
    mag = np.exp(rng.uniform(math.log(1e-5), math.log(0.4), (4, n)))
    keep = rng.random((4, n)) < 0.7
    c[1:] = mag * rng.choice([-1.0, 1.0], (4, n)) * keep
    s = np.linspace(0.0, 1.0, 2049)[:, None]
    enters = np.zeros(n, bool)
    for b in range(0, n, 10_000):   # 2049 x 10k samples at a time
        th = theta_of(c[:, None, b:b + 10_000], s)
        enters[b:b + 10_000] = (np.abs(np.cos(th)) > math.cos(tc)).any(axis=0)
    g = gate(L, sc, c)
    missed = np.flatnonzero(enters & ~g)
    assert enters.sum() > 10_000 and (~g).sum() > 10_000
    assert missed.size == 0, (missed[:5], c[:, missed[:5]].T)


def shape(L, sc, r, cs):
    r = np.ascontiguousarray(np.broadcast_to(r, np.broadcast(r, cs).shape), dtype=np.float64)
    cs = np.ascontiguousarray(np.broadcast_to(cs, r.shape), dtype=np.float64)
    out = np.zeros(r.shape)
    L.oracle_cone_shape(C.byref(sc), r.size, _p(r), _p(cs), _p(out))
    return out


def pin_profile(L):
    """Support, (r0/r)^2 law in the core, symmetry, C^2 edges."""
    sc = _scene()
    tc, r0, r1 = sc.jet_cone_angle, sc.jet_cone_r0, sc.jet_cone_r1
    rng = np.random.default_rng(5)
    # support
    r = rng.uniform(0.5 * r0, 1.2 * r1, 20000)
    th = rng.uniform(0, math.pi, 20000)
    f = shape(L, sc, r, np.cos(th))
    outside = (r <= r0) | (r >= r1) | (np.minimum(th, math.pi - th) >= tc)
    assert np.all(f[outside] == 0.0) and np.all(f[~outside] >= 0.0)
    assert np.all(f <= (r0 / r) ** 2 * (1 + 1e-15))
    # core law
    r = np.geomspace(1.5 * r0, 0.85 * r1, 400)
    for th0 in (0.0, 0.3 * tc, 0.8 * tc * (1 - 1e-12)):
        for cs in (math.cos(th0), -math.cos(th0)):
            f = shape(L, sc, r, cs)
            assert np.all(np.abs(f * (r / r0) ** 2 - 1) < 1e-15), th0
    # symmetry
    r = rng.uniform(r0, r1, 5000)
    cs = rng.uniform(math.cos(tc) - 0.01, 1.0, 5000)
    assert np.array_equal(shape(L, sc, r, cs), shape(L, sc, r, -cs))

    # C^2 edges: along theta at r = 40 (both taper edges) and along r on the
    # axis (the four ramp edges); h = shape * (r/r0)^2 = taper x ramps
    def h_theta(t):
        return shape(L, sc, np.full_like(t, 40.0), np.cos(t)) * (40.0 / r0) ** 2

    def h_r(x):
        return shape(L, sc, x, np.ones_like(x)) * (x / r0) ** 2

    edges = [(h_theta, tc, -1, 0.2 * tc), (h_theta, 0.8 * tc, +1, 0.2 * tc),
             (h_r, r0, +1, 0.5 * r0), (h_r, 1.5 * r0, -1, 0.5 * r0),
             (h_r, 0.85 * r1, +1, 0.15 * r1), (h_r, r1, -1, 0.15 * r1)]
    for fn, xe, side, width in edges:
        # side: the direction from the edge INTO the transition
        flat = fn(np.array([xe - side * 1e-3 * width, xe - side * 0.3 * width]))
        level = round(flat[1])   # the plateau: 0 outside, 1 in the core
        assert np.all(np.abs(flat - level) < 1e-15), ("not flat outside the transition", xe, flat)
        d = np.array([1e-3, 2e-3, 4e-3]) * width
        dev = np.abs(fn(xe + side * d) - level)
        assert dev[0] > 0, ("no transition", xe)
        for k in range(2):
            ratio = dev[k + 1] / dev[k]
            assert 7.5 < ratio < 8.5, ("edge not C^2 (deviation not cubic)", xe, ratio)


def _golden_intensity():
    with open(os.path.join(HERE, "golden", "intensity_reference.json")) as f:
        return json.load(f)


def _h(xs):
    return np.array([float.fromhex(x) for x in xs])


def pin_intensity(L):
    """sigma (g T)^4 / pi and T(r) bit for bit against the reference build."""
    doc = _golden_intensity()
    n_checked = 0
    for sd in doc["scenes"]:
        sc = S.disc_frame(8, 8)
        sc.M_kg, sc.spin, sc.disc_r_out, sc.eddington, sc.disc_sense = (
            sd["M_kg"], sd["spin"], sd["disc_r_out"], sd["eddington"], sd["disc_sense"])
        r, g, T, I = _h(sd["r"]), _h(sd["g"]), _h(sd["T"]), np.array([_h(row) for row in sd["I"]])
        Tm = np.array([L.oracle_disc_temperature(C.byref(sc), float(x)) for x in r])
        assert np.array_equal(Tm, T), (sd["name"], np.flatnonzero(Tm != T))
        rr = np.ascontiguousarray(np.repeat(r, len(g)))
        gg = np.ascontiguousarray(np.tile(g, len(r)))
        cls = np.ones(len(rr), np.uint8)
        out = np.zeros(len(rr))
        L.oracle_disc_intensity(C.byref(sc), len(rr), _p(cls), _p(rr), _p(gg), _p(out))
        assert np.array_equal(out, I.reshape(-1)), (sd["name"], np.flatnonzero(out != I.reshape(-1))[:5])
        # only disc hits with g > 0 count
        cls[:] = 4
        L.oracle_disc_intensity(C.byref(sc), len(rr), _p(cls), _p(rr), _p(gg), _p(out))
        assert not out.any()
        n_checked += np.count_nonzero(I)
    assert n_checked > 500


PINS = {"velocity": pin_velocity, "doppler": pin_doppler, "gate": pin_gate, "profile": pin_profile,
        "intensity": pin_intensity}


@pytest.fixture(scope="module")
def L():
    return bind(oracle.lib())


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(L, name):
    PINS[name](L)


def test_intensity_live_against_reference():
    """The reference's own bolometric function and disc temperature, live."""
    if not oracle.ref_available():
        pytest.skip("reference build (oracle/_ref) not present")
    R = oracle.ref()
    R.bhref_blackbody_bolometric.restype = D
    R.bhref_blackbody_bolometric.argtypes = [D]
    R.bhref_disc_temperature.restype = D
    R.bhref_disc_temperature.argtypes = [P, D]
    L = bind(oracle.lib())
    rng = np.random.default_rng(2)
    for spin, edd in ((0.7, 0.1), (0.998, 0.5), (0.2, 1e-3), (-0.5, 0.3)):
        sc = S.disc_frame(8, 8, spin=spin, eddington=edd)
        r = rng.uniform(1.0, 25.0, 300)
        g = rng.uniform(0.01, 3.0, 300)
        cls = np.ones(300, np.uint8)
        out = np.zeros(300)
        L.oracle_disc_intensity(C.byref(sc), 300, _p(cls), _p(r), _p(g), _p(out))
        for i in range(300):
            T = R.bhref_disc_temperature(C.byref(sc), r[i])
            assert L.oracle_disc_temperature(C.byref(sc), r[i]) == T or (math.isnan(T))
            want = R.bhref_blackbody_bolometric(g[i] * T) if T > 0 else 0.0
            assert out[i] == want, (spin, r[i], g[i], out[i], want)


# ---- planted mistakes --------------------------------------------------------
def _func_span(src, signature):
    i = src.index(signature)
    j = src.index("\n    }\n", i)
    return i, j


MUTANTS = [
    # (name, function signature or None for the whole file, old, new, pin that must fail)
    ("no sqrt(Delta/Sigma) in u^r", "void cone_velocity(", "G * beta * std::sqrt(p.D / p.S)", "G * beta",
     "velocity"),
    ("lapse without Sigma", "void cone_velocity(", "std::sqrt(p.S * p.D / p.A)", "std::sqrt(p.D / p.A)",
     "velocity"),
    ("frame dragging sign", "void cone_velocity(", "u[3] = G * omega / lapse", "u[3] = -G * omega / lapse",
     "velocity"),
    ("beta from 1/Gamma", "void cone_velocity(", "1.0 - 1.0 / (G * G)", "1.0 - 1.0 / G", "velocity"),
    ("gate c3/8", "bool cone_gate(", "std::fabs(c3) / 4.0", "std::fabs(c3) / 8.0", "gate"),
    ("gate c4 coefficient", "bool cone_gate(", "4.0 * std::fabs(c4) / 27.0", "std::fabs(c4) / 27.0", "gate"),
    ("gate forgets the southern cone", "bool cone_gate(", "std::min(c1, 3.14159265358979323846 - c1)", "c1",
     "gate"),
    ("cubic smooth step", "static double smooth_c2(", "t * t * t * (10.0 + t * (6.0 * t - 15.0))",
     "t * t * (3.0 - 2.0 * t)", "profile"),
    ("profile r^-1", "double cone_shape(", "return q * q * ta * ramp;", "return q * ta * ramp;", "profile"),
    ("taper edge at 0.9", "double cone_shape(", "std::cos(0.8 * cone_angle)", "std::cos(0.9 * cone_angle)",
     "profile"),
    ("g^3 intensity", "void oracle_disc_intensity(", "kSigmaSB * gT * gT * gT * gT / kPi",
     "kSigmaSB * gT * gT * gT / kPi", "intensity"),
]


def _build_mutant(tmp, idx, sig, old, new):
    with open(os.path.join(os.path.dirname(oracle.LIB), "..", "src", "oracle.cpp")) as f:
        src = f.read()
    i, j = _func_span(src, sig)
    body = src[i:j]
    assert body.count(old) == 1, (sig, old)
    src = src[:i] + body.replace(old, new) + src[j:]
    d = os.path.join(tmp, f"m{idx}")
    os.makedirs(d)
    srcdir = os.path.join(os.path.dirname(oracle.LIB), "..", "src")
    for name in ("oracle.h",):
        with open(os.path.join(srcdir, name)) as f, open(os.path.join(d, name), "w") as g:
            g.write(re.sub(r'#include "\.\./\.\./include/bhx\.h"',
                           '#include "' + os.path.join(os.path.dirname(HERE), "include", "bhx.h") + '"',
                           f.read()))
    path = os.path.join(d, "oracle.cpp")
    with open(path, "w") as f:
        f.write(src)
    so = os.path.join(d, "libmut.so")
    return subprocess.Popen(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-pthread",
                             path, "-o", so], stderr=subprocess.PIPE), so


def test_pins_catch_planted_mistakes():
    with tempfile.TemporaryDirectory() as tmp:
        jobs = [_build_mutant(tmp, k, sig, old, new) for k, (_, sig, old, new, _) in enumerate(MUTANTS)]
        for (proc, _), m in zip(jobs, MUTANTS):
            _, err = proc.communicate()
            assert proc.returncode == 0, (m[0], err.decode()[-2000:])
        for (_, so), (name, _, _, _, pin) in zip(jobs, MUTANTS):
            L = bind(C.CDLL(so))
            with pytest.raises(AssertionError):
                PINS[pin](L)
