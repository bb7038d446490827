"""Write tests/golden/intensity_reference.json from the reference build.

The differentiable render's loss input is the bolometric intensity of a disc
hit, spec::blackbody_bolometric (src/spectrum.cpp:24-26) of the observed
temperature g T, with T the Novikov-Thorne temperature of the disc at the hit
radius (NovikovThorneDisc::temperature, src/disc.h:100-104).  This script
evaluates both through oracle/_ref/libbhref.so (the reference compiled from
its sources) on a grid of radii and redshifts for the disc scenes the adjoint
is used on, and stores inputs and outputs (as hex doubles, bit-exact).

    python tests/golden/make_golden_intensity.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from blackhole_amd import scene as S  # noqa: E402


def scenes():
    """(name, scene) pairs: the differentiation scene and its target, the
    quasar disc, a retrograde disc and a low-rate disc."""
    out = [("diff_a0.7", S.disc_frame(64, 64)),
           ("target_a0.9", S.disc_frame(64, 64, spin=0.9, incl_deg=50.0)),
           ("quasar", S.quasar_diff_frame(30.0)),
           ("retrograde", S.disc_frame(64, 64, spin=0.5))]
    out[3][1].disc_sense = -1
    lo = S.disc_frame(64, 64, eddington=1e-4)
    lo.M_kg = 4.0e6 * S.M_SUN
    out.append(("low_rate", lo))
    return out


def main():
    if not oracle.ref_available():
        raise SystemExit("reference shim not built: make -C oracle (needs /root/reference)")
    R = oracle.ref()
    R.bhref_blackbody_bolometric.restype = C.c_double
    R.bhref_blackbody_bolometric.argtypes = [C.c_double]
    R.bhref_disc_temperature.restype = C.c_double
    R.bhref_disc_temperature.argtypes = [C.c_void_p, C.c_double]
    g = np.array([0.05, 0.31, 0.7, 1.0, 1.37, 2.9])
    doc = {"note": "written by tests/golden/make_golden_intensity.py from oracle/_ref", "scenes": []}
    for name, sc in scenes():
        r_isco = oracle.lib().oracle_isco(sc.spin, sc.disc_sense)
        r = np.concatenate([[r_isco * 0.99, r_isco, r_isco * (1 + 1e-9), r_isco * 1.01],
                            np.geomspace(r_isco * 1.05, sc.disc_r_out, 25), [sc.disc_r_out * 1.01]])
        T = np.array([R.bhref_disc_temperature(C.byref(sc), float(x)) for x in r])
        I = np.array([[R.bhref_blackbody_bolometric(float(gg * t)) for gg in g] for t in T])
        doc["scenes"].append({
            "name": name, "M_kg": sc.M_kg, "spin": sc.spin, "disc_r_out": sc.disc_r_out,
            "eddington": sc.eddington, "disc_sense": sc.disc_sense,
            "r": [float(x).hex() for x in r], "g": [float(x).hex() for x in g],
            "T": [float(x).hex() for x in T], "I": [[float(x).hex() for x in row] for row in I]})
    with open(os.path.join(HERE, "intensity_reference.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", sum(len(s["r"]) for s in doc["scenes"]), "radii")


if __name__ == "__main__":
    main()
