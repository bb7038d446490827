"""Scene construction: RenderConfig defaults and the CLI's scene-setup rules.

Mirrors the reference's configuration layer so that a scene is described the
way a user of the reference describes it:
  * RenderConfig defaults (src/render.h:15-53), Camera (src/scene.h:343-351),
    Jet defaults (src/jet.h:105-145), NeutronStar defaults (src/neutronstar.h:43-67),
    HotSpot defaults (src/scene.h:306-313);
  * the inclination clamp to [0.5, 179.5] degrees (src/main.cpp:810);
  * the escape-radius rule r_escape = max(3 r_cam + 100, ...) (src/main.cpp:882-890).
The numerical host set-up (Jet::configure, TOV) is native code in the C ABI
library (blackhole_amd/csrc/host_physics.cpp); it makes no GPU call.
"""
import ctypes as C
import math

from .abi import Scene

M_SUN = 1.98847e30


def default_scene():
    """A bhx_scene holding the reference's RenderConfig defaults."""
    s = Scene()
    s.M_kg = 10.0 * M_SUN
    s.spin = 0.9
    s.cam_r, s.cam_theta, s.cam_phi, s.cam_fov = 60.0, math.pi / 2.0, 0.0, 0.6
    s.cam_yaw = s.cam_pitch = s.cam_roll = 0.0
    s.width, s.height, s.sqrt_spp = 1280, 720, 2
    s.disc_enabled, s.disc_r_out, s.eddington, s.disc_sense = 1, 24.0, 0.1, 1
    s.stars_enabled, s.star_count, s.seed = 1, 250000, 20240921
    s.sun_enabled = 0
    s.sun_radius, s.sun_T_eff = 1.0, 5772.0
    s.hs_enabled, s.hs_r, s.hs_phi0, s.hs_sigma, s.hs_contrast, s.hs_omega, s.hs_sense = (
        0, 8.0, 0.0, 0.6, 25.0, 0.0, 1)
    s.ns_enabled = 0
    s.ns_M_kg, s.ns_R_m, s.ns_R, s.ns_T_eff = 1.4 * M_SUN, 12000.0, 5.8, 1.0e6
    s.ns_spin_hz, s.ns_Omega = 0.0, 0.0
    s.ns_caps, s.ns_cap_tilt, s.ns_cap_radius, s.ns_cap_T = 0, math.radians(60.0), math.radians(20.0), 3.0e6
    s.ns_cap_phase0, s.ns_two_caps = 0.0, 1
    s.jet_enabled = 0
    s.jet_z_base, s.jet_z_top, s.jet_R_base, s.jet_k = 3.0, 120.0, 1.6, 0.58
    s.jet_gamma_max, s.jet_z_sat, s.jet_alpha = 10.0, 80.0, 0.7
    s.jet_u_peak, s.jet_u_wid, s.jet_spine, s.jet_spine_wid, s.jet_zeta = 0.82, 0.17, 0.06, 0.40, 2.0
    s.jet_a = s.jet_Omega_F = s.jet_j_nu0 = s.jet_r_bound = 0.0
    s.observer_time = 0.0
    s.rtol, s.atol, s.r_escape, s.max_steps = 1e-9, 1e-11, 4000.0, 20000
    s.threads = 0
    return s


def inclination(deg):
    """Camera theta for an inclination in degrees (src/main.cpp:810)."""
    return min(max(deg, 0.5), 179.5) * math.pi / 180.0


def escape_radius(s):
    """r_escape as the CLI sets it (src/main.cpp:882-890)."""
    far = s.cam_r * 3.0 + 100.0
    if s.sun_enabled:
        sep = math.sqrt(sum(c * c for c in s.sun_centre))
        far = max(far, (sep + s.sun_radius) * 3.0)
    if s.jet_enabled:
        far = max(far, s.jet_r_bound * 1.5)
    return far


def plain_scene(*, spin, incl_deg, r_cam, width, height, fov, disc_r_out, eddington,
                stars=False, star_count=250000, rtol=1e-9, atol=1e-11, max_steps=20000,
                sqrt_spp=1, M_kg=10.0 * M_SUN, seed=20240921):
    s = default_scene()
    s.M_kg = M_kg
    s.spin = spin
    s.cam_r = r_cam
    s.cam_theta = inclination(incl_deg)
    s.cam_fov = fov
    s.width, s.height, s.sqrt_spp = width, height, sqrt_spp
    s.disc_enabled, s.disc_r_out, s.eddington = 1, disc_r_out, eddington
    s.stars_enabled, s.star_count = int(stars), star_count
    s.seed = seed
    s.rtol, s.atol, s.max_steps = rtol, atol, max_steps
    s.r_escape = escape_radius(s)
    return s


def plain_frame(width=256, height=256):
    """Plain FP64 render: a=0.9, 10 Msun, r=500, 80 deg, +-20M image plane,
    NT disc ISCO-20M at 0.1 Eddington, no stars, rtol 1e-10 / atol 1e-12."""
    return plain_scene(spin=0.9, incl_deg=80.0, r_cam=500.0, width=width, height=height,
                       fov=2.0 * math.atan(20.0 / 500.0), disc_r_out=20.0, eddington=0.1,
                       rtol=1e-10, atol=1e-12)


def starry_frame(width=1024, height=1024, stars=True):
    """Plain FP64 render at 1024^2: a=0.998, 60 deg, r=1000, NT disc ISCO-50M at
    0.3 Eddington, 250k-star sky (+-55M image plane, 10 Msun)."""
    return plain_scene(spin=0.998, incl_deg=60.0, r_cam=1000.0, width=width, height=height,
                       fov=2.0 * math.atan(0.055), disc_r_out=50.0, eddington=0.3, stars=stars)


def uniform_sky(count=250000, seed=20240921):
    """The synthetic sky of the 1024^2 frame: count stars uniform on the
    sphere, magnitudes from N(<m) ~ 10^(0.45 m) up to m = 8, B-V ~ N(0.6, 0.4)
    (bhx_uniform_sky).  Returns host arrays (x, y, z, T, irradiance) for
    Tracer.set_catalogue."""
    import numpy as np
    from . import native
    arrs = [np.empty(count) for _ in range(5)]
    native.check(None, native.lib().bhx_uniform_sky(int(count), int(seed),
                                                    *[a.ctypes.data_as(C.c_void_p) for a in arrs]),
                 "bhx_uniform_sky")
    return arrs


def sky_frame(width=1024, height=1024):
    """The 1024^2 frame as configured: a = 0.998, 60 deg, r = 1000 M, NT disc
    ISCO-50 M at 0.3 Eddington, +-55 M image plane, 10 Msun, rtol 1e-10 /
    atol 1e-12 (the tolerances of the 256^2 frame), stars on; its sky is the
    catalogue of uniform_sky(250000, seed), loaded with Tracer.set_catalogue."""
    return plain_scene(spin=0.998, incl_deg=60.0, r_cam=1000.0, width=width, height=height,
                       fov=2.0 * math.atan(0.055), disc_r_out=50.0, eddington=0.3, stars=True,
                       rtol=1e-10, atol=1e-12)


def setup_jet(s, phi_flux=50.0, eps_rad=0.02):
    """Normalise the scene's jets in place (src/main.cpp:399-423, src/jet.h:261-311);
    clears jet_enabled where the reference's CLI switches the jets off.  phi_flux
    and eps_rad are the Jet defaults (src/jet.h:140-141)."""
    from . import native
    native.check(None, native.lib().bhx_setup_jet(C.byref(s), float(phi_flux), float(eps_rad)),
                 "bhx_setup_jet")
    return s


def setup_conical_jet(s, half_angle_deg=5.0, r0=6.0, r1=500.0, gamma=10.0, alpha=0.7,
                      phi_flux=50.0, eps_rad=0.02):
    """Give the scene a conical jet model of this package (no reference counterpart): two
    cones of half opening half_angle_deg about the spin axis from r0 to r1
    (GM/c^2), radial outflow at Lorentz factor gamma, comoving emissivity
    ~ (r0/r)^2 nu^-alpha, normalised like the reference's jets (the radiated
    fraction eps_rad of the Blandford-Znajek power)."""
    from . import native
    s.jet_enabled, s.jet_model = 1, 1
    s.jet_cone_angle, s.jet_cone_r0, s.jet_cone_r1 = math.radians(half_angle_deg), r0, r1
    s.jet_gamma_max, s.jet_alpha = gamma, alpha
    native.check(None, native.lib().bhx_setup_conical_jet(C.byref(s), float(phi_flux), float(eps_rad)),
                 "bhx_setup_conical_jet")
    return s


def polytrope_star(K_si, gamma, M_kg):
    """(R_m, M_kg) of the stable star of mass M_kg for P = K rho^gamma
    (src/tov.h:205-218); BhxError when the equation of state cannot hold it."""
    from . import native
    R, M = C.c_double(), C.c_double()
    rc = native.lib().bhx_polytrope_star(float(K_si), float(gamma), float(M_kg), C.byref(R), C.byref(M))
    if rc != 0:
        raise native.BhxError(f"polytrope K={K_si} Gamma={gamma} cannot support M={M_kg} kg")
    return R.value, M.value


def neutron_star_scene(s, R_m, M_kg, *, frame=True):
    """Turn s into a neutron-star scene for a solved star (src/main.cpp:503-520):
    Schwarzschild exterior, no disc, no jet; with frame=True the camera is put
    at 40 R and the field of view fitted to the lensed radius R/sqrt(1-2/R)."""
    from . import native
    native.check(None, native.lib().bhx_setup_neutron_star(C.byref(s), float(R_m), float(M_kg)),
                 "bhx_setup_neutron_star")
    if frame:
        s.cam_r = 40.0 * s.ns_R
        s.cam_fov = 5.0 * math.atan2(s.ns_R / math.sqrt(1.0 - 2.0 / s.ns_R), s.cam_r)
    s.r_escape = escape_radius(s)
    return s


# The Gamma = 2, K = 100 polytrope of relativistic-star test problems (G = c =
# M_sun = 1) in SI units for P = K rho^2, rho the total mass-energy density.
G_SI, C_SI = 6.67430e-11, 2.99792458e8
POLYTROPE_K100_SI = 100.0 * (G_SI * M_SUN / C_SI ** 2) ** 3 * C_SI ** 2 / M_SUN


def survey_batch(seed=20240921, n=64, n_ns=8, side_min=128, side_max=1024):
    """A seeded batch of uneven scenes: spin U[0, 0.998], inclination U[5, 85]
    deg, observer radius log-U[100, 2000] M, square side log-U[side_min,
    side_max], disc at log-U[0.01, 1] Eddington out to 30 M, image plane
    +-25 M; n_ns of them hold instead a neutron star of mass U[1.2, 1.99]
    M_sun on the K = 100 polytrope (hot caps, 200 Hz), framed by the CLI's
    rule (src/main.cpp:514-519) at the drawn distance.  No star field."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns_idx = set(rng.choice(n, size=n_ns, replace=False).tolist()) if n_ns else set()
    out = []
    for i in range(n):
        a = rng.uniform(0.0, 0.998)
        incl = rng.uniform(5.0, 85.0)
        r_cam = math.exp(rng.uniform(math.log(100.0), math.log(2000.0)))
        side = int(round(math.exp(rng.uniform(math.log(side_min), math.log(side_max)))))
        edd = math.exp(rng.uniform(math.log(0.01), math.log(1.0)))
        m_ns = rng.uniform(1.2, 1.99)
        s = plain_scene(spin=a, incl_deg=incl, r_cam=r_cam, width=side, height=side,
                        fov=2.0 * math.atan(25.0 / r_cam), disc_r_out=30.0, eddington=edd)
        if i in ns_idx:
            R_m, M_kg = polytrope_star(POLYTROPE_K100_SI, 2.0, m_ns * M_SUN)
            s.ns_spin_hz, s.ns_caps = 200.0, 1
            neutron_star_scene(s, R_m, M_kg, frame=False)
            b_app = s.ns_R / math.sqrt(1.0 - 2.0 / s.ns_R)
            s.cam_fov = 5.0 * math.atan2(b_app, s.cam_r)
        out.append(s)
    return out


def disc_frame(width=512, height=512, spin=0.7, incl_deg=45.0, eddington=0.1, r_cam=500.0,
               rtol=1e-10, atol=1e-12):
    """The differentiation scene: NT disc ISCO-20M, no stars, camera at r_cam
    (default 500 M) and inclination incl_deg over a +-20 M image plane;
    r_escape fixed at 3 * 500 + 100 so that varying r_cam does not move the
    escape sphere.  Tolerances 1e-10 / 1e-12: the derivative of the
    integrator with its step sizes frozen approaches the converged derivative
    like the tolerance to a power above one (at 1e-9 single rays differ from
    it by ~1e-5, at 1e-10 by ~4e-7)."""
    s = plain_scene(spin=spin, incl_deg=incl_deg, r_cam=r_cam, width=width, height=height,
                    fov=2.0 * math.atan(20.0 / 500.0), disc_r_out=20.0, eddington=eddington,
                    rtol=rtol, atol=atol)
    s.r_escape = 3.0 * 500.0 + 100.0
    return s


def quasar_frame(incl_deg, width=1024, height=1024, stars=True, half_extent=1050.0):
    """One frame of the quasar sweep: a = 0.998, NT disc ISCO-1000M at 0.5
    Eddington, the conical jet (5 deg half opening, Gamma 10, (r0/r)^2 out to
    500 M), 250k-star sky; camera at 2500 M (outside the disc) framing
    +-half_extent M (default the whole disc)."""
    s = plain_scene(spin=0.998, incl_deg=incl_deg, r_cam=2500.0, width=width, height=height,
                    fov=2.0 * math.atan(half_extent / 2500.0), disc_r_out=1000.0, eddington=0.5,
                    stars=stars)
    setup_conical_jet(s)
    s.r_escape = escape_radius(s)
    return s


def quasar_diff_frame(incl_deg, width=64, height=64, half_extent=120.0, rtol=1e-10, atol=1e-12):
    """The quasar frame as a differentiation scene: no stars (the adjoint's
    intensity is the disc's and the jet's), tolerances 1e-10 / 1e-12 and
    r_escape fixed at 3 * 2500 + 100 so that varying r_cam does not move the
    escape sphere."""
    s = quasar_frame(incl_deg, width, height, stars=False, half_extent=half_extent)
    s.rtol, s.atol = rtol, atol
    s.r_escape = 3.0 * 2500.0 + 100.0
    return s


def quasar_sweep(n_frames=120, width=1024, height=1024, stars=True):
    """Inclination swept linearly 0 -> 90 deg over n_frames (clamped at 0.5
    deg like the CLI, src/main.cpp:810)."""
    return [quasar_frame(90.0 * i / (n_frames - 1), width, height, stars) for i in range(n_frames)]
