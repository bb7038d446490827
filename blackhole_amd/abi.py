"""ctypes mirror of include/bhx.h (the C ABI of the HIP tracer).

Field order and types must match the C header exactly; tests/test_abi.py
checks sizes and offsets against the compiled library.
"""
import ctypes as C

BHX_HORIZON, BHX_DISC, BHX_SUN, BHX_STAR, BHX_SKY, BHX_EXHAUSTED = range(6)
CLASS_NAMES = ("horizon", "disc", "sun", "star", "sky", "exhausted")
TM_LINEAR, TM_REINHARD, TM_ACES, TM_LOG, TM_HDR = range(5)

_SCENE_DOUBLES = [
    "M_kg", "spin", "cam_r", "cam_theta", "cam_phi", "cam_fov", "cam_yaw", "cam_pitch",
    "cam_roll", "disc_r_out", "eddington", "observer_time", "rtol", "atol", "r_escape",
]


class Scene(C.Structure):
    """bhx_scene: RenderConfig (src/render.h:15-53) flattened."""
    _fields_ = (
        [(n, C.c_double) for n in _SCENE_DOUBLES]
        + [("sun_centre", C.c_double * 3), ("sun_radius", C.c_double), ("sun_T_eff", C.c_double)]
        + [(n, C.c_double) for n in ("hs_r", "hs_phi0", "hs_sigma", "hs_contrast", "hs_omega")]
        + [(n, C.c_double) for n in ("ns_M_kg", "ns_R_m", "ns_R", "ns_T_eff", "ns_spin_hz",
                                     "ns_Omega", "ns_cap_tilt", "ns_cap_radius", "ns_cap_T",
                                     "ns_cap_phase0")]
        + [(n, C.c_double) for n in ("jet_z_base", "jet_z_top", "jet_R_base", "jet_k",
                                     "jet_gamma_max", "jet_z_sat", "jet_alpha", "jet_u_peak",
                                     "jet_u_wid", "jet_spine", "jet_spine_wid", "jet_zeta",
                                     "jet_a", "jet_Omega_F", "jet_j_nu0", "jet_r_bound")]
        + [("seed", C.c_uint64)]
        + [(n, C.c_int32) for n in ("width", "height", "sqrt_spp", "max_steps",
                                    "disc_enabled", "disc_sense", "stars_enabled", "star_count",
                                    "sun_enabled", "hs_enabled", "hs_sense", "ns_enabled",
                                    "ns_caps", "ns_two_caps", "jet_enabled", "threads")]
        + [("jet_model", C.c_int32), ("jet_reserved", C.c_int32)]
        + [(n, C.c_double) for n in ("jet_cone_angle", "jet_cone_r0", "jet_cone_r1")]
    )

    def copy(self):
        s = Scene()
        C.pointer(s)[0] = self
        return s

    @property
    def n_rays(self):
        S = max(1, self.sqrt_spp)
        return self.width * self.height * S * S

    @property
    def n_pixels(self):
        return self.width * self.height


class Stats(C.Structure):
    """bhx_stats: RenderStats (src/render.h:55-69) + attempt/RHS counters."""
    _fields_ = (
        [(n, C.c_int64) for n in ("rays", "steps", "captured", "disc_hits", "sun_hits",
                                  "star_hits", "jet_rays", "escaped", "exhausted")]
        + [(n, C.c_double) for n in ("max_norm_error", "max_carter_drift", "seconds",
                                     "total_flux")]
        + [("attempts", C.c_int64), ("rhs_evals", C.c_int64), ("trace_ms", C.c_double),
           ("shade_ms", C.c_double), ("order_ms", C.c_double)]
    )

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Rays(C.Structure):
    _fields_ = [("cls", C.c_void_p), ("y", C.c_void_p), ("jet", C.c_void_p),
                ("steps", C.c_void_p), ("attempts", C.c_void_p), ("norm_err", C.c_void_p),
                ("carter", C.c_void_p)]


class Shaded(C.Structure):
    _fields_ = [("xyz", C.c_void_p), ("ray_xyz", C.c_void_p), ("g", C.c_void_p),
                ("dir", C.c_void_p)]


class Launch(C.Structure):
    _fields_ = [("block", C.c_int32), ("waves_per_simd", C.c_int32), ("chunk", C.c_int32),
                ("refill", C.c_int32), ("queue_order", C.c_int32)]


ORDER_CENTRE, ORDER_COST = 1, 2   # bhx_launch.queue_order


class Bytes(C.Structure):
    """bhx_bytes: a byte stream allocated by the library (bhx_bytes_free)."""
    _fields_ = [("data", C.POINTER(C.c_uint8)), ("size", C.c_size_t)]
