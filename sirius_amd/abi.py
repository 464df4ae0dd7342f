"""ctypes mirror of include/sirius_hip.h (the C-ABI of libsirius_hip.so).

Struct layouts must match the header field for field; tests/test_abi.py checks sizes and
offsets against the compiled library's own sizeof/offsetof report.
"""
import ctypes as C

import numpy as np

ABI_VERSION = 1

OK, E_INVALID, E_DEVICE, E_NOMEM, E_STATE, E_UNSUPPORTED, E_IO = 0, -1, -2, -3, -4, -5, -6
METRIC_KERR_SCHILD, METRIC_MORRIS_THORNE, METRIC_ALCUBIERRE = 0, 1, 2
INTEGRATOR_DP45, INTEGRATOR_SYMPLECTIC = 0, 1
PRECISION_FP64, PRECISION_FP32_ESCALATE = 0, 1
EMISSION_NONE, EMISSION_THIN_DISK, EMISSION_VOLUMETRIC = 0, 1, 2
SCHEDULE_GRID, SCHEDULE_PERSISTENT = 0, 1
RAY_ESCAPED, RAY_CAPTURED, RAY_DISK, RAY_MAX_STEPS, RAY_SPIRAL, RAY_FAILED = 0, 1, 2, 3, 4, 5
FLAG_ESCALATED, FLAG_SURROGATE, FLAG_RENORM, FLAG_AUDITED, FLAG_FAR_SIDE, FLAG_CRITICAL = 1, 2, 4, 8, 16, 32

MLP_IN, MLP_HIDDEN, MLP_LAYERS, MLP_OUT = 5, 256, 4, 6
IMAGE_PPM, IMAGE_PNG, IMAGE_EXR, IMAGE_PFM, IMAGE_RAW, IMAGE_PPM_SRGB = 0, 1, 2, 3, 4, 5
RAYS_FIELDS = ("cls", "flags", "theta", "phi", "lambda", "disk_r", "redshift", "null_drift", "steps",
               "intensity", "state")  # bit i of SIRIUS_RAYS_* is RAYS_FIELDS[i]
RAYS_ALL = (1 << len(RAYS_FIELDS)) - 1


class Metric(C.Structure):
    _fields_ = [
        ("family", C.c_int32), ("reserved0", C.c_int32),
        ("M", C.c_double), ("a", C.c_double), ("Q", C.c_double), ("Lambda", C.c_double),
        ("throat_b0", C.c_double),
        ("warp_v", C.c_double), ("warp_R", C.c_double), ("warp_sigma", C.c_double),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("r_obs", C.c_double), ("theta_obs", C.c_double), ("phi_obs", C.c_double),
        ("fov_deg", C.c_double),
        ("yaw", C.c_double), ("pitch", C.c_double), ("roll", C.c_double),
        ("disk_r_in", C.c_double), ("disk_r_out", C.c_double),
        ("width", C.c_int32), ("height", C.c_int32),
        ("metric_index", C.c_int32), ("reserved0", C.c_int32),
    ]


class TraceOpts(C.Structure):
    _fields_ = [
        ("integrator", C.c_int32), ("precision", C.c_int32),
        ("emission", C.c_int32), ("schedule", C.c_int32),
        ("rtol", C.c_double), ("atol", C.c_double),
        ("h_init", C.c_double), ("h_min", C.c_double), ("h_max", C.c_double),
        ("far_slope", C.c_double),
        ("symplectic_h", C.c_double), ("symplectic_order", C.c_int32),
        ("max_steps", C.c_int32),
        ("r_escape", C.c_double), ("horizon_factor", C.c_double),
        ("n_bins", C.c_int32), ("use_surrogate", C.c_int32),
        ("x_min", C.c_double), ("x_max", C.c_double),
        ("escalate_margin", C.c_double), ("escalate_band", C.c_double), ("surrogate_band", C.c_double),
        ("surrogate_pmin", C.c_double), ("audit_fraction", C.c_double),
        ("audit_seed", C.c_uint64),
        ("block_size", C.c_int32), ("blocks_per_cu", C.c_int32),
        ("refill_min", C.c_int32), ("reserved0", C.c_int32),
        ("vol_j0", C.c_double), ("vol_h_over_r", C.c_double), ("vol_kappa", C.c_double),
        ("vol_r_min", C.c_double), ("vol_r_max", C.c_double), ("vol_tau_max", C.c_double),
        ("escalate_disk_edge", C.c_double), ("escalate_rtol", C.c_double), ("critical_band", C.c_double),
        ("step_radial_exp", C.c_double),
        ("surrogate_err_max", C.c_double), ("surrogate_margin", C.c_double), ("precision_audit", C.c_double),
    ]


class RayOutput(C.Structure):
    _fields_ = [
        ("cls", C.c_void_p), ("flags", C.c_void_p),
        ("theta", C.c_void_p), ("phi", C.c_void_p), ("lambda_", C.c_void_p),
        ("disk_r", C.c_void_p), ("redshift", C.c_void_p), ("null_drift", C.c_void_p),
        ("steps", C.c_void_p), ("intensity", C.c_void_p), ("state", C.c_void_p),
        ("on_device", C.c_int32), ("reserved0", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64), ("steps_accepted", C.c_uint64), ("steps_rejected", C.c_uint64),
        ("steps_landing", C.c_uint64), ("rays_escalated", C.c_uint64),
        ("rays_surrogate", C.c_uint64), ("rays_fallback", C.c_uint64),
        ("rays_audited", C.c_uint64), ("steps_bulk", C.c_uint64), ("steps_escalated", C.c_uint64),
        ("kernel_ms", C.c_double), ("escalate_ms", C.c_double), ("surrogate_ms", C.c_double),
        ("total_ms", C.c_double), ("flops", C.c_double), ("flops_escalated", C.c_double),
        ("null_drift_max", C.c_double), ("null_drift_median", C.c_double),
        ("audit_ms", C.c_double), ("audit_dir_err_max", C.c_double), ("audit_class_mismatch", C.c_uint64),
        ("surrogate_flops", C.c_double),
        ("rays_candidates", C.c_uint64), ("rays_gated", C.c_uint64),
        ("precision_audited", C.c_uint64), ("precision_class_mismatch", C.c_uint64),
        ("precision_dir_err_max", C.c_double), ("precision_dir_err_median", C.c_double),
        ("precision_int_err_max", C.c_double), ("precision_int_err_median", C.c_double),
        ("precision_audit_ms", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Star(C.Structure):
    """SiriusHipStar (the reference's StarEntry field order)."""
    _fields_ = [("dir", C.c_float * 3), ("distance_pc", C.c_float), ("magnitude", C.c_float),
                ("color_bv", C.c_float), ("temperature_K", C.c_float), ("reserved", C.c_float)]


class Starfield(C.Structure):
    """SiriusHipStarfield (generator and sampler parameters)."""
    _fields_ = [("star_count", C.c_uint32), ("seed", C.c_uint32),
                ("min_distance_pc", C.c_float), ("max_distance_pc", C.c_float),
                ("magnitude_limit", C.c_float), ("aperture_mm", C.c_float),
                ("focus_distance_pc", C.c_float), ("brightness_scale", C.c_float),
                ("enable_dof", C.c_int32), ("reserved", C.c_int32)]


BUNDLE_TERMINATED = 1


class RayBundle(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("use_full_riemann", C.c_int32), ("initial_size", C.c_float),
                ("min_bundle_size", C.c_float), ("max_ellipse_ratio", C.c_float), ("reserved0", C.c_int32)]


class Shading(C.Structure):
    """SiriusHipShading (framebuffer shading of launch)."""
    _fields_ = [("samples_per_pixel", C.c_int32), ("temperature_model", C.c_int32), ("disk_temperature", C.c_double),
                ("disk_boost", C.c_double), ("max_steps_level", C.c_double)]


class BundleOutput(C.Structure):
    _fields_ = [("cls", C.c_void_p), ("flags", C.c_void_p), ("jac", C.c_void_p), ("beam", C.c_void_p),
                ("on_device", C.c_int32), ("reserved0", C.c_int32)]


class ImageOpts(C.Structure):
    _fields_ = [("format", C.c_int32), ("exposure", C.c_float), ("zlib_level", C.c_int32), ("threads", C.c_int32)]


class ImageStats(C.Structure):
    _fields_ = [("convert_ms", C.c_double), ("copy_ms", C.c_double), ("pack_ms", C.c_double),
                ("total_ms", C.c_double), ("payload_bytes", C.c_uint64), ("file_bytes", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def metric_ks(M=1.0, a=0.0, Q=0.0, Lambda=0.0):
    return Metric(family=METRIC_KERR_SCHILD, M=M, a=a, Q=Q, Lambda=Lambda)


def metric_mt(b0=1.0):
    return Metric(family=METRIC_MORRIS_THORNE, M=0.0, throat_b0=b0)


def metric_warp(v=0.5, R=1.0, sigma=8.0):
    return Metric(family=METRIC_ALCUBIERRE, M=0.0, warp_v=v, warp_R=R, warp_sigma=sigma)


def camera(r_obs, incl_deg, fov_deg, width, height, disk=(0.0, 30.0), metric_index=0,
           phi_deg=0.0, yaw=0.0, pitch=0.0, roll=0.0):
    return Camera(r_obs=r_obs, theta_obs=np.deg2rad(incl_deg), phi_obs=np.deg2rad(phi_deg),
                  fov_deg=fov_deg, yaw=yaw, pitch=pitch, roll=roll,
                  disk_r_in=disk[0], disk_r_out=disk[1], width=width, height=height,
                  metric_index=metric_index)


class HostRays:
    """Host numpy buffers for n rays and a RayOutput pointing at them."""

    def __init__(self, n, n_bins, with_state=False):
        self.n, self.n_bins = n, n_bins
        self.cls = np.zeros(n, np.uint8)
        self.flags = np.zeros(n, np.uint8)
        self.theta = np.zeros(n, np.float32)
        self.phi = np.zeros(n, np.float32)
        self.lam = np.zeros(n, np.float32)
        self.disk_r = np.zeros(n, np.float32)
        self.redshift = np.zeros(n, np.float32)
        self.null_drift = np.zeros(n, np.float32)
        self.steps = np.zeros(n, np.uint32)
        self.intensity = np.zeros((n_bins, n), np.float32)
        self.state = np.zeros((n, 8), np.float64) if with_state else None

    def output(self):
        p = lambda a: a.ctypes.data if a is not None else None
        return RayOutput(cls=p(self.cls), flags=p(self.flags), theta=p(self.theta),
                         phi=p(self.phi), lambda_=p(self.lam), disk_r=p(self.disk_r),
                         redshift=p(self.redshift), null_drift=p(self.null_drift),
                         steps=p(self.steps), intensity=p(self.intensity),
                         state=p(self.state), on_device=0)

    def take(self, idx):
        """A copy restricted to the given ray indices."""
        out = HostRays(len(idx), self.n_bins, self.state is not None)
        for k in ("cls", "flags", "theta", "phi", "lam", "disk_r", "redshift", "null_drift",
                  "steps"):
            setattr(out, k, getattr(self, k)[idx].copy())
        out.intensity = self.intensity[:, idx].copy()
        if self.state is not None:
            out.state = self.state[idx].copy()
        return out

    def direction(self):
        """Unit escape directions (n, 3) from (theta, phi)."""
        th, ph = self.theta.astype(np.float64), self.phi.astype(np.float64)
        return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
