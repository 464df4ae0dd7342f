"""Python mirror of the HIP backend: the IAccelerator method set
(src/Sirius.Render/Acceleration/Backend/ACIB001A.h:148-189) plus the batched physics entry
point, all calls going through the C-ABI of libsirius_hip.so (include/sirius_hip.h).

There is no CPU fallback: if the library is missing or the device is unusable, every call
raises SiriusHipError.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsirius_hip.so")
_lib = None


class SiriusHipError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"sirius_hip error {rc}: {msg}")
        self.rc = rc


def lib():
    """Load libsirius_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SiriusHipError(abi.E_STATE, f"{LIB_PATH} not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    H, i, d = C.c_void_p, C.c_int, C.c_double
    L.sirius_hip_create.restype = H
    L.sirius_hip_destroy.argtypes = [H]
    L.sirius_hip_initialize.argtypes = [H, i, i]
    L.sirius_hip_cleanup.argtypes = [H]
    L.sirius_hip_is_initialized.argtypes = [H]
    L.sirius_hip_last_error.restype = C.c_char_p
    L.sirius_hip_last_error.argtypes = [H]
    L.sirius_hip_device_info.argtypes = [H, C.c_char_p, i, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                         C.POINTER(i), C.POINTER(i)]
    L.sirius_hip_launch.argtypes = [H, C.POINTER(abi.Camera), C.POINTER(abi.Metric), C.POINTER(abi.TraceOpts)]
    L.sirius_hip_resize.argtypes = [H, i, i]
    L.sirius_hip_get_frame_buffer.restype = C.POINTER(C.c_float)
    L.sirius_hip_get_frame_buffer.argtypes = [H]
    L.sirius_hip_get_frame_buffer_size.restype = C.c_size_t
    L.sirius_hip_get_frame_buffer_size.argtypes = [H]
    L.sirius_hip_upload_background.argtypes = [H, C.c_void_p, i, i]
    L.sirius_hip_synchronize.argtypes = [H]
    L.sirius_hip_fetch_results.argtypes = [H, C.POINTER(abi.RayOutput)]
    L.sirius_hip_default_opts.argtypes = [C.POINTER(abi.TraceOpts), i]
    L.sirius_hip_default_opts.restype = None
    L.sirius_hip_trace_batch.argtypes = [H, C.POINTER(abi.Camera), i, C.POINTER(abi.Metric), i,
                                         C.POINTER(abi.TraceOpts), C.POINTER(abi.RayOutput),
                                         C.POINTER(abi.Stats)]
    L.sirius_hip_trace_batch_noalloc.argtypes = L.sirius_hip_trace_batch.argtypes
    L.sirius_hip_critical_distance.argtypes = [H, C.POINTER(abi.Camera), i, C.POINTER(abi.Metric), i,
                                               C.c_void_p, i]
    L.sirius_hip_autotune.argtypes = [H, C.POINTER(abi.Camera), i, C.POINTER(abi.Metric), i,
                                      C.POINTER(abi.TraceOpts), C.POINTER(d)]
    L.sirius_hip_mlp_param_count.restype = C.c_size_t
    L.sirius_hip_mlp_init.argtypes = [H, C.c_uint64]
    L.sirius_hip_mlp_get_params.argtypes = [H, C.c_void_p, C.c_size_t]
    L.sirius_hip_mlp_set_params.argtypes = [H, C.c_void_p, C.c_size_t]
    L.sirius_hip_mlp_forward.argtypes = [H, C.c_void_p, C.c_void_p, i]
    L.sirius_hip_mlp_loss_grad.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p, i, C.c_void_p,
                                           C.POINTER(C.c_float)]
    L.sirius_hip_mlp_train_step.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p, i, C.c_float,
                                            C.POINTER(C.c_float)]
    L.sirius_hip_surrogate_features.argtypes = [H, C.POINTER(abi.Camera), i, C.POINTER(abi.Metric), i,
                                                C.c_void_p, C.c_void_p]
    L.sirius_hip_surrogate_train.argtypes = [H, C.POINTER(abi.Camera), i, C.POINTER(abi.Metric), i,
                                             C.POINTER(abi.TraceOpts), i, i, C.c_float, C.c_uint64,
                                             C.POINTER(C.c_float)]
    L.sirius_hip_starfield_default.argtypes = [C.POINTER(abi.Starfield)]
    L.sirius_hip_starfield_default.restype = None
    L.sirius_hip_starfield_generate.argtypes = [C.POINTER(abi.Starfield), C.c_void_p, i, C.POINTER(C.c_int)]
    L.sirius_hip_set_starfield.argtypes = [H, i, C.POINTER(abi.Starfield), C.c_void_p, i]
    L.sirius_hip_sample_starfield.argtypes = [H, i, C.c_void_p, i, C.c_void_p]
    L.sirius_hip_last_result_size.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.sirius_hip_default_ray_bundle.argtypes = [C.POINTER(abi.RayBundle)]
    L.sirius_hip_default_ray_bundle.restype = None
    L.sirius_hip_set_ray_bundle.argtypes = [H, C.POINTER(abi.RayBundle)]
    L.sirius_hip_default_shading.argtypes = [C.POINTER(abi.Shading)]
    L.sirius_hip_default_shading.restype = None
    L.sirius_hip_set_shading.argtypes = [H, C.POINTER(abi.Shading)]
    L.sirius_hip_trace_bundle.argtypes = [H, C.POINTER(abi.Camera), i, C.POINTER(abi.Metric), i,
                                          C.POINTER(abi.TraceOpts), C.POINTER(abi.BundleOutput)]
    L.sirius_hip_default_image_opts.argtypes = [C.POINTER(abi.ImageOpts), i]
    L.sirius_hip_default_image_opts.restype = None
    L.sirius_hip_encode_image.argtypes = [H, C.c_void_p, i, i, i, C.POINTER(abi.ImageOpts),
                                          C.POINTER(C.POINTER(C.c_ubyte)), C.POINTER(C.c_size_t),
                                          C.POINTER(abi.ImageStats)]
    L.sirius_hip_write_image.argtypes = [H, C.c_char_p, C.c_void_p, i, i, i, C.POINTER(abi.ImageOpts),
                                         C.POINTER(abi.ImageStats)]
    L.sirius_hip_write_rays.argtypes = [H, C.c_char_p, C.c_uint32, C.POINTER(abi.Camera), i,
                                        C.POINTER(abi.Metric), i, C.POINTER(abi.TraceOpts)]
    _lib = L
    return L


def ray_bundle_default():
    b = abi.RayBundle()
    lib().sirius_hip_default_ray_bundle(C.byref(b))
    return b


def shading_default():
    s = abi.Shading()
    lib().sirius_hip_default_shading(C.byref(s))
    return s


def starfield_default():
    """The reference's StarfieldConfig defaults (host only)."""
    cfg = abi.Starfield()
    lib().sirius_hip_starfield_default(C.byref(cfg))
    return cfg


def starfield_generate(cfg):
    """The catalogue of cfg as an array of abi.Star (host only: no device needed)."""
    L = lib()
    n = C.c_int(0)
    rc = L.sirius_hip_starfield_generate(C.byref(cfg), None, 0, C.byref(n))
    if rc:
        raise SiriusHipError(rc, "starfield_generate")
    stars = (abi.Star * max(n.value, 1))()
    rc = L.sirius_hip_starfield_generate(C.byref(cfg), stars, n.value, C.byref(n))
    if rc:
        raise SiriusHipError(rc, "starfield_generate")
    return stars, n.value


def default_opts(precision=abi.PRECISION_FP64):
    o = abi.TraceOpts()
    lib().sirius_hip_default_opts(C.byref(o), precision)
    return o


class HipAccelerator:
    """IAccelerator method set over the C-ABI (getType/getCapabilities/initialise/...)."""

    BACKEND_HIP = 3  # BackendType::HIP (ACIB001A.h:28)

    def __init__(self):
        self._L = lib()
        self._h = self._L.sirius_hip_create()
        if not self._h:
            raise SiriusHipError(abi.E_DEVICE, "sirius_hip_create failed")

    def close(self):
        if self._h:
            self._L.sirius_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != abi.OK:
            raise SiriusHipError(rc, self._L.sirius_hip_last_error(self._h).decode())

    # ---- IAccelerator ----
    def getType(self):
        return self.BACKEND_HIP

    def getCapabilities(self):
        name = C.create_string_buffer(256)
        tot, fr = C.c_size_t(), C.c_size_t()
        cu, mt = C.c_int(), C.c_int()
        self._check(self._L.sirius_hip_device_info(self._h, name, 256, C.byref(tot), C.byref(fr),
                                                   C.byref(cu), C.byref(mt)))
        return dict(deviceName=name.value.decode(), totalMemory=tot.value, freeMemory=fr.value,
                    supportsFP64=True, multiprocessorCount=cu.value, maxThreadsPerBlock=mt.value)

    def initialise(self, width, height):
        self._check(self._L.sirius_hip_initialize(self._h, width, height))
        return True

    def isInitialised(self):
        return bool(self._L.sirius_hip_is_initialized(self._h))

    def launch(self, cam, metric, opts):
        self._check(self._L.sirius_hip_launch(self._h, C.byref(cam), C.byref(metric), C.byref(opts)))

    def getFrameBuffer(self):
        p = self._L.sirius_hip_get_frame_buffer(self._h)
        n = self._L.sirius_hip_get_frame_buffer_size(self._h) // 4
        if not p:
            raise SiriusHipError(abi.E_STATE, "no framebuffer (initialise first)")
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def getFrameBufferSize(self):
        return self._L.sirius_hip_get_frame_buffer_size(self._h)

    def uploadBackground(self, rgba):
        rgba = np.ascontiguousarray(rgba, np.uint8)
        h, w = rgba.shape[:2]
        self._check(self._L.sirius_hip_upload_background(self._h, rgba.ctypes.data, w, h))
        return True

    def set_starfield(self, side, cfg, stars, n=None):
        """Installs a star catalogue (abi.Star array) as celestial sphere `side` (0 sky / near
        side, 1 a wormhole's far side); n = 0 removes it."""
        n = len(stars) if n is None else n
        self._check(self._L.sirius_hip_set_starfield(self._h, side, C.byref(cfg) if cfg is not None else None,
                                                     stars if n else None, n))

    def sample_starfield(self, side, dirs):
        """Star light (n, 3) of sphere `side` at unit directions dirs (n, 3), on the device."""
        dirs = np.ascontiguousarray(dirs, np.float32)
        rgb = np.zeros((len(dirs), 3), np.float32)
        self._check(self._L.sirius_hip_sample_starfield(self._h, side, dirs.ctypes.data, len(dirs), rgb.ctypes.data))
        return rgb

    # ---- ray bundles ----
    def set_ray_bundle(self, enabled=True, initial_size=None, min_bundle_size=None, max_ellipse_ratio=None):
        """Ray-bundle parameters of later launches (LaunchParams::rayBundle); None keeps the default."""
        b = ray_bundle_default()
        b.enabled = int(enabled)
        for k, v in (("initial_size", initial_size), ("min_bundle_size", min_bundle_size),
                     ("max_ellipse_ratio", max_ellipse_ratio)):
            if v is not None:
                setattr(b, k, v)
        self._check(self._L.sirius_hip_set_ray_bundle(self._h, C.byref(b)))
        return b

    def set_shading(self, samples_per_pixel=None, temperature_model=None, disk_temperature=None, disk_boost=None,
                    max_steps_level=None):
        """Framebuffer shading of later launches (SiriusHipShading); None keeps the default."""
        sh = shading_default()
        for k, v in (("samples_per_pixel", samples_per_pixel), ("temperature_model", temperature_model),
                     ("disk_temperature", disk_temperature), ("disk_boost", disk_boost),
                     ("max_steps_level", max_steps_level)):
            if v is not None:
                setattr(sh, k, v)
        self._check(self._L.sirius_hip_set_shading(self._h, C.byref(sh)))
        return sh

    def trace_bundle(self, cams, metrics, opts, resident=False):
        """Every pixel's ray bundle: dict of cls, flags, jac (4, n) and beam (3, n); with resident
        the results stay in library device memory (nothing is copied, None is returned)."""
        cams, metrics = list(cams), list(metrics)
        ca = (abi.Camera * len(cams))(*cams)
        ma = (abi.Metric * len(metrics))(*metrics)
        if resident:
            bo = abi.BundleOutput(on_device=1)
            self._check(self._L.sirius_hip_trace_bundle(self._h, ca, len(cams), ma, len(metrics), C.byref(opts),
                                                        C.byref(bo)))
            return None
        n = sum(c.width * c.height for c in cams)
        out = {"cls": np.zeros(n, np.uint8), "flags": np.zeros(n, np.uint8),
               "jac": np.zeros((4, n), np.float32), "beam": np.zeros((3, n), np.float32)}
        bo = abi.BundleOutput(cls=out["cls"].ctypes.data, flags=out["flags"].ctypes.data,
                              jac=out["jac"].ctypes.data, beam=out["beam"].ctypes.data, on_device=0)
        self._check(self._L.sirius_hip_trace_bundle(self._h, ca, len(cams), ma, len(metrics), C.byref(opts),
                                                    C.byref(bo)))
        return out

    # ---- image writers and the ray result dump ----
    def _image_args(self, fmt, rgba, exposure, zlib_level, threads):
        o = abi.ImageOpts()
        self._L.sirius_hip_default_image_opts(C.byref(o), fmt)
        o.exposure, o.zlib_level, o.threads = exposure, zlib_level, threads
        if rgba is None:  # the framebuffer of the last launch
            return o, None, 0, 0, 0, None
        if hasattr(rgba, "data_ptr"):  # device memory (float32, contiguous, (h, w, 4))
            h, w = int(rgba.shape[0]), int(rgba.shape[1])
            return o, rgba.data_ptr(), w, h, 1, rgba
        a = np.ascontiguousarray(rgba, np.float32)
        h, w = a.shape[:2]
        return o, a.ctypes.data, w, h, 0, a

    def encode_image(self, fmt, rgba=None, exposure=1.0, zlib_level=-1, threads=0):
        """The file bytes of a frame ((h, w, 4) float RGBA on the host, or device memory with
        data_ptr() and shape; None:
        the framebuffer) in format `fmt` (abi.IMAGE_*), and the call's ImageStats."""
        o, ptr, w, h, dev, keep = self._image_args(fmt, rgba, exposure, zlib_level, threads)
        data, size, st = C.POINTER(C.c_ubyte)(), C.c_size_t(), abi.ImageStats()
        self._check(self._L.sirius_hip_encode_image(self._h, ptr, w, h, dev, C.byref(o), C.byref(data),
                                                    C.byref(size), C.byref(st)))
        return C.string_at(data, size.value), st

    def write_image(self, path, fmt, rgba=None, exposure=1.0, zlib_level=-1, threads=0):
        o, ptr, w, h, dev, keep = self._image_args(fmt, rgba, exposure, zlib_level, threads)
        st = abi.ImageStats()
        self._check(self._L.sirius_hip_write_image(self._h, os.fsencode(path), ptr, w, h, dev, C.byref(o),
                                                   C.byref(st)))
        return st

    def write_rays(self, path, cams=(), metrics=(), opts=None, fields=abi.RAYS_ALL):
        """Dumps the last trace's results with the cameras, metrics and options that made them
        (read back with sirius_amd.raydump.read_rays)."""
        ca = (abi.Camera * max(len(cams), 1))(*cams)
        ma = (abi.Metric * max(len(metrics), 1))(*metrics)
        self._check(self._L.sirius_hip_write_rays(self._h, os.fsencode(path), fields, ca, len(cams), ma,
                                                  len(metrics), C.byref(opts) if opts is not None else None))

    def synchronise(self):
        self._check(self._L.sirius_hip_synchronize(self._h))

    def cleanup(self):
        self._L.sirius_hip_cleanup(self._h)

    def getLastError(self):
        return self._L.sirius_hip_last_error(self._h).decode()

    # ---- physics ----
    def trace_batch(self, cams, metrics, opts, with_state=False, device_out=None):
        """Trace every pixel of the cameras. Returns (abi.HostRays, stats dict), or with
        device_out (an abi.RayOutput of device pointers, on_device=1) (None, stats)."""
        cams, metrics = list(cams), list(metrics)
        ca = (abi.Camera * len(cams))(*cams)
        ma = (abi.Metric * len(metrics))(*metrics)
        st = abi.Stats()
        if device_out is not None:
            self._check(self._L.sirius_hip_trace_batch(self._h, ca, len(cams), ma, len(metrics),
                                                       C.byref(opts), C.byref(device_out), C.byref(st)))
            return None, st.as_dict()
        n = sum(c.width * c.height for c in cams)
        out = abi.HostRays(n, opts.n_bins, with_state)
        ro = out.output()
        self._check(self._L.sirius_hip_trace_batch(self._h, ca, len(cams), ma, len(metrics),
                                                   C.byref(opts), C.byref(ro), C.byref(st)))
        return out, st.as_dict()

    def trace_resident(self, cams, metrics, opts, stats=True):
        """Trace with every result left in library device memory (nothing copied back).
        Returns the stats dict (None with stats=False: no statistics pass and no read-back);
        fetch_results() copies the results afterwards."""
        cams, metrics = list(cams), list(metrics)
        ca = (abi.Camera * len(cams))(*cams)
        ma = (abi.Metric * len(metrics))(*metrics)
        st = abi.Stats()
        ro = abi.RayOutput(on_device=1)
        self._check(self._L.sirius_hip_trace_batch(self._h, ca, len(cams), ma, len(metrics),
                                                   C.byref(opts), C.byref(ro),
                                                   C.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def fetch_results(self, with_state=False):
        """Host copy of the results the library holds: those of the last trace_resident,
        trace_batch or launch, sized by the library's own count."""
        n, nb = C.c_int64(), C.c_int32()
        self._check(self._L.sirius_hip_last_result_size(self._h, C.byref(n), C.byref(nb)))
        out = abi.HostRays(n.value, nb.value, with_state)
        ro = out.output()
        self._check(self._L.sirius_hip_fetch_results(self._h, C.byref(ro)))
        return out

    def critical_distance(self, cams, metrics):
        cams, metrics = list(cams), list(metrics)
        ca = (abi.Camera * len(cams))(*cams)
        ma = (abi.Metric * len(metrics))(*metrics)
        n = sum(c.width * c.height for c in cams)
        out = np.zeros(n, np.float32)
        self._check(self._L.sirius_hip_critical_distance(self._h, ca, len(cams), ma, len(metrics),
                                                         out.ctypes.data, 0))
        return out

    # ---- surrogate network (x, y, cls, dirs, grad: device buffers, given as integer pointers) ----
    def mlp_init(self, seed):
        self._check(self._L.sirius_hip_mlp_init(self._h, seed))

    def mlp_param_count(self):
        return int(self._L.sirius_hip_mlp_param_count())

    def mlp_get_params(self):
        p = np.zeros(self.mlp_param_count(), np.float32)
        self._check(self._L.sirius_hip_mlp_get_params(self._h, p.ctypes.data, p.size))
        return p

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self._L.sirius_hip_mlp_set_params(self._h, p.ctypes.data, p.size))

    def mlp_forward(self, x_dev, y_dev, n):
        self._check(self._L.sirius_hip_mlp_forward(self._h, x_dev, y_dev, n))

    def mlp_loss_grad(self, x_dev, cls_dev, dirs_dev, n, grad_dev):
        loss = C.c_float()
        self._check(self._L.sirius_hip_mlp_loss_grad(self._h, x_dev, cls_dev, dirs_dev, n, grad_dev, C.byref(loss)))
        return loss.value

    def mlp_train_step(self, x_dev, cls_dev, dirs_dev, n, lr):
        loss = C.c_float()
        self._check(self._L.sirius_hip_mlp_train_step(self._h, x_dev, cls_dev, dirs_dev, n, C.c_float(lr),
                                                      C.byref(loss)))
        return loss.value

    def surrogate_features(self, cams, metrics, x_dev, n0_dev):
        """Network inputs x[n*5] and initial directions n0[n*3] of every ray (device buffers)."""
        cams, metrics = list(cams), list(metrics)
        ca = (abi.Camera * len(cams))(*cams)
        ma = (abi.Metric * len(metrics))(*metrics)
        self._check(self._L.sirius_hip_surrogate_features(self._h, ca, len(cams), ma, len(metrics), x_dev, n0_dev))

    def surrogate_train(self, cams, metrics, label_opts, epochs, batch, lr, seed):
        """Train the network on the cameras' rays (labels from the integrator); returns the last
        epoch's mean batch loss."""
        cams, metrics = list(cams), list(metrics)
        ca = (abi.Camera * len(cams))(*cams)
        ma = (abi.Metric * len(metrics))(*metrics)
        loss = C.c_float()
        self._check(self._L.sirius_hip_surrogate_train(self._h, ca, len(cams), ma, len(metrics), C.byref(label_opts),
                                                       epochs, batch, C.c_float(lr), seed, C.byref(loss)))
        return loss.value

    def autotune(self, cams, metrics, opts):
        cams, metrics = list(cams), list(metrics)
        ca = (abi.Camera * len(cams))(*cams)
        ma = (abi.Metric * len(metrics))(*metrics)
        best = C.c_double()
        self._check(self._L.sirius_hip_autotune(self._h, ca, len(cams), ma, len(metrics),
                                                C.byref(opts), C.byref(best)))
        return best.value
