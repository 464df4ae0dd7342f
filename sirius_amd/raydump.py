"""Reader of the ray result dump written by sirius_hip_write_rays (layout in include/sirius_hip.h):
the per-ray arrays of a trace together with the camera, metric and option records that made
them, so that an audit can re-trace any sampled ray.
"""
import ctypes as C

import numpy as np

from . import abi

MAGIC = b"SIRRAYS1"
_DTYPES = {"cls": np.uint8, "flags": np.uint8, "theta": np.float32, "phi": np.float32, "lambda": np.float32,
           "disk_r": np.float32, "redshift": np.float32, "null_drift": np.float32, "steps": np.uint32,
           "intensity": np.float32, "state": np.float64}


def _records(buf, pos, n, cls, size):
    if n and size != C.sizeof(cls):
        raise ValueError(f"{cls.__name__} record is {size} bytes, this reader's is {C.sizeof(cls)}")
    out = [cls.from_buffer_copy(buf, pos + i * size) for i in range(n)]
    return out, pos + n * size


def read_rays(path):
    """Returns a dict: 'n_rays', 'n_bins', 'cameras' (abi.Camera list), 'metrics' (abi.Metric
    list), 'opts' (abi.TraceOpts or None) and one numpy array per field present (intensity shaped
    (n_bins, n_rays), state (n_rays, 8))."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] != MAGIC:
        raise ValueError(f"{path}: not a ray result dump")
    head_bytes, n_bins = np.frombuffer(buf, np.uint32, 2, 8)
    n_rays = int(np.frombuffer(buf, np.uint64, 1, 16)[0])
    fields, = np.frombuffer(buf, np.uint32, 1, 24)
    n_cams, = np.frombuffer(buf, np.int32, 1, 28)
    cam_size, = np.frombuffer(buf, np.uint32, 1, 32)
    n_met, = np.frombuffer(buf, np.int32, 1, 36)
    met_size, opt_size = np.frombuffer(buf, np.uint32, 2, 40)
    out = {"n_rays": n_rays, "n_bins": int(n_bins)}
    pos = int(head_bytes)
    out["cameras"], pos = _records(buf, pos, int(n_cams), abi.Camera, int(cam_size))
    out["metrics"], pos = _records(buf, pos, int(n_met), abi.Metric, int(met_size))
    opts, pos = _records(buf, pos, 1 if opt_size else 0, abi.TraceOpts, int(opt_size))
    out["opts"] = opts[0] if opts else None
    for bit, name in enumerate(abi.RAYS_FIELDS):
        if not (int(fields) >> bit) & 1:
            continue
        pos += (8 - pos % 8) % 8
        count = n_rays * (int(n_bins) if name == "intensity" else 8 if name == "state" else 1)
        a = np.frombuffer(buf, _DTYPES[name], count, pos)
        pos += a.nbytes
        if name == "intensity":
            a = a.reshape(int(n_bins), n_rays)
        elif name == "state":
            a = a.reshape(n_rays, 8)
        out[name] = a
    if pos != len(buf):
        raise ValueError(f"{path}: {len(buf) - pos} bytes left after the last field")
    return out
