"""Step control of the DP45 tracer, on the CPU.

1. The Dormand-Prince coefficients the kernels compile (read from sirius_amd/csrc/geodesic.hpp,
   not retyped here) satisfy what the mathematics fixes: the 5th-order solution has local error
   O(h^6), the embedded estimate O(h^5), and the dense output, with which the FP32 path locates
   and lands events, is a continuous extension of order 4 (local error O(h^5) inside the step,
   where the cubic Hermite interpolant alone is O(h^4)), matching the step's end values and end
   slopes. A wrong or mistyped coefficient drops one of these orders.
2. The inward step proposal (step_radial_exp) of the oracle: fewer rejected attempts on the
   1080p Kerr frame's sample, same answers within the FP32 path's tolerance against the frame's
   FP64 golden columns.
"""
import os
import re
from fractions import Fraction

import numpy as np

from oracle import oracle_py as O
from sirius_amd import abi

HDR = os.path.join(os.path.dirname(__file__), "..", "sirius_amd", "csrc", "geodesic.hpp")


def _coefficients():
    src = open(HDR).read()
    body = src[src.index("void dp45_step("):src.index("SIRIUS_HD __forceinline__ T dense_eval(")]
    vals = {}
    for name, expr in re.findall(r"\b([abeq]\d{1,2})\s*=\s*(-?[\d.]+\s*/\s*[\d.]+|-?[\d.]+)", body):
        num, _, den = expr.partition("/")
        vals[name] = float(Fraction(num.strip()) / (Fraction(den.strip()) if den else 1))
    return vals


def _tableau():
    v = _coefficients()
    A = [[v.get(f"a{i}{j}", 0.0) for j in range(1, i)] for i in range(1, 7)]
    b = [v.get(f"b{i}", 0.0) for i in range(1, 7)]
    e = [v.get(f"e{i}", 0.0) for i in range(1, 8)]
    q = [v.get(f"q{i}", 0.0) for i in range(1, 8)]
    return A, b, e, q


def _kepler(y):
    x, u = y[:2], y[2:]
    return np.concatenate([u, -x / np.linalg.norm(x) ** 3])


def _step(y, h):
    A, b, e, q = _tableau()
    K = [_kepler(y)]
    for i in range(1, 6):
        K.append(_kepler(y + h * sum(A[i][j] * K[j] for j in range(i))))
    y1 = y + h * sum(b[j] * K[j] for j in range(6))
    K.append(_kepler(y1))
    err = h * sum(e[j] * K[j] for j in range(7))
    r5 = h * sum(q[j] * K[j] for j in range(7))
    return y1, K, err, r5


def _dense(y0, y1, k0, k1, r5, h, th):
    d = y1 - y0
    b = h * k0 - d
    c = d - h * k1 - b
    return y0 + th * (d + (1 - th) * (b + th * (c + (1 - th) * r5)))


def _exact(y0, t, n=4000):
    y, h = y0.copy(), t / n
    for _ in range(n):
        y = _step(y, h)[0]
    return y


def _slope(hs, errs):
    return np.polyfit(np.log(hs), np.log(errs), 1)[0]


def test_header_parses_all_coefficients():
    v = _coefficients()
    for name in ("a21", "a65", "b1", "b6", "e1", "e7", "q1", "q3", "q7"):
        assert name in v, name
    assert abs(sum(v[f"b{i}"] for i in (1, 3, 4, 5, 6)) - 1.0) < 1e-15  # consistency
    assert abs(sum(v.get(f"e{i}", 0.0) for i in range(1, 8))) < 1e-15   # estimate of a difference


def test_dp45_orders_and_dense_output():
    y0 = np.array([1.0, 0.0, 0.0, 1.1])  # eccentric bound orbit
    hs = np.array([0.16, 0.08, 0.04, 0.02])
    sol, est, mid, herm, ends = [], [], [], [], []
    for h in hs:
        y1, K, err, r5 = _step(y0, h)
        sol.append(np.linalg.norm(y1 - _exact(y0, h)))
        est.append(np.linalg.norm(err))
        ym = _exact(y0, 0.5 * h)
        mid.append(np.linalg.norm(_dense(y0, y1, K[0], K[6], r5, h, 0.5) - ym))
        herm.append(np.linalg.norm(_dense(y0, y1, K[0], K[6], 0.0 * r5, h, 0.5) - ym))
        # end values and end slopes (central differences of the interpolant in th)
        eps = 1e-6
        s0 = (_dense(y0, y1, K[0], K[6], r5, h, eps) - _dense(y0, y1, K[0], K[6], r5, h, -eps)) / (2 * eps * h)
        s1 = (_dense(y0, y1, K[0], K[6], r5, h, 1 + eps) - _dense(y0, y1, K[0], K[6], r5, h, 1 - eps)) / (2 * eps * h)
        ends.append(max(np.abs(_dense(y0, y1, K[0], K[6], r5, h, 0.0) - y0).max(),
                        np.abs(_dense(y0, y1, K[0], K[6], r5, h, 1.0) - y1).max(),
                        np.abs(s0 - K[0]).max(), np.abs(s1 - K[6]).max()))
    assert _slope(hs, sol) > 5.6, sol        # local error O(h^6)
    assert 4.6 < _slope(hs, est) < 5.4, est  # embedded estimate O(h^5)
    assert _slope(hs, mid) > 4.7, mid        # dense output of order 4
    assert _slope(hs, herm) < 4.3, herm      # the r5 term is what lifts the Hermite interpolant
    assert mid[-1] < 0.2 * herm[-1]
    assert max(ends) < 1e-6


def test_inward_step_proposal_cuts_rejections_not_accuracy():
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_trace_c2.npz"))
    cam = abi.camera(float(d["r_obs"]), float(d["incl_deg"]), float(d["fov"]), int(d["width"]), int(d["height"]),
                     disk=tuple(d["disk"]))
    m = abi.metric_ks(1.0, float(d["a"]))
    pix = d["pix"][:1500]
    keep = np.abs(d["crit"][:1500]) > 0.5
    res = {}
    for ex in (0.0, 2.0):
        o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
        o.n_bins = int(d["n_bins"])
        o.step_radial_exp = ex
        r, st = O.trace([cam], [m], o, ray_ids=pix)
        res[ex] = (r, st)
        assert np.array_equal(r.cls[keep], d["fp64_cls"][:1500][keep])
        esc = keep & (r.cls == 0)
        th, ph = d["fp64_theta"][:1500][esc].astype(float), d["fp64_phi"][:1500][esc].astype(float)
        v = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
        u = r.direction()[esc]
        u = u / np.linalg.norm(u, axis=1, keepdims=True)
        assert np.arccos(np.clip((u * v).sum(1), -1, 1)).max() < 1e-5
        dk = keep & (r.cls == 2)
        I0, I1 = d["fp64_I"][:, :1500][:, dk], r.intensity[:, dk]
        rel = np.abs(I1 - I0).max(0) / np.maximum(I0.max(0), 1e-2 * np.median(I0.max(0)))
        assert rel.max() < 1e-3, (ex, rel.max())
    rej0 = res[0.0][1]["steps_rejected"] / len(pix)
    rej2 = res[2.0][1]["steps_rejected"] / len(pix)
    att0 = (res[0.0][1]["steps_accepted"] + res[0.0][1]["steps_rejected"]) / len(pix)
    att2 = (res[2.0][1]["steps_accepted"] + res[2.0][1]["steps_rejected"]) / len(pix)
    assert rej2 < 0.6 * rej0 and att2 < 0.92 * att0, (rej0, rej2, att0, att2)
