"""Celestial spheres on the device, through the C-ABI.

1. sirius_hip_sample_starfield against the reference's own sampled light (golden fixture,
   StarfieldGenerator::sampleStarfield at directions on, near and away from stars, with and
   without depth-of-field blur). The device repeats the reference's view-dependent float steps
   unfused and in its order, so the set of lit directions is identical and the light agrees to
   the device's acosf / expf rounding (bar 1e-5 relative).
2. The two-sided wormhole scene shaded through initialise / launch with a sphere on each side:
   the star light added to each escaped pixel is the oracle's sample of its side's catalogue at
   the pixel's escape direction (formed as the device forms it), pixel for pixel, and a
   far-side pixel gets nothing from the near sphere.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from sirius_amd import abi
from sirius_amd.accelerator import starfield_default, starfield_generate
from workloads import WORMHOLE_SPHERE_SEEDS, sphere_catalogue, wormhole_scene

pytestmark = pytest.mark.gpu

CFG = dict(aperture_mm=50.0, focus_distance_pc=100.0, brightness_scale=1.0, enable_dof=1)


def _cfg(seed, **kw):
    c = starfield_default()
    c.star_count, c.min_distance_pc, c.max_distance_pc, c.magnitude_limit, c.seed = 4096, 1.0, 1000.0, 20.0, seed
    for k, v in {**CFG, **kw}.items():
        setattr(c, k, v)
    return c


def _stars(rows):
    arr = (abi.Star * len(rows))()
    np.frombuffer(arr, np.float32).reshape(-1, 8)[:] = rows
    return arr


@pytest.mark.parametrize("key,cat,kw", [("rgb_dof", "near", {}),
                                        ("rgb_nodof", "near", dict(aperture_mm=0.0, brightness_scale=2.0, enable_dof=0)),
                                        ("rgb_far_dof", "far", {})])
def test_device_sampler_matches_reference(acc, golden, key, cat, kw):
    d = golden("ref_starfield.npz")
    print(f"[case] starfield sample {key}", flush=True)
    acc.set_starfield(0, _cfg(42, **kw), _stars(d[f"cat_{cat}"]))
    got = acc.sample_starfield(0, d["dirs"])
    ref = d[key]
    lit = ref.max(1) > 0
    assert np.array_equal(got.max(1) > 0, lit)
    rel = np.abs(got - ref).max(1)[lit] / ref.max(1)[lit]
    assert rel.max() < 1e-5, rel.max()
    acc.set_starfield(0, None, None, 0)
    assert not acc.sample_starfield(0, d["dirs"][:8]).any()


@pytest.mark.parametrize("kind", ["generator", "exp1"])
def test_wormhole_spheres_shade_escaped_rays(acc, kind):
    """kind "generator": the reference generator's catalogues (seeds 42 and 43); "exp1": the
    scene's own spheres, 4096 stars each with magnitudes ~ Exp(1) (seeds 0x51121008 and
    0x51121009, tests/workloads.py)."""
    cam, m = wormhole_scene(96)
    w = h = 96
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.emission = abi.EMISSION_NONE
    acc.initialise(w, h)
    # a black sky (the session's handle may hold an earlier test's background): escaped pixels
    # then start at exactly 0, and star light of ~1e-7 is not lost against a bright background
    acc.uploadBackground(np.zeros((1, 1, 4), np.uint8))
    if kind == "generator":
        near, n_near = starfield_generate(_cfg(42))
        far, n_far = starfield_generate(_cfg(43))
    else:
        rows = [sphere_catalogue(seed) for seed in WORMHOLE_SPHERE_SEEDS]
        near, n_near = _stars(rows[0]), len(rows[0])
        far, n_far = _stars(rows[1]), len(rows[1])
    print(f"[case] wormhole launch with two spheres ({kind})", flush=True)
    acc.launch(cam, m, o)
    base = acc.getFrameBuffer().reshape(-1, 4)[:, :3].copy()
    acc.set_starfield(0, _cfg(42), near, n_near)
    acc.set_starfield(1, _cfg(43), far, n_far)
    acc.launch(cam, m, o)
    lit = acc.getFrameBuffer().reshape(-1, 4)[:, :3] - base
    rays, _ = acc.trace_batch([cam], [m], o)
    esc = rays.cls == abi.RAY_ESCAPED
    far_side = (rays.flags & abi.FLAG_FAR_SIDE) != 0
    assert (esc & far_side).sum() > 100 and (esc & ~far_side).sum() > 100
    assert np.all(lit[~esc] == 0)
    # the escape direction as the device forms it: FP64 sine and cosine of the float angles,
    # each component rounded once to float
    th, ph = rays.theta.astype(np.float64), rays.phi.astype(np.float64)
    dirs = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1).astype(np.float32)
    cat_near = np.frombuffer(near, np.float32).reshape(-1, 8)[:n_near]
    cat_far = np.frombuffer(far, np.float32).reshape(-1, 8)[:n_far]
    want = np.zeros_like(lit)
    sel_n, sel_f = esc & ~far_side, esc & far_side
    want[sel_n] = O.starfield_sample(cat_near, dirs[sel_n])
    want[sel_f] = O.starfield_sample(cat_far, dirs[sel_f])
    assert want.max() > 0 and lit.max() > 0
    # the same lit pixels, and the light to the device's acos / exp rounding; a pixel whose
    # direction component falls within a double rounding of a float midpoint could differ by an
    # ulp of direction (and, next to a point-like star, be lit on one side only), which at ~1e-8
    # per component no frame of this size meets
    assert np.array_equal(lit.max(1) > 0, want.max(1) > 0)
    on = want.max(1) > 0
    rel = np.abs(lit - want).max(1)[on] / want.max(1)[on]
    assert rel.max() < 1e-5, rel.max()
    # each side sees only its own sphere: without sphere 0 the near side is dark, the far side
    # unchanged
    acc.set_starfield(0, None, None, 0)
    acc.launch(cam, m, o)
    lit1 = acc.getFrameBuffer().reshape(-1, 4)[:, :3] - base
    assert np.all(lit1[sel_n] == 0) and np.array_equal(lit1[sel_f], lit[sel_f])
    acc.set_starfield(1, None, None, 0)
