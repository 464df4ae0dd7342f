"""Star field (celestial spheres), on the CPU.

The golden fixture tests/golden/ref_starfield.npz holds the reference's own output
(StarfieldGenerator::generate and ::sampleStarfield, src/Sirius.Core/Environment/PHSF001A.h,
compiled in place into oracle/_ref by tests/golden/make_golden.py): three catalogues (the
wormhole scene's two spheres, seeds 42 and 43, and a configuration whose magnitude limit the
reference clamps) and sampled star light at directions on stars, at 1e-5 to 1e-2 rad from
stars and at random, with and without the depth-of-field blur.

Pins: the oracle's restatement (oracle/oracle_py.py) and the library's host generator
(sirius_hip_starfield_generate, no device involved) reproduce the catalogues bit for bit; the
oracle's sampler reproduces the sampled light to float rounding.
"""
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from sirius_amd import abi

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_starfield.npz")
CATS = {"near": (4096, 1.0, 1000.0, 20.0, 42), "far": (4096, 1.0, 1000.0, 20.0, 43),
        "clamped": (2000, 1.0, 10000.0, 25.0, 7)}
SAMPLES = {"rgb_dof": ("near", dict(aperture_mm=50.0, focus_pc=100.0, brightness=1.0, dof=True)),
           "rgb_nodof": ("near", dict(aperture_mm=0.0, focus_pc=100.0, brightness=2.0, dof=False)),
           "rgb_far_dof": ("far", dict(aperture_mm=50.0, focus_pc=100.0, brightness=1.0, dof=True))}


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", sorted(CATS))
def test_oracle_catalogue_is_the_references(gold, name):
    assert np.array_equal(O.starfield_generate(*CATS[name]), gold[f"cat_{name}"])


@pytest.mark.parametrize("name", sorted(CATS))
def test_library_catalogue_is_the_references(gold, name):
    from sirius_amd.accelerator import starfield_default, starfield_generate
    count, dmin, dmax, maglim, seed = CATS[name]
    cfg = starfield_default()
    cfg.star_count, cfg.min_distance_pc, cfg.max_distance_pc = count, dmin, dmax
    cfg.magnitude_limit, cfg.seed = maglim, seed
    stars, n = starfield_generate(cfg)
    got = np.frombuffer(stars, np.float32).reshape(-1, 8)[:n]
    assert np.array_equal(got, gold[f"cat_{name}"])


def test_library_defaults_and_clamps():
    from sirius_amd.accelerator import starfield_default
    c = starfield_default()
    assert (c.star_count, c.seed, c.min_distance_pc, c.max_distance_pc, c.magnitude_limit) == (100000, 42, 1.0, 10000.0, 12.0)
    assert (c.aperture_mm, c.focus_distance_pc, c.brightness_scale, c.enable_dof) == (50.0, 100.0, 1.0, 1)
    import ctypes
    assert ctypes.sizeof(abi.Star) == 32 and ctypes.sizeof(abi.Starfield) == 40


@pytest.mark.parametrize("key", sorted(SAMPLES))
def test_oracle_sampler_matches_reference(gold, key):
    cat, kw = SAMPLES[key]
    got = O.starfield_sample(gold[f"cat_{cat}"], gold["dirs"], **kw)
    ref = gold[key]
    lit = ref.max(1) > 0
    assert np.array_equal(got.max(1) > 0, lit)
    rel = np.abs(got - ref).max(1)[lit] / ref.max(1)[lit]
    assert rel.max() < 2e-6, rel.max()
    # directions on a star (the first 64) nearly all see it: in float, acos of a cosine one
    # ulp below 1 is ~3.5e-4 rad, so a star whose spread is ~1e-6 rad (one near the focus
    # distance) can miss the very direction it was drawn at - in the reference too
    if cat == "near" and kw["dof"]:  # drawn on the near sphere's stars; without blur a quarter miss
        assert lit[:64].mean() > 0.9


def test_sampler_closed_forms(gold):
    """A star alone: at its own direction with no blur the light is 10^(-0.4 m) times its
    colour times the brightness (the spread is exactly 1); at 5.3 sigma and beyond it is 0."""
    star = gold["cat_near"][:1]
    d0 = star[:, :3] / np.linalg.norm(star[:, :3], axis=1, keepdims=True)
    rgb = O.starfield_sample(star, d0.astype(np.float32), aperture_mm=0.0, dof=False, brightness=3.0)
    inten = np.float32(10.0) ** (np.float32(-0.4) * star[0, 4]) * np.float32(3.0)
    assert abs(rgb[0].max() - inten) <= 1e-6 * inten  # the colour is normalised to its largest channel
    # a star at 1 pc with a 50 mm aperture: sigma = 0.05 |1/d - 1/100| = 0.0495 rad, where float
    # acos resolves the angle (near 1 it steps by ~3.5e-4 rad)
    far = np.array([[0.6, 0.0, 0.8, 1.0, 5.0, 0.6, 6500.0, 0.0]], np.float32)
    sig = 50.0 * abs(1.0 / far[0, 3] - 1.0 / 100.0) * 1e-3
    v = far[0, :3].astype(np.float64)
    t = np.cross(v, [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t)
    for k, expect_lit in ((0.5, True), (6.0, False)):
        a = k * sig
        d = (np.cos(a) * v + np.sin(a) * t)[None].astype(np.float32)
        got = O.starfield_sample(far, d)[0].max()
        assert (got > 0) == expect_lit, (k, got)
        if expect_lit:
            want = np.float32(10.0) ** (np.float32(-0.4) * far[0, 4]) * np.exp(-0.5 * k * k)
            assert abs(got - want) < 1e-3 * want
