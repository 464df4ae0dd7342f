"""The shading restatement (oracle/oracle_shade.py) against the reference's own disk-colour
functions (tests/golden/ref_shading.npz: blackbodyToRGB, applyColorMode TrueColor and
applyLimbDarkening from oracle/_ref, the reference compiled in place), and against what the
physics fixes: Wien's displacement of the colour, the white point near 6500 K and the limb
law's end points."""
import numpy as np

from oracle import oracle_shade as S


def test_blackbody_matches_the_reference(golden):
    d = golden("ref_shading.npz")
    got = S.blackbody_rgb(d["T"])
    # same steps and arithmetic widths; the C library's exp and pow may differ from numpy's by
    # an ulp, which the float sums carry to a few float ulps of the normalised colour
    np.testing.assert_allclose(got, d["bb_rgb"], rtol=0, atol=2e-6)


def test_true_color_and_limb_darkening_match_the_reference(golden):
    d = golden("ref_shading.npz")
    got = S.true_color(d["tc_T"], d["tc_g"], d["tc_I"])
    np.testing.assert_allclose(got, d["tc_rgb"], rtol=3e-6, atol=1e-6)
    got = S.limb_darkening(d["limb_in"], d["limb_cos"])
    np.testing.assert_array_equal(got, d["limb_rgb"])


def test_blackbody_colour_follows_the_physics():
    rgb = S.blackbody_rgb([3000.0, 6500.0, 25000.0])
    assert np.all(rgb.max(1) == 1.0)
    assert rgb[0, 0] > rgb[0, 1] > rgb[0, 2]      # cool: red
    assert rgb[2, 2] == 1.0 and rgb[2, 0] < rgb[2, 2]  # hot: blue
    assert np.all(rgb[1] > 0.8)                   # near the D65 white point: nearly neutral
    # the colour moves monotonically blue-ward with temperature (Wien)
    T = np.geomspace(2000.0, 40000.0, 60)
    br = S.blackbody_rgb(T)
    ratio = br[:, 2] / br[:, 0]
    assert np.all(np.diff(ratio) > 0)


def test_limb_law_end_points():
    rgb = np.ones((3, 3), np.float32)
    out = S.limb_darkening(rgb, [1.0, 1e-12, 0.0])
    assert np.allclose(out[0], 1.0)                   # face-on: no darkening
    u = S.limb_coefficients()
    assert np.allclose(out[1], 1.0 / (1.0 + u), rtol=1e-7)  # grazing: 1 / (1 + u)
    assert np.all(out[2] == 0.0)                      # edge-on and below: no light
    assert u[0] < u[1] < u[2]                          # blue darkens most


def test_disk_radiance_scales_as_g_to_the_eighth_at_fixed_colour():
    # T_obs fixed (T_e g constant), so the colour and t^4 are fixed: what remains is g^4 (TrueColor)
    # times g^4 (beaming)
    g = np.array([0.5, 1.0], np.float32)
    te = np.array([12000.0 / 0.5, 12000.0])
    out = S.disk_radiance(te, g, [1.0, 1.0])
    assert np.allclose(out[0] / out[1], 0.5 ** 8, rtol=1e-5)
