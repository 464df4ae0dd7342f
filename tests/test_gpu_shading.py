"""The launch framebuffer's shading through the C-ABI, against the shading restatement
(oracle/oracle_shade.py, pinned to the reference's own colour functions) applied to the same
launch's ray results: RenderSession::shadeDiskHit's colour chain for thin-disk hits
(SNRS001A.cpp:315-434) under both temperature models, and shadePixel's g x g sample grid
(SNRS001A.cpp:480-560).

Bar: relative 2e-5 per channel (floored at 1e-6 of the frame's brightest value): the device and
the restatement take the same steps in the same widths; their exp, pow and the Page-Thorne flux
maximum differ in the last bits."""
import numpy as np
import pytest

from oracle import oracle_py as O
from oracle import oracle_shade as S
from sirius_amd import abi
from sirius_amd.accelerator import default_opts

pytestmark = pytest.mark.gpu


def _expected(acc, cam, m, rays, sh, r_in, r_out):
    """Per-ray radiance [n, 3] of the launch's results by the restatement."""
    n = len(rays.cls)
    out = np.zeros((n, 3), np.float32)
    dk = np.nonzero(rays.cls == abi.RAY_DISK)[0]
    te = np.zeros(len(dk))
    cos_i = np.zeros(len(dk))
    for j, i in enumerate(dk):
        r = float(rays.disk_r[i])
        if sh.temperature_model == 1:
            te[j] = sh.disk_temperature * O.lib().oracle_disk_temperature_ratio(r, m.a, r_in, r_out)
        else:
            te[j] = sh.disk_temperature * (r_in / r) ** 0.75
        k, _, _ = O.rhs(m, rays.state[i, :4], rays.state[i, 4:])
        cos_i[j] = abs(k[3]) / np.linalg.norm(k[1:])
    out[dk] = S.disk_radiance(te, rays.redshift[dk], cos_i, sh.disk_boost)
    stuck = (rays.cls == abi.RAY_MAX_STEPS) | (rays.cls == abi.RAY_FAILED)
    out[stuck] = np.float32(sh.max_steps_level)
    return out


def _compare(fb, want):
    scale = float(np.abs(want).max())
    err = np.abs(fb - want) / np.maximum(np.abs(want), 1e-6 * scale)
    return float(err.max()), scale


@pytest.mark.parametrize("model,spp", [(1, 1), (0, 4), (1, 9)])
def test_launch_shading_matches_restatement(acc, model, spp):
    w, h = 48, 27
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, w, h, disk=(0.0, 30.0))
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    r_in, r_out = O.lib().oracle_isco(0.9), 30.0
    try:
        sh = acc.set_shading(samples_per_pixel=spp, temperature_model=model, disk_temperature=40000.0)
        acc.initialise(w, h)
        print(f"[case] launch shading model {model} spp {spp} {w}x{h}", flush=True)
        acc.launch(cam, m, o)
        fb = acc.getFrameBuffer().reshape(h, w, 4).copy()
        rays = acc.fetch_results(with_state=True)
        gs = int(np.floor(np.sqrt(spp)))
        assert len(rays.cls) == w * h * gs * gs
        per_ray = _expected(acc, cam, m, rays, sh, r_in, r_out)
        want = S.pixel_average(per_ray.reshape(h * gs, w * gs, 3), w, h, gs)
        err, scale = _compare(fb[..., :3], want)
        print(f"[shading] disk rays {(rays.cls == abi.RAY_DISK).sum()}, brightest {scale:.4g}, "
              f"max relative error {err:.2e}", flush=True)
        assert scale > 0 and err < 2e-5
        assert np.all(fb[..., 3] == 1.0)
    finally:
        acc.set_shading()


def test_launch_temperature_scale_and_models_differ(acc):
    """The colour follows LaunchConfig::diskTemperature: a hotter disk is bluer, and the two
    temperature models light the inner disk differently (Novikov-Thorne's zero-torque edge is
    dark, Shakura-Sunyaev's hottest)."""
    w, h = 48, 27
    m = abi.metric_ks(1.0, 0.5)
    cam = abi.camera(40.0, 60.0, 25.0, w, h, disk=(0.0, 20.0))
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    acc.initialise(w, h)
    frames = {}
    try:
        for key, (model, T) in {"nt_cool": (1, 8000.0), "nt_hot": (1, 60000.0), "ss": (0, 60000.0)}.items():
            acc.set_shading(temperature_model=model, disk_temperature=T)
            print(f"[case] launch {key}", flush=True)
            acc.launch(cam, m, o)
            frames[key] = acc.getFrameBuffer().reshape(h, w, 4)[..., :3].copy()
    finally:
        acc.set_shading()
    rays = acc.fetch_results()
    disk = rays.cls.reshape(h, w) == abi.RAY_DISK
    cool, hot = frames["nt_cool"][disk], frames["nt_hot"][disk]
    lit = (cool.max(1) > 0) & (hot.max(1) > 0)
    assert lit.sum() > 100
    assert np.median(cool[lit, 2] / cool[lit, 0]) < np.median(hot[lit, 2] / hot[lit, 0])
    assert not np.allclose(frames["ss"][disk], frames["nt_hot"][disk])


def test_shading_settings_are_validated(acc):
    with pytest.raises(Exception):
        acc.set_shading(samples_per_pixel=0)
    with pytest.raises(Exception):
        acc.set_shading(temperature_model=2)
    with pytest.raises(Exception):
        acc.set_shading(disk_temperature=-1.0)
    acc.set_shading()
