"""The command-line driver end to end on the reference's published case:
barbara, defaults (sigma 20, gamma 2, seed 20250910, symmetric set) gives
noisy PSNR 22.1056 and denoised PSNR 25.0736 (README:7, :208), with the
reference's stdout lines and the three .hdr outputs."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG, SYMMETRIC, kernel_path

pytestmark = pytest.mark.gpu


def test_driver_barbara(tmp_path):
    out = tmp_path / "barbara_denoised.hdr"
    r = subprocess.run([os.path.join(PKG, "bin", "lpa_ici_denoise"), "-i",
                        os.path.join(GOLDEN, "images", "barbara.png"), "-k", kernel_path(SYMMETRIC), "-o", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert "Loaded image with a width of 512 px, a height of 512 px and (imported) 1 channels" in lines
    assert any(l.startswith("Time taken for directional loops + aggregation:") for l in lines)
    assert any(l.startswith("Time taken for lpa function execution:") for l in lines)
    info = json.loads(lines[-1])
    assert round(info["psnr_noisy"], 4) == 22.1056
    assert round(info["psnr_denoised"], 4) == 25.0736
    assert "Calculated PSNR of img_denoised: 25.0736" in r.stdout
    for f in (out, tmp_path / "gray_image.hdr", tmp_path / "gray_image_noisy.hdr"):
        assert f.exists() and f.stat().st_size > 512 * 512 * 4


def test_driver_debug_outputs(tmp_path):
    """--debug writes the reference's per-direction debugging outputs: each
    direction's selected estimate and its scale map s/(S-1)."""
    from test_host import decode_hdr
    dbg = tmp_path / "dbg"
    dbg.mkdir()
    out = tmp_path / "cameraman_denoised.hdr"
    r = subprocess.run([os.path.join(PKG, "bin", "lpa_ici_denoise"), "-i",
                        os.path.join(GOLDEN, "images", "cameraman.png"), "-k", kernel_path(SYMMETRIC), "-o", str(out),
                        "--debug", str(dbg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Saved 32 per-direction estimates and scale maps" in r.stdout
    for d in range(16):
        est = decode_hdr(dbg / f"gray_image_denoised_theta_{d}.hdr")[1]
        smap = decode_hdr(dbg / f"scale_map_theta_{d}.hdr")[1]
        assert est.shape == smap.shape == (256, 256)
        # s/(S-1) for S = 7, within the RGBE mantissa's 1/256 of the exponent
        levels = np.arange(7) / 6.0
        assert np.all(np.abs(smap[..., None] - levels).min(axis=-1) <= 1.0 / 128)
        assert 0.0 <= est.min() and est.max() <= 1.5
    den = decode_hdr(out)[1]
    assert abs(float(den.mean()) - float(decode_hdr(dbg / "gray_image_denoised_theta_0.hdr")[1].mean())) < 0.05


def test_driver_rgb(tmp_path):
    """--channels 3: RGB in and out; each channel denoised in one batched
    launch; the PSNR over all samples improves as for gray."""
    from test_host import decode_hdr, write_png
    import lpa_ici
    rgb = np.stack([(lpa_ici.synth_image(200, 150, s) * 255).astype(np.uint8) for s in (1, 2, 3)], axis=-1)
    src = tmp_path / "rgb.png"
    write_png(src, rgb)
    out = tmp_path / "rgb_denoised.hdr"
    r = subprocess.run([os.path.join(PKG, "bin", "lpa_ici_denoise"), "-i", str(src), "-k", kernel_path(SYMMETRIC),
                        "-o", str(out), "--channels", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Loaded image with a width of 200 px, a height of 150 px and (imported) 3 channels" in r.stdout
    info = json.loads(r.stdout.splitlines()[-1])
    assert info["psnr_denoised"] > info["psnr_noisy"] + 4
    head, _ = decode_hdr(out)
    data = open(out, "rb").read()
    q = np.frombuffer(data.partition(b"\n\n")[2].partition(b"\n")[2], np.uint8).reshape(150, 200, 4)
    assert not np.array_equal(q[..., 0], q[..., 1])  # the channels differ
