"""Host substrate and C ABI checks that need no GPU: loader, variances,
noise, PSNR, image I/O, the run-length packer, kernel-set validation, the
exported symbols and the driver's command line."""
import ctypes
import gzip
import hashlib
import json
import os
import re
import struct
import subprocess
import zlib

import numpy as np
import pytest

import lpa_ici
import oracle
from conftest import GOLDEN, PKG, REF_KERNEL_DIR, REPO, SEED, SIGMA, SYMMETRIC, kernel_path, kernels

DRIVER = os.path.join(PKG, "bin", "lpa_ici_denoise")
SETS = sorted(json.load(open(os.path.join(GOLDEN, "kernels.json"))).keys())
needs_ref = pytest.mark.skipif(not oracle.ref_available("v0_0") or not os.path.isdir(REF_KERNEL_DIR),
                               reason="reference build (oracle/_ref) or tree not present")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


# ---------------------------------------------------------------------------
# kernel files
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", SETS)
def test_loader_matches_reference_fixture(name):
    want = json.load(open(os.path.join(GOLDEN, "kernels.json")))[name]
    ks = lpa_ici.load_kernels(kernel_path(name))
    assert (ks.D, ks.S) == (want["D"], want["S"])
    assert ks.rows.tolist() == want["rows"] and ks.cols.tolist() == want["cols"]
    assert sha(ks.weights, ks.offsets, ks.rows, ks.cols) == want["sha256"]


@needs_ref
@pytest.mark.parametrize("name", SETS)
def test_loader_matches_reference_build_plain_text(name):
    ks = lpa_ici.load_kernels(os.path.join(REF_KERNEL_DIR, name))
    D, S, w, off, rows, cols = oracle.ref_load_kernels(os.path.join(REF_KERNEL_DIR, name))
    assert (ks.D, ks.S) == (D, S)
    assert np.array_equal(ks.weights.view(np.uint32), w.view(np.uint32))
    assert np.array_equal(ks.offsets, off) and np.array_equal(ks.rows, rows) and np.array_equal(ks.cols, cols)


def test_default_set_structure(ks):
    """Appendix facts of the default set: 16x7, crops 1..63, 6008 nonzeros,
    every scale-0 kernel the 1x1 identity."""
    assert (ks.D, ks.S) == (16, 7)
    assert ks.rows.reshape(16, 7)[0].tolist() == [1, 3, 7, 15, 31, 47, 63]
    assert int((ks.weights != 0).sum()) == 6008
    assert np.all(ks.weights[ks.offsets[0::7]] == 1.0)


def write_kernel_file(path, text, gz=False):
    if gz:
        with gzip.open(path, "wt") as f:
            f.write(text)
    else:
        open(path, "w").write(text)
    return str(path)


GOOD = "2 2\n1 1\n1.0\n3 3\n0 0 0 0 .5 .5 0 0 0\n1 1\n1.0\n5 5\n" + " ".join(["0.04"] * 25) + "\n"


@pytest.mark.parametrize("text,message", [
    ("x y\n", "invalid header"),
    ("0 3\n", "invalid header"),
    ("1 1\n0 3\n", "invalid size of kernel 0, 0"),
    ("1 2\n1 1\n1.0\n3 3\n0.1 0.2\n", "missing weights of kernel 0, 1"),
    ("2 1\n3 3\n" + "0 " * 9 + "\n1 1\n1.0\n", "is too small to crop"),
])
def test_loader_errors(tmp_path, text, message):
    p = write_kernel_file(tmp_path / "k.txt", text)
    with pytest.raises(lpa_ici.LpaError) as e:
        lpa_ici.load_kernels(p)
    assert message in str(e.value) and e.value.status == 1
    if oracle.ref_available("v0_0"):
        assert oracle.ref_load_kernels(p) is None  # the reference rejects it too


def test_loader_missing_file_and_gzip(tmp_path):
    with pytest.raises(lpa_ici.LpaError) as e:
        lpa_ici.load_kernels(str(tmp_path / "nope.txt"))
    assert e.value.status == 5 and "Failed to open" in str(e.value)
    a = lpa_ici.load_kernels(write_kernel_file(tmp_path / "k.txt", GOOD))
    b = lpa_ici.load_kernels(write_kernel_file(tmp_path / "k.txt.gz", GOOD, gz=True))
    assert np.array_equal(a.weights, b.weights) and a.rows.tolist() == [1, 3, 1, 3]  # dir 1 cropped to 3x3
    assert np.allclose(b.kernel(1, 1), 0.04)


def test_variances_fixture_and_closed_form(ks):
    want = json.load(open(os.path.join(GOLDEN, "variances.json")))
    v = ks.variances(SIGMA)
    assert [float(a).hex() for a in v] == want["variances_hex"]
    # uniform wedges: sum(w^2) = 1/n up to the 6-decimal printing of 1/n
    n = np.array([(ks.kernel(d, s) != 0).sum() for d in range(16) for s in range(7)])
    s2 = np.float32(SIGMA) ** 2
    assert np.allclose(v / s2, 1.0 / n, rtol=2e-4)


# ---------------------------------------------------------------------------
# noise, PSNR, images
# ---------------------------------------------------------------------------

def test_noise_matches_reference_fixture():
    want = json.load(open(os.path.join(GOLDEN, "noise.json")))
    nz = lpa_ici.make_noisy(np.zeros((512, 512), np.float32), SIGMA, SEED)
    assert [float(a).hex() for a in nz.ravel()[:64]] == want["first"]
    assert sha(nz) == want["sha256_512x512"]
    assert abs(nz.std() - SIGMA) < 1e-3 * 5
    assert not np.array_equal(nz, lpa_ici.make_noisy(np.zeros((512, 512), np.float32), SIGMA, SEED + 1))


def test_psnr_formula():
    rng = np.random.default_rng(3)
    a = rng.random(1000).astype(np.float32)
    b = (a + rng.normal(0, 0.05, 1000)).astype(np.float32)
    d = (a - b).astype(np.float32)
    mse = np.sum((d * d).astype(np.float32).astype(np.float64)) / a.size
    assert lpa_ici.psnr(a, b) == 10.0 * np.log10(1.0 / mse)


@pytest.mark.parametrize("name", ["barbara", "cameraman"])
def test_png_loader_matches_stb_fixture(name):
    want = json.load(open(os.path.join(GOLDEN, "images.json")))[name]
    img = lpa_ici.load_image(os.path.join(GOLDEN, "images", name + ".png"))
    assert list(img.shape) == want["shape"] and sha(img) == want["sha256_loaded"]
    assert sha(lpa_ici.make_noisy(img, SIGMA, SEED)) == want["sha256_noisy"]
    assert lpa_ici.psnr(img, lpa_ici.make_noisy(img, SIGMA, SEED)) == want["psnr_noisy"]


def png_bytes(arr, ctype):
    h, w = arr.shape[:2]
    raw = b"".join(b"\x00" + arr[r].tobytes() for r in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def test_png_rgb_to_gray(tmp_path):
    rng = np.random.default_rng(0)
    rgb = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    p = tmp_path / "rgb.png"
    p.write_bytes(png_bytes(rgb, 2))
    g = (rgb[..., 0].astype(int) * 77 + rgb[..., 1].astype(int) * 150 + rgb[..., 2].astype(int) * 29) >> 8
    img = lpa_ici.load_image(str(p))
    expect = np.power((g.astype(np.float32) / np.float32(255.0)), np.float32(2.2))
    assert np.allclose(img, expect, rtol=1e-6, atol=0)
    if oracle.ref_available("v0_0"):
        assert np.array_equal(img, oracle.ref_loadf(str(p)))


def test_png_errors(tmp_path):
    (tmp_path / "bad.png").write_bytes(b"not a png")
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.load_image(str(tmp_path / "bad.png"))
    with pytest.raises(lpa_ici.LpaError) as e:
        lpa_ici.load_image(str(tmp_path / "missing.png"))
    assert e.value.status == 5


def decode_hdr(path):
    data = open(path, "rb").read()
    head, _, body = data.partition(b"\n\n")
    dims, _, pix = body.partition(b"\n")
    m = re.match(rb"-Y (\d+) \+X (\d+)", dims)
    h, w = int(m.group(1)), int(m.group(2))
    q = np.frombuffer(pix, np.uint8).reshape(h, w, 4).astype(np.float64)
    val = np.where(q[..., 3] > 0, (q[..., 0] + 0.5) * np.ldexp(1.0, (q[..., 3] - 136).astype(int)), 0.0)
    return head, val


def test_hdr_writer(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.random((6, 9)).astype(np.float32)
    img[0, 0], img[0, 1] = -0.3, 0.0
    p = str(tmp_path / "o.hdr")
    lpa_ici.write_hdr(p, img)
    head, val = decode_hdr(p)
    assert head.startswith(b"#?RADIANCE") and b"FORMAT=32-bit_rle_rgbe" in head
    assert val[0, 0] == 0 and val[0, 1] == 0  # negatives and zero clamp to 0
    mask = img > 0
    assert np.all(np.abs(val[mask] - img[mask]) <= img[mask] / 128 + 1e-9)


def test_synth_image_deterministic():
    a = lpa_ici.synth_image(300, 200, 5)
    assert np.array_equal(a, lpa_ici.synth_image(300, 200, 5))
    assert not np.array_equal(a, lpa_ici.synth_image(300, 200, 6))
    assert a.min() >= 0 and a.max() <= 1 and a.shape == (200, 300)
    assert np.array_equal(golden_synth(), lpa_ici.synth_image(64, 64, 1))


def golden_synth():
    return np.load(os.path.join(GOLDEN, "synth64_clean.npy"))


# ---------------------------------------------------------------------------
# run-length layout and kernel-set validation (host side; no device needed)
# ---------------------------------------------------------------------------

def dense_from_runs(ks, r):
    out = np.zeros_like(ks.weights)
    for k in range(ks.K):
        R, C, o = ks.rows[k], ks.cols[k], ks.offsets[k]
        m = out[o:o + R * C].reshape(R, C)
        for i in range(r["kernel_first_run"][k], r["kernel_first_run"][k] + r["kernel_num_runs"][k]):
            row, col, ln, wo = r["run_row"][i] + R // 2, r["run_col"][i] + C // 2, r["run_len"][i], r["run_woff"][i]
            m[row, col:col + ln] = r["weights"][wo:wo + ln]
    return out


@pytest.mark.parametrize("name,runs,weights,interior", [
    (SYMMETRIC, 1052, 6008, 0),
    ("lpa_kernels_m_1_0_d_4_square.txt", None, None, 16),
    ("lpa_kernels_python_16_tapered.txt", 1194, 12320, 48),
])
def test_run_length_round_trip(name, runs, weights, interior):
    """Rebuilding the dense kernels from the runs gives them back exactly
    (max |difference| 0, the check of check_sparse_kernels,
    src/v4_1_lpa_ici_2D_sparse_shared.cu:879-921; -0.0 entries of a file
    come back as +0.0)."""
    ks = kernels(name)
    r = ks.runs()
    back = dense_from_runs(ks, r)
    assert np.abs(back - ks.weights).max() == 0.0
    nz = ks.weights != 0
    assert np.array_equal(back[nz].view(np.uint32), ks.weights[nz].view(np.uint32))
    info = ks.info()
    if runs is not None:
        assert info["num_runs"] == runs and info["num_weights"] == weights
    assert int((r["weights"] == 0).sum()) == interior
    assert info["num_nonzero"] == int((ks.weights != 0).sum()) == int(r["nnz"].sum())
    assert info["max_radius"] <= 31
    # sigma^2 * sumsq is compute_kernel_variances' value, bit for bit
    s2 = np.float32(SIGMA) * np.float32(SIGMA)
    assert np.array_equal((s2 * r["sumsq"]).astype(np.float32), ks.variances(SIGMA))


def test_ring_form_detection():
    """The default set: uniform weights, nested supports, 3172 ring taps
    (6008 nonzeros); the m=(2,0) gaussian set is not in ring form."""
    info = kernels(SYMMETRIC).info()
    assert info["ring_form"] == 1 and info["num_ring_taps"] == 3172
    for d in range(16):
        prev = set()
        for s in range(7):
            k = kernels(SYMMETRIC).kernel(d, s)
            pos = {(r - k.shape[0] // 2, c - k.shape[1] // 2) for r, c in zip(*np.nonzero(k))}
            assert prev <= pos and len(set(k[k != 0].tolist())) == 1
            prev = pos
    assert kernels("lpa_kernels_m_2_0_d_16_h12_1_2_4_8_16_24_32_gaussian.txt").info()["ring_form"] == 0


def test_kset_rejects_radius_over_31():
    """The reference exits above a 63x63 kernel (src/v4_1_lpa_ici_2D_sparse_shared.cu:1189-1195);
    here it is LPA_ETOOBIG, and a non-positive size is LPA_EINVAL."""
    w = np.zeros(65 * 65, np.float32)
    w[65 * 32 + 32] = 1.0
    off, r65, r0, c5 = (np.array([v], np.int32) for v in (0, 65, 0, 5))
    h = ctypes.c_void_p()
    lib = lpa_ici.lib()
    st = lib.lpa_kset_create(1, 1, w.ctypes.data, off.ctypes.data, r65.ctypes.data, r65.ctypes.data,
                             ctypes.byref(h))
    assert st == 2 and "max supported is 63x63" in lib.lpa_last_error().decode()
    st = lib.lpa_kset_create(1, 1, w.ctypes.data, off.ctypes.data, r0.ctypes.data, c5.ctypes.data, ctypes.byref(h))
    assert st == 1 and not h.value


def test_forward_rejects_bad_batches():
    """Batches are checked before any device work: a side above 65535 or more
    than 2^31 - 1 pixels is LPA_ETOOBIG; a non-positive side, a pitch below the
    width, an empty batch or a null set is LPA_EINVAL (the pointers are never
    dereferenced)."""
    ks = lpa_ici.load_kernels()
    lib = lpa_ici.lib()
    cases = [((65536, 1, 0), 2), ((1, 65536, 0), 2), ((65535, 32769, 0), 2), ((0, 5, 0), 1), ((5, -1, 0), 1),
             ((10, 10, 9), 1)]
    for (w, h, pitch), want in cases:
        img = lpa_ici.LpaImage(0x1000, 0x2000, None, w, h, pitch, 20.0 / 255.0, 2.0, None, None)
        arr = (lpa_ici.LpaImage * 1)(img)
        for call in (lambda: lib.lpa_ici_forward(ks.handle, arr, 1, 0, None),
                     lambda: lib.lpa_ici_forward_noalloc(ks.handle, arr, 1, 0, None)):
            assert call() == want, (w, h, pitch, lib.lpa_last_error().decode())
    assert "pitch" in lib.lpa_last_error().decode()
    # an empty batch, a null image array and a null kernel set are refused
    arr = (lpa_ici.LpaImage * 1)(lpa_ici.LpaImage(0x1000, 0x2000, None, 8, 8, 0, 20.0 / 255.0, 2.0, None, None))
    assert lib.lpa_ici_forward(ks.handle, arr, 0, 0, None) == 1 and "no images" in lib.lpa_last_error().decode()
    assert lib.lpa_ici_forward(ks.handle, None, 1, 0, None) == 1
    assert lib.lpa_ici_forward(None, arr, 1, 0, None) == 1 and "kernel set" in lib.lpa_last_error().decode()
    assert lib.lpa_ici_forward_noalloc(ks.handle, arr, 0, 0, None) == 1
    assert lib.lpa_ici_backward(ks.handle, arr, None, None, None, 0, None) == 1


def test_exports_every_header_symbol():
    header = open(os.path.join(REPO, "include", "lpa_ici.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(lpa_[a-z_0-9]+)\s*\(", header, re.M))
    assert len(names) >= 20
    lib = lpa_ici.lib()
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(lpa_ici._lib.SIGNATURES)


# ---------------------------------------------------------------------------
# driver command line (fails before touching the GPU)
# ---------------------------------------------------------------------------

def driver(*args):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=60)


def test_driver_help_and_errors(tmp_path):
    r = driver("-h")
    assert r.returncode == 0 and "-k, --kernels <file>" in r.stdout and "--seed <value>" in r.stdout
    r = driver("-s", "-3")
    assert r.returncode == 1 and "-s needs a positive number, got '-3'" in r.stderr
    r = driver("--seed", "12x")
    assert r.returncode == 1 and "--seed needs an unsigned integer" in r.stderr
    r = driver("--bogus", "1")
    assert r.returncode == 1 and "unknown option '--bogus'" in r.stderr
    r = driver("-g")
    assert r.returncode == 1 and "missing value for '-g'" in r.stderr
    r = driver("-i", str(tmp_path / "nope.png"))
    assert r.returncode == 1 and "Error in loading the image" in r.stderr
    assert f"Output: {tmp_path}/nope_lpa_ici_denoise_gpu.hdr" in r.stdout
    img = os.path.join(GOLDEN, "images", "cameraman.png")
    r = driver("-i", img, "-k", str(tmp_path / "missing.txt"), "-o", str(tmp_path / "out.hdr"))
    assert r.returncode == 1 and "Failed to open" in r.stderr
    assert (tmp_path / "gray_image.hdr").exists() and (tmp_path / "gray_image_noisy.hdr").exists()
    assert "Calculated PSNR of host_img_gray_noisy: 22.1245" in r.stdout


def test_driver_usage_and_echo_text(tmp_path):
    """The usage lines the reference prints (src/Config.hpp:27-37), column for
    column, with the two added flags among them; the echo block and the
    default output name (src/Config.hpp:79-89)."""
    r = driver("--help")
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    assert lines[0] == f"Usage: {DRIVER} [options]"
    assert lines[1:] == [
        "  -i, --image <file>     input image (default data/barbara.png)",
        "  -k, --kernels <file>   LPA kernel file (default "
        "kernels/lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt)",
        "  -o, --output <file>    denoised .hdr output (default <image dir>/<image name>_<version>.hdr)",
        "  -s, --sigma <value>    noise standard deviation on the 0-255 scale (default 20)",
        "  -g, --gamma <value>    ICI threshold (default 2)",
        "      --seed <value>     noise seed (default 20250910)",
        "      --debug <dir>      also write each direction's estimate and scale map as .hdr in <dir>",
        "  -c, --channels <1|3>   1: gray (default); 3: RGB, every channel denoised on its own",
        "  -h, --help             show this message",
    ]
    # an unknown flag without a value reads as a missing value; an error is
    # "Error: ..." and a blank line on stderr, the usage on stdout
    r = driver("--bogus")
    assert r.returncode == 1 and r.stderr.startswith("Error: unknown option or missing value for '--bogus'\n\n")
    assert r.stdout.splitlines()[0] == f"Usage: {DRIVER} [options]"
    r = driver("-c", "2")
    assert r.returncode == 1 and "-c needs 1 or 3, got '2'" in r.stderr
    r = driver("-s", "0")
    assert r.returncode == 1 and "-s needs a positive number, got '0'" in r.stderr
    # a value that looks like a flag is a value; the output name keeps every
    # dot of the stem but the last
    img = str(tmp_path / "a.b.png")
    r = driver("--image", img, "-s", "10", "-g", "1.5", "--seed", "7", "-k", "-h")
    assert r.returncode == 1 and "Error in loading the image" in r.stderr
    assert r.stdout.splitlines()[:4] == [f"Image: {img}", "Kernels: -h", f"Output: {tmp_path}/a.b_lpa_ici_denoise_gpu.hdr",
                                         "Sigma: 10/255, gamma: 1.5, seed: 7"]
    r = driver("-i", "noext", "-o", str(tmp_path / "o.hdr"))
    assert f"Output: {tmp_path}/o.hdr" in r.stdout
    r = driver("-i", "noext")
    assert "Output: noext_lpa_ici_denoise_gpu.hdr" in r.stdout


def write_png(path, img):
    """8-bit PNG (gray [H, W] or RGB [H, W, 3]), filter 0, via zlib."""
    import struct
    import zlib
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    ctype = 0 if img.ndim == 2 else 2
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)
    data = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
    open(path, "wb").write(data)


def test_load_image_channels(tmp_path):
    """stbi_loadf semantics for 1 and 3 channels: a gray PNG loaded as RGB is
    its gray load replicated; an RGB PNG loaded as gray is the gray load of
    the PNG holding (77R + 150G + 29B) >> 8; both bit for bit."""
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, size=(7, 11, 3), dtype=np.uint8)
    gray = rng.integers(0, 256, size=(7, 11), dtype=np.uint8)
    write_png(tmp_path / "rgb.png", rgb)
    write_png(tmp_path / "gray.png", gray)
    folded = ((rgb[..., 0].astype(np.int32) * 77 + rgb[..., 1].astype(np.int32) * 150 +
               rgb[..., 2].astype(np.int32) * 29) >> 8).astype(np.uint8)
    write_png(tmp_path / "folded.png", folded)
    g1 = lpa_ici.load_image(str(tmp_path / "gray.png"))
    g3 = lpa_ici.load_image(str(tmp_path / "gray.png"), channels=3)
    assert g3.shape == (7, 11, 3) and all(np.array_equal(g3[..., c], g1) for c in range(3))
    r1 = lpa_ici.load_image(str(tmp_path / "rgb.png"))
    assert np.array_equal(r1, lpa_ici.load_image(str(tmp_path / "folded.png")))
    r3 = lpa_ici.load_image(str(tmp_path / "rgb.png"), channels=3)
    # each RGB channel is the gray curve at that channel's 8-bit value
    for c in range(3):
        write_png(tmp_path / f"c{c}.png", rgb[..., c])
        assert np.array_equal(r3[..., c], lpa_ici.load_image(str(tmp_path / f"c{c}.png")))
    with pytest.raises(lpa_ici.LpaError):
        lpa_ici.load_image(str(tmp_path / "rgb.png"), channels=2)


def test_hdr_writer_rgb(tmp_path):
    """Shared exponent per pixel: every component within 1/128 of the largest."""
    rng = np.random.default_rng(4)
    img = rng.random((5, 8, 3)).astype(np.float32) * 3
    img[0, 0] = [-1.0, 0.0, 0.0]
    p = str(tmp_path / "c.hdr")
    lpa_ici.write_hdr(p, img)
    data = open(p, "rb").read()
    _, _, body = data.partition(b"\n\n")
    _, _, pix = body.partition(b"\n")
    q = np.frombuffer(pix, np.uint8).reshape(5, 8, 4).astype(np.float64)
    scale = np.where(q[..., 3] > 0, np.ldexp(1.0, (q[..., 3] - 136).astype(int)), 0.0)
    val = (q[..., :3] + 0.5) * scale[..., None] * (q[..., :3] > 0)
    m = img.max(axis=2, keepdims=True)
    ok = m[..., 0] > 0
    assert np.all(np.abs(val - np.maximum(img, 0))[ok] <= (m / 128)[ok] + 1e-9)
    assert np.all(q[0, 0] == 0)
