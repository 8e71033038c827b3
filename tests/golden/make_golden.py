"""Write the golden fixtures under tests/golden/ from the reference's own CPU
code (oracle/_ref/libref_v0_0.so, built from the reference's sources by
`make -C oracle ref`).  Run where the reference tree exists:

    python tests/golden/make_golden.py

Every fixture is data: inputs and the reference's outputs for them.
  kernels.json        per shipped kernel set: D, S, crop sizes, nonzeros and
                      the SHA-256 of the loader's weights/offsets/rows/cols
  variances.json      compute_kernel_variances(symmetric set, sigma=20/255)
  noise.json          create_noisy_image: first samples and SHA-256 of a
                      512x512 plane of zeros + noise (seed 20250910)
  images.json         stbi_loadf of barbara/cameraman (SHA-256), noisy and
                      denoised PSNR of both with the reference defaults
  cameraman_*.npy     noisy input and v0_0 output, cameraman, defaults
  synth64_*.npy       a 64x64 synthetic image, its noisy copy and v0_0
                      outputs with three kernel sets
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "lpa-ici-cuda_amd"))

import oracle  # noqa: E402
import lpa_ici  # noqa: E402

REF_KERNELS = "/root/reference/kernels"
REF_DATA = "/root/reference/data"
SIGMA = 20.0 / 255.0
SEED = 20250910
GAMMA = 2.0
SYNTH_SETS = {
    "symmetric": "lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt",
    "m20_gaussian": "lpa_kernels_m_2_0_d_16_h12_1_2_4_8_16_24_32_gaussian.txt",
    "d4_square": "lpa_kernels_m_1_0_d_4_square.txt",
}


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


class RefSet:
    """Kernel arrays in the shape the oracle wrappers take."""

    def __init__(self, D, S, w, off, rows, cols):
        self.D, self.S, self.weights, self.offsets, self.rows, self.cols = D, S, w, off, rows, cols


def load_ref_set(name):
    r = oracle.ref_load_kernels(os.path.join(REF_KERNELS, name))
    assert r is not None, name
    return RefSet(*r)


def main():
    assert oracle.ref_available("v0_0"), "build oracle/_ref first: make -C oracle ref"
    out = {}
    for name in sorted(os.listdir(REF_KERNELS)):
        ks = load_ref_set(name)
        out[name] = {
            "D": ks.D, "S": ks.S,
            "rows": ks.rows.tolist(), "cols": ks.cols.tolist(),
            "num_weights": int(ks.weights.size), "nonzero": int((ks.weights != 0).sum()),
            "sha256": sha(ks.weights, ks.offsets, ks.rows, ks.cols),
        }
    json.dump(out, open(os.path.join(HERE, "kernels.json"), "w"), indent=1)

    sym = load_ref_set(SYNTH_SETS["symmetric"])
    var = oracle.ref_variances(SIGMA, sym.D * sym.S)
    json.dump({"sigma": SIGMA, "set": SYNTH_SETS["symmetric"],
               "variances_hex": [float(v).hex() for v in var]},
              open(os.path.join(HERE, "variances.json"), "w"), indent=1)

    zeros = np.zeros((512, 512), np.float32)
    nz = oracle.ref_noisy(zeros, SIGMA, SEED)
    json.dump({"sigma": SIGMA, "seed": SEED, "first": [float(v).hex() for v in nz.ravel()[:64]],
               "sha256_512x512": sha(nz)}, open(os.path.join(HERE, "noise.json"), "w"), indent=1)

    images = {}
    for name in ("barbara", "cameraman"):
        clean = oracle.ref_loadf(os.path.join(REF_DATA, name + ".png"))
        noisy = oracle.ref_noisy(clean, SIGMA, SEED)
        den = oracle.ref_denoise(noisy, sym, var, GAMMA)
        images[name] = {"shape": list(clean.shape), "sha256_loaded": sha(clean), "sha256_noisy": sha(noisy),
                        "sha256_denoised": sha(den), "psnr_noisy": oracle.ref_psnr(clean, noisy),
                        "psnr_denoised": oracle.ref_psnr(clean, den)}
        if name == "cameraman":
            np.save(os.path.join(HERE, "cameraman_noisy.npy"), noisy)
            np.save(os.path.join(HERE, "cameraman_v0_0.npy"), den)
    json.dump(images, open(os.path.join(HERE, "images.json"), "w"), indent=1)

    clean = lpa_ici.synth_image(64, 64, 1)
    noisy = oracle.ref_noisy(clean, SIGMA, SEED)
    np.save(os.path.join(HERE, "synth64_clean.npy"), clean)
    np.save(os.path.join(HERE, "synth64_noisy.npy"), noisy)
    for tag, name in SYNTH_SETS.items():
        ks = load_ref_set(name)
        v = oracle.ref_variances(SIGMA, ks.D * ks.S)
        np.save(os.path.join(HERE, f"synth64_v0_0_{tag}.npy"), oracle.ref_denoise(noisy, ks, v, GAMMA))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
