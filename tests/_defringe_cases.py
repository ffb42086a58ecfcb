"""Synthetic defringe scenes and the behavioural checks of the reference's own
defringe unit tests (tests/test_defringe.py:51-423 of the reference), written
against a `defringe(img, radius=10, strength=10, defringe_green=False)` callable
with the standalone (out-of-place) semantics."""
import numpy as np


def purple_fringe(h=64, w=64, width=2):
    img = np.zeros((h, w, 3), np.float32)
    mid = w // 2
    img[:, :mid, :] = 0.95
    img[:, mid:, :] = 0.15
    img[:, mid:mid + width, 0] = 0.55
    img[:, mid:mid + width, 1] = 0.10
    img[:, mid:mid + width, 2] = 0.55
    return img


def red_cloth(h=64, w=64):
    img = np.zeros((h, w, 3), np.float32)
    mid = w // 2
    img[:, :mid] = (0.8, 0.2, 0.25)
    img[:, mid:] = (0.6, 0.15, 0.2)
    return img


def red_cloth_white_stripe(h=96, w=96):
    img = np.zeros((h, w, 3), np.float32)
    img[:, :] = (0.82, 0.18, 0.34)
    img[:, w // 2 - 3:w // 2 + 3, :] = 0.92
    return img


def green_fringe(h=64, w=64, width=2):
    img = np.zeros((h, w, 3), np.float32)
    mid = w // 2
    img[:, :mid, :] = 0.92
    img[:, mid:, :] = 0.12
    img[:, mid:mid + width] = (0.08, 0.60, 0.08)
    return img


def purple_flower(h=64, w=64):
    img = np.full((h, w, 3), 0.20, np.float32)
    img[12:52, 12:52] = (0.45, 0.10, 0.45)
    return img


def micro_fringe(h=64, w=64):
    img = np.zeros((h, w, 3), np.float32)
    mid = w // 2
    img[:, :mid, :] = 0.98
    img[:, mid:, :] = 0.05
    img[:, mid] = (0.50, 0.05, 0.52)
    return img


def neutral_noise(h=64, w=64, seed=42):
    g = np.random.default_rng(seed).random((h, w, 1)).astype(np.float32) * 0.3 + 0.2
    return np.repeat(g, 3, axis=2)


def broad_purple(h=64, w=64):
    img = np.zeros((h, w, 3), np.float32)
    img[:, :] = (0.75, 0.10, 0.75)
    img[0, :, :] = 0.98
    return img


def check_all(defringe):
    """Run every invariant; returns a list of failure messages (empty = pass)."""
    fails = []

    def expect(cond, msg):
        if not cond:
            fails.append(msg)

    img = purple_fringe()
    out = defringe(img, strength=1.0)
    for col in (32, 33):
        r_in, g_in, b_in = img[32, col]
        r_o, g_o, b_o = out[32, col]
        expect(max(0, r_o - g_o) < max(0, r_in - g_in), f"purple R excess not reduced at {col}")
        expect(max(0, b_o - g_o) < max(0, b_in - g_in), f"purple B excess not reduced at {col}")
    r_o, g_o, b_o = out[32, 32]
    expect(r_o <= g_o + 0.02 and b_o <= g_o + 0.02, "full-strength purple: R/B above G+0.02")

    img = red_cloth()
    out = defringe(img, strength=10.0)
    expect(abs(img[32, 10, 0] - out[32, 10, 0]) < 0.02, "red cloth changed")

    img = red_cloth_white_stripe()
    out = defringe(img, strength=10.0)
    x = img.shape[1] // 2 - 4
    expect(float(np.abs(out[8:-8, x - 2:x + 1] - img[8:-8, x - 2:x + 1]).max()) < 0.02,
           "red cloth next to white stripe changed")

    img = green_fringe()
    out = defringe(img, strength=1.0, defringe_green=True)
    for col in (32, 33):
        r_in, g_in, b_in = img[32, col]
        r_o, g_o, b_o = out[32, col]
        expect(max(0, g_o - (r_o + b_o) / 2) < max(0, g_in - (r_in + b_in) / 2),
               f"green excess not reduced at {col}")

    img = purple_fringe()
    out = defringe(img)
    expect(np.allclose(out[32, :28], img[32, :28], atol=1e-5), "highlight side changed")
    expect(np.allclose(out[32, 37:], img[32, 37:], atol=1e-5), "shadow side changed")

    img = neutral_noise()
    expect(np.allclose(defringe(img, strength=1.0), img, atol=1e-5), "neutral image changed")

    img = purple_flower()
    out = defringe(img, strength=1.0)
    expect(np.allclose(out[12:52, 12:52], img[12:52, 12:52], atol=1e-5), "purple flower changed")

    img = micro_fringe()
    out = defringe(img, strength=1.0)
    expect(max(0, out[32, 32, 0] - out[32, 32, 1]) < max(0, img[32, 32, 0] - img[32, 32, 1]),
           "micro fringe not reduced")

    img = purple_fringe()
    out = defringe(img)
    expect(out.shape == img.shape and out.dtype == np.float32, "shape/dtype")
    out = defringe(img, strength=10.0)
    expect(float(out.min()) >= -1e-6 and float(out.max()) <= 1 + 1e-6, "range [0,1]")

    img = purple_fringe(width=4)
    out = defringe(img, strength=0.5)
    expect(float(np.abs(np.diff(np.abs(out[32, :, 0] - img[32, :, 0]))).max()) < 0.50, "mask aliasing")

    img = purple_fringe()
    r_full = defringe(img, strength=1.0)[32, 32, 0]
    r_half = defringe(img, strength=0.5)[32, 32, 0]
    expect(r_full <= r_half <= img[32, 32, 0] + 1e-4, "partial strength not monotone")

    img = broad_purple()
    out = defringe(img)
    changed = np.sum(np.abs(out[10:55, 10:55] - img[10:55, 10:55]) > 1e-5)
    expect(changed / (45 * 45) < 0.30, "broad purple changed")
    return fails
