"""HIP develop path vs the CPU restatement (oracle), through the C ABI.

Integer and AMaZE results must be bit-exact; the detail tone map (box sums
evaluated in a different order) must stay within max-abs 2e-4 and PSNR 70 dB.
"""
import numpy as np
import pytest

from _common import (RGGB, oracle_amaze, oracle_detail_tonemap, oracle_tone_mapping,
                     oracle_white_balance, place_native_u16, psnr)
from libraw_enhanced_amd import scene

pytestmark = pytest.mark.gpu

BGGR, GRBG, GBRG = 0x16161616, 0x61616161, 0x49494949


@pytest.fixture(scope="module")
def acc():
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    a = HIPAccelerator(0)
    yield a
    a.close()


def _bayer(h, w, seed, filters=RGGB):
    rgb = scene.render(h, w, seed)
    cfa = scene.mosaic(rgb, seed, filters)
    return scene.cfa_to_rgb3(cfa, filters)


@pytest.mark.parametrize("h,w", [(64, 80), (192, 256), (333, 517), (1024, 1536), (1000, 1500), (130, 700), (258, 141)])
def test_amaze_bit_exact(acc, h, w):
    print(f"amaze {h}x{w}", flush=True)
    raw3 = _bayer(h, w, scene.BASE_SEED + h)
    got = acc.demosaic_bayer_amaze(raw3, RGGB, (1, 1, 1, 1), 1.0)
    want = oracle_amaze(raw3)
    bad = np.count_nonzero(got != want)
    assert bad == 0, f"{bad} values differ, max-abs {np.abs(got - want).max()}"


# ---------------- micro-contrast ----------------
from test_micro_contrast_cpu import _oracle_mc  # noqa: E402


@pytest.mark.parametrize("kw", [dict(threshold=0.4), dict(threshold=0.7, strength=2.0, target=0.1),
                                dict(threshold=1.3, target=0.04, mask=True)])
def test_micro_contrast_bit_exact(acc, kw):
    print(f"micro contrast {kw}", flush=True)
    h, w = 480, 640
    img = scene.render(h, w, scene.BASE_SEED + 31) * 1.3
    mask = scene.render(h, w, scene.BASE_SEED + 32)[:, :, 1] if kw.get("mask") else None
    args = (kw["threshold"], kw.get("strength", 8.0), kw.get("target", 0.06))
    got = acc.enhance_micro_contrast(img, *args, mask=mask)
    want = _oracle_mc(img, *args, mask=mask)
    bad = np.count_nonzero(got != want)
    assert bad == 0, f"{bad} values differ, max-abs {np.abs(got - want).max()}"


@pytest.mark.parametrize("h,w", [(1, 1), (3, 5), (40, 3), (33, 47)])
def test_micro_contrast_tiny_frames(acc, h, w):
    print(f"micro contrast tiny {h}x{w}", flush=True)
    img = scene.render(h, w, scene.BASE_SEED + h * w) * 1.1
    assert np.array_equal(acc.enhance_micro_contrast(img, 0.2), _oracle_mc(img, 0.2))


def test_micro_contrast_standalone_wrapper(acc):
    from libraw_enhanced_amd.accelerator import LibRawWrapper
    img = scene.render(64, 96, scene.BASE_SEED + 33) * 1.5
    lw = LibRawWrapper(0)
    assert np.array_equal(lw.enhance_micro_contrast(img), _oracle_mc(img, 1.0, 8.0, 0.06))


def test_amaze_nyquist_patterns(acc):
    """Period-2 stripes, a checkerboard and a chirp up to 0.5 cycles/px: the
    content that AMaZE's Nyquist test flags, so its refinement paths run."""
    print("amaze nyquist patterns", flush=True)
    h, w = 400, 600
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    lum = np.empty((h, w), np.float32)
    q = w // 4
    lum[:, :q] = 0.5 + 0.4 * ((x[:, :q].astype(np.int32) & 1) * 2 - 1)
    lum[:, q:2 * q] = 0.5 + 0.4 * ((y[:, q:2 * q].astype(np.int32) & 1) * 2 - 1)
    lum[:, 2 * q:3 * q] = 0.5 + 0.4 * (((x + y)[:, 2 * q:3 * q].astype(np.int32) & 1) * 2 - 1)
    xx, yy = x[:, 3 * q:] - 3 * q, y[:, 3 * q:]
    lum[:, 3 * q:] = 0.5 + 0.4 * np.cos(np.pi * (xx * xx + yy * yy) / (2.0 * max(q, h)))
    tint = np.stack([0.9 + 0.1 * np.sin(y / 37.0), np.ones_like(y), 0.8 + 0.2 * np.cos(x / 53.0)], axis=2)
    rgb = np.ascontiguousarray(lum[:, :, None] * tint, np.float32)
    cfa = scene.mosaic(rgb, scene.BASE_SEED, RGGB, None, 0.0, 0.0)
    raw3 = scene.cfa_to_rgb3(cfa, RGGB)
    got = acc.demosaic_bayer_amaze(raw3, RGGB, (1, 1, 1, 1), 1.0)
    want = oracle_amaze(raw3)
    bad = np.count_nonzero(got != want)
    assert bad == 0, f"{bad} values differ, max-abs {np.abs(got - want).max()}"


@pytest.mark.parametrize("filters", [BGGR, GRBG, GBRG])
def test_amaze_patterns_and_gain(acc, filters):
    h, w = 300, 420
    raw3 = _bayer(h, w, scene.BASE_SEED + 7, filters) * 3.0
    cam_mul = (2.1, 1.0, 1.5, 1.0)  # clip_pt = 1/1.4
    got = acc.demosaic_bayer_amaze(raw3, filters, cam_mul, 3.0)
    want = oracle_amaze(raw3, filters, cam_mul, 3.0)
    assert np.count_nonzero(got != want) == 0


def test_tone_mapping_bit_exact(acc):
    rgb = scene.render(257, 263, scene.BASE_SEED) * 2.5
    got = acc.tone_mapping(rgb, 1.3)
    assert np.array_equal(got, oracle_tone_mapping(rgb, 1.3))


@pytest.mark.parametrize("h,w", [(1, 40), (37, 29), (480, 640)])
def test_detail_tonemap_tolerance(acc, h, w):
    rgb = scene.render(h, w, scene.BASE_SEED + w) * 1.6
    got = acc.apply_detail_preserving_tonemap(rgb)
    want = oracle_detail_tonemap(rgb)
    err = float(np.abs(got - want).max())
    assert err <= 2e-4, err
    assert psnr(got, want) >= 70.0


@pytest.mark.parametrize("h,w", [(1, 1), (5, 70), (130, 64), (300, 129), (481, 641), (1000, 300)])
def test_detail_tonemap_out_of_place(acc, h, w):
    """With a separate output the guide comes from the input pixels instead of
    a guide plane: the same bits as in place, within tolerance of the oracle
    (frames across strip and segment edges)."""
    from libraw_enhanced_amd.accelerator import native
    print(f"detail TM out of place {h}x{w}", flush=True)
    rgb = scene.render(h, w, scene.BASE_SEED + h + 3 * w) * 1.6
    lib = native()
    d_in = acc.dev_from_host(rgb)
    d_out = acc.dev_alloc(rgb.nbytes)
    assert lib.lre_hip_dev_detail_tonemap(acc.handle, d_in.ptr, d_out.ptr, w, h, None) == 0
    acc.synchronize()
    got = d_out.download((h, w, 3))
    assert np.array_equal(d_in.download((h, w, 3)), rgb)  # the input is left as it was
    inplace = acc.apply_detail_preserving_tonemap(rgb)
    assert np.array_equal(got, inplace), f"{np.count_nonzero(got != inplace)} values differ from in place"
    want = oracle_detail_tonemap(rgb)
    assert float(np.abs(got - want).max()) <= 2e-4
    assert psnr(got, want) >= 70.0


def test_white_balance_bit_exact(acc):
    h, w = 123, 211
    cfa = scene.bayer_scene(h, w, scene.BASE_SEED)
    u16 = scene.quantize_u16(cfa)
    img4 = place_native_u16(u16, RGGB)
    wb = (2.1, 1.0, 1.5, 1.0)
    got = acc.apply_white_balance(img4, wb, RGGB)
    assert np.array_equal(got, oracle_white_balance(img4, wb, RGGB))
    xcfa = scene.mosaic(scene.render(h, w, 5), 5, 0, scene.XTRANS)
    x16 = scene.quantize_u16(xcfa, 0, scene.XTRANS)
    ximg4 = place_native_u16(x16, 9, scene.XTRANS)
    got = acc.apply_white_balance(ximg4, wb, 9, scene.XTRANS)
    assert np.array_equal(got, oracle_white_balance(ximg4, wb, 9, scene.XTRANS))


# ---------------- defringe ----------------
import ctypes  # noqa: E402

import _defringe_cases as cases  # noqa: E402
from _common import oracle, p  # noqa: E402


def _oracle_defringe(img, inplace, radius=10.0, edge=0.1, chroma=0.15, strength=10.0, green=False, gscale=0.3):
    h, w, _ = img.shape
    a = np.ascontiguousarray(img, np.float32).copy()
    out = a if inplace else np.zeros_like(a)
    assert oracle().oracle_defringe(p(a), p(out), w, h, ctypes.c_float(radius), ctypes.c_float(edge),
                                    ctypes.c_float(chroma), ctypes.c_float(strength), int(green),
                                    ctypes.c_float(gscale)) == 0
    return out


def _gpu_defringe(acc, img, inplace, radius=10.0, edge=0.1, chroma=0.15, strength=10.0, green=False, gscale=0.3):
    a = np.ascontiguousarray(img, np.float32).copy()
    return acc.defringe(a, a if inplace else None, radius, edge, chroma, strength, green, gscale)


@pytest.mark.parametrize("inplace", [False, True])
# edge=1e-7 narrows a smoothstep ramp below 2^-20, which selects the kernels that divide;
# radius 20 (21 taps a side) takes the separable blur passes and the support pass of its own
@pytest.mark.parametrize("kw", [dict(), dict(green=True), dict(radius=6.0, strength=2.0, green=True, gscale=0.5),
                                dict(edge=1e-7, green=True), dict(radius=20.0, green=True)])
def test_defringe_tolerance(acc, inplace, kw):
    h, w = 480, 720
    img = oracle_amaze(scene.cfa_to_rgb3(scene.bayer_scene(h, w, scene.BASE_SEED + 21))) * 1.8
    want = _oracle_defringe(img, inplace, **kw)
    got = _gpu_defringe(acc, img, inplace, **kw)
    err = float(np.abs(got - want).max())
    ndiff = int(np.count_nonzero(np.any(got != want, axis=2)))
    print(f"defringe {kw} inplace={inplace}: max-abs {err:.3g}, {ndiff} px differ in the last bits")
    assert err <= 2e-4
    assert psnr(got, want) >= 70.0
    assert ndiff < 0.01 * h * w


def test_defringe_reference_invariants_on_gpu(acc):
    def f(img, radius=10.0, strength=10.0, defringe_green=False):
        return _gpu_defringe(acc, img, False, radius=radius, strength=strength, green=defringe_green)
    assert cases.check_all(f) == []


def test_develop_chain_matches_oracle(acc):
    """AMaZE -> defringe (in place) -> detail tone map, all resident on the device."""
    from libraw_enhanced_amd.pipeline import DevelopChain
    h, w = 600, 900
    cfa = scene.bayer_scene(h, w, scene.BASE_SEED + 5)
    d_cfa = acc.dev_from_host(cfa)
    chain = DevelopChain(acc, h, w, ["amaze", "defringe", "detail_tm"])
    chain.run(d_cfa)
    got = chain.result()
    want = oracle_amaze(scene.cfa_to_rgb3(cfa))
    want = _oracle_defringe(want, True)
    want = oracle_detail_tonemap(want)
    err = float(np.abs(got - want).max())
    assert err <= 2e-4, err
    assert psnr(got, want) >= 70.0


# ---------------- chromatic aberration ----------------
from test_ca_cpu import axial as _axial, ca_scene as _ca_scene, lateral as _lateral  # noqa: E402


@pytest.mark.parametrize("kw", [dict(), dict(radius=3, eps=1e-3, strength=1.0)])
def test_axial_ca_tolerance(acc, kw):
    img = _ca_scene(520, 780)
    want = _axial(oracle(), "oracle_axial_ca", img, **kw)
    got = acc.ca_axial_cleanup(img, kw.get("radius", 6), kw.get("eps", 1e-4), kw.get("strength", 0.3))
    err = float(np.abs(got - want).max())
    print(f"axial {kw}: max-abs {err:.3g}, {int(np.count_nonzero(got != want))} values differ")
    assert err <= 2e-4 and psnr(got, want) >= 70.0


@pytest.mark.parametrize("kw", [dict(), dict(cell=48, iters=5, shift=3.0, levels=2)])
def test_lateral_ca_tolerance(acc, kw):
    img = _ca_scene(800, 1200)
    want = _lateral(oracle(), "oracle_lateral_ca", img, **kw)
    got = acc.ca_register_lateral(img, kw.get("cell", 96), kw.get("iters", 3), kw.get("shift", 6.0),
                                  kw.get("conf", 0.02), kw.get("levels", 3))
    err = float(np.abs(got - want).max())
    print(f"lateral {kw}: max-abs {err:.3g}, {int(np.count_nonzero(got != want))} values differ")
    assert err <= 2e-4 and psnr(got, want) >= 70.0


# ---------------- X-Trans ----------------
from test_xtrans_cpu import oracle_xtrans as _oracle_xtrans, xtrans_raw3 as _xtrans_raw3  # noqa: E402


@pytest.mark.parametrize("h,w", [(120, 150), (300, 500), (640, 960), (1040, 1560), (777, 1001)])
def test_xtrans_bit_exact(acc, h, w):
    raw3 = _xtrans_raw3(h, w, scene.BASE_SEED + w)
    print(f"xtrans {h}x{w}", flush=True)
    got = acc.demosaic_xtrans_3pass(raw3, scene.XTRANS, 1.0)
    want = _oracle_xtrans(raw3)
    bad = np.any(got != want, axis=2)
    assert not bad.any(), f"{int(bad.sum())} px differ, max-abs {np.abs(got - want).max()}"


# ---------------- integer preprocessing ----------------
from test_ingest_cpu import black_level as _black_level, ingest as _ingest  # noqa: E402


def test_black_level_bit_exact(acc):
    raw = scene.quantize_u16(scene.bayer_scene(150, 210, 4))
    img4 = place_native_u16(raw, RGGB)
    for kw in (dict(black=512, maximum=16383), dict(cblack=(500, 510, 520, 530), maximum=16000),
               dict(cblack=(3, 0, 1, 0), pos=np.array([[1, 2, 3], [4, 5, 6]]), maximum=16383)):
        want = _black_level(img4, **kw)
        got = acc.apply_black_level_correction(img4.copy(), kw.get("cblack", (0, 0, 0, 0)), kw.get("black", 0),
                                               kw["maximum"], kw.get("pos"))
        assert np.array_equal(got, want), kw


@pytest.mark.parametrize("filters,xtrans", [(RGGB, None), (9, scene.XTRANS)])
def test_ingest_u16_bit_exact(acc, filters, xtrans):
    import ctypes as ct
    from libraw_enhanced_amd.accelerator import native
    h, w = 333, 517
    cfa = scene.bayer_scene(h, w, 9) if xtrans is None else scene.xtrans_scene(h, w, 9, lateral=None)
    raw = scene.quantize_u16(cfa, filters if xtrans is None else 0, xtrans)
    want, want_max = _ingest(raw, filters, xtrans)
    d_raw = acc.dev_from_host(raw)
    d_cfa = acc.dev_alloc(h * w * 4)
    d_max = acc.dev_alloc(4)
    cb = (ct.c_int * 4)(0, 0, 0, 0)
    wb = np.array([2.1, 1.0, 1.5, 1.0], np.float32)
    xt = None if xtrans is None else np.ascontiguousarray(xtrans.reshape(-1), np.int8)
    rc = native().lre_hip_dev_ingest_u16(acc.handle, d_raw.ptr, d_cfa.ptr, w, h, filters,
                                        None if xt is None else xt.ctypes.data_as(ct.POINTER(ct.c_int8)),
                                        cb, 512, 16383, wb.ctypes.data_as(ct.POINTER(ct.c_float)), d_max.ptr, None)
    assert rc == 0
    acc.synchronize()
    assert np.array_equal(d_cfa.download((h, w)), want)
    assert float(d_max.download((1,))[0]) == want_max


# ---------------- output epilogue ----------------
from test_epilogue_cpu import MATRICES, _img, oracle_color, oracle_gamma, oracle_quantize  # noqa: E402


def test_color_space_bit_exact(acc):
    img = _img(120, 170)
    for t in MATRICES:
        assert np.array_equal(acc.convert_color_space(img, t), oracle_color(img, t))


@pytest.mark.parametrize("cs,gp,gs", [(0, 2.222, 4.5), (1, 0, 0), (2, 0, 0), (4, 0, 0), (7, 0, 0), (8, 0, 0)])
def test_gamma_tolerance(acc, cs, gp, gs):
    img = _img(120, 170, seed=cs)
    got = acc.gamma_correct(img, gp, gs, cs)
    want = oracle_gamma(img, gp, gs, cs)
    # powf is evaluated in double on the GPU (correctly rounded); the C library's
    # powf may differ by one unit in the last place, which the sRGB / Rec.2020
    # offsets (a * x - (a - 1)) can widen to a few units near zero
    ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    err = float(np.abs(got - want).max())
    print(f"gamma cs={cs}: {int(np.count_nonzero(ulp))} of {ulp.size} values differ, "
          f"max {int(ulp.max())} ulp, max-abs {err:.3g}")
    assert int(ulp.max()) <= 4 and np.count_nonzero(ulp) <= ulp.size // 100
    assert err <= 1e-6


def test_quantize_and_fused_epilogue(acc):
    import ctypes as ct
    from libraw_enhanced_amd.accelerator import native
    img = _img(97, 131)
    h, w, _ = img.shape
    t = np.ascontiguousarray(MATRICES[1]).reshape(12)
    for bits in (8, 16):
        assert np.array_equal(acc.quantize(img, bits), oracle_quantize(img, bits))
    # fused (matrix -> gamma -> quantize) equals the separate GPU stages
    staged = acc.gamma_correct(acc.convert_color_space(img, MATRICES[1]), 0, 0, 1)
    d_in = acc.dev_from_host(img)
    for bits, dt in ((32, np.float32), (8, np.uint8), (16, np.uint16)):
        d_out = acc.dev_alloc(h * w * 3 * np.dtype(dt).itemsize)
        rc = native().lre_hip_dev_epilogue(acc.handle, d_in.ptr, d_out.ptr, w, h,
                                          t.ctypes.data_as(ct.POINTER(ct.c_float)), 0.0, 0.0, 1, bits, None)
        assert rc == 0
        acc.synchronize()
        got = d_out.download((h, w, 3), dt)
        want = staged if bits == 32 else acc.quantize(staged, bits)
        assert np.array_equal(got, want), bits


# ---------------- green matching, zero removal ----------------
from test_ingest_cpu import (FOUR_COLOR, _native_channels, oracle_green_matching,  # noqa: E402
                             oracle_remove_zeroes)


def _img4_from(vals, filters):
    h, w = vals.shape
    ch = _native_channels(h, w, filters)
    img = np.zeros((h, w, 4), np.uint16)
    for k in range(4):
        img[..., k][ch == k] = vals[ch == k]
    return img


def test_green_matching_bit_exact(acc):
    rng = np.random.Generator(np.random.PCG64(11))
    for h, w in ((64, 90), (301, 257)):
        base = 1000 + rng.integers(-12, 13, (h, w))          # flat with a little noise
        base[h // 3: h // 2, w // 4: w // 2] = rng.integers(200, 4000, (h // 2 - h // 3, w // 2 - w // 4))
        base[5, 6] = 4090                                     # saturated region
        img = _img4_from(base.astype(np.uint16), FOUR_COLOR)
        want = oracle_green_matching(img, FOUR_COLOR, 4095)
        changed = int(np.count_nonzero(want != img))
        got = acc.apply_green_matching(img.copy(), FOUR_COLOR, 4095)
        print(f"green matching {h}x{w}: {changed} samples corrected")
        assert changed > 0 and np.array_equal(got, want)


@pytest.mark.parametrize("density", [0.002, 0.05, 0.6])
def test_remove_zeroes_bit_exact(acc, density):
    rng = np.random.Generator(np.random.PCG64(int(density * 1000)))
    h, w = 203, 311
    vals = rng.integers(1, 16000, (h, w)).astype(np.uint16)
    vals[rng.random((h, w)) < density] = 0                 # dead pixels, clusters, mostly dead
    vals[100:110, 50:70] = 0                               # a block: long raster-order chains
    for filters in (RGGB, FOUR_COLOR):
        img = _img4_from(vals, filters)
        want = oracle_remove_zeroes(img, filters)
        got = acc.remove_zeroes(img.copy(), filters)
        print(f"remove_zeroes density {density} filters {filters:#x}: "
              f"{int(np.count_nonzero(want != img))} samples replaced")
        assert np.array_equal(got, want)


from test_ingest_cpu import oracle_auto_wb, oracle_scale_aber  # noqa: E402


def test_auto_wb_bit_exact(acc):
    rng = np.random.Generator(np.random.PCG64(21))
    for filters, xt in ((RGGB, None), (FOUR_COLOR, None), (9, scene.XTRANS)):
        vals = rng.integers(0, 16384, (517, 333)).astype(np.uint16)
        if xt is None:
            img = _img4_from(vals, filters)
        else:
            cx = np.asarray(xt)[np.arange(517)[:, None] % 6, np.arange(333)[None, :] % 6]
            img = np.zeros((517, 333, 4), np.uint16)
            for k in range(3):
                img[..., k][cx == k] = vals[cx == k]
        got = acc.auto_wb(img, filters, xt)
        want = oracle_auto_wb(img, filters, xt)
        assert np.array_equal(got, want), (filters, got, want)


@pytest.mark.parametrize("aber", [(1.0008, 1, 0.9993, 1), (1.05, 1, 0.95, 1), (0.9, 1, 1.0, 1)])
def test_scale_aber_bit_exact(acc, aber):
    rng = np.random.Generator(np.random.PCG64(5))
    img = rng.integers(0, 65536, (301, 411, 4)).astype(np.uint16)
    want = oracle_scale_aber(img, aber)
    got = acc.scale_aber(img.copy(), aber)
    assert np.array_equal(got, want)
    assert np.count_nonzero(want != img) > 0


# ---------------- tiny and ragged frames ----------------
# Frames smaller than a tile, one pixel wide or tall, and odd sizes: the tile
# walks, the border passes and the clamped window loads must still match.

@pytest.mark.parametrize("h,w", [(1, 1), (2, 3), (5, 7), (17, 9), (31, 33), (1, 200), (200, 2)])
def test_amaze_tiny_frames(acc, h, w):
    print(f"amaze tiny {h}x{w}", flush=True)
    raw3 = _bayer(h, w, scene.BASE_SEED + 3 * h + w)
    got = acc.demosaic_bayer_amaze(raw3, RGGB, (1, 1, 1, 1), 1.0)
    want = oracle_amaze(raw3)
    assert np.array_equal(got, want), f"max-abs {np.abs(got - want).max()}"


@pytest.mark.parametrize("h,w", [(1, 1), (6, 6), (21, 30), (23, 47), (40, 200), (200, 23)])
def test_xtrans_tiny_frames(acc, h, w):
    print(f"xtrans tiny {h}x{w}", flush=True)
    raw3 = _xtrans_raw3(h, w, scene.BASE_SEED + 5 * h + w)
    got = acc.demosaic_xtrans_3pass(raw3, scene.XTRANS, 1.0)
    want = _oracle_xtrans(raw3)
    assert np.array_equal(got, want), f"max-abs {np.abs(got - want).max()}"


@pytest.mark.parametrize("h,w", [(1, 1), (3, 5), (1, 64), (40, 3), (33, 47)])
def test_rgb_stages_tiny_frames(acc, h, w):
    print(f"rgb stages tiny {h}x{w}", flush=True)
    img = _ca_scene(h, w)
    checks = [
        (_gpu_defringe(acc, img, False), _oracle_defringe(img, False)),
        (_gpu_defringe(acc, img, True), _oracle_defringe(img, True)),
        (acc.ca_axial_cleanup(img, 6, 1e-4, 0.3), _axial(oracle(), "oracle_axial_ca", img)),
        (acc.ca_register_lateral(img, 96, 3, 6.0, 0.02, 3), _lateral(oracle(), "oracle_lateral_ca", img)),
        (acc.apply_detail_preserving_tonemap(img), oracle_detail_tonemap(img)),
    ]
    for k, (got, want) in enumerate(checks):
        err = float(np.abs(got - want).max())
        assert err <= 2e-4, (k, err)
