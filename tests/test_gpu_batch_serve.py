"""Batched develop and request loop on the GPU, through the C ABI.

Every image of a mixed batch (Bayer and X-Trans, float and uint16 input,
several sizes) must come out bit-identical to the same stages run one image at
a time through the per-image entry points; a captured graph must reproduce the
eager batch on every replay; a served request must give the same bytes alone
and in a mix of concurrent requests; one batch image is also checked against
the oracle chain within max-abs 2e-4 and PSNR 70 dB.
"""
import ctypes as ct

import numpy as np
import pytest

from _common import RGGB, oracle, oracle_amaze, oracle_detail_tonemap, psnr
from libraw_enhanced_amd import scene

pytestmark = pytest.mark.gpu

WB = (2.1, 1.0, 1.5, 1.0)
BLACK, WHITE = 512, 16383


@pytest.fixture(scope="module")
def acc():
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    a = HIPAccelerator(0)
    yield a
    a.close()


def _spec():
    # (kind, h, w, filters, input)
    return [("bayer", 240, 320, RGGB, "f32"), ("bayer", 300, 420, 0x16161616, "f32"),
            ("xtrans", 240, 330, 9, "f32"), ("bayer", 256, 384, RGGB, "u16"),
            ("xtrans", 252, 330, 9, "u16"), ("bayer", 130, 700, RGGB, "f32")]


def _inputs(k, kind, h, w, filters, inp):
    seed = scene.BASE_SEED + 100 + k
    if kind == "bayer":
        cfa = scene.mosaic(scene.render(h, w, seed), seed, filters)
        xt = None
    else:
        cfa = scene.xtrans_scene(h, w, seed, lateral=None)
        xt = scene.XTRANS
    if inp == "u16":
        return scene.quantize_u16(cfa, filters if xt is None else 0, xt), xt
    return np.ascontiguousarray(cfa, np.float32), xt


class _Set:
    """Device buffers of the mixed batch."""

    def __init__(self, acc, spec=None):
        from libraw_enhanced_amd import develop as dv
        self.items = []
        for k, (kind, h, w, filters, inp) in enumerate(spec or _spec()):
            data, xt = _inputs(k, kind, h, w, filters, inp)
            d_in = acc.dev_from_host(data)
            d_rgb = acc.dev_alloc(h * w * 12)
            if inp == "u16":
                im = dv.make_image(w, h, filters=filters, xtrans=xt, cam_mul=WB, d_raw=d_in,
                                   black=BLACK, white=WHITE, wb=WB, d_rgb=d_rgb)
            else:
                im = dv.make_image(w, h, filters=filters, xtrans=xt, d_cfa=d_in, maximum_value=1.0,
                                   d_rgb=d_rgb)
            self.items.append(dict(kind=kind, h=h, w=w, filters=filters, inp=inp, data=data, xt=xt,
                                   d_in=d_in, d_rgb=d_rgb, im=im))

    def images(self):
        return [it["im"] for it in self.items]

    def results(self):
        return [it["d_rgb"].download((it["h"], it["w"], 3)) for it in self.items]


def _single(acc, it, stages, p):
    """The same stages one image at a time through the per-image entry points."""
    from libraw_enhanced_amd.accelerator import native
    lib, ctx = native(), acc.handle
    h, w, filters = it["h"], it["w"], it["filters"]
    d_rgb = acc.dev_alloc(h * w * 12)
    xt = None if it["xt"] is None else np.ascontiguousarray(it["xt"].reshape(-1), np.int8)
    xtp = None if xt is None else xt.ctypes.data_as(ct.POINTER(ct.c_int8))
    if it["inp"] == "u16":
        d_cfa = acc.dev_alloc(h * w * 4)
        d_max = acc.dev_alloc(4)
        wb = np.array(it.get("wb", WB), np.float32)
        # an input already black-subtracted (preprocessed) is ingested with black 0
        d_raw = it.get("d_pre", it["d_in"])
        black_in = 0 if "d_pre" in it else BLACK
        cb = (ct.c_int * 4)(*((0, 0, 0, 0) if "d_pre" in it else it.get("cblack", (0, 0, 0, 0))))
        assert lib.lre_hip_dev_ingest_u16(ctx, d_raw.ptr, d_cfa.ptr, w, h, filters, xtp, cb, black_in,
                                          WHITE, wb.ctypes.data_as(ct.POINTER(ct.c_float)), d_max.ptr,
                                          None) == 0
        acc.synchronize()
        dmax = float(d_max.download((1,))[0])
        maxv = float(np.float32(WHITE - BLACK))
        if 0 < dmax < maxv and dmax > maxv * np.float32(0.75):
            maxv = dmax
        cam = np.array(WB, np.float32)
    else:
        d_cfa, maxv, cam = it["d_in"], 1.0, np.ones(4, np.float32)
    d_mask = None
    for s in stages:
        if s == "demosaic":
            if filters == 9:
                rc = lib.lre_hip_dev_xtrans3(ctx, d_cfa.ptr, d_rgb.ptr, w, h, xtp, maxv, None)
            else:
                rc = lib.lre_hip_dev_amaze(ctx, d_cfa.ptr, d_rgb.ptr, w, h, filters,
                                           cam.ctypes.data_as(ct.POINTER(ct.c_float)), maxv, None)
        elif s == "axial_ca":
            rc = lib.lre_hip_dev_axial_ca(ctx, d_rgb.ptr, d_rgb.ptr, w, h, p.axial_radius,
                                          p.axial_epsilon, p.axial_strength, None)
        elif s == "lateral_ca":
            rc = lib.lre_hip_dev_lateral_ca(ctx, d_rgb.ptr, d_rgb.ptr, w, h, p.lateral_cell,
                                            p.lateral_iterations, p.lateral_max_shift,
                                            p.lateral_min_confidence, p.lateral_levels, None)
        elif s == "defringe":
            rc = lib.lre_hip_dev_defringe(ctx, d_rgb.ptr, d_rgb.ptr, w, h, p.defringe_radius,
                                          p.defringe_edge_threshold, p.defringe_chroma_threshold,
                                          p.defringe_strength, p.defringe_green, p.defringe_green_scale,
                                          None)
        elif s == "detail_tm":
            rc = lib.lre_hip_dev_detail_tonemap(ctx, d_rgb.ptr, d_rgb.ptr, w, h, None)
        elif s == "tone_mapping":
            rc = lib.lre_hip_dev_tone_mapping(ctx, d_rgb.ptr, d_rgb.ptr, w, h, p.tone_after_scale, None)
        elif s == "color_space":
            t = np.ctypeslib.as_array(p.color_transform).astype(np.float32)
            rc = lib.lre_hip_dev_convert_color_space(ctx, d_rgb.ptr, d_rgb.ptr, w, h,
                                                     t.ctypes.data_as(ct.POINTER(ct.c_float)), None)
        elif s == "gamma":
            rc = lib.lre_hip_dev_gamma_correct(ctx, d_rgb.ptr, d_rgb.ptr, w, h, p.gamma_power,
                                               p.gamma_slope, p.output_color_space, None)
        elif s == "highlight":
            d_mask = acc.dev_alloc(h * w * 4)
            wb3 = np.array(WB[:3], np.float32)
            wbp = wb3.ctypes.data_as(ct.POINTER(ct.c_float)) if it["inp"] == "u16" else None
            rc = lib.lre_hip_dev_recover_highlights(ctx, d_rgb.ptr, d_rgb.ptr, w, h,
                                                    ct.c_float(p.highlight_threshold), d_mask.ptr, wbp, None,
                                                    None)
        elif s == "micro_contrast":
            thr = np.float32(p.highlight_threshold) + np.float32(0.3)  # float, as the runtime adds it
            rc = lib.lre_hip_dev_micro_contrast(ctx, d_rgb.ptr, d_rgb.ptr, w, h, ct.c_float(thr),
                                                ct.c_float(p.micro_strength), ct.c_float(p.micro_target_contrast),
                                                d_mask.ptr if d_mask is not None else None, None)
        assert rc == 0, s
    acc.synchronize()
    return d_rgb.download((h, w, 3))


CHAIN = ("demosaic", "axial_ca", "defringe", "detail_tm")


@pytest.mark.parametrize("lanes", [1, 3])
def test_batch_matches_per_image(acc, lanes):
    from libraw_enhanced_amd import develop as dv
    print(f"batch lanes={lanes}", flush=True)
    S = _Set(acc)
    p = dv.ChainParams.defaults()
    b = dv.BatchDeveloper(0, lanes)
    b.run(S.images(), CHAIN, p)
    b.synchronize()
    got = S.results()
    for it, g in zip(S.items, got):
        want = _single(acc, it, CHAIN, p)
        bad = int(np.count_nonzero(g != want))
        assert bad == 0, f"{it['kind']} {it['h']}x{it['w']} {it['inp']}: {bad} values differ"
    b.close()


def test_batch_and_serve_ingest_levels(acc):
    """Bayer uint16 input through the batch and the request loop: per-colour
    black levels, a fourth colour (second green) with its own multiplier, and
    samples below black and above white, bit-identical to the per-image ingest
    and AMaZE entry points."""
    from libraw_enhanced_amd import develop as dv
    from test_ingest_cpu import FOUR_COLOR
    wb4 = (2.1, 1.0, 1.5, 0.95)
    cases = [(RGGB, (0, 0, 0, 0), 250, 330), (RGGB, (3, 0, 7, 0), 161, 517),
             (FOUR_COLOR, (2, 5, 1, 9), 210, 300), (0x16161616, (4, 0, 0, 6), 97, 131)]
    rng = np.random.default_rng(scene.BASE_SEED + 7)
    items = []
    for k, (filters, cblack, h, w) in enumerate(cases):
        seed = scene.BASE_SEED + 200 + k
        raw = scene.quantize_u16(scene.mosaic(scene.render(h, w, seed), seed, filters), filters)
        flat = raw.reshape(-1)
        pick = rng.choice(flat.size, flat.size // 50, replace=False)
        flat[pick[: pick.size // 2]] = rng.integers(0, BLACK, pick.size // 2, dtype=np.uint16)
        flat[pick[pick.size // 2:]] = rng.integers(WHITE - 40, 65536, pick.size - pick.size // 2, dtype=np.uint16)
        print(f"ingest levels filters={filters:#x} cblack={cblack} {h}x{w}", flush=True)
        d_in, d_rgb = acc.dev_from_host(raw), acc.dev_alloc(h * w * 12)
        im = dv.make_image(w, h, filters=filters, cam_mul=WB, d_raw=d_in, black=BLACK, white=WHITE,
                           cblack=cblack, wb=wb4, d_rgb=d_rgb)
        items.append(dict(kind="bayer", h=h, w=w, filters=filters, inp="u16", data=raw, xt=None, d_in=d_in,
                          d_rgb=d_rgb, im=im, cblack=cblack, wb=wb4))
    b = dv.BatchDeveloper(0, 2)
    b.run([it["im"] for it in items], ("demosaic",))
    b.synchronize()
    server = dv.DevelopServer(0, lanes=2, hbm_budget_bytes=1 << 30, stages=("demosaic",))
    outs = [np.full((it["h"], it["w"], 3), np.nan, np.float32) for it in items]
    for k, it in enumerate(items):
        server.submit(k, it["im"], it["data"].ctypes.data, outs[k].ctypes.data, keep=(it["data"], outs[k]))
    done = 0
    while done < len(items):
        for c in server.poll(timeout_ms=1000):
            assert c["status"] == 0, c
            done += 1
    server.close()
    for k, it in enumerate(items):
        want = _single(acc, it, ("demosaic",), None)
        got = it["d_rgb"].download((it["h"], it["w"], 3))
        assert np.array_equal(got, want), f"case {k}: batch differs from the per-image path"
        assert np.array_equal(outs[k], want), f"case {k}: served result differs from the per-image path"
    b.close()


@pytest.mark.parametrize("group,budget", [(False, 0), (True, 0), (True, 1), ("auto", 0)])
def test_batch_grouping_modes(acc, group, budget):
    """Group mode (one launch per stage over a group), with the default budget
    (one group) and a 1-byte budget (one image per group), and per-image lanes:
    all bit-identical to the per-image entry points, also for a stage list
    without the demosaic run over images already in RGB."""
    from libraw_enhanced_amd import develop as dv
    print(f"grouping group={group} budget={budget}", flush=True)
    p = dv.ChainParams.defaults(defringe_green=1)
    S = _Set(acc)
    b = dv.BatchDeveloper(0, 2)
    assert b.grouping() == {"group_mode": "auto", "post_budget_bytes": 0}
    b.set_grouping(group, budget)
    assert b.grouping() == {"group_mode": group, "post_budget_bytes": budget}
    stages = ("demosaic", "axial_ca", "defringe", "detail_tm", "color_space", "gamma")
    b.run(S.images(), stages, p)
    b.synchronize()
    got = S.results()
    for it, g in zip(S.items, got):
        want = _single(acc, it, stages, p)
        bad = int(np.count_nonzero(g != want))
        assert bad == 0, f"{it['kind']} {it['h']}x{it['w']} {it['inp']}: {bad} values differ"
    # post stages only, in place over the demosaicked images
    post = ("defringe", "axial_ca", "tone_mapping", "detail_tm")
    b.run(S.images(), ("demosaic",), p)
    b.synchronize()
    base = S.results()
    b.run(S.images(), post, p)
    b.synchronize()
    from libraw_enhanced_amd.accelerator import native
    for it, x, g in zip(S.items, base, S.results()):
        h, w = it["h"], it["w"]
        d = acc.dev_from_host(np.ascontiguousarray(x))
        lib, ctx = native(), acc.handle
        for st in post:
            if st == "defringe":
                rc = lib.lre_hip_dev_defringe(ctx, d.ptr, d.ptr, w, h, p.defringe_radius,
                                              p.defringe_edge_threshold, p.defringe_chroma_threshold,
                                              p.defringe_strength, p.defringe_green, p.defringe_green_scale, None)
            elif st == "axial_ca":
                rc = lib.lre_hip_dev_axial_ca(ctx, d.ptr, d.ptr, w, h, p.axial_radius, p.axial_epsilon,
                                              p.axial_strength, None)
            elif st == "tone_mapping":
                rc = lib.lre_hip_dev_tone_mapping(ctx, d.ptr, d.ptr, w, h, p.tone_after_scale, None)
            else:
                rc = lib.lre_hip_dev_detail_tonemap(ctx, d.ptr, d.ptr, w, h, None)
            assert rc == 0, st
        acc.synchronize()
        want = d.download((h, w, 3))
        assert np.array_equal(g, want), (post, it["kind"], it["h"], it["w"])
    # a graph of the grouped batch replays to the same bits
    b.capture(S.images(), stages, p)
    for it in S.items:
        it["d_rgb"].upload(np.full((it["h"], it["w"], 3), np.nan, np.float32))
    acc.synchronize()
    b.replay()
    b.synchronize()
    for e, g in zip(got, S.results()):
        assert np.array_equal(e, g)
    b.close()


@pytest.mark.parametrize("group", [False, True])
def test_batch_defringe_blur_tiles(acc, group):
    """The defringe blur tile height (48 or 96 rows, or chosen by size) leaves
    the bits unchanged, in lane and in group mode."""
    from libraw_enhanced_amd import develop as dv
    p = dv.ChainParams.defaults(defringe_green=1)
    stages = ("demosaic", "defringe", "detail_tm")
    S = _Set(acc, _spec()[:3])
    want = [_single(acc, it, stages, p) for it in S.items]
    b = dv.BatchDeveloper(0, 2)
    b.set_grouping(group)
    for rows in (48, 96, 0):
        print(f"blur rows {rows} group={group}", flush=True)
        b.set_tiles(rows)
        assert b.tiles() == {"defringe_blur_rows": rows}
        b.run(S.images(), stages, p)
        b.synchronize()
        for it, g, w in zip(S.items, S.results(), want):
            assert np.array_equal(g, w), (rows, it["kind"])
    from libraw_enhanced_amd.accelerator import LibRawInvalidArgument
    with pytest.raises(LibRawInvalidArgument):
        b.set_tiles(64)
    b.close()


FOUR = 0xB4B4B4B4  # R G / G2 B: the pattern green matching acts on


def _preprocessed(raw, filters, zero, green):
    """The reference's order on its four-channel raw image, by the oracle
    (core/libraw_wrapper.cpp:1496-1547): zero removal, black subtraction and
    clip, green matching; the sites' own channel back as a plane."""
    from test_ingest_cpu import black_level, oracle_green_matching, oracle_remove_zeroes
    h, w = raw.shape
    r, c = np.arange(h)[:, None], np.arange(w)[None, :]
    ch = ((filters >> ((((r << 1) & 14) | (c & 1)) << 1)) & 3).astype(np.int64)
    img4 = np.zeros((h, w, 4), np.uint16)
    np.put_along_axis(img4, ch[..., None], raw[..., None], axis=2)
    if zero:
        img4 = oracle_remove_zeroes(img4, filters)
    img4 = black_level(img4, black=BLACK, maximum=WHITE)
    if green:
        img4 = oracle_green_matching(img4, filters, WHITE)
    return np.ascontiguousarray(np.take_along_axis(img4, ch[..., None], axis=2)[..., 0])


def _with_zeros(raw, seed):
    """Dead sites: single zeros, and runs of zeros of one channel close enough
    that a later one's window holds an earlier one (raster-order refill)."""
    out = raw.copy()
    rng = np.random.Generator(np.random.PCG64(seed))
    h, w = out.shape
    for _ in range(40):
        out[rng.integers(0, h), rng.integers(0, w)] = 0
    for _ in range(6):
        y, x = int(rng.integers(4, h - 8)), int(rng.integers(4, w - 8))
        out[y, x:x + 6:2] = 0
        out[y + 2, x:x + 6:2] = 0
    return out


@pytest.mark.parametrize("stages", [("remove_zeroes", "demosaic", "axial_ca", "defringe", "detail_tm"),
                                    ("green_matching", "demosaic", "defringe", "detail_tm"),
                                    ("remove_zeroes", "green_matching", "demosaic", "detail_tm")])
def test_batch_and_serve_raw_preprocessing(acc, stages):
    """Zero removal and green matching ahead of the demosaic, in the batch (lane
    and group mode, graph replay) and the request loop, against the oracle's
    preprocessing on the four-channel image followed by the per-image entry
    points: bit-identical.  X-Trans images skip both stages."""
    from libraw_enhanced_amd import develop as dv
    print(f"preprocessing {stages}", flush=True)
    spec = [("bayer", 240, 320, FOUR, "u16"), ("bayer", 256, 384, RGGB, "u16"), ("xtrans", 252, 330, 9, "u16"),
            ("bayer", 200, 260, FOUR, "f32")]
    S = _Set(acc, spec)
    zero, green = "remove_zeroes" in stages, "green_matching" in stages
    for k, it in enumerate(S.items):
        if it["inp"] != "u16":
            continue
        raw = _with_zeros(it["data"], 500 + k) if zero else it["data"]
        it["d_in"].upload(raw)
        it["raw"] = raw
        if it["filters"] != 9:  # X-Trans images skip both stages
            it["d_pre"] = acc.dev_from_host(_preprocessed(raw, it["filters"], zero, green))
    post = tuple(s for s in stages if s not in ("remove_zeroes", "green_matching"))
    want = [_single(acc, it, post, dv.ChainParams.defaults()) for it in S.items]
    # the stages act on these images (the comparison is not vacuous)
    four = S.items[0]
    if zero:
        assert np.count_nonzero(four["raw"] == 0) > 0
    if green:
        assert not np.array_equal(_preprocessed(four["raw"], FOUR, zero, True),
                                  _preprocessed(four["raw"], FOUR, zero, False))
    b = dv.BatchDeveloper(0, 2)
    for group in (False, True):
        b.set_grouping(group)
        b.run(S.images(), stages)
        b.synchronize()
        for it, g, w in zip(S.items, S.results(), want):
            assert np.array_equal(g, w), (stages, group, it["kind"], it["filters"], it["inp"])
    b.capture(S.images(), stages)
    for it in S.items:
        it["d_rgb"].upload(np.full((it["h"], it["w"], 3), np.nan, np.float32))
    acc.synchronize()
    b.replay()
    b.synchronize()
    for it, g, w in zip(S.items, S.results(), want):
        assert np.array_equal(g, w), (stages, "replay", it["kind"])
    b.close()
    # the request loop (uint16 requests only)
    server = dv.DevelopServer(0, lanes=2, hbm_budget_bytes=8 << 30, stages=stages)
    outs = {}
    for k, it in enumerate(S.items):
        if it["inp"] != "u16":
            continue
        im = dv.make_image(it["w"], it["h"], filters=it["filters"], xtrans=it["xt"], cam_mul=WB, black=BLACK,
                           white=WHITE, wb=WB)
        src = np.ascontiguousarray(it["raw"])
        out = np.zeros((it["h"], it["w"], 3), np.float32)
        outs[k] = (src, out)
        server.submit(k, im, src.ctypes.data, out.ctypes.data)
    done = 0
    while done < len(outs):
        for c in server.poll(timeout_ms=2000):
            assert c["status"] == 0, c
            done += 1
    server.close()
    for k, (_, out) in outs.items():
        assert np.array_equal(out, want[k]), (stages, "served", S.items[k]["kind"])


@pytest.mark.parametrize("kw", [dict(defringe_edge_threshold=1e-7), dict(defringe_radius=20.0)])
def test_batch_defringe_without_group_pass(acc, kw):
    """Defringe parameters outside the multi-image pass's range (a ramp width
    below 2^-20; a blur of more than 16 taps) fall back to per-image lanes,
    same bits."""
    from libraw_enhanced_amd import develop as dv
    print(f"defringe fallback {kw}", flush=True)
    p = dv.ChainParams.defaults(**kw)
    S = _Set(acc, _spec()[:2])
    b = dv.BatchDeveloper(0, 2)
    stages = ("demosaic", "defringe", "detail_tm")
    b.run(S.images(), stages, p)
    b.synchronize()
    for it, g in zip(S.items, S.results()):
        assert np.array_equal(g, _single(acc, it, stages, p)), it["kind"]
    b.close()


def test_batch_other_stage_lists(acc):
    from libraw_enhanced_amd import develop as dv
    p = dv.ChainParams.defaults(defringe_green=1, axial_radius=3, output_color_space=2)
    p.color_transform[:] = [1.62, -0.49, -0.13, 0.0, -0.18, 1.35, -0.17, 0.0, 0.02, -0.31, 1.29, 0.0]
    spec = _spec()[:3]
    # the detail tone map followed by the matrix and/or gamma runs them inside
    # its store, and the matrix followed by gamma shares one pass: same bits
    for stages in [("demosaic",), ("demosaic", "lateral_ca", "tone_mapping"),
                   ("demosaic", "defringe", "axial_ca", "detail_tm"),
                   ("demosaic", "detail_tm", "color_space", "gamma"),
                   ("demosaic", "detail_tm", "gamma"), ("demosaic", "detail_tm", "color_space"),
                   ("demosaic", "color_space", "gamma"), ("demosaic", "gamma", "color_space")]:
        print(f"stages {stages}", flush=True)
        S = _Set(acc, spec)
        b = dv.BatchDeveloper(0, 2)
        b.run(S.images(), stages, p)
        b.synchronize()
        for it, g in zip(S.items, S.results()):
            assert np.array_equal(g, _single(acc, it, stages, p)), (stages, it["kind"])
        b.close()


def test_fused_epilogue_srgb(acc):
    from libraw_enhanced_amd import develop as dv
    p = dv.ChainParams.defaults(output_color_space=1)
    p.color_transform[:] = [1.2, -0.1, -0.1, 0.0, -0.05, 1.1, -0.05, 0.0, 0.0, -0.2, 1.2, 0.0]
    stages = ("demosaic", "axial_ca", "detail_tm", "color_space", "gamma")
    S = _Set(acc, _spec()[3:5])
    b = dv.BatchDeveloper(0, 1)
    b.run(S.images(), stages, p)
    b.synchronize()
    for it, g in zip(S.items, S.results()):
        assert np.array_equal(g, _single(acc, it, stages, p)), it["kind"]
    b.close()


def test_batch_highlight_stages(acc):
    """postprocess's highlight_mode order on the device: recovery (keeping its
    mask), tone map, micro-contrast blended by that mask; and micro-contrast
    alone, gated on green."""
    from libraw_enhanced_amd import develop as dv
    p = dv.ChainParams.defaults(highlight_threshold=0.8)
    for stages in [("demosaic", "highlight", "detail_tm", "micro_contrast"),
                   ("demosaic", "highlight", "tone_mapping", "micro_contrast"),
                   ("demosaic", "micro_contrast"), ("demosaic", "highlight")]:
        print(f"stages {stages}", flush=True)
        S = _Set(acc, _spec()[:5])
        b = dv.BatchDeveloper(0, 2)
        b.run(S.images(), stages, p)
        b.synchronize()
        for it, g in zip(S.items, S.results()):
            want = _single(acc, it, stages, p)
            bad = int(np.count_nonzero(g != want))
            assert bad == 0, f"{stages} {it['kind']} {it['inp']}: {bad} values differ"
        b.close()


def test_batch_graph_replays(acc):
    from libraw_enhanced_amd import develop as dv
    S = _Set(acc)
    b = dv.BatchDeveloper(0, 2)
    b.run(S.images(), CHAIN)
    b.synchronize()
    eager = S.results()
    b.capture(S.images(), CHAIN)
    for rep in range(3):
        print(f"replay {rep}", flush=True)
        for it in S.items:  # clobber the outputs between replays
            it["d_rgb"].upload(np.full((it["h"], it["w"], 3), np.nan, np.float32))
        acc.synchronize()
        b.replay()
        b.synchronize()
        for e, g in zip(eager, S.results()):
            assert np.array_equal(e, g), rep
        # eager work between replays
        b.run(S.images()[:2], ("demosaic",))
        b.synchronize()
    b.close()


def test_batch_autotune_and_config(acc):
    from libraw_enhanced_amd import develop as dv
    S = _Set(acc, _spec()[:3])
    b = dv.BatchDeveloper(0, 2)
    ms = b.autotune(S.images(), CHAIN)
    cfg = b.config()
    assert ms > 0 and 1 <= cfg["lanes"] <= 4 and cfg["amaze_slots_per_cu"] in (1, 2, 4)
    assert b.tiles()["defringe_blur_rows"] in (0, 48, 96)
    from libraw_enhanced_amd.accelerator import LibRawInvalidArgument
    with pytest.raises(LibRawInvalidArgument):
        b.set_config(0, 8, 6)
    with pytest.raises(LibRawInvalidArgument):
        b.run(S.images(), ("axial_ca", "demosaic"))
    # raw preprocessing: before the demosaic, zero removal first, once each
    for bad in [("green_matching", "remove_zeroes", "demosaic"), ("remove_zeroes",),
                ("remove_zeroes", "axial_ca"), ("demosaic", "green_matching"),
                ("remove_zeroes", "remove_zeroes", "demosaic")]:
        with pytest.raises(LibRawInvalidArgument):
            b.run(S.images(), bad)
    b.close()


def test_batch_image_matches_oracle_chain(acc):
    """u16 Bayer input: ingest + adjust_maximum + AMaZE + axial + defringe + detail TM."""
    from libraw_enhanced_amd import develop as dv
    from test_ca_cpu import axial as _axial
    from test_gpu_develop import _oracle_defringe
    from test_ingest_cpu import ingest as _ingest
    spec = [("bayer", 256, 384, RGGB, "u16")]
    S = _Set(acc, spec)
    b = dv.BatchDeveloper(0, 1)
    b.run(S.images(), CHAIN)
    b.synchronize()
    got = S.results()[0]
    it = S.items[0]
    cfa, dmax = _ingest(it["data"], RGGB, None, black=BLACK, maximum=WHITE, wb=WB)
    o = oracle()
    o.oracle_adjust_maximum.restype = ct.c_float
    maxv = o.oracle_adjust_maximum(ct.c_float(dmax), ct.c_float(WHITE - BLACK), ct.c_float(0.75))
    want = oracle_amaze(scene.cfa_to_rgb3(cfa), RGGB, WB, maxv)
    want = _axial(o, "oracle_axial_ca", want, inplace=True)
    want = _oracle_defringe(want, True)
    want = oracle_detail_tonemap(want)
    err = float(np.abs(got - want).max())
    print(f"batch u16 chain vs oracle: max-abs {err:.3g}")
    assert err <= 2e-4 and psnr(got, want) >= 70.0


def test_serve_alone_and_in_the_mix(acc):
    from libraw_enhanced_amd import develop as dv
    spec = _spec()
    reqs = []
    for k, (kind, h, w, filters, _) in enumerate(spec):
        raw, xt = _inputs(k, kind, h, w, filters, "u16")
        im = dv.make_image(w, h, filters=filters, xtrans=xt, cam_mul=WB, black=BLACK, white=WHITE, wb=WB)
        reqs.append((im, np.ascontiguousarray(raw), h, w))

    def serve(server, order):
        outs = {k: np.full((reqs[k][2], reqs[k][3], 3), np.nan, np.float32) for k in order}
        for k in order:
            im, raw, h, w = reqs[k]
            server.submit(k, im, raw.ctypes.data, outs[k].ctypes.data, keep=(raw, outs[k]))
        done = 0
        while done < len(order):
            for c in server.poll(timeout_ms=1000):
                assert c["status"] == 0, c
                assert c["t_submit"] <= c["t_admit"] <= c["t_done"]
                done += 1
        return outs

    alone = dv.DevelopServer(0, lanes=1, hbm_budget_bytes=8 << 30)
    solo = {k: serve(alone, [k])[k] for k in range(len(reqs))}
    alone.close()
    # all at once, 3 lanes, a budget that admits two or three at a time
    mix = dv.DevelopServer(0, lanes=3, hbm_budget_bytes=64 << 20)
    outs = serve(mix, list(range(len(reqs))) * 1)
    _, peak = mix.footprint()
    mix.close()
    for k in range(len(reqs)):
        assert np.array_equal(solo[k], outs[k]), f"request {k} differs alone vs in the mix"
    assert peak > 0
    # the batch path gives the same bytes too
    S = _Set(acc, [(kind, h, w, f, "u16") for (kind, h, w, f, _) in spec])
    b = dv.BatchDeveloper(0, 2)
    b.run(S.images(), CHAIN)
    b.synchronize()
    for k, g in enumerate(S.results()):
        assert np.array_equal(g, solo[k]), f"request {k}: batch and served results differ"
    b.close()


def test_batch_graph_dropped_when_buffers_grow(acc):
    """A captured graph holds the arenas and lane scratch as kernel arguments:
    an eager run that must reallocate them drops the graph, so replay fails
    instead of launching on freed memory; a new capture replays bit-exact."""
    from libraw_enhanced_amd import develop as dv
    from libraw_enhanced_amd.accelerator import LibRawInvalidArgument
    small = _Set(acc, _spec()[:2])
    b = dv.BatchDeveloper(0, 2)
    b.run(small.images(), CHAIN)
    b.synchronize()
    eager = small.results()
    b.capture(small.images(), CHAIN)
    b.replay()
    b.synchronize()
    for e, g in zip(eager, small.results()):
        assert np.array_equal(e, g)
    # larger images: more lane scratch and more demosaic tiles than the graph's
    big = _Set(acc, [("bayer", 640, 960, RGGB, "f32"), ("xtrans", 600, 900, 9, "u16")])
    print("eager batch that grows the buffers", flush=True)
    b.run(big.images(), CHAIN)
    b.synchronize()
    with pytest.raises(LibRawInvalidArgument):
        b.replay()
    for it in small.items:
        it["d_rgb"].upload(np.full((it["h"], it["w"], 3), np.nan, np.float32))
    acc.synchronize()
    b.capture(small.images(), CHAIN)
    b.replay()
    b.synchronize()
    for e, g in zip(eager, small.results()):
        assert np.array_equal(e, g)
    b.close()


def test_serve_rejects_bad_lists_and_isolates_requests(acc):
    """Stage lists a request cannot run, and stage parameters the launchers
    reject, fail at creation; an invalid request fails at submission; the
    requests around it come out as they do alone."""
    from libraw_enhanced_amd import develop as dv
    from libraw_enhanced_amd.accelerator import LibRawInvalidArgument
    for stages, kw in [(("axial_ca", "defringe"), {}), ((), {}),
                       (CHAIN, {"axial_radius": 17}), (CHAIN, {"defringe_radius": float("nan")})]:
        with pytest.raises(LibRawInvalidArgument):
            dv.DevelopServer(0, lanes=1, hbm_budget_bytes=1 << 30, stages=stages,
                             params=dv.ChainParams.defaults(**kw))
    kind, h, w, filters, _ = _spec()[0]
    raw, xt = _inputs(0, kind, h, w, filters, "u16")
    raw = np.ascontiguousarray(raw)
    im = dv.make_image(w, h, filters=filters, xtrans=xt, cam_mul=WB, black=BLACK, white=WHITE, wb=WB)
    bad = dv.make_image(w, h, filters=filters, xtrans=xt, cam_mul=WB, black=WHITE, white=BLACK, wb=WB)
    outs = [np.full((h, w, 3), np.nan, np.float32) for _ in range(3)]
    server = dv.DevelopServer(0, lanes=2, hbm_budget_bytes=1 << 30)
    server.submit(0, im, raw.ctypes.data, outs[0].ctypes.data, keep=(raw, outs[0]))
    with pytest.raises(LibRawInvalidArgument):
        server.submit(1, bad, raw.ctypes.data, outs[1].ctypes.data)
    server.submit(2, im, raw.ctypes.data, outs[2].ctypes.data, keep=(raw, outs[2]))
    done = {}
    while len(done) < 2:
        for c in server.poll(timeout_ms=1000):
            done[c["id"]] = c["status"]
    server.close()
    assert done == {0: 0, 2: 0}
    assert np.isnan(outs[1]).all()
    assert np.array_equal(outs[0], outs[2]) and not np.isnan(outs[0]).any()


def test_context_scratch_across_streams(acc):
    """Device-resident calls of one context on two streams share its scratch:
    the second call waits for the first, and both results equal the same calls
    on one stream."""
    from libraw_enhanced_amd.accelerator import native
    hip = ct.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [ct.POINTER(ct.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ct.c_void_p]
    other = ct.c_void_p()
    assert hip.hipStreamCreate(ct.byref(other)) == 0
    lib, ctx = native(), acc.handle
    h, w = 300, 420
    img = np.ascontiguousarray(scene.render(h, w, scene.BASE_SEED + 7), np.float32)
    d_a, d_b = acc.dev_from_host(img), acc.dev_from_host(img)
    d_x, d_y = acc.dev_alloc(img.nbytes), acc.dev_alloc(img.nbytes)
    args = (ct.c_float(10.0), ct.c_float(0.1), ct.c_float(0.15), ct.c_float(10.0), 0, ct.c_float(0.3))
    # reference: both on the context stream
    assert lib.lre_hip_dev_defringe(ctx, d_a.ptr, d_x.ptr, w, h, *args, None) == 0
    assert lib.lre_hip_dev_axial_ca(ctx, d_b.ptr, d_y.ptr, w, h, 6, ct.c_float(1e-4), ct.c_float(0.3), None) == 0
    acc.synchronize()
    want_x, want_y = d_x.download((h, w, 3)), d_y.download((h, w, 3))
    for d in (d_x, d_y):
        d.upload(np.full((h, w, 3), np.nan, np.float32))
    acc.synchronize()
    # the same two calls on two streams, back to back
    assert lib.lre_hip_dev_defringe(ctx, d_a.ptr, d_x.ptr, w, h, *args, other) == 0
    assert lib.lre_hip_dev_axial_ca(ctx, d_b.ptr, d_y.ptr, w, h, 6, ct.c_float(1e-4), ct.c_float(0.3), None) == 0
    acc.synchronize()
    hip.hipStreamSynchronize.argtypes = [ct.c_void_p]
    assert hip.hipStreamSynchronize(other) == 0
    assert np.array_equal(d_x.download((h, w, 3)), want_x)
    assert np.array_equal(d_y.download((h, w, 3)), want_y)
    assert hip.hipStreamDestroy(other) == 0
