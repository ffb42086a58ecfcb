// Many images on one device: the batched develop (multi-image tile queues for
// the demosaics, remaining stages spread over streams, optional HIP graph) and
// the request loop (FIFO admission by device footprint, one worker per stream).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "lre_common.h"
#include "lre_internal.h"

namespace {

constexpr int MAX_LANES = 8;

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

bool valid_stage(int s) { return s >= LRE_STAGE_DEMOSAIC && s <= LRE_STAGE_GREEN_MATCHING; }
bool pre_stage(int s) { return s == LRE_STAGE_REMOVE_ZEROES || s == LRE_STAGE_GREEN_MATCHING; }

// leading raw-preprocessing stages (zero removal, green matching) of a list
int pre_count(const int *stages, int ns) {
    int k = 0;
    while (k < ns && pre_stage(stages[k])) k++;
    return k;
}

bool has_stage(const int *stages, int ns, int st) {
    for (int i = 0; i < ns; i++)
        if (stages[i] == st) return true;
    return false;
}

// A stage list with a highlight stage keeps the image's highlight mask at the
// start of the lane scratch, below what every stage uses.
size_t mask_bytes(const lre_image &im, const int *stages, int ns) {
    return has_stage(stages, ns, LRE_STAGE_HIGHLIGHT) ? align256((size_t)im.width * im.height * 4) : 0;
}

int check_stages(const int *stages, int ns) {
    if (ns < 0 || (ns > 0 && !stages)) {
        lre_set_error("invalid stage list");
        return LRE_ERR_ARG;
    }
    for (int i = 0; i < ns; i++) {
        if (!valid_stage(stages[i])) {
            lre_set_error("unknown stage code %d", stages[i]);
            return LRE_ERR_ARG;
        }
    }
    // [zero removal] [green matching] demosaic ..., or a list without these
    const int k = pre_count(stages, ns);
    if (k == 2 && !(stages[0] == LRE_STAGE_REMOVE_ZEROES && stages[1] == LRE_STAGE_GREEN_MATCHING)) {
        lre_set_error("zero removal comes before green matching, once each");
        return LRE_ERR_ARG;
    }
    if (k > 2 || (k > 0 && (k == ns || stages[k] != LRE_STAGE_DEMOSAIC))) {
        lre_set_error("zero removal and green matching go first, once each, followed by the demosaic");
        return LRE_ERR_ARG;
    }
    for (int i = k + 1; i < ns; i++)
        if (stages[i] == LRE_STAGE_DEMOSAIC || pre_stage(stages[i])) {
            lre_set_error("the demosaic stage must come first, after zero removal and green matching");
            return LRE_ERR_ARG;
        }
    if (k == 0 && ns > 0 && stages[0] != LRE_STAGE_DEMOSAIC && has_stage(stages, ns, LRE_STAGE_DEMOSAIC)) {
        lre_set_error("the demosaic stage must come first");
        return LRE_ERR_ARG;
    }
    return LRE_OK;
}

// Raw preprocessing of one uint16 image before the demosaic
// (core/libraw_wrapper.cpp:1496-1547): zero removal on the raw samples, black
// subtraction and clip, green matching, then white balance into the float CFA
// plane with the data maximum, and adjust_maximum.  Without zero removal and
// green matching the first three run fused in the ingest kernel.  work: a
// uint16 plane (the raw samples are copied there unless it is d_raw itself),
// work2: a second plane for green matching, zscr: zero removal scratch.
// Bayer only for the two optional stages (X-Trans images skip them: green
// matching is a no-op there in the reference).
int enqueue_ingest(const lre_image &im, const uint16_t *d_raw, bool zero, bool green, uint16_t *work,
                   uint16_t *work2, void *zscr, float *cfa, float *stat, hipStream_t s) {
    const int w = im.width, h = im.height;
    const signed char *xt = im.filters == 9 ? im.xtrans : nullptr;
    const bool pre = (zero || green) && im.filters != 9;
    int rc;
    if (!pre) {
        rc = lre_launch_ingest_u16(d_raw, cfa, w, h, im.filters, xt, im.cblack, im.black, im.white, im.wb, stat, s);
    } else {
        if (work != d_raw)
            LRE_HIP_CHECK(hipMemcpyAsync(work, d_raw, (size_t)w * h * 2, hipMemcpyDeviceToDevice, s));
        if (zero && (rc = lre_launch_remove_zeroes_plane(work, w, h, im.filters, zscr, s))) return rc;
        if ((rc = lre_launch_black_plane(work, w, h, im.filters, xt, im.cblack, im.black, im.white, s))) return rc;
        const uint16_t *src = work;
        if (green) {
            if ((rc = lre_launch_green_matching_plane(work, work2, w, h, im.filters, im.white, s))) return rc;
            src = work2;
        }
        const int no_black[4] = {0, 0, 0, 0};
        rc = lre_launch_ingest_u16(src, cfa, w, h, im.filters, xt, no_black, 0, im.white, im.wb, stat, s);
    }
    if (rc) return rc;
    return lre_launch_adjust_maximum(stat, (float)(im.white - im.black), im.adjust_threshold, stat + 1, s);
}

// device bytes of that preprocessing beyond the raw samples (0 without it)
size_t pre_bytes(const lre_image &im, bool zero, bool green, bool own_raw) {
    if (!(zero || green) || im.filters == 9) return 0;
    const size_t plane = align256((size_t)im.width * im.height * 2);
    return (own_raw ? 0 : plane) + (green ? plane : 0) +
           (zero ? align256(lre_remove_zeroes_scratch_bytes(im.width, im.height)) : 0);
}

// The parameters of the stages in the list, checked before anything is
// enqueued (a launcher that rejects them later would leave a partial chain).
int check_params(const int *stages, int ns, const lre_chain_params &p) {
    int rc = LRE_OK;
    for (int i = 0; i < ns && !rc; i++) {
        if (stages[i] == LRE_STAGE_AXIAL_CA) rc = lre_check_axial_params(p.axial_radius);
        else if (stages[i] == LRE_STAGE_DEFRINGE) rc = lre_check_defringe_params(p.defringe_radius);
    }
    return rc;
}

int check_image(const lre_image &im, bool demosaic) {
    if (im.width < 1 || im.height < 1 || !im.d_rgb) {
        lre_set_error("image needs a positive size and an output buffer");
        return LRE_ERR_ARG;
    }
    if (demosaic) {
        if (!im.d_cfa && !im.d_raw) {
            lre_set_error("image has neither a float nor a uint16 CFA plane");
            return LRE_ERR_ARG;
        }
        if (im.d_cfa && !(im.maximum_value > 0.f)) {
            lre_set_error("maximum_value must be positive");
            return LRE_ERR_ARG;
        }
        if (!im.d_cfa && !(im.white > im.black)) {
            lre_set_error("white level must exceed the black level");
            return LRE_ERR_ARG;
        }
    }
    return LRE_OK;
}

// device scratch one image needs for the stages after the demosaic
size_t post_scratch(const lre_image &im, const int *stages, int ns, const lre_chain_params &p) {
    size_t m = 0;
    for (int i = 0; i < ns; i++) {
        size_t b = 0;
        switch (stages[i]) {
        case LRE_STAGE_LATERAL_CA: b = lre_lateral_scratch_bytes(im.width, im.height, p.lateral_cell); break;
        case LRE_STAGE_AXIAL_CA: b = lre_axial_scratch_bytes(im.width, im.height); break;
        case LRE_STAGE_DEFRINGE: b = lre_defringe_scratch_bytes(im.width, im.height); break;
        case LRE_STAGE_DETAIL_TM: b = lre_detail_tm_scratch_bytes(im.width, im.height); break;
        case LRE_STAGE_HIGHLIGHT:  // a copy of the input, then the stage's own scratch
            b = align256((size_t)im.width * im.height * 12) + lre_highlight_scratch_bytes(im.width, im.height);
            break;
        case LRE_STAGE_MICRO_CONTRAST: b = lre_micro_contrast_scratch_bytes(im.width, im.height); break;
        default: break;
        }
        m = std::max(m, b);
    }
    return mask_bytes(im, stages, ns) + m;
}

// one post-demosaic stage, in place on the image's RGB buffer (the develop
// pipeline calls every one of them in place, core/libraw_wrapper.cpp:1776-1818)
int run_post_stage(int stage, const lre_image &im, const lre_chain_params &p, void *scratch,
                   hipStream_t s, int blur_rows = 0) {
    float *rgb = im.d_rgb;
    const int w = im.width, h = im.height;
    switch (stage) {
    case LRE_STAGE_LATERAL_CA:
        return lre_launch_lateral_ca(rgb, rgb, w, h, p.lateral_cell, p.lateral_iterations,
                                     p.lateral_max_shift, p.lateral_min_confidence,
                                     p.lateral_levels, scratch, s);
    case LRE_STAGE_AXIAL_CA:
        return lre_launch_axial_ca(rgb, rgb, w, h, p.axial_radius, p.axial_epsilon,
                                   p.axial_strength, scratch, s);
    case LRE_STAGE_DEFRINGE:
        return lre_launch_defringe(rgb, rgb, w, h, p.defringe_radius, p.defringe_edge_threshold,
                                   p.defringe_chroma_threshold, p.defringe_strength,
                                   p.defringe_green, p.defringe_green_scale, scratch, s, blur_rows);
    case LRE_STAGE_DETAIL_TM:
        return lre_launch_detail_tonemap(rgb, rgb, w, h, scratch, s);
    case LRE_STAGE_TONE_MAPPING:
        return lre_launch_tone_mapping(rgb, rgb, (size_t)w * h * 3, p.tone_after_scale, s);
    case LRE_STAGE_COLOR_SPACE:
        return lre_launch_epilogue(rgb, rgb, (size_t)w * h, p.color_transform, 0.f, 0.f, -1, 32, s);
    case LRE_STAGE_GAMMA:
        return lre_launch_epilogue(rgb, rgb, (size_t)w * h, nullptr, p.gamma_power, p.gamma_slope,
                                   p.output_color_space, 32, s);
    default:
        return LRE_OK;
    }
}

// The stages from index `from` on.  A colour matrix and/or gamma stage that
// directly follows the detail tone map runs inside its store, and a gamma
// stage that directly follows the colour matrix shares its pass.  The results
// are the same bits as one pass per stage (same float operations on the same
// values); the fused forms save a 12 B/px read and write per stage.
int run_post_chain(const int *stages, int from, int ns, const lre_image &im, const lre_chain_params &p,
                   void *lane_scratch, hipStream_t s, int blur_rows = 0) {
    const size_t mb = mask_bytes(im, stages, ns), npx = (size_t)im.width * im.height;
    float *mask = mb ? static_cast<float *>(lane_scratch) : nullptr;
    void *scratch = static_cast<char *>(lane_scratch) + mb;
    bool mask_ready = false;
    for (int si = from; si < ns; si++) {
        const int st = stages[si];
        if (st == LRE_STAGE_HIGHLIGHT) {
            // in place: the input is copied once, as the reference rewrites its buffer
            float *copy = static_cast<float *>(scratch);
            LRE_HIP_CHECK(hipMemcpyAsync(copy, im.d_rgb, npx * 12, hipMemcpyDeviceToDevice, s));
            const bool u16 = !im.d_cfa && im.d_raw;
            if (const int rc = lre_launch_recover_highlights(copy, im.d_rgb, im.width, im.height,
                                                             p.highlight_threshold, mask, u16 ? im.wb : nullptr,
                                                             nullptr, static_cast<char *>(scratch) +
                                                                          align256(npx * 12),
                                                             s))
                return rc;
            mask_ready = true;
            continue;
        }
        if (st == LRE_STAGE_MICRO_CONTRAST) {
            if (const int rc = lre_launch_micro_contrast(im.d_rgb, im.d_rgb, im.width, im.height,
                                                         p.highlight_threshold + 0.3f, p.micro_strength,
                                                         p.micro_target_contrast, mask_ready ? mask : nullptr,
                                                         scratch, s))
                return rc;
            continue;
        }
        if (st == LRE_STAGE_DETAIL_TM || st == LRE_STAGE_COLOR_SPACE) {
            int j = si + 1;
            bool matrix = st == LRE_STAGE_COLOR_SPACE;
            if (st == LRE_STAGE_DETAIL_TM && j < ns && stages[j] == LRE_STAGE_COLOR_SPACE) {
                matrix = true;
                j++;
            }
            const bool gamma = j < ns && stages[j] == LRE_STAGE_GAMMA;
            if (gamma) j++;
            if (j > si + 1) {
                const lre::Epilogue E = lre_make_epilogue(matrix ? p.color_transform : nullptr, p.gamma_power,
                                                          p.gamma_slope, gamma ? p.output_color_space : -1);
                const size_t npx = (size_t)im.width * im.height;
                const int rc = st == LRE_STAGE_DETAIL_TM
                                   ? lre_launch_detail_tonemap(im.d_rgb, im.d_rgb, im.width, im.height, scratch, s, &E)
                                   : lre_launch_epilogue(im.d_rgb, im.d_rgb, npx, p.color_transform, p.gamma_power,
                                                         p.gamma_slope, p.output_color_space, 32, s);
                if (rc) return rc;
                si = j - 1;
                continue;
            }
        }
        if (const int rc = run_post_stage(st, im, p, scratch, s, blur_rows)) return rc;
    }
    return LRE_OK;
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    // grow-only; the caller has made sure no queued work still uses the old buffer
    int ensure(size_t want) {
        if (want <= bytes) return LRE_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            lre_set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
            return LRE_ERR_NOMEM;
        }
        bytes = want;
        return LRE_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want) {
        if (want <= bytes) return LRE_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            lre_set_error("hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
            return LRE_ERR_NOMEM;
        }
        bytes = want;
        return LRE_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf scratch;
};

int lane_open(Lane &L) {
    LRE_HIP_CHECK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    LRE_HIP_CHECK(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    return LRE_OK;
}

void lane_close(Lane &L) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    L.scratch.release();
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    L.stream = nullptr;
    L.done = nullptr;
}

int device_cus(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        cus = 256;
    return cus > 0 ? cus : 256;
}

// longest-processing-time-first assignment of images to lanes
std::vector<std::vector<int>> assign_lanes(const lre_image *imgs, int n, int lanes) {
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (size_t)imgs[a].width * imgs[a].height > (size_t)imgs[b].width * imgs[b].height;
    });
    std::vector<std::vector<int>> out(lanes);
    std::vector<size_t> load(lanes, 0);
    for (int i : order) {
        int best = 0;
        for (int k = 1; k < lanes; k++)
            if (load[k] < load[best]) best = k;
        out[best].push_back(i);
        load[best] += (size_t)imgs[i].width * imgs[i].height;
    }
    return out;
}

}  // namespace

// ---------------------------------------------------------------------------
// batch

struct lre_batch {
    int device = 0;
    int num_cus = 256;
    int lanes = 2;
    int amaze_per_cu = 2;
    int xtrans_per_cu = 4;
    Lane lane[MAX_LANES];
    DevBuf amaze_arena, xtrans_arena, ingest_ws, d_tab_a, d_tab_x;
    HostBuf h_tab_a, h_tab_x;
    hipEvent_t ev_fork = nullptr, ev_ingest = nullptr, ev_demosaic[2] = {nullptr, nullptr};
    hipEvent_t ev_tables = nullptr;  // the last table upload has run
    bool tables_pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // tables of the captured graph (kept apart from those of eager runs)
    DevBuf g_tab_a, g_tab_x;
    // post-demosaic stages: one multi-image pass per stage over groups of
    // images whose scratch fits post_budget (group mode), or per image over
    // `lanes` streams
    int group_mode = 2;      // 0 off, 1 on, 2 by image size
    int blur_rows = 0;       // defringe blur tile height: 48, 96 or 0 (by image size)
    size_t post_budget = 0;  // 0: DEFAULT_POST_BUDGET, at most a quarter of the device's memory
    DevBuf post_scratch, d_tab_post, g_tab_post;
    HostBuf h_tab_post;
};

namespace {

int batch_sync_all(lre_batch *b) {
    for (int k = 0; k < MAX_LANES; k++)
        if (b->lane[k].stream) LRE_HIP_CHECK(hipStreamSynchronize(b->lane[k].stream));
    return LRE_OK;
}

void drop_graph(lre_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Default stage-scratch budget of group mode.  A group's launches already fill
// the device at a fraction of it (defringe, the largest, takes 60 B per pixel:
// 8 GiB holds some 140 MP of images at once).
constexpr size_t DEFAULT_POST_BUDGET = (size_t)8 << 30;

// group mode 2 groups batches whose mean image is below this many pixels
// (measured on the default chain: 256 x 0.3 MP, grouped 1.6x faster than the
// best lane count; 64 x 3.1 MP equal; 16 x 12 MP, grouped 6% slower)
constexpr size_t GROUP_AUTO_PX = 2000000;

// one multi-image pass of a group: where its table sits and its grid sizes
struct GroupPass {
    int stage;               // LRE_STAGE_AXIAL_CA, _DEFRINGE or _DETAIL_TM
    size_t tab_off = 0;      // into the post table buffer
    int grids[6] = {0, 0, 0, 0, 0, 0};
    bool epi = false;        // detail TM with a fused matrix / gamma
    bool matrix = false, gamma = false;
    int rows = 0;            // defringe: blur tile height
};
// images whose post-demosaic scratch is held at once, run on one lane
struct Group {
    int lane = 0;
    size_t base = 0;                  // the lane's region of the post scratch
    std::vector<int> imgs;
    std::vector<size_t> scratch_off;  // per image, from base
    size_t scratch_bytes = 0;
    std::vector<GroupPass> passes;
};

struct Plan {
    std::vector<lre_demosaic_job> bayer, xtrans;
    int tiles_a = 0, tiles_x = 0, slots_a = 0, slots_x = 0;
    bool demosaic = false;
    std::vector<size_t> ingest_off;  // per image offset into the ingest workspace (or SIZE_MAX)
    size_t ingest_bytes = 0;
    int first = 0;                   // index of the first post-demosaic stage
    bool zero = false, green = false;  // raw preprocessing stages
    size_t lane_scratch = 0;
    bool grouped = false;            // post stages as multi-image passes
    std::vector<Group> groups;
    size_t post_scratch = 0, post_tab_bytes = 0;
};

// Whether every post-demosaic stage of the list has a multi-image pass (or is
// a pointwise stage, run per image in the group's stream order).
bool groupable(const int *stages, int from, int ns, const lre_chain_params &p) {
    for (int i = from; i < ns; i++) {
        switch (stages[i]) {
        case LRE_STAGE_AXIAL_CA: case LRE_STAGE_DETAIL_TM: case LRE_STAGE_TONE_MAPPING:
        case LRE_STAGE_COLOR_SPACE: case LRE_STAGE_GAMMA:
            break;
        case LRE_STAGE_DEFRINGE:
            if (!lre_defringe_multi_supported(p.defringe_radius, p.defringe_edge_threshold,
                                              p.defringe_chroma_threshold, p.defringe_strength, p.defringe_green,
                                              p.defringe_green_scale))
                return false;
            break;
        default:
            return false;
        }
    }
    return true;
}

// scratch one image needs in group mode (the largest of its passes)
size_t group_scratch(const lre_image &im, const int *stages, int from, int ns) {
    size_t m = 0;
    for (int i = from; i < ns; i++) {
        size_t b = 0;
        switch (stages[i]) {
        case LRE_STAGE_AXIAL_CA: b = lre_axial_scratch_bytes(im.width, im.height); break;
        case LRE_STAGE_DEFRINGE: b = lre_defringe_multi_scratch_bytes(im.width, im.height); break;
        case LRE_STAGE_DETAIL_TM: b = lre_detail_tm_scratch_bytes(im.width, im.height); break;
        default: break;
        }
        m = std::max(m, b);
    }
    return align256(m);
}

size_t device_memory(int device) {
    size_t free_b = 0, total = 0;
    (void)hipSetDevice(device);
    if (hipMemGetInfo(&free_b, &total) != hipSuccess) return (size_t)64 << 30;
    return total;
}

// The images are spread over the lanes as in per-image mode (largest first,
// to the least loaded lane); each lane's images, in image order, form groups
// as large as the lane's share of the budget allows (an image larger than it
// forms a group of its own).  Groups of one lane run one after another and
// share the lane's region of the post scratch.
void plan_groups(const lre_image *imgs, int n, const int *stages, int from, int ns, size_t budget, int lanes,
                 int blur_rows, Plan &P) {
    P.groups.clear();
    P.post_scratch = P.post_tab_bytes = 0;
    const size_t share = budget / (size_t)lanes;
    Group g;
    size_t lane_max = 0;
    auto close = [&]() {
        if (g.imgs.empty()) return;
        const int m = (int)g.imgs.size();
        for (int si = from; si < ns; si++) {
            const int st = stages[si];
            GroupPass gp;
            gp.stage = st;
            if (st == LRE_STAGE_DETAIL_TM) {
                // a colour matrix and/or gamma right after it run in its store
                int j = si + 1;
                if (j < ns && stages[j] == LRE_STAGE_COLOR_SPACE) { gp.matrix = true; j++; }
                if (j < ns && stages[j] == LRE_STAGE_GAMMA) { gp.gamma = true; j++; }
                gp.epi = gp.matrix || gp.gamma;
                si = j - 1;
            }
            if (st == LRE_STAGE_DEFRINGE) {
                size_t tiles = 0;
                for (int i : g.imgs) tiles += lre_defringe_tall_tiles(imgs[i].width, imgs[i].height);
                gp.rows = lre_defringe_blur_rows(blur_rows, tiles);
            }
            size_t tb = 0;
            if (st == LRE_STAGE_AXIAL_CA) tb = lre_axial_multi_table_bytes(m);
            else if (st == LRE_STAGE_DEFRINGE) tb = lre_defringe_multi_table_bytes(m);
            else if (st == LRE_STAGE_DETAIL_TM) tb = lre_detail_tm_multi_table_bytes(m);
            gp.tab_off = P.post_tab_bytes;
            P.post_tab_bytes += align256(tb);
            g.passes.push_back(gp);
        }
        lane_max = std::max(lane_max, g.scratch_bytes);
        P.groups.push_back(std::move(g));
        g = Group();
    };
    auto per_lane = assign_lanes(imgs, n, lanes);
    for (int k = 0; k < lanes; k++) {
        std::sort(per_lane[k].begin(), per_lane[k].end());
        lane_max = 0;
        for (int i : per_lane[k]) {
            const size_t b = group_scratch(imgs[i], stages, from, ns);
            if (!g.imgs.empty() && g.scratch_bytes + b > share) close();
            if (g.imgs.empty()) {
                g.lane = k;
                g.base = P.post_scratch;
            }
            g.imgs.push_back(i);
            g.scratch_off.push_back(g.scratch_bytes);
            g.scratch_bytes += b;
        }
        close();
        P.post_scratch += lane_max;
    }
}

int make_plan(lre_batch *b, const lre_image *imgs, int n, const int *stages, int ns,
              const lre_chain_params &p, Plan &P) {
    int rc = check_stages(stages, ns);
    if (!rc) rc = check_params(stages, ns, p);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !imgs)) {
        lre_set_error("invalid image list");
        return LRE_ERR_ARG;
    }
    const int k = pre_count(stages, ns);
    P.demosaic = k < ns && stages[k] == LRE_STAGE_DEMOSAIC;
    P.first = P.demosaic ? k + 1 : 0;
    P.zero = has_stage(stages, k, LRE_STAGE_REMOVE_ZEROES);
    P.green = has_stage(stages, k, LRE_STAGE_GREEN_MATCHING);
    P.ingest_off.assign(n, SIZE_MAX);
    for (int i = 0; i < n; i++) {
        if ((rc = check_image(imgs[i], P.demosaic))) return rc;
        P.lane_scratch = std::max(P.lane_scratch, post_scratch(imgs[i], stages, ns, p));
        if (!P.demosaic) continue;
        const lre_image &im = imgs[i];
        lre_demosaic_job j{};
        j.d_rgb = im.d_rgb;
        j.w = im.width;
        j.h = im.height;
        j.filters = im.filters;
        std::memcpy(j.xtrans, im.xtrans, 36);
        std::memcpy(j.cam_mul, im.cam_mul, sizeof(j.cam_mul));
        if (im.d_cfa) {
            j.d_cfa = im.d_cfa;
            j.maximum_value = im.maximum_value;
            j.d_maxv = nullptr;
        } else {
            // workspace: float CFA plane + {data_max, maximum}
            // float CFA plane + {data_max, maximum}, then the preprocessing planes
            P.ingest_off[i] = P.ingest_bytes;
            P.ingest_bytes += align256((size_t)im.width * im.height * 4) + 256 + pre_bytes(im, P.zero, P.green, false);
            j.maximum_value = (float)(im.white - im.black);
        }
        (im.filters == 9 ? P.xtrans : P.bayer).push_back(j);
    }
    const int from = P.first;
    // By image size: one launch per stage over many images pays where images
    // are small (launch and tail bound); large images fill the device alone
    // and keep their stages' data warm in the caches when each one runs its
    // whole chain before the next on a lane.
    size_t px = 0;
    for (int i = 0; i < n; i++) px += (size_t)imgs[i].width * imgs[i].height;
    const bool want = b->group_mode == 1 || (b->group_mode == 2 && n > 0 && px / (size_t)n < GROUP_AUTO_PX);
    P.grouped = want && ns > from && n > 0 && groupable(stages, from, ns, p);
    if (P.grouped) {
        const size_t budget =
            b->post_budget ? b->post_budget : std::min(DEFAULT_POST_BUDGET, device_memory(b->device) / 4);
        plan_groups(imgs, n, stages, from, ns, budget, b->lanes, b->blur_rows, P);
        P.lane_scratch = 0;
    }
    for (auto &j : P.bayer) P.tiles_a += lre_amaze_tiles(j.w, j.h);
    for (auto &j : P.xtrans) P.tiles_x += lre_xtrans_tiles(j.w, j.h);
    P.slots_a = std::max(1, std::min(P.tiles_a, b->num_cus * b->amaze_per_cu));
    P.slots_x = std::max(1, std::min(P.tiles_x, b->num_cus * b->xtrans_per_cu));
    return LRE_OK;
}

// Sizes every buffer for the plan.  Growing a buffer waits for the device.
int ensure_buffers(lre_batch *b, const Plan &P, int n) {
    const size_t need_a = P.bayer.empty() ? 0 : lre_amaze_multi_scratch_bytes(P.slots_a);
    const size_t need_x = P.xtrans.empty() ? 0 : lre_xtrans_multi_scratch_bytes(P.slots_x);
    bool grow = P.demosaic && (need_a > b->amaze_arena.bytes || need_x > b->xtrans_arena.bytes ||
                               P.ingest_bytes > b->ingest_ws.bytes ||
                               lre_amaze_multi_table_bytes(std::max(n, 1)) > b->d_tab_a.bytes ||
                               lre_xtrans_multi_table_bytes(std::max(n, 1)) > b->d_tab_x.bytes);
    for (int k = 0; k < b->lanes; k++) grow = grow || P.lane_scratch > b->lane[k].scratch.bytes;
    grow = grow || (P.grouped && (P.post_scratch > b->post_scratch.bytes || P.post_tab_bytes > b->d_tab_post.bytes ||
                                  P.post_tab_bytes > b->h_tab_post.bytes));
    if (!grow) return LRE_OK;
    int rc = batch_sync_all(b);
    if (rc) return rc;
    // A captured graph holds the arenas, workspace and lane scratch as kernel
    // arguments: once they are reallocated it must not be launched again.
    drop_graph(b);
    for (int k = 0; k < b->lanes; k++)
        if ((rc = b->lane[k].scratch.ensure(P.lane_scratch))) return rc;
    if (P.grouped) {
        if ((rc = b->post_scratch.ensure(P.post_scratch))) return rc;
        if ((rc = b->d_tab_post.ensure(P.post_tab_bytes))) return rc;
        if ((rc = b->h_tab_post.ensure(P.post_tab_bytes))) return rc;
    }
    if (P.demosaic) {
        if ((rc = b->amaze_arena.ensure(need_a))) return rc;
        if ((rc = b->xtrans_arena.ensure(need_x))) return rc;
        if (P.ingest_bytes && (rc = b->ingest_ws.ensure(P.ingest_bytes))) return rc;
        if ((rc = b->d_tab_a.ensure(lre_amaze_multi_table_bytes(std::max(n, 1))))) return rc;
        if ((rc = b->d_tab_x.ensure(lre_xtrans_multi_table_bytes(std::max(n, 1))))) return rc;
        if ((rc = b->h_tab_a.ensure(lre_amaze_multi_table_bytes(std::max(n, 1))))) return rc;
        if ((rc = b->h_tab_x.ensure(lre_xtrans_multi_table_bytes(std::max(n, 1))))) return rc;
    }
    return LRE_OK;
}

// Fill the per-image tables and copy them to d_a / d_x (stream-ordered on lane 0).
int upload_tables(lre_batch *b, Plan &P, const lre_image *imgs, int n, void *d_a, void *d_x, void *d_post,
                  const lre_chain_params *params, bool synchronous) {
    // the pinned tables may still be read by the previous upload
    if (b->tables_pending) LRE_HIP_CHECK(hipEventSynchronize(b->ev_tables));
    b->tables_pending = false;
    // point u16 images at their workspace planes
    char *ws = static_cast<char *>(b->ingest_ws.p);
    size_t ib = 0, ix = 0;
    for (int i = 0; i < n && P.demosaic; i++) {
        lre_demosaic_job &j = imgs[i].filters == 9 ? P.xtrans[ix++] : P.bayer[ib++];
        if (P.ingest_off[i] == SIZE_MAX) continue;
        // the float CFA plane, then {data maximum, demosaic maximum}
        j.d_cfa = reinterpret_cast<float *>(ws + P.ingest_off[i]);
        j.d_maxv = reinterpret_cast<float *>(ws + P.ingest_off[i] + align256((size_t)j.w * j.h * 4)) + 1;
    }
    hipStream_t s = b->lane[0].stream;
    if (P.demosaic && !P.bayer.empty()) {
        lre_amaze_multi_fill(P.bayer.data(), (int)P.bayer.size(), b->h_tab_a.p);
        LRE_HIP_CHECK(hipMemcpyAsync(d_a, b->h_tab_a.p, lre_amaze_multi_table_bytes((int)P.bayer.size()),
                                     hipMemcpyHostToDevice, s));
    }
    if (P.demosaic && !P.xtrans.empty()) {
        int t = lre_xtrans_multi_fill(P.xtrans.data(), (int)P.xtrans.size(), b->h_tab_x.p);
        if (t < 0) return t;
        LRE_HIP_CHECK(hipMemcpyAsync(d_x, b->h_tab_x.p, lre_xtrans_multi_table_bytes((int)P.xtrans.size()),
                                     hipMemcpyHostToDevice, s));
    }
    if (P.grouped && P.post_tab_bytes) {
        if (!d_post) {
            lre_set_error("no device buffer for the post-stage tables");
            return LRE_ERR_ARG;
        }
        char *h = static_cast<char *>(b->h_tab_post.p);
        char *scr = static_cast<char *>(b->post_scratch.p);
        const lre_chain_params &p = *params;
        for (Group &g : P.groups) {
            std::vector<lre_post_job> jobs(g.imgs.size());
            for (size_t k = 0; k < g.imgs.size(); k++) {
                const lre_image &im = imgs[g.imgs[k]];
                jobs[k] = lre_post_job{im.d_rgb, im.d_rgb, im.width, im.height, scr + g.base + g.scratch_off[k]};
            }
            const int m = (int)jobs.size();
            for (GroupPass &gp : g.passes) {
                int rc = LRE_OK;
                if (gp.stage == LRE_STAGE_AXIAL_CA)
                    rc = lre_axial_multi_fill(jobs.data(), m, p.axial_radius, p.axial_epsilon, p.axial_strength,
                                              h + gp.tab_off, gp.grids);
                else if (gp.stage == LRE_STAGE_DEFRINGE)
                    rc = lre_defringe_multi_fill(jobs.data(), m, h + gp.tab_off, gp.grids, gp.rows);
                else if (gp.stage == LRE_STAGE_DETAIL_TM)
                    rc = lre_detail_tm_multi_fill(jobs.data(), m, h + gp.tab_off, gp.grids);
                if (rc) return rc;
            }
        }
        LRE_HIP_CHECK(hipMemcpyAsync(d_post, b->h_tab_post.p, P.post_tab_bytes, hipMemcpyHostToDevice, s));
    }
    LRE_HIP_CHECK(hipEventRecord(b->ev_tables, s));
    b->tables_pending = true;
    if (synchronous) {
        LRE_HIP_CHECK(hipEventSynchronize(b->ev_tables));
        b->tables_pending = false;
    }
    return LRE_OK;
}

// Enqueue the whole batch.  Tables must already be on the device (d_a, d_x).
// the post-demosaic stages of group mode: per group, on its lane's stream,
// one multi-image pass per stage (pointwise stages per image)
int enqueue_groups(const Plan &P, const lre_image *imgs, const lre_chain_params &p, const void *d_post,
                   const Lane *L) {
    const char *d = static_cast<const char *>(d_post);
    for (const Group &g : P.groups) {
        const int m = (int)g.imgs.size();
        hipStream_t s = L[g.lane].stream;
        for (const GroupPass &gp : g.passes) {
            int rc = LRE_OK;
            switch (gp.stage) {
            case LRE_STAGE_AXIAL_CA:
                rc = lre_launch_axial_multi(d + gp.tab_off, m, p.axial_radius, gp.grids, s);
                break;
            case LRE_STAGE_DEFRINGE:
                rc = lre_launch_defringe_multi(d + gp.tab_off, m, gp.grids, p.defringe_radius,
                                               p.defringe_edge_threshold, p.defringe_chroma_threshold,
                                               p.defringe_strength, p.defringe_green, p.defringe_green_scale, gp.rows,
                                               s);
                break;
            case LRE_STAGE_DETAIL_TM: {
                const lre::Epilogue E = lre_make_epilogue(gp.matrix ? p.color_transform : nullptr, p.gamma_power,
                                                          p.gamma_slope, gp.gamma ? p.output_color_space : -1);
                rc = lre_launch_detail_tm_multi(d + gp.tab_off, m, gp.grids, s, gp.epi ? &E : nullptr);
                break;
            }
            default:
                // pointwise stages: per image (a matrix followed by gamma shares one pass,
                // as in the per-image chain)
                for (int i : g.imgs) {
                    const lre_image &im = imgs[i];
                    const size_t npx = (size_t)im.width * im.height;
                    if (gp.stage == LRE_STAGE_COLOR_SPACE && &gp + 1 < g.passes.data() + g.passes.size() &&
                        (&gp + 1)->stage == LRE_STAGE_GAMMA)
                        rc = lre_launch_epilogue(im.d_rgb, im.d_rgb, npx, p.color_transform, p.gamma_power,
                                                 p.gamma_slope, p.output_color_space, 32, s);
                    else if (gp.stage == LRE_STAGE_GAMMA && &gp > g.passes.data() &&
                             (&gp - 1)->stage == LRE_STAGE_COLOR_SPACE)
                        rc = LRE_OK;  // done with the matrix
                    else
                        rc = run_post_stage(gp.stage, im, p, nullptr, s);
                    if (rc) return rc;
                }
                break;
            }
            if (rc) return rc;
        }
    }
    return LRE_OK;
}

int enqueue(lre_batch *b, const Plan &P, const lre_image *imgs, int n, const int *stages, int ns,
            const lre_chain_params &p, const void *d_a, void *d_x, const void *d_post) {
    Lane *L = b->lane;
    const int lanes = b->lanes;
    hipStream_t s0 = L[0].stream;
    LRE_HIP_CHECK(hipEventRecord(b->ev_fork, s0));
    for (int k = 1; k < lanes; k++) LRE_HIP_CHECK(hipStreamWaitEvent(L[k].stream, b->ev_fork, 0));
    int rc;
    if (P.demosaic) {
        // uint16 preprocessing + device-side maximum (lane 0)
        char *ws = static_cast<char *>(b->ingest_ws.p);
        for (int i = 0; i < n; i++) {
            if (P.ingest_off[i] == SIZE_MAX) continue;
            const lre_image &im = imgs[i];
            const size_t px = (size_t)im.width * im.height;
            char *base = ws + P.ingest_off[i];
            float *cfa = reinterpret_cast<float *>(base);
            float *stat = reinterpret_cast<float *>(base + align256(px * 4));
            // preprocessing region: the working copy of the raw samples, the
            // green matching output, the zero removal scratch
            char *pre = base + align256(px * 4) + 256;
            uint16_t *work = reinterpret_cast<uint16_t *>(pre);
            uint16_t *work2 = reinterpret_cast<uint16_t *>(pre + align256(px * 2));
            void *zscr = pre + align256(px * 2) * (P.green ? 2 : 1);
            if ((rc = enqueue_ingest(im, im.d_raw, P.zero, P.green, work, work2, zscr, cfa, stat, s0))) return rc;
        }
        hipStream_t sx = lanes > 1 ? L[1].stream : s0;
        if (!P.xtrans.empty() && sx != s0) {
            LRE_HIP_CHECK(hipEventRecord(b->ev_ingest, s0));
            LRE_HIP_CHECK(hipStreamWaitEvent(sx, b->ev_ingest, 0));
        }
        if (!P.bayer.empty() &&
            (rc = lre_launch_amaze_multi(P.bayer.data(), (int)P.bayer.size(), d_a, P.tiles_a,
                                         P.slots_a, b->amaze_arena.p, s0)))
            return rc;
        if (!P.xtrans.empty() &&
            (rc = lre_launch_xtrans_multi(P.xtrans.data(), (int)P.xtrans.size(), b->h_tab_x.p, d_x,
                                          P.tiles_x, P.slots_x, b->xtrans_arena.p, sx)))
            return rc;
        // every lane waits for both demosaic queues
        LRE_HIP_CHECK(hipEventRecord(b->ev_demosaic[0], s0));
        if (sx != s0) LRE_HIP_CHECK(hipEventRecord(b->ev_demosaic[1], sx));
        for (int k = 0; k < lanes; k++) {
            if (k != 0) LRE_HIP_CHECK(hipStreamWaitEvent(L[k].stream, b->ev_demosaic[0], 0));
            if (sx != s0 && L[k].stream != sx)
                LRE_HIP_CHECK(hipStreamWaitEvent(L[k].stream, b->ev_demosaic[1], 0));
        }
    }
    const int first = P.first;
    if (P.grouped) {
        if ((rc = enqueue_groups(P, imgs, p, d_post, L))) return rc;
    } else if (ns > first) {
        auto plan = assign_lanes(imgs, n, lanes);
        for (int k = 0; k < lanes; k++)
            for (int i : plan[k])
                if ((rc = run_post_chain(stages, first, ns, imgs[i], p, L[k].scratch.p, L[k].stream,
                                         b->blur_rows)))
                    return rc;
    }
    for (int k = 1; k < lanes; k++) {
        LRE_HIP_CHECK(hipEventRecord(L[k].done, L[k].stream));
        LRE_HIP_CHECK(hipStreamWaitEvent(s0, L[k].done, 0));
    }
    return LRE_OK;
}

int batch_run_eager(lre_batch *b, const lre_image *imgs, int n, const int *stages, int ns,
                    const lre_chain_params *params) {
    lre_chain_params p;
    if (params) p = *params;
    else lre_chain_params_default(&p);
    Plan P;
    int rc = make_plan(b, imgs, n, stages, ns, p, P);
    if (rc) return rc;
    if ((rc = ensure_buffers(b, P, n))) return rc;
    if ((P.demosaic || P.grouped) &&
        (rc = upload_tables(b, P, imgs, n, b->d_tab_a.p, b->d_tab_x.p, b->d_tab_post.p, &p, false)))
        return rc;
    return enqueue(b, P, imgs, n, stages, ns, p, b->d_tab_a.p, b->d_tab_x.p, b->d_tab_post.p);
}

}  // namespace

extern "C" {

void lre_chain_params_default(lre_chain_params *p) {
    if (!p) return;
    p->defringe_radius = 10.0f;
    p->defringe_edge_threshold = 0.1f;
    p->defringe_chroma_threshold = 0.15f;
    p->defringe_strength = 10.0f;
    p->defringe_green = 0;
    p->defringe_green_scale = 0.3f;
    p->axial_radius = 6;
    p->axial_epsilon = 1e-4f;
    p->axial_strength = 0.3f;
    p->lateral_cell = 96;
    p->lateral_iterations = 3;
    p->lateral_max_shift = 6.0f;
    p->lateral_min_confidence = 0.02f;
    p->lateral_levels = 3;
    p->tone_after_scale = 1.0f;
    for (int k = 0; k < 12; k++) p->color_transform[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    p->gamma_power = 0.0f;  // ProcessingParams defaults (core/accelerator.h:106, :123-124)
    p->gamma_slope = 0.0f;
    p->output_color_space = 1;
    p->highlight_threshold = 1.0f;  // compute_default_threshold's fallback (core/libraw_wrapper.cpp:3004-3011)
    p->micro_strength = 8.0f;       // postprocess's call (:1686, :1698-1705)
    p->micro_target_contrast = 0.04f;
}

int lre_hip_batch_create(lre_batch **out, int device, int lanes) {
    if (!out) return LRE_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        lre_set_error("no HIP device %d", device);
        return LRE_ERR_NODEV;
    }
    if (lanes < 1 || lanes > MAX_LANES) {
        lre_set_error("lanes must be 1..%d", MAX_LANES);
        return LRE_ERR_ARG;
    }
    LRE_HIP_CHECK(hipSetDevice(device));
    auto *b = new lre_batch();
    b->device = device;
    b->num_cus = device_cus(device);
    b->lanes = lanes;
    int rc = LRE_OK;
    for (int k = 0; k < MAX_LANES && !rc; k++) rc = lane_open(b->lane[k]);
    hipEvent_t *evs[] = {&b->ev_fork, &b->ev_ingest, &b->ev_demosaic[0], &b->ev_demosaic[1], &b->ev_tables};
    for (auto *e : evs)
        if (!rc && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) rc = LRE_ERR_HIP;
    if (rc) {
        lre_hip_batch_destroy(b);
        return rc;
    }
    *out = b;
    return LRE_OK;
}

int lre_hip_batch_destroy(lre_batch *b) {
    if (!b) return LRE_OK;
    (void)hipSetDevice(b->device);
    batch_sync_all(b);
    drop_graph(b);
    for (int k = 0; k < MAX_LANES; k++) lane_close(b->lane[k]);
    DevBuf *bufs[] = {&b->amaze_arena, &b->xtrans_arena, &b->ingest_ws, &b->d_tab_a, &b->d_tab_x,
                      &b->g_tab_a, &b->g_tab_x, &b->post_scratch, &b->d_tab_post, &b->g_tab_post};
    for (auto *d : bufs) d->release();
    b->h_tab_a.release();
    b->h_tab_x.release();
    b->h_tab_post.release();
    hipEvent_t evs[] = {b->ev_fork, b->ev_ingest, b->ev_demosaic[0], b->ev_demosaic[1], b->ev_tables};
    for (auto e : evs)
        if (e) (void)hipEventDestroy(e);
    delete b;
    return LRE_OK;
}

int lre_hip_batch_run(lre_batch *b, const lre_image *imgs, int n, const int *stages, int ns,
                      const lre_chain_params *params) {
    if (!b) return LRE_ERR_ARG;
    LRE_HIP_CHECK(hipSetDevice(b->device));
    return batch_run_eager(b, imgs, n, stages, ns, params);
}

int lre_hip_batch_synchronize(lre_batch *b) {
    if (!b) return LRE_ERR_ARG;
    return batch_sync_all(b);
}

int lre_hip_batch_capture(lre_batch *b, const lre_image *imgs, int n, const int *stages, int ns,
                          const lre_chain_params *params) {
    if (!b) return LRE_ERR_ARG;
    LRE_HIP_CHECK(hipSetDevice(b->device));
    lre_chain_params p;
    if (params) p = *params;
    else lre_chain_params_default(&p);
    // one eager run sizes every buffer (no allocation may happen under capture)
    int rc = batch_run_eager(b, imgs, n, stages, ns, &p);
    if (rc) return rc;
    if ((rc = batch_sync_all(b))) return rc;
    drop_graph(b);
    Plan P;
    if ((rc = make_plan(b, imgs, n, stages, ns, p, P))) return rc;
    if ((rc = ensure_buffers(b, P, n))) return rc;
    if (P.demosaic || P.grouped) {
        if ((rc = b->g_tab_a.ensure(lre_amaze_multi_table_bytes(std::max(n, 1))))) return rc;
        if ((rc = b->g_tab_x.ensure(lre_xtrans_multi_table_bytes(std::max(n, 1))))) return rc;
        if (P.grouped && (rc = b->g_tab_post.ensure(std::max<size_t>(P.post_tab_bytes, 256)))) return rc;
        // the graph's own tables, uploaded now and read by every replay
        if ((rc = upload_tables(b, P, imgs, n, b->g_tab_a.p, b->g_tab_x.p, b->g_tab_post.p, &p, true))) return rc;
    }
    hipStream_t s0 = b->lane[0].stream;
    LRE_HIP_CHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    rc = enqueue(b, P, imgs, n, stages, ns, p, b->g_tab_a.p, b->g_tab_x.p, b->g_tab_post.p);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s0, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        lre_set_error("stream capture failed: %s", hipGetErrorString(e));
        return LRE_ERR_HIP;
    }
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        lre_set_error("graph instantiation failed: %s", hipGetErrorString(e));
        return LRE_ERR_HIP;
    }
    return LRE_OK;
}

int lre_hip_batch_replay(lre_batch *b) {
    if (!b || !b->exec) {
        lre_set_error("no captured batch");
        return LRE_ERR_ARG;
    }
    LRE_HIP_CHECK(hipSetDevice(b->device));
    LRE_HIP_CHECK(hipGraphLaunch(b->exec, b->lane[0].stream));
    return LRE_OK;
}

int lre_hip_batch_set_config(lre_batch *b, int lanes, int amaze_per_cu, int xtrans_per_cu) {
    if (!b) return LRE_ERR_ARG;
    if (lanes < 1 || lanes > MAX_LANES || amaze_per_cu < 1 || amaze_per_cu > 16 ||
        xtrans_per_cu < 1 || xtrans_per_cu > 16) {
        lre_set_error("config out of range (lanes 1..%d, slots per CU 1..16)", MAX_LANES);
        return LRE_ERR_ARG;
    }
    int rc = batch_sync_all(b);
    if (rc) return rc;
    drop_graph(b);
    b->lanes = lanes;
    b->amaze_per_cu = amaze_per_cu;
    b->xtrans_per_cu = xtrans_per_cu;
    return LRE_OK;
}

int lre_hip_batch_set_grouping(lre_batch *b, int group_mode, size_t post_budget_bytes) {
    if (!b) return LRE_ERR_ARG;
    if (group_mode < 0 || group_mode > 2) {
        lre_set_error("group_mode must be 0, 1 or 2");
        return LRE_ERR_ARG;
    }
    int rc = batch_sync_all(b);
    if (rc) return rc;
    drop_graph(b);
    b->group_mode = group_mode;
    b->post_budget = post_budget_bytes;
    return LRE_OK;
}

int lre_hip_batch_set_tiles(lre_batch *b, int defringe_blur_rows) {
    if (!b) return LRE_ERR_ARG;
    if (defringe_blur_rows != 0 && defringe_blur_rows != 48 && defringe_blur_rows != 96) {
        lre_set_error("defringe blur rows must be 0, 48 or 96");
        return LRE_ERR_ARG;
    }
    int rc = batch_sync_all(b);
    if (rc) return rc;
    drop_graph(b);
    b->blur_rows = defringe_blur_rows;
    return LRE_OK;
}

int lre_hip_batch_get_tiles(lre_batch *b, int *defringe_blur_rows) {
    if (!b) return LRE_ERR_ARG;
    if (defringe_blur_rows) *defringe_blur_rows = b->blur_rows;
    return LRE_OK;
}

int lre_hip_batch_get_grouping(lre_batch *b, int *group_mode, size_t *post_budget_bytes) {
    if (!b) return LRE_ERR_ARG;
    if (group_mode) *group_mode = b->group_mode;
    if (post_budget_bytes) *post_budget_bytes = b->post_budget;
    return LRE_OK;
}

int lre_hip_batch_get_config(lre_batch *b, int *lanes, int *amaze_per_cu, int *xtrans_per_cu) {
    if (!b) return LRE_ERR_ARG;
    if (lanes) *lanes = b->lanes;
    if (amaze_per_cu) *amaze_per_cu = b->amaze_per_cu;
    if (xtrans_per_cu) *xtrans_per_cu = b->xtrans_per_cu;
    return LRE_OK;
}

int lre_hip_batch_autotune(lre_batch *b, const lre_image *imgs, int n, const int *stages, int ns,
                           const lre_chain_params *params, int reps, float *best_ms) {
    if (!b) return LRE_ERR_ARG;
    LRE_HIP_CHECK(hipSetDevice(b->device));
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    LRE_HIP_CHECK(hipEventCreate(&e0));
    LRE_HIP_CHECK(hipEventCreate(&e1));
    // a warm run (sizes the buffers), then `reps` timed runs of the current shape
    auto time_it = [&](float &ms) -> int {
        int rc = batch_run_eager(b, imgs, n, stages, ns, params);
        if (!rc) rc = batch_sync_all(b);
        if (rc) return rc;
        if (hipEventRecord(e0, b->lane[0].stream) != hipSuccess) return LRE_ERR_HIP;
        for (int r = 0; r < reps && !rc; r++) rc = batch_run_eager(b, imgs, n, stages, ns, params);
        if (rc) return rc;
        if (hipEventRecord(e1, b->lane[0].stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return LRE_ERR_HIP;
        ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)reps;
        return LRE_OK;
    };
    // launch shapes: group mode, lanes, workgroups per CU of the demosaic queues
    struct Cand { int group, lanes, a, x; };
    std::vector<Cand> cands;
    for (int group : {1, 0})
        for (int lanes : {1, 2, 3, 4})
            for (int a : {1, 2, 4}) cands.push_back({group, lanes, a, 2 * a});
    float best = 1e30f;
    Cand pick = {b->group_mode, b->lanes, b->amaze_per_cu, b->xtrans_per_cu};
    const size_t budget = b->post_budget;
    int rc = LRE_OK;
    for (const Cand &c : cands) {
        float ms = 0.f;
        if ((rc = lre_hip_batch_set_grouping(b, c.group, budget))) break;
        if ((rc = lre_hip_batch_set_config(b, c.lanes, c.a, c.x))) break;
        if ((rc = time_it(ms))) break;
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    if (!rc) rc = lre_hip_batch_set_grouping(b, pick.group, budget);
    if (!rc) rc = lre_hip_batch_set_config(b, pick.lanes, pick.a, pick.x);
    // then the defringe blur tile height on that shape (same bits either way)
    int rows = b->blur_rows;
    for (int cand : {48, 96}) {
        float ms = 0.f;
        if (rc || (rc = lre_hip_batch_set_tiles(b, cand)) || (rc = time_it(ms))) break;
        if (ms < best) {
            best = ms;
            rows = cand;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_ms) *best_ms = best;
    return lre_hip_batch_set_tiles(b, rows);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// request loop

namespace {

using Clock = std::chrono::steady_clock;

// Device memory for admitted requests, cached by exact size so a steady mix of
// request sizes stops calling hipMalloc after warm-up.
struct Pool {
    std::multimap<size_t, void *> free_blocks;
    size_t cached = 0;
    // cap: bytes the cache plus admitted requests may hold before the cache is given back
    void *get(size_t bytes, size_t in_use, size_t cap) {
        auto it = free_blocks.find(bytes);
        if (it != free_blocks.end()) {
            void *p = it->second;
            free_blocks.erase(it);
            cached -= bytes;
            return p;
        }
        if (cached + in_use > cap) trim();
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            // give back the cache and retry once
            trim();
            if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        }
        return p;
    }
    void put(void *p, size_t bytes) {
        free_blocks.emplace(bytes, p);
        cached += bytes;
    }
    void trim() {
        for (auto &kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
        cached = 0;
    }
};

}  // namespace

struct lre_server {
    int device = 0;
    int lanes = 1;
    size_t budget = 0;
    std::vector<int> stages;
    lre_chain_params params{};
    int demosaic_per_cu = 2;  // persistent workgroups per CU of one request's tile queue
    Clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::pair<lre_request, double>> pending;
    std::deque<lre_completion> done;
    size_t in_use = 0, peak = 0;
    bool stop = false;
    Pool pool;
    std::vector<std::thread> workers;
    std::vector<Lane> lane;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

struct Footprint {
    size_t raw, cfa, rgb, scratch, pre, total;
};

Footprint footprint(const lre_server *S, const lre_image &im) {
    Footprint f{};
    const size_t px = (size_t)im.width * im.height;
    f.raw = align256(px * 2);
    f.cfa = align256(px * 4) + 256;
    f.rgb = align256(px * 12);
    f.scratch = align256(post_scratch(im, S->stages.data(), (int)S->stages.size(), S->params));
    const int k = pre_count(S->stages.data(), (int)S->stages.size());
    // the raw samples are the request's own copy: preprocessed in place
    f.pre = pre_bytes(im, has_stage(S->stages.data(), k, LRE_STAGE_REMOVE_ZEROES),
                      has_stage(S->stages.data(), k, LRE_STAGE_GREEN_MATCHING), true);
    f.total = f.raw + f.cfa + f.rgb + f.scratch + f.pre;
    return f;
}

// the demosaic queue of one request: its own arena, sized by tiles
size_t demosaic_arena(const lre_image &im, int num_cus, int per_cu, int *slots) {
    if (im.filters == 9) {
        const int t = lre_xtrans_tiles(im.width, im.height);
        *slots = std::max(1, std::min(t, num_cus * per_cu));
        return align256(lre_xtrans_multi_scratch_bytes(*slots)) +
               align256(lre_xtrans_multi_table_bytes(1));
    }
    const int t = lre_amaze_tiles(im.width, im.height);
    *slots = std::max(1, std::min(t, num_cus * per_cu));
    return align256(lre_amaze_multi_scratch_bytes(*slots)) + align256(lre_amaze_multi_table_bytes(1));
}

// One request on one stream, start to finish.  Returns an LRE status.  May
// return early with work still queued on the lane's stream: serve_one waits.
int serve_one_enqueue(lre_server *S, Lane &L, const lre_request &rq, char *dev, const Footprint &f,
                      size_t arena_bytes, int slots, void *h_tab) {
    const lre_image &im = rq.image;
    hipStream_t s = L.stream;
    const size_t px = (size_t)im.width * im.height;
    uint16_t *d_raw = reinterpret_cast<uint16_t *>(dev);
    float *d_cfa = reinterpret_cast<float *>(dev + f.raw);
    float *d_stat = reinterpret_cast<float *>(dev + f.raw + align256(px * 4));
    float *d_rgb = reinterpret_cast<float *>(dev + f.raw + f.cfa);
    void *d_scr = dev + f.raw + f.cfa + f.rgb;
    char *d_arena = dev + f.total;
    void *d_tab = d_arena + (arena_bytes - align256(im.filters == 9 ? lre_xtrans_multi_table_bytes(1)
                                                                    : lre_amaze_multi_table_bytes(1)));
    int rc;
    LRE_HIP_CHECK(hipMemcpyAsync(d_raw, rq.h_raw, px * 2, hipMemcpyHostToDevice, s));
    const int k = pre_count(S->stages.data(), (int)S->stages.size());
    const bool zero = has_stage(S->stages.data(), k, LRE_STAGE_REMOVE_ZEROES);
    const bool green = has_stage(S->stages.data(), k, LRE_STAGE_GREEN_MATCHING);
    char *d_pre = dev + f.raw + f.cfa + f.rgb + f.scratch;  // green matching output, then zero removal scratch
    uint16_t *work2 = reinterpret_cast<uint16_t *>(d_pre);
    void *zscr = d_pre + (green ? align256(px * 2) : 0);
    if ((rc = enqueue_ingest(im, d_raw, zero, green, d_raw, work2, zscr, d_cfa, d_stat, s))) return rc;
    lre_demosaic_job j{};
    j.d_cfa = d_cfa;
    j.d_rgb = d_rgb;
    j.w = im.width;
    j.h = im.height;
    j.filters = im.filters;
    std::memcpy(j.xtrans, im.xtrans, 36);
    std::memcpy(j.cam_mul, im.cam_mul, sizeof(j.cam_mul));
    j.maximum_value = (float)(im.white - im.black);
    j.d_maxv = d_stat + 1;
    const bool has_demosaic = k < (int)S->stages.size() && S->stages[k] == LRE_STAGE_DEMOSAIC;
    if (has_demosaic) {
        if (im.filters == 9) {
            const int t = lre_xtrans_multi_fill(&j, 1, h_tab);
            if (t < 0) return t;
            // no tile on a frame with a side of 22 px or less: the pixels inside
            // the 8-px frame stay as in a new (zeroed) output buffer
            if (t == 0) LRE_HIP_CHECK(hipMemsetAsync(d_rgb, 0, px * 12, s));
            LRE_HIP_CHECK(hipMemcpyAsync(d_tab, h_tab, lre_xtrans_multi_table_bytes(1), hipMemcpyHostToDevice, s));
            if ((rc = lre_launch_xtrans_multi(&j, 1, h_tab, d_tab, t, slots, d_arena, s))) return rc;
        } else {
            const int t = lre_amaze_multi_fill(&j, 1, h_tab);
            LRE_HIP_CHECK(hipMemcpyAsync(d_tab, h_tab, lre_amaze_multi_table_bytes(1), hipMemcpyHostToDevice, s));
            if ((rc = lre_launch_amaze_multi(&j, 1, d_tab, t, slots, d_arena, s))) return rc;
        }
    }
    lre_image out = im;
    out.d_rgb = d_rgb;
    if ((rc = run_post_chain(S->stages.data(), has_demosaic ? k + 1 : 0, (int)S->stages.size(), out, S->params,
                             d_scr, s)))
        return rc;
    LRE_HIP_CHECK(hipMemcpyAsync(rq.h_rgb, d_rgb, px * 12, hipMemcpyDeviceToHost, s));
    return LRE_OK;
}

// Every exit waits for the lane's stream, so that neither the device block
// (returned to the pool) nor the caller's host buffers are still in use by a
// copy or kernel of this request when its completion is posted.
int serve_one(lre_server *S, Lane &L, const lre_request &rq, char *dev, const Footprint &f,
              size_t arena_bytes, int slots, void *h_tab) {
    const int rc = serve_one_enqueue(S, L, rq, dev, f, arena_bytes, slots, h_tab);
    const hipError_t e = hipStreamSynchronize(L.stream);
    if (rc) return rc;
    if (e != hipSuccess) {
        lre_set_error("request stream: %s", hipGetErrorString(e));
        return LRE_ERR_HIP;
    }
    return LRE_OK;
}

void worker_main(lre_server *S, int k) {
    (void)hipSetDevice(S->device);
    Lane &L = S->lane[k];
    const int num_cus = device_cus(S->device);
    void *h_tab = nullptr;
    const size_t tab_bytes = std::max(lre_amaze_multi_table_bytes(1), lre_xtrans_multi_table_bytes(1));
    if (hipHostMalloc(&h_tab, tab_bytes, hipHostMallocDefault) != hipSuccess) h_tab = nullptr;
    for (;;) {
        lre_request rq;
        double t_submit;
        Footprint f;
        size_t arena = 0, need = 0;
        int slots = 1;
        void *dev = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // FIFO: only the head request may be admitted
            for (;;) {
                if (S->stop) {
                    if (h_tab) (void)hipHostFree(h_tab);
                    return;
                }
                if (!S->pending.empty()) {
                    const lre_request &head = S->pending.front().first;
                    f = footprint(S, head.image);
                    arena = demosaic_arena(head.image, num_cus, S->demosaic_per_cu, &slots);
                    need = f.total + arena;
                    if (S->in_use + need <= S->budget || S->in_use == 0) break;
                }
                S->cv_work.wait(lk);
            }
            rq = S->pending.front().first;
            t_submit = S->pending.front().second;
            S->pending.pop_front();
            dev = S->pool.get(need, S->in_use + need, S->budget);
            S->in_use += need;
            S->peak = std::max(S->peak, S->in_use);
        }
        lre_completion c{};
        c.id = rq.id;
        c.t_submit = t_submit;
        c.t_admit = S->now();
        if (!dev) {
            lre_set_error("device allocation of %zu bytes failed", need);
            c.status = LRE_ERR_NOMEM;
        } else if (!h_tab) {
            c.status = LRE_ERR_NOMEM;
        } else {
            c.status = serve_one(S, L, rq, static_cast<char *>(dev), f, arena, slots, h_tab);
        }
        c.t_done = S->now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (dev) S->pool.put(dev, need);
            S->in_use -= need;
            S->done.push_back(c);
        }
        S->cv_work.notify_all();
        S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int lre_hip_serve_create(lre_server **out, int device, int lanes, size_t budget, const int *stages,
                         int ns, const lre_chain_params *params) {
    if (!out) return LRE_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        lre_set_error("no HIP device %d", device);
        return LRE_ERR_NODEV;
    }
    if (lanes < 1 || lanes > MAX_LANES || budget == 0) {
        lre_set_error("lanes must be 1..%d and the budget positive", MAX_LANES);
        return LRE_ERR_ARG;
    }
    int rc = check_stages(stages, ns);
    if (rc) return rc;
    // requests carry a uint16 CFA plane only: the chain must demosaic it first
    // (after zero removal and green matching, when listed)
    const int k = pre_count(stages, ns);
    if (k >= ns || stages[k] != LRE_STAGE_DEMOSAIC) {
        lre_set_error("a served stage list must start with the demosaic stage");
        return LRE_ERR_ARG;
    }
    lre_chain_params p;
    if (params) p = *params;
    else lre_chain_params_default(&p);
    if ((rc = check_params(stages, ns, p))) return rc;
    LRE_HIP_CHECK(hipSetDevice(device));
    auto *S = new lre_server();
    S->device = device;
    S->lanes = lanes;
    S->budget = budget;
    S->stages.assign(stages, stages + ns);
    S->params = p;
    S->t0 = Clock::now();
    S->lane.resize(lanes);
    for (int k = 0; k < lanes && !rc; k++) rc = lane_open(S->lane[k]);
    if (rc) {
        for (auto &L : S->lane) lane_close(L);
        delete S;
        return rc;
    }
    for (int k = 0; k < lanes; k++) S->workers.emplace_back(worker_main, S, k);
    *out = S;
    return LRE_OK;
}

int lre_hip_serve_submit(lre_server *S, const lre_request *rq) {
    if (!S || !rq || !rq->h_raw || !rq->h_rgb || rq->image.width < 1 || rq->image.height < 1 ||
        !(rq->image.white > rq->image.black)) {
        lre_set_error("invalid request");
        return LRE_ERR_ARG;
    }
    if (rq->image.filters == 9 && 2 * rq->image.width + 2 > 32767) {
        lre_set_error("X-Trans width %d exceeds the 16-bit hexagon offsets", rq->image.width);
        return LRE_ERR_ARG;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->pending.emplace_back(*rq, S->now());
    }
    S->cv_work.notify_all();
    return LRE_OK;
}

int lre_hip_serve_poll(lre_server *S, lre_completion *out, int max, int timeout_ms) {
    if (!S || !out || max < 1) return LRE_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && timeout_ms > 0)
        S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !S->done.empty(); });
    int k = 0;
    while (k < max && !S->done.empty()) {
        out[k++] = S->done.front();
        S->done.pop_front();
    }
    return k;
}

double lre_hip_serve_now(lre_server *S) { return S ? S->now() : 0.0; }

int lre_hip_serve_footprint(lre_server *S, size_t *current, size_t *peak) {
    if (!S) return LRE_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    if (current) *current = S->in_use;
    if (peak) *peak = S->peak;
    return LRE_OK;
}

int lre_hip_serve_destroy(lre_server *S) {
    if (!S) return LRE_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    (void)hipSetDevice(S->device);
    for (auto &L : S->lane) lane_close(L);
    S->pool.trim();
    delete S;
    return LRE_OK;
}

void *lre_hip_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

int lre_hip_host_free(void *p) {
    if (p) LRE_HIP_CHECK(hipHostFree(p));
    return LRE_OK;
}

}  // extern "C"
