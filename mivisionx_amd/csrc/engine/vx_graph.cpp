// vx_graph.cpp -- graphs and nodes: construction, verification, target affinity, fusion into the
// vision supernode, execution on one HIP stream per graph, and vxScheduleGraph / vxWaitGraph.
//
// Follows the reference's graph life cycle (vx_api.cpp:2877-3055 vxVerifyGraph/vxProcessGraph/
// vxScheduleGraph/vxWaitGraph; ago_interface.cpp:2154-2529 agoExecuteGraph; target resolution
// ago_drama_alloc.cpp:565-727), re-planned for one GPU: every node runs on the GPU, nodes are
// enqueued back to back with no host synchronisation between them (the reference synchronises at
// every CPU/GPU level change, ago_interface.cpp:2294-2309), and when a graph contains the RGB
// vision chain (ColorConvert -> ChannelExtract(Y) -> ScaleImage / GaussianPyramid / Canny / Harris)
// that chain runs as one fused supernode (mvx_VisionGraph_*), which the reference reserves
// (ago_util_hip.cpp:740-744) but never launches.  Every Gaussian3x3 -> Sobel3x3 -> Magnitude + Phase
// chain over a graph input (optionally with a half-size bilinear ScaleImage of the same input) is
// collected into a second supernode that runs all such chains, of any image sizes, as one
// persistent launch replayed from a HIP graph (mvx_ChainBatch_*).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <map>
#include <set>

#include "engine.hpp"

using namespace mvxe;

namespace mvxe {
vx_status prepare_node(_vx_node *n, hipStream_t st);

struct FusedVision {
    _vx_node *cc = nullptr, *ex = nullptr, *scale = nullptr, *pyr = nullptr, *canny = nullptr, *harris = nullptr;
    mvx_vision_params_t params{};
    mvx_vision_frame_t frame{};
    DevBuf ws;
    size_t wsSize = 0;
    DevBuf count;  // corner count when the Harris node has no num_corners scalar
    mvx_vision_lane_t *lane = nullptr;  // this graph's own side stream for the pyramid / pick stages
};

struct ChainNodes {
    _vx_node *gauss, *sobel, *mag, *phase, *scale;  // scale may be null
    _vx_image *in, *blur, *gx, *gy, *magOut, *phaseOut, *half;
    bool merged;  // virtual gx / gy read only by mag and phase: the merged SobelMagnitudePhase rules
};

struct FusedChains {
    std::vector<ChainNodes> chains;
    std::vector<mvx_chain_image_t> images;
    mvx_chain_plan_t plan{};
    DevBuf ws;
    void *graphExec = nullptr;  // instantiated HIP graph holding the batched launch
};
}  // namespace mvxe

static uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void perf_add(vx_perf_t &p, uint64_t beg, uint64_t end)
{
    p.beg = beg;
    p.end = end;
    p.tmp = end - beg;
    p.sum += p.tmp;
    p.num++;
    p.avg = p.sum / p.num;
    p.min = p.num == 1 ? p.tmp : std::min(p.min, p.tmp);
    p.max = std::max(p.max, p.tmp);
}

// the object a parameter stands for in dependency analysis: pyramid levels belong to their pyramid
static _vx_reference *root_of(_vx_reference *r)
{
    if (is_ref(r, VX_TYPE_IMAGE) && ((_vx_image *)r)->parent) return ((_vx_image *)r)->parent;
    return r;
}

namespace mvxe {

_vx_node *create_node(_vx_graph *graph, const KernelDef *def, std::initializer_list<_vx_reference *> params)
{
    if (!is_ref(graph, VX_TYPE_GRAPH) || !def) return nullptr;
    if (graph->scheduled) return nullptr;   // a vxScheduleGraph worker may be executing it
    std::lock_guard<std::mutex> lk(graph->runLock);
    _vx_node *n = new _vx_node(graph->context);
    n->kernel = def;
    n->graph = graph;
    n->params.assign(def->params.size(), nullptr);
    size_t i = 0;
    for (_vx_reference *p : params) {
        if (i >= n->params.size()) break;
        if (p && !is_ref(p)) { delete n; return nullptr; }
        n->params[i] = p;
        if (p) p->intCount++;
        i++;
    }
    n->intCount = 1;  // held by the graph
    add_ref(graph->context, n);
    graph->nodes.push_back(n);
    invalidate_graph(graph);
    return n;
}

void invalidate_graph(_vx_graph *g)
{
    for (_vx_node *n : g->nodes) {
        deinit_user_node(n);
        n->local.release();
        n->localSize = 0;
        n->fused = FUSED_NONE;
        n->mergedMag = n->mergedPhase = nullptr;
    }
    if (g->chains) {
        if (g->chains->graphExec) mvx_graph_destroy(g->chains->graphExec);
        g->chains->ws.release();
        delete g->chains;
        g->chains = nullptr;
    }
    g->chainsAt = -1;
    if (g->fused) {
        if (g->fused->lane) mvx_VisionGraph_lane_destroy(g->fused->lane);
        g->fused->ws.release();
        g->fused->count.release();
        delete g->fused;
        g->fused = nullptr;
    }
    g->fusedAt = -1;
    g->order.clear();
    g->writtenArrays.clear();
    g->writtenScalars.clear();
    g->verified = false;
    g->state = VX_GRAPH_STATE_UNVERIFIED;
}

}  // namespace mvxe

// ------------------------------------------------------------------------------ fusion -----------

static std::vector<_vx_node *> consumers(_vx_graph *g, _vx_reference *data)
{
    std::vector<_vx_node *> out;
    for (_vx_node *n : g->nodes)
        for (size_t i = 0; i < n->params.size(); i++)
            if (n->params[i] && root_of(n->params[i]) == root_of(data) && n->kernel->params[i].dir != VX_OUTPUT) {
                out.push_back(n);
                break;
            }
    return out;
}

static void fused_params(FusedVision *f)
{
    _vx_image *rgb = (_vx_image *)f->cc->params[0];
    _vx_image *scaled = (_vx_image *)f->scale->params[1];
    _vx_pyramid *pyr = (_vx_pyramid *)f->pyr->params[1];
    _vx_threshold *t = (_vx_threshold *)f->canny->params[1];
    _vx_array *corners = (_vx_array *)f->harris->params[6];
    mvx_vision_params_t &p = f->params;
    p.width = rgb->width;
    p.height = rgb->height;
    p.scaledWidth = scaled->width;
    p.scaledHeight = scaled->height;
    p.levels = (uint32_t)pyr->levels;
    p.cannyLower = t->lower.U32;
    p.cannyUpper = t->upper.U32;
    p.harrisThreshold = ((_vx_scalar *)f->harris->params[1])->value.f32;
    p.harrisMinDistance = ((_vx_scalar *)f->harris->params[2])->value.f32;
    p.harrisK = ((_vx_scalar *)f->harris->params[3])->value.f32;
    p.cornerCapacity = (uint32_t)corners->capacity;
}

// find ColorConvert(RGB -> virtual YUV) -> ChannelExtract(Y, virtual) whose luma feeds exactly one
// bilinear ScaleImage, one HALF GaussianPyramid, one 3x3 L1 Canny and one 3x3/3 Harris node
static FusedVision *find_fusion(_vx_graph *g)
{
    for (_vx_node *cc : g->nodes) {
        if (cc->kernel->id != VX_KERNEL_COLOR_CONVERT) continue;
        _vx_image *yuv = image_of(cc->params[1]);
        if (!yuv || !yuv->isVirtual) continue;
        std::vector<_vx_node *> c1 = consumers(g, yuv);
        if (c1.size() != 1 || c1[0]->kernel->id != VX_KERNEL_CHANNEL_EXTRACT) continue;
        _vx_node *ex = c1[0];
        // the extracted channel must be the luma plane, at the RGB image's size
        _vx_scalar *chan = is_ref(ex->params[1], VX_TYPE_SCALAR) ? (_vx_scalar *)ex->params[1] : nullptr;
        if (!chan || (chan->value.e != VX_CHANNEL_Y && chan->value.e != VX_CHANNEL_0)) continue;
        _vx_image *luma = image_of(ex->params[2]);
        _vx_image *rgbIn = image_of(cc->params[0]);
        if (!luma || !luma->isVirtual || !rgbIn || rgbIn->format != VX_DF_IMAGE_RGB) continue;
        if (luma->width != rgbIn->width || luma->height != rgbIn->height || luma->format != VX_DF_IMAGE_U8) continue;
        FusedVision f;
        f.cc = cc;
        f.ex = ex;
        bool ok = true;
        for (_vx_node *n : consumers(g, luma)) {
            _vx_node **slot = n->kernel->id == VX_KERNEL_SCALE_IMAGE ? &f.scale : n->kernel->id == VX_KERNEL_GAUSSIAN_PYRAMID ? &f.pyr
                            : n->kernel->id == VX_KERNEL_CANNY_EDGE_DETECTOR ? &f.canny : n->kernel->id == VX_KERNEL_HARRIS_CORNERS ? &f.harris
                            : nullptr;
            if (!slot || *slot || n->params[0] != luma) { ok = false; break; }
            *slot = n;
        }
        if (!ok || !f.scale || !f.pyr || !f.canny || !f.harris) continue;
        {
            // the tile pass computes the bilinear, border-UNDEFINED scaling only
            _vx_scalar *interp = is_ref(f.scale->params[2], VX_TYPE_SCALAR) ? (_vx_scalar *)f.scale->params[2] : nullptr;
            if (!interp || interp->value.e != VX_INTERPOLATION_BILINEAR || f.scale->border.mode != VX_BORDER_UNDEFINED) continue;
            // and the gradient-3, L1-norm Canny only
            _vx_scalar *gs = is_ref(f.canny->params[2], VX_TYPE_SCALAR) ? (_vx_scalar *)f.canny->params[2] : nullptr;
            _vx_scalar *nt = is_ref(f.canny->params[3], VX_TYPE_SCALAR) ? (_vx_scalar *)f.canny->params[3] : nullptr;
            if (!gs || !nt || gs->value.i32 != 3 || nt->value.e != VX_NORM_L1) continue;
            // and the gradient-3, block-3 Harris only
            _vx_scalar *hg = is_ref(f.harris->params[4], VX_TYPE_SCALAR) ? (_vx_scalar *)f.harris->params[4] : nullptr;
            _vx_scalar *hb = is_ref(f.harris->params[5], VX_TYPE_SCALAR) ? (_vx_scalar *)f.harris->params[5] : nullptr;
            if (!hg || !hb || hg->value.i32 != 3 || hb->value.i32 != 3) continue;
        }
        if (((_vx_pyramid *)f.pyr->params[1])->levels > 8) continue;
        fused_params(&f);
        if (mvx_VisionGraph_supported(&f.params) != 1) continue;
        return new FusedVision(f);
    }
    return nullptr;
}

static vx_status fused_prepare(_vx_graph *g)
{
    FusedVision *f = g->fused;
    fused_params(f);
    _vx_pyramid *pyr = (_vx_pyramid *)f->pyr->params[1];
    _vx_image *rgb = (_vx_image *)f->cc->params[0], *scaled = (_vx_image *)f->scale->params[1], *edges = (_vx_image *)f->canny->params[4];
    _vx_array *corners = (_vx_array *)f->harris->params[6];
    _vx_scalar *num = is_ref(f->harris->params[7], VX_TYPE_SCALAR) ? (_vx_scalar *)f->harris->params[7] : nullptr;
    mvx_vision_frame_t &fr = f->frame;
    memset(&fr, 0, sizeof(fr));
    fr.rgb = rgb->plane_ptr(0);
    fr.rgbStride = rgb->planes[0].stride;
    for (size_t l = 0; l < pyr->images.size(); l++) {
        fr.pyramid[l] = pyr->images[l]->plane_ptr(0);
        fr.pyramidStride[l] = pyr->images[l]->planes[0].stride;
    }
    fr.scaled = scaled->plane_ptr(0);
    fr.scaledStride = scaled->planes[0].stride;
    fr.edges = edges->plane_ptr(0);
    fr.edgesStride = edges->planes[0].stride;
    fr.corners = corners->items_ptr();
    fr.numItems = corners->count_ptr();
    if (num) {
        fr.numCorners = (uint64_t *)num->dev.ptr();
    } else {
        vx_status s = f->count.alloc(8);
        if (s) return s;
        fr.numCorners = (uint64_t *)f->count.ptr();
    }
    if (!f->lane && mvx_VisionGraph_lane_create(&f->lane) != 0) f->lane = nullptr;   // without one, every stage runs in order
    f->wsSize = mvx_VisionGraph_workspace_size(&f->params, 1);
    vx_status s = f->ws.alloc(f->wsSize);
    if (s) return s;
    const int r = mvx_VisionGraph_prepare(g->stream, &f->params, &fr, 1, f->ws.ptr(), f->wsSize);
    if (r) {
        log(g, r, "ERROR: vxVerifyGraph: fused vision supernode prepare returned %d\n", r);
        return r;
    }
    return VX_SUCCESS;
}

// -------------------------------------------------------------------- batched chains -----------
// Node roles, Khronos form (inputs first) or AMD form (output first, ago_kernel_list.cpp:339-474).

static bool gauss_io(_vx_node *n, _vx_image **in, _vx_image **out)
{
    if (n->kernel->id == VX_KERNEL_GAUSSIAN_3x3) { *in = image_of(n->params[0]); *out = image_of(n->params[1]); return true; }
    if (n->kernel->id == AMD_GAUSSIAN_U8_U8_3x3) { *in = image_of(n->params[1]); *out = image_of(n->params[0]); return true; }
    return false;
}

static bool sobel_io(_vx_node *n, _vx_image **in, _vx_image **gx, _vx_image **gy)
{
    if (n->kernel->id == VX_KERNEL_SOBEL_3x3) {
        *in = image_of(n->params[0]); *gx = image_of(n->params[1]); *gy = image_of(n->params[2]);
        return true;
    }
    if (n->kernel->id == AMD_SOBEL_S16S16_U8_3x3_GXY) {
        *gx = image_of(n->params[0]); *gy = image_of(n->params[1]); *in = image_of(n->params[2]);
        return true;
    }
    return false;
}

// Magnitude (phase = false) or Phase (phase = true) reading (gx, gy)
static bool gradpair_io(_vx_node *n, bool phase, _vx_image **gx, _vx_image **gy, _vx_image **out)
{
    const vx_enum khr = phase ? VX_KERNEL_PHASE : VX_KERNEL_MAGNITUDE, amd = phase ? AMD_PHASE_U8_S16S16 : AMD_MAGNITUDE_S16_S16S16;
    if (n->kernel->id == khr) {
        *gx = image_of(n->params[0]); *gy = image_of(n->params[1]); *out = image_of(n->params[2]);
        return true;
    }
    if (n->kernel->id == amd) {
        *out = image_of(n->params[0]); *gx = image_of(n->params[1]); *gy = image_of(n->params[2]);
        return true;
    }
    return false;
}

static bool scale_io(_vx_node *n, _vx_image **in, _vx_image **out)
{
    if (n->kernel->id == VX_KERNEL_SCALE_IMAGE) {
        _vx_scalar *interp = is_ref(n->params[2], VX_TYPE_SCALAR) ? (_vx_scalar *)n->params[2] : nullptr;
        if (!interp || interp->value.e != VX_INTERPOLATION_BILINEAR || n->border.mode != VX_BORDER_UNDEFINED) return false;
        *in = image_of(n->params[0]); *out = image_of(n->params[1]);
        return true;
    }
    if (n->kernel->id == AMD_SCALE_IMAGE_U8_U8_BILINEAR) { *in = image_of(n->params[1]); *out = image_of(n->params[0]); return true; }
    return false;
}

static bool written_in_graph(_vx_graph *g, _vx_reference *r)
{
    for (_vx_node *n : g->nodes)
        for (size_t i = 0; i < n->params.size(); i++)
            if (n->params[i] && n->kernel->params[i].dir != VX_INPUT && root_of(n->params[i]) == root_of(r)) return true;
    return false;
}

// The node-merge rules for Sobel (ago_drama_remove.cpp:155-182): Sobel_S16S16_U8_3x3_GXY whose gx and gy
// are both virtual and read by no node other than one Magnitude and/or one Phase node, each reading
// (gx, gy) in that order, becomes SobelMagnitudePhase, SobelMagnitude or SobelPhase.  Returns false when
// the rules do not apply; otherwise *mag / *phase are the absorbed nodes (one may be null).
static bool sobel_merge_targets(_vx_graph *g, _vx_node *sobel, _vx_node **mag, _vx_node **phase)
{
    _vx_image *in = nullptr, *gx = nullptr, *gy = nullptr;
    *mag = *phase = nullptr;
    if (!sobel_io(sobel, &in, &gx, &gy) || !gx || !gy || gx == gy || !gx->isVirtual || !gy->isVirtual) return false;
    std::set<_vx_node *> readers;
    for (_vx_node *n : consumers(g, gx)) readers.insert(n);
    for (_vx_node *n : consumers(g, gy)) readers.insert(n);
    for (_vx_node *n : readers) {
        _vx_image *a = nullptr, *b = nullptr, *o = nullptr;
        if (gradpair_io(n, false, &a, &b, &o) && a == gx && b == gy && !*mag) *mag = n;
        else if (gradpair_io(n, true, &a, &b, &o) && a == gx && b == gy && !*phase) *phase = n;
        else return false;   // the gradients are used outside the rule
    }
    return *mag || *phase;
}

// every Gaussian3x3 (input: a U8 graph input of even size; output: virtual, read only by one
// Sobel3x3 with both outputs) whose gradients feed one Magnitude and one Phase node, plus a bilinear
// ScaleImage of the same input to exactly half size when the graph has one
static FusedChains *find_chains(_vx_graph *g)
{
    FusedChains fc;
    std::set<_vx_node *> used;
    for (_vx_node *gs : g->order) {
        _vx_image *in = nullptr, *blur = nullptr;
        if (gs->fused || !gauss_io(gs, &in, &blur) || !in || !blur || !blur->isVirtual) continue;
        if (in->format != VX_DF_IMAGE_U8 || in->parent || in->uniform || written_in_graph(g, in)) continue;
        if (in->width < 2 || in->height < 2 || (in->width & 1) || (in->height & 1)) continue;
        std::vector<_vx_node *> c = consumers(g, blur);
        _vx_image *sin = nullptr, *gx = nullptr, *gy = nullptr;
        if (c.size() != 1 || c[0]->fused || !sobel_io(c[0], &sin, &gx, &gy) || sin != blur || !gx || !gy) continue;
        ChainNodes ch{gs, c[0], nullptr, nullptr, nullptr, in, blur, gx, gy, nullptr, nullptr, nullptr, false};
        for (_vx_node *n : consumers(g, gx)) {
            _vx_image *a = nullptr, *b = nullptr, *o = nullptr;
            if (n->fused) continue;
            if (gradpair_io(n, false, &a, &b, &o) && a == gx && b == gy && !ch.mag) { ch.mag = n; ch.magOut = o; }
            else if (gradpair_io(n, true, &a, &b, &o) && a == gx && b == gy && !ch.phase) { ch.phase = n; ch.phaseOut = o; }
        }
        if (!ch.mag || !ch.phase) continue;
        {
            _vx_node *mm = nullptr, *mp = nullptr;
            if (g->allowMerge && sobel_merge_targets(g, ch.sobel, &mm, &mp)) {
                if (mm != ch.mag || mp != ch.phase) continue;
                ch.merged = true;
            }
        }
        for (_vx_node *n : consumers(g, in)) {
            _vx_image *si = nullptr, *so = nullptr;
            if (n->fused || used.count(n) || !scale_io(n, &si, &so) || si != in || !so) continue;
            if (so->format == VX_DF_IMAGE_U8 && so->width == in->width / 2 && so->height == in->height / 2) {
                ch.scale = n;
                ch.half = so;
                break;
            }
        }
        for (_vx_node *n : {ch.gauss, ch.sobel, ch.mag, ch.phase, ch.scale}) if (n) used.insert(n);
        fc.chains.push_back(ch);
    }
    if (fc.chains.empty()) return nullptr;
    return new FusedChains(fc);
}

static vx_status chains_prepare(_vx_graph *g)
{
    FusedChains *f = g->chains;
    f->images.clear();
    for (const ChainNodes &c : f->chains) {
        mvx_chain_image_t im{};
        im.src = c.in->plane_ptr(0); im.width = c.in->width; im.height = c.in->height; im.srcStride = c.in->planes[0].stride;
        if (!c.merged) {
            im.gx = (int16_t *)c.gx->plane_ptr(0); im.gxStride = c.gx->planes[0].stride;
            im.gy = (int16_t *)c.gy->plane_ptr(0); im.gyStride = c.gy->planes[0].stride;
        }
        im.mag = (int16_t *)c.magOut->plane_ptr(0); im.magStride = c.magOut->planes[0].stride;
        im.phase = c.phaseOut->plane_ptr(0); im.phaseStride = c.phaseOut->planes[0].stride;
        if (c.half) { im.half = c.half->plane_ptr(0); im.halfStride = c.half->planes[0].stride; }
        f->images.push_back(im);
    }
    const uint32_t n = (uint32_t)f->images.size();
    vx_status s = f->ws.alloc(mvx_ChainBatch_workspace_size(n));
    if (s) return s;
    // the autotuner times every tile shape and grid on these very images (MVX_CHAIN_AUTOTUNE=0: the
    // default 64x32 tile at full occupancy)
    const char *tune = getenv("MVX_CHAIN_AUTOTUNE");
    int r = (tune && tune[0] == '0')
                ? mvx_ChainBatch_prepare(f->images.data(), n, 4, 0, f->ws.ptr(), f->ws.size, &f->plan)
                : mvx_ChainBatch_autotune(g->stream, f->images.data(), n, 2, f->ws.ptr(), f->ws.size, &f->plan, nullptr);
    if (!r) r = mvx_ChainBatch_graph_create(&f->plan, &f->graphExec);
    if (r) {
        log(g, r, "ERROR: vxVerifyGraph: chain batch supernode prepare returned %d\n", r);
        return r;
    }
    log(g, VX_SUCCESS, "INFO: vxVerifyGraph: chain batch supernode: %u chains (gaussian_3x3, sobel_3x3, magnitude, phase, "
                       "scale_image) in one launch, tile %ux%u, %u workgroups\n", n, f->plan.tileWidth, f->plan.tileHeight,
        f->plan.workgroups);
    return VX_SUCCESS;
}

// ------------------------------------------------------------------------------ verification -----

static const char *short_name(const KernelDef *k)
{
    const char *p = strrchr(k->name, '.');
    return p ? p + 1 : k->name;
}

namespace mvxe {

vx_status verify_graph(_vx_graph *g)
{
    invalidate_graph(g);
    _vx_context *ctx = g->context;
    // 1. parameters present and of the kernel's types
    for (_vx_node *n : g->nodes) {
        for (size_t i = 0; i < n->params.size(); i++) {
            const ParamSpec &sp = n->kernel->params[i];
            _vx_reference *r = n->params[i];
            if (!r) {
                if (sp.optional) continue;
                log(g, VX_ERROR_NOT_SUFFICIENT, "ERROR: vxVerifyGraph: %s: parameter #%zu missing\n", n->kernel->name, i);
                return VX_ERROR_NOT_SUFFICIENT;
            }
            if (r->type != sp.type) {
                log(g, VX_ERROR_INVALID_TYPE, "ERROR: vxVerifyGraph: %s: parameter #%zu is a %s, expected %s\n", n->kernel->name, i,
                    type_name(r->type), type_name(sp.type));
                return VX_ERROR_INVALID_TYPE;
            }
            if (r->isVirtual && r->scope && r->scope != g) {
                log(g, VX_ERROR_INVALID_SCOPE, "ERROR: vxVerifyGraph: %s: virtual object of another graph\n", n->kernel->name);
                return VX_ERROR_INVALID_SCOPE;
            }
        }
    }
    // 2. single writer per object, then topological order (Kahn)
    std::map<_vx_reference *, _vx_node *> writer;
    for (_vx_node *n : g->nodes)
        for (size_t i = 0; i < n->params.size(); i++)
            if (n->params[i] && n->kernel->params[i].dir != VX_INPUT) {
                _vx_reference *d = root_of(n->params[i]);
                if (writer.count(d) && writer[d] != n) {
                    log(g, VX_ERROR_MULTIPLE_WRITERS, "ERROR: vxVerifyGraph: two nodes write %s\n", d->name.empty() ? "an object" : d->name.c_str());
                    return VX_ERROR_MULTIPLE_WRITERS;
                }
                writer[d] = n;
            }
    std::map<_vx_node *, int> indeg;
    std::map<_vx_node *, std::vector<_vx_node *>> next;
    for (_vx_node *n : g->nodes) indeg[n] = 0;
    for (_vx_node *n : g->nodes) {
        std::set<_vx_node *> deps;
        for (size_t i = 0; i < n->params.size(); i++)
            if (n->params[i] && n->kernel->params[i].dir != VX_OUTPUT) {
                auto it = writer.find(root_of(n->params[i]));
                if (it != writer.end() && it->second != n) deps.insert(it->second);
            }
        for (_vx_node *d : deps) { next[d].push_back(n); indeg[n]++; }
    }
    std::vector<_vx_node *> ready;
    for (_vx_node *n : g->nodes) if (!indeg[n]) ready.push_back(n);
    while (!ready.empty()) {
        _vx_node *n = ready.front();
        ready.erase(ready.begin());
        g->order.push_back(n);
        for (_vx_node *m : next[n]) if (--indeg[m] == 0) ready.push_back(m);
    }
    if (g->order.size() != g->nodes.size()) {
        g->order.clear();
        log(g, VX_ERROR_INVALID_GRAPH, "ERROR: vxVerifyGraph: the graph has a cycle\n");
        return VX_ERROR_INVALID_GRAPH;
    }
    // virtual data must be written before it is read
    for (_vx_node *n : g->nodes)
        for (size_t i = 0; i < n->params.size(); i++) {
            _vx_reference *r = n->params[i];
            if (r && r->isVirtual && n->kernel->params[i].dir == VX_INPUT && !writer.count(root_of(r))) {
                log(g, VX_ERROR_INVALID_GRAPH, "ERROR: vxVerifyGraph: %s reads a virtual object no node writes\n", n->kernel->name);
                return VX_ERROR_INVALID_GRAPH;
            }
        }
    // 3. validate in execution order (virtual outputs receive their meta data before their readers)
    for (_vx_node *n : g->order) {
        vx_status s = n->kernel->validate(n);
        n->status = s;
        if (s) return s;
    }
    // 4. target affinity: node > graph > context > AGO_DEFAULT_TARGET (read into the context) > GPU,
    //    against what each kernel's query_target_support answers (ago_drama_alloc.cpp:565-727).  Every
    //    built-in kernel is a GPU kernel, so a CPU request falls back with a warning; a user kernel runs
    //    where it says it can.
    for (_vx_node *n : g->order) {
        vx_status ts = VX_SUCCESS;
        const uint32_t supported = target_support(n, &ts);
        if (ts) return ts;
        if (!(supported & (AGO_TARGET_AFFINITY_CPU | AGO_TARGET_AFFINITY_GPU))) {
            log(n, VX_FAILURE, "ERROR: kernel %s not supported yet\n", n->kernel->name);
            return VX_FAILURE;
        }
        uint32_t want = n->affinity.device_type ? n->affinity.device_type
                      : g->affinity.device_type ? g->affinity.device_type
                      : ctx->affinity.device_type ? ctx->affinity.device_type : AGO_TARGET_AFFINITY_GPU;
        if (want & supported) {
            n->target = want & AGO_TARGET_AFFINITY_GPU ? AGO_TARGET_AFFINITY_GPU : AGO_TARGET_AFFINITY_CPU;
        } else {
            n->target = supported & AGO_TARGET_AFFINITY_GPU ? AGO_TARGET_AFFINITY_GPU : AGO_TARGET_AFFINITY_CPU;
            log(n, VX_SUCCESS, "WARNING: kernel %s not supported on %s -- falling back to %s\n", n->kernel->name,
                want == AGO_TARGET_AFFINITY_CPU ? "CPU" : "GPU", n->target == AGO_TARGET_AFFINITY_GPU ? "GPU" : "CPU");
        }
    }
    propagate_valid_rects(g);
    // 5. fusion.  The node-merge rules change which reference kernel runs (and so its bits); they follow
    //    the reference's NO_NODE_MERGE flag only.  The supernodes compute exactly what their nodes would
    //    and are turned off by either flag or MVX_NO_FUSION.
    g->allowMerge = !(g->optimizerFlags & AGO_GRAPH_OPTIMIZER_FLAG_NO_NODE_MERGE);
    const bool noFusion = (g->optimizerFlags & (AGO_GRAPH_OPTIMIZER_FLAG_NO_NODE_MERGE | AGO_GRAPH_OPTIMIZER_FLAG_NO_SUPERNODE_MERGE)) ||
                          getenv("MVX_NO_FUSION");
    if (!noFusion && (g->fused = find_fusion(g))) {
        FusedVision *f = g->fused;
        for (_vx_node *n : {f->cc, f->ex, f->scale, f->pyr, f->canny, f->harris}) n->fused = FUSED_VISION;
        for (size_t i = 0; i < g->order.size(); i++)
            if (g->order[i] == f->cc) g->fusedAt = (int)i;
        log(g, VX_SUCCESS, "INFO: vxVerifyGraph: fused vision supernode (color_convert, channel_extract, scale_image, "
                           "gaussian_pyramid, canny_edge_detector, harris_corners) %ux%u\n", f->params.width, f->params.height);
    }
    if (!noFusion && (g->chains = find_chains(g))) {
        for (const ChainNodes &c : g->chains->chains)
            for (_vx_node *n : {c.gauss, c.sobel, c.mag, c.phase, c.scale})
                if (n) n->fused = FUSED_CHAIN;
        for (size_t i = 0; i < g->order.size() && g->chainsAt < 0; i++)
            if (g->order[i]->fused == FUSED_CHAIN) g->chainsAt = (int)i;
    }
    // virtual images that no kernel touches: the gradients of merged Sobel nodes and of merged chains
    std::set<_vx_reference *> unallocated;
    if (g->chains)
        for (const ChainNodes &c : g->chains->chains)
            if (c.merged) { unallocated.insert(c.gx); unallocated.insert(c.gy); }
    if (g->allowMerge) {
        for (_vx_node *n : g->order) {
            _vx_node *mm = nullptr, *mp = nullptr;
            if (n->fused || !sobel_merge_targets(g, n, &mm, &mp)) continue;
            if ((mm && mm->fused) || (mp && mp->fused)) continue;
            n->mergedMag = mm;
            n->mergedPhase = mp;
            for (_vx_node *m : {mm, mp}) if (m) m->fused = FUSED_MERGE;
            _vx_image *in = nullptr, *gx = nullptr, *gy = nullptr;
            sobel_io(n, &in, &gx, &gy);
            unallocated.insert(gx);
            unallocated.insert(gy);
            log(g, VX_SUCCESS, "INFO: vxVerifyGraph: merged %s%s%s into %s\n", short_name(n->kernel), mm ? " + magnitude" : "",
                mp ? " + phase" : "", mm && mp ? "SobelMagnitudePhase_S16U8_U8_3x3" : mm ? "SobelMagnitude_S16_U8_3x3" : "SobelPhase_U8_U8_3x3");
        }
    }
    // 6. device memory: every object a node touches, each node's workspace, one stream
    vx_status s = ctx->ensure_hip();
    if (s) return s;
    if (!g->stream && hipStreamCreate(&g->stream) != hipSuccess) return VX_FAILURE;
    std::set<_vx_reference *> seen;
    for (_vx_node *n : g->order) {
        for (size_t i = 0; i < n->params.size(); i++) {
            _vx_reference *r = n->params[i];
            if (!r || seen.count(r)) continue;
            seen.insert(r);
            const bool written = n->kernel->params[i].dir != VX_INPUT;
            if (r->type == VX_TYPE_IMAGE) {
                _vx_image *im = (_vx_image *)r;
                // the fused supernode keeps the virtual YUV image and luma in LDS: nothing to allocate
                if (g->fused && im->isVirtual && (r == g->fused->cc->params[1] || r == g->fused->ex->params[2])) continue;
                if (im->isVirtual && unallocated.count(r)) continue;
                // the chain batch keeps each blurred image in LDS
                if (g->chains && im->isVirtual && n->fused == FUSED_CHAIN) {
                    bool blur = false;
                    for (const ChainNodes &c : g->chains->chains) blur = blur || r == c.blur;
                    if (blur) continue;
                }
                if ((s = im->allocate())) return s;
            } else if (r->type == VX_TYPE_PYRAMID) {
                for (_vx_image *im : ((_vx_pyramid *)r)->images)
                    if ((s = im->allocate())) return s;
            } else if (r->type == VX_TYPE_ARRAY) {
                if ((s = ((_vx_array *)r)->allocate())) return s;
                if (written) g->writtenArrays.push_back((_vx_array *)r);
            } else if (r->type == VX_TYPE_SCALAR && written) {
                if ((s = ((_vx_scalar *)r)->dev.alloc(8))) return s;
                g->writtenScalars.push_back((_vx_scalar *)r);
            }
        }
        if (n->fused) continue;
        if (n->mergedMag || n->mergedPhase) continue;   // the merged kernels need no workspace
        if (n->kernel->initialize) {
            if ((s = n->kernel->initialize(n))) return s;
            if (n->localSize && (s = n->local.alloc(n->localSize))) return s;
            if ((s = prepare_node(n, g->stream))) return s;
        }
    }
    if (g->fused && (s = fused_prepare(g))) return s;
    if (g->chains && (s = chains_prepare(g))) return s;
    if (hipStreamSynchronize(g->stream) != hipSuccess) return VX_FAILURE;
    g->verified = true;
    g->state = VX_GRAPH_STATE_VERIFIED;
    if (getenv("MVX_VERBOSE")) {
        for (_vx_node *n : g->order)
            log(g, VX_SUCCESS, "  node %s%s\n", short_name(n->kernel),
                n->fused == FUSED_VISION ? " (fused vision supernode)" : n->fused == FUSED_CHAIN ? " (chain batch supernode)"
                : n->fused == FUSED_MERGE ? " (merged into its Sobel node)" : (n->mergedMag || n->mergedPhase) ? " (merged kernel)" : "");
    }
    return VX_SUCCESS;
}

vx_status execute_graph(_vx_graph *g)
{
    if (!g->verified) {
        vx_status s = verify_graph(g);
        if (s) { g->state = VX_GRAPH_STATE_ABANDONED; return s; }
    }
    g->state = VX_GRAPH_STATE_RUNNING;
    const uint64_t t0 = now_ns();
    vx_status s = VX_SUCCESS;
    for (size_t i = 0; i < g->order.size() && !s; i++) {
        _vx_node *n = g->order[i];
        if (n->fused == FUSED_CHAIN) {
            if ((int)i != g->chainsAt) continue;
            const int r = mvx_graph_launch(g->chains->graphExec, g->stream);
            if (r) { log(g, r, "ERROR: vxProcessGraph: chain batch supernode returned %d\n", r); s = r; }
            continue;
        }
        if (n->fused == FUSED_MERGE) continue;
        if (n->fused) {
            if ((int)i != g->fusedAt) continue;
            FusedVision *f = g->fused;
            fused_params(f);  // thresholds and scalars are read at every run
            const int r = mvx_VisionGraph_run_lane(g->stream, f->lane, &f->params, 1, f->ws.ptr(), f->wsSize);
            if (r) { log(g, r, "ERROR: vxProcessGraph: fused vision supernode returned %d\n", r); s = r; }
            continue;
        }
        s = n->kernel->execute(n, g->stream);
        n->status = s;
    }
    if (hipStreamSynchronize(g->stream) != hipSuccess && !s) s = VX_FAILURE;
    // host copies of what kernels wrote: array counts and scalar values
    for (_vx_array *a : g->writtenArrays) if (!s) s = a->pull_count();
    for (_vx_scalar *sc : g->writtenScalars)
        if (!s && hipMemcpy(&sc->value, sc->dev.ptr(), sc->item_size(), hipMemcpyDeviceToHost) != hipSuccess) s = VX_FAILURE;
    const uint64_t t1 = now_ns();
    perf_add(g->perf, t0, t1);
    for (_vx_node *n : g->order) perf_add(n->perf, t0, t1);
    g->state = s ? VX_GRAPH_STATE_ABANDONED : VX_GRAPH_STATE_COMPLETED;
    return s;
}

}  // namespace mvxe

// ------------------------------------------------------------------------------ graph API --------

VX_API_ENTRY vx_graph VX_API_CALL vxCreateGraph(vx_context context)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return nullptr;
    _vx_graph *g = new _vx_graph(context);
    add_ref(context, g);
    return g;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseGraph(vx_graph *graph)
{
    if (!graph || !is_ref(*graph, VX_TYPE_GRAPH)) return VX_ERROR_INVALID_REFERENCE;
    _vx_graph *g = *graph;
    if (g->worker.joinable()) g->worker.join();
    release_ref(g, true);
    *graph = nullptr;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxVerifyGraph(vx_graph graph)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return VX_ERROR_INVALID_REFERENCE;
    std::lock_guard<std::mutex> lk(graph->runLock);
    return verify_graph(graph);
}

VX_API_ENTRY vx_status VX_API_CALL vxProcessGraph(vx_graph graph)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return VX_ERROR_INVALID_REFERENCE;
    if (graph->scheduled) return VX_ERROR_GRAPH_SCHEDULED;
    std::lock_guard<std::mutex> lk(graph->runLock);
    graph->lastStatus = execute_graph(graph);
    return graph->lastStatus;
}

VX_API_ENTRY vx_status VX_API_CALL vxScheduleGraph(vx_graph graph)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return VX_ERROR_INVALID_REFERENCE;
    if (graph->scheduled) return VX_ERROR_GRAPH_SCHEDULED;
    if (graph->worker.joinable()) graph->worker.join();
    graph->scheduled = true;
    // one worker per graph, as the reference's per-graph thread (ago_interface.cpp:30-50, 2822-2875)
    graph->worker = std::thread([graph]() {
        std::lock_guard<std::mutex> lk(graph->runLock);
        graph->lastStatus = execute_graph(graph);
    });
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxWaitGraph(vx_graph graph)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return VX_ERROR_INVALID_REFERENCE;
    if (!graph->scheduled) return VX_FAILURE;
    if (graph->worker.joinable()) graph->worker.join();
    graph->scheduled = false;
    return graph->lastStatus;
}

VX_API_ENTRY vx_bool VX_API_CALL vxIsGraphVerified(vx_graph graph)
{
    return is_ref(graph, VX_TYPE_GRAPH) && graph->verified ? vx_true_e : vx_false_e;
}

#define CHECK_SIZE(T) if (size != sizeof(T)) return VX_ERROR_INVALID_PARAMETERS
#define PUT(T, v) do { CHECK_SIZE(T); *(T *)ptr = (T)(v); } while (0)

VX_API_ENTRY vx_status VX_API_CALL vxQueryGraph(vx_graph graph, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_GRAPH_NUMNODES: PUT(vx_uint32, graph->nodes.size()); break;
    case VX_GRAPH_PERFORMANCE: CHECK_SIZE(vx_perf_t); *(vx_perf_t *)ptr = graph->perf; break;
    case VX_GRAPH_NUMPARAMETERS: PUT(vx_uint32, 0); break;
    case VX_GRAPH_STATE: PUT(vx_enum, graph->state); break;
    case VX_GRAPH_ATTRIBUTE_AMD_AFFINITY: CHECK_SIZE(AgoTargetAffinityInfo); *(AgoTargetAffinityInfo *)ptr = graph->affinity; break;
    case VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS: PUT(vx_uint32, graph->optimizerFlags); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

// Changing what a verified graph runs releases its device state (invalidate_graph).  A graph that a
// vxScheduleGraph worker may be executing is refused; otherwise the change waits for any
// vxProcessGraph / vxVerifyGraph in progress on another thread (runLock).
VX_API_ENTRY vx_status VX_API_CALL vxSetGraphAttribute(vx_graph graph, vx_enum attribute, const void *ptr, vx_size size)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    if (graph->scheduled) return VX_ERROR_GRAPH_SCHEDULED;
    switch (attribute) {
    case VX_GRAPH_ATTRIBUTE_AMD_AFFINITY: {
        CHECK_SIZE(AgoTargetAffinityInfo);
        std::lock_guard<std::mutex> lk(graph->runLock);
        graph->affinity = *(const AgoTargetAffinityInfo *)ptr;
        invalidate_graph(graph);
        break;
    }
    case VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS: {
        CHECK_SIZE(vx_uint32);
        std::lock_guard<std::mutex> lk(graph->runLock);
        graph->optimizerFlags = *(const vx_uint32 *)ptr;
        invalidate_graph(graph);
        break;
    }
    case VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT:
        CHECK_SIZE(AgoGraphImportInfo);
        return import_graph_text(graph, (const AgoGraphImportInfo *)ptr);
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

// ------------------------------------------------------------------------------ node API ---------

VX_API_ENTRY vx_node VX_API_CALL vxCreateGenericNode(vx_graph graph, vx_kernel kernel)
{
    if (!is_ref(graph, VX_TYPE_GRAPH) || !is_ref(kernel, VX_TYPE_KERNEL)) return nullptr;
    if (kernel->def->user && !kernel->def->user->finalized) return nullptr;   // vxFinalizeKernel first
    return create_node(graph, kernel->def, {});
}

VX_API_ENTRY vx_status VX_API_CALL vxSetParameterByIndex(vx_node node, vx_uint32 index, vx_reference value)
{
    if (!is_ref(node, VX_TYPE_NODE)) return VX_ERROR_INVALID_REFERENCE;
    if (index >= node->params.size()) return VX_ERROR_INVALID_PARAMETERS;
    if (value && !is_ref(value)) return VX_ERROR_INVALID_REFERENCE;
    if (value && value->type != node->kernel->params[index].type) return VX_ERROR_INVALID_TYPE;
    _vx_graph *g = node->graph;
    if (g->scheduled) return VX_ERROR_GRAPH_SCHEDULED;
    std::lock_guard<std::mutex> lk(g->runLock);
    // the reference refuses parameter changes on a verified graph (vx_api.cpp:3809-3810)
    if (g->verified) return VX_ERROR_NOT_SUPPORTED;
    if (value) value->intCount++;
    if (node->params[index]) release_ref(node->params[index], false);
    node->params[index] = value;
    invalidate_graph(g);
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryNode(vx_node node, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(node, VX_TYPE_NODE)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_NODE_STATUS: PUT(vx_status, node->status); break;
    case VX_NODE_PERFORMANCE: CHECK_SIZE(vx_perf_t); *(vx_perf_t *)ptr = node->perf; break;
    case VX_NODE_BORDER: CHECK_SIZE(vx_border_t); *(vx_border_t *)ptr = node->border; break;
    case VX_NODE_LOCAL_DATA_SIZE: PUT(vx_size, node->kernel->user ? (node->localHost.empty() ? node->userLocalSize : node->localHost.size())
                                                               : node->localSize); break;
    case VX_NODE_PARAMETERS: PUT(vx_uint32, node->params.size()); break;
    case VX_NODE_ATTRIBUTE_AMD_AFFINITY: CHECK_SIZE(AgoTargetAffinityInfo); *(AgoTargetAffinityInfo *)ptr = node->affinity; break;
    case VX_NODE_ATTRIBUTE_AMD_HIP_STREAM: PUT(hipStream_t, node->graph->stream); break;
    case VX_NODE_LOCAL_DATA_PTR: PUT(void *, node->localPtr); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxSetNodeAttribute(vx_node node, vx_enum attribute, const void *ptr, vx_size size)
{
    if (!is_ref(node, VX_TYPE_NODE)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    // a user kernel's initialize callback hands its local data over while the graph verifies
    if (attribute == VX_NODE_LOCAL_DATA_PTR) {
        CHECK_SIZE(void *);
        node->localPtr = *(void *const *)ptr;
        return VX_SUCCESS;
    }
    if (node->graph->scheduled) return VX_ERROR_GRAPH_SCHEDULED;
    std::lock_guard<std::mutex> lk(node->graph->runLock);
    switch (attribute) {
    case VX_NODE_LOCAL_DATA_SIZE:
        CHECK_SIZE(vx_size);
        if (!node->kernel->user) return VX_ERROR_NOT_SUPPORTED;
        node->userLocalSize = *(const vx_size *)ptr;
        break;
    case VX_NODE_BORDER: CHECK_SIZE(vx_border_t); node->border = *(const vx_border_t *)ptr; break;
    case VX_NODE_ATTRIBUTE_AMD_AFFINITY: CHECK_SIZE(AgoTargetAffinityInfo); node->affinity = *(const AgoTargetAffinityInfo *)ptr; break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    invalidate_graph(node->graph);
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseNode(vx_node *node)
{
    if (!node || !is_ref(*node, VX_TYPE_NODE)) return VX_ERROR_INVALID_REFERENCE;
    release_ref(*node, true);
    *node = nullptr;
    return VX_SUCCESS;
}

VX_API_ENTRY void VX_API_CALL vxRemoveNode(vx_node *node)
{
    if (!node || !is_ref(*node, VX_TYPE_NODE)) return;
    _vx_node *n = *node;
    _vx_graph *g = n->graph;
    if (g->scheduled) return;
    auto it = std::find(g->nodes.begin(), g->nodes.end(), n);
    if (it != g->nodes.end()) {
        std::lock_guard<std::mutex> lk(g->runLock);
        invalidate_graph(g);
        g->nodes.erase(it);
        release_ref(n, false);
    }
    release_ref(n, true);
    *node = nullptr;
}
