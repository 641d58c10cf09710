// vx_objects.cpp -- context, references and data objects of the MI355X OpenVX engine.
//
// Public API of amd_openvx/openvx/api/vx_api.cpp (context :77-460, images :1006-1960, scalars
// :2170-2300, thresholds :5940-6420, arrays :6900-7400, pyramids :7470-7560, remaps :7840-7990,
// matrices :6460-6700) restated over HBM-resident objects.  Storage follows the reference's
// layout rules so that kernels see the same bytes: row stride ALIGN16(width * bytes per pixel)
// (ago_util.cpp:1107), zero-initialised buffers with zero guard bytes on both sides
// (ago_util.cpp:187-200), array count word in front of the items (ago_internal.h:171).
#include <dlfcn.h>

#include "engine.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

using namespace mvxe;

static void destroy(_vx_reference *r);

namespace mvxe {

void log(_vx_reference *ref, vx_status status, const char *fmt, ...)
{
    char msg[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    _vx_context *ctx = ref ? ref->context : nullptr;
    if (ctx && ctx->logCallback)
        ctx->logCallback(ctx, ref, status, msg);
    else if (status != VX_SUCCESS || getenv("MVX_VERBOSE"))
        fputs(msg, stderr);
}

vx_status DevBuf::alloc(size_t bytes)
{
    if (base) return VX_SUCCESS;
    const size_t total = bytes + 2 * GUARD;
    if (hipMalloc((void **)&base, total) != hipSuccess) {
        base = nullptr;
        return VX_ERROR_NO_MEMORY;
    }
    if (hipMemset(base, 0, total) != hipSuccess) {
        (void)hipFree(base);
        base = nullptr;
        return VX_FAILURE;
    }
    size = bytes;
    return VX_SUCCESS;
}

void DevBuf::release()
{
    if (base) (void)hipFree(base);
    base = nullptr;
    size = 0;
}

bool is_ref(const void *p, vx_enum type)
{
    if (!p) return false;
    const _vx_reference *r = (const _vx_reference *)p;
    if (r->magic != MAGIC) return false;
    return type == 0 || r->type == type;
}

void add_ref(_vx_context *ctx, _vx_reference *r)
{
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    ctx->refs.push_back(r);
}

void release_ref(_vx_reference *r, bool external)
{
    if (!r) return;
    if (external) r->extCount--;
    else r->intCount--;
    if (r->extCount <= 0 && r->intCount <= 0) ::destroy(r);
}

_vx_image *image_of(_vx_reference *r) { return is_ref(r, VX_TYPE_IMAGE) ? (_vx_image *)r : nullptr; }

_vx_scalar *make_scalar(_vx_context *ctx, vx_enum type, const void *value)
{
    _vx_scalar *s = new _vx_scalar(ctx, type);
    if (value) memcpy(&s->value, value, s->item_size());
    add_ref(ctx, s);
    return s;
}

}  // namespace mvxe

static void destroy(_vx_reference *r)
{
    _vx_context *ctx = r->context;
    if (r->type == VX_TYPE_CONTEXT) return;  // only vxReleaseContext ends a context
    {
        std::lock_guard<std::recursive_mutex> g(ctx->lock);
        for (size_t i = 0; i < ctx->refs.size(); i++)
            if (ctx->refs[i] == r) { ctx->refs.erase(ctx->refs.begin() + i); break; }
    }
    switch (r->type) {
    case VX_TYPE_IMAGE: ((_vx_image *)r)->dev.release(); break;
    case VX_TYPE_SCALAR: ((_vx_scalar *)r)->dev.release(); break;
    case VX_TYPE_ARRAY: ((_vx_array *)r)->dev.release(); break;
    case VX_TYPE_REMAP: ((_vx_remap *)r)->devFloat.release(); ((_vx_remap *)r)->devFixed.release(); break;
    case VX_TYPE_PYRAMID:
        for (_vx_image *img : ((_vx_pyramid *)r)->images) release_ref(img, false);
        break;
    case VX_TYPE_NODE: {
        _vx_node *n = (_vx_node *)r;
        mvxe::deinit_user_node(n);
        for (_vx_reference *p : n->params) release_ref(p, false);
        n->local.release();
        break;
    }
    case VX_TYPE_GRAPH: {
        _vx_graph *g = (_vx_graph *)r;
        if (g->worker.joinable()) g->worker.join();
        mvxe::invalidate_graph(g);
        std::vector<_vx_node *> nodes = g->nodes;
        g->nodes.clear();
        for (_vx_node *n : nodes) release_ref(n, false);
        if (g->stream) (void)hipStreamDestroy(g->stream);
        break;
    }
    default: break;
    }
    r->magic = 0;
    delete r;
}

// ------------------------------------------------------------------------------ context ----------

vx_status _vx_context::ensure_hip()
{
    if (hipReady) return VX_SUCCESS;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        log(this, VX_ERROR_NOT_SUPPORTED, "ERROR: no HIP device: this engine runs its kernels on an MI355X (gfx950) GPU\n");
        return VX_ERROR_NOT_SUPPORTED;
    }
    int dev = affinity.device_type == AGO_TARGET_AFFINITY_GPU ? (int)affinity.device_info : 0;
    if (hipDevice >= 0) dev = hipDevice;
    if (dev >= n) dev = 0;
    if (hipSetDevice(dev) != hipSuccess) return VX_FAILURE;
    hipDevice = dev;
    hipReady = true;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_context VX_API_CALL vxCreateContext(void)
{
    _vx_context *ctx = new _vx_context();
    // AGO_DEFAULT_TARGET (ago_drama_alloc.cpp) selects the default affinity; this engine runs GPU only
    if (const char *t = getenv("AGO_DEFAULT_TARGET")) {
        if (!strcmp(t, "CPU")) ctx->affinity.device_type = AGO_TARGET_AFFINITY_CPU;
        else if (!strcmp(t, "GPU")) ctx->affinity.device_type = AGO_TARGET_AFFINITY_GPU;
    }
    for (const KernelDef &k : kernel_table()) {
        _vx_kernel *kr = new _vx_kernel(ctx, &k);
        kr->name = k.name;
        kr->extCount = 0;
        kr->intCount = 1;
        ctx->kernels.push_back(kr);
    }
    return ctx;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseContext(vx_context *context)
{
    if (!context || !is_ref(*context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    _vx_context *ctx = *context;
    // holders before what they hold: graphs, nodes, parameters, pyramids, then everything else
    const vx_enum passes[] = {VX_TYPE_GRAPH, VX_TYPE_NODE, VX_TYPE_PARAMETER, VX_TYPE_PYRAMID, 0};
    for (vx_enum passType : passes) {
        std::vector<_vx_reference *> refs;
        {
            std::lock_guard<std::recursive_mutex> g(ctx->lock);
            refs = ctx->refs;
        }
        for (_vx_reference *r : refs) {
            if (passType && r->type != passType) continue;
            bool alive = false;
            {
                std::lock_guard<std::recursive_mutex> g(ctx->lock);
                for (_vx_reference *x : ctx->refs) alive |= (x == r);
            }
            if (!alive) continue;
            r->extCount = 0;
            r->intCount = 0;
            destroy(r);
        }
    }
    for (_vx_kernel *k : ctx->kernels) { k->magic = 0; delete k; }
    for (_vx_kernel *k : ctx->removedKernels) { k->magic = 0; delete k; }
    for (const ModuleInfo &m : ctx->modules)
        if (m.handle) dlclose(m.handle);
    ctx->magic = 0;
    delete ctx;
    *context = nullptr;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_context VX_API_CALL vxGetContext(vx_reference reference)
{
    return is_ref(reference) ? reference->context : nullptr;
}

VX_API_ENTRY vx_status VX_API_CALL vxGetStatus(vx_reference reference)
{
    return is_ref(reference) ? VX_SUCCESS : VX_ERROR_INVALID_REFERENCE;
}

#define CHECK_SIZE(T) if (size != sizeof(T)) return VX_ERROR_INVALID_PARAMETERS
#define PUT(T, v) do { CHECK_SIZE(T); *(T *)ptr = (T)(v); } while (0)

VX_API_ENTRY vx_status VX_API_CALL vxQueryContext(vx_context context, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_CONTEXT_VENDOR_ID: PUT(vx_uint16, VX_ID_AMD); break;
    case VX_CONTEXT_VERSION: PUT(vx_uint16, VX_VERSION); break;
    case VX_CONTEXT_UNIQUE_KERNELS: PUT(vx_uint32, context->kernels.size()); break;
    case VX_CONTEXT_MODULES: {
        vx_uint32 n = 0;
        for (const ModuleInfo &m : context->modules) n += m.path.empty() ? 0 : 1;
        PUT(vx_uint32, n);
        break;
    }
    case VX_CONTEXT_REFERENCES: PUT(vx_uint32, context->refs.size()); break;
    case VX_CONTEXT_OPTICAL_FLOW_MAX_WINDOW_DIMENSION: PUT(vx_size, 9); break;
    case VX_CONTEXT_IMPLEMENTATION:
        if (size < VX_MAX_IMPLEMENTATION_NAME) return VX_ERROR_INVALID_PARAMETERS;
        snprintf((char *)ptr, size, "mivisionx-mi355x %s", mvx_version());
        break;
    case VX_CONTEXT_IMMEDIATE_BORDER: CHECK_SIZE(vx_border_t); *(vx_border_t *)ptr = context->immediateBorder; break;
    case VX_CONTEXT_ATTRIBUTE_AMD_AFFINITY: CHECK_SIZE(AgoTargetAffinityInfo); *(AgoTargetAffinityInfo *)ptr = context->affinity; break;
    case VX_CONTEXT_ATTRIBUTE_AMD_HIP_DEVICE: {
        vx_status s = context->ensure_hip();
        if (s) return s;
        PUT(vx_int32, context->hipDevice);
        break;
    }
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxSetContextAttribute(vx_context context, vx_enum attribute, const void *ptr, vx_size size)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_CONTEXT_IMMEDIATE_BORDER: CHECK_SIZE(vx_border_t); context->immediateBorder = *(const vx_border_t *)ptr; break;
    case VX_CONTEXT_ATTRIBUTE_AMD_AFFINITY: CHECK_SIZE(AgoTargetAffinityInfo); context->affinity = *(const AgoTargetAffinityInfo *)ptr; break;
    case VX_CONTEXT_ATTRIBUTE_AMD_HIP_DEVICE:
        CHECK_SIZE(vx_int32);
        if (context->hipReady) return VX_ERROR_NOT_SUPPORTED;  // objects may already live on the first device
        context->hipDevice = *(const vx_int32 *)ptr;
        break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY void VX_API_CALL vxRegisterLogCallback(vx_context context, vx_log_callback_f callback, vx_bool)
{
    if (is_ref(context, VX_TYPE_CONTEXT)) context->logCallback = callback;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryReference(vx_reference ref, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(ref)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_REFERENCE_COUNT: PUT(vx_uint32, ref->extCount.load()); break;
    case VX_REFERENCE_TYPE: PUT(vx_enum, ref->type); break;
    case VX_REFERENCE_NAME: CHECK_SIZE(vx_char *); *(const vx_char **)ptr = ref->name.c_str(); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxSetReferenceName(vx_reference ref, const vx_char *name)
{
    if (!is_ref(ref)) return VX_ERROR_INVALID_REFERENCE;
    ref->name = name ? name : "";
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxRetainReference(vx_reference ref)
{
    if (!is_ref(ref)) return VX_ERROR_INVALID_REFERENCE;
    ref->extCount++;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseReference(vx_reference *ref_ptr)
{
    if (!ref_ptr || !is_ref(*ref_ptr)) return VX_ERROR_INVALID_REFERENCE;
    if ((*ref_ptr)->type == VX_TYPE_CONTEXT) return vxReleaseContext((vx_context *)ref_ptr);
    if ((*ref_ptr)->type == VX_TYPE_KERNEL) { *ref_ptr = nullptr; return VX_SUCCESS; }
    release_ref(*ref_ptr, true);
    *ref_ptr = nullptr;
    return VX_SUCCESS;
}

template <typename T> static vx_status release_typed(T **p, vx_enum type)
{
    if (!p || !is_ref(*p, type)) return VX_ERROR_INVALID_REFERENCE;
    release_ref(*p, true);
    *p = nullptr;
    return VX_SUCCESS;
}

// ------------------------------------------------------------------------------ kernels ----------

VX_API_ENTRY vx_kernel VX_API_CALL vxGetKernelByName(vx_context context, const vx_char *name)
{
    if (!is_ref(context, VX_TYPE_CONTEXT) || !name) return nullptr;
    for (_vx_kernel *k : context->kernels)
        if (!strcmp(k->def->name, name)) return k;
    log(context, VX_ERROR_INVALID_PARAMETERS, "ERROR: vxGetKernelByName: kernel not supported: %s\n", name);
    return nullptr;
}

VX_API_ENTRY vx_kernel VX_API_CALL vxGetKernelByEnum(vx_context context, vx_enum kernel_enum)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return nullptr;
    for (_vx_kernel *k : context->kernels)
        if (k->def->id == kernel_enum) return k;
    return nullptr;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryKernel(vx_kernel kernel, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(kernel, VX_TYPE_KERNEL)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_KERNEL_PARAMETERS: PUT(vx_uint32, kernel->def->params.size()); break;
    case VX_KERNEL_NAME:
        if (size < strlen(kernel->def->name) + 1) return VX_ERROR_INVALID_PARAMETERS;
        strcpy((char *)ptr, kernel->def->name);
        break;
    case VX_KERNEL_ENUM: PUT(vx_enum, kernel->def->id); break;
    case VX_KERNEL_LOCAL_DATA_SIZE: PUT(vx_size, kernel->def->user ? kernel->def->user->localDataSize : 0); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseKernel(vx_kernel *kernel)
{
    if (!kernel || !is_ref(*kernel, VX_TYPE_KERNEL)) return VX_ERROR_INVALID_REFERENCE;
    *kernel = nullptr;  // kernels live as long as their context
    return VX_SUCCESS;
}

VX_API_ENTRY vx_parameter VX_API_CALL vxGetKernelParameterByIndex(vx_kernel kernel, vx_uint32 index)
{
    if (!is_ref(kernel, VX_TYPE_KERNEL) || index >= kernel->def->params.size()) return nullptr;
    _vx_parameter *p = new _vx_parameter(kernel->context);
    p->kernel = kernel;
    p->node = nullptr;
    p->index = index;
    add_ref(kernel->context, p);
    return p;
}

VX_API_ENTRY vx_parameter VX_API_CALL vxGetParameterByIndex(vx_node node, vx_uint32 index)
{
    if (!is_ref(node, VX_TYPE_NODE) || index >= node->kernel->params.size()) return nullptr;
    _vx_parameter *p = new _vx_parameter(node->context);
    p->kernel = nullptr;
    for (_vx_kernel *k : node->context->kernels)
        if (k->def == node->kernel) p->kernel = k;
    p->node = node;
    p->index = index;
    add_ref(node->context, p);
    return p;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryParameter(vx_parameter param, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(param, VX_TYPE_PARAMETER)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    const ParamSpec &spec = param->kernel->def->params[param->index];
    switch (attribute) {
    case VX_PARAMETER_INDEX: PUT(vx_uint32, param->index); break;
    case VX_PARAMETER_DIRECTION: PUT(vx_enum, spec.dir); break;
    case VX_PARAMETER_TYPE: PUT(vx_enum, spec.type); break;
    case VX_PARAMETER_STATE: PUT(vx_enum, spec.optional ? VX_PARAMETER_STATE_OPTIONAL : VX_PARAMETER_STATE_REQUIRED); break;
    case VX_PARAMETER_REF: {
        CHECK_SIZE(vx_reference);
        _vx_reference *r = param->node ? param->node->params[param->index] : nullptr;
        if (r) r->extCount++;
        *(vx_reference *)ptr = r;
        break;
    }
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseParameter(vx_parameter *param) { return release_typed(param, VX_TYPE_PARAMETER); }

// ------------------------------------------------------------------------------ images -----------

bool _vx_image::layout()
{
    planes.clear();
    auto add = [&](uint32_t w, uint32_t h, uint32_t bpp) { planes.push_back(ImagePlane{w, h, bpp, align16(w * bpp), 0}); };
    const uint32_t w = width, h = height, cw = (width + 1) / 2, ch = (height + 1) / 2;
    switch (format) {
    case VX_DF_IMAGE_U8: add(w, h, 1); break;
    case VX_DF_IMAGE_U16: case VX_DF_IMAGE_S16: add(w, h, 2); break;
    case VX_DF_IMAGE_U32: case VX_DF_IMAGE_S32: add(w, h, 4); break;
    case VX_DF_IMAGE_RGB: add(w, h, 3); break;
    case VX_DF_IMAGE_RGBX: add(w, h, 4); break;
    case VX_DF_IMAGE_UYVY: case VX_DF_IMAGE_YUYV: add(w, h, 2); break;
    case VX_DF_IMAGE_IYUV: add(w, h, 1); add(cw, ch, 1); add(cw, ch, 1); break;
    case VX_DF_IMAGE_NV12: case VX_DF_IMAGE_NV21: add(w, h, 1); add(cw, ch, 2); break;
    case VX_DF_IMAGE_YUV4: add(w, h, 1); add(w, h, 1); add(w, h, 1); break;
    default: return false;
    }
    size_t off = 0;
    for (ImagePlane &p : planes) {
        p.offset = off;
        off = align256(off + (size_t)p.stride * p.height);
    }
    return true;
}

size_t _vx_image::bytes() const
{
    if (planes.empty()) return 0;
    const ImagePlane &p = planes.back();
    return align256(p.offset + (size_t)p.stride * p.height);
}

static void fill_pattern(const _vx_image *img, uint32_t plane, std::vector<uint8_t> &host)
{
    const ImagePlane &p = img->planes[plane];
    host.assign((size_t)p.stride * p.height, 0);
    uint8_t px[4] = {0, 0, 0, 0};
    const vx_pixel_value_t &v = img->uniformValue;
    switch (img->format) {
    case VX_DF_IMAGE_U8: px[0] = v.U8; break;
    case VX_DF_IMAGE_U16: memcpy(px, &v.U16, 2); break;
    case VX_DF_IMAGE_S16: memcpy(px, &v.S16, 2); break;
    case VX_DF_IMAGE_U32: memcpy(px, &v.U32, 4); break;
    case VX_DF_IMAGE_S32: memcpy(px, &v.S32, 4); break;
    case VX_DF_IMAGE_RGB: memcpy(px, v.RGB, 3); break;
    case VX_DF_IMAGE_RGBX: memcpy(px, v.RGBX, 4); break;
    case VX_DF_IMAGE_IYUV: case VX_DF_IMAGE_YUV4: px[0] = v.YUV[plane]; break;
    case VX_DF_IMAGE_NV12: if (plane == 0) px[0] = v.YUV[0]; else { px[0] = v.YUV[1]; px[1] = v.YUV[2]; } break;
    case VX_DF_IMAGE_NV21: if (plane == 0) px[0] = v.YUV[0]; else { px[0] = v.YUV[2]; px[1] = v.YUV[1]; } break;
    case VX_DF_IMAGE_UYVY: px[0] = v.YUV[1]; px[1] = v.YUV[0]; break;  // pairs alternate U/V; one value each
    case VX_DF_IMAGE_YUYV: px[0] = v.YUV[0]; px[1] = v.YUV[1]; break;
    default: break;
    }
    for (uint32_t y = 0; y < p.height; y++)
        for (uint32_t x = 0; x < p.width; x++)
            memcpy(&host[(size_t)y * p.stride + (size_t)x * p.bpp], px, p.bpp);
}

vx_status _vx_image::allocate()
{
    if (parent && !dev.base && !external) {
        // pyramid levels are allocated individually, each with its own guard bytes
    }
    if (external || dev.base) return VX_SUCCESS;
    if (planes.empty() && !layout()) return VX_ERROR_INVALID_FORMAT;
    if (width == 0 || height == 0) return VX_ERROR_INVALID_DIMENSION;
    vx_status s = context->ensure_hip();
    if (s) return s;
    if ((s = dev.alloc(bytes())) != VX_SUCCESS) return s;
    if (uniform) {
        std::vector<uint8_t> host;
        for (uint32_t p = 0; p < planes.size(); p++) {
            fill_pattern(this, p, host);
            if (hipMemcpy(plane_ptr(p), host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) return VX_FAILURE;
        }
    }
    return VX_SUCCESS;
}

uint8_t *_vx_image::plane_ptr(uint32_t p)
{
    if (p >= planes.size()) return nullptr;
    uint8_t *base = external ? external : dev.ptr();
    return base ? base + planes[p].offset : nullptr;
}

static _vx_image *new_image(_vx_context *ctx, uint32_t w, uint32_t h, vx_df_image fmt, bool isVirtual, _vx_graph *scope)
{
    _vx_image *img = new _vx_image(ctx);
    img->width = w;
    img->height = h;
    img->format = fmt;
    img->isVirtual = isVirtual;
    img->scope = scope;
    if (fmt != VX_DF_IMAGE_VIRT && !img->layout()) {
        delete img;
        return nullptr;
    }
    add_ref(ctx, img);
    return img;
}

VX_API_ENTRY vx_image VX_API_CALL vxCreateImage(vx_context context, vx_uint32 width, vx_uint32 height, vx_df_image color)
{
    if (!is_ref(context, VX_TYPE_CONTEXT) || width == 0 || height == 0 || color == VX_DF_IMAGE_VIRT) return nullptr;
    return new_image(context, width, height, color, false, nullptr);
}

VX_API_ENTRY vx_image VX_API_CALL vxCreateVirtualImage(vx_graph graph, vx_uint32 width, vx_uint32 height, vx_df_image color)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return nullptr;
    return new_image(graph->context, width, height, color, true, graph);
}

VX_API_ENTRY vx_image VX_API_CALL vxCreateUniformImage(vx_context context, vx_uint32 width, vx_uint32 height, vx_df_image color,
                                                       const vx_pixel_value_t *value)
{
    if (!value) return nullptr;
    _vx_image *img = vxCreateImage(context, width, height, color);
    if (!img) return nullptr;
    img->uniform = true;
    img->uniformValue = *value;
    return img;
}

VX_API_ENTRY vx_image VX_API_CALL vxCreateImageFromHandle(vx_context context, vx_df_image color, const vx_imagepatch_addressing_t addrs[],
                                                          void *const ptrs[], vx_enum memory_type)
{
    if (!is_ref(context, VX_TYPE_CONTEXT) || !addrs || !ptrs) return nullptr;
    _vx_image *img = vxCreateImage(context, addrs[0].dim_x, addrs[0].dim_y, color);
    if (!img) return nullptr;
    if (memory_type == VX_MEMORY_TYPE_HIP) {
        // zero-copy: the caller's device planes become the image; every plane must sit at the
        // offset this layout expects from plane 0, with a stride that is a multiple of 16
        for (uint32_t p = 0; p < img->planes.size(); p++) {
            if (addrs[p].stride_y <= 0 || (addrs[p].stride_y & 15) || addrs[p].stride_x != (vx_int32)img->planes[p].bpp ||
                (uint32_t)addrs[p].stride_y < img->planes[p].width * img->planes[p].bpp || ((uintptr_t)ptrs[p] & 15)) {
                log(img, VX_ERROR_INVALID_PARAMETERS, "ERROR: vxCreateImageFromHandle: plane %u stride/alignment not supported\n", p);
                release_ref(img, true);
                return nullptr;
            }
            img->planes[p].stride = (uint32_t)addrs[p].stride_y;
            img->planes[p].offset = (size_t)((uint8_t *)ptrs[p] - (uint8_t *)ptrs[0]);
        }
        img->external = (uint8_t *)ptrs[0];
        return img;
    }
    if (memory_type != VX_MEMORY_TYPE_HOST) {
        release_ref(img, true);
        return nullptr;
    }
    // host handle: the pixels are copied into HBM once, here
    for (uint32_t p = 0; p < img->planes.size(); p++) {
        vx_rectangle_t rect{0, 0, img->width, img->height};
        if (vxCopyImagePatch(img, &rect, p, &addrs[p], ptrs[p], VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST) != VX_SUCCESS) {
            release_ref(img, true);
            return nullptr;
        }
    }
    return img;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryImage(vx_image image, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(image, VX_TYPE_IMAGE)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_IMAGE_WIDTH: PUT(vx_uint32, image->width); break;
    case VX_IMAGE_HEIGHT: PUT(vx_uint32, image->height); break;
    case VX_IMAGE_FORMAT: PUT(vx_df_image, image->format); break;
    case VX_IMAGE_PLANES: PUT(vx_size, image->planes.size()); break;
    case VX_IMAGE_SPACE: PUT(vx_enum, image->space); break;
    case VX_IMAGE_RANGE: PUT(vx_enum, image->range); break;
    case VX_IMAGE_MEMORY_TYPE: PUT(vx_enum, image->external ? (vx_enum)VX_MEMORY_TYPE_HIP : (vx_enum)VX_MEMORY_TYPE_NONE); break;
    case VX_IMAGE_IS_UNIFORM: PUT(vx_bool, image->uniform ? vx_true_e : vx_false_e); break;
    case VX_IMAGE_UNIFORM_VALUE: CHECK_SIZE(vx_pixel_value_t); *(vx_pixel_value_t *)ptr = image->uniformValue; break;
    case VX_IMAGE_ATTRIBUTE_AMD_GPU_BUFFER_STRIDE:
        if (image->planes.empty()) return VX_ERROR_NOT_ALLOCATED;
        PUT(vx_uint32, image->planes[0].stride);
        break;
    case VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER: {
        CHECK_SIZE(void *);
        vx_status s = image->allocate();
        if (s) return s;
        *(void **)ptr = image->plane_ptr(0);
        break;
    }
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxSetImageAttribute(vx_image image, vx_enum attribute, const void *ptr, vx_size size)
{
    if (!is_ref(image, VX_TYPE_IMAGE)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_IMAGE_SPACE: CHECK_SIZE(vx_enum); image->space = *(const vx_enum *)ptr; break;
    case VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER:
        // the image now lives in the caller's device buffer (same plane layout as this image's)
        CHECK_SIZE(void *);
        if (((uintptr_t) * (void *const *)ptr) & 15) return VX_ERROR_INVALID_PARAMETERS;
        image->dev.release();
        image->external = *(uint8_t *const *)ptr;
        break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseImage(vx_image *image) { return release_typed(image, VX_TYPE_IMAGE); }

static vx_status plane_rect(_vx_image *img, const vx_rectangle_t *rect, uint32_t plane, uint32_t &x0, uint32_t &y0, uint32_t &w,
                            uint32_t &h)
{
    if (plane >= img->planes.size()) return VX_ERROR_INVALID_PARAMETERS;
    vx_rectangle_t r = rect ? *rect : vx_rectangle_t{0, 0, img->width, img->height};
    if (r.start_x >= r.end_x || r.start_y >= r.end_y || r.end_x > img->width || r.end_y > img->height) return VX_ERROR_INVALID_PARAMETERS;
    const ImagePlane &p = img->planes[plane];
    const uint32_t sx = img->width / p.width, sy = img->height / p.height;  // 1 or 2 for chroma planes
    x0 = r.start_x / sx;
    y0 = r.start_y / sy;
    w = (r.end_x - r.start_x + sx - 1) / sx;
    h = (r.end_y - r.start_y + sy - 1) / sy;
    w = std::min(w, p.width - x0);
    h = std::min(h, p.height - y0);
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxCopyImagePatch(vx_image image, const vx_rectangle_t *image_rect, vx_uint32 plane,
                                                   const vx_imagepatch_addressing_t *user_addr, void *user_ptr, vx_enum usage,
                                                   vx_enum user_mem_type)
{
    if (!is_ref(image, VX_TYPE_IMAGE)) return VX_ERROR_INVALID_REFERENCE;
    if (!user_addr || !user_ptr || (usage != VX_READ_ONLY && usage != VX_WRITE_ONLY)) return VX_ERROR_INVALID_PARAMETERS;
    if (user_mem_type != VX_MEMORY_TYPE_HOST && user_mem_type != VX_MEMORY_TYPE_HIP) return VX_ERROR_INVALID_PARAMETERS;
    uint32_t x0, y0, w, h;
    vx_status s = plane_rect(image, image_rect, plane, x0, y0, w, h);
    if (s) return s;
    if ((s = image->allocate())) return s;
    const ImagePlane &p = image->planes[plane];
    if (user_addr->stride_x != (vx_int32)p.bpp || user_addr->stride_y < (vx_int32)(w * p.bpp)) return VX_ERROR_INVALID_PARAMETERS;
    uint8_t *dev = image->plane_ptr(plane) + (size_t)y0 * p.stride + (size_t)x0 * p.bpp;
    const bool hipUser = user_mem_type == VX_MEMORY_TYPE_HIP;
    hipError_t e;
    if (usage == VX_READ_ONLY)
        e = hipMemcpy2D(user_ptr, user_addr->stride_y, dev, p.stride, (size_t)w * p.bpp, h,
                        hipUser ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost);
    else
        e = hipMemcpy2D(dev, p.stride, user_ptr, user_addr->stride_y, (size_t)w * p.bpp, h,
                        hipUser ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
    if (e == hipSuccess && hipUser) e = hipDeviceSynchronize();
    if (usage == VX_WRITE_ONLY) image->uniform = false;
    return e == hipSuccess ? VX_SUCCESS : VX_FAILURE;
}

VX_API_ENTRY vx_status VX_API_CALL vxMapImagePatch(vx_image image, const vx_rectangle_t *rect, vx_uint32 plane, vx_map_id *map_id,
                                                  vx_imagepatch_addressing_t *addr, void **ptr, vx_enum usage, vx_enum mem_type,
                                                  vx_uint32 flags)
{
    if (!is_ref(image, VX_TYPE_IMAGE)) return VX_ERROR_INVALID_REFERENCE;
    if (!map_id || !addr || !ptr) return VX_ERROR_INVALID_PARAMETERS;
    uint32_t x0, y0, w, h;
    vx_status s = plane_rect(image, rect, plane, x0, y0, w, h);
    if (s) return s;
    if ((s = image->allocate())) return s;
    const ImagePlane &p = image->planes[plane];
    ImageMap m;
    m.id = image->nextMap++;
    m.plane = plane;
    m.rect = vx_rectangle_t{x0, y0, x0 + w, y0 + h};
    m.usage = usage;
    m.addr = vx_imagepatch_addressing_t{w, h, (vx_int32)p.bpp, 0, VX_SCALE_UNITY, VX_SCALE_UNITY, 1, 1, 0};
    m.addr.scale_x = VX_SCALE_UNITY * p.width / image->width;
    m.addr.scale_y = VX_SCALE_UNITY * p.height / image->height;
    uint8_t *dev = image->plane_ptr(plane) + (size_t)y0 * p.stride + (size_t)x0 * p.bpp;
    if (mem_type == VX_MEMORY_TYPE_HIP) {
        // device mapping: no copy, the caller works on HBM directly
        if (hipDeviceSynchronize() != hipSuccess) return VX_FAILURE;
        m.addr.stride_y = (vx_int32)p.stride;
        *ptr = dev;
    } else {
        const uint32_t pitch = (flags & VX_NOGAP_X) ? w * p.bpp : p.stride;
        m.addr.stride_y = (vx_int32)pitch;
        m.host.assign((size_t)pitch * h, 0);
        if (usage == VX_READ_ONLY || usage == VX_READ_AND_WRITE)
            if (hipMemcpy2D(m.host.data(), pitch, dev, p.stride, (size_t)w * p.bpp, h, hipMemcpyDeviceToHost) != hipSuccess) return VX_FAILURE;
        *ptr = m.host.data();
    }
    *addr = m.addr;
    *map_id = m.id;
    image->maps.push_back(std::move(m));
    if (usage != VX_READ_ONLY) image->uniform = false;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxUnmapImagePatch(vx_image image, vx_map_id map_id)
{
    if (!is_ref(image, VX_TYPE_IMAGE)) return VX_ERROR_INVALID_REFERENCE;
    for (size_t i = 0; i < image->maps.size(); i++) {
        ImageMap &m = image->maps[i];
        if (m.id != map_id) continue;
        vx_status s = VX_SUCCESS;
        if (!m.host.empty() && (m.usage == VX_WRITE_ONLY || m.usage == VX_READ_AND_WRITE)) {
            const ImagePlane &p = image->planes[m.plane];
            uint8_t *dev = image->plane_ptr(m.plane) + (size_t)m.rect.start_y * p.stride + (size_t)m.rect.start_x * p.bpp;
            if (hipMemcpy2D(dev, p.stride, m.host.data(), m.addr.stride_y, (size_t)(m.rect.end_x - m.rect.start_x) * p.bpp,
                            m.rect.end_y - m.rect.start_y, hipMemcpyHostToDevice) != hipSuccess)
                s = VX_FAILURE;
        }
        image->maps.erase(image->maps.begin() + i);
        return s;
    }
    return VX_ERROR_INVALID_PARAMETERS;
}

// the image's valid rectangle: the whole image, until graph verification propagates the rectangles of
// the kernels that write it (vx_api.cpp:1946; agoComputeImageValidRectangleOutputs)
VX_API_ENTRY vx_status VX_API_CALL vxGetValidRegionImage(vx_image image, vx_rectangle_t *rect)
{
    if (!is_ref(image, VX_TYPE_IMAGE)) return VX_ERROR_INVALID_REFERENCE;
    if (!rect) return VX_ERROR_INVALID_PARAMETERS;
    *rect = image->validRectSet ? image->validRect : vx_rectangle_t{0, 0, image->width, image->height};
    return VX_SUCCESS;
}

// ------------------------------------------------------------------------------ scalars ----------

size_t _vx_scalar::item_size() const
{
    switch (dataType) {
    case VX_TYPE_CHAR: case VX_TYPE_INT8: case VX_TYPE_UINT8: return 1;
    case VX_TYPE_INT16: case VX_TYPE_UINT16: return 2;
    case VX_TYPE_INT32: case VX_TYPE_UINT32: case VX_TYPE_FLOAT32: case VX_TYPE_ENUM: case VX_TYPE_DF_IMAGE: case VX_TYPE_BOOL: return 4;
    case VX_TYPE_INT64: case VX_TYPE_UINT64: case VX_TYPE_FLOAT64: return 8;
    case VX_TYPE_SIZE: return sizeof(vx_size);
    default: return 0;
    }
}

VX_API_ENTRY vx_scalar VX_API_CALL vxCreateScalar(vx_context context, vx_enum data_type, const void *ptr)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return nullptr;
    _vx_scalar probe(context, data_type);
    if (!probe.item_size()) return nullptr;
    probe.magic = 0;
    return make_scalar(context, data_type, ptr);
}

VX_API_ENTRY vx_scalar VX_API_CALL vxCreateVirtualScalar(vx_graph graph, vx_enum data_type)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return nullptr;
    _vx_scalar *s = vxCreateScalar(graph->context, data_type, nullptr);
    if (s) { s->isVirtual = true; s->scope = graph; }
    return s;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryScalar(vx_scalar scalar, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(scalar, VX_TYPE_SCALAR)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    if (attribute != VX_SCALAR_TYPE) return VX_ERROR_NOT_SUPPORTED;
    PUT(vx_enum, scalar->dataType);
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxCopyScalar(vx_scalar scalar, void *user_ptr, vx_enum usage, vx_enum user_mem_type)
{
    if (!is_ref(scalar, VX_TYPE_SCALAR)) return VX_ERROR_INVALID_REFERENCE;
    if (!user_ptr || user_mem_type != VX_MEMORY_TYPE_HOST) return VX_ERROR_INVALID_PARAMETERS;
    // a scalar that kernels write has a device copy, which is current while a graph runs (a CPU user
    // node reads and writes it between GPU nodes)
    const size_t n = scalar->item_size();
    if (usage == VX_READ_ONLY) {
        if (scalar->dev.ptr() && hipMemcpy(&scalar->value, scalar->dev.ptr(), n, hipMemcpyDeviceToHost) != hipSuccess) return VX_FAILURE;
        memcpy(user_ptr, &scalar->value, n);
    } else if (usage == VX_WRITE_ONLY) {
        memcpy(&scalar->value, user_ptr, n);
        if (scalar->dev.ptr() && hipMemcpy(scalar->dev.ptr(), &scalar->value, n, hipMemcpyHostToDevice) != hipSuccess) return VX_FAILURE;
    } else {
        return VX_ERROR_INVALID_PARAMETERS;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseScalar(vx_scalar *scalar) { return release_typed(scalar, VX_TYPE_SCALAR); }

// ------------------------------------------------------------------------------ thresholds -------

VX_API_ENTRY vx_threshold VX_API_CALL vxCreateThreshold(vx_context c, vx_enum thresh_type, vx_enum data_type)
{
    if (!is_ref(c, VX_TYPE_CONTEXT)) return nullptr;
    if (thresh_type != VX_THRESHOLD_TYPE_BINARY && thresh_type != VX_THRESHOLD_TYPE_RANGE) return nullptr;
    if (data_type != VX_TYPE_UINT8 && data_type != VX_TYPE_INT16 && data_type != VX_TYPE_UINT16 && data_type != VX_TYPE_INT32) return nullptr;
    _vx_threshold *t = new _vx_threshold(c, thresh_type);
    t->dataType = data_type;
    t->inputFormat = data_type == VX_TYPE_UINT8 ? VX_DF_IMAGE_U8 : data_type == VX_TYPE_INT16 ? VX_DF_IMAGE_S16
                   : data_type == VX_TYPE_UINT16 ? VX_DF_IMAGE_U16 : VX_DF_IMAGE_S32;
    t->trueValue.U32 = 255;
    add_ref(c, t);
    return t;
}

VX_API_ENTRY vx_threshold VX_API_CALL vxCreateThresholdForImage(vx_context context, vx_enum thresh_type, vx_df_image input_format,
                                                               vx_df_image output_format)
{
    const vx_enum dt = input_format == VX_DF_IMAGE_U8 ? VX_TYPE_UINT8 : input_format == VX_DF_IMAGE_S16 ? VX_TYPE_INT16
                     : input_format == VX_DF_IMAGE_U16 ? VX_TYPE_UINT16 : input_format == VX_DF_IMAGE_S32 ? VX_TYPE_INT32 : 0;
    if (!dt) return nullptr;
    _vx_threshold *t = vxCreateThreshold(context, thresh_type, dt);
    if (t) {
        t->inputFormat = input_format;
        t->outputFormat = output_format;
    }
    return t;
}

static void set_pixel(vx_pixel_value_t &dst, const void *ptr, vx_size size)
{
    memset(&dst, 0, sizeof(dst));
    memcpy(&dst, ptr, std::min<vx_size>(size, sizeof(dst)));
}

VX_API_ENTRY vx_status VX_API_CALL vxSetThresholdAttribute(vx_threshold thresh, vx_enum attribute, const void *ptr, vx_size size)
{
    if (!is_ref(thresh, VX_TYPE_THRESHOLD)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr || !size) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_THRESHOLD_THRESHOLD_VALUE: set_pixel(thresh->value, ptr, size); break;
    case VX_THRESHOLD_THRESHOLD_LOWER: set_pixel(thresh->lower, ptr, size); break;
    case VX_THRESHOLD_THRESHOLD_UPPER: set_pixel(thresh->upper, ptr, size); break;
    case VX_THRESHOLD_TRUE_VALUE: set_pixel(thresh->trueValue, ptr, size); break;
    case VX_THRESHOLD_FALSE_VALUE: set_pixel(thresh->falseValue, ptr, size); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryThreshold(vx_threshold thresh, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(thresh, VX_TYPE_THRESHOLD)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    auto get = [&](const vx_pixel_value_t &v) {
        if (size > sizeof(v)) return VX_ERROR_INVALID_PARAMETERS;
        memcpy(ptr, &v, size);
        return VX_SUCCESS;
    };
    switch (attribute) {
    case VX_THRESHOLD_TYPE: PUT(vx_enum, thresh->threshType); break;
    case VX_THRESHOLD_DATA_TYPE: PUT(vx_enum, thresh->dataType); break;
    case VX_THRESHOLD_INPUT_FORMAT: PUT(vx_df_image, thresh->inputFormat); break;
    case VX_THRESHOLD_OUTPUT_FORMAT: PUT(vx_df_image, thresh->outputFormat); break;
    case VX_THRESHOLD_THRESHOLD_VALUE: return get(thresh->value);
    case VX_THRESHOLD_THRESHOLD_LOWER: return get(thresh->lower);
    case VX_THRESHOLD_THRESHOLD_UPPER: return get(thresh->upper);
    case VX_THRESHOLD_TRUE_VALUE: return get(thresh->trueValue);
    case VX_THRESHOLD_FALSE_VALUE: return get(thresh->falseValue);
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxCopyThresholdValue(vx_threshold thresh, vx_pixel_value_t *value_ptr, vx_enum usage, vx_enum user_mem_type)
{
    if (!is_ref(thresh, VX_TYPE_THRESHOLD)) return VX_ERROR_INVALID_REFERENCE;
    if (!value_ptr || user_mem_type != VX_MEMORY_TYPE_HOST || thresh->threshType != VX_THRESHOLD_TYPE_BINARY) return VX_ERROR_INVALID_PARAMETERS;
    if (usage == VX_READ_ONLY) *value_ptr = thresh->value;
    else if (usage == VX_WRITE_ONLY) thresh->value = *value_ptr;
    else return VX_ERROR_INVALID_PARAMETERS;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxCopyThresholdRange(vx_threshold thresh, vx_pixel_value_t *lower_value_ptr,
                                                       vx_pixel_value_t *upper_value_ptr, vx_enum usage, vx_enum user_mem_type)
{
    if (!is_ref(thresh, VX_TYPE_THRESHOLD)) return VX_ERROR_INVALID_REFERENCE;
    if (!lower_value_ptr || !upper_value_ptr || user_mem_type != VX_MEMORY_TYPE_HOST || thresh->threshType != VX_THRESHOLD_TYPE_RANGE)
        return VX_ERROR_INVALID_PARAMETERS;
    if (usage == VX_READ_ONLY) { *lower_value_ptr = thresh->lower; *upper_value_ptr = thresh->upper; }
    else if (usage == VX_WRITE_ONLY) { thresh->lower = *lower_value_ptr; thresh->upper = *upper_value_ptr; }
    else return VX_ERROR_INVALID_PARAMETERS;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseThreshold(vx_threshold *thresh) { return release_typed(thresh, VX_TYPE_THRESHOLD); }

// ------------------------------------------------------------------------------ arrays -----------

static vx_size item_size_of(vx_enum t)
{
    switch (t) {
    case VX_TYPE_CHAR: case VX_TYPE_INT8: case VX_TYPE_UINT8: return 1;
    case VX_TYPE_INT16: case VX_TYPE_UINT16: return 2;
    case VX_TYPE_INT32: case VX_TYPE_UINT32: case VX_TYPE_FLOAT32: case VX_TYPE_ENUM: case VX_TYPE_BOOL: return 4;
    case VX_TYPE_INT64: case VX_TYPE_UINT64: case VX_TYPE_FLOAT64: case VX_TYPE_SIZE: return 8;
    case VX_TYPE_RECTANGLE: return sizeof(vx_rectangle_t);
    case VX_TYPE_KEYPOINT: return sizeof(vx_keypoint_t);
    case VX_TYPE_COORDINATES2D: return sizeof(vx_coordinates2d_t);
    case VX_TYPE_COORDINATES2DF: return sizeof(vx_coordinates2df_t);
    case VX_TYPE_COORDINATES3D: return 12;
    default: return 0;
    }
}

vx_status _vx_array::allocate()
{
    if (dev.base) return VX_SUCCESS;
    if (!capacity || !itemSize) return VX_ERROR_INVALID_DIMENSION;
    vx_status s = context->ensure_hip();
    if (s) return s;
    return dev.alloc(256 + capacity * itemSize);
}

vx_status _vx_array::pull_count()
{
    if (!dev.base) return VX_SUCCESS;
    uint32_t n = 0;
    if (hipMemcpy(&n, count_ptr(), 4, hipMemcpyDeviceToHost) != hipSuccess) return VX_FAILURE;
    numItems = std::min<vx_size>(n, capacity);
    return VX_SUCCESS;
}

vx_status _vx_array::push_count()
{
    vx_status s = allocate();
    if (s) return s;
    const uint32_t n = (uint32_t)numItems;
    return hipMemcpy(count_ptr(), &n, 4, hipMemcpyHostToDevice) == hipSuccess ? VX_SUCCESS : VX_FAILURE;
}

VX_API_ENTRY vx_array VX_API_CALL vxCreateArray(vx_context context, vx_enum item_type, vx_size capacity)
{
    if (!is_ref(context, VX_TYPE_CONTEXT) || !capacity) return nullptr;
    const vx_size isz = item_size_of(item_type);
    if (!isz) return nullptr;
    _vx_array *a = new _vx_array(context, item_type, isz, capacity);
    add_ref(context, a);
    return a;
}

VX_API_ENTRY vx_array VX_API_CALL vxCreateVirtualArray(vx_graph graph, vx_enum item_type, vx_size capacity)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return nullptr;
    const vx_size isz = item_size_of(item_type);
    if (!isz) return nullptr;
    _vx_array *a = new _vx_array(graph->context, item_type, isz, capacity);
    a->isVirtual = true;
    a->scope = graph;
    add_ref(graph->context, a);
    return a;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryArray(vx_array arr, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(arr, VX_TYPE_ARRAY)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_ARRAY_ITEMTYPE: PUT(vx_enum, arr->itemType); break;
    case VX_ARRAY_NUMITEMS: PUT(vx_size, arr->numItems); break;
    case VX_ARRAY_CAPACITY: PUT(vx_size, arr->capacity); break;
    case VX_ARRAY_ITEMSIZE: PUT(vx_size, arr->itemSize); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxAddArrayItems(vx_array arr, vx_size count, const void *ptr, vx_size stride)
{
    if (!is_ref(arr, VX_TYPE_ARRAY)) return VX_ERROR_INVALID_REFERENCE;
    if (!count) return VX_SUCCESS;
    if (!ptr || stride < arr->itemSize) return VX_ERROR_INVALID_PARAMETERS;
    if (arr->numItems + count > arr->capacity) return VX_FAILURE;
    vx_status s = arr->allocate();
    if (s) return s;
    if (hipMemcpy2D(arr->items_ptr() + arr->numItems * arr->itemSize, arr->itemSize, ptr, stride, arr->itemSize, count,
                    hipMemcpyHostToDevice) != hipSuccess)
        return VX_FAILURE;
    arr->numItems += count;
    return arr->push_count();
}

VX_API_ENTRY vx_status VX_API_CALL vxTruncateArray(vx_array arr, vx_size new_num_items)
{
    if (!is_ref(arr, VX_TYPE_ARRAY)) return VX_ERROR_INVALID_REFERENCE;
    if (new_num_items > arr->numItems) return VX_ERROR_INVALID_PARAMETERS;
    arr->numItems = new_num_items;
    return arr->push_count();
}

VX_API_ENTRY vx_status VX_API_CALL vxCopyArrayRange(vx_array arr, vx_size range_start, vx_size range_end, vx_size user_stride,
                                                   void *user_ptr, vx_enum usage, vx_enum user_mem_type)
{
    if (!is_ref(arr, VX_TYPE_ARRAY)) return VX_ERROR_INVALID_REFERENCE;
    if (!user_ptr || range_start >= range_end || range_end > arr->numItems || user_stride < arr->itemSize ||
        (user_mem_type != VX_MEMORY_TYPE_HOST && user_mem_type != VX_MEMORY_TYPE_HIP))
        return VX_ERROR_INVALID_PARAMETERS;
    vx_status s = arr->allocate();
    if (s) return s;
    const bool hipUser = user_mem_type == VX_MEMORY_TYPE_HIP;
    uint8_t *dev = arr->items_ptr() + range_start * arr->itemSize;
    const vx_size n = range_end - range_start;
    hipError_t e;
    if (usage == VX_READ_ONLY)
        e = hipMemcpy2D(user_ptr, user_stride, dev, arr->itemSize, arr->itemSize, n, hipUser ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost);
    else if (usage == VX_WRITE_ONLY)
        e = hipMemcpy2D(dev, arr->itemSize, user_ptr, user_stride, arr->itemSize, n, hipUser ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
    else
        return VX_ERROR_INVALID_PARAMETERS;
    return e == hipSuccess ? VX_SUCCESS : VX_FAILURE;
}

VX_API_ENTRY vx_status VX_API_CALL vxMapArrayRange(vx_array arr, vx_size range_start, vx_size range_end, vx_map_id *map_id,
                                                  vx_size *stride, void **ptr, vx_enum usage, vx_enum mem_type, vx_uint32)
{
    if (!is_ref(arr, VX_TYPE_ARRAY)) return VX_ERROR_INVALID_REFERENCE;
    if (!map_id || !stride || !ptr || range_start >= range_end || range_end > arr->numItems || arr->mapId) return VX_ERROR_INVALID_PARAMETERS;
    vx_status s = arr->allocate();
    if (s) return s;
    arr->mapId = 1;
    arr->mapStart = range_start;
    arr->mapEnd = range_end;
    arr->mapUsage = usage;
    *stride = arr->itemSize;
    *map_id = arr->mapId;
    uint8_t *dev = arr->items_ptr() + range_start * arr->itemSize;
    if (mem_type == VX_MEMORY_TYPE_HIP) {
        arr->mapHost.clear();
        *ptr = dev;
        return hipDeviceSynchronize() == hipSuccess ? VX_SUCCESS : VX_FAILURE;
    }
    arr->mapHost.assign((range_end - range_start) * arr->itemSize, 0);
    if (usage != VX_WRITE_ONLY && hipMemcpy(arr->mapHost.data(), dev, arr->mapHost.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return VX_FAILURE;
    *ptr = arr->mapHost.data();
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxUnmapArrayRange(vx_array arr, vx_map_id map_id)
{
    if (!is_ref(arr, VX_TYPE_ARRAY)) return VX_ERROR_INVALID_REFERENCE;
    if (!arr->mapId || map_id != arr->mapId) return VX_ERROR_INVALID_PARAMETERS;
    vx_status s = VX_SUCCESS;
    if (!arr->mapHost.empty() && arr->mapUsage != VX_READ_ONLY)
        if (hipMemcpy(arr->items_ptr() + arr->mapStart * arr->itemSize, arr->mapHost.data(), arr->mapHost.size(), hipMemcpyHostToDevice) != hipSuccess)
            s = VX_FAILURE;
    arr->mapHost.clear();
    arr->mapId = 0;
    return s;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseArray(vx_array *arr) { return release_typed(arr, VX_TYPE_ARRAY); }

// ------------------------------------------------------------------------------ pyramids ---------

bool _vx_pyramid::build_levels()
{
    for (_vx_image *img : images) release_ref(img, false);
    images.clear();
    if (width == 0 || height == 0) return true;  // virtual pyramid: levels come at verification
    uint32_t w = width, h = height;
    // level sizes of ago_util.cpp:1312-1334: ceil(scale * previous) (ORB: from the level 4 steps up)
    const float orb[4] = {VX_SCALE_PYRAMID_ORB, VX_SCALE_PYRAMID_ORB * VX_SCALE_PYRAMID_ORB,
                          VX_SCALE_PYRAMID_ORB * VX_SCALE_PYRAMID_ORB * VX_SCALE_PYRAMID_ORB, VX_SCALE_PYRAMID_HALF};
    for (vx_size l = 0; l < levels; l++) {
        _vx_image *img = new _vx_image(context);
        img->width = w;
        img->height = h;
        img->format = format;
        img->isVirtual = isVirtual;
        img->scope = scope;
        img->parent = this;
        img->extCount = 0;
        img->intCount = 1;
        if (!img->layout()) { delete img; return false; }
        add_ref(context, img);
        images.push_back(img);
        if (scale == VX_SCALE_PYRAMID_ORB) {
            w = (uint32_t)ceilf(orb[l & 3] * images[l & ~(vx_size)3]->width);
            h = (uint32_t)ceilf(orb[l & 3] * images[l & ~(vx_size)3]->height);
        } else {
            w = (uint32_t)ceilf(scale * w);
            h = (uint32_t)ceilf(scale * h);
        }
    }
    return true;
}

static _vx_pyramid *new_pyramid(_vx_context *ctx, vx_size levels, vx_float32 scale, uint32_t w, uint32_t h, vx_df_image fmt,
                                _vx_graph *scope)
{
    if (!levels || !(scale > 0.0f && scale < 1.0f)) return nullptr;
    _vx_pyramid *p = new _vx_pyramid(ctx);
    p->levels = levels;
    p->scale = scale;
    p->width = w;
    p->height = h;
    p->format = fmt;
    p->isVirtual = scope != nullptr;
    p->scope = scope;
    if (!p->build_levels()) { delete p; return nullptr; }
    add_ref(ctx, p);
    return p;
}

VX_API_ENTRY vx_pyramid VX_API_CALL vxCreatePyramid(vx_context context, vx_size levels, vx_float32 scale, vx_uint32 width,
                                                   vx_uint32 height, vx_df_image format)
{
    if (!is_ref(context, VX_TYPE_CONTEXT) || !width || !height || format == VX_DF_IMAGE_VIRT) return nullptr;
    return new_pyramid(context, levels, scale, width, height, format, nullptr);
}

VX_API_ENTRY vx_pyramid VX_API_CALL vxCreateVirtualPyramid(vx_graph graph, vx_size levels, vx_float32 scale, vx_uint32 width,
                                                          vx_uint32 height, vx_df_image format)
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return nullptr;
    return new_pyramid(graph->context, levels, scale, width, height, format, graph);
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryPyramid(vx_pyramid pyr, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(pyr, VX_TYPE_PYRAMID)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_PYRAMID_LEVELS: PUT(vx_size, pyr->levels); break;
    case VX_PYRAMID_SCALE: PUT(vx_float32, pyr->scale); break;
    case VX_PYRAMID_WIDTH: PUT(vx_uint32, pyr->width); break;
    case VX_PYRAMID_HEIGHT: PUT(vx_uint32, pyr->height); break;
    case VX_PYRAMID_FORMAT: PUT(vx_df_image, pyr->format); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_image VX_API_CALL vxGetPyramidLevel(vx_pyramid pyr, vx_uint32 index)
{
    if (!is_ref(pyr, VX_TYPE_PYRAMID) || index >= pyr->images.size()) return nullptr;
    pyr->images[index]->extCount++;
    return pyr->images[index];
}

VX_API_ENTRY vx_status VX_API_CALL vxReleasePyramid(vx_pyramid *pyr) { return release_typed(pyr, VX_TYPE_PYRAMID); }

// ------------------------------------------------------------------------------ matrices ---------

VX_API_ENTRY vx_matrix VX_API_CALL vxCreateMatrix(vx_context c, vx_enum data_type, vx_size columns, vx_size rows)
{
    if (!is_ref(c, VX_TYPE_CONTEXT) || !columns || !rows) return nullptr;
    vx_size isz = data_type == VX_TYPE_FLOAT32 || data_type == VX_TYPE_INT32 ? 4 : data_type == VX_TYPE_UINT8 ? 1 : 0;
    if (!isz) return nullptr;
    _vx_matrix *m = new _vx_matrix(c);
    m->dataType = data_type;
    m->columns = columns;
    m->rows = rows;
    m->data.assign(columns * rows * isz, 0);
    add_ref(c, m);
    return m;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryMatrix(vx_matrix mat, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(mat, VX_TYPE_MATRIX)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_MATRIX_TYPE: PUT(vx_enum, mat->dataType); break;
    case VX_MATRIX_ROWS: PUT(vx_size, mat->rows); break;
    case VX_MATRIX_COLUMNS: PUT(vx_size, mat->columns); break;
    case VX_MATRIX_SIZE: PUT(vx_size, mat->data.size()); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxCopyMatrix(vx_matrix mat, void *user_ptr, vx_enum usage, vx_enum user_mem_type)
{
    if (!is_ref(mat, VX_TYPE_MATRIX)) return VX_ERROR_INVALID_REFERENCE;
    if (!user_ptr || user_mem_type != VX_MEMORY_TYPE_HOST) return VX_ERROR_INVALID_PARAMETERS;
    if (usage == VX_READ_ONLY) memcpy(user_ptr, mat->data.data(), mat->data.size());
    else if (usage == VX_WRITE_ONLY) memcpy(mat->data.data(), user_ptr, mat->data.size());
    else return VX_ERROR_INVALID_PARAMETERS;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseMatrix(vx_matrix *mat) { return release_typed(mat, VX_TYPE_MATRIX); }

// ------------------------------------------------------------------------------ remaps -----------

VX_API_ENTRY vx_remap VX_API_CALL vxCreateRemap(vx_context context, vx_uint32 src_width, vx_uint32 src_height, vx_uint32 dst_width,
                                               vx_uint32 dst_height)
{
    if (!is_ref(context, VX_TYPE_CONTEXT) || !src_width || !src_height || !dst_width || !dst_height) return nullptr;
    _vx_remap *r = new _vx_remap(context);
    r->srcWidth = src_width;
    r->srcHeight = src_height;
    r->dstWidth = dst_width;
    r->dstHeight = dst_height;
    r->map.assign((size_t)dst_width * dst_height * 2, 0.0f);
    add_ref(context, r);
    return r;
}

VX_API_ENTRY vx_status VX_API_CALL vxSetRemapPoint(vx_remap table, vx_uint32 dst_x, vx_uint32 dst_y, vx_float32 src_x, vx_float32 src_y)
{
    if (!is_ref(table, VX_TYPE_REMAP)) return VX_ERROR_INVALID_REFERENCE;
    if (dst_x >= table->dstWidth || dst_y >= table->dstHeight) return VX_ERROR_INVALID_PARAMETERS;
    float *p = &table->map[((size_t)dst_y * table->dstWidth + dst_x) * 2];
    p[0] = src_x;
    p[1] = src_y;
    table->dirty = true;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxGetRemapPoint(vx_remap table, vx_uint32 dst_x, vx_uint32 dst_y, vx_float32 *src_x, vx_float32 *src_y)
{
    if (!is_ref(table, VX_TYPE_REMAP)) return VX_ERROR_INVALID_REFERENCE;
    if (dst_x >= table->dstWidth || dst_y >= table->dstHeight || !src_x || !src_y) return VX_ERROR_INVALID_PARAMETERS;
    const float *p = &table->map[((size_t)dst_y * table->dstWidth + dst_x) * 2];
    *src_x = p[0];
    *src_y = p[1];
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxCopyRemapPatch(vx_remap remap, const vx_rectangle_t *rect, vx_size user_stride_y, void *user_ptr,
                                                   vx_enum user_coordinate_type, vx_enum usage, vx_enum user_mem_type)
{
    if (!is_ref(remap, VX_TYPE_REMAP)) return VX_ERROR_INVALID_REFERENCE;
    if (!rect || !user_ptr || user_coordinate_type != VX_TYPE_COORDINATES2DF || user_mem_type != VX_MEMORY_TYPE_HOST ||
        rect->start_x >= rect->end_x || rect->start_y >= rect->end_y || rect->end_x > remap->dstWidth || rect->end_y > remap->dstHeight ||
        user_stride_y < (rect->end_x - rect->start_x) * sizeof(vx_coordinates2df_t))
        return VX_ERROR_INVALID_PARAMETERS;
    const size_t rowBytes = (size_t)(rect->end_x - rect->start_x) * sizeof(vx_coordinates2df_t);
    for (uint32_t y = rect->start_y; y < rect->end_y; y++) {
        float *m = &remap->map[((size_t)y * remap->dstWidth + rect->start_x) * 2];
        uint8_t *u = (uint8_t *)user_ptr + (size_t)(y - rect->start_y) * user_stride_y;
        if (usage == VX_READ_ONLY) memcpy(u, m, rowBytes);
        else if (usage == VX_WRITE_ONLY) memcpy(m, u, rowBytes);
        else return VX_ERROR_INVALID_PARAMETERS;
    }
    if (usage == VX_WRITE_ONLY) remap->dirty = true;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxQueryRemap(vx_remap table, vx_enum attribute, void *ptr, vx_size size)
{
    if (!is_ref(table, VX_TYPE_REMAP)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
    switch (attribute) {
    case VX_REMAP_SOURCE_WIDTH: PUT(vx_uint32, table->srcWidth); break;
    case VX_REMAP_SOURCE_HEIGHT: PUT(vx_uint32, table->srcHeight); break;
    case VX_REMAP_DESTINATION_WIDTH: PUT(vx_uint32, table->dstWidth); break;
    case VX_REMAP_DESTINATION_HEIGHT: PUT(vx_uint32, table->dstHeight); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    return VX_SUCCESS;
}

vx_status _vx_remap::sync(hipStream_t stream)
{
    if (!dirty && devFixed.base) return VX_SUCCESS;
    vx_status s = context->ensure_hip();
    if (!s) s = devFloat.alloc(map.size() * sizeof(float));
    if (!s) s = devFixed.alloc((size_t)dstWidth * dstHeight * 4);
    if (s) return s;
    // the float table goes up once per change; the fixed-point table is derived from it on the GPU
    if (hipMemcpyAsync(devFloat.ptr(), map.data(), map.size() * sizeof(float), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return VX_FAILURE;
    const int r = mvx_Remap_FixedMap_F32(stream, dstWidth, dstHeight, devFixed.ptr(), dstWidth * 4, (const float *)devFloat.ptr(),
                                         dstWidth * 8, srcWidth, srcHeight);
    if (r) return r;
    dirty = false;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxReleaseRemap(vx_remap *table) { return release_typed(table, VX_TYPE_REMAP); }
