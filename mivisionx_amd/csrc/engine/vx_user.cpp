// vx_user.cpp -- the plug-in surface of the engine: modules loaded with vxLoadKernels, user kernels
// (vxAddUserKernel, vxAddParameterToKernel, vxSetKernelAttribute, vxFinalizeKernel, vxRemoveKernel),
// meta formats for their validate callbacks, the query_target_support decision for every node, and
// the valid-rectangle propagation behind vxGetValidRegionImage.
//
// Follows the reference's interface (agoLoadModule ago_interface.cpp:1128-1177; vx_api.cpp:2325 and
// 2584-2810; target resolution ago_drama_alloc.cpp:587-624; valid rectangles
// agoComputeImageValidRectangleOutputs ago_interface.cpp:1735-1790) so that modules written for the
// reference -- amd_rpp's style of kernel, for one -- load and run unchanged.  What differs is where
// data lives: every object is in HBM.  A user kernel that declares GPU support (or asks for GPU buffer
// access) is handed device pointers and the graph's HIP stream and enqueues its own work; one that
// runs on the CPU is called after the graph's stream has drained and reaches its data through the
// map / copy API, which copies synchronously.
#include <dlfcn.h>

#include <cmath>

#include "engine.hpp"

using namespace mvxe;

namespace {

bool is_user(_vx_kernel *k) { return k && k->def && k->def->user; }

vx_status user_validate(_vx_node *n);
vx_status user_initialize(_vx_node *n);
vx_status user_execute(_vx_node *n, hipStream_t st);

}  // namespace

// ------------------------------------------------------------------------------ modules ----------

VX_API_ENTRY vx_status VX_API_CALL vxLoadKernels(vx_context context, const vx_char *module)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    if (!module || !module[0]) return VX_ERROR_INVALID_PARAMETERS;
    std::lock_guard<std::recursive_mutex> lk(context->lock);
    const std::string path = std::string("lib") + module + ".so";
    for (const ModuleInfo &m : context->modules)
        if (m.path == path) {
            log(context, VX_SUCCESS, "WARNING: kernels already loaded from %s\n", path.c_str());
            return VX_SUCCESS;
        }
    void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        log(context, VX_ERROR_INVALID_MODULE, "ERROR: Unable to load module %s (%s)\n", path.c_str(), dlerror());
        return VX_ERROR_INVALID_MODULE;
    }
    auto publish = (vx_publish_kernels_f)dlsym(h, "vxPublishKernels");
    if (!publish) {
        log(context, VX_ERROR_INVALID_MODULE, "ERROR: vxPublishKernels symbol missing in %s\n", path.c_str());
        dlclose(h);
        return VX_ERROR_INVALID_MODULE;
    }
    context->modules.push_back(ModuleInfo{module, path, h});
    const size_t before = context->kernels.size();
    context->importingModule = (int)context->modules.size() - 1;
    const vx_status s = publish(context);
    context->importingModule = -1;
    if (s == VX_SUCCESS)
        log(context, VX_SUCCESS, "OK: loaded %zu kernels from %s\n", context->kernels.size() - before, path.c_str());
    else
        log(context, s, "ERROR: vxPublishKernels(%s) failed (%d)\n", module, s);
    return s;
}

VX_API_ENTRY vx_status VX_API_CALL vxUnloadKernels(vx_context context, const vx_char *module)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    if (!module) return VX_ERROR_INVALID_PARAMETERS;
    std::lock_guard<std::recursive_mutex> lk(context->lock);
    const std::string path = std::string("lib") + module + ".so";
    for (size_t i = 0; i < context->modules.size(); i++) {
        if (context->modules[i].path != path) continue;
        vx_status s = VX_SUCCESS;
        if (auto unpublish = (vx_unpublish_kernels_f)dlsym(context->modules[i].handle, "vxUnpublishKernels")) s = unpublish(context);
        // kernels the module published and did not remove itself leave the lookup list now; the
        // objects stay until the context ends (nodes may still point at them)
        for (size_t k = 0; k < context->kernels.size();) {
            _vx_kernel *kr = context->kernels[k];
            if (is_user(kr) && kr->user.module == (int)i) {
                context->removedKernels.push_back(kr);
                context->kernels.erase(context->kernels.begin() + k);
            } else {
                k++;
            }
        }
        // the module's code stays mapped: callbacks of kernels that nodes still hold live in it
        context->modules[i].path.clear();
        return s;
    }
    return VX_ERROR_INVALID_PARAMETERS;
}

VX_API_ENTRY vx_status VX_API_CALL vxAllocateUserKernelId(vx_context context, vx_enum *pKernelEnumId)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    if (!pKernelEnumId) return VX_ERROR_INVALID_PARAMETERS;
    std::lock_guard<std::recursive_mutex> lk(context->lock);
    if (context->nextUserKernelId >= 4096) return VX_ERROR_NO_RESOURCES;
    *pKernelEnumId = VX_KERNEL_BASE(VX_ID_USER, 0) + context->nextUserKernelId++;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxAllocateUserKernelLibraryId(vx_context context, vx_enum *pLibraryId)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    if (!pLibraryId) return VX_ERROR_INVALID_PARAMETERS;
    std::lock_guard<std::recursive_mutex> lk(context->lock);
    if (context->nextUserLibraryId >= 256) return VX_ERROR_NO_RESOURCES;
    *pLibraryId = context->nextUserLibraryId++;
    return VX_SUCCESS;
}

// ------------------------------------------------------------------------------ user kernels -----

VX_API_ENTRY vx_kernel VX_API_CALL vxAddUserKernel(vx_context context, const vx_char name[VX_MAX_KERNEL_NAME], vx_enum enumeration,
                                                   vx_kernel_f func_ptr, vx_uint32 numParams, vx_kernel_validate_f validate,
                                                   vx_kernel_initialize_f init, vx_kernel_deinitialize_f deinit)
{
    if (!is_ref(context, VX_TYPE_CONTEXT) || !name || !func_ptr || !validate || numParams == 0 || numParams > MAX_PARAMS) return nullptr;
    std::lock_guard<std::recursive_mutex> lk(context->lock);
    for (_vx_kernel *k : context->kernels)
        if (k->def->id == enumeration || !strcmp(k->def->name, name)) {
            log(context, VX_ERROR_INVALID_PARAMETERS, "ERROR: vxAddUserKernel: a kernel named %s or with enum 0x%x exists\n", name,
                (unsigned)enumeration);
            return nullptr;
        }
    _vx_kernel *k = new _vx_kernel(context, nullptr);
    k->user.name = name;
    k->user.func = func_ptr;
    k->user.validate = validate;
    k->user.init = init;
    k->user.deinit = deinit;
    k->user.module = context->importingModule;
    k->own.id = enumeration;
    k->own.name = k->user.name.c_str();
    k->own.params.assign(numParams, ParamSpec{0, 0, false});
    k->own.validate = user_validate;
    k->own.initialize = user_initialize;
    k->own.execute = user_execute;
    k->own.targets = AGO_TARGET_AFFINITY_CPU;
    k->own.user = &k->user;
    k->def = &k->own;
    k->name = name;
    k->extCount = 1;
    k->intCount = 1;
    context->kernels.push_back(k);
    return k;
}

VX_API_ENTRY vx_status VX_API_CALL vxAddParameterToKernel(vx_kernel kernel, vx_uint32 index, vx_enum dir, vx_enum data_type, vx_enum state)
{
    if (!is_ref(kernel, VX_TYPE_KERNEL)) return VX_ERROR_INVALID_REFERENCE;
    if (!is_user(kernel) || kernel->user.finalized || index >= kernel->own.params.size()) return VX_ERROR_INVALID_PARAMETERS;
    if (dir != VX_INPUT && dir != VX_OUTPUT && dir != VX_BIDIRECTIONAL) return VX_ERROR_INVALID_PARAMETERS;
    if (state != VX_PARAMETER_STATE_REQUIRED && state != VX_PARAMETER_STATE_OPTIONAL) return VX_ERROR_INVALID_PARAMETERS;
    if (data_type == VX_TYPE_DELAY && dir != VX_INPUT) return VX_ERROR_INVALID_PARAMETERS;
    kernel->own.params[index] = ParamSpec{dir, data_type, state == VX_PARAMETER_STATE_OPTIONAL};
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxFinalizeKernel(vx_kernel kernel)
{
    if (!is_ref(kernel, VX_TYPE_KERNEL) || !is_user(kernel) || kernel->user.finalized) return VX_ERROR_INVALID_REFERENCE;
    for (const ParamSpec &p : kernel->own.params)
        if (!p.type) return VX_ERROR_INVALID_REFERENCE;   // every parameter must have been added
    kernel->user.finalized = true;
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxRemoveKernel(vx_kernel kernel)
{
    if (!is_ref(kernel, VX_TYPE_KERNEL)) return VX_ERROR_INVALID_REFERENCE;
    if (!is_user(kernel)) return VX_ERROR_INVALID_PARAMETERS;   // built-in kernels stay
    _vx_context *ctx = kernel->context;
    std::lock_guard<std::recursive_mutex> lk(ctx->lock);
    for (size_t i = 0; i < ctx->kernels.size(); i++)
        if (ctx->kernels[i] == kernel) {
            ctx->kernels.erase(ctx->kernels.begin() + i);
            ctx->removedKernels.push_back(kernel);   // nodes may still use it; freed with the context
            kernel->user.finalized = false;
            return VX_SUCCESS;
        }
    return VX_ERROR_INVALID_PARAMETERS;
}

VX_API_ENTRY vx_status VX_API_CALL vxSetKernelAttribute(vx_kernel kernel, vx_enum attribute, const void *ptr, vx_size size)
{
    if (!is_ref(kernel, VX_TYPE_KERNEL)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr || !is_user(kernel)) return VX_ERROR_INVALID_PARAMETERS;
    UserKernel &u = kernel->user;
    switch (attribute) {
    case VX_KERNEL_LOCAL_DATA_SIZE:
        if (size != sizeof(vx_size)) return VX_ERROR_INVALID_PARAMETERS;
        u.localDataSize = *(const vx_size *)ptr;
        return VX_SUCCESS;
    case VX_KERNEL_ATTRIBUTE_AMD_QUERY_TARGET_SUPPORT:
        if (size != sizeof(void *)) return VX_ERROR_INVALID_PARAMETERS;
        if (u.finalized) return VX_ERROR_NOT_SUPPORTED;
        memcpy(&u.queryTarget, ptr, sizeof(void *));
        return VX_SUCCESS;
    case VX_KERNEL_ATTRIBUTE_AMD_GPU_BUFFER_ACCESS_ENABLE:
        if (size != sizeof(vx_bool)) return VX_ERROR_INVALID_PARAMETERS;
        if (u.finalized) return VX_ERROR_NOT_SUPPORTED;
        u.gpuBufferAccess = *(const vx_bool *)ptr == vx_true_e;
        return VX_SUCCESS;
    default:
        return VX_ERROR_NOT_SUPPORTED;
    }
}

// ------------------------------------------------------------------------------ meta formats -----

VX_API_ENTRY vx_status VX_API_CALL vxSetMetaFormatAttribute(vx_meta_format meta, vx_enum attribute, const void *ptr, vx_size size)
{
    if (!is_ref(meta, VX_TYPE_META_FORMAT)) return VX_ERROR_INVALID_REFERENCE;
    if (!ptr) return VX_ERROR_INVALID_PARAMETERS;
#define META_SET(T, field) do { if (size != sizeof(T)) return VX_ERROR_INVALID_PARAMETERS; meta->field = *(const T *)ptr; } while (0)
    switch (attribute) {
    case VX_IMAGE_WIDTH: META_SET(vx_uint32, width); break;
    case VX_IMAGE_HEIGHT: META_SET(vx_uint32, height); break;
    case VX_IMAGE_FORMAT: META_SET(vx_df_image, format); break;
    case VX_ARRAY_ITEMTYPE: META_SET(vx_enum, itemType); break;
    case VX_ARRAY_CAPACITY: META_SET(vx_size, capacity); break;
    case VX_SCALAR_TYPE: META_SET(vx_enum, scalarType); break;
    case VX_PYRAMID_LEVELS: META_SET(vx_size, levels); break;
    case VX_PYRAMID_SCALE: META_SET(vx_float32, scale); break;
    case VX_PYRAMID_WIDTH: META_SET(vx_uint32, width); break;
    case VX_PYRAMID_HEIGHT: META_SET(vx_uint32, height); break;
    case VX_PYRAMID_FORMAT: META_SET(vx_df_image, format); break;
    case VX_VALID_RECT_CALLBACK: META_SET(vx_kernel_image_valid_rectangle_f, validRect); break;
    default: return VX_ERROR_NOT_SUPPORTED;
    }
#undef META_SET
    return VX_SUCCESS;
}

VX_API_ENTRY vx_status VX_API_CALL vxSetMetaFormatFromReference(vx_meta_format meta, vx_reference exemplar)
{
    if (!is_ref(meta, VX_TYPE_META_FORMAT) || !is_ref(exemplar)) return VX_ERROR_INVALID_REFERENCE;
    switch (exemplar->type) {
    case VX_TYPE_IMAGE: {
        _vx_image *im = (_vx_image *)exemplar;
        meta->width = im->width; meta->height = im->height; meta->format = im->format;
        break;
    }
    case VX_TYPE_ARRAY: meta->itemType = ((_vx_array *)exemplar)->itemType; meta->capacity = ((_vx_array *)exemplar)->capacity; break;
    case VX_TYPE_SCALAR: meta->scalarType = ((_vx_scalar *)exemplar)->dataType; break;
    case VX_TYPE_PYRAMID: {
        _vx_pyramid *p = (_vx_pyramid *)exemplar;
        meta->levels = p->levels; meta->scale = p->scale; meta->width = p->width; meta->height = p->height; meta->format = p->format;
        break;
    }
    default: return VX_ERROR_NOT_SUPPORTED;
    }
    if (meta->objType && meta->objType != exemplar->type) return VX_ERROR_INVALID_TYPE;
    return VX_SUCCESS;
}

// ------------------------------------------------------------------------------ user nodes -------

namespace {

vx_status user_fail(_vx_node *n, vx_status s, const char *what)
{
    log(n, s, "ERROR: %s: %s\n", n->kernel->name, what);
    return s;
}

// applies the meta data a validate callback set to output parameter i: virtual objects take it, others
// must already match it (the reference's agoVerifyNode, ago_interface.cpp:1347-1460)
vx_status apply_meta(_vx_node *n, uint32_t i, const _vx_meta_format &m)
{
    _vx_reference *r = n->params[i];
    switch (r->type) {
    case VX_TYPE_IMAGE: {
        _vx_image *im = (_vx_image *)r;
        if (im->isVirtual && (!im->width || !im->height || im->format == VX_DF_IMAGE_VIRT)) {
            if (!im->width) im->width = m.width;
            if (!im->height) im->height = m.height;
            if (im->format == VX_DF_IMAGE_VIRT) im->format = m.format;
            if (!im->layout()) return user_fail(n, VX_ERROR_INVALID_FORMAT, "output image format");
        }
        if (m.format != VX_DF_IMAGE_VIRT && im->format != m.format) return user_fail(n, VX_ERROR_INVALID_FORMAT, "output image format");
        if ((m.width && im->width != m.width) || (m.height && im->height != m.height))
            return user_fail(n, VX_ERROR_INVALID_DIMENSION, "output image size");
        return VX_SUCCESS;
    }
    case VX_TYPE_ARRAY: {
        _vx_array *a = (_vx_array *)r;
        if (a->isVirtual && !a->capacity) a->capacity = m.capacity;
        if (m.itemType && a->itemType != m.itemType) return user_fail(n, VX_ERROR_INVALID_TYPE, "output array item type");
        if (m.capacity && a->capacity < m.capacity) return user_fail(n, VX_ERROR_INVALID_DIMENSION, "output array capacity");
        return VX_SUCCESS;
    }
    case VX_TYPE_SCALAR:
        if (m.scalarType && ((_vx_scalar *)r)->dataType != m.scalarType) return user_fail(n, VX_ERROR_INVALID_TYPE, "output scalar type");
        return VX_SUCCESS;
    case VX_TYPE_PYRAMID: {
        _vx_pyramid *p = (_vx_pyramid *)r;
        if (p->isVirtual && (!p->width || !p->height)) {
            p->width = m.width;
            p->height = m.height;
            if (p->format == VX_DF_IMAGE_VIRT) p->format = m.format;
            if (!p->build_levels()) return user_fail(n, VX_ERROR_INVALID_FORMAT, "output pyramid");
        }
        if ((m.levels && p->levels != m.levels) || (m.width && p->width != m.width) || (m.height && p->height != m.height))
            return user_fail(n, VX_ERROR_INVALID_DIMENSION, "output pyramid");
        return VX_SUCCESS;
    }
    default:
        return VX_SUCCESS;
    }
}

vx_status user_validate(_vx_node *n)
{
    UserKernel *u = n->kernel->user;
    const uint32_t num = (uint32_t)n->params.size();
    std::vector<_vx_meta_format *> metas(num);
    std::vector<vx_meta_format> handles(num);
    for (uint32_t i = 0; i < num; i++) {
        metas[i] = new _vx_meta_format(n->context);
        metas[i]->objType = n->kernel->params[i].type;
        handles[i] = metas[i];
    }
    vx_status s = u->validate(n, (const vx_reference *)n->params.data(), num, handles.data());
    if (s) log(n, s, "ERROR: %s: validate callback returned %d\n", n->kernel->name, s);
    n->validRectCb.assign(num, nullptr);
    for (uint32_t i = 0; i < num && !s; i++) {
        if (n->kernel->params[i].dir != VX_OUTPUT || !n->params[i]) continue;
        s = apply_meta(n, i, *metas[i]);
        n->validRectCb[i] = metas[i]->validRect;
    }
    for (_vx_meta_format *m : metas) delete m;
    return s;
}

vx_status user_initialize(_vx_node *n)
{
    UserKernel *u = n->kernel->user;
    n->localSize = 0;   // no device workspace: user kernels get host local data
    const vx_size bytes = n->userLocalSize ? n->userLocalSize : u->localDataSize;
    if (bytes && !n->localPtr) {
        n->localHost.assign(bytes, 0);
        n->localPtr = n->localHost.data();
    }
    if (u->init) {
        const vx_status s = u->init(n, (const vx_reference *)n->params.data(), (vx_uint32)n->params.size());
        if (s) return user_fail(n, s, "initialize callback failed");
    }
    n->userInitialized = true;
    return VX_SUCCESS;
}

// GPU nodes and GPU-buffer nodes enqueue on the graph's stream (VX_NODE_ATTRIBUTE_AMD_HIP_STREAM); CPU
// nodes run after the stream has drained, as the reference synchronises at a GPU -> CPU level change
// (ago_interface.cpp:2337-2365)
vx_status user_execute(_vx_node *n, hipStream_t st)
{
    UserKernel *u = n->kernel->user;
    const bool device = n->target == AGO_TARGET_AFFINITY_GPU || u->gpuBufferAccess;
    if (!device) {
        if (hipStreamSynchronize(st) != hipSuccess) return user_fail(n, VX_FAILURE, "stream synchronisation before a CPU node");
        // item counts that earlier nodes of this graph wrote on the device
        for (size_t i = 0; i < n->params.size(); i++)
            if (n->params[i] && n->params[i]->type == VX_TYPE_ARRAY && n->kernel->params[i].dir != VX_OUTPUT &&
                ((_vx_array *)n->params[i])->dev.ptr())
                if (((_vx_array *)n->params[i])->pull_count()) return user_fail(n, VX_FAILURE, "array count");
    }
    const vx_status s = u->func(n, (const vx_reference *)n->params.data(), (vx_uint32)n->params.size());
    if (s) log(n, s, "ERROR: kernel %s exec failed (%d)\n", n->kernel->name, s);
    return s;
}

}  // namespace

namespace mvxe {

void deinit_user_node(_vx_node *n)
{
    if (!n->userInitialized) return;
    n->userInitialized = false;
    UserKernel *u = n->kernel->user;
    if (u && u->deinit) (void)u->deinit(n, (const vx_reference *)n->params.data(), (vx_uint32)n->params.size());
    if (!n->localHost.empty() && n->localPtr == n->localHost.data()) n->localPtr = nullptr;
    n->localHost.clear();
}

// query_target_support (ago_drama_alloc.cpp:587-624): built-in kernels answer from their table row;
// a user kernel through its callback, or CPU only without one
uint32_t target_support(_vx_node *n, vx_status *status)
{
    *status = VX_SUCCESS;
    UserKernel *u = n->kernel->user;
    if (!u) return n->kernel->targets;
    if (!u->queryTarget) return AGO_TARGET_AFFINITY_CPU;
    vx_uint32 t = 0;
    const vx_status s = u->queryTarget(n->graph, n, vx_false_e, t);
    if (s) {
        log(n, s, "ERROR: kernel %s: query_target_support_f(*,*,0,*) => %d\n", n->kernel->name, s);
        *status = s;
        return 0;
    }
    if (t & ~(uint32_t)(AGO_TARGET_AFFINITY_CPU | AGO_TARGET_AFFINITY_GPU)) {
        log(n, VX_FAILURE, "ERROR: kernel %s: query_target_support_f returned unsupported affinity flags: 0x%08x\n", n->kernel->name, t);
        *status = VX_FAILURE;
        return 0;
    }
    return t;
}

}  // namespace mvxe

// ------------------------------------------------------------------------------ valid rectangles -

namespace mvxe {

// agoComputeImageValidRectangleOutputs (ago_interface.cpp:1735-1790): in execution order, built-in
// kernels apply their valid_rect_callback rule; a user node calls the VX_VALID_RECT_CALLBACK its validate
// callback set for an output image with the valid rectangles of all its input images (pyramid levels
// one by one), and otherwise resets the output to the whole image (the reference's user kernels carry
// AGO_KERNEL_FLAG_VALID_RECT_RESET, vx_api.cpp:2601).
void propagate_valid_rects(_vx_graph *g)
{
    for (_vx_node *n : g->order) {
        if (!n->kernel->user) {
            if (n->kernel->validRect) n->kernel->validRect(n);
            continue;
        }
        std::vector<_vx_image *> inImgs;
        for (size_t i = 0; i < n->params.size(); i++) {
            _vx_reference *r = n->params[i];
            if (!r || n->kernel->params[i].dir != VX_INPUT) continue;
            if (r->type == VX_TYPE_IMAGE) inImgs.push_back((_vx_image *)r);
            else if (r->type == VX_TYPE_PYRAMID)
                for (_vx_image *l : ((_vx_pyramid *)r)->images) inImgs.push_back(l);
        }
        std::vector<vx_rectangle_t> inRects(inImgs.size());
        std::vector<const vx_rectangle_t *> inPtrs(inImgs.size());
        for (size_t k = 0; k < inImgs.size(); k++) {
            vxGetValidRegionImage(inImgs[k], &inRects[k]);
            inPtrs[k] = &inRects[k];
        }
        for (size_t i = 0; i < n->params.size(); i++) {
            _vx_reference *r = n->params[i];
            if (!r || n->kernel->params[i].dir != VX_OUTPUT) continue;
            std::vector<_vx_image *> outs;
            if (r->type == VX_TYPE_IMAGE) outs.push_back((_vx_image *)r);
            else if (r->type == VX_TYPE_PYRAMID) outs = ((_vx_pyramid *)r)->images;
            else continue;
            vx_kernel_image_valid_rectangle_f cb = i < n->validRectCb.size() ? n->validRectCb[i] : nullptr;
            std::vector<vx_rectangle_t> rects(outs.size());
            std::vector<vx_rectangle_t *> outPtrs(outs.size());
            for (size_t k = 0; k < outs.size(); k++) {
                rects[k] = vx_rectangle_t{0, 0, outs[k]->width, outs[k]->height};
                outPtrs[k] = &rects[k];
            }
            if (cb && cb(n, (vx_uint32)i, inPtrs.data(), outPtrs.data()) != VX_SUCCESS)
                log(n, VX_FAILURE, "WARNING: %s: valid rectangle callback of output %zu failed\n", n->kernel->name, i);
            for (size_t k = 0; k < outs.size(); k++) {
                outs[k]->validRect = rects[k];
                outs[k]->validRectSet = true;
            }
        }
    }
}

}  // namespace mvxe
