// user_module.hip -- a small OpenVX user-kernel module for the plug-in tests (test infrastructure, not
// part of the product).  It is written against the public headers only, the way a third-party module
// (amd_rpp's kernels, for instance) is written against the reference's, and exports vxPublishKernels /
// vxUnpublishKernels for vxLoadKernels("mvx_test_user").
//
//   com.example.invert       GPU (query_target_support): out = 255 - in, a HIP kernel launched on the
//                            node's stream (VX_NODE_ATTRIBUTE_AMD_HIP_STREAM) over the images' device
//                            buffers (VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER); its output's valid rectangle is
//                            the input's shrunk by 2 (VX_VALID_RECT_CALLBACK)
//   com.example.add_sat      CPU: out = min(a + b, 255) through vxMapImagePatch, counts its runs in the
//                            node's local data (VX_KERNEL_LOCAL_DATA_SIZE) and reports them in a scalar
//   com.example.copy_gpubuf  CPU with VX_KERNEL_ATTRIBUTE_AMD_GPU_BUFFER_ACCESS_ENABLE: a device copy
//                            enqueued on the node's stream
#include <VX/vx.h>
#include <vx_ext_amd.h>
#include <hip/hip_runtime.h>

#include <cstring>

#define TEST_KERNEL(n) (VX_KERNEL_BASE(VX_ID_USER, 0x3) + (n))

__global__ void k_invert(const uint8_t *in, uint32_t inStride, uint8_t *out, uint32_t outStride, uint32_t w, uint32_t h)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < w && y < h) out[(size_t)y * outStride + x] = (uint8_t)(255u - in[(size_t)y * inStride + x]);
}

static vx_status VX_CALLBACK rect_shrink2(vx_node, vx_uint32, const vx_rectangle_t *const in[], vx_rectangle_t *const out[])
{
    out[0]->start_x = in[0]->start_x + 2;
    out[0]->start_y = in[0]->start_y + 2;
    out[0]->end_x = in[0]->end_x - 2;
    out[0]->end_y = in[0]->end_y - 2;
    return VX_SUCCESS;
}

static vx_status size_u8(vx_reference r, vx_uint32 *w, vx_uint32 *h)
{
    vx_df_image f = 0;
    if (vxQueryImage((vx_image)r, VX_IMAGE_WIDTH, w, sizeof(*w)) || vxQueryImage((vx_image)r, VX_IMAGE_HEIGHT, h, sizeof(*h)) ||
        vxQueryImage((vx_image)r, VX_IMAGE_FORMAT, &f, sizeof(f)))
        return VX_ERROR_INVALID_PARAMETERS;
    return f == VX_DF_IMAGE_U8 ? VX_SUCCESS : VX_ERROR_INVALID_FORMAT;
}

static vx_status VX_CALLBACK validate_unary(vx_node, const vx_reference p[], vx_uint32 num, vx_meta_format metas[])
{
    vx_uint32 w = 0, h = 0;
    vx_status s = num == 2 ? size_u8(p[0], &w, &h) : VX_ERROR_INVALID_PARAMETERS;
    if (s) return s;
    const vx_df_image f = VX_DF_IMAGE_U8;
    vxSetMetaFormatAttribute(metas[1], VX_IMAGE_WIDTH, &w, sizeof(w));
    vxSetMetaFormatAttribute(metas[1], VX_IMAGE_HEIGHT, &h, sizeof(h));
    return vxSetMetaFormatAttribute(metas[1], VX_IMAGE_FORMAT, &f, sizeof(f));
}

static vx_status VX_CALLBACK validate_invert(vx_node node, const vx_reference p[], vx_uint32 num, vx_meta_format metas[])
{
    vx_status s = validate_unary(node, p, num, metas);
    if (s) return s;
    vx_kernel_image_valid_rectangle_f cb = rect_shrink2;
    return vxSetMetaFormatAttribute(metas[1], VX_VALID_RECT_CALLBACK, &cb, sizeof(cb));
}

static vx_status VX_CALLBACK invert_exec(vx_node node, const vx_reference *p, vx_uint32)
{
    hipStream_t st = nullptr;
    void *in = nullptr, *out = nullptr;
    vx_uint32 w = 0, h = 0, is = 0, os = 0;
    if (vxQueryNode(node, VX_NODE_ATTRIBUTE_AMD_HIP_STREAM, &st, sizeof(st)) || size_u8(p[0], &w, &h) ||
        vxQueryImage((vx_image)p[0], VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER, &in, sizeof(in)) ||
        vxQueryImage((vx_image)p[1], VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER, &out, sizeof(out)) ||
        vxQueryImage((vx_image)p[0], VX_IMAGE_ATTRIBUTE_AMD_GPU_BUFFER_STRIDE, &is, sizeof(is)) ||
        vxQueryImage((vx_image)p[1], VX_IMAGE_ATTRIBUTE_AMD_GPU_BUFFER_STRIDE, &os, sizeof(os)) || !in || !out)
        return VX_FAILURE;
    hipLaunchKernelGGL(k_invert, dim3((w + 255) / 256, h), dim3(256), 0, st, (const uint8_t *)in, is, (uint8_t *)out, os, w, h);
    return hipGetLastError() == hipSuccess ? VX_SUCCESS : VX_FAILURE;
}

static vx_status VX_CALLBACK query_gpu(vx_graph, vx_node, vx_bool, vx_uint32 &supported)
{
    supported = AGO_TARGET_AFFINITY_GPU;
    return VX_SUCCESS;
}

struct AddState { vx_uint32 magic, runs; };

static vx_status VX_CALLBACK validate_add(vx_node, const vx_reference p[], vx_uint32 num, vx_meta_format metas[])
{
    vx_uint32 w = 0, h = 0, w2 = 0, h2 = 0;
    if (num != 4) return VX_ERROR_INVALID_PARAMETERS;
    vx_status s = size_u8(p[0], &w, &h);
    if (!s) s = size_u8(p[1], &w2, &h2);
    if (s) return s;
    if (w != w2 || h != h2) return VX_ERROR_INVALID_DIMENSION;
    const vx_df_image f = VX_DF_IMAGE_U8;
    const vx_enum t = VX_TYPE_UINT32;
    vxSetMetaFormatAttribute(metas[2], VX_IMAGE_WIDTH, &w, sizeof(w));
    vxSetMetaFormatAttribute(metas[2], VX_IMAGE_HEIGHT, &h, sizeof(h));
    vxSetMetaFormatAttribute(metas[2], VX_IMAGE_FORMAT, &f, sizeof(f));
    return vxSetMetaFormatAttribute(metas[3], VX_SCALAR_TYPE, &t, sizeof(t));
}

static vx_status VX_CALLBACK add_init(vx_node node, const vx_reference *, vx_uint32)
{
    AddState *s = nullptr;
    if (vxQueryNode(node, VX_NODE_LOCAL_DATA_PTR, &s, sizeof(s)) || !s) return VX_FAILURE;
    s->magic = 0xC0FFEEu;
    s->runs = 0;
    return VX_SUCCESS;
}

static vx_status VX_CALLBACK add_deinit(vx_node node, const vx_reference *, vx_uint32)
{
    AddState *s = nullptr;
    if (vxQueryNode(node, VX_NODE_LOCAL_DATA_PTR, &s, sizeof(s)) || !s || s->magic != 0xC0FFEEu) return VX_FAILURE;
    s->magic = 0;
    return VX_SUCCESS;
}

static vx_status VX_CALLBACK add_exec(vx_node node, const vx_reference *p, vx_uint32)
{
    AddState *st = nullptr;
    if (vxQueryNode(node, VX_NODE_LOCAL_DATA_PTR, &st, sizeof(st)) || !st || st->magic != 0xC0FFEEu) return VX_FAILURE;
    vx_uint32 w = 0, h = 0;
    if (size_u8(p[0], &w, &h)) return VX_FAILURE;
    vx_rectangle_t r = {0, 0, w, h};
    vx_map_id ma, mb, mo;
    vx_imagepatch_addressing_t aa, ab, ao;
    void *pa = nullptr, *pb = nullptr, *po = nullptr;
    if (vxMapImagePatch((vx_image)p[0], &r, 0, &ma, &aa, &pa, VX_READ_ONLY, VX_MEMORY_TYPE_HOST, VX_NOGAP_X)) return VX_FAILURE;
    if (vxMapImagePatch((vx_image)p[1], &r, 0, &mb, &ab, &pb, VX_READ_ONLY, VX_MEMORY_TYPE_HOST, VX_NOGAP_X)) return VX_FAILURE;
    if (vxMapImagePatch((vx_image)p[2], &r, 0, &mo, &ao, &po, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST, VX_NOGAP_X)) return VX_FAILURE;
    for (vx_uint32 y = 0; y < h; y++)
        for (vx_uint32 x = 0; x < w; x++) {
            const unsigned v = ((const uint8_t *)pa)[y * aa.stride_y + x] + ((const uint8_t *)pb)[y * ab.stride_y + x];
            ((uint8_t *)po)[y * ao.stride_y + x] = (uint8_t)(v > 255u ? 255u : v);
        }
    vxUnmapImagePatch((vx_image)p[0], ma);
    vxUnmapImagePatch((vx_image)p[1], mb);
    if (vxUnmapImagePatch((vx_image)p[2], mo)) return VX_FAILURE;
    st->runs++;
    return vxCopyScalar((vx_scalar)p[3], &st->runs, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST);
}

static vx_status VX_CALLBACK copy_exec(vx_node node, const vx_reference *p, vx_uint32)
{
    hipStream_t st = nullptr;
    void *in = nullptr, *out = nullptr;
    vx_uint32 w = 0, h = 0, is = 0, os = 0;
    if (vxQueryNode(node, VX_NODE_ATTRIBUTE_AMD_HIP_STREAM, &st, sizeof(st)) || size_u8(p[0], &w, &h) ||
        vxQueryImage((vx_image)p[0], VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER, &in, sizeof(in)) ||
        vxQueryImage((vx_image)p[1], VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER, &out, sizeof(out)) ||
        vxQueryImage((vx_image)p[0], VX_IMAGE_ATTRIBUTE_AMD_GPU_BUFFER_STRIDE, &is, sizeof(is)) ||
        vxQueryImage((vx_image)p[1], VX_IMAGE_ATTRIBUTE_AMD_GPU_BUFFER_STRIDE, &os, sizeof(os)) || !in || !out)
        return VX_FAILURE;
    return hipMemcpy2DAsync(out, os, in, is, w, h, hipMemcpyDeviceToDevice, st) == hipSuccess ? VX_SUCCESS : VX_FAILURE;
}

#define CHECK(x) do { vx_status _s = (x); if (_s != VX_SUCCESS) return _s; } while (0)

extern "C" vx_status VX_API_CALL vxPublishKernels(vx_context ctx)
{
    vx_kernel k = vxAddUserKernel(ctx, "com.example.invert", TEST_KERNEL(1), invert_exec, 2, validate_invert, nullptr, nullptr);
    if (!k) return VX_FAILURE;
    amd_kernel_query_target_support_f q = query_gpu;
    CHECK(vxSetKernelAttribute(k, VX_KERNEL_ATTRIBUTE_AMD_QUERY_TARGET_SUPPORT, &q, sizeof(q)));
    CHECK(vxAddParameterToKernel(k, 0, VX_INPUT, VX_TYPE_IMAGE, VX_PARAMETER_STATE_REQUIRED));
    CHECK(vxAddParameterToKernel(k, 1, VX_OUTPUT, VX_TYPE_IMAGE, VX_PARAMETER_STATE_REQUIRED));
    CHECK(vxFinalizeKernel(k));

    k = vxAddUserKernel(ctx, "com.example.add_sat", TEST_KERNEL(2), add_exec, 4, validate_add, add_init, add_deinit);
    if (!k) return VX_FAILURE;
    vx_size local = sizeof(AddState);
    CHECK(vxSetKernelAttribute(k, VX_KERNEL_LOCAL_DATA_SIZE, &local, sizeof(local)));
    CHECK(vxAddParameterToKernel(k, 0, VX_INPUT, VX_TYPE_IMAGE, VX_PARAMETER_STATE_REQUIRED));
    CHECK(vxAddParameterToKernel(k, 1, VX_INPUT, VX_TYPE_IMAGE, VX_PARAMETER_STATE_REQUIRED));
    CHECK(vxAddParameterToKernel(k, 2, VX_OUTPUT, VX_TYPE_IMAGE, VX_PARAMETER_STATE_REQUIRED));
    CHECK(vxAddParameterToKernel(k, 3, VX_OUTPUT, VX_TYPE_SCALAR, VX_PARAMETER_STATE_REQUIRED));
    CHECK(vxFinalizeKernel(k));

    k = vxAddUserKernel(ctx, "com.example.copy_gpubuf", TEST_KERNEL(3), copy_exec, 2, validate_unary, nullptr, nullptr);
    if (!k) return VX_FAILURE;
    vx_bool on = vx_true_e;
    CHECK(vxSetKernelAttribute(k, VX_KERNEL_ATTRIBUTE_AMD_GPU_BUFFER_ACCESS_ENABLE, &on, sizeof(on)));
    CHECK(vxAddParameterToKernel(k, 0, VX_INPUT, VX_TYPE_IMAGE, VX_PARAMETER_STATE_REQUIRED));
    CHECK(vxAddParameterToKernel(k, 1, VX_OUTPUT, VX_TYPE_IMAGE, VX_PARAMETER_STATE_REQUIRED));
    return vxFinalizeKernel(k);
}

extern "C" vx_status VX_API_CALL vxUnpublishKernels(vx_context ctx)
{
    for (const char *name : {"com.example.invert", "com.example.add_sat", "com.example.copy_gpubuf"}) {
        vx_kernel k = vxGetKernelByName(ctx, name);
        if (k) CHECK(vxRemoveKernel(k));
    }
    return VX_SUCCESS;
}
