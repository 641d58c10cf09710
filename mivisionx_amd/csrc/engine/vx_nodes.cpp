// vx_nodes.cpp -- node constructors (vx*Node) and immediate-mode functions (vxu*) of the vision-node
// set, with the Khronos parameter order of the reference's api/vx_nodes.cpp:78-704 and api/vxu.cpp.
// Enum and integer arguments become scalars owned by the node, as in the reference.
#include "engine.hpp"

using namespace mvxe;

namespace {

const KernelDef *def_of(vx_enum id)
{
    for (const KernelDef &k : kernel_table())
        if (k.id == id) return &k;
    return nullptr;
}

// a node whose scalars created here are owned by the node alone
vx_node make(vx_graph graph, vx_enum id, std::initializer_list<_vx_reference *> params, std::initializer_list<_vx_reference *> owned = {})
{
    if (!is_ref(graph, VX_TYPE_GRAPH)) return nullptr;
    vx_node n = create_node(graph, def_of(id), params);
    for (_vx_reference *r : owned) release_ref(r, true);
    return n;
}

_vx_scalar *senum(vx_graph g, vx_enum v) { return is_ref(g, VX_TYPE_GRAPH) ? make_scalar(g->context, VX_TYPE_ENUM, &v) : nullptr; }
_vx_scalar *sint(vx_graph g, vx_int32 v) { return is_ref(g, VX_TYPE_GRAPH) ? make_scalar(g->context, VX_TYPE_INT32, &v) : nullptr; }
_vx_scalar *ssize(vx_graph g, vx_size v) { return is_ref(g, VX_TYPE_GRAPH) ? make_scalar(g->context, VX_TYPE_SIZE, &v) : nullptr; }

// vxu: a one-node graph, run to completion with the context's immediate border mode
template <typename F> vx_status immediate(vx_context context, F build)
{
    if (!is_ref(context, VX_TYPE_CONTEXT)) return VX_ERROR_INVALID_REFERENCE;
    vx_graph g = vxCreateGraph(context);
    if (!g) return VX_ERROR_NO_MEMORY;
    vx_node n = build(g);
    vx_status s = n ? VX_SUCCESS : VX_ERROR_INVALID_PARAMETERS;
    if (!s) s = vxSetNodeAttribute(n, VX_NODE_BORDER, &context->immediateBorder, sizeof(vx_border_t));
    if (!s) s = vxVerifyGraph(g);
    if (!s) s = vxProcessGraph(g);
    if (n) vxReleaseNode(&n);
    vxReleaseGraph(&g);
    return s;
}

}  // namespace

VX_API_ENTRY vx_node VX_API_CALL vxColorConvertNode(vx_graph graph, vx_image input, vx_image output)
{
    return make(graph, VX_KERNEL_COLOR_CONVERT, {input, output});
}

VX_API_ENTRY vx_node VX_API_CALL vxChannelExtractNode(vx_graph graph, vx_image input, vx_enum channel, vx_image output)
{
    _vx_scalar *c = senum(graph, channel);
    return make(graph, VX_KERNEL_CHANNEL_EXTRACT, {input, c, output}, {c});
}

VX_API_ENTRY vx_node VX_API_CALL vxSobel3x3Node(vx_graph graph, vx_image input, vx_image output_x, vx_image output_y)
{
    return make(graph, VX_KERNEL_SOBEL_3x3, {input, output_x, output_y});
}

VX_API_ENTRY vx_node VX_API_CALL vxMagnitudeNode(vx_graph graph, vx_image grad_x, vx_image grad_y, vx_image mag)
{
    return make(graph, VX_KERNEL_MAGNITUDE, {grad_x, grad_y, mag});
}

VX_API_ENTRY vx_node VX_API_CALL vxPhaseNode(vx_graph graph, vx_image grad_x, vx_image grad_y, vx_image orientation)
{
    return make(graph, VX_KERNEL_PHASE, {grad_x, grad_y, orientation});
}

VX_API_ENTRY vx_node VX_API_CALL vxScaleImageNode(vx_graph graph, vx_image src, vx_image dst, vx_enum type)
{
    _vx_scalar *t = senum(graph, type);
    return make(graph, VX_KERNEL_SCALE_IMAGE, {src, dst, t}, {t});
}

VX_API_ENTRY vx_node VX_API_CALL vxBox3x3Node(vx_graph graph, vx_image input, vx_image output)
{
    return make(graph, VX_KERNEL_BOX_3x3, {input, output});
}

VX_API_ENTRY vx_node VX_API_CALL vxGaussian3x3Node(vx_graph graph, vx_image input, vx_image output)
{
    return make(graph, VX_KERNEL_GAUSSIAN_3x3, {input, output});
}

VX_API_ENTRY vx_node VX_API_CALL vxMedian3x3Node(vx_graph graph, vx_image input, vx_image output)
{
    return make(graph, VX_KERNEL_MEDIAN_3x3, {input, output});
}

VX_API_ENTRY vx_node VX_API_CALL vxGaussianPyramidNode(vx_graph graph, vx_image input, vx_pyramid gaussian)
{
    return make(graph, VX_KERNEL_GAUSSIAN_PYRAMID, {input, gaussian});
}

VX_API_ENTRY vx_node VX_API_CALL vxCannyEdgeDetectorNode(vx_graph graph, vx_image input, vx_threshold hyst, vx_int32 gradient_size,
                                                        vx_enum norm_type, vx_image output)
{
    _vx_scalar *g = sint(graph, gradient_size), *n = senum(graph, norm_type);
    return make(graph, VX_KERNEL_CANNY_EDGE_DETECTOR, {input, hyst, g, n, output}, {g, n});
}

VX_API_ENTRY vx_node VX_API_CALL vxHarrisCornersNode(vx_graph graph, vx_image input, vx_scalar strength_thresh, vx_scalar min_distance,
                                                    vx_scalar sensitivity, vx_int32 gradient_size, vx_int32 block_size,
                                                    vx_array corners, vx_scalar num_corners)
{
    _vx_scalar *g = sint(graph, gradient_size), *b = sint(graph, block_size);
    return make(graph, VX_KERNEL_HARRIS_CORNERS, {input, strength_thresh, min_distance, sensitivity, g, b, corners, num_corners}, {g, b});
}

VX_API_ENTRY vx_node VX_API_CALL vxFastCornersNode(vx_graph graph, vx_image input, vx_scalar strength_thresh,
                                                  vx_bool nonmax_suppression, vx_array corners, vx_scalar num_corners)
{
    vx_bool nm = nonmax_suppression;
    _vx_scalar *b = is_ref(graph, VX_TYPE_GRAPH) ? make_scalar(graph->context, VX_TYPE_BOOL, &nm) : nullptr;
    return make(graph, VX_KERNEL_FAST_CORNERS, {input, strength_thresh, b, corners, num_corners}, {b});
}

VX_API_ENTRY vx_node VX_API_CALL vxWarpAffineNode(vx_graph graph, vx_image input, vx_matrix matrix, vx_enum type, vx_image output)
{
    _vx_scalar *t = senum(graph, type);
    return make(graph, VX_KERNEL_WARP_AFFINE, {input, matrix, t, output}, {t});
}

VX_API_ENTRY vx_node VX_API_CALL vxWarpPerspectiveNode(vx_graph graph, vx_image input, vx_matrix matrix, vx_enum type, vx_image output)
{
    _vx_scalar *t = senum(graph, type);
    return make(graph, VX_KERNEL_WARP_PERSPECTIVE, {input, matrix, t, output}, {t});
}

VX_API_ENTRY vx_node VX_API_CALL vxRemapNode(vx_graph graph, vx_image input, vx_remap table, vx_enum policy, vx_image output)
{
    _vx_scalar *p = senum(graph, policy);
    return make(graph, VX_KERNEL_REMAP, {input, table, p, output}, {p});
}

VX_API_ENTRY vx_node VX_API_CALL vxHalfScaleGaussianNode(vx_graph graph, vx_image input, vx_image output, vx_int32 kernel_size)
{
    _vx_scalar *k = sint(graph, kernel_size);
    return make(graph, VX_KERNEL_HALFSCALE_GAUSSIAN, {input, output, k}, {k});
}

VX_API_ENTRY vx_node VX_API_CALL vxOpticalFlowPyrLKNode(vx_graph graph, vx_pyramid old_images, vx_pyramid new_images, vx_array old_points,
                                                       vx_array new_points_estimates, vx_array new_points, vx_enum termination,
                                                       vx_scalar epsilon, vx_scalar num_iterations, vx_scalar use_initial_estimate,
                                                       vx_size window_dimension)
{
    _vx_scalar *t = senum(graph, termination), *w = ssize(graph, window_dimension);
    return make(graph, VX_KERNEL_OPTICAL_FLOW_PYR_LK,
                {old_images, new_images, old_points, new_points_estimates, new_points, t, epsilon, num_iterations, use_initial_estimate, w},
                {t, w});
}

// ------------------------------------------------------------------------------ immediate mode ---

VX_API_ENTRY vx_status VX_API_CALL vxuColorConvert(vx_context c, vx_image input, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxColorConvertNode(g, input, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuChannelExtract(vx_context c, vx_image input, vx_enum channel, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxChannelExtractNode(g, input, channel, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuSobel3x3(vx_context c, vx_image input, vx_image output_x, vx_image output_y)
{
    return immediate(c, [&](vx_graph g) { return vxSobel3x3Node(g, input, output_x, output_y); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuMagnitude(vx_context c, vx_image grad_x, vx_image grad_y, vx_image mag)
{
    return immediate(c, [&](vx_graph g) { return vxMagnitudeNode(g, grad_x, grad_y, mag); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuPhase(vx_context c, vx_image grad_x, vx_image grad_y, vx_image orientation)
{
    return immediate(c, [&](vx_graph g) { return vxPhaseNode(g, grad_x, grad_y, orientation); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuScaleImage(vx_context c, vx_image src, vx_image dst, vx_enum type)
{
    return immediate(c, [&](vx_graph g) { return vxScaleImageNode(g, src, dst, type); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuBox3x3(vx_context c, vx_image input, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxBox3x3Node(g, input, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuGaussian3x3(vx_context c, vx_image input, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxGaussian3x3Node(g, input, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuMedian3x3(vx_context c, vx_image input, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxMedian3x3Node(g, input, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuGaussianPyramid(vx_context c, vx_image input, vx_pyramid gaussian)
{
    return immediate(c, [&](vx_graph g) { return vxGaussianPyramidNode(g, input, gaussian); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuCannyEdgeDetector(vx_context c, vx_image input, vx_threshold hyst, vx_int32 gradient_size,
                                                       vx_enum norm_type, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxCannyEdgeDetectorNode(g, input, hyst, gradient_size, norm_type, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuFastCorners(vx_context c, vx_image input, vx_scalar strength_thresh, vx_bool nonmax_suppression,
                                                 vx_array corners, vx_scalar num_corners)
{
    return immediate(c, [&](vx_graph g) { return vxFastCornersNode(g, input, strength_thresh, nonmax_suppression, corners, num_corners); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuHarrisCorners(vx_context c, vx_image input, vx_scalar strength_thresh, vx_scalar min_distance,
                                                   vx_scalar sensitivity, vx_int32 gradient_size, vx_int32 block_size, vx_array corners,
                                                   vx_scalar num_corners)
{
    return immediate(c, [&](vx_graph g) {
        return vxHarrisCornersNode(g, input, strength_thresh, min_distance, sensitivity, gradient_size, block_size, corners, num_corners);
    });
}

VX_API_ENTRY vx_status VX_API_CALL vxuWarpAffine(vx_context c, vx_image input, vx_matrix matrix, vx_enum type, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxWarpAffineNode(g, input, matrix, type, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuWarpPerspective(vx_context c, vx_image input, vx_matrix matrix, vx_enum type, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxWarpPerspectiveNode(g, input, matrix, type, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuHalfScaleGaussian(vx_context c, vx_image input, vx_image output, vx_int32 kernel_size)
{
    return immediate(c, [&](vx_graph g) { return vxHalfScaleGaussianNode(g, input, output, kernel_size); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuRemap(vx_context c, vx_image input, vx_remap table, vx_enum policy, vx_image output)
{
    return immediate(c, [&](vx_graph g) { return vxRemapNode(g, input, table, policy, output); });
}

VX_API_ENTRY vx_status VX_API_CALL vxuOpticalFlowPyrLK(vx_context c, vx_pyramid old_images, vx_pyramid new_images, vx_array old_points,
                                                      vx_array new_points_estimates, vx_array new_points, vx_enum termination,
                                                      vx_scalar epsilon, vx_scalar num_iterations, vx_scalar use_initial_estimate,
                                                      vx_size window_dimension)
{
    return immediate(c, [&](vx_graph g) {
        return vxOpticalFlowPyrLKNode(g, old_images, new_images, old_points, new_points_estimates, new_points, termination, epsilon,
                                      num_iterations, use_initial_estimate, window_dimension);
    });
}
