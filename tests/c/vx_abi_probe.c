/* Compiled as C99 by tests/test_engine_cpu.py: the public headers are valid C and their ABI facts
 * (structure sizes, enumeration values) are the OpenVX 1.3 / AMD extension ones. */
#include <stdio.h>
#include <VX/vx.h>
#include <VX/vxu.h>
#include <vx_ext_amd.h>

#define P(name, value) printf("%s %lld\n", name, (long long)(value))

int main(void)
{
    P("sizeof_keypoint", sizeof(vx_keypoint_t));
    P("sizeof_addressing", sizeof(vx_imagepatch_addressing_t));
    P("sizeof_pixel_value", sizeof(vx_pixel_value_t));
    P("sizeof_border", sizeof(vx_border_t));
    P("sizeof_affinity", sizeof(AgoTargetAffinityInfo));
    P("sizeof_perf", sizeof(vx_perf_t));
    P("sizeof_rectangle", sizeof(vx_rectangle_t));
    P("sizeof_coordinates2df", sizeof(vx_coordinates2df_t));
    P("VX_DF_IMAGE_U8", VX_DF_IMAGE_U8);
    P("VX_DF_IMAGE_RGB", VX_DF_IMAGE_RGB);
    P("VX_DF_IMAGE_IYUV", VX_DF_IMAGE_IYUV);
    P("VX_KERNEL_GAUSSIAN_3x3", VX_KERNEL_GAUSSIAN_3x3);
    P("VX_KERNEL_CANNY_EDGE_DETECTOR", VX_KERNEL_CANNY_EDGE_DETECTOR);
    P("VX_KERNEL_HARRIS_CORNERS", VX_KERNEL_HARRIS_CORNERS);
    P("VX_KERNEL_REMAP", VX_KERNEL_REMAP);
    P("VX_NORM_L1", VX_NORM_L1);
    P("VX_INTERPOLATION_BILINEAR", VX_INTERPOLATION_BILINEAR);
    P("VX_THRESHOLD_TYPE_RANGE", VX_THRESHOLD_TYPE_RANGE);
    P("VX_BORDER_CONSTANT", VX_BORDER_CONSTANT);
    P("VX_READ_ONLY", VX_READ_ONLY);
    P("VX_IMAGE_WIDTH", VX_IMAGE_WIDTH);
    P("VX_ARRAY_NUMITEMS", VX_ARRAY_NUMITEMS);
    P("VX_NODE_BORDER", VX_NODE_BORDER);
    P("VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT", VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT);
    P("VX_NODE_ATTRIBUTE_AMD_AFFINITY", VX_NODE_ATTRIBUTE_AMD_AFFINITY);
    P("VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER", VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER);
    P("VX_MEMORY_TYPE_HIP", VX_MEMORY_TYPE_HIP);
    P("AGO_TARGET_AFFINITY_GPU", AGO_TARGET_AFFINITY_GPU);
    P("VX_ERROR_INVALID_GRAPH", VX_ERROR_INVALID_GRAPH);
    return 0;
}
