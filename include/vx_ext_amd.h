/*
 * vx_ext_amd.h -- the AMD OpenVX extension subset the MI355X engine keeps for drop-in use:
 * target affinity (AgoTargetAffinityInfo, AGO_TARGET_AFFINITY_*), graph import from GDF text,
 * HIP device buffers on images and the HIP stream of a node.  Values and layouts are those of the
 * reference's amd_openvx/openvx/include/vx_ext_amd.h (AGO_TARGET_AFFINITY_CPU/GPU :43-47,
 * AgoTargetAffinityInfo :358-364, AgoGraphImportInfo :382-390, attribute enums :126-222), so that
 * applications and runvx-style tools written against the reference compile and behave the same.
 */
#ifndef _VX_EXT_AMD_H_
#define _VX_EXT_AMD_H_

#include <VX/vx.h>

#ifdef __cplusplus
extern "C" {
#endif

/* target affinity: AgoTargetAffinityInfo.device_type */
#define AGO_TARGET_AFFINITY_CPU 0x0010
#define AGO_TARGET_AFFINITY_GPU 0x0020

typedef struct {
    vx_uint32 device_type;  /* AGO_TARGET_AFFINITY_CPU or AGO_TARGET_AFFINITY_GPU (0 = inherit) */
    vx_uint32 device_info;  /* GPU index */
    vx_uint32 group;        /* fusion group (nodes of one group may be launched as one kernel) */
    vx_uint32 reserved;
} AgoTargetAffinityInfo;

/* GDF text import: "file <name>" loads a file, otherwise the text is parsed as is.  $1..$num_ref in
 * the text name ref[0..num_ref-1]; for every data object whose description carries a third
 * ':'-separated field (runvx I/O such as ":READ,file"), data_registry_callback_f is called with the
 * object, its name and that field. */
typedef struct {
    vx_char *text;
    vx_uint32 num_ref;
    vx_reference *ref;
    vx_int32 dumpToConsole;
    void(VX_CALLBACK *data_registry_callback_f)(void *obj, vx_reference ref, const char *name, const char *app_params);
    void *data_registry_callback_obj;
} AgoGraphImportInfo;

/* user-kernel attributes (vx_ext_amd.h:149-160 of the reference) */
enum vx_kernel_attribute_amd_e {
    VX_KERNEL_ATTRIBUTE_AMD_QUERY_TARGET_SUPPORT = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_KERNEL) + 0x01,
    VX_KERNEL_ATTRIBUTE_AMD_GPU_BUFFER_ACCESS_ENABLE = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_KERNEL) + 0x05,
};

enum vx_context_attribute_amd_e {
    VX_CONTEXT_ATTRIBUTE_AMD_AFFINITY = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_CONTEXT) + 0x02,
    VX_CONTEXT_ATTRIBUTE_AMD_HIP_DEVICE = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_CONTEXT) + 0x07,
};

enum vx_graph_attribute_amd_e {
    VX_GRAPH_ATTRIBUTE_AMD_AFFINITY = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_GRAPH) + 0x01,
    VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_GRAPH) + 0x02,
    VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_GRAPH) + 0x04,
};

enum vx_node_attribute_amd_e {
    VX_NODE_ATTRIBUTE_AMD_AFFINITY = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_NODE) + 0x01,
    VX_NODE_ATTRIBUTE_AMD_HIP_STREAM = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_NODE) + 0x03,
};

enum vx_image_attribute_amd_e {
    VX_IMAGE_ATTRIBUTE_AMD_GPU_BUFFER_STRIDE = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_IMAGE) + 0x04,
    VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER = VX_ATTRIBUTE_BASE(VX_ID_AMD, VX_TYPE_IMAGE) + 0x06,
};

enum vx_memory_type_amd_e {
    VX_MEMORY_TYPE_HIP = VX_ENUM_BASE(VX_ID_KHRONOS, VX_ENUM_MEMORY_TYPE) + 0x3,
};

/* VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS bits (values of ago/ago_internal.h:126,128).  Either one
 * makes every node run on its own kernels: no fused supernode. */
#define AGO_GRAPH_OPTIMIZER_FLAG_NO_NODE_MERGE 0x00000008
#define AGO_GRAPH_OPTIMIZER_FLAG_NO_SUPERNODE_MERGE 0x00000020

#ifdef __cplusplus
}

/* A user kernel's target-support callback (vx_ext_amd.h:436-438 of the reference): sets
 * supported_target_affinity to AGO_TARGET_AFFINITY_CPU, _GPU or both.  A kernel without one runs on the
 * CPU: its execute callback is called on the host after the graph's stream has drained, and reads and
 * writes its objects through the map / copy API.  A GPU kernel's execute callback runs on the host too,
 * without a drain, and enqueues device work on VX_NODE_ATTRIBUTE_AMD_HIP_STREAM over
 * VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER pointers; so does a CPU kernel with
 * VX_KERNEL_ATTRIBUTE_AMD_GPU_BUFFER_ACCESS_ENABLE set. */
typedef vx_status(VX_CALLBACK *amd_kernel_query_target_support_f)(vx_graph graph, vx_node node, vx_bool use_opencl_1_2,
                                                                  vx_uint32 &supported_target_affinity);
#endif

#endif
