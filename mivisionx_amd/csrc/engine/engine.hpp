// engine.hpp -- internal objects of the MI355X OpenVX engine (libmvx_openvx.so).
//
// The engine is the host side of the vision-node path: the vx*/vxu* API, the kernel table with its
// validate / initialize / execute commands and target affinity, graph verification and execution,
// and the GDF text import.  It mirrors the reference AGO engine's object model
// (amd_openvx/openvx/ago/ago_internal.h: AgoData, AgoNode, AgoGraph, AgoKernel) with one
// difference that shapes everything else: every image, array and scalar that a node reads or
// writes lives in HBM, and every kernel runs on the GPU.  Host access (map / copy / query) is a
// synchronous copy, graph execution enqueues on one HIP stream per graph and only vxProcessGraph /
// vxWaitGraph synchronise.
#pragma once

#include <VX/vx.h>
#include <VX/vxu.h>
#include <vx_ext_amd.h>
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mvx_kernels.h"

namespace mvxe {

constexpr uint32_t MAGIC = 0x4558564du;  // "MVXE"
constexpr size_t GUARD = 64;              // zero bytes before and after every device buffer
constexpr uint32_t MAX_PARAMS = 12;

inline uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }
inline size_t align256(size_t v) { return (v + 255u) & ~(size_t)255u; }

void log(struct ::_vx_reference *ref, vx_status status, const char *fmt, ...);

// device memory owned by the engine: zero-filled, GUARD zero bytes on both sides
struct DevBuf {
    uint8_t *base = nullptr;  // hipMalloc'd
    size_t size = 0;          // usable bytes (after the front guard)
    uint8_t *ptr() const { return base ? base + GUARD : nullptr; }
    vx_status alloc(size_t bytes);
    void release();
};

struct ParamSpec {
    vx_enum dir;    // VX_INPUT / VX_OUTPUT / VX_BIDIRECTIONAL
    vx_enum type;   // VX_TYPE_*
    bool optional;
};

struct Node;
struct UserKernel;
using ValidateFn = vx_status (*)(::_vx_node *node);
using InitFn = vx_status (*)(::_vx_node *node);
using ExecFn = vx_status (*)(::_vx_node *node, hipStream_t stream);
using ValidRectFn = void (*)(::_vx_node *node);

// one row of the kernel table (the role of ago_kernel_list.cpp's entries and the command
// protocol of agoKernel_* functions: validate fills output meta data, initialize sizes the node's
// device workspace, execute enqueues on the graph's stream and never synchronises)
struct KernelDef {
    vx_enum id;
    const char *name;
    std::vector<ParamSpec> params;
    ValidateFn validate;
    InitFn initialize;  // may be null (no workspace)
    ExecFn execute;
    uint32_t targets;   // the query_target_support answer of a built-in kernel: AGO_TARGET_AFFINITY_GPU
    ValidRectFn validRect = nullptr;  // valid_rect_callback: output valid rectangles from the inputs' (null: unchanged)
    UserKernel *user = nullptr;       // set for kernels added with vxAddUserKernel
};

// a kernel added with vxAddUserKernel (vx_api.cpp:2584-2810 of the reference): its callbacks and attributes
struct UserKernel {
    vx_kernel_f func = nullptr;
    vx_kernel_validate_f validate = nullptr;
    vx_kernel_initialize_f init = nullptr;
    vx_kernel_deinitialize_f deinit = nullptr;
    amd_kernel_query_target_support_f queryTarget = nullptr;  // null: CPU only
    bool gpuBufferAccess = false;   // VX_KERNEL_ATTRIBUTE_AMD_GPU_BUFFER_ACCESS_ENABLE
    vx_size localDataSize = 0;      // VX_KERNEL_LOCAL_DATA_SIZE: host bytes given to each node
    bool finalized = false;
    int module = -1;                // index in the context's module list, -1 when added by the application
    std::string name;
};

// the targets a node's kernel supports (its query_target_support command); 0 with *status set on error
uint32_t target_support(::_vx_node *node, vx_status *status);
// output valid rectangles of every node, in execution order (agoComputeImageValidRectangleOutputs)
void propagate_valid_rects(::_vx_graph *graph);
// user kernels: release what initialize made (deinitialize callback, host local data)
void deinit_user_node(::_vx_node *node);

const std::vector<KernelDef> &kernel_table();

// AMD low-level kernel enums (ago_kernels.h, enum vx_kernel_amd_e)
enum : vx_enum {
    AMD_REMAP_U8_U8_NEAREST = 0xd000e4,
    AMD_REMAP_U8_U8_NEAREST_CONSTANT = 0xd000e5,
    AMD_REMAP_U8_U8_BILINEAR = 0xd000e6,
    AMD_REMAP_U8_U8_BILINEAR_CONSTANT = 0xd000e7,
    AMD_WARP_AFFINE_U8_U8_NEAREST = 0xd000e8,
    AMD_WARP_AFFINE_U8_U8_NEAREST_CONSTANT = 0xd000e9,
    AMD_WARP_AFFINE_U8_U8_BILINEAR = 0xd000ea,
    AMD_WARP_AFFINE_U8_U8_BILINEAR_CONSTANT = 0xd000eb,
    AMD_WARP_PERSPECTIVE_U8_U8_NEAREST = 0xd000ec,
    AMD_WARP_PERSPECTIVE_U8_U8_NEAREST_CONSTANT = 0xd000ed,
    AMD_WARP_PERSPECTIVE_U8_U8_BILINEAR = 0xd000ee,
    AMD_WARP_PERSPECTIVE_U8_U8_BILINEAR_CONSTANT = 0xd000ef,
    AMD_MAGNITUDE_S16_S16S16 = 0xd0006f,
    AMD_PHASE_U8_S16S16 = 0xd00070,
    AMD_CHANNEL_COPY_U8_U8 = 0xd00071,
    AMD_COLOR_CONVERT_Y_RGB = 0xd000a2,
    AMD_BOX_U8_U8_3x3 = 0xd000b0,
    AMD_MEDIAN_U8_U8_3x3 = 0xd000b3,
    AMD_GAUSSIAN_U8_U8_3x3 = 0xd000b4,
    AMD_FAST_CORNERS_XY_U8_SUPRESSION = 0xd000cc,
    AMD_FAST_CORNERS_XY_U8_NOSUPRESSION = 0xd000cd,
    AMD_SCALE_GAUSSIAN_HALF_U8_U8_3x3 = 0xd000b5,
    AMD_SCALE_GAUSSIAN_HALF_U8_U8_5x5 = 0xd000b6,
    AMD_SOBEL_S16S16_U8_3x3_GXY = 0xd000c3,
    AMD_SOBEL_PHASE_U8_U8_3x3 = 0xd000b8,
    AMD_SOBEL_MAGNITUDE_S16_U8_3x3 = 0xd000c1,
    AMD_SOBEL_MAGNITUDE_PHASE_S16U8_U8_3x3 = 0xd000c2,
    AMD_CANNY_SOBEL_U16_U8_3x3_L1NORM = 0xd000da,
    AMD_CANNY_SOBEL_U16_U8_3x3_L2NORM = 0xd000db,
    AMD_CANNY_SOBEL_U16_U8_5x5_L1NORM = 0xd000dc,
    AMD_CANNY_SOBEL_U16_U8_5x5_L2NORM = 0xd000dd,
    AMD_CANNY_SOBEL_U16_U8_7x7_L1NORM = 0xd000de,
    AMD_CANNY_SOBEL_U16_U8_7x7_L2NORM = 0xd000df,
    AMD_SCALE_IMAGE_U8_U8_NEAREST = 0xd000f0,
    AMD_SCALE_IMAGE_U8_U8_BILINEAR = 0xd000f1,
    AMD_SCALE_IMAGE_U8_U8_BILINEAR_REPLICATE = 0xd000f2,
    AMD_SCALE_IMAGE_U8_U8_BILINEAR_CONSTANT = 0xd000f3,
    AMD_SCALE_IMAGE_U8_U8_AREA = 0xd000f4,
};

// what executes a node: its own execute command, one of the graph's supernodes, or (FUSED_MERGE) the
// merged kernel of the Sobel node whose virtual gradients it reads (the reference's node-merge rules)
enum FusedKind : int { FUSED_NONE = 0, FUSED_VISION = 1, FUSED_CHAIN = 2, FUSED_MERGE = 3 };

}  // namespace mvxe

struct _vx_reference {
    uint32_t magic = mvxe::MAGIC;
    vx_enum type;
    struct _vx_context *context;
    struct _vx_graph *scope = nullptr;  // owning graph of virtual objects
    std::atomic<int> extCount{1};
    std::atomic<int> intCount{0};
    bool isVirtual = false;
    std::string name;
    _vx_reference(vx_enum t, _vx_context *c) : type(t), context(c) {}
    virtual ~_vx_reference() {}
};

struct ModuleInfo {        // a module loaded by vxLoadKernels
    std::string name, path;
    void *handle;
};

struct _vx_context : _vx_reference {
    std::recursive_mutex lock;
    std::vector<ModuleInfo> modules;
    int importingModule = -1;         // module whose vxPublishKernels is running
    std::vector<struct _vx_kernel *> removedKernels;   // vxRemoveKernel'd, freed with the context
    vx_enum nextUserKernelId = 0, nextUserLibraryId = 1;
    std::vector<_vx_reference *> refs;  // every live object
    AgoTargetAffinityInfo affinity{};
    vx_border_t immediateBorder{VX_BORDER_UNDEFINED, {}};
    vx_log_callback_f logCallback = nullptr;
    int hipDevice = -1;
    bool hipReady = false;
    std::vector<struct _vx_kernel *> kernels;
    _vx_context() : _vx_reference(VX_TYPE_CONTEXT, nullptr) { context = this; }
    vx_status ensure_hip();
};

struct ImagePlane {
    uint32_t width, height, bpp, stride;
    size_t offset;  // from the image's device base
};

struct ImageMap {
    vx_map_id id;
    uint32_t plane;
    vx_rectangle_t rect;
    vx_enum usage;
    std::vector<uint8_t> host;
    vx_imagepatch_addressing_t addr;
};

struct _vx_image : _vx_reference {
    uint32_t width = 0, height = 0;
    vx_df_image format = VX_DF_IMAGE_VIRT;
    vx_enum space = VX_COLOR_SPACE_BT709, range = VX_CHANNEL_RANGE_FULL;
    std::vector<ImagePlane> planes;
    mvxe::DevBuf dev;
    uint8_t *external = nullptr;  // VX_MEMORY_TYPE_HIP handle: caller-owned device memory
    bool uniform = false;
    vx_pixel_value_t uniformValue{};
    struct _vx_pyramid *parent = nullptr;
    vx_rectangle_t validRect{0, 0, 0, 0};
    bool validRectSet = false;           // false: the whole image is valid
    std::vector<ImageMap> maps;
    vx_map_id nextMap = 1;
    _vx_image(_vx_context *c) : _vx_reference(VX_TYPE_IMAGE, c) {}
    bool layout();                       // derive planes from width/height/format
    vx_status allocate();                // device memory (idempotent)
    uint8_t *plane_ptr(uint32_t p);      // device address of plane p (null before allocation)
    size_t bytes() const;
};

struct _vx_scalar : _vx_reference {
    vx_enum dataType;
    union { vx_uint64 u64; vx_int32 i32; vx_uint32 u32; vx_float32 f32; vx_float64 f64; vx_enum e; vx_size sz;
            vx_uint8 u8; vx_bool b; vx_df_image df; } value{};
    mvxe::DevBuf dev;        // device shadow for scalars that kernels write (8 bytes)
    bool deviceWritten = false;
    _vx_scalar(_vx_context *c, vx_enum t) : _vx_reference(VX_TYPE_SCALAR, c), dataType(t) {}
    size_t item_size() const;
};

struct _vx_threshold : _vx_reference {
    vx_enum threshType;
    vx_enum dataType = VX_TYPE_UINT8;
    vx_df_image inputFormat = VX_DF_IMAGE_U8, outputFormat = VX_DF_IMAGE_U8;
    vx_pixel_value_t value{}, lower{}, upper{}, trueValue{}, falseValue{};
    _vx_threshold(_vx_context *c, vx_enum t) : _vx_reference(VX_TYPE_THRESHOLD, c), threshType(t) {}
};

struct _vx_array : _vx_reference {
    vx_enum itemType;
    vx_size itemSize, capacity;
    vx_size numItems = 0;     // host copy of the device count, refreshed after graphs write it
    mvxe::DevBuf dev;         // [count u32 | pad to 256 | items]
    std::vector<uint8_t> mapHost;
    vx_map_id mapId = 0;
    vx_size mapStart = 0, mapEnd = 0;
    vx_enum mapUsage = 0;
    _vx_array(_vx_context *c, vx_enum t, vx_size isz, vx_size cap)
        : _vx_reference(VX_TYPE_ARRAY, c), itemType(t), itemSize(isz), capacity(cap) {}
    vx_status allocate();
    uint32_t *count_ptr() { return (uint32_t *)dev.ptr(); }
    uint8_t *items_ptr() { return dev.ptr() ? dev.ptr() + 256 : nullptr; }
    vx_status pull_count();   // numItems <- device count (synchronous)
    vx_status push_count();   // device count <- numItems (synchronous)
};

struct _vx_pyramid : _vx_reference {
    vx_size levels;
    vx_float32 scale;
    uint32_t width, height;
    vx_df_image format;
    std::vector<_vx_image *> images;
    _vx_pyramid(_vx_context *c) : _vx_reference(VX_TYPE_PYRAMID, c) {}
    bool build_levels();
};

struct _vx_matrix : _vx_reference {
    vx_enum dataType;
    vx_size columns, rows;
    std::vector<uint8_t> data;
    _vx_matrix(_vx_context *c) : _vx_reference(VX_TYPE_MATRIX, c) {}
};

struct _vx_remap : _vx_reference {
    uint32_t srcWidth, srcHeight, dstWidth, dstHeight;
    std::vector<float> map;  // dst h x w x {x, y}, as set through the API
    bool dirty = true;       // host table changed since the device copies were made
    mvxe::DevBuf devFloat;   // the float table in HBM
    mvxe::DevBuf devFixed;   // {u16 x, u16 y} 13.3 fixed point, dstWidth*4 bytes per row
    _vx_remap(_vx_context *c) : _vx_reference(VX_TYPE_REMAP, c) {}
    vx_status sync(hipStream_t stream);  // refresh the device tables when dirty (enqueues the conversion)
};

struct _vx_kernel : _vx_reference {
    const mvxe::KernelDef *def;
    mvxe::KernelDef own{};           // a user kernel's own table row (def == &own)
    mvxe::UserKernel user;
    _vx_kernel(_vx_context *c, const mvxe::KernelDef *d) : _vx_reference(VX_TYPE_KERNEL, c), def(d) {}
};

// the meta data a user kernel's validate callback gives each output (vxSetMetaFormatAttribute)
struct _vx_meta_format : _vx_reference {
    vx_enum objType = 0;
    uint32_t width = 0, height = 0;
    vx_df_image format = VX_DF_IMAGE_VIRT;
    vx_enum itemType = 0, scalarType = 0;
    vx_size capacity = 0, levels = 0;
    vx_float32 scale = 0.f;
    vx_kernel_image_valid_rectangle_f validRect = nullptr;
    _vx_meta_format(_vx_context *c) : _vx_reference(VX_TYPE_META_FORMAT, c) {}
    ~_vx_meta_format() { magic = 0; }
};

struct _vx_parameter : _vx_reference {
    _vx_kernel *kernel;
    struct _vx_node *node;
    uint32_t index;
    _vx_parameter(_vx_context *c) : _vx_reference(VX_TYPE_PARAMETER, c) {}
};

struct _vx_node : _vx_reference {
    const mvxe::KernelDef *kernel;
    struct _vx_graph *graph;
    std::vector<_vx_reference *> params;
    vx_border_t border{VX_BORDER_UNDEFINED, {}};
    AgoTargetAffinityInfo affinity{};
    vx_status status = VX_SUCCESS;
    vx_perf_t perf{};
    mvxe::DevBuf local;       // node workspace sized by the kernel's initialize command
    size_t localSize = 0;
    uint32_t target = AGO_TARGET_AFFINITY_GPU;
    int fused = mvxe::FUSED_NONE;  // executed by one of the graph's supernodes (FusedKind)
    // a Sobel node whose virtual gx / gy only these nodes read: it runs the merged kernel
    // (SobelMagnitudePhase / SobelMagnitude / SobelPhase) and writes their outputs itself
    struct _vx_node *mergedMag = nullptr, *mergedPhase = nullptr;
    uint32_t variant = 0;     // kernel choice made by the validate command (per kernel; 0 = default)
    // user kernels: host local data (VX_NODE_LOCAL_DATA_SIZE / _PTR), initialize state, valid-rect callbacks
    void *localPtr = nullptr;
    vx_size userLocalSize = 0;
    std::vector<uint8_t> localHost;
    bool userInitialized = false;
    std::vector<vx_kernel_image_valid_rectangle_f> validRectCb;
    _vx_node(_vx_context *c) : _vx_reference(VX_TYPE_NODE, c) {}
};

namespace mvxe {
struct FusedVision;  // the RGB vision supernode (vx_graph.cpp)
struct FusedChains;  // the batched Gaussian -> Sobel -> Magnitude/Phase (+ScaleImage) supernode (vx_graph.cpp)
}

struct _vx_graph : _vx_reference {
    std::vector<_vx_node *> nodes;
    std::vector<_vx_node *> order;  // topological execution order after verification
    bool verified = false;
    vx_enum state = VX_GRAPH_STATE_UNVERIFIED;
    hipStream_t stream = nullptr;
    AgoTargetAffinityInfo affinity{};
    uint32_t optimizerFlags = 0;
    bool allowMerge = true;           // node-merge rules on (no AGO_GRAPH_OPTIMIZER_FLAG_NO_NODE_MERGE)
    vx_perf_t perf{};
    mvxe::FusedVision *fused = nullptr;
    int fusedAt = -1;                 // position in `order` where the supernode runs
    mvxe::FusedChains *chains = nullptr;
    int chainsAt = -1;                // position in `order` where the chain batch runs
    std::vector<_vx_array *> writtenArrays;
    std::vector<_vx_scalar *> writtenScalars;
    // vxScheduleGraph worker
    std::thread worker;
    std::mutex runLock;               // held while the graph verifies or executes, and while it is invalidated
    std::atomic<bool> scheduled{false};
    vx_status lastStatus = VX_SUCCESS;
    _vx_graph(_vx_context *c) : _vx_reference(VX_TYPE_GRAPH, c) {}
};

namespace mvxe {
bool is_ref(const void *p, vx_enum type = 0);
void add_ref(_vx_context *ctx, _vx_reference *r);
void release_ref(_vx_reference *r, bool external);
_vx_image *image_of(_vx_reference *r);
vx_status import_graph_text(_vx_graph *graph, const AgoGraphImportInfo *info);
_vx_node *create_node(_vx_graph *graph, const KernelDef *def, std::initializer_list<_vx_reference *> params);
_vx_scalar *make_scalar(_vx_context *ctx, vx_enum type, const void *value);
vx_status verify_graph(_vx_graph *graph);
vx_status execute_graph(_vx_graph *graph);
void invalidate_graph(_vx_graph *graph);
const char *type_name(vx_enum type);
vx_enum enum_from_name(const char *name, bool *ok);
vx_df_image format_from_name(const char *name);
}  // namespace mvxe
