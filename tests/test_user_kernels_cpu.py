"""The plug-in surface on the CPU side: a user-kernel module (tests/hip/user_module.hip, built by
__graft_entry__.build()) loaded with vxLoadKernels, the kernels it publishes with vxAddUserKernel /
vxAddParameterToKernel / vxSetKernelAttribute / vxFinalizeKernel, their parameter queries, the
refusals the reference makes (duplicate kernels, changes after finalization, missing modules), the
target decision made from their query_target_support callbacks, and vxUnloadKernels.  Nothing here
runs a kernel."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE = os.path.join(ROOT, "tests", "hip", "libmvx_test_user.so")
c = ctypes


@pytest.fixture(scope="module")
def vx():
    from mivisionx_amd import openvx
    if not os.path.exists(MODULE):
        pytest.skip("the test module is built by __graft_entry__.build()")
    L = openvx.lib()
    # loaded once by path, so that vxLoadKernels("mvx_test_user") finds it by its soname
    c.CDLL(MODULE)
    for name, (res, args) in {
        "vxLoadKernels": (c.c_int32, [c.c_void_p, c.c_char_p]), "vxUnloadKernels": (c.c_int32, [c.c_void_p, c.c_char_p]),
        "vxAddUserKernel": (c.c_void_p, [c.c_void_p, c.c_char_p, c.c_int32, c.c_void_p, c.c_uint32, c.c_void_p, c.c_void_p, c.c_void_p]),
        "vxAddParameterToKernel": (c.c_int32, [c.c_void_p, c.c_uint32, c.c_int32, c.c_int32, c.c_int32]),
        "vxFinalizeKernel": (c.c_int32, [c.c_void_p]), "vxRemoveKernel": (c.c_int32, [c.c_void_p]),
        "vxSetKernelAttribute": (c.c_int32, [c.c_void_p, c.c_int32, c.c_void_p, c.c_size_t]),
        "vxQueryKernel": (c.c_int32, [c.c_void_p, c.c_int32, c.c_void_p, c.c_size_t]),
        "vxGetKernelParameterByIndex": (c.c_void_p, [c.c_void_p, c.c_uint32]),
        "vxQueryParameter": (c.c_int32, [c.c_void_p, c.c_int32, c.c_void_p, c.c_size_t]),
        "vxReleaseParameter": (c.c_int32, [c.POINTER(c.c_void_p)]),
        "vxAllocateUserKernelId": (c.c_int32, [c.c_void_p, c.POINTER(c.c_int32)]),
    }.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return openvx


@pytest.fixture()
def ctx(vx):
    L = vx.lib()
    h = L.vxCreateContext()
    yield h
    hh = c.c_void_p(h)
    L.vxReleaseContext(c.byref(hh))


KERNEL_BASE = ((0xFFE << 20) | (0x3 << 12)) - (1 << 32)   # VX_KERNEL_BASE(VX_ID_USER, 3) as a signed vx_enum
VX_INPUT, VX_OUTPUT = 0x0, 0x1          # VX_ENUM_BASE(VX_ID_KHRONOS, VX_ENUM_DIRECTION = 0) + 0 / + 1
REQUIRED = 0x8000                      # VX_ENUM_BASE(VX_ID_KHRONOS, VX_ENUM_PARAMETER_STATE = 8)
VX_TYPE_IMAGE, VX_TYPE_SCALAR = 0x80F, 0x80D
VX_KERNEL_PARAMETERS, VX_KERNEL_ENUM, VX_KERNEL_LOCAL_DATA_SIZE = 0x80400, 0x80402, 0x80403
VX_PARAMETER_DIRECTION, VX_PARAMETER_TYPE = 0x80501, 0x80502
VX_CONTEXT_MODULES = 0x80103
VX_ERROR_INVALID_MODULE = -11


def _query_u32(fn, obj, attr):
    v = c.c_uint32()
    assert fn(obj, attr, c.byref(v), 4) == 0
    return v.value


def test_module_publishes_its_kernels(vx, ctx):
    L = vx.lib()
    assert L.vxLoadKernels(ctx, b"mvx_test_user") == 0
    assert _query_u32(L.vxQueryContext, ctx, VX_CONTEXT_MODULES) == 1
    assert L.vxLoadKernels(ctx, b"mvx_test_user") == 0        # a second load is a warning, not an error
    for i, (name, nparams) in enumerate((("com.example.invert", 2), ("com.example.add_sat", 4), ("com.example.copy_gpubuf", 2))):
        k = L.vxGetKernelByName(ctx, name.encode())
        assert k and L.vxGetKernelByEnum(ctx, KERNEL_BASE + 1 + i) == k
        assert _query_u32(L.vxQueryKernel, k, VX_KERNEL_PARAMETERS) == nparams
        assert _query_u32(L.vxQueryKernel, k, VX_KERNEL_ENUM) == (KERNEL_BASE + 1 + i) & 0xFFFFFFFF
    k = L.vxGetKernelByName(ctx, b"com.example.add_sat")
    sz = c.c_size_t()
    assert L.vxQueryKernel(k, VX_KERNEL_LOCAL_DATA_SIZE, c.byref(sz), c.sizeof(sz)) == 0 and sz.value == 8
    dirs = []
    for i in range(4):
        p = L.vxGetKernelParameterByIndex(k, i)
        dirs.append((_query_u32(L.vxQueryParameter, p, VX_PARAMETER_DIRECTION), _query_u32(L.vxQueryParameter, p, VX_PARAMETER_TYPE)))
        pp = c.c_void_p(p)
        L.vxReleaseParameter(c.byref(pp))
    assert dirs == [(VX_INPUT, VX_TYPE_IMAGE), (VX_INPUT, VX_TYPE_IMAGE), (VX_OUTPUT, VX_TYPE_IMAGE), (VX_OUTPUT, VX_TYPE_SCALAR)]
    # vxUnpublishKernels removes them; the module no longer counts
    assert L.vxUnloadKernels(ctx, b"mvx_test_user") == 0
    assert not L.vxGetKernelByName(ctx, b"com.example.invert")
    assert _query_u32(L.vxQueryContext, ctx, VX_CONTEXT_MODULES) == 0


def test_missing_module_is_an_invalid_module(vx, ctx):
    assert vx.lib().vxLoadKernels(ctx, b"no_such_module_anywhere") == VX_ERROR_INVALID_MODULE


def test_user_kernel_rules(vx, ctx):
    """duplicate names / enums are refused; parameters cannot change after vxFinalizeKernel, and a kernel
    is finalized only when every parameter has been added; built-in kernels take no user attributes"""
    L = vx.lib()
    KFUNC = c.CFUNCTYPE(c.c_int32, c.c_void_p, c.c_void_p, c.c_uint32)
    VFUNC = c.CFUNCTYPE(c.c_int32, c.c_void_p, c.c_void_p, c.c_uint32, c.c_void_p)
    fn, val = KFUNC(lambda *a: 0), VFUNC(lambda *a: 0)
    kid = c.c_int32()
    assert L.vxAllocateUserKernelId(ctx, c.byref(kid)) == 0 and ((kid.value & 0xFFFFFFFF) >> 20) == 0xFFE
    k = L.vxAddUserKernel(ctx, b"com.test.k", kid.value, c.cast(fn, c.c_void_p), 2, c.cast(val, c.c_void_p), None, None)
    assert k
    assert not L.vxAddUserKernel(ctx, b"com.test.k", kid.value + 1, c.cast(fn, c.c_void_p), 2, c.cast(val, c.c_void_p), None, None)
    assert not L.vxAddUserKernel(ctx, b"com.test.other", kid.value, c.cast(fn, c.c_void_p), 2, c.cast(val, c.c_void_p), None, None)
    assert not L.vxAddUserKernel(ctx, b"org.khronos.openvx.gaussian_3x3", kid.value + 2, c.cast(fn, c.c_void_p), 2,
                                 c.cast(val, c.c_void_p), None, None)
    assert L.vxAddParameterToKernel(k, 0, VX_INPUT, VX_TYPE_IMAGE, REQUIRED) == 0
    assert L.vxFinalizeKernel(k) != 0                         # parameter 1 is missing
    assert L.vxAddParameterToKernel(k, 2, VX_INPUT, VX_TYPE_IMAGE, REQUIRED) != 0   # beyond numParams
    assert L.vxAddParameterToKernel(k, 1, VX_OUTPUT, VX_TYPE_IMAGE, REQUIRED) == 0
    assert L.vxFinalizeKernel(k) == 0
    assert L.vxAddParameterToKernel(k, 1, VX_OUTPUT, VX_TYPE_IMAGE, REQUIRED) != 0  # finalized
    gk = L.vxGetKernelByName(ctx, b"org.khronos.openvx.gaussian_3x3")
    sz = c.c_size_t(16)
    assert L.vxSetKernelAttribute(gk, VX_KERNEL_LOCAL_DATA_SIZE, c.byref(sz), c.sizeof(sz)) != 0
    assert L.vxRemoveKernel(gk) != 0 and L.vxRemoveKernel(k) == 0
    assert not L.vxGetKernelByName(ctx, b"com.test.k")


def test_target_decision_from_query_target_support(vx, ctx):
    """a GPU-only user kernel asked to run on the CPU falls back to the GPU with the reference's warning;
    the CPU user kernel is placed on the CPU even in a GPU-affinity graph.  Logged before any device work."""
    L = vx.lib()
    assert L.vxLoadKernels(ctx, b"mvx_test_user") == 0
    messages = []
    cb = vx.LOG_CB(lambda c_, ref, status, t: messages.append(t.decode()))
    L.vxRegisterLogCallback(ctx, cb, 0)
    g = L.vxCreateGraph(ctx)
    # the graph asks for the GPU explicitly: the reference engine that other tests load as their checker sets
    # AGO_DEFAULT_TARGET=CPU in this process's environment
    text = ("affinity GPU\n"
            "data a = image:64,32,U008\ndata b = image:64,32,U008\ndata s = image:64,32,U008\ndata n = scalar:UINT32,0\n"
            "node com.example.invert a b attr:AFFINITY:CPU\nnode com.example.add_sat a b s n\n")
    assert vx.import_gdf(g, text) == 0
    L.vxVerifyGraph(g)    # without a GPU this stops after the target decision
    log = "".join(messages)
    assert "kernel com.example.invert not supported on CPU -- falling back to GPU" in log, log
    assert "com.example.add_sat not supported on GPU -- falling back to CPU" in log, log
