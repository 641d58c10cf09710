"""The OpenVX engine on the CPU side: exported API, header ABI, GDF import, graph validation.

No kernel runs here; graphs that would need the GPU must fail loudly (there is no CPU fallback).
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(*headers):
    names = set()
    for h in headers:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"VX_API_CALL\s+(vx\w+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def vx():
    from mivisionx_amd import openvx
    return openvx


def test_engine_exports_every_declared_function(vx):
    lib = vx.lib()
    names = _declared("VX/vx.h", "VX/vxu.h")
    assert len(names) >= 100
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_compile_as_c99_with_openvx_abi(tmp_path):
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "c", "vx_abi_probe.c"), "-o", str(exe)])
    got = dict(line.split() for line in subprocess.check_output([str(exe)]).decode().splitlines())
    got = {k: int(v) for k, v in got.items()}
    # structure layouts of the OpenVX 1.3 headers and vx_ext_amd.h
    assert got["sizeof_keypoint"] == 28 and got["sizeof_addressing"] == 32 and got["sizeof_pixel_value"] == 16
    assert got["sizeof_border"] == 20 and got["sizeof_affinity"] == 16 and got["sizeof_perf"] == 64
    assert got["sizeof_rectangle"] == 16 and got["sizeof_coordinates2df"] == 8
    # enumeration values: fourcc codes, kernel ids, enum / attribute bases
    assert got["VX_DF_IMAGE_U8"] == 0x38303055 and got["VX_DF_IMAGE_RGB"] == 0x32424752 and got["VX_DF_IMAGE_IYUV"] == 0x56555949
    assert (got["VX_KERNEL_GAUSSIAN_3x3"], got["VX_KERNEL_CANNY_EDGE_DETECTOR"], got["VX_KERNEL_HARRIS_CORNERS"],
            got["VX_KERNEL_REMAP"]) == (0x13, 0x1B, 0x25, 0x28)
    assert got["VX_NORM_L1"] == 0x10000 and got["VX_INTERPOLATION_BILINEAR"] == 0x4001
    assert got["VX_THRESHOLD_TYPE_RANGE"] == 0xB001 and got["VX_BORDER_CONSTANT"] == 0xC001 and got["VX_READ_ONLY"] == 0x11001
    assert got["VX_IMAGE_WIDTH"] == 0x80F00 and got["VX_ARRAY_NUMITEMS"] == 0x80E01 and got["VX_NODE_BORDER"] == 0x80302
    assert got["VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT"] == 0xD80202 and got["VX_NODE_ATTRIBUTE_AMD_AFFINITY"] == 0xD80301
    assert got["VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER"] == 0xD80F06 and got["VX_MEMORY_TYPE_HIP"] == 0xE003
    assert got["AGO_TARGET_AFFINITY_GPU"] == 0x20 and got["VX_ERROR_INVALID_GRAPH"] == -18


def test_kernel_names_and_enums_resolve(vx):
    lib = vx.lib()
    ctx = lib.vxCreateContext()
    try:
        for name in ["org.khronos.openvx.gaussian_3x3", "org.khronos.openvx.canny_edge_detector",
                     "org.khronos.openvx.harris_corners", "com.amd.openvx.Gaussian_U8_U8_3x3",
                     "com.amd.openvx.ScaleGaussianHalf_U8_U8_5x5"]:
            assert lib.vxGetKernelByName(ctx, name.encode()), name
        assert lib.vxGetKernelByEnum(ctx, 0x13)               # VX_KERNEL_GAUSSIAN_3x3
        assert lib.vxGetKernelByEnum(ctx, 0xd000b4)           # VX_KERNEL_AMD_GAUSSIAN_U8_U8_3x3
    finally:
        h = ctypes.c_void_p(ctx)
        lib.vxReleaseContext(ctypes.byref(h))


GDF = """
# 4K vision graph (RGB -> luma -> scale / pyramid / Canny / Harris)
data rgb = image:3840,2160,RGB2:READ,frame.rgb
data yuv = image-virtual:3840,2160,IYUV
data luma = virtual-image:3840,2160,U008
data scaled = image:1920,1080,U008:WRITE,scaled.u8
data pyr = pyramid:5,HALF,3840,2160,U008
data edges = image:3840,2160,U008
data hyst = threshold:RANGE,U008,U008:INIT,80,100
data grad = scalar:INT32,3
data strength = scalar:FLOAT32,0.0005
data mind = scalar:FLOAT32,5.0
data sens = scalar:FLOAT32,0.04
data corners = array:KEYPOINT,65536
data count = scalar:SIZE,0
node org.khronos.openvx.color_convert rgb yuv
node org.khronos.openvx.channel_extract yuv !CHANNEL_Y luma
node org.khronos.openvx.scale_image luma scaled !BILINEAR
node org.khronos.openvx.gaussian_pyramid luma pyr
node org.khronos.openvx.canny_edge_detector luma hyst grad !NORM_L1 edges \\
     attr:BORDER_MODE:UNDEFINED
node org.khronos.openvx.harris_corners luma strength mind sens grad grad corners count attr:AFFINITY:GPU
"""


def test_gdf_import_builds_the_graph_and_reports_io(vx):
    lib = vx.lib()
    ctx = lib.vxCreateContext()
    try:
        g = lib.vxCreateGraph(ctx)
        seen = []
        assert vx.import_gdf(g, GDF, callback=lambda ref, name, params: seen.append((name, params))) == 0
        n = ctypes.c_uint32()
        assert lib.vxQueryGraph(g, vx.VX_GRAPH_NUMNODES, ctypes.byref(n), 4) == 0
        assert n.value == 6
        assert ("rgb", "READ,frame.rgb") in seen and ("scaled", "WRITE,scaled.u8") in seen
        assert len(seen) == 2   # INIT is applied by the importer itself
    finally:
        h = ctypes.c_void_p(ctx)
        lib.vxReleaseContext(ctypes.byref(h))


@pytest.mark.parametrize("text,status", [
    ("data a = image:64,64,U008\nnode org.khronos.openvx.no_such_kernel a a\n", -10),
    ("data a = image:64,64,U008\nnode org.khronos.openvx.gaussian_3x3 a undefined_b\n", -10),
    ("data a = image:64,64,U008\nfrobnicate a\n", -3),
    ("data a = tensor:4,1,1,1,U008\n", -1),
])
def test_gdf_import_rejects_bad_text(vx, text, status):
    lib = vx.lib()
    ctx = lib.vxCreateContext()
    try:
        g = lib.vxCreateGraph(ctx)
        assert vx.import_gdf(g, text) == status
    finally:
        h = ctypes.c_void_p(ctx)
        lib.vxReleaseContext(ctypes.byref(h))


def _verify(vx, text):
    lib = vx.lib()
    ctx = lib.vxCreateContext()
    try:
        g = lib.vxCreateGraph(ctx)
        assert vx.import_gdf(g, text) == 0
        return lib.vxVerifyGraph(g)
    finally:
        h = ctypes.c_void_p(ctx)
        lib.vxReleaseContext(ctypes.byref(h))


def test_validation_errors_come_before_any_device_work(vx):
    # wrong input format
    assert _verify(vx, "data a = image:64,64,S016\ndata b = image:64,64,U008\n"
                       "node org.khronos.openvx.gaussian_3x3 a b\n") == vx.VX_ERROR_INVALID_FORMAT
    # output size mismatch
    assert _verify(vx, "data a = image:64,64,U008\ndata b = image:32,64,U008\n"
                       "node org.khronos.openvx.box_3x3 a b\n") == vx.VX_ERROR_INVALID_DIMENSION
    # two writers of one image
    assert _verify(vx, "data a = image:64,64,U008\ndata b = image:64,64,U008\n"
                       "node org.khronos.openvx.gaussian_3x3 a b\nnode org.khronos.openvx.box_3x3 a b\n") == vx.VX_ERROR_MULTIPLE_WRITERS
    # a cycle
    assert _verify(vx, "data a = image:64,64,U008\ndata b = image:64,64,U008\n"
                       "node org.khronos.openvx.gaussian_3x3 a b\nnode org.khronos.openvx.box_3x3 b a\n") == vx.VX_ERROR_INVALID_GRAPH
    # a variant this engine does not provide is refused, not emulated
    assert _verify(vx, "data a = image:64,64,U008\ndata h = threshold:RANGE,U008,U008:INIT,10,20\n"
                       "data g = scalar:INT32,5\ndata e = image:64,64,U008\n"
                       "node org.khronos.openvx.canny_edge_detector a h g !NORM_L1 e\n") == vx.VX_ERROR_NOT_SUPPORTED


def test_valid_graph_without_a_gpu_fails_loudly(vx):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    status = _verify(vx, "data a = image:64,64,U008\ndata b = image:64,64,U008\nnode org.khronos.openvx.gaussian_3x3 a b\n")
    assert status == vx.VX_ERROR_NOT_SUPPORTED


def test_runvx_cli_reports_usage_and_bad_files():
    from mivisionx_amd import openvx
    assert os.path.exists(openvx.RUNVX)
    r = subprocess.run([openvx.RUNVX], capture_output=True)
    assert r.returncode == 1 and b"usage" in r.stderr
    r = subprocess.run([openvx.RUNVX, "/nonexistent.gdf"], capture_output=True)
    assert r.returncode == 1


def test_runvx_md5_matches_rfc1321(tmp_path):
    """runvx's image checksums are MD5 (the reference's CHasher); the engine's own implementation
    against hashlib on RFC 1321's test suite and on lengths around the 64-byte block edges"""
    import hashlib
    exe = tmp_path / "md5_probe"
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(ROOT, "tests", "c", "md5_probe.cpp"), "-o", str(exe)], check=True)
    msgs = ["", "a", "abc", "message digest", "abcdefghijklmnopqrstuvwxyz",
            "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789", "1234567890" * 8] + \
           ["x" * n for n in (55, 56, 57, 63, 64, 65, 119, 120, 128, 1000)]
    out = subprocess.run([str(exe)] + msgs, capture_output=True, check=True).stdout.decode().split()
    assert out == [hashlib.md5(m.encode()).hexdigest() for m in msgs]


def _verify_log(vx, text, flags=None):
    """vxVerifyGraph of a GDF with the log captured; the optimizer's choices (merges, supernodes) are logged
    before any device work, so they are visible without a GPU"""
    lib = vx.lib()
    ctx = lib.vxCreateContext()
    messages = []
    cb = vx.LOG_CB(lambda c_, ref, status, t: messages.append(t.decode()))
    try:
        lib.vxRegisterLogCallback(ctx, cb, 0)
        g = lib.vxCreateGraph(ctx)
        if flags is not None:
            f = ctypes.c_uint32(flags)
            assert lib.vxSetGraphAttribute(g, vx.VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS, ctypes.byref(f), 4) == 0
        assert vx.import_gdf(g, text) == 0
        status = lib.vxVerifyGraph(g)
        return status, "".join(messages)
    finally:
        h = ctypes.c_void_p(ctx)
        lib.vxReleaseContext(ctypes.byref(h))


SOBEL_VIRTUAL = """data a = image:333,64,U008
data gx = image-virtual:333,64,S016
data gy = image-virtual:333,64,S016
data m = image:333,64,S016
data p = image:333,64,U008
node org.khronos.openvx.sobel_3x3 a gx gy
"""


def test_sobel_merge_rules_select_the_merged_kernels(vx):
    """ago_drama_remove.cpp:155-182: virtual gx / gy read only by Magnitude and/or Phase merge into one kernel"""
    both = SOBEL_VIRTUAL + "node org.khronos.openvx.magnitude gx gy m\nnode org.khronos.openvx.phase gx gy p\n"
    _, log = _verify_log(vx, both)
    assert "into SobelMagnitudePhase_S16U8_U8_3x3" in log, log
    _, log = _verify_log(vx, SOBEL_VIRTUAL + "node org.khronos.openvx.magnitude gx gy m\n")
    assert "into SobelMagnitude_S16_U8_3x3" in log, log
    _, log = _verify_log(vx, SOBEL_VIRTUAL + "node org.khronos.openvx.phase gx gy p\n")
    assert "into SobelPhase_U8_U8_3x3" in log, log
    # the flag that turns node merging off
    _, log = _verify_log(vx, both, flags=0x08)
    assert "merged" not in log, log
    # gy read by another node: the gradients are not solitary, nothing merges
    extra = both + "data m2 = image:333,64,S016\nnode org.khronos.openvx.magnitude gy gx m2\n"
    _, log = _verify_log(vx, extra)
    assert "merged" not in log, log
    # non-virtual gradients never merge
    _, log = _verify_log(vx, both.replace("image-virtual:333,64,S016", "image:333,64,S016"))
    assert "merged" not in log, log


VISION = """data rgb = image:640,480,RGB2
data yuv = image-virtual:640,480,IYUV
data luma = image-virtual:0,0,VIRT
data scaled = image:320,240,U008
data pyr = pyramid:4,HALF,640,480,U008
data edges = image:640,480,U008
data hyst = threshold:RANGE,U008,U008:INIT,80,100
data grad = scalar:INT32,3
data strength = scalar:FLOAT32,0.0005
data mind = scalar:FLOAT32,5.0
data sens = scalar:FLOAT32,0.04
data corners = array:KEYPOINT,1000
data count = scalar:SIZE,0
node org.khronos.openvx.color_convert rgb yuv
node org.khronos.openvx.channel_extract yuv !CHANNEL_{ch} luma
node org.khronos.openvx.scale_image luma scaled !BILINEAR
node org.khronos.openvx.gaussian_pyramid luma pyr
node org.khronos.openvx.canny_edge_detector luma hyst grad !NORM_L1 edges
node org.khronos.openvx.harris_corners luma strength mind sens grad grad corners count
"""


def test_vision_supernode_takes_the_luma_channel_only(vx):
    _, log = _verify_log(vx, VISION.format(ch="Y"))
    assert "fused vision supernode" in log, log
    # channel U of IYUV is a quarter-size plane: not the supernode's luma (its consumers then fail their size checks)
    _, log = _verify_log(vx, VISION.format(ch="U").replace("pyramid:4,HALF,640,480", "pyramid:4,HALF,320,240")
                         .replace("image:640,480,U008\ndata hyst", "image:320,240,U008\ndata hyst")
                         .replace("image:320,240,U008\ndata pyr", "image:160,120,U008\ndata pyr"))
    assert "fused vision supernode" not in log, log


def test_scheduled_graph_refuses_changes(vx):
    """while a vxScheduleGraph worker may run the graph, setters that would free its device state are refused"""
    lib = vx.lib()
    ctx = lib.vxCreateContext()
    try:
        g = lib.vxCreateGraph(ctx)
        assert vx.import_gdf(g, "data a = image:64,64,U008\ndata b = image:64,64,U008\n"
                                "node org.khronos.openvx.gaussian_3x3 a b\n") == 0
        assert lib.vxScheduleGraph(g) == 0
        f = ctypes.c_uint32(0)
        assert lib.vxSetGraphAttribute(g, vx.VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS, ctypes.byref(f), 4) == -21  # VX_ERROR_GRAPH_SCHEDULED
        lib.vxWaitGraph(g)   # the worker's status (no GPU here: a verification failure) is not the point
        assert lib.vxSetGraphAttribute(g, vx.VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS, ctypes.byref(f), 4) == 0
    finally:
        h = ctypes.c_void_p(ctx)
        lib.vxReleaseContext(ctypes.byref(h))
