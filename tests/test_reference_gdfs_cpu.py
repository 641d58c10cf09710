"""The reference's own GDF test graphs for this node set import unchanged through the engine's GDF
reader (tests/amd_openvx_gdfs/**, run by the reference with runOpenVX.py).  CPU only: the graphs are
parsed and built, not verified.  Skipped where the reference tree is absent (the GPU box)."""
import ctypes
import os

import pytest

REF = "/root/reference/tests/amd_openvx_gdfs"
# every GDF of the reference's suite whose nodes are in this node set
NODE_SET = [
    "arithmetic/Magnitude_S16_S16S16.gdf", "arithmetic/Phase_U8_S16S16.gdf", "color/ChannelCopy_U8_U8.gdf",
    "filter/Box_U8_U8_3x3.gdf", "filter/Gaussian_U8_U8_3x3.gdf", "filter/Median_U8_U8_3x3.gdf",
    "filter/Sobel_S16S16_U8_3x3_GXY.gdf", "filter/Sobel_S16_U8_3x3_GX.gdf", "filter/Sobel_S16_U8_3x3_GY.gdf",
    "geometric/Remap_U8_U8_Bilinear.gdf", "geometric/Remap_U8_U8_Bilinear_Constant.gdf",
    "geometric/ScaleImage_U8_U8_Bilinear.gdf", "geometric/WarpAffine_U8_U8_Bilinear.gdf",
    "geometric/WarpAffine_U8_U8_Bilinear_Constant.gdf", "geometric/WarpPerspective_U8_U8_Bilinear_Constant.gdf",
    "vision/Canny_3x3_L1NORM.gdf", "vision/Harris_3x3.gdf",
    "vision_profile/07_box.gdf", "vision_profile/08_canny.gdf", "vision_profile/18_gaussian.gdf",
    "vision_profile/19_harrisCorners.gdf", "vision_profile/20_halfScaleGaussian.gdf", "vision_profile/23_magnitude.gdf",
    "vision_profile/25_median.gdf", "vision_profile/31_phase.gdf", "vision_profile/32_gaussianPyramid.gdf",
    "vision_profile/34_scaleImage.gdf", "vision_profile/35_sobel.gdf", "vision_profile/39_warpAffine.gdf",
    "vision_profile/40_warpPerspective.gdf",
    # the further kernel families
    "geometric/Remap_U8_U8_Nearest.gdf", "geometric/Remap_U8_U8_Nearest_Constant.gdf", "geometric/ScaleImage_U8_U8_Area.gdf",
    "geometric/ScaleImage_U8_U8_Bilinear_Constant.gdf", "geometric/ScaleImage_U8_U8_Bilinear_Replicate.gdf",
    "geometric/ScaleImage_U8_U8_Nearest.gdf", "geometric/WarpAffine_U8_U8_Nearest.gdf", "geometric/WarpAffine_U8_U8_Nearest_Constant.gdf",
    "geometric/WarpPerspective_U8_U8_Nearest_Constant.gdf", "vision/Canny_3x3_L2NORM.gdf", "vision/Canny_5x5_L1NORM.gdf",
    "vision/Canny_5x5_L2NORM.gdf", "vision/Canny_7x7_L1NORM.gdf", "vision/Canny_7x7_L2NORM.gdf",
    "vision/FastCorners_XY_U8_NoSupression.gdf", "vision/FastCorners_XY_U8_Supression.gdf", "vision/Harris_5x5.gdf",
    "vision/Harris_7x7.gdf", "color/ColorConvert_RGB_IYUV.gdf", "color/ColorConvert_RGB_NV12.gdf", "vision_profile/11_colorConvert.gdf",
    "vision_profile/17_fastCorners.gdf", "vision_profile/29_opticalFlowLK.gdf", "vision_profile/33_remap.gdf",
    "color/ChannelExtract_U8_U24_Pos0.gdf", "color/ChannelExtract_U8_U24_Pos1.gdf", "color/ChannelExtract_U8_U24_Pos2.gdf",
    "color/ChannelExtract_U8_U16_Pos0.gdf", "color/ChannelExtract_U8_U16_Pos1.gdf", "vision_profile/10_channelExtract.gdf",
]


@pytest.mark.parametrize("rel", NODE_SET)
def test_reference_gdf_imports(rel):
    path = os.path.join(REF, rel)
    if not os.path.exists(path):
        pytest.skip("the reference tree is not present")
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    messages = []
    cb = vx.LOG_CB(lambda ctx_, ref, status, text: messages.append(text.decode()))
    L.vxRegisterLogCallback(ctx, cb, 0)
    try:
        g = L.vxCreateGraph(ctx)
        io = []
        status = vx.import_gdf(g, open(path).read(), callback=lambda ref, name, params: io.append((name, params)))
        assert status == 0, messages
        n = ctypes.c_uint32(0)
        assert L.vxQueryGraph(g, vx.VX_GRAPH_NUMNODES, ctypes.byref(n), 4) == 0
        assert n.value >= 1
    finally:
        cc = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(cc))
