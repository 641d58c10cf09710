// traffic_probe.hip -- two kernels of known HBM traffic that turn profiler byte counters (FETCH_SIZE,
// WRITE_SIZE per dispatch) into bytes, run as a plain HIP program (no PyTorch in the process):
//   k_copy16    a 16-byte-per-lane plane copy: reads and writes `bytes` once each;
//   k_rowload   the access shape of the fused vision tile pass: one dword load per lane at a 3-byte
//               lane stride along an RGB row (consecutive lanes' dwords overlap), one byte stored per
//               lane -- reads every RGB byte of a row once per wave (3 B/px), writes 1 B/px.
// Usage: traffic_probe [launches]   (default 3 of each; inputs are 8192 x 4096 RGB = 96 MiB)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                             \
        }                                                                         \
    } while (0)

constexpr int W = 8192, H = 4096;

__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

// lane = pixel column; 64 pixels per wave; the last pixel of a row reads its dword one byte earlier so
// the load stays inside the row (the pixel's bytes are then bits 8..31)
__global__ void __launch_bounds__(256) k_rowload(const uint8_t *__restrict__ rgb, uint8_t *__restrict__ out)
{
    const int x = (int)(blockIdx.x * 256 + threadIdx.x), y = (int)blockIdx.y;
    if (x >= W) return;
    const bool last = x == W - 1;
    const uint8_t *p = rgb + (size_t)y * 3 * W + 3 * x - (last ? 1 : 0);
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    if (last) v >>= 8;
    out[(size_t)y * W + x] = (uint8_t)((v & 0xff) + ((v >> 8) & 0xff) + ((v >> 16) & 0xff));
}

int main(int argc, char **argv)
{
    const int launches = argc > 1 ? atoi(argv[1]) : 3;
    const size_t rgbBytes = (size_t)W * H * 3, outBytes = (size_t)W * H;
    uint8_t *rgb = nullptr, *out = nullptr, *copy = nullptr;
    CHECK(hipMalloc(&rgb, rgbBytes));
    CHECK(hipMalloc(&out, outBytes));
    CHECK(hipMalloc(&copy, rgbBytes));
    CHECK(hipMemset(rgb, 0x5a, rgbBytes));
    CHECK(hipDeviceSynchronize());
    const size_t n16 = rgbBytes / 16;
    for (int i = 0; i < launches; i++) {
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, 0, (const uint4 *)rgb, (uint4 *)copy, n16);
        CHECK(hipGetLastError());
    }
    for (int i = 0; i < launches; i++) {
        hipLaunchKernelGGL(k_rowload, dim3(W / 256, H), dim3(256), 0, 0, rgb, out);
        CHECK(hipGetLastError());
    }
    CHECK(hipDeviceSynchronize());
    printf("traffic_probe: k_copy16 reads and writes %zu bytes; k_rowload reads %zu bytes, writes %zu bytes; %d launches each\n",
           rgbBytes, rgbBytes, outBytes, launches);
    CHECK(hipFree(rgb));
    CHECK(hipFree(out));
    CHECK(hipFree(copy));
    return 0;
}
