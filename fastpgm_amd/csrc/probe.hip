// Throughput probes for the roofline of the counting kernels (gfx950), used by bench.py.
//
// The counting sinks are not bound by HBM: the per-lane sinks issue one LDS atomic add per
// row and test into lane-private counters, the device-memory sink one device-memory
// atomic add per row.  These kernels measure what the device sustains for exactly those
// access patterns, so that each sink's rate can be stated as a fraction of a measured peak:
//   k_probe_lds_atomics: every lane adds into its own 32-bit words of a [cell][lane] LDS
//     array (64 distinct words per wave instruction, 32 KiB per workgroup as in the lane
//     sink), 16 adds per step to addresses formed before the loop: no other work;
//   k_probe_dev_atomics: every lane adds 1 to scattered words of a device-memory table
//     (a multiplicative hash of its index), as the device-memory sink does.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kWG = 256;
constexpr int kCells = 32;  // [cell][lane] words per wave pair: 32 * 64 * 4 waves * 4 B = 32 KiB

__global__ __launch_bounds__(kWG) void k_probe_lds_atomics(int steps, uint32_t *sink)
{
    __shared__ uint32_t h[kCells * kWG];
    const int tid = threadIdx.x;
    for (int i = tid; i < kCells * kWG; i += kWG) h[i] = 0;
    __syncthreads();
    // 16 distinct cells per lane, addresses formed once: the loop is LDS atomics only
    uint32_t *p[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) p[j] = &h[(((uint32_t)tid * 7u + (uint32_t)j * 5u) & (kCells - 1)) * kWG + tid];
    for (int s = 0; s < steps; ++s) {
#pragma unroll
        for (int j = 0; j < 16; ++j) atomicAdd(p[j], 1u);
    }
    __syncthreads();
    // the sums stay observable
    uint32_t t = 0;
    for (int c = 0; c < kCells; ++c) t += h[c * kWG + tid];
    if (t == 0xFFFFFFFFu) sink[blockIdx.x] = t;
}

__global__ __launch_bounds__(kWG) void k_probe_dev_atomics(int steps, uint32_t *table, uint32_t mask)
{
    uint32_t key = ((uint32_t)blockIdx.x * kWG + threadIdx.x) * 2654435761u;
    for (int s = 0; s < steps; ++s) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            key = key * 1664525u + 1013904223u;
            atomicAdd(&table[key & mask], 1u);
        }
    }
}

double time_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

}  // namespace

extern "C" {

// LDS atomic adds per second over the whole device (lane-private words, no conflicts), and
// the same per CU per clock (clock from the device properties).  Returns 0 on success.
int fpgm_probe_lds_atomics(double *adds_per_s, double *per_cu_clock)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
    const int grid = prop.multiProcessorCount * 16, steps = 4096;
    uint32_t *sink = nullptr;
    hipEvent_t e0, e1;
    if (hipMalloc((void **)&sink, sizeof(uint32_t) * grid) != hipSuccess) return 1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_probe_lds_atomics, dim3(grid), dim3(kWG), 0, 0, 16, sink);  // warm-up
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(k_probe_lds_atomics, dim3(grid), dim3(kWG), 0, 0, steps, sink);
    (void)hipEventRecord(e1, 0);
    const hipError_t e = hipEventSynchronize(e1);
    const double ms = time_ms(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    if (e != hipSuccess || ms <= 0.0) return 1;
    const double adds = (double)grid * kWG * steps * 16;
    *adds_per_s = adds / (ms * 1e-3);
    *per_cu_clock = *adds_per_s / ((double)prop.multiProcessorCount * prop.clockRate * 1e3);
    return 0;
}

// Scattered device-memory atomic adds per second into a table of table_words words.
int fpgm_probe_dev_atomics(int64_t table_words, double *adds_per_s)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
    uint32_t mask = 1;
    while ((int64_t)mask * 2 <= table_words) mask *= 2;
    uint32_t *table = nullptr;
    if (hipMalloc((void **)&table, sizeof(uint32_t) * mask) != hipSuccess) return 1;
    (void)hipMemset(table, 0, sizeof(uint32_t) * mask);
    const int grid = prop.multiProcessorCount * 8, steps = 256;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_probe_dev_atomics, dim3(grid), dim3(kWG), 0, 0, 4, table, mask - 1);
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(k_probe_dev_atomics, dim3(grid), dim3(kWG), 0, 0, steps, table, mask - 1);
    (void)hipEventRecord(e1, 0);
    const hipError_t e = hipEventSynchronize(e1);
    const double ms = time_ms(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(table);
    if (e != hipSuccess || ms <= 0.0) return 1;
    *adds_per_s = (double)grid * kWG * steps * 8 / (ms * 1e-3);
    return 0;
}

}  // extern "C"
