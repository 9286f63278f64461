// LDS atomic throughput on the device: clocks per 64-lane wave instruction and CU for
// 64-bit and 32-bit integer adds and 32-bit float adds under different address patterns
// (random cells, one bank or slot per lane, a few hot cells).  The histogram kernels'
// layouts are chosen from these rates.  Every index is masked to the 8,192-entry LDS
// array; one output word per thread keeps the work live.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
typedef unsigned long long u64;
constexpr int kCells = 8192;
constexpr int kIters = 2048;

template <int MODE>
__global__ __launch_bounds__(512) void k_probe(u64 *out, uint32_t seed)
{
    __shared__ u64 lds64[kCells];
    uint32_t *lds32 = reinterpret_cast<uint32_t *>(lds64);
    float *ldsf = reinterpret_cast<float *>(lds64);
    for (int i = threadIdx.x; i < kCells; i += blockDim.x) lds64[i] = 0;
    __syncthreads();
    uint32_t x = seed ^ (blockIdx.x * 7919u + threadIdx.x * 104729u);
    uint32_t const lane = threadIdx.x & 63u;
    for (int it = 0; it < kIters; ++it) {
        x = x * 1664525u + 1013904223u;
        uint32_t const r = x >> 19; // 13 random bits
        if (MODE == 0) atomicAdd(&lds64[r & 8191u], (u64)1);                      // u64 random
        if (MODE == 1) atomicAdd(&lds64[((r & 511u) << 4) | (lane & 15u)], (u64)1); // u64 bank-pair = lane%16
        if (MODE == 2) atomicAdd(&lds32[r & 16383u], 1u);                          // u32 random
        if (MODE == 3) atomicAdd(&lds32[((r & 511u) << 5) | (lane & 31u)], 1u);   // u32 bank = lane%32
        if (MODE == 4) atomicAdd(&lds64[((r & 255u) << 5) | (lane & 31u)], (u64)1); // u64, lane%32 slot
        if (MODE == 5) atomicAdd(&lds64[lane], (u64)1);                            // u64 distinct fixed
        if (MODE == 6) atomicAdd(&lds64[((r & 127u) << 6) | lane], (u64)1);        // u64, lane slot of 64
        if (MODE == 7) { uint32_t v = lds32[r & 16383u]; x ^= v; }                 // u32 random read
        if (MODE == 8) atomicAdd(&ldsf[r & 16383u], 1.0F);                         // f32 random
        if (MODE == 9) atomicAdd(&ldsf[((r & 511u) << 5) | (lane & 31u)], 1.0F);  // f32 bank = lane%32
        if (MODE == 10) atomicAdd(&ldsf[(r & 15u) << 5], 1.0F);                    // f32, 16 addresses, one bank
        if (MODE == 11) atomicAdd(&lds64[(((r & 255u) << 5) | (lane & 31u)) & 8191u], (u64)1); // u64 32-slot
        if (MODE == 12) atomicAdd(&lds64[((r & 17u) << 5) | (lane & 31u)], (u64)1); // u64 few bins, lane slots
    }
    __syncthreads();
    out[blockIdx.x * blockDim.x + threadIdx.x] = lds64[threadIdx.x & (kCells - 1)] + x;
}

template <int MODE> float run(u64 *d, int blocks)
{
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipLaunchKernelGGL(k_probe<MODE>, dim3(blocks), dim3(512), 0, 0, d, 1u);
    hipEventRecord(a);
    for (int k = 0; k < 5; ++k) hipLaunchKernelGGL(k_probe<MODE>, dim3(blocks), dim3(512), 0, 0, d, 2u + k);
    hipEventRecord(b);
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    return ms / 5;
}

int main()
{
    int blocks = 256 * 8;
    u64 *d = nullptr;
    if (hipMalloc(&d, sizeof(u64) * blocks * 512) != hipSuccess) return 1;
    hipDeviceProp_t p;
    hipGetDeviceProperties(&p, 0);
    double const clk = p.clockRate * 1e3; // Hz
    int const cus = p.multiProcessorCount;
    const char *names[] = {"u64 random", "u64 lane%16 pair", "u32 random", "u32 lane%32 bank", "u64 lane%32 slot",
                           "u64 fixed per lane", "u64 lane%64 slot", "u32 random read", "f32 random",
                           "f32 lane%32 bank", "f32 16 addr 1 bank", "u64 32-slot again", "u64 18 bins x 32 slots"};
    float ms[13];
    ms[0] = run<0>(d, blocks); ms[1] = run<1>(d, blocks); ms[2] = run<2>(d, blocks); ms[3] = run<3>(d, blocks);
    ms[4] = run<4>(d, blocks); ms[5] = run<5>(d, blocks); ms[6] = run<6>(d, blocks); ms[7] = run<7>(d, blocks);
    ms[8] = run<8>(d, blocks); ms[9] = run<9>(d, blocks); ms[10] = run<10>(d, blocks); ms[11] = run<11>(d, blocks);
    ms[12] = run<12>(d, blocks);
    double const instr = (double)blocks * 8 * kIters; // wave instructions
    printf("cus %d clock %.0f MHz\n", cus, clk / 1e6);
    for (int m = 0; m < 13; ++m) {
        double const per_cu_clk = (ms[m] * 1e-3) * clk * cus; // CU-clocks
        printf("%-22s %8.3f ms  %6.2f clk per wave-instruction per CU\n", names[m], ms[m], per_cu_clk / instr);
    }
    hipFree(d);
    return 0;
}
