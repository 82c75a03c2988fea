// device_rates.hip -- measured peak rates of the device, the ceilings the scatter and shading
// kernels are compared against:
//   - global atomic adds (f32, u32, u64) at hashed addresses over 64M dwords (spread), 1M dwords
//     (L2-resident) and 4,096 dwords (contended): 8192 x 256 threads x 16 adds;
//   - VALU issue: fp32 fma (8 independent chains per thread, no packed math) and v_exp_f32
//     (transcendental), as wave instructions per second.
// Each case is launched twice and the second launch is timed.
//
// usage: device_rates
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                               \
        }                                                                           \
    } while (0)

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// KIND 0: f32, 1: u32, 2: u64 (index halved, so every access stays inside the buffer)
template <int KIND>
__global__ void __launch_bounds__(256) k_atomic_rate(void* buf, uint32_t mask, int iters) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    for (int i = 0; i < iters; i++) {
        const uint32_t a = mix32(t * 131u + (uint32_t)i) & mask;
        if (KIND == 0) atomicAdd((float*)buf + a, 1.0f);
        else if (KIND == 1) atomicAdd((unsigned int*)buf + a, 1u);
        else atomicAdd((unsigned long long*)buf + (a >> 1), 1ull);
    }
}

// 8 independent chains per thread; OP 0: fma, 1: exp2 (v_exp_f32).  The result is stored so the
// work stays live.
template <int OP>
__global__ void __launch_bounds__(256) k_valu_rate(float* out, int iters, float x, float y) {
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = (float)(threadIdx.x + j) * 1e-3f;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (OP == 0) a[j] = __builtin_fmaf(a[j], x, y);
            else a[j] = __builtin_amdgcn_exp2f(a[j] * -1e-3f);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; j++) s += a[j];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

int main() {
    const size_t bytes = (size_t)256 << 20;   // 64M dwords
    void* buf = nullptr;
    CHECK(hipMalloc(&buf, bytes));
    CHECK(hipMemset(buf, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int blocks = 8192, iters = 16;
    const double n = (double)blocks * 256 * iters;
    const char* kinds[3] = {"f32", "u32", "u64"};
    const uint32_t masks[3] = {(64u << 20) - 1, (1u << 20) - 1, 4095u};
    const char* ranges[3] = {"64M dwords (spread)", "1M dwords (L2-resident)", "4096 dwords (contended)"};
    for (int m = 0; m < 3; m++)
        for (int k = 0; k < 3; k++) {
            float ms = 0.f;
            for (int rep = 0; rep < 2; rep++) {
                CHECK(hipEventRecord(e0));
                if (k == 0) k_atomic_rate<0><<<blocks, 256>>>(buf, masks[m], iters);
                else if (k == 1) k_atomic_rate<1><<<blocks, 256>>>(buf, masks[m], iters);
                else k_atomic_rate<2><<<blocks, 256>>>(buf, masks[m], iters);
                CHECK(hipGetLastError());
                CHECK(hipEventRecord(e1));
                CHECK(hipEventSynchronize(e1));
                CHECK(hipEventElapsedTime(&ms, e0, e1));
            }
            std::printf("atomic %s add, %-26s %8.3f ms  %7.2f G atomics/s\n", kinds[k], ranges[m], ms, n / (ms * 1e-3) / 1e9);
        }
    // VALU: 16384 workgroups (64 waves per CU), 8 chains x 2048 iterations per thread
    const int vblocks = 16384, viters = 2048;
    const double winst = (double)vblocks * 4 * 8 * viters;   // wave instructions of the chains
    for (int op = 0; op < 2; op++) {
        float ms = 0.f;
        for (int rep = 0; rep < 2; rep++) {
            CHECK(hipEventRecord(e0));
            if (op == 0) k_valu_rate<0><<<vblocks, 256>>>((float*)buf, viters, 0.999f, 1e-4f);
            else k_valu_rate<1><<<vblocks, 256>>>((float*)buf, viters, 0.999f, 1e-4f);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        std::printf("valu %s: %8.3f ms  %8.2f T wave-instructions/s\n", op == 0 ? "fp32 fma" : "v_exp_f32", ms, winst / (ms * 1e-3) / 1e12);
    }
    CHECK(hipFree(buf));
    return 0;
}
