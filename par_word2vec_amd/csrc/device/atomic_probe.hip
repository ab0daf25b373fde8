// atomic_probe.hip -- request-size probe of the packed-bf16 atomic path.
//
// Measures what the write-backs of the trainer are bound by: the L2's
// atomic unit.  A [555001][320] bf16 table (the 1B-shaped workload's W_out:
// 555,001 rows, d = 300 padded to 320) takes global_atomic_pk_add_bf16 from
// 2,048 workgroups of 4 waves; in each instruction the 64 lanes cover R rows
// x (256 / R) contiguous bytes at random rows (R = 16, 8, 4, 2, 1), over a
// 4,096-row hot set (L2-resident) and over the whole table.  Plain 32-bit
// stores of the same shapes are timed for comparison.  The rates printed are
// instructions and lane operations per second of the second of two launches
// (HIP events); with `rocprofv3 --pmc TCC_ATOMIC_sum` the same run shows the
// L2's atomic requests per launch (one per 64-byte row segment).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

constexpr int kRowElems = 320;
constexpr int kVocab = 555001;
constexpr int kBlocks = 2048;
constexpr int kThreads = 256;
constexpr int kIters = 2000;

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

__device__ __forceinline__ uint32_t mix32(uint32_t x) {  // integer hash (lowbias32)
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// R rows per instruction; STORE: plain 32-bit stores instead of atomics.
// Every address is inside the table: row < hot <= kVocab, and the 2 * (64 / R)
// elements a row group covers start at a multiple of that width below 320.
template <int R, bool STORE>
__global__ void probe(uint16_t *w, int hot, int iters) {
    constexpr int kLanesPerRow = 64 / R;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int group = lane / kLanesPerRow, sub = lane % kLanesPerRow;
    for (int it = 0; it < iters; it++) {
        const uint32_t h = mix32(wave * 7919u + (uint32_t)it * 104729u + (uint32_t)group * 31u);
        const uint32_t row = h % (uint32_t)hot;
        const uint32_t seg = (h >> 20) % (uint32_t)(kRowElems / (2 * kLanesPerRow));
        const uint32_t elem = seg * 2 * kLanesPerRow + 2 * sub;
        uint32_t *p = reinterpret_cast<uint32_t *>(w + (size_t)row * kRowElems + elem);
        if (STORE) {
            *p = 0x00010001u;
        } else {
            (void)__builtin_amdgcn_global_atomic_fadd_v2bf16(reinterpret_cast<bf16x2_t *>(p),
                                                             __builtin_bit_cast(bf16x2_t, 0x00010001u));
        }
    }
}

template <int R, bool STORE>
void run(uint16_t *w, int hot) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    probe<R, STORE><<<kBlocks, kThreads>>>(w, hot, kIters);  // warm-up
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(a));
    probe<R, STORE><<<kBlocks, kThreads>>>(w, hot, kIters);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, a, b));
    const double insts = (double)kBlocks * (kThreads / 64) * kIters;
    std::printf("  %-6s %2d rows x %3d B: %7.2f ms  %6.2f G inst/s  %7.1f G lane-ops/s\n", STORE ? "store" : "atomic",
                R, 256 / R, ms, insts / ms / 1e6, insts * 64 / ms / 1e6);
    std::fflush(stdout);
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
}

}  // namespace

int main() {
    uint16_t *w = nullptr;
    const size_t bytes = (size_t)kVocab * kRowElems * sizeof(uint16_t);
    CHECK(hipMalloc(&w, bytes));
    CHECK(hipMemset(w, 0, bytes));
    for (int hot : {4096, kVocab}) {
        std::printf("rows %d\n", hot);
        run<16, false>(w, hot);
        run<8, false>(w, hot);
        run<4, false>(w, hot);
        run<2, false>(w, hot);
        run<1, false>(w, hot);
        run<8, true>(w, hot);
        run<2, true>(w, hot);
    }
    CHECK(hipFree(w));
    return 0;
}
