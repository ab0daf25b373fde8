// sgns_internal.h -- argument blocks and launchers shared between the kernel
// translation units and the C-ABI layer (not part of the public interface).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgns {

// Write-back policies of the matrix-core instance (par_sgns_write_back).
enum WbPolicy { kWbPolicyAtomic = 0, kWbPolicyStoreOut = 1, kWbPolicyStoreIn = 2, kWbPolicyStore = 3, kWbPolicyHot = 4 };

// Device counters, one slot each (uint64), reset before every training call.
enum Counter { kCntTokens = 0, kCntSentences, kCntPairs, kCntBatches, kCntNegatives, kCntRepPairs, kCntSlots = 8 };

struct DetArgs {
    const int32_t *tokens;
    const uint8_t *b;
    const int64_t *sent_off;
    const float *sent_alpha;
    const int64_t *batch_off;
    const int64_t *neg_off;
    const int32_t *negs;
    const float *exp_table;
    float *w_in, *w_out;
    float *m_in, *m_out, *m_grad, *m_out_upd, *m_in_new;  // scratch
    int32_t ld, d, W, B, kmax;
    int32_t compute_loss;
    int64_t s_begin, s_end;
    double *loss;                 // [2]
    unsigned long long *counters; // [kCntSlots]
};

struct FastArgs {
    const int32_t *tokens;
    const int64_t *sent_off;
    const int64_t *sent_raw;
    const int32_t *order;         // persistent queue order (may be null)
    const int64_t *queue_raw;     // raw words dequeued before each queue slot (with order)
    const uint2 *alias;           // alias table: {keep probability (float bits), alias} per bucket
    void *w_in, *w_out;
    int32_t ld, d, nq, W, R, NEG, B, V, kmax;
    int32_t lds_per_wave;         // bytes
    int32_t lds_negs;             // byte offset of the batch's negatives in a wave's LDS (after the snapshot)
    int32_t tab_off, tab_log2;    // the token hash set of a long sentence: byte offset, log2 slots (0: none)
    int32_t persistent;
    int32_t grab;                 // persistent: queue slots (sentences) a wave takes per dequeue, >= 1
    int32_t rounding;             // bf16 write-back: 0 stochastic, 1 nearest
    int32_t rep_mfma;             // bf16 tables: repulsion products on the matrix cores
    int32_t write_back;           // matrix-core instance: WbPolicy
    uint32_t hot_bytes;           // kWbPolicyHot: rows at byte offsets >= this take plain stores
    int32_t compute_loss;
    int32_t iter, epoch;
    float a0;
    uint64_t T, seed;
    int64_t s_begin, s_end;
    unsigned long long *queue;
    double *loss;
    unsigned long long *counters;
    // optional exact schedule (window shrink per centre, learning rate of
    // centre 0 / the rest, negatives per batch) replacing the Philox draws
    const uint8_t *plan_b;
    const float *plan_alpha;
    const int64_t *plan_batch_off;
    const int64_t *plan_neg_off;
    const int32_t *plan_negs;
};

struct WindowArgs {
    const int32_t *centre, *ctx, *neg;
    int64_t n_windows;
    int32_t nctx, nneg, d, ld, V;
    const void *w_in, *w_out;
    float *g_in, *g_out;          // adjoint outputs (dense, [V][ld])
    double *loss;
};

hipError_t launch_det(const DetArgs &a, hipStream_t st);
hipError_t launch_fast(const FastArgs &a, int bf16, int np, int waves_per_block, int blocks, hipStream_t st);
// Instance groups of the training kernel, one translation unit each (fast_inst_*.hip).
hipError_t fast_vector_fp32(const FastArgs &a, int store_out, int np, int wpb, int blocks, hipStream_t st);
hipError_t fast_vector_bf16(const FastArgs &a, int store_out, int np, int wpb, int blocks, hipStream_t st);
hipError_t fast_mfma_atomic(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st);
hipError_t fast_mfma_store_out(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st);
hipError_t fast_mfma_store(const FastArgs &a, int store_in_only, int np, int wpb, int blocks, hipStream_t st);
hipError_t fast_mfma_hot(const FastArgs &a, int np, int wpb, int blocks, hipStream_t st);
hipError_t launch_reset(unsigned long long *counters, int n_counters, double *loss, unsigned long long *queue,
                        hipStream_t st);
hipError_t launch_sample(const uint2 *alias, int V, uint64_t seed, int64_t sent, int64_t n, int32_t *out,
                         hipStream_t st);
hipError_t launch_window(const WindowArgs &a, int bf16, int np, int adjoint, int flags, hipStream_t st);

// Smallest power of two >= r (r >= 1): the position->slot table is indexed
// by position & (pow2_ceil(R) - 1).
__host__ __device__ inline int pow2_ceil(int r) {
    int p = 1;
    while (p < r) p <<= 1;
    return p;
}

// Matrix-core repulsion (bf16 tables): a batch of at most kMfI*16 positions
// is processed in groups of kMfU*16 unique negatives.  Its LDS holds the batch
// snapshot (B rows of ld + kMfPad bf16 and one zero row that tile rows past
// the batch read; the pad spreads the four row groups of a transposed read
// over distinct banks), one 16 x 16 bf16 tile for transposing the gradient
// (rows of kMfTS) and two staging tiles of 16 rows x 32 dimensions of N
// (rows of kMfSS: the four row groups of a column read fall on distinct banks).
// Two tiles per group (32 negatives): a batch of short sentences draws at most
// 4 x negative of them at a sentence's edges, and every tile of the group holds
// registers whether it is used or not (four tiles: 250 VGPRs, two: 206).  A
// batch with more unique negatives takes one W_in write-back per group.
constexpr int kMfU = 2;
constexpr int kMfI = 2;
constexpr int kMfPad = 4;
constexpr int kMfTS = 20;
constexpr int kMfSS = 36;
__host__ __device__ inline int mf_snap_stride(int ld) { return ld + kMfPad; }
inline int mf_bytes(int ld, int B) {
    return (B + 1) * mf_snap_stride(ld) * 2 + 16 * kMfTS * 2 + 2 * 16 * kMfSS * 2;
}

// A wave's LDS holds, during a sentence's attraction, the ring (fp32 value +
// loaded value per pair and slot) and, once the ring has been flushed, the
// repulsion state over the same bytes: the batch-row snapshot at offset 0,
// then the batch's negatives (drawn, unique, multiplicities: 3 x kmax int32).
// Offset of the negatives = the snapshot's size.
inline int fast_lds_negs(int nq, int bf16, int B, int ld, int mfma) {
    const int e = bf16 ? 4 : 8;
    int snap = B * nq * (mfma ? e : 8);  // matrix path: the batch rows; vector path: their fp32 W_in deltas
    if (mfma && mf_bytes(ld, B) > snap) snap = mf_bytes(ld, B);
    return (snap + 15) & ~15;
}

inline int fast_lds_per_wave(int R, int nq, int bf16, int kmax, int B, int ld, int mfma) {
    const int e = bf16 ? 4 : 8;
    // + one zero pair per lane after the ring: a lane whose last pair lies
    // past the row reads and writes it in place of that pair
    const int ring = R * nq * (8 + e) + 64 * 8;
    const int rep = fast_lds_negs(nq, bf16, B, ld, mfma) + 3 * kmax * 4;
    return ((ring > rep ? ring : rep) + 15) & ~15;
}

}  // namespace sgns
