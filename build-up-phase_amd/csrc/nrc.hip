// nrc.hip — neural radiance cache (Muller, Rousselle, Novak & Keller 2021) on gfx950 MFMA.
//
// Network 96 -> 64 -> 64 -> 64 -> 64 -> 64 -> 16, ReLU on hidden layers, BF16 operands, FP32 accumulation
// (v_mfma_f32_16x16x32_bf16), FP32 master weights, Adam.
//
// Layout trick: a layer is computed transposed, H^T (out x 16 samples) = W (out x in) * A^T (in x 16 samples),
// with the weights as the MFMA A operand and 16 samples as the N dimension.  The accumulator of output tile t
// holds, in lane l, rows 16t + 4(l/16) + i (i < 4) of sample l % 16.  For the next layer's B operand, lane l
// must hold 8 k-values of sample l % 16 at k-group l / 16; we take the 4 rows of tile 2s and the 4 rows of
// tile 2s + 1 that the lane already holds, i.e. the k order of step s is
//     kperm(s, g, j) = 32s + 4g + j (j < 4),   32s + 16 + 4g + (j - 4) (j >= 4)
// and pre-permute the next layer's weight columns the same way (fragment packing below).  Activations never
// leave registers between layers.
//
// Training: one 256-thread block = 4 waves x 16 samples.  Forward activations are also written to LDS as
// A^T[unit][64 samples] (bf16); the backward data gradient uses W^T as the A operand (same trick); the weight
// gradient dW = G^T (out x 64) * A (64 x in) is an MFMA GEMM over the 64 samples of the block whose A operand
// (G^T rows) and B operand (A^T rows) are 16-byte LDS reads.  Each block writes its partial gradient and loss; a
// second kernel sums the partials in block order (deterministic), applies Adam and writes the updated weights'
// BF16 fragments.
#include <cmath>

#include "nrc.h"
#include "rtx_internal.h"

namespace rtx {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int in_dim(int l) { return l == 0 ? 96 : 64; }
__host__ __device__ constexpr int out_dim(int l) { return l == 5 ? 16 : 64; }
__host__ __device__ constexpr int w_off(int l) { return l == 0 ? 0 : 6208 + (l - 1) * 4160; }
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + out_dim(l) * in_dim(l); }
// forward fragments (tiles over out, steps over in) and backward fragments (tiles over in, steps over out)
__host__ __device__ constexpr int fwd_base(int l) { return l == 0 ? 0 : 12 + (l - 1) * 8; }
__host__ __device__ constexpr int bwd_base(int l) { return (l - 1) * 8; }  // l = 1..5
constexpr int kFwdFrags = 46;
constexpr int kBwdFrags = 36;
static_assert(b_off(5) + 16 == (int)RTX_NRC_PARAM_COUNT, "parameter layout");

__device__ __forceinline__ int kperm(int s, int g, int j) { return 32 * s + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4)); }

// LDS rows of the training kernel: [row][64 samples] BF16, the 8 16-byte chunks of a row permuted by the row's
// low 3 bits (chunk c stored at c ^ (row & 7)), so that the weight-gradient operand reads (16 rows x one chunk)
// spread over the banks instead of all landing in the same 16 bytes of every 128-byte row.
__device__ __forceinline__ int lds_idx(int row, int col) { return row * 64 + (col ^ ((row & 7) << 3)); }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct Norm {
    float lo[3];
    float inv[3];
};

// Which samples one training step takes.  Host-sized steps: samples [0, n).  Device-sized steps (count_dev set):
// batch b of nb = ceil(count / batch_size) interleaved batches, samples b, b + nb, b + 2 nb, ... (n of them); a
// step with b >= nb does nothing.
struct StepShape {
    const uint32_t* count_dev;
    uint32_t n;
    uint32_t batch;
    uint32_t batch_size;
};

__device__ __forceinline__ bool step_samples(const StepShape& sh, uint32_t& n, uint32_t& stride, uint32_t& offset) {
    if (!sh.count_dev) {
        n = sh.n;
        stride = 1;
        offset = 0;
        return n > 0;
    }
    const uint32_t count = *sh.count_dev;
    const uint32_t nb = (count + sh.batch_size - 1) / sh.batch_size;
    if (sh.batch >= nb) return false;
    n = (count - sh.batch + nb - 1) / nb;
    stride = nb;
    offset = sh.batch;
    return true;
}

__device__ __forceinline__ void load_query(const rtx_nrc_query* src, rtx_nrc_query& q) {
    const float4* p = (const float4*)src;
    const float4 a = p[0], b = p[1], c = p[2];
    q.position[0] = a.x; q.position[1] = a.y; q.position[2] = a.z;
    q.normal[0] = a.w; q.normal[1] = b.x; q.normal[2] = b.y;
    q.direction[0] = b.z; q.direction[1] = b.w; q.direction[2] = c.x;
    q.albedo[0] = c.y; q.albedo[1] = c.z; q.albedo[2] = c.w;
}

// B fragments of the input layer for the lane's sample: k = 32s + 8g + j, i.e. the lane fills the 8-feature chunk
// c = 4s + g at step s.  Chunks 0..8 are frequency chunks (axis c / 3, frequencies 4(c % 3) .. 4(c % 3) + 3, sin
// and cos interleaved), so each lane evaluates one sin/cos pair and three angle doublings per step; chunk 9
// holds normal, direction and albedo[0..1], chunk 10 albedo[2], chunk 11 zeros.
__device__ __forceinline__ void encode(const rtx_nrc_query& q, const Norm& nm, int g, bool valid, bf16x8* in0,
                                       __bf16* at0, int col) {
    float pn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) pn[a] = (q.position[a] - nm.lo[a]) * nm.inv[a];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int c = 4 * s + g;
        float v[8];
        if (c < 9) {
            const int a = c / 3, k0 = 4 * (c - 3 * a);
            const float p = a == 0 ? pn[0] : (a == 1 ? pn[1] : pn[2]);
            // sin / cos of 2^k0 pi p, then three angle doublings: sin 2x = 2 sin x cos x, cos 2x = (cos x - sin x)
            // (cos x + sin x); each doubling at most doubles the absolute error (a few FP32 ulps after three, far
            // below the BF16 rounding of the features)
            float sn, cs;
            sincospif(ldexpf(p, k0), &sn, &cs);
            v[0] = sn;
            v[1] = cs;
#pragma unroll
            for (int i = 1; i < 4; ++i) {
                const float s2 = 2.0f * sn * cs;
                const float c2 = (cs - sn) * (cs + sn);
                sn = s2;
                cs = c2;
                v[2 * i] = sn;
                v[2 * i + 1] = cs;
            }
        } else {
            const float c9[8] = {q.normal[0],    q.normal[1],    q.normal[2], q.direction[0],
                                 q.direction[1], q.direction[2], q.albedo[0], q.albedo[1]};
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = c == 9 ? c9[j] : (c == 10 && j == 0 ? q.albedo[2] : 0.0f);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            in0[s][j] = (__bf16)(valid ? v[j] : 0.0f);
            if (at0) at0[lds_idx(8 * c + j, col)] = in0[s][j];
        }
    }
}

// ReLU and repack 4 output tiles into the 2 B fragments of the next layer (k order kperm).  The values are
// rounded to BF16 first and the ReLU is taken on the BF16 bit patterns as 16-bit integers (packed integer max with
// 0): negative values, -0 included, have the sign bit set and become +0, the others are unchanged, which is
// bit for bit the rounding of max(0, x) for every x but NaN.  A NaN with the sign bit clear stays NaN (fmaxf(0, x)
// would give 0), so a network whose activations diverged shows NaN in its outputs and loss instead of hiding it.
__device__ __forceinline__ void relu_pack(const f32x4* acc, bf16x8* out) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        u32x4 w;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const f32x2 f = {acc[2 * s + (q >> 1)][2 * (q & 1)], acc[2 * s + (q >> 1)][2 * (q & 1) + 1]};
            const i16x2 h = __builtin_bit_cast(i16x2, __builtin_convertvector(f, bf16x2));
            w[q] = __builtin_bit_cast(unsigned, __builtin_elementwise_max(h, i16x2{0, 0}));
        }
        out[s] = __builtin_bit_cast(bf16x8, w);
    }
}

// Store the activations held in B fragments (kperm order) into LDS A^T[unit][64 samples].
__device__ __forceinline__ void store_act(const bf16x8* in, __bf16* at, int g, int col) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) at[lds_idx(kperm(s, g, j), col)] = in[s][j];
}

// Inference, LDS-staged: a persistent block copies the 46 forward fragments (46 KB) into LDS once, then each wave
// runs two 16-query tiles per pass so that every fragment read from LDS feeds two MFMAs (weights are the A operand,
// queries the N dimension).  Each tile sees exactly the operations of the training kernel's forward pass.
constexpr int kInferLdsFrags = kFwdFrags * 64;  // bf16x8 entries

template <int L>
__device__ __forceinline__ void layer_forward2(const bf16x8* __restrict__ frag, const float* __restrict__ params,
                                               const bf16x8* inA, const bf16x8* inB, f32x4* accA, f32x4* accB,
                                               int lane) {
    constexpr int T = out_dim(L) / 16, S = in_dim(L) / 32;
    const int g = lane >> 4;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        // the bias is the accumulator's initial value (no separate adds after the products)
        f32x4 cA, cB;
#pragma unroll
        for (int i = 0; i < 4; ++i) cA[i] = cB[i] = params[b_off(L) + 16 * t + 4 * g + i];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bf16x8 a = frag[(fwd_base(L) + t * S + s) * 64 + lane];
            cA = mfma(a, inA[s], cA);
            cB = mfma(a, inB[s], cB);
        }
        accA[t] = cA;
        accB[t] = cB;
    }
}

__global__ __launch_bounds__(256) void nrc_infer_lds_kernel(const bf16x8* __restrict__ fwd,
                                                            const float* __restrict__ params,
                                                            const rtx_nrc_query* __restrict__ queries, uint32_t n,
                                                            const uint32_t* __restrict__ n_dev, Norm nm,
                                                            float* __restrict__ out_rgb, InferGather gt) {
    __shared__ bf16x8 frag[kInferLdsFrags];
    for (int i = threadIdx.x; i < kInferLdsFrags; i += blockDim.x) frag[i] = fwd[i];
    __syncthreads();
    if (n_dev) n = min(n, *n_dev);
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, g = lane >> 4;
    const uint64_t pairs = ((uint64_t)n + 31) / 32;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t p = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < pairs; p += waves) {
        // an offset the compiler cannot see through, so the fragment reads stay in the loop (hoisted, the 46
        // fragments would take 184 VGPRs and the kernel would run at one wave per SIMD)
        int lane_f = lane;
        asm volatile("" : "+v"(lane_f));
        const uint64_t iA = p * 32 + c, iB = iA + 16;
        const bool vA = iA < n, vB = iB < n;
        rtx_nrc_query qA{}, qB{};
        if (vA) load_query(queries + iA, qA);
        if (vB) load_query(queries + iB, qB);
        bf16x8 inA[3], inB[3];
        encode(qA, nm, g, vA, inA, nullptr, 0);
        encode(qB, nm, g, vB, inB, nullptr, 0);
        bf16x8 aA[2], aB[2];
        f32x4 accA[4], accB[4];
        layer_forward2<0>(frag, params, inA, inB, accA, accB, lane_f);
        relu_pack(accA, aA);
        relu_pack(accB, aB);
        layer_forward2<1>(frag, params, aA, aB, accA, accB, lane_f);
        relu_pack(accA, aA);
        relu_pack(accB, aB);
        layer_forward2<2>(frag, params, aA, aB, accA, accB, lane_f);
        relu_pack(accA, aA);
        relu_pack(accB, aB);
        layer_forward2<3>(frag, params, aA, aB, accA, accB, lane_f);
        relu_pack(accA, aA);
        relu_pack(accB, aB);
        layer_forward2<4>(frag, params, aA, aB, accA, accB, lane_f);
        relu_pack(accA, aA);
        relu_pack(accB, aB);
        layer_forward2<5>(frag, params, aA, aB, accA, accB, lane_f);
        if (g == 0) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint64_t i = h ? iB : iA;
                if (!(h ? vB : vA)) continue;
                const f32x4& acc = h ? accB[0] : accA[0];
                const float y0 = fmaxf(0.0f, acc[0]), y1 = fmaxf(0.0f, acc[1]), y2 = fmaxf(0.0f, acc[2]);
                if (gt.rad) {
                    const uint32_t pid = gt.pid[i];
                    const float4 t = gt.thr[pid];
                    float4 r = gt.rad[pid];
                    r.x += t.x * y0;
                    r.y += t.y * y1;
                    r.z += t.z * y2;
                    gt.rad[pid] = r;
                } else {
                    out_rgb[3 * i] = y0;
                    out_rgb[3 * i + 1] = y1;
                    out_rgb[3 * i + 2] = y2;
                }
            }
        }
    }
}

// LDS of the training kernel (64 samples): A^T of the layer inputs and G^T of the current layer.
struct TrainLds {
    __bf16 at0[96 * 64];
    __bf16 at[5][64 * 64];
    __bf16 gt[64 * 64];
    float red[8];
};

// Weight gradient of layer L for the block's 64 samples: dW tile (t, u) = G^T rows * A columns.
template <int L>
__device__ __forceinline__ void weight_grad(TrainLds& sm, const __bf16* at, float* __restrict__ part, int wave,
                                            int lane) {
    constexpr int T = out_dim(L) / 16, U = in_dim(L) / 16, N = T * U;
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int k = wave; k < N; k += 4) {
        const int t = k / U, u = k % U;
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8 a = *(const bf16x8*)(sm.gt + lds_idx(16 * t + r, 32 * s + 8 * g));
            const bf16x8 b = *(const bf16x8*)(at + lds_idx(16 * u + r, 32 * s + 8 * g));
            c = mfma(a, b, c);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) part[w_off(L) + (16 * t + 4 * g + i) * in_dim(L) + 16 * u + r] = c[i];
    }
    // bias gradient: sum over the 64 samples of G^T rows; 4 lanes per row (16 samples each), then a shuffle sum
    const int row = threadIdx.x >> 2, q = threadIdx.x & 3;
    if (row < out_dim(L)) {
        float sum = 0.0f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bf16x8 v = *(const bf16x8*)(sm.gt + lds_idx(row, 16 * q + 8 * h));
#pragma unroll
            for (int j = 0; j < 8; ++j) sum += (float)v[j];
        }
        sum += __shfl_xor(sum, 1);
        sum += __shfl_xor(sum, 2);
        if (q == 0) part[b_off(L) + row] = sum;
    }
}

// The training kernel keeps a layer's fragments (and forward biases) in registers, loaded one layer ahead so that
// their L2 latency overlaps the previous layer's work: acc[t] = W_L (tile t) * in + bias.
template <int L>
__device__ __forceinline__ void load_fwd_layer(const bf16x8* __restrict__ fwd, const float* __restrict__ params,
                                               int lane, bf16x8* f, float* b) {
    constexpr int T = out_dim(L) / 16, S = in_dim(L) / 32;
    const int g = lane >> 4;
#pragma unroll
    for (int i = 0; i < T * S; ++i) f[i] = fwd[(fwd_base(L) + i) * 64 + lane];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) b[4 * t + i] = params[b_off(L) + 16 * t + 4 * g + i];
}

template <int L>
__device__ __forceinline__ void layer_forward_regs(const bf16x8* f, const float* b, const bf16x8* in, f32x4* acc) {
    constexpr int T = out_dim(L) / 16, S = in_dim(L) / 32;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        // the bias is the accumulator's initial value, as in inference
        f32x4 c;
#pragma unroll
        for (int i = 0; i < 4; ++i) c[i] = b[4 * t + i];
#pragma unroll
        for (int s = 0; s < S; ++s) c = mfma(f[t * S + s], in[s], c);
        acc[t] = c;
    }
}

template <int L>
__device__ __forceinline__ void load_bwd_layer(const bf16x8* __restrict__ bwd, int lane, bf16x8* f) {
    constexpr int S = (L == 5) ? 1 : 2;
#pragma unroll
    for (int i = 0; i < 4 * S; ++i) f[i] = bwd[(bwd_base(L) + i) * 64 + lane];
}

// Backward data gradient through layer L (1..5): G_{L-1} = (W_L^T G_L) masked by the ReLU of layer L-1; f holds
// the layer's backward fragments.
template <int L>
__device__ __forceinline__ void back_data(const bf16x8* f, const bf16x8* gin, const bf16x8* act, f32x4* gout) {
    constexpr int S = (L == 5) ? 1 : 2;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < S; ++s) c = mfma(f[t * S + s], gin[s], c);
        gout[t] = c;
    }
    // ReLU mask: element (t, i) sits at fragment s = t / 2, j = (t % 2) * 4 + i of the layer input
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (!((float)act[t >> 1][(t & 1) * 4 + i] > 0.0f)) gout[t][i] = 0.0f;
}

__device__ __forceinline__ void store_grad(const f32x4* gv, int tiles, __bf16* gt, int g, int col, bf16x8* frag) {
    // C-layout gradient -> LDS G^T and -> B fragments (kperm order) for the next backward step
    for (int t = 0; t < tiles; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) gt[lds_idx(16 * t + 4 * g + i, col)] = (__bf16)gv[t][i];
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int t = 2 * s + (j >> 2);
            frag[s][j] = t < tiles ? (__bf16)gv[t][j & 3] : (__bf16)0.0f;
        }
}

__global__ __launch_bounds__(256) void nrc_train_kernel(const bf16x8* __restrict__ fwd, const bf16x8* __restrict__ bwd,
                                                        const float* __restrict__ params,
                                                        const rtx_nrc_sample* __restrict__ samples, StepShape shape,
                                                        Norm nm, float* __restrict__ partials,
                                                        float* __restrict__ loss_part) {
    __shared__ TrainLds sm;
    uint32_t n, stride, offset;
    // a step that does not run, and blocks past the step's samples, leave at once (the whole block alike)
    if (!step_samples(shape, n, stride, offset) || blockIdx.x * 64u >= n) return;
    const float inv_n = 1.0f / (float)n;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int col = wave * 16 + c;  // sample column within the block
    const uint64_t idx = (uint64_t)blockIdx.x * 64 + col;
    const bool valid = idx < n;
    float* part = partials + (uint64_t)blockIdx.x * RTX_NRC_PARAM_COUNT;
    bf16x8 fa[12], fb[8];  // fragments of the current and the next layer
    float ba[16], bb[16];
    load_fwd_layer<0>(fwd, params, lane, fa, ba);

    rtx_nrc_sample smp{};
    if (valid) {
        const uint64_t si = idx * stride + offset;
        load_query(&samples[si].query, smp.query);
        smp.target[0] = samples[si].target[0];
        smp.target[1] = samples[si].target[1];
        smp.target[2] = samples[si].target[2];
    }
    // forward, keeping every layer input in registers (B fragments) and in LDS (A^T)
    bf16x8 in0[3];
    encode(smp.query, nm, g, valid, in0, sm.at0, col);
    bf16x8 acts[5][2];
    f32x4 acc[4];
    load_fwd_layer<1>(fwd, params, lane, fb, bb);
    layer_forward_regs<0>(fa, ba, in0, acc);
    relu_pack(acc, acts[0]);
    store_act(acts[0], sm.at[0], g, col);
    load_fwd_layer<2>(fwd, params, lane, fa, ba);
    layer_forward_regs<1>(fb, bb, acts[0], acc);
    relu_pack(acc, acts[1]);
    store_act(acts[1], sm.at[1], g, col);
    load_fwd_layer<3>(fwd, params, lane, fb, bb);
    layer_forward_regs<2>(fa, ba, acts[1], acc);
    relu_pack(acc, acts[2]);
    store_act(acts[2], sm.at[2], g, col);
    load_fwd_layer<4>(fwd, params, lane, fa, ba);
    layer_forward_regs<3>(fb, bb, acts[2], acc);
    relu_pack(acc, acts[3]);
    store_act(acts[3], sm.at[3], g, col);
    load_fwd_layer<5>(fwd, params, lane, fb, bb);
    layer_forward_regs<4>(fa, ba, acts[3], acc);
    relu_pack(acc, acts[4]);
    store_act(acts[4], sm.at[4], g, col);
    load_bwd_layer<5>(bwd, lane, fa);
    f32x4 out[1];
    layer_forward_regs<5>(fb, bb, acts[4], out);

    // relative L2 loss on RGB (rows 0..2 of the output tile, held by k-group 0)
    f32x4 gy[1] = {{0.0f, 0.0f, 0.0f, 0.0f}};
    float loss = 0.0f;
    if (valid && g == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float y = out[0][i], t = smp.target[i];
            const float den = y * y + 0.01f;
            gy[0][i] = 2.0f * (y - t) / den * inv_n;
            loss += (y - t) * (y - t) / den;
        }
    }
    bf16x8 gfrag[2];
    store_grad(gy, 1, sm.gt, g, col, gfrag);
    __syncthreads();
    load_bwd_layer<4>(bwd, lane, fb);
    weight_grad<5>(sm, sm.at[4], part, wave, lane);
    f32x4 gv[4];
    back_data<5>(fa, gfrag, acts[4], gv);
    __syncthreads();
    store_grad(gv, 4, sm.gt, g, col, gfrag);
    __syncthreads();
    load_bwd_layer<3>(bwd, lane, fa);
    weight_grad<4>(sm, sm.at[3], part, wave, lane);
    back_data<4>(fb, gfrag, acts[3], gv);
    __syncthreads();
    store_grad(gv, 4, sm.gt, g, col, gfrag);
    __syncthreads();
    load_bwd_layer<2>(bwd, lane, fb);
    weight_grad<3>(sm, sm.at[2], part, wave, lane);
    back_data<3>(fa, gfrag, acts[2], gv);
    __syncthreads();
    store_grad(gv, 4, sm.gt, g, col, gfrag);
    __syncthreads();
    load_bwd_layer<1>(bwd, lane, fa);
    weight_grad<2>(sm, sm.at[1], part, wave, lane);
    back_data<2>(fb, gfrag, acts[1], gv);
    __syncthreads();
    store_grad(gv, 4, sm.gt, g, col, gfrag);
    __syncthreads();
    weight_grad<1>(sm, sm.at[0], part, wave, lane);
    back_data<1>(fa, gfrag, acts[0], gv);
    __syncthreads();
    store_grad(gv, 4, sm.gt, g, col, gfrag);
    __syncthreads();
    weight_grad<0>(sm, sm.at0, part, wave, lane);

    // block loss
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0) sm.red[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (sm.red[0] + sm.red[1] + sm.red[2] + sm.red[3]) * inv_n;
}

// BF16 copy of weight p into its forward fragment slot and, for layers 1..5, its backward fragment slot: the
// inverse of the gathers in nrc_pack_kernel (biases are read in FP32 and have no slot).
__device__ __forceinline__ void scatter_weight(int p, float w, __bf16* __restrict__ fwd, __bf16* __restrict__ bwd) {
    int l = 5;
    while (l > 0 && p < w_off(l)) --l;
    const int local = p - w_off(l);
    if (local >= out_dim(l) * in_dim(l)) return;
    const int row = local / in_dim(l), col = local % in_dim(l);
    const __bf16 b = (__bf16)w;
    {  // forward: tile t = row / 16, step s = col / 32; k order 8g + j (input layer) or kperm
        const int S = in_dim(l) / 32, t = row >> 4, r = row & 15, s = col >> 5, k = col & 31;
        const int g = l == 0 ? k >> 3 : (k & 15) >> 2;
        const int j = l == 0 ? k & 7 : ((k >> 4) << 2) + (k & 3);
        fwd[((fwd_base(l) + t * S + s) * 64 + g * 16 + r) * 8 + j] = b;
    }
    if (l >= 1) {  // backward: tile t = col / 16, step s = row / 32, k order kperm over output rows
        const int S = l == 5 ? 1 : 2, t = col >> 4, r = col & 15, s = row >> 5, k = row & 31;
        const int g = (k & 15) >> 2, j = ((k >> 4) << 2) + (k & 3);
        bwd[((bwd_base(l) + t * S + s) * 64 + g * 16 + r) * 8 + j] = b;
    }
}

// Sum the block partials and take one Adam step (writing the BF16 fragments of the updated weights), or only
// write the gradient.  A block covers 16 parameters with 16 threads each; thread k sums partials k, k + 16, ...
// and a fixed LDS tree adds the 16 slices (deterministic).  The extra last block sums the per-block losses the
// same way.
__global__ __launch_bounds__(256) void nrc_adam_kernel(const float* __restrict__ partials, StepShape shape,
                                                       float* __restrict__ params, float* __restrict__ m,
                                                       float* __restrict__ v, float lr, float b1, float b2, float eps,
                                                       const uint32_t* __restrict__ steps, float* __restrict__ grad_out,
                                                       __bf16* __restrict__ fwd, __bf16* __restrict__ bwd,
                                                       const float* __restrict__ loss_part,
                                                       float* __restrict__ loss_out) {
    __shared__ float s[16][17];
    uint32_t n, stride, offset;
    if (!step_samples(shape, n, stride, offset)) return;
    const uint32_t nblk = (n + 63) / 64;
    if (blockIdx.x == gridDim.x - 1) {
        __shared__ float ls[256];
        float a = 0.0f;
        for (uint32_t b = threadIdx.x; b < nblk; b += 256) a += loss_part[b];
        ls[threadIdx.x] = a;
        __syncthreads();
        for (uint32_t h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) ls[threadIdx.x] += ls[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) *loss_out = ls[0];
        return;
    }
    const uint32_t pi = threadIdx.x & 15u, k = threadIdx.x >> 4;
    const uint32_t p = blockIdx.x * 16u + pi;
    float acc = 0.0f;
    if (p < RTX_NRC_PARAM_COUNT)
        for (uint32_t b = k; b < nblk; b += 16) acc += partials[(uint64_t)b * RTX_NRC_PARAM_COUNT + p];
    s[k][pi] = acc;
    __syncthreads();
    for (uint32_t h = 8; h > 0; h >>= 1) {
        if (k < h) s[k][pi] += s[k + h][pi];
        __syncthreads();
    }
    if (k != 0 || p >= RTX_NRC_PARAM_COUNT) return;
    const float gsum = s[0][pi];
    if (grad_out) {
        grad_out[p] = gsum;
        return;
    }
    // bias corrections of step t = steps taken + this batch's index + 1 (FP64, rounded)
    const double t = (double)(*steps + shape.batch + 1u);
    const float c1 = (float)(1.0 / (1.0 - pow((double)b1, t)));
    const float c2 = (float)(1.0 / (1.0 - pow((double)b2, t)));
    const float mm = b1 * m[p] + (1.0f - b1) * gsum;
    const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    const float w = params[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    params[p] = w;
    scatter_weight((int)p, w, fwd, bwd);
}

// Rebuild the BF16 MFMA fragments from the FP32 master parameters.
__global__ __launch_bounds__(256) void nrc_pack_kernel(const float* __restrict__ params, bf16x8* __restrict__ fwd,
                                                       bf16x8* __restrict__ bwd) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = (kFwdFrags + kBwdFrags) * 64;
    if (id >= total) return;
    const int lane = id & 63, frag = id >> 6;
    const int r = lane & 15, g = lane >> 4;
    bf16x8 out;
    if (frag < kFwdFrags) {
        int l = 0;
        while (l < 5 && frag >= fwd_base(l + 1)) ++l;
        const int S = in_dim(l) / 32;
        const int local = frag - fwd_base(l), t = local / S, s = local % S;
        const int row = 16 * t + r;
        for (int j = 0; j < 8; ++j) {
            const int colk = l == 0 ? 32 * s + 8 * g + j : kperm(s, g, j);
            out[j] = (__bf16)params[w_off(l) + row * in_dim(l) + colk];
        }
        fwd[frag * 64 + lane] = out;
    } else {
        const int bf = frag - kFwdFrags;
        int l = 1;
        while (l < 5 && bf >= bwd_base(l + 1)) ++l;
        const int S = l == 5 ? 1 : 2;
        const int local = bf - bwd_base(l), t = local / S, s = local % S;
        const int colin = 16 * t + r;
        for (int j = 0; j < 8; ++j) {
            const int rowout = kperm(s, g, j);
            out[j] = rowout < out_dim(l) ? (__bf16)params[w_off(l) + rowout * in_dim(l) + colin] : (__bf16)0.0f;
        }
        bwd[bf * 64 + lane] = out;
    }
}

}  // namespace

uint64_t nrc_fragment_bytes() { return (uint64_t)(kFwdFrags + kBwdFrags) * 64 * sizeof(bf16x8); }

hipError_t nrc_pack(NrcState& st, hipStream_t s) {
    nrc_pack_kernel<<<((kFwdFrags + kBwdFrags) * 64 + 255) / 256, 256, 0, s>>>(st.params, (bf16x8*)st.fwd,
                                                                               (bf16x8*)st.bwd);
    return hipGetLastError();
}

static Norm make_norm(const NrcState& st) {
    Norm nm;
    for (int a = 0; a < 3; ++a) {
        nm.lo[a] = st.lo[a];
        const float e = st.hi[a] - st.lo[a];
        nm.inv[a] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    return nm;
}

hipError_t nrc_infer(const NrcState& st, const rtx_nrc_query* q, uint32_t n, float* out, hipStream_t s,
                     const uint32_t* n_dev, const InferGather* gather) {
    if (n == 0) return hipSuccess;
    // persistent LDS-staged kernel: up to 3 blocks per CU (46 KB of fragments each), each wave looping over pairs
    // of 16-query tiles
    const uint64_t pairs = ((uint64_t)n + 31) / 32;
    const uint64_t want = (pairs + 3) / 4;
    const unsigned blocks = (unsigned)std::min<uint64_t>(want, (uint64_t)st.cu_count * 3);
    nrc_infer_lds_kernel<<<blocks, 256, 0, s>>>((const bf16x8*)st.fwd, st.params, q, n, n_dev, make_norm(st), out,
                                                 gather ? *gather : InferGather{});
    return hipGetLastError();
}

// Adds the steps just taken to the step counter: one (host-sized step) or ceil(count / batch_size).
__global__ void nrc_advance_kernel(uint32_t* steps, const uint32_t* count_dev, uint32_t batch_size) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *steps += count_dev ? (*count_dev + batch_size - 1) / batch_size : 1u;
}

// One training step (gradient kernel + Adam kernel) of the given shape; the grid covers the largest batch.
static hipError_t launch_step(NrcState& st, const rtx_nrc_sample* samples, const StepShape& shape, uint32_t nblk,
                              float* grad_out, float* loss_out, hipStream_t s) {
    float* loss = loss_out ? loss_out : st.loss;
    nrc_train_kernel<<<nblk, 256, 0, s>>>((const bf16x8*)st.fwd, (const bf16x8*)st.bwd, st.params, samples, shape,
                                          make_norm(st), st.partials, st.loss_part);
    // one block per 16 parameters + the loss block
    const unsigned pb = (RTX_NRC_PARAM_COUNT + 15) / 16 + 1;
    if (grad_out) {
        nrc_adam_kernel<<<pb, 256, 0, s>>>(st.partials, shape, st.params, st.m, st.v, 0, 0, 0, 0, st.steps_dev,
                                           grad_out, nullptr, nullptr, st.loss_part, loss);
        return hipGetLastError();
    }
    nrc_adam_kernel<<<pb, 256, 0, s>>>(st.partials, shape, st.params, st.m, st.v, st.lr, st.beta1, st.beta2, st.eps,
                                       st.steps_dev, nullptr, (__bf16*)st.fwd, (__bf16*)st.bwd, st.loss_part, loss);
    return hipGetLastError();
}

hipError_t nrc_step(NrcState& st, const rtx_nrc_sample* samples, uint32_t n, float* grad_out, float* loss_out,
                    hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t nblk = (n + 63) / 64;
    if (nblk > st.max_blocks) return hipErrorInvalidValue;
    hipError_t e = launch_step(st, samples, StepShape{nullptr, n, 0u, st.batch_size}, nblk, grad_out, loss_out, s);
    if (e || grad_out) return e;
    nrc_advance_kernel<<<1, 64, 0, s>>>(st.steps_dev, nullptr, st.batch_size);
    return hipGetLastError();
}

hipError_t nrc_steps_device(NrcState& st, const rtx_nrc_sample* samples, const uint32_t* count_dev,
                            uint32_t max_batches, hipStream_t s) {
    for (uint32_t b = 0; b < max_batches; ++b) {
        hipError_t e = launch_step(st, samples, StepShape{count_dev, 0u, b, st.batch_size}, st.max_blocks, nullptr,
                                   nullptr, s);
        if (e) return e;
    }
    nrc_advance_kernel<<<1, 64, 0, s>>>(st.steps_dev, count_dev, st.batch_size);
    return hipGetLastError();
}

}  // namespace rtx
