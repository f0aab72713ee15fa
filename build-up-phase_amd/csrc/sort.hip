// sort.hip — stable onesweep LSD radix sort of (key, uint32 value) pairs for gfx950.
//
// Method: Adinets & Merrill 2022, "Onesweep: A Faster Least Significant Digit Radix Sort for GPUs".
//   1. one pass over the keys builds the 256-bin histogram of every 8-bit digit (LDS-privatised);
//   2. one 256-thread block turns the histograms into digit start offsets;
//   3. per digit pass, each block takes a 4096-key tile (tile id from an atomic ticket, so every tile it
//      waits for belongs to a block that is already running), ranks its keys stably in LDS using 64-lane
//      ballots, publishes its per-digit counts and resolves its global offsets with decoupled look-back,
//      stages the tile in LDS in digit order and writes it out in runs.
// The look-back spin is bounded: past the limit the block raises kErrSortTimeout and finishes.
#include "rtx_internal.h"
#include "sort.h"

namespace rtx {

namespace {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kTile = kSortBlock * kSortItems;
constexpr int kRadix = 256;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr int kSpinLimit = 1 << 22;

template <typename K>
__global__ __launch_bounds__(256) void hist_kernel(const K* __restrict__ keys, uint32_t n, int passes,
                                                   uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[8][kRadix];
    for (int i = threadIdx.x; i < 8 * kRadix; i += blockDim.x) (&h[0][0])[i] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        K k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p][(uint32_t)(k >> (8 * p)) & 255u], 1u);
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p)
        for (int d = threadIdx.x; d < kRadix; d += blockDim.x)
            if (h[p][d]) atomicAdd(&hist[p * kRadix + d], h[p][d]);
}

__global__ __launch_bounds__(256) void scan_hist_kernel(const uint32_t* __restrict__ hist, int passes,
                                                        uint32_t* __restrict__ start) {
    __shared__ uint32_t s[kRadix];
    int t = threadIdx.x;
    for (int p = 0; p < passes; ++p) {
        uint32_t c = hist[p * kRadix + t];
        s[t] = c;
        __syncthreads();
        for (int off = 1; off < kRadix; off <<= 1) {
            uint32_t v = t >= off ? s[t - off] : 0u;
            __syncthreads();
            s[t] += v;
            __syncthreads();
        }
        start[p * kRadix + t] = s[t] - c;
        __syncthreads();
    }
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__device__ __forceinline__ uint32_t status_load(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename K>
__global__ __launch_bounds__(kSortBlock) void onesweep_kernel(const K* __restrict__ keys_in,
                                                              const uint32_t* __restrict__ vals_in,
                                                              K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                              uint32_t n, int shift,
                                                              const uint32_t* __restrict__ digit_start,
                                                              uint32_t* __restrict__ status,
                                                              uint32_t* __restrict__ tile_counter,
                                                              uint32_t* __restrict__ err_flags) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_cnt[4][kRadix];
    __shared__ uint32_t s_wave_off[4][kRadix];
    __shared__ uint32_t s_run[kRadix];
    __shared__ uint32_t s_excl[kRadix];
    __shared__ uint32_t s_global[kRadix];
    __shared__ K s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];

    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    if (t == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = t; i < 4 * kRadix; i += kSortBlock) (&s_wave_cnt[0][0])[i] = 0;
    s_run[t] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tile_start = (uint64_t)tile * kTile;

    K key[kSortItems];
    uint32_t val[kSortItems];
    uint32_t rank[kSortItems];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        uint64_t idx = tile_start + (uint64_t)i * kSortBlock + t;
        bool valid = idx < n;
        key[i] = valid ? keys_in[idx] : (K)0;
        val[i] = valid ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
    }

#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        uint64_t idx = tile_start + (uint64_t)i * kSortBlock + t;
        bool valid = idx < n;
        uint32_t d = (uint32_t)(key[i] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bits = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bits : ~bits;
        }
        uint32_t lr = (uint32_t)__popcll(peers & lt_mask);
        if (valid && lr == 0) s_wave_cnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t c0 = s_wave_cnt[0][t], c1 = s_wave_cnt[1][t], c2 = s_wave_cnt[2][t], c3 = s_wave_cnt[3][t];
            uint32_t run = s_run[t];
            s_wave_off[0][t] = run;
            s_wave_off[1][t] = run + c0;
            s_wave_off[2][t] = run + c0 + c1;
            s_wave_off[3][t] = run + c0 + c1 + c2;
            s_run[t] = run + c0 + c1 + c2 + c3;
            s_wave_cnt[0][t] = 0;
            s_wave_cnt[1][t] = 0;
            s_wave_cnt[2][t] = 0;
            s_wave_cnt[3][t] = 0;
        }
        __syncthreads();
        rank[i] = valid ? s_wave_off[w][d] + lr : 0xFFFFFFFFu;
    }

    // tile-local exclusive prefix over digits (positions inside the staged tile)
    const uint32_t count = s_run[t];
    s_excl[t] = count;
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        uint32_t v = t >= off ? s_excl[t - off] : 0u;
        __syncthreads();
        s_excl[t] += v;
        __syncthreads();
    }
    s_excl[t] -= count;

    // decoupled look-back for digit t
    {
        uint32_t* st = status + (uint64_t)tile * kRadix + t;
        uint32_t excl = 0;
        if (tile == 0) {
            status_store(st, kFlagInc | count);
        } else {
            status_store(st, kFlagAgg | count);
            int64_t p = (int64_t)tile - 1;
            int spins = 0;
            while (p >= 0) {
                uint32_t v = status_load(status + (uint64_t)p * kRadix + t);
                uint32_t f = v & ~kValMask;
                if (f == 0u) {
                    if (++spins > kSpinLimit) {
                        atomicOr(err_flags, kErrSortTimeout);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += v & kValMask;
                if (f == kFlagInc) break;
                --p;
            }
            status_store(st, kFlagInc | ((excl + count) & kValMask));
        }
        s_global[t] = digit_start[t] + excl;
    }
    __syncthreads();

#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        if (rank[i] != 0xFFFFFFFFu) {
            uint32_t d = (uint32_t)(key[i] >> shift) & 255u;
            uint32_t pos = s_excl[d] + rank[i];
            s_keys[pos] = key[i];
            s_vals[pos] = val[i];
        }
    }
    __syncthreads();
    const uint64_t rem = (uint64_t)n - tile_start;
    const uint32_t tile_n = rem < (uint64_t)kTile ? (uint32_t)rem : (uint32_t)kTile;
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        uint32_t pos = (uint32_t)i * kSortBlock + t;
        if (pos < tile_n) {
            K k = s_keys[pos];
            uint32_t d = (uint32_t)(k >> shift) & 255u;
            uint32_t o = s_global[d] + (pos - s_excl[d]);
            if (o < n) {  // only false after a look-back timeout (reported through err_flags)
                keys_out[o] = k;
                vals_out[o] = s_vals[pos];
            }
        }
    }
}

}  // namespace

uint64_t sort_scratch_bytes(uint32_t n, int key_bytes) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    int passes = key_bytes;  // 8-bit digits
    uint64_t b = 0;
    b += align_up((uint64_t)n * key_bytes, 256);       // alternate keys
    b += align_up((uint64_t)n * 4, 256);               // alternate values
    b += align_up((uint64_t)passes * kRadix * 4, 256); // histograms
    b += align_up((uint64_t)passes * kRadix * 4, 256); // digit starts
    b += align_up((uint64_t)passes * 4, 256);          // tile counters
    b += align_up((uint64_t)passes * tiles * kRadix * 4, 256);  // look-back status
    return b;
}

template <typename K>
static hipError_t sort_pairs_impl(K* keys, uint32_t* vals, uint32_t n, int passes, void* scratch,
                                  uint32_t* err_flags, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint8_t* p = (uint8_t*)scratch;
    K* keys_alt = (K*)p;
    p += align_up((uint64_t)n * sizeof(K), 256);
    uint32_t* vals_alt = (uint32_t*)p;
    p += align_up((uint64_t)n * 4, 256);
    uint32_t* hist = (uint32_t*)p;
    p += align_up((uint64_t)sizeof(K) * kRadix * 4, 256);
    uint32_t* start = (uint32_t*)p;
    p += align_up((uint64_t)sizeof(K) * kRadix * 4, 256);
    uint32_t* counters = (uint32_t*)p;
    p += align_up((uint64_t)sizeof(K) * 4, 256);
    uint32_t* status = (uint32_t*)p;
    // hist, start, counters and status are contiguous: clear them with one fill
    uint64_t clear_words = ((uint8_t*)status - (uint8_t*)hist) / 4 + (uint64_t)passes * tiles * kRadix;
    fill_u32_kernel<<<1024, 256, 0, stream>>>(hist, clear_words, 0u);
    int hist_blocks = (int)std::min<uint64_t>(1024, (n + 4095) / 4096);
    hist_kernel<K><<<hist_blocks, 256, 0, stream>>>(keys, n, passes, hist);
    scan_hist_kernel<<<1, 256, 0, stream>>>(hist, passes, start);
    K* kin = keys;
    K* kout = keys_alt;
    uint32_t* vin = nullptr;
    uint32_t* vout = vals_alt;
    for (int pass = 0; pass < passes; ++pass) {
        onesweep_kernel<K><<<(unsigned)tiles, kSortBlock, 0, stream>>>(
            kin, vin, kout, vout, n, 8 * pass, start + pass * kRadix, status + (uint64_t)pass * tiles * kRadix,
            counters + pass, err_flags);
        std::swap(kin, kout);
        vin = vout;
        vout = (vout == vals_alt) ? vals : vals_alt;
    }
    // an even number of passes leaves keys in `keys` and values in `vals`
    return hipGetLastError();
}

hipError_t sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t n, void* scratch, uint32_t* err_flags,
                          hipStream_t stream) {
    return sort_pairs_impl<uint32_t>(keys, vals, n, 4, scratch, err_flags, stream);
}

hipError_t sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint32_t n, void* scratch, uint32_t* err_flags,
                          hipStream_t stream) {
    return sort_pairs_impl<uint64_t>(keys, vals, n, 8, scratch, err_flags, stream);
}

}  // namespace rtx
