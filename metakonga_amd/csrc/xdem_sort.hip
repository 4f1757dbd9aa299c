// xdem_sort.hip — stable LSD radix sort of (cell hash, index) pairs and the
// cell start/end scan, for gfx950.
//
// Replaces thrust::sort_by_key + cudaMemset + reorderDataAndFindCellStart_kernel
// (source/mphysics_cuda_def.cu:80-108, include/mphysics_cuda_impl.cuh:166-212).
// The reference sort is a stable radix sort, so its output order is the unique
// (hash ascending, input position ascending) order; any stable sort reproduces
// it bit for bit.
//
// Design (wave64 / LDS):
//  * 8-bit digits, only ceil(log2(ncell)) key bits are sorted (21 bits for the
//    reference 128^3 grid: 3 passes).
//  * per pass: tile histogram -> one-block exclusive scan (digit-major, so the
//    scan yields each tile's global digit offsets) -> stable scatter.
//  * within a wave, lanes holding the same digit are found with one __ballot per
//    digit bit; the lowest such lane updates the wave's LDS digit counter, so
//    nearly-sorted input (the common case: particles keep their cell between
//    steps) costs no LDS atomics and no serialised conflicts.
//  * the scatter stages the tile in LDS in digit order and writes each digit run
//    with consecutive lanes, so global writes are coalesced.
#include <mutex>

#include "../../include/xdem.h"
#include "xdem_common.hpp"

namespace xdem {

namespace {

constexpr int RADIX = 256;
constexpr int SORT_THREADS = 1024;  // 16 waves: a 4096-key tile with 4 keys per lane
constexpr int SORT_ITEMS = 4;
constexpr int TILE = SORT_THREADS * SORT_ITEMS;
constexpr int WAVES = SORT_THREADS / 64;

__device__ inline uint64_t lanemask_lt() {
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// Lanes (among `active`) whose digit equals this lane's digit.
__device__ inline uint64_t match_digit(uint32_t d, int nbits, uint64_t active) {
    uint64_t peers = active;
    for (int b = 0; b < nbits; ++b) {
        int bit = (d >> b) & 1;
        uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// Element e of the tile held by (wave w, lane l, item i): wave-striped so that
// the lanes of one item row read consecutive addresses and the processing order
// (i, lane) is the input order within each wave.
__device__ inline uint32_t tile_elem(uint32_t w, uint32_t i, uint32_t lane) {
    return w * (64 * SORT_ITEMS) + i * 64 + lane;
}

__global__ __launch_bounds__(SORT_THREADS) void k_hist(const uint32_t* __restrict__ keys,
                                                       uint32_t n, int shift, int nbits,
                                                       uint32_t* __restrict__ hist,
                                                       uint32_t ntiles) {
    __shared__ uint32_t wcnt[WAVES * RADIX];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int i = t; i < WAVES * RADIX; i += SORT_THREADS) wcnt[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TILE;
    const uint32_t dmask = (1u << nbits) - 1;
    for (int i = 0; i < SORT_ITEMS; ++i) {
        uint32_t e = base + tile_elem(w, i, lane);
        bool valid = e < n;
        uint32_t d = valid ? (keys[e] >> shift) & dmask : 0;
        uint64_t active = __ballot(valid);
        if (active == 0) continue;  // wave-uniform
        uint64_t peers = match_digit(d, nbits, active);
        if (valid && (peers & lanemask_lt()) == 0) wcnt[w * RADIX + d] += __popcll(peers);
    }
    __syncthreads();
    if (t < RADIX) {
        uint32_t s = 0;
        for (int k = 0; k < WAVES; ++k) s += wcnt[k * RADIX + t];
        hist[t * ntiles + blockIdx.x] = s;
    }
}

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    uint32_t lane = __lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Exclusive scan of one digit's row of tile counts (block d scans
// hist[d*ntiles .. (d+1)*ntiles) in place) and the digit's total.  One block
// per digit: 256 rows in parallel instead of one block over all of them.
constexpr int SCAN_THREADS = 256;
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_digits(uint32_t* __restrict__ hist,
                                                              uint32_t ntiles,
                                                              uint32_t* __restrict__ dtot) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    __shared__ uint32_t carry_s;
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < ntiles; base += SCAN_THREADS) {
        uint32_t e = base + t;
        uint32_t v = e < ntiles ? row[e] : 0;
        uint32_t inc = wave_incl_scan(v);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t before = carry_s;
        for (uint32_t k = 0; k < w; ++k) before += wsum[k];
        if (e < ntiles) row[e] = before + inc - v;
        __syncthreads();
        if (t == SCAN_THREADS - 1) carry_s = before + inc;
        __syncthreads();
    }
    if (t == 0) dtot[blockIdx.x] = carry_s;
}

__global__ __launch_bounds__(SORT_THREADS) void k_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t n, int shift, int nbits,
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ dtot, uint32_t ntiles) {
    __shared__ uint32_t wcnt[WAVES * RADIX];
    __shared__ uint32_t dbase[RADIX];
    __shared__ uint32_t lstart[RADIX];
    __shared__ uint32_t gstart[RADIX];
    __shared__ uint32_t wtot[WAVES];
    __shared__ uint32_t lkeys[TILE];
    __shared__ uint32_t lvals[TILE];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int i = t; i < WAVES * RADIX; i += SORT_THREADS) wcnt[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TILE;
    const uint32_t dmask = (1u << nbits) - 1;
    uint32_t key[SORT_ITEMS], val[SORT_ITEMS], rank[SORT_ITEMS];
    const uint64_t lt = lanemask_lt();
    for (int i = 0; i < SORT_ITEMS; ++i) {
        uint32_t e = base + tile_elem(w, i, lane);
        bool valid = e < n;
        key[i] = valid ? kin[e] : 0xFFFFFFFFu;
        val[i] = valid ? vin[e] : 0;
        uint32_t d = (key[i] >> shift) & dmask;
        uint64_t active = __ballot(valid);
        rank[i] = 0xFFFFFFFFu;
        if (active == 0) continue;
        uint64_t peers = match_digit(d, nbits, active);
        if (valid) {
            uint32_t c = wcnt[w * RADIX + d];
            rank[i] = c + __popcll(peers & lt);
            if ((peers & lt) == 0) wcnt[w * RADIX + d] = c + __popcll(peers);
        }
    }
    __syncthreads();
    // per digit: exclusive offsets over waves, and the tile total
    uint32_t total = 0;
    if (t < RADIX) {
        for (int k = 0; k < WAVES; ++k) {
            uint32_t c = wcnt[k * RADIX + t];
            wcnt[k * RADIX + t] = total;
            total += c;
        }
        gstart[t] = offsets[t * ntiles + blockIdx.x];
    }
    // digit bases: exclusive scan of the digit totals (every block redoes it;
    // the first RADIX threads, one digit each)
    {
        const uint32_t v = t < RADIX ? dtot[t] : 0u;
        const uint32_t inc = wave_incl_scan(v);
        if (lane == 63 && t < RADIX) wtot[w] = inc;
        __syncthreads();
        if (t < RADIX) {
            uint32_t before = 0;
            for (uint32_t k = 0; k < w; ++k) before += wtot[k];
            dbase[t] = before + inc - v;
            gstart[t] += dbase[t];
        }
        __syncthreads();
    }
    // exclusive scan of the tile's totals over the digits
    const uint32_t incl = wave_incl_scan(total);
    if (lane == 63 && t < RADIX) wtot[w] = incl;
    __syncthreads();
    if (t < RADIX) {
        uint32_t before = 0;
        for (uint32_t k = 0; k < w; ++k) before += wtot[k];
        lstart[t] = before + incl - total;
    }
    __syncthreads();
    for (int i = 0; i < SORT_ITEMS; ++i) {
        if (rank[i] == 0xFFFFFFFFu) continue;
        uint32_t d = (key[i] >> shift) & dmask;
        uint32_t p = lstart[d] + wcnt[w * RADIX + d] + rank[i];
        lkeys[p] = key[i];
        lvals[p] = val[i];
    }
    __syncthreads();
    const uint32_t tile_n = min((uint32_t)TILE, n - base);
    for (int i = 0; i < SORT_ITEMS; ++i) {
        uint32_t j = i * SORT_THREADS + t;
        if (j < tile_n) {
            uint32_t k = lkeys[j];
            uint32_t d = (k >> shift) & dmask;
            uint32_t o = gstart[d] + (j - lstart[d]);
            kout[o] = k;
            vout[o] = lvals[j];
        }
    }
}

// reorderDataAndFindCellStart_kernel (mphysics_cuda_impl.cuh:166-212): run
// boundaries of the sorted hashes give cstart/cend; sorted_index = index.
__global__ __launch_bounds__(512) void k_cell_ranges(const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals,
                                                     uint32_t* __restrict__ cstart,
                                                     uint32_t* __restrict__ cend,
                                                     uint32_t* __restrict__ sorted_index,
                                                     uint32_t n) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    uint32_t h = keys[id];
    if (id == 0 || h != keys[id - 1]) {
        cstart[h] = id;
        if (id > 0) cend[keys[id - 1]] = id;
    }
    if (id == n - 1) cend[h] = id + 1;
    sorted_index[id] = vals[id];
}

// Fill (vector stores when p is 16-byte aligned).
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t count, int aligned) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t n4 = aligned ? count / 4 : 0;
    uint4* p4 = reinterpret_cast<uint4*>(p);
    for (size_t k = i; k < n4; k += stride) p4[k] = make_uint4(v, v, v, v);
    for (size_t k = n4 * 4 + i; k < count; k += stride) p[k] = v;
}

}  // namespace

uint32_t bits_for(uint32_t ncell) {
    uint32_t b = 0;
    while (b < 32 && (uint64_t(1) << b) < ncell) ++b;
    return b;
}

int32_t fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s) {
    if (count == 0) return XDEM_OK;
    int aligned = ((uintptr_t)p & 15) == 0;
    size_t blocks = (count / 4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, count,
                       aligned);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

// 32-bit words of scratch the sort of n keys uses: ping-pong keys and values,
// the per-tile digit histogram and the digit totals.  Every word is written
// before it is read within one sort, so the scratch needs no clearing.
size_t sort_scratch_words(uint32_t n) {
    return 2 * (size_t)n + (size_t)RADIX * div_up(n, TILE) + RADIX + 64;
}

// The sort with a caller-provided scratch of sort_scratch_words(n) words.
int32_t sort_pairs_in(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits,
                      hipStream_t s, uint32_t* ws) {
    if (n <= 1 || key_bits == 0) return XDEM_OK;
    const uint32_t ntiles = div_up(n, TILE);
    const size_t hist_n = (size_t)RADIX * ntiles;
    uint32_t* tk = ws;
    uint32_t* tv = tk + n;
    uint32_t* hist = tv + n;
    uint32_t* dtot = hist + hist_n;
    uint32_t* src_k = keys;
    uint32_t* src_v = vals;
    uint32_t* dst_k = tk;
    uint32_t* dst_v = tv;
    const int passes = (int)div_up(key_bits, 8);
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        int nbits = (int)key_bits - shift < 8 ? (int)key_bits - shift : 8;
        hipLaunchKernelGGL(k_hist, dim3(ntiles), dim3(SORT_THREADS), 0, s, src_k, n, shift, nbits,
                           hist, ntiles);
        XDEM_LAUNCHED();
        hipLaunchKernelGGL(k_scan_digits, dim3(RADIX), dim3(SCAN_THREADS), 0, s, hist, ntiles, dtot);
        XDEM_LAUNCHED();
        hipLaunchKernelGGL(k_scatter, dim3(ntiles), dim3(SORT_THREADS), 0, s, src_k, src_v, dst_k,
                           dst_v, n, shift, nbits, hist, dtot, ntiles);
        XDEM_LAUNCHED();
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    if (src_k != keys) {
        XDEM_HIP(hipMemcpyAsync(keys, src_k, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
        XDEM_HIP(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    return XDEM_OK;
}

int32_t sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, hipStream_t s) {
    if (n <= 1 || key_bits == 0) return XDEM_OK;
    XDEM_TRY(workspace(0).reserve(sizeof(uint32_t) * sort_scratch_words(n)));
    return sort_pairs_in(keys, vals, n, key_bits, s, static_cast<uint32_t*>(workspace(0).ptr));
}

int32_t cell_ranges(const uint32_t* keys, const uint32_t* vals, uint32_t* cstart, uint32_t* cend,
                    uint32_t* sorted_index, uint32_t n, hipStream_t s) {
    if (n == 0) return XDEM_OK;
    hipLaunchKernelGGL(k_cell_ranges, dim3(div_up(n, 512)), dim3(512), 0, s, keys, vals, cstart,
                       cend, sorted_index, n);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

// ------------------------------------------- no-allocation form of the reorder
// The scratch of xdem_reorder_data_and_find_cell_start_noalloc: a file-scope
// device array (each device holds its own copy for the life of the process),
// sized for XDEM_NOALLOC_MAX_KEYS keys.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
constexpr size_t NOALLOC_WORDS = 2 * (size_t)XDEM_NOALLOC_MAX_KEYS +
                                 (size_t)RADIX * ((XDEM_NOALLOC_MAX_KEYS + TILE - 1) / TILE) +
                                 RADIX + 64;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[MAX_DEVICES] = {};
}  // namespace

// Device address of `scratch` on the current device (looked up once per device).
int32_t noalloc_scratch(uint32_t** out) {
    int dev = 0;
    XDEM_HIP(hipGetDevice(&dev));
    XDEM_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        XDEM_REQUIRE(e == hipSuccess && p, "hipGetSymbolAddress(scratch) failed on device %d: %s",
                     dev, hipGetErrorString(e));
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[dev];
    return XDEM_OK;
}

}  // namespace xdem

extern "C" int32_t xdem_reorder_data_and_find_cell_start_noalloc(uint32_t* hash, uint32_t* index,
                                                                 uint32_t* cstart, uint32_t* cend,
                                                                 uint32_t* sorted_index,
                                                                 uint32_t n, uint32_t ncell) {
    using namespace xdem;
    XDEM_REQUIRE(ncell > 0, "ncell == 0");
    XDEM_REQUIRE(n <= XDEM_NOALLOC_MAX_KEYS,
                 "n = %u exceeds the %u keys of the no-allocation scratch (use "
                 "xdem_reorder_data_and_find_cell_start)",
                 n, (uint32_t)XDEM_NOALLOC_MAX_KEYS);
    XDEM_REQUIRE(n == 0 || (hash && index && sorted_index), "null array");
    XDEM_REQUIRE(cstart && cend, "null cell arrays");
    uint32_t* ws = nullptr;
    XDEM_TRY(noalloc_scratch(&ws));
    hipStream_t s = stream();
    XDEM_TRY(sort_pairs_in(hash, index, n, bits_for(ncell), s, ws));
    XDEM_TRY(fill_u32(cstart, 0xFFFFFFFFu, ncell, s));
    XDEM_TRY(fill_u32(cend, 0u, ncell, s));
    XDEM_TRY(cell_ranges(hash, index, cstart, cend, sorted_index, n, s));
    return XDEM_OK;
}

extern "C" int32_t xdem_reorder_data_and_find_cell_start(uint32_t* hash, uint32_t* index,
                                                         uint32_t* cstart, uint32_t* cend,
                                                         uint32_t* sorted_index, uint32_t n,
                                                         uint32_t ncell) {
    using namespace xdem;
    XDEM_REQUIRE(ncell > 0, "ncell == 0");
    XDEM_REQUIRE(n == 0 || (hash && index && sorted_index), "null array");
    XDEM_REQUIRE(cstart && cend, "null cell arrays");
    hipStream_t s = stream();
    XDEM_TRY(sort_pairs(hash, index, n, bits_for(ncell), s));
    // cudaMemset(cstart, 0xff) / cudaMemset(cend, 0) — mphysics_cuda_def.cu:97-98
    XDEM_TRY(fill_u32(cstart, 0xFFFFFFFFu, ncell, s));
    XDEM_TRY(fill_u32(cend, 0u, ncell, s));
    XDEM_TRY(cell_ranges(hash, index, cstart, cend, sorted_index, n, s));
    return XDEM_OK;
}
