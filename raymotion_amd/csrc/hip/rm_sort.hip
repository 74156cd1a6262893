// Device scan and radix sort of the GPU tree build (rm_build.hip): the
// exclusive sum that compacts the surviving clusters of each clustering round
// in order, and the stable sort of (Morton code, triangle) pairs that puts the
// triangles in spatial order.
//
// Both work on tiles of 2048 elements, one per 256-thread block (8 per lane).
// The scan: per-tile sums, one block scanning the tile sums with a running
// carry, then each tile scanned again with its offset. The sort: least
// significant digit first, seven passes of 9 bits over key bits [0, 63); a
// pass counts the digits of every tile, scans the digit-major table of counts
// (digit 0 of every tile, then digit 1, ...), which gives each (digit, tile)
// its first output position, and scatters the tile. Inside a tile the 4 waves
// take 512 consecutive elements each, 64 per row; a lane's rank among the
// elements of its digit before it comes from ballots (the lanes of its row with
// the same 9 digit bits, below it) and a per-wave running count per digit in
// LDS, so the order of equal digits is the input order: every pass is stable
// and the sort orders by key, ties by input position.
#include "rm_sort.h"

namespace rmh {
namespace {

constexpr int kT = 256;              // threads per block
constexpr int kPer = 8;              // elements per thread
constexpr int kTile = kT * kPer;     // elements per tile
constexpr int kDigitBits = 9;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kPasses = 7;           // 7 x 9 = 63 key bits

__host__ __device__ inline unsigned tiles_of(int n) { return unsigned((size_t(n) + kTile - 1) / kTile); }

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    const int lane = int(__lane_id());
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, unsigned(o));
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive prefix of `v` over the block's threads (thread order), and the
// block total in *total. Every thread calls it.
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kT / 64];
    const int lane = int(__lane_id()), wave = int(threadIdx.x) / 64;
    const uint32_t inc = wave_inclusive_sum(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kT / 64; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();  // wsum is reused by the next call
    *total = all;
    return before + inc - v;
}

// The thread's 8 consecutive elements of tile `tile` (0 beyond n).
__device__ __forceinline__ void load8(const uint32_t* __restrict__ in, int n, unsigned tile, uint32_t v[kPer]) {
    const size_t i0 = size_t(tile) * kTile + size_t(threadIdx.x) * kPer;
#pragma unroll
    for (int j = 0; j < kPer; ++j) v[j] = i0 + j < size_t(n) ? in[i0 + j] : 0u;
}

__global__ __launch_bounds__(kT) void k_scan_tiles(const uint32_t* __restrict__ in, int n, uint32_t* __restrict__ sums) {
    uint32_t v[kPer];
    load8(in, n, blockIdx.x, v);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) s += v[j];
    uint32_t total;
    (void)block_exclusive_sum(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block: exclusive sum of the `count` tile sums in place, chunk by chunk.
__global__ __launch_bounds__(kT) void k_scan_carry(uint32_t* __restrict__ sums, int count) {
    uint32_t carry = 0;
    for (unsigned c = 0; c < tiles_of(count); ++c) {
        uint32_t v[kPer];
        load8(sums, count, c, v);
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) s += v[j];
        uint32_t total;
        uint32_t run = carry + block_exclusive_sum(s, &total);
        const size_t i0 = size_t(c) * kTile + size_t(threadIdx.x) * kPer;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            if (i0 + j < size_t(count)) sums[i0 + j] = run;
            run += v[j];
        }
        carry += total;
    }
}

// Exclusive sum of each tile plus its offset (sums[tile], or 0 without sums).
// `out` may be `in`: a block reads its whole tile before writing it.
__global__ __launch_bounds__(kT) void k_scan_apply(const uint32_t* in, int n, const uint32_t* __restrict__ sums,
                                                   uint32_t* out) {
    uint32_t v[kPer];
    load8(in, n, blockIdx.x, v);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) s += v[j];
    uint32_t total;
    uint32_t run = block_exclusive_sum(s, &total) + (sums ? sums[blockIdx.x] : 0u);
    const size_t i0 = size_t(blockIdx.x) * kTile + size_t(threadIdx.x) * kPer;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        if (i0 + j < size_t(n)) out[i0 + j] = run;
        run += v[j];
    }
}

// Digit counts of tile b: hist[d * tiles + b].
__global__ __launch_bounds__(kT) void k_sort_count(const uint64_t* __restrict__ keys, int n, int shift,
                                                   uint32_t* __restrict__ hist, unsigned tiles) {
    __shared__ uint32_t h[kDigits];
    for (int d = int(threadIdx.x); d < kDigits; d += kT) h[d] = 0;
    __syncthreads();
    const size_t base = size_t(blockIdx.x) * kTile;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const size_t i = base + size_t(r) * kT + threadIdx.x;
        if (i < size_t(n)) atomicAdd(&h[int(keys[i] >> shift) & (kDigits - 1)], 1u);
    }
    __syncthreads();
    for (int d = int(threadIdx.x); d < kDigits; d += kT) hist[size_t(d) * tiles + blockIdx.x] = h[d];
}

// Scatter of tile b to the positions that the scanned counts (first[d * tiles
// + b]: where digit d of tile b starts) and the in-tile ranks give.
__global__ __launch_bounds__(kT) void k_sort_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
                                                     int shift, const uint32_t* __restrict__ first, unsigned tiles) {
    __shared__ uint32_t cnt[kT / 64][kDigits];
    const int lane = int(__lane_id()), wave = int(threadIdx.x) / 64;
    for (int d = int(threadIdx.x); d < kDigits; d += kT)
#pragma unroll
        for (int w = 0; w < kT / 64; ++w) cnt[w][d] = 0;
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    const size_t base = size_t(blockIdx.x) * kTile + size_t(wave) * (kPer * 64);
    uint64_t k[kPer];
    uint32_t v[kPer], rank[kPer];
    int dg[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const size_t i = base + size_t(r) * 64 + size_t(lane);
        const bool valid = i < size_t(n);
        k[r] = valid ? kin[i] : 0ull;
        v[r] = valid ? vin[i] : 0u;
        const int d = int(k[r] >> shift) & (kDigits - 1);
        dg[r] = valid ? d : -1;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const unsigned long long m = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const uint32_t ahead = uint32_t(__popcll(peers & below));
        uint32_t old = 0;
        if (valid) old = cnt[wave][d];
        rank[r] = old + ahead;
        if (valid && ahead == 0) cnt[wave][d] = old + uint32_t(__popcll(peers));
    }
    __syncthreads();
    // per digit: the tile's first position, then each wave's share in order
    for (int d = int(threadIdx.x); d < kDigits; d += kT) {
        uint32_t run = first[size_t(d) * tiles + blockIdx.x];
#pragma unroll
        for (int w = 0; w < kT / 64; ++w) {
            const uint32_t c = cnt[w][d];
            cnt[w][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        if (dg[r] < 0) continue;
        const uint32_t p = cnt[wave][dg[r]] + rank[r];
        kout[p] = k[r];
        vout[p] = v[r];
    }
}

}  // namespace

size_t scan_tmp_bytes(int n) { return (size_t(tiles_of(n > 0 ? n : 1)) * 4 + 255) & ~size_t(255); }

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int n, void* tmp, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned tiles = tiles_of(n);
    uint32_t* sums = static_cast<uint32_t*>(tmp);
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kT), 0, st, in, n, nullptr, out);
    } else {
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kT), 0, st, in, n, sums);
        hipLaunchKernelGGL(k_scan_carry, dim3(1), dim3(kT), 0, st, sums, int(tiles));
        hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kT), 0, st, in, n, sums, out);
    }
    return hipGetLastError();
}

size_t sort_tmp_bytes(int n) {
    const size_t counts = size_t(tiles_of(n > 0 ? n : 1)) * kDigits;
    return ((counts * 4 + 255) & ~size_t(255)) + scan_tmp_bytes(int(counts));
}

hipError_t sort_pairs_63(uint64_t* keys_in, uint32_t* vals_in, uint64_t* keys_out, uint32_t* vals_out, int n,
                         void* tmp, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned tiles = tiles_of(n);
    const size_t counts = size_t(tiles) * kDigits;
    if (counts >= (size_t(1) << 31)) return hipErrorInvalidValue;
    uint32_t* hist = static_cast<uint32_t*>(tmp);
    void* scan_tmp = static_cast<char*>(tmp) + ((counts * 4 + 255) & ~size_t(255));
    // An odd number of passes: the last one writes keys_out / vals_out.
    static_assert(kPasses % 2 == 1, "passes alternate from the input buffers to the output buffers");
    uint64_t* ks[2] = {keys_in, keys_out};
    uint32_t* vs[2] = {vals_in, vals_out};
    for (int p = 0; p < kPasses; ++p) {
        const int src = p & 1, dst = src ^ 1;
        const int shift = p * kDigitBits;
        hipLaunchKernelGGL(k_sort_count, dim3(tiles), dim3(kT), 0, st, ks[src], n, shift, hist, tiles);
        if (hipError_t e = exclusive_scan(hist, hist, int(counts), scan_tmp, st); e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(kT), 0, st, ks[src], vs[src], ks[dst], vs[dst], n, shift,
                           hist, tiles);
    }
    return hipGetLastError();
}

}  // namespace rmh
