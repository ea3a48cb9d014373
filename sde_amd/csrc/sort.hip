// Key/value radix sort of the visibilities, written for CDNA4.
//
// The cell sort orders the visibilities by tile-major cell key; the values
// are the original index and the weight of each point.  The sort is
// stable, so points of one cell keep their original order and the sorted
// order is deterministic (so is every sum taken over it in a fixed order).  Only bits
// [0, nbits) order the keys; bits above ride along (bit 31 carries a flag
// of the prep pass).
//
// Least-significant-digit first, 8 bits a pass.  A pass is a reduce-then-
// scan over tiles of kTile keys, with no waits between workgroups:
//   k_sort_count       each tile's digit counts (one workgroup per tile; LDS
//                      atomics into 16 sub-histograms, so that runs of one
//                      digit in neighbouring lanes spread over banks);
//   k_sort_chunk       per chunk of kChunkTiles tiles: each tile's counts
//                      replaced by their exclusive prefix within the chunk,
//                      and the chunk's totals;
//   k_sort_chunk_scan  one workgroup: the chunk totals replaced by the output
//                      slot of each chunk's first key of every digit (the
//                      digit's start over all keys plus the earlier chunks);
//   k_sort_scatter     one workgroup per tile: the keys ranked stably in the
//                      tile (per wave: the lanes of equal digit found as the
//                      OR of their lane bits in an LDS mask per digit -- the
//                      same mask whatever order the ORs land in -- and a
//                      running count per wave; across waves: LDS prefix counts
//                      in input order), staged in LDS in digit order and
//                      written out as one contiguous run per digit from the
//                      tile's slot of that digit.  Workgroups are dealt to the
//                      XCDs in turn, so workgroup b runs tile
//                      xcd_tile(b): each XCD takes a contiguous range of
//                      tiles, whose runs of one digit are neighbours in the
//                      output and fill whole cache lines in that XCD's L2.
// Every pass's counts come from kernels, so the sort needs no memset and
// can be captured in a graph and replayed.  The counts are 32-bit: n < 2^31.
// The pass that writes the keys may count the first pass's digits itself,
// tile by tile into the same scratch (sort_tile_counts); the sort then
// starts at k_sort_chunk.
#include <algorithm>

#include "internal.h"

namespace {

constexpr int kBits = 8;
constexpr int kBins = 1 << kBits;
constexpr int kThr = 1024;              // threads of a scatter workgroup (16 waves)
constexpr int kWaves = kThr / 64;
constexpr int kRounds = 4;              // keys per thread
constexpr int kTile = kThr * kRounds;   // keys per tile
constexpr int kSeg = 64 * kRounds;      // keys of one wave: a contiguous segment of the tile
constexpr int kCountThr = 256;          // threads of a count workgroup
constexpr int kCountKeys = kTile / kCountThr;
constexpr int kGroupTiles = 32;         // tiles per thread group of k_sort_chunk
constexpr int kChunkTiles = 4 * kGroupTiles;
constexpr int kXcd = 8;                 // XCDs of the MI355X
static_assert(kWaves * kBins <= kTile, "the match masks fit in the value staging area");
static_assert(kTile == kSortTileKeys && kBins == kSortBins, "the tiles and digits a prep pass counts");

struct SortTemp {
    uint32_t *keys;   // n: intermediate keys
    uint64_t *vals;   // n: intermediate values
    unsigned *tcount; // tiles * kBins: tile counts, then their prefix within the chunk
    unsigned *csum;   // chunks * kBins: chunk totals, then each chunk's first output slot
};

__host__ __device__ inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }
inline int sort_tiles(int64_t n) { return (int)((n + kTile - 1) / kTile); }
inline int sort_chunks(int tiles) { return (tiles + kChunkTiles - 1) / kChunkTiles; }
inline int sort_passes(int nbits) { return (nbits + kBits - 1) / kBits; }

SortTemp carve(void *temp, int64_t n) {
    char *p = static_cast<char *>(temp);
    const int tiles = sort_tiles(n);
    SortTemp t;
    t.keys = reinterpret_cast<uint32_t *>(p);
    p += align256(sizeof(uint32_t) * (size_t)n);
    t.vals = reinterpret_cast<uint64_t *>(p);
    p += align256(sizeof(uint64_t) * (size_t)n);
    t.tcount = reinterpret_cast<unsigned *>(p);
    p += align256(sizeof(unsigned) * (size_t)tiles * kBins);
    t.csum = reinterpret_cast<unsigned *>(p);
    return t;
}

size_t temp_bytes_for(int64_t n) {
    const int tiles = sort_tiles(n);
    return align256(sizeof(uint32_t) * (size_t)n) + align256(sizeof(uint64_t) * (size_t)n) +
           align256(sizeof(unsigned) * (size_t)tiles * kBins) +
           align256(sizeof(unsigned) * (size_t)sort_chunks(tiles) * kBins);
}

__device__ __forceinline__ unsigned digit_of(uint32_t key, int shift, unsigned mask) { return (key >> shift) & mask; }

// The tile a scatter workgroup sorts: workgroups are dealt to the kXcd XCDs
// in turn (b % kXcd), and XCD x takes the contiguous range of tiles
// [x*q + min(x, r), ...) of q or q + 1 tiles, in the order its workgroups
// start.  A bijection of [0, T) whatever the XCD count; only the speed
// depends on it.
__device__ __forceinline__ int xcd_tile(int b, int T) {
    const int q = T / kXcd, r = T % kXcd, x = b % kXcd, j = b / kXcd;
    return x * q + min(x, r) + j;
}

template <bool VEC>
__global__ __launch_bounds__(kCountThr) void k_sort_count(const uint32_t *__restrict__ keys, int64_t n, int shift,
                                                         int bits, unsigned *__restrict__ tcount) {
    constexpr int kSub = 16;  // sub-histograms: (wave, lane & 3)
    constexpr int kStride = kBins + 1;  // one word apart on the banks for each sub-histogram
    __shared__ unsigned h[kSub][kStride];
    for (int k = threadIdx.x; k < kSub * kStride; k += kCountThr) (&h[0][0])[k] = 0u;
    __syncthreads();
    const unsigned mask = (1u << bits) - 1u;
    const int lane = threadIdx.x & 63, sub = (threadIdx.x >> 6) * 4 + (lane & 3);
    const int64_t base = (int64_t)blockIdx.x * kTile;
    uint32_t key[kCountKeys];
    bool valid[kCountKeys];
    if (VEC) {  // keys 16-byte aligned: four keys a load
#pragma unroll
        for (int r = 0; r < kCountKeys / 4; r++) {
            const int64_t i = base + ((int64_t)r * kCountThr + threadIdx.x) * 4;
            if (i + 3 < n) {
                const uint4 x = *reinterpret_cast<const uint4 *>(keys + i);
                key[4 * r] = x.x;
                key[4 * r + 1] = x.y;
                key[4 * r + 2] = x.z;
                key[4 * r + 3] = x.w;
#pragma unroll
                for (int e = 0; e < 4; e++) valid[4 * r + e] = true;
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    valid[4 * r + e] = i + e < n;
                    key[4 * r + e] = valid[4 * r + e] ? keys[i + e] : 0u;
                }
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < kCountKeys; r++) {
            const int64_t i = base + (int64_t)r * kCountThr + threadIdx.x;
            valid[r] = i < n;
            key[r] = valid[r] ? keys[i] : 0u;
        }
    }
#pragma unroll
    for (int r = 0; r < kCountKeys; r++)
        if (valid[r]) atomicAdd(&h[sub][digit_of(key[r], shift, mask)], 1u);
    __syncthreads();
    unsigned s = 0;
#pragma unroll
    for (int k = 0; k < kSub; k++) s += h[k][threadIdx.x];
    tcount[(size_t)blockIdx.x * kBins + threadIdx.x] = s;
}

// Thread (group g, digit d) holds the counts of kGroupTiles tiles of the chunk.
__global__ __launch_bounds__(1024) void k_sort_chunk(unsigned *__restrict__ tcount, int tiles,
                                                    unsigned *__restrict__ csum) {
    __shared__ unsigned gs[4][kBins];
    const int d = threadIdx.x & (kBins - 1), g = threadIdx.x >> 8;
    const int t0 = blockIdx.x * kChunkTiles + g * kGroupTiles;
    unsigned c[kGroupTiles], s = 0;
#pragma unroll
    for (int k = 0; k < kGroupTiles; k++) {
        c[k] = t0 + k < tiles ? tcount[(size_t)(t0 + k) * kBins + d] : 0u;
        s += c[k];
    }
    gs[g][d] = s;
    __syncthreads();
    unsigned run = 0;
    for (int q = 0; q < g; q++) run += gs[q][d];
#pragma unroll
    for (int k = 0; k < kGroupTiles; k++) {
        if (t0 + k < tiles) tcount[(size_t)(t0 + k) * kBins + d] = run;
        run += c[k];
    }
    if (g == 3) csum[(size_t)blockIdx.x * kBins + d] = run;
}

// One workgroup; thread (quarter q, digit d) walks a quarter of the chunks.
__global__ __launch_bounds__(1024) void k_sort_chunk_scan(unsigned *__restrict__ csum, int nchunks) {
    constexpr int B = 16;
    __shared__ unsigned part[4][kBins];
    __shared__ unsigned dbase[kBins];
    const int d = threadIdx.x & (kBins - 1), q = threadIdx.x >> 8;
    const int per = (nchunks + 3) / 4;
    const int c0 = min(nchunks, q * per), c1 = min(nchunks, c0 + per);
    unsigned s = 0;
    for (int c = c0; c < c1; c += B) {
        unsigned x[B];
#pragma unroll
        for (int k = 0; k < B; k++) x[k] = c + k < c1 ? csum[(size_t)(c + k) * kBins + d] : 0u;
#pragma unroll
        for (int k = 0; k < B; k++) s += x[k];
    }
    part[q][d] = s;
    __syncthreads();
    // each digit's start over all keys: exclusive scan of the digit totals
    if (threadIdx.x < 64) {
        unsigned c[4], tot = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int dd = threadIdx.x * 4 + k;
            c[k] = part[0][dd] + part[1][dd] + part[2][dd] + part[3][dd];
            tot += c[k];
        }
        unsigned incl = tot;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned y = __shfl_up(incl, off);
            if ((int)threadIdx.x >= off) incl += y;
        }
        unsigned e = incl - tot;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            dbase[threadIdx.x * 4 + k] = e;
            e += c[k];
        }
    }
    __syncthreads();
    unsigned run = dbase[d];
    for (int qq = 0; qq < q; qq++) run += part[qq][d];
    for (int c = c0; c < c1; c += B) {
        unsigned x[B];
#pragma unroll
        for (int k = 0; k < B; k++) x[k] = c + k < c1 ? csum[(size_t)(c + k) * kBins + d] : 0u;
#pragma unroll
        for (int k = 0; k < B; k++) {
            if (c + k < c1) csum[(size_t)(c + k) * kBins + d] = run;
            run += x[k];
        }
    }
}

// One pass: digit bits [shift, shift + bits) of keys_in/vals_in, in the
// order of the keys, to keys_out/vals_out.  Wave w of the tile holds keys
// [w*kSeg, (w+1)*kSeg) of it, round r of them at r*64 + lane: the tile's
// input order is (wave, round, lane).
template <bool WTV>
__global__ __launch_bounds__(kThr) void k_sort_scatter(const uint32_t *__restrict__ keys_in,
                                                      const uint64_t *__restrict__ vals_in,
                                                      const float *__restrict__ vals_wt,
                                                      uint32_t *__restrict__ keys_out,
                                                      uint64_t *__restrict__ vals_out, int64_t n, int shift, int bits,
                                                      const unsigned *__restrict__ tcount,
                                                      const unsigned *__restrict__ csum) {
    // wcnt[w * kBins + d]: keys of digit d in wave w so far, then the
    // exclusive prefix of those counts over the waves
    __shared__ unsigned short wcnt[kWaves * kBins];
    __shared__ uint32_t skey[kTile];
    __shared__ uint64_t sval[kTile];
    __shared__ unsigned dstart[kBins];  // first local slot of each digit (tile in digit order)
    __shared__ unsigned goff[kBins];    // output slot of the tile's first key of each digit
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned mask = (1u << bits) - 1u;
    for (int k = threadIdx.x; k < kWaves * kBins; k += kThr) wcnt[k] = 0;
    __syncthreads();
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t seg = (int64_t)tile * kTile + wave * kSeg;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned short *wc = wcnt + wave * kBins;
    // the wave's match masks of the 256 digits, in the value staging area
    // (unused until the staging below)
    unsigned long long *mk = reinterpret_cast<unsigned long long *>(sval) + wave * kBins;
    for (int k = lane; k < kBins; k += 64) mk[k] = 0ull;

    uint32_t key[kRounds];
    uint64_t val[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const int64_t i = seg + r * 64 + lane;
        key[r] = i < n ? keys_in[i] : 0u;
        if (WTV) val[r] = i < n ? ((uint64_t)(uint32_t)i | ((uint64_t)__float_as_uint(vals_wt[i]) << 32)) : 0ull;
        else val[r] = i < n ? vals_in[i] : 0ull;
    }
    // rank within the wave: the lanes of equal digit in this round (the OR of
    // their lane bits), after the wave's keys of that digit in earlier
    // rounds.  Every lane of a digit reads the mask and the running count
    // before its first lane clears the one and adds to the other: the LDS
    // accesses of one wave complete in program order.
    unsigned info[kRounds];  // digit | rank in the wave << 16; digit kBins: past the end
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const bool valid = seg + r * 64 + lane < n;
        const unsigned d = digit_of(key[r], shift, mask);
        if (valid) atomicOr(&mk[d], 1ull << lane);
        const unsigned long long m = valid ? mk[d] : 0ull;
        const unsigned rank = (unsigned)__popcll(m & below);
        const unsigned prior = valid ? wc[d] : 0u;
        if (valid && rank == 0) {
            mk[d] = 0ull;
            wc[d] = (unsigned short)(prior + (unsigned)__popcll(m));
        }
        info[r] = valid ? (d | ((prior + rank) << 16)) : (unsigned)kBins;
    }
    __syncthreads();

    // per digit: exclusive prefix over the waves, the tile's count and its
    // output slot
    if (threadIdx.x < kBins) {
        const int d = threadIdx.x;
        unsigned run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const unsigned c = wcnt[w * kBins + d];
            wcnt[w * kBins + d] = (unsigned short)run;
            run += c;
        }
        dstart[d] = run;  // the tile's count of digit d, scanned below
        goff[d] = csum[(size_t)(tile / kChunkTiles) * kBins + d] + tcount[(size_t)tile * kBins + d];
    }
    __syncthreads();
    // local start of each digit: exclusive scan of the tile's counts
    if (threadIdx.x < 64) {
        unsigned c[4], s = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            c[q] = dstart[threadIdx.x * 4 + q];
            s += c[q];
        }
        unsigned incl = s;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned y = __shfl_up(incl, off);
            if ((int)threadIdx.x >= off) incl += y;
        }
        unsigned e = incl - s;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            dstart[threadIdx.x * 4 + q] = e;
            e += c[q];
        }
    }
    __syncthreads();
    // stage the tile in digit order
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const unsigned d = info[r] & 0xffffu;
        if (d < (unsigned)kBins) {
            const unsigned slot = dstart[d] + wcnt[wave * kBins + d] + (info[r] >> 16);
            skey[slot] = key[r];
            sval[slot] = val[r];
        }
    }
    __syncthreads();
    // write out: each digit's keys as one contiguous run
    const int64_t left = n - (int64_t)tile * kTile;
    const int cnt_tile = left < kTile ? (int)left : kTile;
    for (int s = threadIdx.x; s < cnt_tile; s += kThr) {
        const uint32_t k = skey[s];
        const unsigned d = digit_of(k, shift, mask);
        const unsigned o = goff[d] + (unsigned)(s - (int)dstart[d]);
        keys_out[o] = k;
        vals_out[o] = sval[s];
    }
}

}  // namespace

unsigned *sort_tile_counts(void *temp, int64_t n) { return carve(temp, n > 0 ? n : 1).tcount; }

size_t sort_temp_bytes(int64_t n, int nbits) {
    if (n < 0 || n >= (int64_t)1 << 31 || nbits < 1 || nbits > 32) return 0;
    return temp_bytes_for(n > 0 ? n : 1);
}

hipError_t launch_radix_sort(hipStream_t s, void *temp, size_t temp_bytes, const uint32_t *keys_in,
                             uint32_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out, int64_t n,
                             int nbits, const float *vals_wt, bool counted0) {
    if (n < 0 || n >= (int64_t)1 << 31 || nbits < 1 || nbits > 32) return hipErrorInvalidValue;
    if (counted0 && nbits < kBits) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    if (temp_bytes < temp_bytes_for(n)) return hipErrorInvalidValue;
    SortTemp t = carve(temp, n);
    const int tiles = sort_tiles(n), chunks = sort_chunks(tiles), np = sort_passes(nbits);
    // ping-pong so the last pass writes keys_out/vals_out
    const uint32_t *ki = keys_in;
    const uint64_t *vi = vals_in;
    for (int p = 0; p < np; p++) {
        const bool to_out = ((np - 1 - p) % 2) == 0;
        uint32_t *ko = to_out ? keys_out : t.keys;
        uint64_t *vo = to_out ? vals_out : t.vals;
        const int shift = p * kBits, bits = std::min(kBits, nbits - shift);
        if (p == 0 && counted0) {
            // the prep pass counted this pass's digits
        } else if (reinterpret_cast<uintptr_t>(ki) % 16 == 0)
            hipLaunchKernelGGL(k_sort_count<true>, dim3(tiles), dim3(kCountThr), 0, s, ki, n, shift, bits, t.tcount);
        else
            hipLaunchKernelGGL(k_sort_count<false>, dim3(tiles), dim3(kCountThr), 0, s, ki, n, shift, bits, t.tcount);
        hipLaunchKernelGGL(k_sort_chunk, dim3(chunks), dim3(1024), 0, s, t.tcount, tiles, t.csum);
        hipLaunchKernelGGL(k_sort_chunk_scan, dim3(1), dim3(1024), 0, s, t.csum, chunks);
        if (p == 0 && vals_wt)  // the first pass forms the values
            hipLaunchKernelGGL(k_sort_scatter<true>, dim3(tiles), dim3(kThr), 0, s, ki, vi, vals_wt, ko, vo, n, shift,
                               bits, t.tcount, t.csum);
        else
            hipLaunchKernelGGL(k_sort_scatter<false>, dim3(tiles), dim3(kThr), 0, s, ki, vi, (const float *)nullptr, ko,
                               vo, n, shift, bits, t.tcount, t.csum);
        ki = ko;
        vi = vo;
    }
    return hipGetLastError();
}
