// Stable LSD radix sort of (uint32 key, uint32 value) pairs for gfx950, with 8-, 9- or 10-bit
// digits (the narrowest width with the fewest passes: keys of 17-18 bits sort in 2 passes of
// 9 bits, 19-20 bits in 2 of 10, 25-27 bits in 3 of 9, 28-30 bits in 3 of 10).
//
// Per pass (digit d = (key >> shift) & (NB - 1), NB = 2^digit bits):
//   k_rs_hist    : one workgroup per tile of RS_TILE keys builds an NB-bin LDS histogram
//                  and stores it digit-major: hist[d * n_tiles + tile].
//   k_rs_scan    : one workgroup per digit turns its row into an exclusive prefix over
//                  tiles and stores the digit total.
//   k_rs_scatter : each tile ranks its keys stably (wave ballots give the rank among equal
//                  digits inside a wave, LDS counters carry it across waves and rounds),
//                  regroups the tile by digit in LDS and writes each digit run to
//                  global = digit_prefix[d] + hist[d][tile] + local rank, so consecutive
//                  lanes write consecutive addresses.
// Stability: items are ranked in index order (round r covers tile items r*256 .. r*256+255,
// lane order inside a round), hence equal keys keep their input order.

#include "common.h"

namespace nsph {

constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_WAVES = RS_BLOCK / 64;

// Exclusive scan of one value per thread over a 256-thread block. Returns the exclusive
// prefix; *total receives the block sum. Uses `tmp` (>= RS_WAVES entries) in LDS.
__device__ inline uint32_t block_excl_scan_256(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    uint32_t wave_base = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
        uint32_t s = tmp[w];
        if (w < wave) wave_base += s;
        sum += s;
    }
    __syncthreads();
    *total = sum;
    return wave_base + x - v;
}

// Exclusive scan over NB bins (NB / 256 consecutive bins per thread): out[b] = sum in[< b].
// in and out may be the same array; get(b) reads bin b.
template <int NB, class Get, class Put>
__device__ inline void bins_excl_scan(Get get, Put put, uint32_t* tmp) {
    constexpr int PER = NB / RS_BLOCK;
    const int t = threadIdx.x;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        v[i] = get(t * PER + i);
        sum += v[i];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan_256(sum, tmp, &tot);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        put(t * PER + i, run);
        run += v[i];
    }
}

template <int DB>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint32_t* __restrict__ keys, int n,
                                                      int shift, uint32_t* __restrict__ hist,
                                                      int n_tiles, const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    constexpr int NB = 1 << DB;
    __shared__ uint32_t h[NB];
    const int t = threadIdx.x;
    for (int b = t; b < NB; b += RS_BLOCK) h[b] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        int idx = base + i * RS_BLOCK + t;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (uint32_t)(NB - 1)], 1u);
    }
    __syncthreads();
    for (int b = t; b < NB; b += RS_BLOCK) hist[(size_t)b * n_tiles + blockIdx.x] = h[b];
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scan(uint32_t* __restrict__ hist, int n_tiles,
                                                      uint32_t* __restrict__ digit_tot,
                                                      const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    __shared__ uint32_t tmp[RS_WAVES];
    uint32_t* row = hist + (size_t)blockIdx.x * n_tiles;
    uint32_t carry = 0;
    for (int base = 0; base < n_tiles; base += RS_BLOCK) {
        int i = base + threadIdx.x;
        uint32_t v = i < n_tiles ? row[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan_256(v, tmp, &tot);
        if (i < n_tiles) row[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) digit_tot[blockIdx.x] = carry;
}

template <int DB>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ digit_tot, int n_tiles,
    const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    constexpr int NB = 1 << DB;
    constexpr uint32_t MASK = (uint32_t)(NB - 1);
    __shared__ uint32_t s_gbase[NB];   // global start of this tile's run of digit d
    __shared__ uint32_t s_lstart[NB];  // start of digit d inside the regrouped tile
    __shared__ uint32_t s_run[NB];     // items of digit d ranked so far
    __shared__ uint32_t s_wcnt[RS_WAVES][NB];
    __shared__ uint32_t s_keys[RS_TILE];
    __shared__ uint32_t s_vals[RS_TILE];
    __shared__ uint32_t tmp[RS_WAVES];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int base = blockIdx.x * RS_TILE;
    const int n_valid = min(RS_TILE, n - base);

    bins_excl_scan<NB>([&](int b) { return digit_tot[b]; },
                       [&](int b, uint32_t pre) { s_gbase[b] = pre + hist[(size_t)b * n_tiles + blockIdx.x]; },
                       tmp);
    for (int b = t; b < NB; b += RS_BLOCK) {
        s_run[b] = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) s_wcnt[w][b] = 0;
    }

    uint32_t key[RS_ITEMS], val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        int idx = base + i * RS_BLOCK + t;
        bool ok = idx < n;
        key[i] = ok ? kin[idx] : 0xFFFFFFFFu;  // invalid items rank last in the top digit
        val[i] = ok ? vin[idx] : 0u;
    }
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const uint32_t d = (key[i] >> shift) & MASK;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const int bit = (d >> b) & 1;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t r_wave = (uint32_t)__popcll(peers & lt_mask);
        if (r_wave == 0) s_wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t r = s_run[d] + r_wave;
        for (int w = 0; w < wave; ++w) r += s_wcnt[w][d];
        rank[i] = r;
        __syncthreads();
        for (int b = t; b < NB; b += RS_BLOCK) {
            uint32_t add = 0;
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) {
                add += s_wcnt[w][b];
                s_wcnt[w][b] = 0;
            }
            s_run[b] += add;
        }
        __syncthreads();
    }

    bins_excl_scan<NB>([&](int b) { return s_run[b]; }, [&](int b, uint32_t pre) { s_lstart[b] = pre; }, tmp);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const uint32_t d = (key[i] >> shift) & MASK;
        const uint32_t pos = s_lstart[d] + rank[i];
        s_keys[pos] = key[i];
        s_vals[pos] = val[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const int p = i * RS_BLOCK + t;
        if (p < n_valid) {
            const uint32_t k = s_keys[p];
            const uint32_t d = (k >> shift) & MASK;
            const uint32_t g = s_gbase[d] + (uint32_t)p - s_lstart[d];
            kout[g] = k;
            vout[g] = s_vals[p];
        }
    }
}

int radix_reserve(RadixSort& rs, int n) {
    if (n <= rs.capacity) return NSPH_OK;
    const int n_tiles = (n + RS_TILE - 1) / RS_TILE;
    NSPH_TRY(buf_ensure(rs.keys_alt, (size_t)n * 4));
    NSPH_TRY(buf_ensure(rs.vals_alt, (size_t)n * 4));
    NSPH_TRY(buf_ensure(rs.hist, (size_t)n_tiles * 1024 * 4));
    NSPH_TRY(buf_ensure(rs.digit_tot, 1024 * 4));
    rs.capacity = n;
    return NSPH_OK;
}

// Passes over `bits` key bits and the digit width: the narrowest of 8, 9 and 10 bits that
// needs the fewest passes (a wider digit costs more per pass: more bins to scan and clear).
static int rs_passes(int bits, int* digit_bits) {
    if (bits <= 0) {
        *digit_bits = 8;
        return 0;
    }
    const int p8 = (bits + 7) / 8, p9 = (bits + 8) / 9, p10 = (bits + 9) / 10;
    if (p8 <= p10) {
        *digit_bits = 8;
        return p8;
    }
    *digit_bits = p9 <= p10 ? 9 : 10;
    return p10;
}

void radix_result(const RadixSort& rs, uint32_t* keys, uint32_t* vals, int n, int bits, uint32_t** out_keys,
                  uint32_t** out_vals) {
    int db = 8;
    const bool alt = n > 0 && (rs_passes(bits, &db) & 1) != 0;
    *out_keys = alt ? bp<uint32_t>(rs.keys_alt) : keys;
    *out_vals = alt ? bp<uint32_t>(rs.vals_alt) : vals;
}

int radix_sort(RadixSort& rs, uint32_t* keys, uint32_t* vals, int n, int bits,
               hipStream_t stream, uint32_t** out_keys, uint32_t** out_vals, const int* guard) {
    NSPH_CHECK(n <= rs.capacity, NSPH_ERR_CAPACITY, "radix sort: n=%d above capacity %d", n,
               rs.capacity);
    uint32_t *ka = keys, *va = vals, *kb = bp<uint32_t>(rs.keys_alt), *vb = bp<uint32_t>(rs.vals_alt);
    if (n <= 0) {
        *out_keys = ka;
        *out_vals = va;
        return NSPH_OK;
    }
    const int n_tiles = (n + RS_TILE - 1) / RS_TILE;
    int db = 8;
    const int passes = rs_passes(bits, &db);
    for (int p = 0; p < passes; ++p) {
        const int shift = db * p;
        if (db == 10) {
            hipLaunchKernelGGL(k_rs_hist<10>, dim3(n_tiles), dim3(RS_BLOCK), 0, stream, ka, n, shift,
                               bp<uint32_t>(rs.hist), n_tiles, guard);
            hipLaunchKernelGGL(k_rs_scan, dim3(1024), dim3(RS_BLOCK), 0, stream,
                               bp<uint32_t>(rs.hist), n_tiles, bp<uint32_t>(rs.digit_tot), guard);
            hipLaunchKernelGGL(k_rs_scatter<10>, dim3(n_tiles), dim3(RS_BLOCK), 0, stream, ka, va, kb,
                               vb, n, shift, bp<uint32_t>(rs.hist), bp<uint32_t>(rs.digit_tot), n_tiles, guard);
        } else if (db == 9) {
            hipLaunchKernelGGL(k_rs_hist<9>, dim3(n_tiles), dim3(RS_BLOCK), 0, stream, ka, n, shift,
                               bp<uint32_t>(rs.hist), n_tiles, guard);
            hipLaunchKernelGGL(k_rs_scan, dim3(512), dim3(RS_BLOCK), 0, stream,
                               bp<uint32_t>(rs.hist), n_tiles, bp<uint32_t>(rs.digit_tot), guard);
            hipLaunchKernelGGL(k_rs_scatter<9>, dim3(n_tiles), dim3(RS_BLOCK), 0, stream, ka, va, kb,
                               vb, n, shift, bp<uint32_t>(rs.hist), bp<uint32_t>(rs.digit_tot), n_tiles, guard);
        } else {
            hipLaunchKernelGGL(k_rs_hist<8>, dim3(n_tiles), dim3(RS_BLOCK), 0, stream, ka, n, shift,
                               bp<uint32_t>(rs.hist), n_tiles, guard);
            hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(RS_BLOCK), 0, stream,
                               bp<uint32_t>(rs.hist), n_tiles, bp<uint32_t>(rs.digit_tot), guard);
            hipLaunchKernelGGL(k_rs_scatter<8>, dim3(n_tiles), dim3(RS_BLOCK), 0, stream, ka, va, kb,
                               vb, n, shift, bp<uint32_t>(rs.hist), bp<uint32_t>(rs.digit_tot), n_tiles, guard);
        }
        NSPH_HIP(hipGetLastError());
        uint32_t* t0 = ka; ka = kb; kb = t0;
        uint32_t* t1 = va; va = vb; vb = t1;
    }
    *out_keys = ka;
    *out_vals = va;
    return NSPH_OK;
}

void radix_free(RadixSort& rs) {
    buf_free(rs.keys_alt);
    buf_free(rs.vals_alt);
    buf_free(rs.hist);
    buf_free(rs.digit_tot);
    rs.capacity = 0;
}

int bits_for(uint32_t max_key) {
    int b = 0;
    while (b < 32 && (max_key >> b) != 0u) ++b;
    return b;
}

}  // namespace nsph
