// sph_sort.hip -- atomic-free stable LSD radix sort of (key, value) uint32
// pairs and an exclusive scan, for the periodic cell build.
//
// Digit counting and stable ranking use wavefront ballots: for an 8-bit
// digit, eight __ballot calls give each lane the mask of lanes holding the
// same digit (a wave-level "match"); popcount below the lane gives its rank
// and the lowest lane of each group owns that digit's count.  No atomics,
// global or LDS, so the result is deterministic and the order stable.
#include "sph_sort.hpp"

namespace sph {

static constexpr int RS_THREADS = 256;           // 4 waves of 64
static constexpr int RS_WAVES = RS_THREADS / 64;
static constexpr int RS_ITEMS = RADIX_TILE_KEYS / RS_THREADS;  // items per thread per tile
static constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
static_assert(RS_TILE == RADIX_TILE_KEYS, "a tile is whole items per thread");
static constexpr int RS_BINS = 256;

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const unsigned lane = threadIdx.x & 63u;
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Mask of lanes whose 8-bit digit equals this lane's digit (valid lanes only).
// Every lane of the wave must call it (ballots are wave-wide).
__device__ __forceinline__ uint64_t wave_match8(unsigned d, bool valid)
{
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Exclusive scan of one value per thread over a 256-thread block; returns the
// prefix and writes the block total to *total.  Shuffle scans within the
// four waves, then each thread adds the totals of the waves below it (two
// barriers; an LDS Hillis-Steele scan took sixteen).
__device__ unsigned block_exscan256(unsigned x, unsigned *lds /*[256]*/, unsigned *total)
{
    static_assert(RS_THREADS == 256, "four waves");
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    unsigned below = 0;
    for (int ww = 0; ww < w; ++ww) below += lds[ww];
    *total = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();  // the caller may reuse lds
    return below + inc - x;
}

// hist[d * nblocks + blk] = count of digit d in tile blk.
// n_dev (may be null): element count on the device, at most n.
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const unsigned *__restrict__ keys, int n,
                                                         const int *__restrict__ n_dev, int shift,
                                                         unsigned *__restrict__ hist, int nblocks)
{
    __shared__ unsigned wc[RS_WAVES][RS_BINS];
    const int tid = threadIdx.x, w = tid >> 6;
    if (n_dev) n = min(n, *n_dev);
    if ((int)blockIdx.x * RS_TILE >= n) {  // empty tile
        for (int d = tid; d < RS_BINS; d += RS_THREADS) hist[(size_t)d * nblocks + blockIdx.x] = 0;
        return;
    }
    for (int i = tid; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&wc[0][0])[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int idx = base + it * RS_THREADS + tid;
        const bool valid = idx < n;
        const unsigned d = valid ? (keys[idx] >> shift) & 0xffu : 0u;
        const uint64_t m = wave_match8(d, valid);
        if (valid && (m & lanemask_lt()) == 0) wc[w][d] += (unsigned)__popcll(m);
    }
    __syncthreads();
    for (int d = tid; d < RS_BINS; d += RS_THREADS) {
        unsigned s = 0;
        for (int ww = 0; ww < RS_WAVES; ++ww) s += wc[ww][d];
        hist[(size_t)d * nblocks + blockIdx.x] = s;
    }
}

// One block per digit: exclusive scan of hist[d][0..nblocks) in place and
// digit total into tot[d].
__global__ __launch_bounds__(RS_THREADS) void k_rs_scan_digits(unsigned *__restrict__ hist,
                                                                int nblocks,
                                                                unsigned *__restrict__ tot)
{
    __shared__ unsigned lds[256];
    unsigned *row = hist + (size_t)blockIdx.x * nblocks;
    const int tid = threadIdx.x;
    const int per = (nblocks + RS_THREADS - 1) / RS_THREADS;
    const int b0 = tid * per;
    unsigned s = 0;
    for (int i = 0; i < per; ++i)
        if (b0 + i < nblocks) s += row[b0 + i];
    unsigned total;
    unsigned run = block_exscan256(s, lds, &total);
    for (int i = 0; i < per; ++i) {
        if (b0 + i < nblocks) {
            unsigned x = row[b0 + i];
            row[b0 + i] = run;
            run += x;
        }
    }
    if (tid == 0) tot[blockIdx.x] = total;
}

// Stable scatter of one tile.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const unsigned *__restrict__ kin, const unsigned *__restrict__ vin,
    unsigned *__restrict__ kout, unsigned *__restrict__ vout, int n, const int *__restrict__ n_dev,
    int shift, const unsigned *__restrict__ hist, const unsigned *__restrict__ tot, int nblocks,
    int identity_vals)
{
    __shared__ unsigned base[RS_BINS];
    __shared__ unsigned wc[RS_WAVES][RS_BINS];
    __shared__ unsigned lds[256];
    const int tid = threadIdx.x, w = tid >> 6;
    if (n_dev) n = min(n, *n_dev);
    if ((int)blockIdx.x * RS_TILE >= n) return;  // empty tile (uniform over the block)
    {
        unsigned total;
        const unsigned pre = block_exscan256(tot[tid], lds, &total);
        base[tid] = pre + hist[(size_t)tid * nblocks + blockIdx.x];
        for (int ww = 0; ww < RS_WAVES; ++ww) wc[ww][tid] = 0;
    }
    __syncthreads();
    const int tile = blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int idx = tile + it * RS_THREADS + tid;
        const bool valid = idx < n;
        const unsigned key = valid ? kin[idx] : 0u;
        const unsigned d = (key >> shift) & 0xffu;
        const uint64_t m = wave_match8(d, valid);
        const unsigned rank = (unsigned)__popcll(m & lanemask_lt());
        if (valid && rank == 0) wc[w][d] = (unsigned)__popcll(m);
        __syncthreads();
        if (valid) {
            unsigned pos = base[d] + rank;
            for (int ww = 0; ww < w; ++ww) pos += wc[ww][d];
            kout[pos] = key;
            vout[pos] = identity_vals ? (unsigned)idx : vin[idx];
        }
        __syncthreads();
        {
            unsigned s = 0;
            for (int ww = 0; ww < RS_WAVES; ++ww) {
                s += wc[ww][tid];
                wc[ww][tid] = 0;
            }
            base[tid] += s;
        }
        __syncthreads();
    }
}

// Sorts (keys_in, implicit identity values) by the low key_bits bits.
// Result lands in (k_out, v_out); k_alt/v_alt are ping-pong buffers.  The
// input keys are left untouched.  temp holds radix_sort_temp_bytes(n).
// With n_dev the element count is read on the device (min(n, *n_dev)): the
// grid covers n and tiles past the count exit at once.
void radix_sort_pairs(const unsigned *keys_in, unsigned *k_out, unsigned *v_out,
                      unsigned *k_alt, unsigned *v_alt, int n, int key_bits, void *temp,
                      hipStream_t stream, const int *n_dev)
{
    if (n <= 0) return;
    const int nb = (n + RS_TILE - 1) / RS_TILE;
    unsigned *hist = (unsigned *)temp;
    unsigned *tot = hist + (size_t)RS_BINS * nb;
    int passes = (key_bits + 7) / 8;
    if (passes < 1) passes = 1;
    // choose the ping-pong order so that the last pass writes k_out / v_out
    const unsigned *kin = keys_in;
    const unsigned *vin = nullptr;
    for (int p = 0; p < passes; ++p) {
        const bool last_to_out = ((passes - 1 - p) % 2) == 0;
        unsigned *ko = last_to_out ? k_out : k_alt;
        unsigned *vo = last_to_out ? v_out : v_alt;
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_THREADS), 0, stream, kin, n, n_dev, shift, hist, nb);
        hipLaunchKernelGGL(k_rs_scan_digits, dim3(RS_BINS), dim3(RS_THREADS), 0, stream, hist, nb, tot);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_THREADS), 0, stream, kin, vin, ko, vo, n,
                           n_dev, shift, hist, tot, nb, p == 0 ? 1 : 0);
        kin = ko;
        vin = vo;
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of uint32 (three phases: tile sums, scan of sums, apply).
static constexpr int SC_THREADS = 256;
static constexpr int SC_ITEMS = 8;
static constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

__global__ __launch_bounds__(SC_THREADS) void k_scan_tiles(const unsigned *__restrict__ in, int n,
                                                            unsigned *__restrict__ sums)
{
    __shared__ unsigned lds[256];
    const int base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
    unsigned s = 0;
    for (int i = 0; i < SC_ITEMS; ++i)
        if (base + i < n) s += in[base + i];
    unsigned total;
    block_exscan256(s, lds, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_top(unsigned *__restrict__ sums, int nt,
                                                          unsigned *__restrict__ grand)
{
    __shared__ unsigned lds[256];
    const int per = (nt + SC_THREADS - 1) / SC_THREADS;
    const int b0 = threadIdx.x * per;
    unsigned s = 0;
    for (int i = 0; i < per; ++i)
        if (b0 + i < nt) s += sums[b0 + i];
    unsigned total;
    unsigned run = block_exscan256(s, lds, &total);
    for (int i = 0; i < per; ++i) {
        if (b0 + i < nt) {
            unsigned x = sums[b0 + i];
            sums[b0 + i] = run;
            run += x;
        }
    }
    if (threadIdx.x == 0 && grand) *grand = total;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_apply(const unsigned *__restrict__ in, int n,
                                                            const unsigned *__restrict__ sums,
                                                            unsigned *__restrict__ out)
{
    __shared__ unsigned lds[256];
    const int base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
    unsigned v[SC_ITEMS];
    unsigned s = 0;
    for (int i = 0; i < SC_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    unsigned total;
    unsigned run = block_exscan256(s, lds, &total) + sums[blockIdx.x];
    for (int i = 0; i < SC_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
}

size_t exscan_temp_bytes(int n)
{
    const int nt = (n + SC_TILE - 1) / SC_TILE;
    return (size_t)(nt > 0 ? nt : 1) * sizeof(unsigned);
}

void exclusive_scan_u32(const unsigned *in, unsigned *out, int n, unsigned *grand, void *temp,
                        hipStream_t stream)
{
    if (n <= 0) return;
    const int nt = (n + SC_TILE - 1) / SC_TILE;
    unsigned *sums = (unsigned *)temp;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SC_THREADS), 0, stream, in, n, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SC_THREADS), 0, stream, sums, nt, grand);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(SC_THREADS), 0, stream, in, n, sums, out);
}

}  // namespace sph
