// sph_sort.hpp -- host launchers of the atomic-free radix sort and scan.
#pragma once
#include <stddef.h>
#include <hip/hip_runtime.h>

namespace sph {

// Keys per radix-sort tile (one 256-thread block, two keys per thread):
// 512-key tiles give a 262,144-key sort 512 blocks per pass, two per CU
// (2048-key tiles left half the CUs idle in the histogram and scatter
// kernels: 137 -> 126 us for the whole C2 build stage).
// (256-key tiles: 130 us.)
constexpr int RADIX_TILE_KEYS = 512;

// Device bytes of a sort's counters for n keys: 256 digit counts per tile
// and the 256 digit totals.
constexpr size_t radix_sort_temp_bytes(int n)
{
    return sizeof(unsigned) * 256 *
           ((size_t)(n > 0 ? (n + RADIX_TILE_KEYS - 1) / RADIX_TILE_KEYS : 1) + 1);
}
void radix_sort_pairs(const unsigned *keys_in, unsigned *k_out, unsigned *v_out,
                      unsigned *k_alt, unsigned *v_alt, int n, int key_bits, void *temp,
                      hipStream_t stream, const int *n_dev = nullptr);

size_t exscan_temp_bytes(int n);
void exclusive_scan_u32(const unsigned *in, unsigned *out, int n, unsigned *grand, void *temp,
                        hipStream_t stream);

}  // namespace sph
