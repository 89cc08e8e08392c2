// sph_sort.hpp -- host launchers of the atomic-free radix sort and scan.
#pragma once
#include <stddef.h>
#include <hip/hip_runtime.h>

namespace sph {

size_t radix_sort_temp_bytes(int n);
void radix_sort_pairs(const unsigned *keys_in, unsigned *k_out, unsigned *v_out,
                      unsigned *k_alt, unsigned *v_alt, int n, int key_bits, void *temp,
                      hipStream_t stream, const int *n_dev = nullptr);

size_t exscan_temp_bytes(int n);
void exclusive_scan_u32(const unsigned *in, unsigned *out, int n, unsigned *grand, void *temp,
                        hipStream_t stream);

}  // namespace sph
