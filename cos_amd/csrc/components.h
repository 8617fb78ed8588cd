// Internal launch interface of the component kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "capon_kernels.h"

namespace cosk {

hipError_t fill_i32(int* p, int n, int v, hipStream_t s);
hipError_t compact_flags(const uint8_t* flags, int n, int* list, int* count, hipStream_t s);
// the same into nb lists by the bucket bk.b[L] of each pixel's row L (rowL[j + Y] for
// output row j = pixel % nzp): list b at list + b * lstride, its length in count[b]
hipError_t compact_flags_rows(const uint8_t* flags, int n, int* list, int* count, const int* rowL, int nzp, int Y,
                              RowBuckets bk, int nb, int lstride, hipStream_t s);
hipError_t solve_lu(const float2* A, const float2* b, float2* x, int N, int batch, hipStream_t s);
hipError_t buildr_beamspace(const float2* xbs, float2* R, float d, int Nb, int K, int Y, int Nx,
                            int Nz, hipStream_t s);
// Capon in the subspace of V [Nb][L] from the conj-stored unnormalised S of every pixel.
hipError_t subspace_capon(const float2* R, const float2* V, const float2* x, int L, int Nb, int M, int Mstride,
                          int off, int Y, int Nx, int Nz, float d, float2* w, float2* z, float* P, hipStream_t s);
hipError_t beamspace_matrix(const float2* R, const float2* B, const float2* x, int L, int Nb, int M,
                            int Y, int Nx, int Nz, float2* w, float2* z, float* P, hipStream_t s);

}  // namespace cosk
