// mfma_layout.hip -- test helper: multiplies one 16x16xK tile with each MFMA builtin the
// contraction kernels use, reading and writing the operands with the lane layout the kernels
// assume (A[l%16][kgroup], B[kgroup][l%16]) and returns the raw accumulator of every lane
// (D[l*4 + e]); the host locates each value in the plain matrix product.  Launched with one wave of 64 lanes; all indexes stay inside the
// 16x16 / 16xK / Kx16 buffers.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// A: 16 x 4 row-major, B: 4 x 16 row-major, D: 16 x 16 row-major
__global__ void k_f64(const double* A, const double* B, double* D) {
    const int l = threadIdx.x;
    d4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(l % 16) * 4 + l / 16], B[(l / 16) * 16 + l % 16],
                                               acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) D[l * 4 + e] = acc[e];
}

__global__ void k_f32(const float* A, const float* B, float* D) {
    const int l = threadIdx.x;
    f4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(l % 16) * 4 + l / 16], B[(l / 16) * 16 + l % 16],
                                               acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) D[l * 4 + e] = acc[e];
}

// A: 16 x 32 row-major bf16 bits, B: 32 x 16 stored column-major (B[j*32 + k]), D: 16 x 16
__global__ void k_bf16(const uint16_t* A, const uint16_t* B, float* D) {
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) {
        a[e] = __builtin_bit_cast(__bf16, A[(l % 16) * 32 + 8 * (l / 16) + e]);
        b[e] = __builtin_bit_cast(__bf16, B[(l % 16) * 32 + 8 * (l / 16) + e]);
    }
    f4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) D[l * 4 + e] = acc[e];
}

extern "C" int mfma_probe(int which, const void* hA, const void* hB, void* hD) {
    size_t na, nb, nd;
    if (which == 0) { na = 64 * 8; nb = 64 * 8; nd = 256 * 8; }
    else if (which == 1) { na = 64 * 4; nb = 64 * 4; nd = 256 * 4; }
    else { na = 512 * 2; nb = 512 * 2; nd = 256 * 4; }
    void *dA, *dB, *dD;
    if (hipMalloc(&dA, na) || hipMalloc(&dB, nb) || hipMalloc(&dD, nd)) return -1;
    if (hipMemcpy(dA, hA, na, hipMemcpyHostToDevice) || hipMemcpy(dB, hB, nb, hipMemcpyHostToDevice))
        return -2;
    if (which == 0) k_f64<<<1, 64>>>((const double*)dA, (const double*)dB, (double*)dD);
    else if (which == 1) k_f32<<<1, 64>>>((const float*)dA, (const float*)dB, (float*)dD);
    else k_bf16<<<1, 64>>>((const uint16_t*)dA, (const uint16_t*)dB, (float*)dD);
    if (hipGetLastError() != hipSuccess) return -3;
    if (hipMemcpy(hD, dD, nd, hipMemcpyDeviceToHost)) return -4;
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dD);
    return 0;
}
