// mfma_layout.hip -- test helpers for the lane layouts the kernels rely on.
// mfma_probe: multiplies one 16x16xK tile with each MFMA builtin the contraction kernels use,
// reading and writing the operands with the lane layout the kernels assume (A[l%16][kgroup],
// B[kgroup][l%16]) and returns the raw accumulator of every lane (D[l*4 + e]); the host locates
// each value in the plain matrix product.
// lane_probe: the DPP / permlane lane exchanges of wave_ops.h (xor distances 1..32, quad
// broadcasts) applied to one value per lane, for the host to compare with the permutation.
// Both launch one wave of 64 lanes; all indexes stay inside their buffers.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../diplodocustransport.jl_amd/csrc/wave_ops.h"

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

// out[r * 64 + l], r = 0..5: xor_lane<1, 2, 4, 8, 16, 32>(in[l]); r = 6..9: quad_lane<0..3>;
// r = 10..15: the same xor distances on the low 32 bits as int (through xor_lane_rt);
// r = 16..19: quad_lane<0..3> on int.
__global__ void k_lanes(const double* in, double* out) {
    const int l = threadIdx.x;
    const double v = in[l];
    out[0 * 64 + l] = dtx::xor_lane<1>(v);
    out[1 * 64 + l] = dtx::xor_lane<2>(v);
    out[2 * 64 + l] = dtx::xor_lane<4>(v);
    out[3 * 64 + l] = dtx::xor_lane<8>(v);
    out[4 * 64 + l] = dtx::xor_lane<16>(v);
    out[5 * 64 + l] = dtx::xor_lane<32>(v);
    out[6 * 64 + l] = dtx::quad_lane<0>(v);
    out[7 * 64 + l] = dtx::quad_lane<1>(v);
    out[8 * 64 + l] = dtx::quad_lane<2>(v);
    out[9 * 64 + l] = dtx::quad_lane<3>(v);
    const unsigned long long u = (unsigned long long)l * 0x9E3779B97F4A7C15ull;
#pragma unroll
    for (int i = 0; i < 6; ++i)
        out[(10 + i) * 64 + l] = (double)(dtx::xor_lane_rt(u, 1 << i) >> 40);
    const int iv = 1000 + 7 * l;
    out[16 * 64 + l] = (double)dtx::quad_lane<0>(iv);
    out[17 * 64 + l] = (double)dtx::quad_lane<1>(iv);
    out[18 * 64 + l] = (double)dtx::quad_lane<2>(iv);
    out[19 * 64 + l] = (double)dtx::quad_lane<3>(iv);
}

extern "C" int lane_probe(const double* h_in, double* h_out) {
    double *d_in, *d_out;
    if (hipMalloc(&d_in, 64 * 8) || hipMalloc(&d_out, 20 * 64 * 8)) return -1;
    if (hipMemcpy(d_in, h_in, 64 * 8, hipMemcpyHostToDevice)) return -2;
    k_lanes<<<1, 64>>>(d_in, d_out);
    if (hipGetLastError() != hipSuccess) return -3;
    if (hipMemcpy(h_out, d_out, 20 * 64 * 8, hipMemcpyDeviceToHost)) return -4;
    (void)hipFree(d_in); (void)hipFree(d_out);
    return 0;
}
