// synth.hip -- device-side synthetic binary operators for the benchmark workloads.
//
// G[k,i,j] = scale * exp(normal(key, k + No*(i + N1*j))), where normal() is Box-Muller (cosine
// branch) of two SplitMix64 uniforms -- the same counter-based scheme as the host generator
// (dtx/synth.py), so a tensor can be generated where it is used.  Losses are the column sums
// L[i,j] = sum_k G[k,i,j], so every input pair conserves particle number.
#include "bf16_util.h"
#include "dtx_internal.h"

namespace dtx {

__device__ inline uint64_t d_splitmix(uint64_t key, uint64_t idx) {
    uint64_t z = key + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline double d_uniform(uint64_t key, uint64_t idx) {
    return ((double)(d_splitmix(key, idx) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ inline double d_lognormal(uint64_t key, uint64_t idx) {
    const double u1 = d_uniform(key, 2 * idx), u2 = d_uniform(key, 2 * idx + 1);
    return exp(sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2));
}

struct SynthArgs {
    void* A;
    const uint32_t* pairs;
    int64_t npairs_pad;
    int No, nktiles, kind, N1, N2, o1, o2, precision;
    uint64_t key;
    double scale;
};

// One thread per (packed row r of k-tile kt, pair p).
__global__ void synth_gain_kernel(SynthArgs s) {
    const int64_t total = (int64_t)s.nktiles * s.npairs_pad * TK;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % TK);
        const int64_t p = (t / TK) % s.npairs_pad;
        const int kt = (int)(t / (TK * s.npairs_pad));
        const int k = kt * TK + r;
        double v = 0.0;
        const uint32_t pr = s.pairs[p];
        if (k < s.No && pr != PAIR_PAD) {
            const int i = (int)(pr & 0xFFFFu) - s.o1, j = (int)(pr >> 16) - s.o2;
            v = s.scale * d_lognormal(s.key, (uint64_t)k + (uint64_t)s.No *
                                                 ((uint64_t)i + (uint64_t)s.N1 * j));
            if (s.kind == CH_TRI && i != j)
                v += s.scale * d_lognormal(s.key, (uint64_t)k + (uint64_t)s.No *
                                                      ((uint64_t)j + (uint64_t)s.N1 * i));
        }
        const int64_t P = s.npairs_pad;
        if (s.precision == DTX_PREC_FP64) {
            ((double*)s.A)[((int64_t)kt * P + p) * TK + r] = v;
        } else if (s.precision == DTX_PREC_FP32) {
            ((float*)s.A)[((int64_t)kt * P + p) * TK + r] = (float)v;
        } else {
            uint16_t hi, lo;
            split_bf16(v, &hi, &lo);
            uint16_t* base = (uint16_t*)s.A + (int64_t)kt * P * TK * 2 + (p / 32) * TK * 64;
            base[r * 32 + (p % 32)] = hi;
            base[TK * 32 + r * 32 + (p % 32)] = lo;
        }
    }
}

// One thread per (i, j): L = sum_k G[k,i,j]; subtract from the loss matrix D (column-major).
__global__ void synth_loss_kernel(double* D, int N, int o1, int N1, int o2, int N2, int No,
                                  uint64_t key, double scale, int loss_on_first) {
    const int64_t total = (int64_t)N1 * N2;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % N1), j = (int)(t / N1);
        double L = 0.0;
        for (int k = 0; k < No; ++k)
            L = L + scale * d_lognormal(key, (uint64_t)k + (uint64_t)No *
                                                 ((uint64_t)i + (uint64_t)N1 * j));
        if (loss_on_first) D[(int64_t)(o1 + i) + (int64_t)N * (o2 + j)] -= L;
        else D[(int64_t)(o2 + j) + (int64_t)N * (o1 + i)] -= L;
    }
}

hipError_t launch_synth_gain_packed(int precision, void* A, const uint32_t* pairs,
                                    int64_t npairs_pad, int No, int nktiles, int kind, int N1,
                                    int N2, uint64_t key, double scale, hipStream_t stream) {
    // The pair table holds cell-local state indices; recover the species offsets from its
    // first entry (pair 0 is (o1, o1) for TRI and (o1, o2) for RECT).
    uint32_t first = 0;
    hipError_t e = hipMemcpy(&first, pairs, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    SynthArgs s;
    s.A = A;
    s.pairs = pairs;
    s.npairs_pad = npairs_pad;
    s.No = No;
    s.nktiles = nktiles;
    s.kind = kind;
    s.N1 = N1;
    s.N2 = N2;
    s.o1 = (int)(first & 0xFFFFu);
    s.o2 = (int)(first >> 16);
    s.precision = precision;
    s.key = key;
    s.scale = scale;
    synth_gain_kernel<<<8192, 256, 0, stream>>>(s);
    return hipGetLastError();
}

hipError_t launch_synth_loss(double* loss, int N, int off1, int N1, int off2, int N2, int No,
                             uint64_t key, double scale, int loss_on_first,
                             hipStream_t stream) {
    const int64_t total = (int64_t)N1 * N2;
    int blocks = (int)std::min<int64_t>((total + 255) / 256, 16384);
    synth_loss_kernel<<<blocks, 256, 0, stream>>>(loss, N, off1, N1, off2, N2, No, key, scale,
                                                  loss_on_first);
    return hipGetLastError();
}

}  // namespace dtx
