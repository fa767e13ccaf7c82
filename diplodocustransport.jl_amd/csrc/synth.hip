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
    const uint32_t* chunks;
    int64_t nchunks;
    int No, nktiles, kind, N1, N2, pad1, pad2, precision;
    uint64_t key;
    double scale;
};

// One thread per packed element (k-tile, chunk, j offset t, row r) of the chunked layout.
__global__ void synth_gain_kernel(SynthArgs s) {
    const int64_t per_tile = s.nchunks * KC * TK;
    const int64_t total = (int64_t)s.nktiles * per_tile;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int kt = (int)(e / per_tile);
        const int64_t rem = e % per_tile;
        const int64_t ch = rem / (KC * TK);
        const int t = (int)((rem / TK) % KC);
        const int r = (int)(rem % TK);
        const int k = kt * TK + r;
        const uint32_t c = s.chunks[ch];
        const int i = (int)(c & 0xFFFFu) - s.pad1;
        const int j = (int)(c >> 16) - (s.kind == CH_TRI ? s.pad1 : s.pad2) + t;
        const bool valid = (s.kind == CH_TRI) ? (j >= i && j < s.N1) : (j < s.N2);
        double v = 0.0;
        if (k < s.No && valid) {
            v = s.scale * d_lognormal(s.key, (uint64_t)k + (uint64_t)s.No *
                                                 ((uint64_t)i + (uint64_t)s.N1 * j));
            if (s.kind == CH_TRI && i != j)
                v += s.scale * d_lognormal(s.key, (uint64_t)k + (uint64_t)s.No *
                                                      ((uint64_t)j + (uint64_t)s.N1 * i));
        }
        const int64_t base = ((int64_t)kt * s.nchunks + ch) * KC * TK;
        if (s.precision == DTX_PREC_FP64) {
            ((double*)s.A)[base + t * TK + r] = v;
        } else if (s.precision == DTX_PREC_FP32) {
            ((float*)s.A)[base + t * TK + r] = (float)v;
        } else {
            uint16_t hi, lo;
            split_bf16(v, &hi, &lo);
            uint16_t* b = (uint16_t*)s.A + 2 * base;
            b[bf16_frag_index(r, t)] = hi;
            b[TK * KC + bf16_frag_index(r, t)] = lo;
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

// One thread per packed element (k-tile of rows i, chunk of j, offset t, row r) of the
// moment-rate matrices: the k-sums of the channel's pair-form gain, plain and energy-weighted.
__global__ void synth_moment_kernel(double* AN, double* AE, int64_t nchunks, int nktiles,
                                    int kind, int No, int N1, int N2, const double* ew,
                                    uint64_t key, double scale) {
    const int64_t total = (int64_t)nktiles * nchunks * KC * TK;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % TK);
        const int t = (int)((e / TK) % KC);
        const int64_t ch = (e / ((int64_t)TK * KC)) % nchunks;
        const int kt = (int)(e / ((int64_t)TK * KC * nchunks));
        const int i = kt * TK + r;
        const int j = (int)(ch * KC) + t;
        const bool valid = i < N1 && j < N2 && (kind != CH_TRI || j >= i);
        double wn = 0.0, we = 0.0;
        if (valid) {
            for (int k = 0; k < No; ++k) {
                double v = scale * d_lognormal(key, (uint64_t)k + (uint64_t)No *
                                                         ((uint64_t)i + (uint64_t)N1 * j));
                if (kind == CH_TRI && i != j)
                    v += scale * d_lognormal(key, (uint64_t)k + (uint64_t)No *
                                                      ((uint64_t)j + (uint64_t)N1 * i));
                wn = wn + v;
                if (ew) we = we + ew[k] * v;
            }
        }
        AN[e] = wn;
        if (AE) AE[e] = we;
    }
}

hipError_t launch_synth_moment_packed(double* AN, double* AE, int64_t nchunks, int nktiles,
                                      int kind, int No, int N1, int N2, const double* ew,
                                      uint64_t key, double scale, hipStream_t stream) {
    const int64_t total = (int64_t)nktiles * nchunks * KC * TK;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 16384);
    synth_moment_kernel<<<blocks, 256, 0, stream>>>(AN, AE, nchunks, nktiles, kind, No, N1, N2,
                                                    ew, key, scale);
    return hipGetLastError();
}

hipError_t launch_synth_gain_packed(int precision, void* A, const uint32_t* chunks,
                                    int64_t nchunks, int No, int nktiles, int kind, int N1,
                                    int N2, int pad1, int pad2, uint64_t key, double scale,
                                    hipStream_t stream) {
    SynthArgs s;
    s.A = A;
    s.chunks = chunks;
    s.nchunks = nchunks;
    s.No = No;
    s.nktiles = nktiles;
    s.kind = kind;
    s.N1 = N1;
    s.N2 = N2;
    s.pad1 = pad1;
    s.pad2 = pad2;
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
