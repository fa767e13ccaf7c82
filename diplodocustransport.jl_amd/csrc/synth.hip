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

// ------------------------------------------------------------- synthetic flux stencils -----
// The rows of the space and momentum flux matrices of the synthetic Cartesian problems, one
// thread per state row r = c N + k.  Each row's entries are formed in the order
// BuildFluxMatrices pushes them (per advection axis: (r, r+), (r, r), (r, r), (r, r-); then the
// synchrotron (r, r+1), (r, r), (r, r), (r, r-1)), duplicates summed from 0.0 in that order as
// sparse() combines them (src/Fluxes/BuildFluxMatrices.jl:119-129), F = X + P summed X first
// (src/SteppingStructs.jl:60), explicit zeros kept, and the row's columns sorted ascending as
// a CSC -> ELL conversion leaves them.  Expressions keep the operation order of the host
// generator (no contraction).

struct FluxRow {
    int64_t col[FLUX_MAX_ROW];
    double val[FLUX_MAX_ROW];
    int cnt;
};

__device__ inline void flux_push(FluxRow& R, int64_t col, double v) {
    for (int e = 0; e < R.cnt; ++e)
        if (R.col[e] == col) {
            R.val[e] = R.val[e] + v;
            return;
        }
    R.col[R.cnt] = col;
    R.val[R.cnt] = 0.0 + v;
    ++R.cnt;
}

// x sqrt(1 - x^2) + asin(x): twice the integral of sqrt(1 - u^2) (the perpendicular velocity)
__device__ inline double perp_primitive(double u) { return u * sqrt(1.0 - u * u) + asin(u); }

// g(u) = 3u - u^3 of the synchrotron force's angular integral
// (src/Fluxes/FluxFunctionsEmissionForces.jl:46-70)
__device__ inline double sync_angular(double u) { return 3.0 * u - pow(u, 3.0); }

// The row r of the matrix a.which into R (and of X into X, P into P for the F merge).
__device__ void flux_row(const FluxSynthArgs& a, int64_t r, FluxRow& R) {
    const int c = (int)(r / a.N), k = (int)(r % a.N);
    int s = 0;
    while (s + 1 < a.S && k >= a.sp_off[s + 1]) ++s;
    const int t = k - a.sp_off[s];
    const int ip = t % a.px[s], iu = (t / a.px[s]) % a.py[s], ih = t / (a.px[s] * a.py[s]);
    const double* pr = a.pr + a.pr_off[s];
    const double* ur = a.ur + a.ur_off[s];
    const double* hr = a.hr + a.hr_off[s];
    const int yz = a.y_num * a.z_num;
    const int cx = c / yz, cy = (c / a.z_num) % a.y_num, cz = c % a.z_num;
    const double dx = fabs(a.xr[cx + 1] - a.xr[cx]), dy = fabs(a.yr[cy + 1] - a.yr[cy]),
                 dz = fabs(a.zr[cz + 1] - a.zr[cz]);
    const double m = a.mass[s];
    const double p0 = pr[ip], p1 = pr[ip + 1], u0 = ur[iu], u1 = ur[iu + 1], h0 = hr[ih],
                 h1 = hr[ih + 1];
    const double du = u1 - u0, dh = h1 - h0;
    FluxRow X, P;
    X.cnt = 0;
    P.cnt = 0;
    if (a.which != DTX_FLUX_P) {
        // upwind advection (the B / C stencils, src/Fluxes/BuildFluxMatrices.jl:708-974) with
        // the bin-averaged velocity component (E1 - E0)/dp times the angular average
        const double E0 = sqrt(m * m + p0 * p0), E1 = sqrt(m * m + p1 * p1);
        const double vbar = (E1 - E0) / (p1 - p0);
        for (int ax = 0; ax < 2; ++ax) {
            const int num = ax == 0 ? a.x_num : a.y_num;
            if (!((a.axes >> ax) & 1) || num == 1) continue;
            double vel, trans;
            int ci;
            if (ax == 0) {
                vel = vbar * (u1 * u1 - u0 * u0) / (2.0 * du);
                trans = dy * dz;
                ci = cx;
            } else {
                vel = vbar * (perp_primitive(u1) - perp_primitive(u0)) / (2.0 * du) *
                      (cos(h0) - cos(h1)) / dh;
                trans = dx * dz;
                ci = cy;
            }
            int cp = ci + 1, cm = ci - 1;
            if (a.periodic) {
                if (cp >= num) cp = 0;
                if (cm < 0) cm = num - 1;
            } else {
                cp = cp < num - 1 ? cp : num - 1;
                cm = cm > 0 ? cm : 0;
            }
            const int cellp = ax == 0 ? cp * yz + cy * a.z_num + cz : cx * yz + cp * a.z_num + cz;
            const int cellm = ax == 0 ? cm * yz + cy * a.z_num + cz : cx * yz + cm * a.z_num + cz;
            const int64_t bp = (int64_t)cellp * a.N + k, bm = (int64_t)cellm * a.N + k;
            const double Bp = a.dt0 * trans * vel, Bm = -Bp;
            const double hpr = Bp > 0.0 ? 0.0 : 1.0, hpl = 1.0 - hpr;
            const double hmr = Bm > 0.0 ? 1.0 : 0.0, hml = 1.0 - hmr;
            if (bp != r) {
                flux_push(X, bp, Bp * hpr);
                flux_push(X, r, Bp * hpl);
            }
            if (bm != r) {
                flux_push(X, r, Bm * hmr);
                flux_push(X, bm, Bm * hml);
            }
        }
    }
    if (a.which != DTX_FLUX_X && a.sync_K > 0.0) {
        // synchrotron radiation reaction along p (the I stencil, :275-413), closed p boundaries
        const double ang = sync_angular(u1) - sync_angular(u0);
        const double vol = a.dt0 * dx * dy * dz;
        auto norm_of = [&](int q) { return (pr[q + 1] - pr[q]) * du * dh; };
        double Ip = 0.0, Im = -0.0;
        if (m != 0.0 && a.charge[s] != 0.0) {
            auto force = [&](double pp) {
                return -a.sync_K / (3.0 * m) * pp * sqrt(m * m + pp * pp) * ang * dh;
            };
            Ip = vol * force(p1);
            Im = -(vol * force(p0));
        }
        const double hpr = Ip > 0.0 ? 0.0 : 1.0, hpl = 1.0 - hpr;
        const double hmr = Im > 0.0 ? 1.0 : 0.0, hml = 1.0 - hmr;
        const double norm = norm_of(ip);
        if (ip < a.px[s] - 1) {
            flux_push(P, r + 1, Ip * hpr / norm_of(ip + 1));
            flux_push(P, r, Ip * hpl / norm);
        }
        if (ip > 0) {
            flux_push(P, r, Im * hmr / norm);
            flux_push(P, r - 1, Im * hml / norm_of(ip - 1));
        }
        // the J stencil along u (:420-557), pushed after I as BuildFluxMatrices pushes them
        // (:117-120): J(u) = K (1/2) m u (u^2 - 1) (h0 - h1) (2 asinh(p1/m) - 2 asinh(p0/m))
        // dt dx dy dz (src/Fluxes/FluxFunctionsEmissionForces.jl:84-150), closed u boundaries
        double Jp = 0.0, Jm = -0.0;
        if (m != 0.0 && a.charge[s] != 0.0) {
            const double ash = 2.0 * asinh(p1 / m) - 2.0 * asinh(p0 / m);
            auto jforce = [&](double uu) {
                return a.sync_K * (0.5 * m * uu * (uu * uu - 1.0) * (h0 - h1) * ash);
            };
            Jp = vol * jforce(u1);
            Jm = -(vol * jforce(u0));
        }
        const double jpr = Jp > 0.0 ? 0.0 : 1.0, jpl = 1.0 - jpr;
        const double jmr = Jm > 0.0 ? 1.0 : 0.0, jml = 1.0 - jmr;
        auto norm_u = [&](int q) { return (p1 - p0) * (ur[q + 1] - ur[q]) * dh; };
        const int px = a.px[s];
        if (iu < a.py[s] - 1) {
            flux_push(P, r + px, Jp * jpr / norm_u(iu + 1));
            flux_push(P, r, Jp * jpl / norm);
        }
        if (iu > 0) {
            flux_push(P, r, Jm * jmr / norm);
            flux_push(P, r - px, Jm * jml / norm_u(iu - 1));
        }
    }
    R.cnt = 0;
    for (int e = 0; e < X.cnt; ++e) flux_push(R, X.col[e], X.val[e]);
    for (int e = 0; e < P.cnt; ++e) flux_push(R, P.col[e], P.val[e]);
    for (int e = 1; e < R.cnt; ++e)  // insertion sort by column
        for (int q = e; q > 0 && R.col[q - 1] > R.col[q]; --q) {
            const int64_t tc = R.col[q];
            R.col[q] = R.col[q - 1];
            R.col[q - 1] = tc;
            const double tv = R.val[q];
            R.val[q] = R.val[q - 1];
            R.val[q - 1] = tv;
        }
}

// Count pass (a.width == 0): the longest row into stats[0] and the entry count into stats[1];
// with a.inv_ap set, also 1/Ap = 1/(dx dy dz) of every row and Vol = dt0 dx dy dz of every
// cell.  Fill pass: the rows into the [width][n] ELL arrays, padded with column -1 and 0.0.
__global__ __launch_bounds__(256) void synth_flux_kernel(FluxSynthArgs a) {
    unsigned long long longest = 0, entries = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.n;
         r += (int64_t)gridDim.x * blockDim.x) {
        FluxRow R;
        flux_row(a, r, R);
        if (a.width == 0) {
            longest = longest > (unsigned long long)R.cnt ? longest : (unsigned long long)R.cnt;
            entries += (unsigned long long)R.cnt;
        } else {
            for (int e = 0; e < a.width; ++e) {
                a.col[(int64_t)e * a.n + r] = e < R.cnt ? (int32_t)R.col[e] : -1;
                a.val[(int64_t)e * a.n + r] = e < R.cnt ? R.val[e] : 0.0;
            }
        }
        if (a.inv_ap) {
            const int c = (int)(r / a.N);
            const int yz = a.y_num * a.z_num;
            const int cx = c / yz, cy = (c / a.z_num) % a.y_num, cz = c % a.z_num;
            const double v = fabs(a.xr[cx + 1] - a.xr[cx]) * fabs(a.yr[cy + 1] - a.yr[cy]) *
                             fabs(a.zr[cz + 1] - a.zr[cz]);
            a.inv_ap[r] = 1.0 / v;
            if (r % a.N == 0) a.vol[c] = a.dt0 * v;
        }
    }
    if (a.width == 0) {
        for (int m = 1; m < 64; m <<= 1) {
            const unsigned long long o = __shfl_xor(longest, m, 64);
            longest = longest > o ? longest : o;
            entries += __shfl_xor(entries, m, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMax(&a.stats[0], longest);
            atomicAdd(&a.stats[1], entries);
        }
    }
}

hipError_t launch_synth_flux(const FluxSynthArgs& a, hipStream_t stream) {
    const int blocks = (int)std::min<int64_t>((a.n + 255) / 256, 8192);
    synth_flux_kernel<<<blocks, 256, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace dtx
