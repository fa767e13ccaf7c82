// runtime.hip -- host runtime behind the dtx C ABI (include/dtx.h).
//
// Replaces the device placement of ForwardEulerStruct / ForwardSymplecticEulerStruct
// (src/SteppingStructs.jl:33-100, 142-227), the functor bodies (src/SteppingMethods.jl:12-107,
// 120-243) and the Solve loop (src/Solver.jl:1-64).  Operators are packed once on the host
// (binary channels in symmetric pair form, CSC -> column-sorted ELL), uploaded, and every step
// is two (ForwardEuler) or three (symplectic) kernel launches on one HIP stream with no host
// synchronisation; per-step diagnostics land in a device ring that the host reads when it
// synchronises.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "bf16_util.h"
#include "dtx_internal.h"
#include "status_text.h"

using namespace dtx;

namespace {

std::mutex g_err_mu;
std::string g_create_error;

struct HostChannel {
    int32_t s_out, s1, s2, kind;
    int32_t No, N1, N2;
    int64_t npairs, npairs_pad;  // the pair list is not padded: npairs_pad == npairs
    std::vector<double> S;  // [pair][row] FP64
    std::vector<uint32_t> pairs;
};

struct TimedEvent {
    hipEvent_t start = nullptr, stop = nullptr;
    int cls = 0;  // 0 binary, 1 update, 2 emission GEMV
};

}  // namespace

// What the host prints for a step once its status record is read: the verbose level it was
// stepped with and the end of its save interval (t_save in the verbose-3 line).
struct StepNote {
    int8_t verbose = 0;
    double t_stop = 0.0;
};

struct dtx_handle {
    dtx_config cfg;
    int32_t N = 0, C = 0, S = 0;
    int64_t n = 0;
    std::vector<int32_t> sp_off, sp_size;
    int32_t precision = 0, method = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // host staging of the binary operator
    bool binary = false;
    bool binary_dirty = false;
    std::map<std::tuple<int, int, int>, HostChannel> hch;
    std::vector<double> loss_host;  // N*N column-major
    bool loss_nonzero = false;
    bool synth_binary = false;      // channels generated on the device

    // padded state coordinates of the contraction's B operand (species start on SPECIES_ALIGN)
    int32_t Np = 0;
    std::vector<int32_t> sp_pad_off, pad_index, unpad;
    int32_t* d_pad_index = nullptr;
    int32_t* d_unpad = nullptr;

    // device: binary operator and its work tables (gain items in the contraction precision,
    // loss items always FP64)
    std::vector<PackedChannel> chans;
    // FP64 moment-rate correction of a reduced-precision gain term: per gain channel the
    // matrices W_N / W_E (LOSS layout, rows of species 1 over the states of species 2), their
    // slots, and the corrected binary sum the update kernels read
    int32_t moment_correction = 1;  // requested (applies to FP32 / BF16x3)
    bool correct_active = false;
    bool fused_update = false;  // corrected Forward Euler steps by euler_update_fused_kernel
    std::vector<PackedChannel> mchans;
    std::vector<double> eweight_host;
    // moment-rate correction: gain and loss sums per state, moment partials, and the
    // coefficients when a kernel of their own computes them (update grids of more than one
    // round of workgroups; otherwise every update workgroup derives them)
    double *d_gsum = nullptr, *d_lsum = nullptr, *d_cpart = nullptr, *d_coef = nullptr;
    int32_t zb = 0, zspan = UPDATE_STATES, tps = 4;
    int32_t *d_ones = nullptr, *d_gslots = nullptr, *d_nslots_s = nullptr, *d_mom_off = nullptr,
            *d_mom_list = nullptr;
    int32_t n_mom_list = 0, max_ns = 0;
    double* d_loss = nullptr;
    bool loss_present = false;
    BinWork* d_items = nullptr;     // loss and gain items of one launch
    // the same items with their partial slots and moment-rate sums in the device's `scratch`
    // array (dtx_step_noalloc), when they fit it
    BinWork* d_items_scratch = nullptr;
    double* scratch = nullptr;      // this device's address of `scratch` (null: unavailable)
    int64_t scratch_need = 0;       // bytes of scratch the no-alloc step needs
    std::vector<double> ring_host;  // host copy of the status ring (read without allocating)
    std::vector<BinWork> items_host;  // host copy (grouped launches)
    int32_t nitems = 0, tc = 16, ldc = 0, max_slots = 0, part_slots = 0;
    double* d_part = nullptr;
    double* d_mpart = nullptr;    // per-cell sums of the moment-rate items [n_mpart][ldc]
    int32_t n_mpart = 0;
    int64_t part_stride = 0;
    int32_t* d_nslot = nullptr;
    int32_t* d_nslot32 = nullptr;  // leading FP32 slots per state (wide split-BF16), or null
    int64_t pairs_total = 0;
    double bin_flops = 0, gain_bytes = 0;

    // device: state and vectors
    double* d_f[2] = {nullptr, nullptr};
    int cur = 0;
    double* d_fT = nullptr;       // FP64, state-major, padded states: [Np][ldc]
    float* d_fT32 = nullptr;      // FP32 contraction: FP32 copy of d_fT
    int num_cu = 256;             // compute units of the device (work-item sizing)
    int wide = 0;                 // wide gain items (C > 64): k tiles per item (BF16x3 2 or 4,
                                  // FP64 2), 0 for 64-cell tiles
    uint16_t* d_fhi = nullptr;    // BF16 split in fragment order (fb_index), ldc x Np
    uint16_t* d_flo = nullptr;
    double *d_invap = nullptr, *d_dfinj = nullptr, *d_vol = nullptr, *d_finit = nullptr;
    double* d_eweight = nullptr;
    int32_t *d_state_species = nullptr, *d_species_off = nullptr;
    double *d_moments = nullptr, *d_mom_part = nullptr, *d_block_part = nullptr,
           *d_status = nullptr;
    // Forward Euler on small grids: steps are finalised in batches (at FIN_BATCH pending steps
    // or before the next host synchronisation); each pending step keeps its block partials in
    // ring slot step % FIN_BATCH and its scalars in d_fin_rec
    int64_t bp_words = 0, mp_words = 0;   // doubles per slot of block_part / mom_part
    double* d_fin_rec = nullptr;          // [FIN_BATCH][FIN_WORDS]
    int fin_count = 0;                    // pending steps
    int64_t fin_first = 0;                // step index of the first pending step
    UpdateArgs fin_base{};                // arguments of the last pending step
    uint8_t* d_active = nullptr;
    EllMatrix ell[3];
    double* d_emi = nullptr;
    double* d_emi_part = nullptr;  // emission GEMV part sums (emi_q > 1)
    int32_t emi_q = 1;
    double *d_fmom = nullptr, *d_dfmom = nullptr, *d_dpflux = nullptr, *d_dbin = nullptr,
           *d_demi = nullptr;
    bool vectors_set = false;
    bool geometry_set = false;      // 1/Ap and Vol written by dtx_synth_flux
    bool emission = false;

    int64_t step_counter = 0;
    int64_t reported = 0;
    std::vector<StepNote> step_notes;  // steps not yet reported (ring order)
    int64_t device_bytes = 0;
    std::vector<void*> allocations;

    bool timing = false;
    int32_t timing_every = 1;   // time the kernels of every timing_every-th step
    std::vector<TimedEvent> events;
    dtx_kernel_times times{};
};

// ---------------------------------------------------------------- no-alloc scratch -------
// Device scratch of dtx_step_noalloc: the contraction's partial slots, then the per-cell sums
// of its moment-rate items.  One array per device, held until the process ends and shared by
// every no-alloc call on that device (calls that overlap on one device race on it).  Host code
// must never use the name: each device's address comes from hipGetSymbolAddress, taken once
// per device with that device current (dtx_create).
constexpr size_t SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[SCRATCH_MAX_DEVICES] = {};
}  // namespace

// This device's address of `scratch` (the device must be current), or null.
static double* scratch_address(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_addr[device] = (double*)ptr;
    }
    return g_scratch_addr[device];
}

// --------------------------------------------------------------------------- helpers ------

static int fail(dtx_handle* h, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    else {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_create_error = buf;
    }
    return code;
}

#define HIP_CHECK(h, expr)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((h), DTX_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                  \
                        hipGetErrorString(e_), __FILE__, __LINE__);                        \
    } while (0)

template <typename T>
static int dalloc(dtx_handle* h, T** p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    size_t bytes = (size_t)count * sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(h, DTX_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes,
                    hipGetErrorString(e));
    }
    h->device_bytes += (int64_t)bytes;
    h->allocations.push_back(*p);
    return DTX_OK;
}

template <typename T>
static void dfree(dtx_handle* h, T** p, int64_t count) {
    if (*p) {
        (void)hipFree(*p);
        auto it = std::find(h->allocations.begin(), h->allocations.end(), (void*)*p);
        if (it != h->allocations.end()) h->allocations.erase(it);
        h->device_bytes -= (int64_t)(count > 0 ? count : 1) * (int64_t)sizeof(T);
        *p = nullptr;
    }
}

// Enqueue the finalisation of the pending steps (one launch for all of them).
static hipError_t flush_finalize(dtx_handle* h) {
    if (h->fin_count == 0) return hipSuccess;
    UpdateArgs base = h->fin_base;
    base.block_part = h->d_block_part;  // the ring
    const int n = h->fin_count;
    h->fin_count = 0;
    return launch_finalize_batch(base, h->bp_words, (int)(h->fin_first % FIN_BATCH), n,
                                 h->stream);
}

// Synchronise the handle's stream after finalising the pending steps, so the status ring and
// the moments are complete when the host reads them.
static hipError_t stream_sync(dtx_handle* h) {
    hipError_t e = flush_finalize(h);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(h->stream);
}

template <typename T>
static int upload(dtx_handle* h, T* dst, const T* src, int64_t count) {
    if (count <= 0) return DTX_OK;
    HIP_CHECK(h, hipMemcpyAsync(dst, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice,
                                h->stream));
    HIP_CHECK(h, stream_sync(h));
    return DTX_OK;
}

static void pair_of(const HostChannel& ch, int64_t p, int* i, int* j) {
    uint32_t v = ch.pairs[p];
    *i = (int)(v & 0xFFFFu);
    *j = (int)(v >> 16);
}

uint64_t dtx::splitmix_key(uint64_t seed, uint64_t stream) {
    uint64_t z = seed + (stream + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Build the pair enumeration of a channel: TRI (i <= j, i outer) or RECT (i in s1 outer).
// Entries are cell-local state indices (i | j << 16), in row-major pair order.
static void build_pairs(const dtx_handle* h, HostChannel& ch) {
    ch.pairs.clear();
    const int o1 = h->sp_off[ch.s1], o2 = h->sp_off[ch.s2];
    if (ch.kind == CH_TRI) {
        for (int i = 0; i < ch.N1; ++i)
            for (int j = i; j < ch.N1; ++j)
                ch.pairs.push_back((uint32_t)(o1 + i) | ((uint32_t)(o1 + j) << 16));
    } else {
        for (int i = 0; i < ch.N1; ++i)
            for (int j = 0; j < ch.N2; ++j)
                ch.pairs.push_back((uint32_t)(o1 + i) | ((uint32_t)(o2 + j) << 16));
    }
    ch.npairs = (int64_t)ch.pairs.size();
    ch.npairs_pad = ch.npairs;
}

// Chunk table of a channel: rows i (padded index) cut into KC-aligned j chunks.  TRI rows start
// at the KC-aligned block holding i; the A entries of j < i or j past the species are zero.
static std::vector<uint32_t> chunk_table(const dtx_handle* h, int kind, int s1, int s2) {
    std::vector<uint32_t> t;
    const int N1 = h->sp_size[s1], N2 = h->sp_size[s2];
    const int P1 = h->sp_pad_off[s1], P2 = h->sp_pad_off[s2];
    if (kind == CH_TRI) {
        for (int i = 0; i < N1; ++i)
            for (int jb = i / KC * KC; jb < N1; jb += KC)
                t.push_back((uint32_t)(P1 + i) | ((uint32_t)(P1 + jb) << 16));
    } else if (kind == CH_RECT) {
        for (int i = 0; i < N1; ++i)
            for (int jb = 0; jb < N2; jb += KC)
                t.push_back((uint32_t)(P1 + i) | ((uint32_t)(P2 + jb) << 16));
    } else {
        for (int jb = 0; jb < h->Np; jb += KC) t.push_back((uint32_t)jb << 16);
    }
    return t;
}

static HostChannel& host_channel(dtx_handle* h, int s_out, int s1, int s2) {
    auto key = std::make_tuple(s_out, s1, s2);
    auto it = h->hch.find(key);
    if (it != h->hch.end()) return it->second;
    HostChannel ch;
    ch.s_out = s_out;
    ch.s1 = s1;
    ch.s2 = s2;
    ch.kind = (s1 == s2) ? CH_TRI : CH_RECT;
    ch.No = h->sp_size[s_out];
    ch.N1 = h->sp_size[s1];
    ch.N2 = h->sp_size[s2];
    build_pairs(h, ch);
    ch.S.assign((size_t)ch.npairs_pad * ch.No, 0.0);
    return h->hch.emplace(key, std::move(ch)).first->second;
}

static void release_channels(dtx_handle* h, std::vector<PackedChannel>& list) {
    for (auto& c : list) {
        if (c.A) {
            (void)hipFree(c.A);
            h->device_bytes -= c.a_bytes;
            auto it = std::find(h->allocations.begin(), h->allocations.end(), c.A);
            if (it != h->allocations.end()) h->allocations.erase(it);
        }
        if (c.chunks) {
            (void)hipFree(c.chunks);
            h->device_bytes -= c.nchunks * 4;
            auto it = std::find(h->allocations.begin(), h->allocations.end(), (void*)c.chunks);
            if (it != h->allocations.end()) h->allocations.erase(it);
        }
    }
    list.clear();
}

static void release_binary_device(dtx_handle* h) {
    release_channels(h, h->chans);
    release_channels(h, h->mchans);
}

static int64_t chunk_elems(int precision) {
    // FP64/FP32: [t][r] (KC x TK);  BF16X3: hi then lo, each in MFMA fragment order
    return (int64_t)KC * TK * (precision == DTX_PREC_BF16X3 ? 2 : 1);
}

static int64_t elem_bytes(int precision) {
    return precision == DTX_PREC_FP64 ? 8 : (precision == DTX_PREC_FP32 ? 4 : 2);
}

// Allocate the device arrays of a channel and upload its chunk table.
static int alloc_channel(dtx_handle* h, PackedChannel& pc, const std::vector<uint32_t>& table) {
    pc.nchunks = (int64_t)table.size();
    pc.a_bytes = chunk_elems(pc.precision) * pc.nchunks * pc.nktiles * elem_bytes(pc.precision);
    hipError_t e = hipMalloc(&pc.A, (size_t)std::max<int64_t>(pc.a_bytes, 1));
    if (e != hipSuccess) {
        pc.A = nullptr;
        return fail(h, DTX_ERR_OOM, "hipMalloc(%lld) for a binary channel failed: %s",
                    (long long)pc.a_bytes, hipGetErrorString(e));
    }
    h->device_bytes += pc.a_bytes;
    h->allocations.push_back(pc.A);
    e = hipMalloc((void**)&pc.chunks, (size_t)std::max<int64_t>(pc.nchunks, 1) * 4);
    if (e != hipSuccess) {
        pc.chunks = nullptr;
        return fail(h, DTX_ERR_OOM, "hipMalloc for a chunk table failed: %s",
                    hipGetErrorString(e));
    }
    h->device_bytes += pc.nchunks * 4;
    h->allocations.push_back(pc.chunks);
    HIP_CHECK(h, hipMemcpy(pc.chunks, table.data(), (size_t)pc.nchunks * 4,
                           hipMemcpyHostToDevice));
    return DTX_OK;
}

static int64_t tri_index(int i, int j, int n) {
    return (int64_t)i * n - (int64_t)i * (i - 1) / 2 + (j - i);
}

// Pack an FP64 pair-ordered channel into the chunked device layout of the handle's precision.
static int pack_upload_channel(dtx_handle* h, const HostChannel& hc) {
    PackedChannel pc;
    pc.s_out = hc.s_out;
    pc.s1 = hc.s1;
    pc.s2 = hc.s2;
    pc.kind = hc.kind;
    pc.precision = h->precision;
    pc.No = hc.No;
    pc.N1 = hc.N1;
    pc.N2 = hc.N2;
    pc.npairs = hc.npairs;
    pc.nktiles = (hc.No + TK - 1) / TK;
    const std::vector<uint32_t> table = chunk_table(h, hc.kind, hc.s1, hc.s2);
    const int P1 = h->sp_pad_off[hc.s1], P2 = h->sp_pad_off[hc.s2];
    const int64_t NCH = (int64_t)table.size();
    const int64_t CE = chunk_elems(h->precision);
    std::vector<uint8_t> buf((size_t)(CE * NCH * pc.nktiles * elem_bytes(h->precision)), 0);
    for (int kt = 0; kt < pc.nktiles; ++kt)
        for (int64_t ch = 0; ch < NCH; ++ch) {
            const int i = (int)(table[ch] & 0xFFFFu) - P1;
            const int jb = (int)(table[ch] >> 16) - (hc.kind == CH_TRI ? P1 : P2);
            const int64_t base = ((int64_t)kt * NCH + ch) * CE;
            for (int t = 0; t < KC; ++t) {
                const int j = jb + t;
                int64_t p = -1;
                if (hc.kind == CH_TRI) {
                    if (j >= i && j < hc.N1) p = tri_index(i, j, hc.N1);
                } else if (j < hc.N2) {
                    p = (int64_t)i * hc.N2 + j;
                }
                for (int r = 0; r < TK; ++r) {
                    const int row = kt * TK + r;
                    const double v = (p >= 0 && row < hc.No) ? hc.S[(size_t)p * hc.No + row] : 0.0;
                    if (h->precision == DTX_PREC_FP64) {
                        ((double*)buf.data())[base + t * TK + r] = v;
                    } else if (h->precision == DTX_PREC_FP32) {
                        ((float*)buf.data())[base + t * TK + r] = (float)v;
                    } else {
                        uint16_t hi, lo;
                        split_bf16(v, &hi, &lo);
                        ((uint16_t*)buf.data())[base + bf16_frag_index(r, t)] = hi;
                        ((uint16_t*)buf.data())[base + TK * KC + bf16_frag_index(r, t)] = lo;
                    }
                }
            }
        }
    int rc = alloc_channel(h, pc, table);
    h->chans.push_back(pc);
    if (rc) return rc;
    HIP_CHECK(h, hipMemcpy(pc.A, buf.data(), (size_t)pc.a_bytes, hipMemcpyHostToDevice));
    return DTX_OK;
}

// Chunk table of a moment-rate channel: the j chunks of species s2 (LOSS layout, no row part).
static std::vector<uint32_t> moment_chunk_table(const dtx_handle* h, int s2) {
    std::vector<uint32_t> t;
    for (int jb = 0; jb < h->sp_size[s2]; jb += KC)
        t.push_back((uint32_t)(h->sp_pad_off[s2] + jb) << 16);
    return t;
}

static PackedChannel moment_channel(const dtx_handle* h, int s_out, int s1, int s2, int kind,
                                    int moment) {
    PackedChannel mc;
    mc.s_out = s_out;
    mc.s1 = s1;
    mc.s2 = s2;
    mc.kind = CH_LOSS;  // rows i of s1, times f_i at the end
    mc.precision = DTX_PREC_FP64;
    mc.No = h->sp_size[s1];
    mc.N1 = h->sp_size[s1];
    mc.N2 = h->sp_size[s2];
    mc.npairs = kind == CH_TRI ? (int64_t)mc.N1 * (mc.N1 + 1) / 2 : (int64_t)mc.N1 * mc.N2;
    mc.nktiles = (mc.N1 + TK - 1) / TK;
    mc.moment = moment;
    return mc;
}

// Moment-rate matrices of a host channel: W_N[i,j] = sum_k S[k,(i,j)] and
// W_E[i,j] = sum_k E_k S[k,(i,j)] in FP64 (k ascending), from the channel's FP64 pair form.
static int build_moment_channels_host(dtx_handle* h, const HostChannel& hc) {
    const bool energy = !h->eweight_host.empty();
    const std::vector<uint32_t> table = moment_chunk_table(h, hc.s2);
    const int64_t nch = (int64_t)table.size();
    const double* ew = energy ? h->eweight_host.data() + h->sp_off[hc.s_out] : nullptr;
    for (int m = 0; m < (energy ? 2 : 1); ++m) {
        PackedChannel mc = moment_channel(h, hc.s_out, hc.s1, hc.s2, hc.kind, m);
        std::vector<double> buf((size_t)(mc.nktiles * nch * KC * TK), 0.0);
        for (int i = 0; i < hc.N1; ++i)
            for (int j = (hc.kind == CH_TRI ? i : 0); j < hc.N2; ++j) {
                const int64_t p = hc.kind == CH_TRI ? tri_index(i, j, hc.N1)
                                                    : (int64_t)i * hc.N2 + j;
                const double* S = hc.S.data() + (size_t)p * hc.No;
                double w = 0.0;
                for (int k = 0; k < hc.No; ++k) w = w + (m ? ew[k] * S[k] : S[k]);
                const int kt = i / TK, r = i % TK, ch = j / KC, t = j % KC;
                buf[(size_t)(((kt * nch + ch) * KC + t) * TK + r)] = w;
            }
        int rc = alloc_channel(h, mc, table);
        h->mchans.push_back(mc);
        if (rc) return rc;
        HIP_CHECK(h, hipMemcpy(mc.A, buf.data(), (size_t)mc.a_bytes, hipMemcpyHostToDevice));
    }
    return DTX_OK;
}

// The same matrices for a channel generated on the device, from its generator key.
static int build_moment_channels_synth(dtx_handle* h, const PackedChannel& pc) {
    const bool energy = h->d_eweight != nullptr;
    const std::vector<uint32_t> table = moment_chunk_table(h, pc.s2);
    PackedChannel mn = moment_channel(h, pc.s_out, pc.s1, pc.s2, pc.kind, 0);
    int rc = alloc_channel(h, mn, table);
    h->mchans.push_back(mn);
    if (rc) return rc;
    double* AE = nullptr;
    if (energy) {
        PackedChannel me = moment_channel(h, pc.s_out, pc.s1, pc.s2, pc.kind, 1);
        rc = alloc_channel(h, me, table);
        h->mchans.push_back(me);
        if (rc) return rc;
        AE = (double*)me.A;
    }
    HIP_CHECK(h, launch_synth_moment_packed((double*)mn.A, AE, mn.nchunks, mn.nktiles, pc.kind,
                                            pc.No, pc.N1, pc.N2,
                                            energy ? h->d_eweight + h->sp_off[pc.s_out] : nullptr,
                                            pc.synth_key, pc.synth_scale, h->stream));
    return DTX_OK;
}

// LOSS channels: one per species, rows of the species against all (padded) states, FP64,
// packed on the device from the loss matrix d_loss.
static int pack_loss_channels(dtx_handle* h) {
    if (!h->d_loss) return DTX_OK;
    for (int s = 0; s < h->S; ++s) {
        PackedChannel pc;
        pc.s_out = s;
        pc.s1 = s;
        pc.s2 = s;
        pc.kind = CH_LOSS;
        pc.precision = DTX_PREC_FP64;
        pc.No = h->sp_size[s];
        pc.N1 = h->sp_size[s];
        pc.N2 = h->N;
        pc.npairs = (int64_t)pc.No * h->N;
        pc.nktiles = (pc.No + TK - 1) / TK;
        int rc = alloc_channel(h, pc, chunk_table(h, CH_LOSS, s, s));
        h->chans.push_back(pc);
        if (rc) return rc;
        HIP_CHECK(h, launch_pack_loss(h->d_loss, h->N, h->d_unpad, h->Np, h->sp_off[s], pc.No,
                                      pc.nktiles, (double*)pc.A, h->stream));
    }
    HIP_CHECK(h, stream_sync(h));
    return DTX_OK;
}

// Build the grouped work tables for the handle's channels and cell count.
static int build_work(dtx_handle* h) {
    dfree(h, &h->d_items, h->nitems);
    dfree(h, &h->d_items_scratch, h->nitems);
    dfree(h, &h->d_part, h->part_stride * h->part_slots);
    dfree(h, &h->d_nslot, h->N);
    dfree(h, &h->d_nslot32, h->N);
    dfree(h, &h->d_gsum, h->n);
    dfree(h, &h->d_lsum, h->n);
    dfree(h, &h->d_cpart, (int64_t)h->C * h->S * h->zb * 3);
    dfree(h, &h->d_coef, (int64_t)h->C * h->S * 2);
    dfree(h, &h->d_ones, h->N);
    dfree(h, &h->d_gslots, h->S);
    dfree(h, &h->d_nslots_s, h->S);
    dfree(h, &h->d_mom_off, 2 * h->S + 1);
    dfree(h, &h->d_mom_list, h->n_mom_list);
    dfree(h, &h->d_mpart, (int64_t)h->n_mpart * h->ldc);
    h->n_mpart = 0;
    h->nitems = 0;
    h->max_slots = h->part_slots = 0;
    h->correct_active = false;
    h->fused_update = false;
    const int C = h->C;
    const int ctiles = (C + h->tc - 1) / h->tc;
    // wide gain items: h->wide k tiles x WIDE cells each
    const int wctiles = (C + WIDE - 1) / WIDE;
    auto gain_ksteps = [&](const PackedChannel& c) {
        return h->wide ? (c.nktiles + h->wide - 1) / h->wide : c.nktiles;
    };
    double total = 0;
    h->pairs_total = 0;
    h->bin_flops = 0;
    h->gain_bytes = 0;
    for (auto& c : h->chans) {
        if (c.kind != CH_LOSS && h->wide) total += (double)gain_ksteps(c) * wctiles * c.nchunks;
        else total += (double)c.nktiles * ctiles * c.nchunks;
        if (c.kind != CH_LOSS) {
            h->pairs_total += c.npairs;
            h->bin_flops += 2.0 * c.No * (double)c.npairs * C;
        } else {
            h->bin_flops += 2.0 * c.No * (double)h->N * C;
        }
        h->gain_bytes += (double)c.a_bytes;
    }
    // LOSS-layout channels (losses, moment rates) are small FP64 GEMVs next to the gain
    // contraction: short ranges keep their items from forming the tail of the launch.  At most
    // 4 chunks next to 64-cell gain items, 16 next to wide ones (measured: each range is a
    // partial slot that is written and read back, and with 4 the wide launches were 3-6%
    // slower at transport_1d and transport_2d, while the 64-cell FP32 launch at transport_1d
    // was 5% faster)
    const int64_t loss_cap = h->wide ? 16 : 4;
    auto loss_range_of = [&](int64_t rg) { return std::min<int64_t>(rg, loss_cap); };
    // wide FP64 workgroups of several groups: LOSS items take the wide shape as well, so every
    // workgroup of the launch holds a CU alike
    const bool wide_loss =
        h->wide && h->precision == DTX_PREC_FP64 && F64_WIDE_GROUPS > 1;
    auto launch_items = [&](int64_t rg) {
        const int64_t lr = loss_range_of(rg);
        int64_t n = 0;
        for (auto& c : h->chans)
            n += c.kind == CH_LOSS
                     ? (c.nchunks + lr - 1) / lr *
                           (wide_loss ? gain_ksteps(c) * wctiles : c.nktiles * ctiles)
                     : (c.nchunks + rg - 1) / rg * gain_ksteps(c) * (h->wide ? wctiles : ctiles);
        for (auto& c : h->mchans) n += (c.nchunks + lr - 1) / lr * c.nktiles * ctiles;
        return n;
    };
    // Chunks per gain item.  64-cell launches (about four resident workgroups per CU) take
    // ranges for about 1024 items.  The wide launch with two k tiles holds two workgroups per
    // CU and runs fastest when all its items fit one round of them: the smallest range with at
    // most 2 x CUs items (at transport_1d, FP64 and split-BF16, a round and a bit more measured
    // 10-20% slower, and fewer, longer items slower too).  With four k tiles (one workgroup per
    // CU) the items cannot fit one round; about three rounds of gain items measured best.  A
    // wide FP64 workgroup of G groups holds a CU alone and covers G chunk shares: one round is
    // CUs items.
    auto clamp_range = [](double r) {
        return std::min<int64_t>(MAX_RANGE, std::max<int64_t>(8, (int64_t)std::ceil(r)));
    };
    int64_t range = clamp_range(total / (h->wide == 4 ? 768 : h->wide ? 512 : 1024));
    if (h->wide == 2) {
        const int64_t round = (wide_loss ? 1 : 2) * (int64_t)h->num_cu;
        for (int64_t rg = 8; rg <= MAX_RANGE; ++rg)
            if (launch_items(rg) <= round) {
                range = rg;
                break;
            }
    }
    const int64_t loss_range = loss_range_of(range);
    std::vector<int> next_slot(h->S, 0), gain_slots(h->S, 0);
    h->part_stride = (int64_t)C * h->N;
    // A unit is (channel, chunk range, k tile); it has one item per cell tile.  Gain slots of a
    // species come first, then its loss slots, and the update kernel adds them in slot order.
    // Moment-rate slots follow all of them.
    struct Unit { BinWork w; int slot; int nct; };
    std::vector<Unit> gain_units, loss_units, moment_units;
    const bool bf16 = h->precision == DTX_PREC_BF16X3;
    auto add_units = [&](const PackedChannel& c, int64_t rg, std::vector<Unit>& out,
                         const std::function<int()>& next) {
        const bool loss = c.kind == CH_LOSS;
        const bool wide = h->wide && (!loss || wide_loss);
        const int prec = loss ? DTX_PREC_FP64 : h->precision;
        const int s_row = c.moment >= 0 ? c.s1 : c.s_out;  // owner of the output rows
        for (int64_t c0 = 0; c0 < c.nchunks; c0 += rg) {
            const int64_t c1 = std::min<int64_t>(c0 + rg, c.nchunks);
            const int slot = next();
            for (int kt = 0; kt < (wide ? gain_ksteps(c) : c.nktiles); ++kt) {
                BinWork w{};
                w.A = c.A;
                w.chunks = c.chunks;
                w.fT = h->d_fT;
                if (prec == DTX_PREC_FP64) w.B = h->d_fT;
                else if (prec == DTX_PREC_FP32) w.B = h->d_fT32;
                else w.B = h->d_fhi;
                w.b_lo = (bf16 && !loss) ? (int64_t)(h->d_flo - h->d_fhi) : 0;
                w.a_ktile_stride = chunk_elems(prec) * c.nchunks;
                w.ktile = kt;
                w.c_begin = (int32_t)c0;
                w.c_end = (int32_t)c1;
                w.row_base = h->sp_off[s_row];
                w.row_pad_base = h->sp_pad_off[s_row];
                w.rows = c.No;
                w.cells = C;
                w.ldc = h->ldc;
                w.ldb = h->ldc / 16;
                w.N = h->N;
                w.kind = c.moment >= 0 ? CH_MOMENT : c.kind;
                w.tc = wide ? WIDE : h->tc;
                out.push_back({w, slot, wide ? wctiles : ctiles});
            }
        }
    };
    for (int pass = 0; pass < 2; ++pass) {
        for (auto& c : h->chans)
            if ((c.kind == CH_LOSS) == (pass == 1))
                add_units(c, pass ? loss_range : range, pass ? loss_units : gain_units,
                          [&]() { return next_slot[c.s_out]++; });
        if (pass == 0) gain_slots = next_slot;  // gain slots of species s: [0, gain_slots[s])
    }
    for (int s = 0; s < h->S; ++s) h->max_slots = std::max(h->max_slots, next_slot[s]);
    // moment-rate items write per-cell sums: one row of ldc values per (moment channel, chunk
    // range, k tile), listed per (species, moment) in summation order
    int n_mpart = 0;
    std::vector<std::vector<int32_t>> mom(2 * h->S);
    for (auto& c : h->mchans)
        add_units(c, loss_range, moment_units, [&]() {
            const int base = n_mpart;
            n_mpart += c.nktiles;
            auto& l = mom[2 * c.s_out + c.moment];
            for (int kt = 0; kt < c.nktiles; ++kt) l.push_back(base + kt);
            return base;
        });
    h->part_slots = h->max_slots;
    h->n_mpart = n_mpart;
    if (n_mpart > 0) {
        int rc = dalloc(h, &h->d_mpart, (int64_t)n_mpart * h->ldc);
        if (rc) return rc;
    }
    if (h->part_slots > 0) {
        int rc = dalloc(h, &h->d_part, h->part_stride * h->part_slots);
        if (rc) return rc;
    }
    // Launch order: gain units, then the short LOSS-layout units.  Workgroups are dispatched to
    // the eight XCDs round-robin, and each XCD has its own L2: units go in groups of eight, and
    // the item of (unit x of the group, cell tile t) is number 8t + x of the group, so all cell
    // tiles of a unit -- which read the same operator block -- run on one XCD at about the same
    // time and fetch the block from HBM once.
    std::vector<const Unit*> order;
    for (const auto* list : {&gain_units, &loss_units, &moment_units})
        for (const Unit& u : *list) order.push_back(&u);
    std::vector<BinWork> items;
    items.reserve(order.size() * ctiles);
    for (size_t g0 = 0; g0 < order.size(); g0 += 8) {
        const size_t g1 = std::min(order.size(), g0 + 8);
        int nct = 0;
        for (size_t x = g0; x < g1; ++x) nct = std::max(nct, order[x]->nct);
        for (int ct = 0; ct < nct; ++ct)
            for (size_t x = g0; x < g1; ++x) {
                if (ct >= order[x]->nct) continue;
                BinWork w = order[x]->w;
                w.ctile = ct;
                w.part = w.kind == CH_MOMENT
                             ? h->d_mpart + (int64_t)(order[x]->slot + w.ktile) * h->ldc
                             : h->d_part + (int64_t)order[x]->slot * h->part_stride;
                items.push_back(w);
            }
    }
    h->nitems = (int32_t)items.size();
    h->items_host = items;
    if (h->nitems) {
        int rc = dalloc(h, &h->d_items, h->nitems);
        if (!rc) rc = upload(h, h->d_items, items.data(), h->nitems);
        if (rc) return rc;
    }
    // the no-alloc step's items: the same work with its outputs in `scratch`
    h->scratch_need = (h->part_stride * h->part_slots + (int64_t)h->n_mpart * h->ldc) * 8;
    if (h->nitems && h->scratch && (size_t)h->scratch_need <= SCRATCH_BYTES) {
        double* mbase = h->scratch + h->part_stride * h->part_slots;
        std::vector<BinWork> si = items;
        for (BinWork& w : si)
            w.part = w.kind == CH_MOMENT ? mbase + (w.part - h->d_mpart)
                                         : h->scratch + (w.part - h->d_part);
        int rc = dalloc(h, &h->d_items_scratch, h->nitems);
        if (!rc) rc = upload(h, h->d_items_scratch, si.data(), h->nitems);
        if (rc) return rc;
    }
    std::vector<int32_t> nslot(h->N, 0);
    for (int s = 0; s < h->S; ++s)
        for (int k = 0; k < h->sp_size[s]; ++k) nslot[h->sp_off[s] + k] = next_slot[s];
    int rc = dalloc(h, &h->d_nslot, h->N);
    if (!rc) rc = upload(h, h->d_nslot, nslot.data(), h->N);
    if (rc) return rc;
    if (h->wide && bf16) {  // wide split-BF16 gain items write FP32 sums: the gain slots
        std::vector<int32_t> n32(h->N, 0);
        for (int s = 0; s < h->S; ++s)
            for (int k = 0; k < h->sp_size[s]; ++k) n32[h->sp_off[s] + k] = gain_slots[s];
        rc = dalloc(h, &h->d_nslot32, h->N);
        if (!rc) rc = upload(h, h->d_nslot32, n32.data(), h->N);
        if (rc) return rc;
    }
    if (h->mchans.empty()) return DTX_OK;
    // correction tables
    std::vector<int32_t> ones(h->N, 1), moff(2 * h->S + 1, 0), mlist;
    for (int q = 0; q < 2 * h->S; ++q) {
        moff[q] = (int32_t)mlist.size();
        mlist.insert(mlist.end(), mom[q].begin(), mom[q].end());
    }
    moff[2 * h->S] = (int32_t)mlist.size();
    h->n_mom_list = (int32_t)mlist.size();
    h->max_ns = 0;
    for (int s = 0; s < h->S; ++s) h->max_ns = std::max(h->max_ns, h->sp_size[s]);
    std::vector<int32_t> nsl(next_slot.begin(), next_slot.end());
    // threads per state in the slot sums: 16 when four would leave the launch with fewer than
    // 65536 threads (single-cell boxes), so each thread has fewer slots to walk
    h->tps = (int64_t)h->C * h->max_ns * 4 < 65536 ? 16 : 4;
    h->zspan = UPDATE_THREADS / h->tps;
    h->zb = (h->max_ns + h->zspan - 1) / h->zspan;
    rc = dalloc(h, &h->d_gsum, h->n);
    if (!rc) rc = dalloc(h, &h->d_lsum, h->n);
    if (!rc) rc = dalloc(h, &h->d_cpart, (int64_t)h->C * h->S * h->zb * 3);
    // one round of update workgroups: about four per CU
    const int64_t update_wgs = (int64_t)h->C * ((h->N + UPDATE_STATES - 1) / UPDATE_STATES);
    if (!rc && update_wgs > 4 * (int64_t)h->num_cu)
        rc = dalloc(h, &h->d_coef, (int64_t)h->C * h->S * 2);
    if (!rc) rc = dalloc(h, &h->d_ones, h->N);
    if (!rc) rc = dalloc(h, &h->d_gslots, h->S);
    if (!rc) rc = dalloc(h, &h->d_nslots_s, h->S);
    if (!rc) rc = dalloc(h, &h->d_mom_off, 2 * h->S + 1);
    if (!rc) rc = dalloc(h, &h->d_mom_list, h->n_mom_list);
    if (!rc) rc = upload(h, h->d_ones, ones.data(), h->N);
    if (!rc) rc = upload(h, h->d_gslots, gain_slots.data(), h->S);
    if (!rc) rc = upload(h, h->d_nslots_s, nsl.data(), h->S);
    if (!rc) rc = upload(h, h->d_mom_off, moff.data(), 2 * h->S + 1);
    if (!rc) rc = upload(h, h->d_mom_list, mlist.data(), h->n_mom_list);
    if (rc) return rc;
    h->correct_active = true;
    // one launch where its grid, one workgroup per (cell, species), is one round on the CUs
    h->fused_update = h->moment_correction != 2 && (int64_t)h->C * h->S <= h->num_cu &&
                      fused_update_fits(h->N, h->S, h->sp_off.data(), h->tps);
    return DTX_OK;
}

static int finalize_binary(dtx_handle* h) {
    if (!h->binary_dirty) return DTX_OK;
    const bool want_correction = h->moment_correction && h->precision != DTX_PREC_FP64;
    if (!h->synth_binary) {
        release_binary_device(h);
        for (auto& kv : h->hch) {
            const HostChannel& hc = kv.second;
            bool any = false;
            for (double v : hc.S)
                if (v != 0.0) { any = true; break; }
            if (!any) continue;
            int rc = pack_upload_channel(h, hc);
            if (!rc && want_correction) rc = build_moment_channels_host(h, hc);
            if (rc) return rc;
        }
        h->hch.clear();
        dfree(h, &h->d_loss, (int64_t)h->N * h->N);
        if (h->loss_nonzero) {
            int rc = dalloc(h, &h->d_loss, (int64_t)h->N * h->N);
            if (rc) return rc;
            rc = upload(h, h->d_loss, h->loss_host.data(), (int64_t)h->N * h->N);
            if (rc) return rc;
        }
        h->loss_host.clear();
        h->loss_host.shrink_to_fit();
    }
    if (h->synth_binary) {
        release_channels(h, h->mchans);
        if (want_correction)
            for (const PackedChannel& pc : h->chans)
                if (pc.synth) {
                    int rc = build_moment_channels_synth(h, pc);
                    if (rc) return rc;
                }
    }
    int rc = pack_loss_channels(h);
    if (rc) return rc;
    h->loss_present = h->d_loss != nullptr;
    dfree(h, &h->d_loss, (int64_t)h->N * h->N);
    rc = build_work(h);
    if (rc) return rc;
    h->binary_dirty = false;
    return DTX_OK;
}

// Publish the current f into the contraction's B-operand copies (padded, per precision).
static int refresh_fT(dtx_handle* h) {
    HIP_CHECK(h, launch_publish_state(h->d_f[h->cur], h->N, h->C, h->ldc, h->d_pad_index,
                                      h->d_fT, h->d_fT32, h->d_fhi, h->d_flo, h->stream));
    return DTX_OK;
}

static int ensure_ready(dtx_handle* h) {
    if (!h->vectors_set)
        return fail(h, DTX_ERR_STATE, "dtx_set_vectors must be called before stepping");
    if (h->binary && h->binary_dirty) {
        int rc = finalize_binary(h);
        if (rc) return rc;
        rc = refresh_fT(h);  // ldc may have been set now
        if (rc) return rc;
    }
    return DTX_OK;
}

// The functor's console lines, worded as src/SteppingMethods.jl:49,87-93 (ForwardEuler) and
// :222-228 (ForwardSymplecticEuler), numbers in Julia's form.
static void print_status(const dtx_handle* h, const double* st, const StepNote& note) {
    const bool euler = h->method == DTX_METHOD_FORWARD_EULER;
    if (euler && st[ST_FINITE] == 0.0) printf("non-finite value in df calculation\n");
    const int verbose = note.verbose;
    if (verbose < 1 || verbose > 3) return;
    const double cr = st[ST_CR];
    // only lines that are printed get formatted (the shortest-digit search is costly)
    if (verbose == 1 && !(cr > 1.0)) return;
    char a[40], b[40], c[40], d[40], e[40], f[40], g[40], t[40], ts[40];
    auto cr_text = [&](char* out, double v) { status_text::julia_float(out, 40, euler ? v : status_text::round_sig3(v)); };
    cr_text(a, cr);
    status_text::julia_float(t, sizeof t, st[ST_TSTART]);
    status_text::julia_float(b, sizeof b, st[ST_DT]);
    if (verbose == 1) {
        printf("Cr = %s, t=%s, dt=%s, system may be unstable\n", a, t, b);
    } else if (verbose == 2) {
        printf(euler ? "\rCr = %s, t=%s, dt=%s\n" : "Cr = %s, t=%s, dt=%s\n", a, t, b);
    } else {
        cr_text(c, st[ST_CR_BIN]);
        cr_text(d, st[ST_CR_EMI]);
        cr_text(e, st[ST_CR_FLUX]);
        status_text::julia_float(ts, sizeof ts, note.t_stop);
        if (euler) {
            printf("Cr = %s,Cr_Bin = %s, Cr_Emi = %s, Cr_Flux = %s, t=%s, t_save =%s, dt=%s\n",
                   a, c, d, e, t, ts, b);
        } else {
            cr_text(f, st[ST_CR_MOM]);
            cr_text(g, st[ST_CR_SPACE]);
            printf("Cr = %s, Cr_Bin = %s, Cr_Emi = %s, Cr_PFlux = %s, Cr_Momentum = %s, "
                   "Cr_Space = %s,  t=%s, t_save=%s , dt=%s\n", a, c, d, e, f, g, t, ts, b);
        }
    }
}

// Read and print the diagnostics of every step not yet reported (the stream must be idle).
static int drain_status(dtx_handle* h) {
    const int64_t pending = h->step_counter - h->reported;
    if (pending <= 0) return DTX_OK;
    std::vector<double>& ring = h->ring_host;  // sized by dtx_create
    HIP_CHECK(h, hipMemcpy(ring.data(), h->d_status, ring.size() * sizeof(double),
                           hipMemcpyDeviceToHost));
    const int64_t first = std::max<int64_t>(h->reported, h->step_counter - STATUS_RING);
    for (int64_t s = first; s < h->step_counter; ++s) {
        const double* st = ring.data() + (s % STATUS_RING) * STATUS_WORDS;
        print_status(h, st, h->step_notes[(size_t)(s % STATUS_RING)]);
    }
    fflush(stdout);
    h->reported = h->step_counter;
    return DTX_OK;
}

static void fill_stats(const double* st, dtx_step_stats* out) {
    memset(out, 0, sizeof *out);
    out->step_index = (int32_t)st[ST_STEP];
    out->finite = (int32_t)st[ST_FINITE];
    out->cr_evaluated = (int32_t)st[ST_CR_EVAL];
    out->cr = st[ST_CR];
    out->cr_bin = st[ST_CR_BIN];
    out->cr_emi = st[ST_CR_EMI];
    out->cr_flux = st[ST_CR_FLUX];
    out->cr_momentum = st[ST_CR_MOM];
    out->cr_space = st[ST_CR_SPACE];
    out->sum_f = st[ST_SUMF];
    out->dt = st[ST_DT];
    out->t_start = st[ST_TSTART];
    out->save = (int32_t)st[ST_SAVE];
}

// Largest number of steps Solve accepts for one save interval (see dtx_solve).
constexpr double kMaxIntervalSteps = 1e8;

static UpdateArgs make_args(dtx_handle* h, double s, int verbose, double t_start, double dt,
                            int save, int which_ell, bool use_scratch = false) {
    UpdateArgs a{};
    a.f = h->d_f[h->cur];
    a.f_out = h->d_f[h->cur ^ 1];
    a.fT = h->d_fT;
    a.fT32 = h->d_fT32;
    a.fhi = h->d_fhi;
    a.flo = h->d_flo;
    a.pad_index = h->d_pad_index;
    a.ldb = h->ldc / 16;
    a.inv_ap = h->d_invap;
    a.df_inj = h->d_dfinj;
    a.vol = h->d_vol;
    if (h->correct_active) {  // gain sums rescaled per (cell, species), plus loss sums
        a.part = h->d_lsum;
        a.nslot = h->d_ones;
        a.nslot32 = nullptr;
        a.gsum = h->d_gsum;
        a.cpart = h->d_cpart;
        a.mpart = h->d_mpart;
        a.mom_off = h->d_mom_off;
        a.mom_list = h->d_mom_list;
        a.zb = h->zb;
        a.zspan = h->zspan;
        a.coef = h->d_coef;
    } else {
        a.part = h->d_part;
        a.nslot = h->d_nslot;
        a.nslot32 = h->d_nslot32;
    }
    a.part_stride = h->part_stride;
    if (use_scratch) {  // the no-alloc step's slots and moment-rate sums
        if (!h->correct_active) a.part = h->scratch;
        else a.mpart = h->scratch + h->part_stride * h->part_slots;
    }
    a.emi = h->d_emi;
    a.emi_part = h->d_emi_part;
    a.emi_q = h->emi_q;
    a.active = h->d_active;
    const EllMatrix& E = h->ell[which_ell];
    a.ell_col = E.col;
    a.ell_val = E.val;
    a.ell_width = E.width;
    a.eweight = h->d_eweight;
    a.state_species = h->d_state_species;
    a.species_off = h->d_species_off;
    a.moments = h->d_moments;
    a.mom_part = h->d_mom_part;
    a.block_part = h->d_block_part;
    a.status = h->d_status;
    a.f_mom = h->d_fmom;
    a.df_mom = h->d_dfmom;
    a.dpflux = h->d_dpflux;
    a.dbin = h->d_dbin;
    a.demi = h->d_demi;
    a.n = h->n;
    a.N = h->N;
    a.C = h->C;
    a.S = h->S;
    a.ldc = h->ldc;
    a.binary = h->binary && h->nitems > 0;
    a.emission = h->emission;
    a.verbose = verbose;
    a.step_index = (int32_t)(h->step_counter % 0x7FFFFFFF);
    a.ring = STATUS_RING;
    a.s = s;
    a.t_start = t_start;
    a.dt = dt;
    a.save = save;
    return a;
}

static int timed_begin(dtx_handle* h, int cls, TimedEvent* ev) {
    *ev = TimedEvent{};
    if (!h->timing || h->step_counter % h->timing_every != 0) return DTX_OK;
    ev->cls = cls;
    HIP_CHECK(h, hipEventCreate(&ev->start));
    HIP_CHECK(h, hipEventCreate(&ev->stop));
    HIP_CHECK(h, hipEventRecord(ev->start, h->stream));
    return DTX_OK;
}

static int timed_end(dtx_handle* h, TimedEvent* ev) {
    if (!ev->start) return DTX_OK;
    HIP_CHECK(h, hipEventRecord(ev->stop, h->stream));
    h->events.push_back(*ev);
    return DTX_OK;
}

// Arguments of the moment-rate correction's slot-sum launch (correct_active handles).
static CorrectArgs make_correct_args(const dtx_handle* h) {
    CorrectArgs ca{};
    ca.part = h->d_part;
    ca.part_stride = h->part_stride;
    ca.gsum = h->d_gsum;
    ca.lsum = h->d_lsum;
    ca.cpart = h->d_cpart;
    ca.zb = h->zb;
    ca.tps = h->tps;
    ca.species_off = h->d_species_off;
    ca.gslots = h->d_gslots;
    ca.nslots = h->d_nslots_s;
    ca.gain_f32 = h->wide && h->precision == DTX_PREC_BF16X3;
    ca.eweight = h->d_eweight;
    ca.N = h->N;
    ca.C = h->C;
    ca.S = h->S;
    return ca;
}

// Enqueue one functor call with an already clipped dt.  use_scratch: the contraction's slots
// and moment-rate sums in the device's `scratch` array, and no kernel timing (no events).
static int enqueue_step(dtx_handle* h, double t_start, double t_stop, double dt, int verbose,
                        int save, bool use_scratch = false) {
    const double s = dt / h->cfg.dt0;  // src/SteppingMethods.jl:27
    const bool bin = h->binary && h->nitems > 0;
    const bool timing = h->timing && !use_scratch;
    // corrected Forward Euler: slot sums, coefficients and update in one launch where it fits
    const bool fused = bin && h->correct_active && h->fused_update &&
                       h->method == DTX_METHOD_FORWARD_EULER;
    CorrectArgs fca{};
    if (fused) {
        fca = make_correct_args(h);
        if (use_scratch) fca.part = h->scratch;
    }
    TimedEvent ev{};
    if (h->emission) {
        // df_Emi of the state the step starts from, read by the update kernels
        int rc = timing ? timed_begin(h, 2, &ev) : DTX_OK;
        if (rc) return rc;
        const UpdateArgs ea = make_args(h, s, verbose, t_start, dt, save, DTX_FLUX_F, use_scratch);
        HIP_CHECK(h, launch_emission(ea, h->stream));
        rc = timing ? timed_end(h, &ev) : DTX_OK;
        if (rc) return rc;
    }
    if (bin) {
        int rc = timing ? timed_begin(h, 0, &ev) : DTX_OK;
        if (rc) return rc;
        const BinWork* items = use_scratch ? h->d_items_scratch : h->d_items;
        if (h->wide)
            HIP_CHECK(h, launch_binary_wide(h->precision, h->wide, items, h->nitems, h->stream));
        else HIP_CHECK(h, launch_binary(h->precision, h->tc, items, h->nitems, h->stream));
        if (h->correct_active && !fused) {
            CorrectArgs ca = make_correct_args(h);
            if (use_scratch) ca.part = h->scratch;
            HIP_CHECK(h, launch_gain_correct(ca, h->stream));
        }
        rc = timing ? timed_end(h, &ev) : DTX_OK;
        if (rc) return rc;
    }
    int rc = timing ? timed_begin(h, 1, &ev) : DTX_OK;
    if (rc) return rc;
    if (h->method == DTX_METHOD_FORWARD_EULER) {
        UpdateArgs a = make_args(h, s, verbose, t_start, dt, save, DTX_FLUX_F, use_scratch);
        if (finalize_is_single(a)) {
            // batched finalisation: this step's block partials and scalars go to its ring slot
            const int slot = (int)(h->step_counter % FIN_BATCH);
            a.block_part = h->d_block_part + slot * h->bp_words;
            a.fin_rec = h->d_fin_rec;
            a.fin_slot = slot;
            if (fused) HIP_CHECK(h, launch_euler_update_fused(a, fca, false, h->stream));
            else HIP_CHECK(h, launch_euler_update_only(a, h->stream));
            if (h->fin_count == 0) h->fin_first = h->step_counter;
            h->fin_count++;
            h->fin_base = a;
            if (h->fin_count == FIN_BATCH) HIP_CHECK(h, flush_finalize(h));
        } else if (fused) {
            HIP_CHECK(h, launch_euler_update_fused(a, fca, true, h->stream));
        } else {
            HIP_CHECK(h, launch_euler_update(a, h->stream));
        }
    } else {
        UpdateArgs a = make_args(h, s, verbose, t_start, dt, save, DTX_FLUX_P, use_scratch);
        HIP_CHECK(h, launch_symplectic_momentum(a, h->stream));
        UpdateArgs b = make_args(h, s, verbose, t_start, dt, save, DTX_FLUX_X, use_scratch);
        HIP_CHECK(h, launch_symplectic_space(b, h->stream));
    }
    rc = timing ? timed_end(h, &ev) : DTX_OK;
    if (rc) return rc;
    h->step_notes[(size_t)(h->step_counter % STATUS_RING)] = StepNote{(int8_t)verbose, t_stop};
    h->cur ^= 1;
    h->step_counter++;
    if (h->step_counter - h->reported >= STATUS_RING / 2) {
        HIP_CHECK(h, stream_sync(h));
        return drain_status(h);
    }
    return DTX_OK;
}

// Free an ELL matrix's arrays and mark it absent.
static void release_ell(dtx_handle* h, EllMatrix& E) {
    dfree(h, &E.col, E.n * E.width);
    dfree(h, &E.val, E.n * E.width);
    E.width = 0;
    E.nnz = 0;
}

// CSC (1-based) -> ELL with slots sorted by column inside each row.
static int set_ell(dtx_handle* h, EllMatrix& E, int64_t n, const int64_t* colptr,
                   const int64_t* rowval, const double* nzval) {
    std::vector<std::vector<std::pair<int64_t, double>>> rows((size_t)n);
    for (int64_t j = 0; j < n; ++j) {
        const int64_t k0 = colptr[j] - 1, k1 = colptr[j + 1] - 1;
        if (k0 < 0 || k1 < k0)
            return fail(h, DTX_ERR_INVALID, "colptr is not a valid 1-based CSC pointer");
        for (int64_t k = k0; k < k1; ++k) {
            const int64_t r = rowval[k] - 1;
            if (r < 0 || r >= n) return fail(h, DTX_ERR_INVALID, "rowval out of range");
            rows[(size_t)r].push_back({j, nzval[k]});
        }
    }
    int64_t width = 0, nnz = 0;
    for (auto& r : rows) {
        width = std::max<int64_t>(width, (int64_t)r.size());
        nnz += (int64_t)r.size();
    }
    release_ell(h, E);
    if (width == 0) return DTX_OK;
    std::vector<int32_t> col((size_t)(width * n), -1);
    std::vector<double> val((size_t)(width * n), 0.0);
    for (int64_t r = 0; r < n; ++r) {
        // column order is already ascending (columns were visited in order)
        auto& v = rows[(size_t)r];
        for (size_t s = 0; s < v.size(); ++s) {
            col[(size_t)s * n + r] = (int32_t)v[s].first;
            val[(size_t)s * n + r] = v[s].second;
        }
    }
    int32_t* dcol = nullptr;
    double* dval = nullptr;
    int rc = dalloc(h, &dcol, width * n);
    if (!rc) rc = dalloc(h, &dval, width * n);
    if (!rc) rc = upload(h, dcol, col.data(), width * n);
    if (!rc) rc = upload(h, dval, val.data(), width * n);
    if (rc) {  // the matrix stays absent; nothing is left allocated
        dfree(h, &dcol, width * n);
        dfree(h, &dval, width * n);
        return rc;
    }
    E.col = dcol;
    E.val = dval;
    E.n = n;
    E.width = (int32_t)width;
    E.nnz = nnz;
    return DTX_OK;
}

// ------------------------------------------------------------------- grouped stepping -------
// Several simulations advanced together: one contraction launch per (kernel precision, cell-tile
// width) over the members' concatenated work tables, one slot-sum and one coefficient launch
// for the members with the FP64 moment-rate correction, and one update launch over a table of
// the members' update arguments.  Each member's work items, reduction order and correction
// arithmetic are the ones it uses alone, so its results do not depend on the group.
//
// Membership changes never synchronise the device: the combined tables live in grow-only
// device buffers and are rewritten by copies on the group stream, which run after every launch
// already enqueued there that reads the previous tables.  The host stages each copy in a
// pinned buffer of a small ring whose reuse waits on the event recorded after its copy.

namespace dtx {

constexpr int GROUP_RING = 8;

struct StagedUpload {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
};

struct Group {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<dtx_handle*> members;
    int precision = 0;
    // static tables of the membership, one device buffer:
    // narrow items (list 0), wide items of 2 / 4 k tiles (lists 2, 3), update block starts,
    // slot-sum argument table and block starts, coefficient block starts
    void* d_static = nullptr;
    size_t static_cap = 0;
    std::vector<void*> retired;        // replaced device buffers, freed with the group
    BinWork* d_items[4] = {nullptr, nullptr, nullptr, nullptr};
    int n_items[4] = {0, 0, 0, 0};
    int32_t* d_block_start = nullptr;
    int total_blocks = 0;
    CorrectArgs* d_ctable = nullptr;
    int32_t* d_cstart = nullptr;
    int n_corr = 0, corr_blocks = 0;
    int32_t* d_coef_start = nullptr;
    int coef_blocks = 0;
    // per-step update argument table
    UpdateArgs* d_table = nullptr;
    size_t table_cap = 0;
    StagedUpload ring[GROUP_RING];
    int ring_pos = 0;
    std::string err;
};

Group* group_create(int device) {
    Group* g = new Group();
    g->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return nullptr;
    }
    for (int q = 0; q < GROUP_RING; ++q)
        (void)hipEventCreateWithFlags(&g->ring[q].ev, hipEventDisableTiming);
    return g;
}

void group_destroy(Group* g) {
    if (!g) return;
    (void)hipStreamSynchronize(g->stream);
    if (g->d_static) (void)hipFree(g->d_static);
    if (g->d_table) (void)hipFree(g->d_table);
    for (void* p : g->retired) (void)hipFree(p);
    for (int q = 0; q < GROUP_RING; ++q) {
        if (g->ring[q].host) (void)hipHostFree(g->ring[q].host);
        (void)hipEventDestroy(g->ring[q].ev);
    }
    (void)hipStreamDestroy(g->stream);
    delete g;
}

const char* group_error(const Group* g) { return g->err.c_str(); }

const dtx_config* handle_config(const dtx_handle* h) { return &h->cfg; }

bool handle_has_vectors(const dtx_handle* h) { return h->vectors_set; }

int binary_host_begin(dtx_handle* h, const char* who) {
    if (h->synth_binary)
        return fail(h, DTX_ERR_STATE, "%s: cannot add host channels to a synthetic operator",
                    who);
    if (h->loss_host.empty()) h->loss_host.assign((size_t)h->N * h->N, 0.0);
    return DTX_OK;
}

BinaryHostView binary_host_channel(dtx_handle* h, int s_out, int s1, int s2) {
    HostChannel& ch = host_channel(h, s_out, s1, s2);
    return BinaryHostView{ch.S.data(), ch.No, ch.N1, ch.N2, ch.kind == CH_TRI};
}

double* binary_host_loss(dtx_handle* h) { return h->loss_host.data(); }

void binary_host_end(dtx_handle* h, bool loss_nonzero) {
    h->loss_nonzero = h->loss_nonzero || loss_nonzero;
    h->binary = true;
    h->binary_dirty = true;
}

int handle_fail(dtx_handle* h, int code, const char* msg) { return fail(h, code, "%s", msg); }

const int32_t* handle_species_offsets(const dtx_handle* h) { return h->sp_off.data(); }

hipStream_t group_stream(const Group* g) { return g->stream; }

int group_sync(Group* g) {
    hipError_t e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        g->err = std::string("group stream: ") + hipGetErrorString(e);
        return DTX_ERR_HIP;
    }
    return DTX_OK;
}

// A device buffer of at least `bytes`, grown geometrically; the buffer it replaces may still be
// read by launches on the group stream, so it is kept until the group is destroyed.
static int grow_device(Group* g, void** buf, size_t* cap, size_t bytes) {
    if (bytes <= *cap && *buf) return DTX_OK;
    const size_t want = std::max(bytes, 2 * *cap) + 4096;
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        g->err = "group table allocation failed";
        return DTX_ERR_OOM;
    }
    if (*buf) g->retired.push_back(*buf);
    *buf = p;
    *cap = want;
    return DTX_OK;
}

// Copy `bytes` from host memory to dst on the group stream through the next pinned staging
// buffer of the ring (waiting for that buffer's previous copy to have run).
static int staged_upload(Group* g, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return DTX_OK;
    StagedUpload& u = g->ring[g->ring_pos++ % GROUP_RING];
    if (u.used && hipEventSynchronize(u.ev) != hipSuccess) {
        g->err = "staging event failed";
        return DTX_ERR_HIP;
    }
    if (bytes > u.cap) {
        if (u.host) (void)hipHostFree(u.host);
        u.host = nullptr;
        u.cap = 0;
        const size_t want = std::max(bytes, (size_t)1 << 16);
        if (hipHostMalloc(&u.host, want, hipHostMallocDefault) != hipSuccess) {
            g->err = "pinned staging allocation failed";
            return DTX_ERR_OOM;
        }
        u.cap = want;
    }
    memcpy(u.host, src, bytes);
    hipError_t e = hipMemcpyAsync(dst, u.host, bytes, hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess) e = hipEventRecord(u.ev, g->stream);
    u.used = true;
    if (e != hipSuccess) {
        g->err = std::string("group table upload: ") + hipGetErrorString(e);
        return DTX_ERR_HIP;
    }
    return DTX_OK;
}

// Make a handle ready for grouped steps from any thread: operators packed and work tables
// built, nothing pending on its own stream.
int group_prepare(dtx_handle* h) {
    if (h->method != DTX_METHOD_FORWARD_EULER)
        return fail(h, DTX_ERR_UNSUPPORTED, "grouped stepping supports ForwardEuler only");
    int rc = ensure_ready(h);
    if (rc) return rc;
    HIP_CHECK(h, stream_sync(h));
    return DTX_OK;
}

// Set the members (each prepared by group_prepare, or already a member).  Enqueues the new
// tables on the group stream; does not wait for the device.
int group_set(Group* g, dtx_handle* const* hs, int count) {
    g->members.assign(hs, hs + count);
    std::vector<BinWork> lists[4];  // cell tiles of 16, of 64, wide items of 2 and 4 k tiles
    std::vector<int32_t> starts, cstart, coef_start;
    std::vector<CorrectArgs> ctable;
    g->total_blocks = 0;
    g->corr_blocks = 0;
    g->coef_blocks = 0;
    for (int m = 0; m < count; ++m) {
        dtx_handle* h = hs[m];
        if (h->method != DTX_METHOD_FORWARD_EULER) {
            g->err = "grouped stepping supports ForwardEuler members only";
            return DTX_ERR_UNSUPPORTED;
        }
        if (m == 0) g->precision = h->precision;
        if (h->precision != g->precision) {
            g->err = "grouped members must share one contraction precision";
            return DTX_ERR_UNSUPPORTED;
        }
        if (!h->vectors_set || (h->binary && h->binary_dirty)) {
            g->err = "grouped member is not prepared";
            return DTX_ERR_STATE;
        }
        const int q = h->wide ? 1 + h->wide / 2 : (h->tc <= 16 ? 0 : 1);
        if (h->binary) lists[q].insert(lists[q].end(), h->items_host.begin(), h->items_host.end());
        starts.push_back(g->total_blocks);
        g->total_blocks += h->C * ((h->N + UPDATE_STATES - 1) / UPDATE_STATES);
        if (h->binary && h->nitems > 0 && h->correct_active) {
            ctable.push_back(make_correct_args(h));
            cstart.push_back(g->corr_blocks);
            g->corr_blocks += h->C * h->S * h->zb;
        }
        coef_start.push_back(g->coef_blocks);
        if (h->binary && h->nitems > 0 && h->correct_active && h->d_coef)
            g->coef_blocks += (int)(((int64_t)h->C * h->S + UPDATE_THREADS / 64 - 1) /
                                    (UPDATE_THREADS / 64));
    }
    g->n_corr = (int)ctable.size();
    // The narrow shapes go into one launch (list 0): GEMV items of single-cell requests and
    // 16-cell items stream their operator (HBM-bound) while 32- and 64-cell items are
    // MFMA-bound, so the two lists are interleaved in proportion to run side by side.  Each
    // member's items keep their order.
    {
        std::vector<BinWork> mixed;
        mixed.reserve(lists[0].size() + lists[1].size());
        const size_t n0 = lists[0].size(), n1 = lists[1].size();
        size_t i0 = 0, i1 = 0;
        while (i0 < n0 || i1 < n1) {
            if (i1 >= n1 || (i0 < n0 && i0 * n1 <= i1 * n0)) mixed.push_back(lists[0][i0++]);
            else mixed.push_back(lists[1][i1++]);
        }
        lists[0].swap(mixed);
        lists[1].clear();
    }
    // one host image of the static tables, 256-byte aligned parts
    std::vector<uint8_t> img;
    auto put = [&](const void* src, size_t bytes) {
        const size_t off = (img.size() + 255) / 256 * 256;
        img.resize(off + bytes);
        if (bytes) memcpy(img.data() + off, src, bytes);
        return off;
    };
    size_t off_items[4], off_start, off_ctable, off_cstart, off_coef;
    for (int q = 0; q < 4; ++q) {
        g->n_items[q] = (int)lists[q].size();
        off_items[q] = put(lists[q].data(), lists[q].size() * sizeof(BinWork));
    }
    off_start = put(starts.data(), starts.size() * 4);
    off_ctable = put(ctable.data(), ctable.size() * sizeof(CorrectArgs));
    off_cstart = put(cstart.data(), cstart.size() * 4);
    off_coef = put(coef_start.data(), coef_start.size() * 4);
    int rc = grow_device(g, &g->d_static, &g->static_cap, img.size());
    if (rc) return rc;
    uint8_t* base = (uint8_t*)g->d_static;
    for (int q = 0; q < 4; ++q) g->d_items[q] = (BinWork*)(base + off_items[q]);
    g->d_block_start = (int32_t*)(base + off_start);
    g->d_ctable = (CorrectArgs*)(base + off_ctable);
    g->d_cstart = (int32_t*)(base + off_cstart);
    g->d_coef_start = (int32_t*)(base + off_coef);
    rc = staged_upload(g, g->d_static, img.data(), img.size());
    if (rc) return rc;
    return grow_device(g, (void**)&g->d_table, &g->table_cap,
                       (size_t)std::max(count, 1) * sizeof(UpdateArgs));
}

// Enqueue one ForwardEuler step (dt[m], no save clipping) of every member.
int group_step(Group* g, const double* dt, int verbose) {
    const int count = (int)g->members.size();
    if (count == 0) return DTX_OK;
    std::vector<UpdateArgs> table((size_t)count);
    for (int m = 0; m < count; ++m) {
        dtx_handle* h = g->members[m];
        table[(size_t)m] = make_args(h, dt[m] / h->cfg.dt0, verbose, 0.0, dt[m], 0, DTX_FLUX_F);
    }
    int rc = staged_upload(g, g->d_table, table.data(), table.size() * sizeof(UpdateArgs));
    if (rc) return rc;
    hipError_t e = launch_binary_mixed(g->precision, g->d_items[0], g->n_items[0], g->stream);
    for (int q = 2; q < 4 && e == hipSuccess; ++q)
        e = launch_binary_wide(g->precision, 2 * (q - 1), g->d_items[q], g->n_items[q], g->stream);
    if (e == hipSuccess)
        e = launch_gain_correct_group(g->d_ctable, g->d_cstart, g->n_corr, g->corr_blocks,
                                      g->stream);
    if (e == hipSuccess)
        e = launch_coef_group(g->d_table, g->d_coef_start, count, g->coef_blocks, g->stream);
    // members with emission blocks: their GEMV, one launch each
    for (int m = 0; m < count && e == hipSuccess; ++m)
        e = launch_emission(table[(size_t)m], g->stream);
    if (e == hipSuccess)
        e = launch_euler_update_group(g->d_table, g->d_block_start, count, g->total_blocks,
                                      g->stream);
    if (e != hipSuccess) {
        g->err = std::string("group step: ") + hipGetErrorString(e);
        return DTX_ERR_HIP;
    }
    for (int m = 0; m < count; ++m) {
        dtx_handle* h = g->members[m];
        h->step_notes[(size_t)(h->step_counter % STATUS_RING)] = StepNote{(int8_t)verbose, 0.0};
        h->cur ^= 1;
        h->step_counter++;
        h->reported = h->step_counter;  // grouped steps are not echoed
    }
    return DTX_OK;
}

}  // namespace dtx

// ------------------------------------------------------------------------ C ABI ----------

extern "C" {

int dtx_create(const dtx_config* cfg, dtx_handle** out) {
    if (!cfg || !out) return fail(nullptr, DTX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->num_species < 1 || cfg->num_species > DTX_MAX_SPECIES)
        return fail(nullptr, DTX_ERR_INVALID, "num_species must be in 1..%d", DTX_MAX_SPECIES);
    if (cfg->x_num < 1 || cfg->y_num < 1 || cfg->z_num < 1)
        return fail(nullptr, DTX_ERR_INVALID, "x_num, y_num, z_num must be positive");
    if (cfg->precision < DTX_PREC_FP64 || cfg->precision > DTX_PREC_BF16X3)
        return fail(nullptr, DTX_ERR_INVALID, "unknown precision %d", cfg->precision);
    if (cfg->method != DTX_METHOD_FORWARD_EULER &&
        cfg->method != DTX_METHOD_FORWARD_SYMPLECTIC_EULER)
        return fail(nullptr, DTX_ERR_INVALID, "unknown method %d", cfg->method);
    if (!(cfg->dt0 > 0.0)) return fail(nullptr, DTX_ERR_INVALID, "dt0 must be positive");
    std::unique_ptr<dtx_handle> h(new dtx_handle());
    h->cfg = *cfg;
    h->S = cfg->num_species;
    h->sp_off.assign(h->S + 1, 0);
    h->sp_size.assign(h->S, 0);
    int64_t N = 0;
    for (int s = 0; s < h->S; ++s) {
        if (cfg->px_num[s] < 1 || cfg->py_num[s] < 1 || cfg->pz_num[s] < 1)
            return fail(nullptr, DTX_ERR_INVALID, "species %d has an empty momentum grid", s);
        h->sp_off[s] = (int32_t)N;
        h->sp_size[s] = cfg->px_num[s] * cfg->py_num[s] * cfg->pz_num[s];
        N += h->sp_size[s];
    }
    h->sp_off[h->S] = (int32_t)N;
    // The update kernels keep two N-length FP64 rows of a cell in LDS (plus 3 KB of
    // reduction scratch) within the 64 KB per-workgroup budget.
    if (N > 3840)
        return fail(nullptr, DTX_ERR_UNSUPPORTED, "N = %lld states per cell exceeds 3840",
                    (long long)N);
    h->N = (int32_t)N;
    const int64_t C = (int64_t)cfg->x_num * cfg->y_num * cfg->z_num;
    if (C > (1 << 30)) return fail(nullptr, DTX_ERR_UNSUPPORTED, "too many cells");
    h->C = (int32_t)C;
    h->n = N * C;
    if (h->n > 0x7FFFFFFF)
        return fail(nullptr, DTX_ERR_UNSUPPORTED, "n = %lld exceeds the 32-bit column index",
                    (long long)h->n);
    h->precision = cfg->precision;
    h->method = cfg->method;
    // one cell: GEMV items that stream the operator (MFMA tiles would be 15/16 padding)
    // 17-32 cells take 32-cell tiles: a 64-cell tile would spend half its MFMAs on padding
    h->tc = h->C == 1 ? 1 : (h->C <= 16) ? 16 : (h->C <= 32) ? 32 : 64;
    // wide gain items above 64 cells: split-BF16 with four k tiles per item from 512 cells on,
    // where the launch still has enough items for one eight-wave workgroup per CU.  FP64 with
    // two k tiles up to 256 cells: with many cells the 64-cell items, four workgroups per CU,
    // measured faster (the wide FP64 item holds two), and with few the halved operator reads win
    // (the FP32 contraction keeps 64-cell tiles)
    if (h->precision == DTX_PREC_BF16X3 && h->C > 64) h->wide = h->C >= 512 ? 4 : 2;
    if (h->precision == DTX_PREC_FP64 && h->C > 64 && h->C <= 256) h->wide = 2;
    const int cell_align = h->wide ? WIDE : std::max(h->tc, 16);
    h->ldc = (h->C + cell_align - 1) / cell_align * cell_align;
    h->sp_pad_off.assign(h->S + 1, 0);
    for (int s = 0; s < h->S; ++s)
        h->sp_pad_off[s + 1] = h->sp_pad_off[s] +
                               (h->sp_size[s] + SPECIES_ALIGN - 1) / SPECIES_ALIGN * SPECIES_ALIGN;
    h->Np = h->sp_pad_off[h->S];
    if (h->Np > 65535)
        return fail(nullptr, DTX_ERR_UNSUPPORTED, "padded state count exceeds 65535");
    h->pad_index.assign((size_t)N, 0);
    h->unpad.assign((size_t)h->Np, -1);
    for (int s = 0; s < h->S; ++s)
        for (int k = 0; k < h->sp_size[s]; ++k) {
            h->pad_index[(size_t)(h->sp_off[s] + k)] = h->sp_pad_off[s] + k;
            h->unpad[(size_t)(h->sp_pad_off[s] + k)] = h->sp_off[s] + k;
        }
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess)
        return fail(nullptr, DTX_ERR_HIP, "hipSetDevice(%d): %s", cfg->device,
                    hipGetErrorString(e));
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return fail(nullptr, DTX_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, cfg->device) ==
                hipSuccess && cus > 0)
            h->num_cu = cus;
    }
    dtx_handle* hp = h.get();
    int rc = 0;
    const int64_t n = h->n;
    rc |= dalloc(hp, &h->d_f[0], n);
    rc |= dalloc(hp, &h->d_f[1], n);
    rc |= dalloc(hp, &h->d_fT, (int64_t)h->ldc * h->Np);
    if (h->precision == DTX_PREC_FP32) rc |= dalloc(hp, &h->d_fT32, (int64_t)h->ldc * h->Np);
    if (h->precision == DTX_PREC_BF16X3) {
        rc |= dalloc(hp, &h->d_fhi, (int64_t)h->ldc * h->Np);
        rc |= dalloc(hp, &h->d_flo, (int64_t)h->ldc * h->Np);
    }
    rc |= dalloc(hp, &h->d_pad_index, N);
    rc |= dalloc(hp, &h->d_unpad, h->Np);
    rc |= dalloc(hp, &h->d_invap, n);
    rc |= dalloc(hp, &h->d_dfinj, n);
    rc |= dalloc(hp, &h->d_vol, C);
    rc |= dalloc(hp, &h->d_finit, n);
    rc |= dalloc(hp, &h->d_state_species, N);
    rc |= dalloc(hp, &h->d_species_off, h->S + 1);
    rc |= dalloc(hp, &h->d_moments, C * h->S * 2);
    const int64_t nseg = (N + UPDATE_STATES - 1) / UPDATE_STATES;  // update workgroups per cell
    // per-block partials, then the fold partials of the finalize kernel
    const int64_t nfold =
        (std::max(C * nseg, C * (int64_t)h->S) + UPDATE_THREADS - 1) / UPDATE_THREADS;
    h->bp_words = (C * nseg + nfold) * BP_WORDS;
    h->mp_words = C * nseg * h->S * 2;
    rc |= dalloc(hp, &h->d_block_part, (int64_t)FIN_BATCH * h->bp_words);
    rc |= dalloc(hp, &h->d_mom_part, h->mp_words);
    rc |= dalloc(hp, &h->d_fin_rec, (int64_t)FIN_BATCH * FIN_WORDS);
    rc |= dalloc(hp, &h->d_status, (int64_t)STATUS_RING * STATUS_WORDS);
    if (h->method == DTX_METHOD_FORWARD_SYMPLECTIC_EULER) {
        rc |= dalloc(hp, &h->d_fmom, n);
        rc |= dalloc(hp, &h->d_dfmom, n);
        rc |= dalloc(hp, &h->d_dpflux, n);
        rc |= dalloc(hp, &h->d_dbin, n);
        rc |= dalloc(hp, &h->d_demi, n);
    }
    if (rc) {
        std::string msg = h->err;
        dtx_destroy(h.release());
        return fail(nullptr, DTX_ERR_OOM, "%s", msg.c_str());
    }
    std::vector<int32_t> ss((size_t)N);
    for (int s = 0; s < h->S; ++s)
        for (int k = 0; k < h->sp_size[s]; ++k) ss[(size_t)(h->sp_off[s] + k)] = s;
    rc = upload(hp, h->d_state_species, ss.data(), N);
    if (!rc) rc = upload(hp, h->d_pad_index, h->pad_index.data(), N);
    if (!rc) rc = upload(hp, h->d_unpad, h->unpad.data(), h->Np);
    if (!rc) rc = upload(hp, h->d_species_off, h->sp_off.data(), h->S + 1);
    if (!rc) {
        e = hipMemsetAsync(h->d_fT, 0, (size_t)h->ldc * h->Np * 8, h->stream);
        if (e == hipSuccess && h->d_fT32)
            e = hipMemsetAsync(h->d_fT32, 0, (size_t)h->ldc * h->Np * 4, h->stream);
        if (e == hipSuccess && h->d_fhi)
            e = hipMemsetAsync(h->d_fhi, 0, (size_t)h->ldc * h->Np * 2, h->stream);
        if (e == hipSuccess && h->d_flo)
            e = hipMemsetAsync(h->d_flo, 0, (size_t)h->ldc * h->Np * 2, h->stream);
        if (e == hipSuccess) e = hipMemsetAsync(h->d_status, 0,
                                                (size_t)STATUS_RING * STATUS_WORDS * 8,
                                                h->stream);
        if (e == hipSuccess) e = hipMemsetAsync(h->d_moments, 0, (size_t)C * h->S * 16,
                                                h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(hp, DTX_ERR_HIP, "init: %s", hipGetErrorString(e));
    }
    if (rc) {
        std::string msg = h->err;
        dtx_destroy(h.release());
        return fail(nullptr, rc, "%s", msg.c_str());
    }
    h->step_notes.assign(STATUS_RING, StepNote{});
    h->ring_host.assign((size_t)STATUS_RING * STATUS_WORDS, 0.0);
    h->scratch = scratch_address(cfg->device);  // device current (hipSetDevice above)
    *out = h.release();
    return DTX_OK;
}

void dtx_destroy(dtx_handle* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto& ev : h->events) {
        (void)hipEventDestroy(ev.start);
        (void)hipEventDestroy(ev.stop);
    }
    release_binary_device(h);
    for (void* p : h->allocations) (void)hipFree(p);
    h->allocations.clear();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* dtx_last_error(const dtx_handle* h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_create_error.c_str();
}

int dtx_set_flux_csc(dtx_handle* h, int32_t which, int64_t n, const int64_t* colptr,
                     const int64_t* rowval, const double* nzval) {
    if (!h) return DTX_ERR_INVALID;
    if (which < DTX_FLUX_F || which > DTX_FLUX_P)
        return fail(h, DTX_ERR_INVALID, "unknown flux matrix %d", which);
    if (n != h->n)
        return fail(h, DTX_ERR_INVALID, "flux matrix is %lld x %lld, expected n = %lld",
                    (long long)n, (long long)n, (long long)h->n);
    if (!colptr || (colptr[n] > 1 && (!rowval || !nzval)))
        return fail(h, DTX_ERR_INVALID, "null CSC array");
    return set_ell(h, h->ell[which], n, colptr, rowval, nzval);
}

int dtx_synth_flux(dtx_handle* h, const dtx_flux_synth* sp) {
    if (!h) return DTX_ERR_INVALID;
    if (!sp || !sp->xr || !sp->yr || !sp->zr || !sp->pr || !sp->ur || !sp->hr)
        return fail(h, DTX_ERR_INVALID, "null bounds array");
    if ((sp->axes & ~3) || !(sp->sync_K >= 0.0) || !std::isfinite(sp->sync_K))
        return fail(h, DTX_ERR_INVALID, "invalid flux spec (axes %d, sync_K %g)", sp->axes,
                    sp->sync_K);
    const dtx_config& cfg = h->cfg;
    FluxSynthArgs a{};
    // one buffer of bounds: x, y, z, then p, u, phi of every species
    std::vector<double> host;
    host.insert(host.end(), sp->xr, sp->xr + cfg.x_num + 1);
    host.insert(host.end(), sp->yr, sp->yr + cfg.y_num + 1);
    host.insert(host.end(), sp->zr, sp->zr + cfg.z_num + 1);
    int64_t np = 0, nu = 0, nh = 0;
    for (int s = 0; s < h->S; ++s) {
        a.pr_off[s] = (int32_t)np;
        a.ur_off[s] = (int32_t)nu;
        a.hr_off[s] = (int32_t)nh;
        np += cfg.px_num[s] + 1;
        nu += cfg.py_num[s] + 1;
        nh += cfg.pz_num[s] + 1;
        a.px[s] = cfg.px_num[s];
        a.py[s] = cfg.py_num[s];
        a.sp_off[s] = h->sp_off[s];
        a.mass[s] = sp->mass[s];
        a.charge[s] = sp->charge[s];
    }
    a.sp_off[h->S] = h->sp_off[h->S];
    const int64_t o_p = (int64_t)host.size();
    host.insert(host.end(), sp->pr, sp->pr + np);
    const int64_t o_u = (int64_t)host.size();
    host.insert(host.end(), sp->ur, sp->ur + nu);
    const int64_t o_h = (int64_t)host.size();
    host.insert(host.end(), sp->hr, sp->hr + nh);
    double* d_bounds = nullptr;
    unsigned long long* d_stats = nullptr;
    const int64_t nb = (int64_t)host.size();
    int rc = dalloc(h, &d_bounds, nb);
    if (!rc) rc = dalloc(h, &d_stats, 2);
    if (!rc) rc = upload(h, d_bounds, host.data(), nb);
    a.xr = d_bounds;
    a.yr = a.xr + cfg.x_num + 1;
    a.zr = a.yr + cfg.y_num + 1;
    a.pr = d_bounds + o_p;
    a.ur = d_bounds + o_u;
    a.hr = d_bounds + o_h;
    a.dt0 = cfg.dt0;
    a.sync_K = sp->sync_K;
    a.n = h->n;
    a.N = h->N;
    a.S = h->S;
    a.x_num = cfg.x_num;
    a.y_num = cfg.y_num;
    a.z_num = cfg.z_num;
    a.axes = sp->axes;
    a.periodic = sp->periodic ? 1 : 0;
    a.stats = d_stats;
    // the matrices the method reads; the first count pass also writes 1/Ap and Vol.  Every exit
    // below frees the temporaries, and a matrix is marked present only when both its arrays
    // exist.
    const int which_list[2] = {DTX_FLUX_X, DTX_FLUX_P};
    const bool euler = cfg.method == DTX_METHOD_FORWARD_EULER;
    const int nwhich = euler ? 1 : 2;
    auto hip_fail = [&](hipError_t e, const char* what) {
        return fail(h, DTX_ERR_HIP, "dtx_synth_flux: %s: %s", what, hipGetErrorString(e));
    };
    for (int q = 0; q < nwhich && !rc; ++q) {
        a.which = euler ? DTX_FLUX_F : which_list[q];
        a.width = 0;
        a.inv_ap = q == 0 ? h->d_invap : nullptr;
        a.vol = q == 0 ? h->d_vol : nullptr;
        unsigned long long st[2] = {0, 0};
        rc = upload(h, d_stats, st, 2);
        if (rc) break;
        hipError_t e = launch_synth_flux(a, h->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(st, d_stats, sizeof(st), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = stream_sync(h);
        if (e != hipSuccess) {
            rc = hip_fail(e, "count pass");
            break;
        }
        EllMatrix& E = h->ell[a.which];
        release_ell(h, E);
        const int32_t width = (int32_t)st[0];
        if (width == 0) continue;
        int32_t* dcol = nullptr;
        double* dval = nullptr;
        rc = dalloc(h, &dcol, (int64_t)width * h->n);
        if (!rc) rc = dalloc(h, &dval, (int64_t)width * h->n);
        if (!rc) {
            a.width = width;
            a.col = dcol;
            a.val = dval;
            a.inv_ap = nullptr;
            a.vol = nullptr;
            e = launch_synth_flux(a, h->stream);
            if (e == hipSuccess) e = stream_sync(h);
            if (e != hipSuccess) rc = hip_fail(e, "fill pass");
        }
        if (rc) {
            dfree(h, &dcol, (int64_t)width * h->n);
            dfree(h, &dval, (int64_t)width * h->n);
            break;
        }
        E.col = dcol;
        E.val = dval;
        E.n = h->n;
        E.width = width;
        E.nnz = (int64_t)st[1];
    }
    dfree(h, &d_bounds, nb);
    dfree(h, &d_stats, 2);
    if (!rc) h->geometry_set = true;
    return rc;
}

int dtx_set_binary_dense(dtx_handle* h, int64_t N, const double* M) {
    if (!h) return DTX_ERR_INVALID;
    if (N != h->N)
        return fail(h, DTX_ERR_INVALID, "M_Bin is (%lld^2 x %lld), expected N = %d",
                    (long long)N, (long long)N, h->N);
    if (!M) return fail(h, DTX_ERR_INVALID, "null M_Bin");
    h->hch.clear();
    h->synth_binary = false;
    h->loss_host.assign((size_t)N * N, 0.0);
    h->loss_nonzero = false;
    const int64_t NN = N * N;
    // Fold M[k,i,j] and M[k,j,i] into one pair coefficient; entries with k == i or k == j
    // (the loss diagonal written by LossMatrix_to_M_Bin!) go to the FP64 loss matrix.
    for (int so = 0; so < h->S; ++so)
        for (int s1 = 0; s1 < h->S; ++s1)
            for (int s2 = s1; s2 < h->S; ++s2) {
                HostChannel& ch = host_channel(h, so, s1, s2);
                const int o = h->sp_off[so];
                for (int64_t p = 0; p < ch.npairs; ++p) {
                    int gi, gj;
                    pair_of(ch, p, &gi, &gj);
                    const double* Mij = M + (int64_t)gi * N + (int64_t)gj * NN;
                    const double* Mji = M + (int64_t)gj * N + (int64_t)gi * NN;
                    double* dst = ch.S.data() + (size_t)p * ch.No;
                    for (int k = 0; k < ch.No; ++k) {
                        const int gk = o + k;
                        double v = (gi == gj) ? Mij[gk] : Mij[gk] + Mji[gk];
                        if (v != 0.0 && (gk == gi || gk == gj)) {
                            const int other = (gk == gi) ? gj : gi;
                            h->loss_host[(size_t)gk + (size_t)N * other] += v;
                            h->loss_nonzero = true;
                            v = 0.0;
                        }
                        dst[k] = v;
                    }
                }
            }
    h->binary = true;
    h->binary_dirty = true;
    return DTX_OK;
}

int dtx_set_binary_csc(dtx_handle* h, int64_t N, const int64_t* colptr, const int64_t* rowval,
                       const double* nzval) {
    if (!h) return DTX_ERR_INVALID;
    if (N != h->N)
        return fail(h, DTX_ERR_INVALID, "M_Bin is (%lld^2 x %lld), expected N = %d",
                    (long long)N, (long long)N, h->N);
    if (!colptr || (colptr[N] > 1 && (!rowval || !nzval)))
        return fail(h, DTX_ERR_INVALID, "null CSC array");
    const int64_t NN = N * N;
    for (int64_t j = 0; j < N; ++j) {  // validate before touching the handle's operator
        const int64_t k0 = colptr[j] - 1, k1 = colptr[j + 1] - 1;
        if (k0 < 0 || k1 < k0)
            return fail(h, DTX_ERR_INVALID, "colptr is not a valid 1-based CSC pointer");
        for (int64_t q = k0; q < k1; ++q)
            if (rowval[q] < 1 || rowval[q] > NN)
                return fail(h, DTX_ERR_INVALID, "rowval out of range");
    }
    h->hch.clear();
    h->synth_binary = false;
    h->loss_host.assign((size_t)N * N, 0.0);
    h->loss_nonzero = false;
    // species of every state, and a pair's position in its channel
    std::vector<int> sp((size_t)N);
    for (int s = 0; s < h->S; ++s)
        for (int k = h->sp_off[s]; k < h->sp_off[s + 1]; ++k) sp[(size_t)k] = s;
    for (int so = 0; so < h->S; ++so)  // every channel exists, as the dense form builds them
        for (int s1 = 0; s1 < h->S; ++s1)
            for (int s2 = s1; s2 < h->S; ++s2) (void)host_channel(h, so, s1, s2);
    // Stored entry M[k,i,j] (row (i-1)N + k, column j): entries with k == i or k == j go to the
    // FP64 loss matrix at (k, the other index); the others to the pair (min, max) of their
    // channel.  Each pair row receives its two halves M[k,i,j] and M[k,j,i] in CSC order,
    // which gives the dense form's sum M[k,lo,hi] + M[k,hi,lo] bit for bit (0 + a + b).
    for (int64_t j = 0; j < N; ++j)
        for (int64_t q = colptr[j] - 1; q < colptr[j + 1] - 1; ++q) {
            const double v = nzval[q];
            if (v == 0.0) continue;
            const int64_t r = rowval[q] - 1;
            const int64_t i = r / N, k = r % N;
            if (k == i || k == j) {
                const int64_t other = (k == i) ? j : i;
                h->loss_host[(size_t)k + (size_t)N * other] += v;
                h->loss_nonzero = true;
                continue;
            }
            int64_t lo = i, hi = j;
            if (sp[(size_t)i] > sp[(size_t)j] || (sp[(size_t)i] == sp[(size_t)j] && i > j)) {
                lo = j;
                hi = i;
            }
            const int so = sp[(size_t)k], s1 = sp[(size_t)lo], s2 = sp[(size_t)hi];
            HostChannel& ch = host_channel(h, so, s1, s2);
            const int a = (int)(lo - h->sp_off[s1]), b = (int)(hi - h->sp_off[s2]);
            const int64_t p = (s1 == s2) ? tri_index(a, b, ch.N1) : (int64_t)a * ch.N2 + b;
            double& dst = ch.S[(size_t)p * ch.No + (size_t)(k - h->sp_off[so])];
            dst = dst + v;
        }
    h->binary = true;
    h->binary_dirty = true;
    return DTX_OK;
}

int dtx_add_binary_gain(dtx_handle* h, int32_t s_out, int32_t s1, int32_t s2, const double* G) {
    if (!h) return DTX_ERR_INVALID;
    if (s_out < 0 || s_out >= h->S || s1 < 0 || s1 >= h->S || s2 < 0 || s2 >= h->S)
        return fail(h, DTX_ERR_INVALID, "species index out of range");
    if (!G) return fail(h, DTX_ERR_INVALID, "null gain block");
    if (h->synth_binary) {
        release_binary_device(h);
        h->synth_binary = false;
    }
    if (h->loss_host.empty()) h->loss_host.assign((size_t)h->N * h->N, 0.0);
    const int No = h->sp_size[s_out], N1 = h->sp_size[s1], N2 = h->sp_size[s2];
    const int a = std::min(s1, s2), b = std::max(s1, s2);
    HostChannel& ch = host_channel(h, s_out, a, b);
    const int oa = h->sp_off[a], ob = h->sp_off[b];
    for (int64_t p = 0; p < ch.npairs; ++p) {
        int gi, gj;
        pair_of(ch, p, &gi, &gj);
        const int i = gi - oa, j = gj - ob;  // i in a, j in b
        double* dst = ch.S.data() + (size_t)p * No;
        const double* g1;
        const double* g2 = nullptr;
        if (s1 == s2) {  // TRI: G[k,i,j] + G[k,j,i]  (i <= j)
            g1 = G + (int64_t)No * (i + (int64_t)N1 * j);
            if (i != j) g2 = G + (int64_t)No * (j + (int64_t)N1 * i);
        } else if (s1 == a) {
            g1 = G + (int64_t)No * (i + (int64_t)N1 * j);
        } else {  // given as (s_out; b, a): G[k, j, i]
            g1 = G + (int64_t)No * (j + (int64_t)N1 * i);
        }
        for (int k = 0; k < No; ++k) dst[k] += g2 ? g1[k] + g2[k] : g1[k];
    }
    (void)N2;
    h->binary = true;
    h->binary_dirty = true;
    return DTX_OK;
}

int dtx_add_binary_loss(dtx_handle* h, int32_t s1, int32_t s2, const double* L) {
    if (!h) return DTX_ERR_INVALID;
    if (s1 < 0 || s1 >= h->S || s2 < 0 || s2 >= h->S)
        return fail(h, DTX_ERR_INVALID, "species index out of range");
    if (!L) return fail(h, DTX_ERR_INVALID, "null loss block");
    if (h->synth_binary)
        return fail(h, DTX_ERR_STATE, "cannot mix host channels with a synthetic operator");
    if (h->loss_host.empty()) h->loss_host.assign((size_t)h->N * h->N, 0.0);
    const int N1 = h->sp_size[s1], N2 = h->sp_size[s2];
    const int o1 = h->sp_off[s1], o2 = h->sp_off[s2];
    for (int c = 0; c < N2; ++c)
        for (int a = 0; a < N1; ++a) {
            const double v = L[a + (int64_t)N1 * c];
            if (v != 0.0) {
                h->loss_host[(size_t)(o1 + a) + (size_t)h->N * (o2 + c)] -= v;
                h->loss_nonzero = true;
            }
        }
    h->binary = true;
    h->binary_dirty = true;
    return DTX_OK;
}

int dtx_set_binary_domain(dtx_handle* h, int64_t count, const int64_t* domain) {
    if (!h) return DTX_ERR_INVALID;
    dfree(h, &h->d_active, h->C);
    if (count < 0) return DTX_OK;
    std::vector<uint8_t> act((size_t)h->C, 0);
    for (int64_t d = 0; d < count; ++d) {
        if (domain[d] < 0 || domain[d] >= h->C)
            return fail(h, DTX_ERR_INVALID, "domain cell %lld out of range",
                        (long long)domain[d]);
        act[(size_t)domain[d]] = 1;
    }
    int rc = dalloc(h, &h->d_active, h->C);
    if (rc) return rc;
    return upload(h, h->d_active, act.data(), h->C);
}

int dtx_set_emission_blocks(dtx_handle* h, const double* blocks) {
    if (!h) return DTX_ERR_INVALID;
    if (!blocks) return fail(h, DTX_ERR_INVALID, "null emission blocks");
    if (h->N > EMI_MAX_N)
        return fail(h, DTX_ERR_UNSUPPORTED, "emission blocks need at most %d states per cell",
                    EMI_MAX_N);
    const int64_t count = (int64_t)h->C * h->N * h->N;
    if (!h->d_emi) {
        int rc = dalloc(h, &h->d_emi, count);
        if (rc) return rc;
    }
    if (!h->d_demi) {  // the GEMV's output (symplectic handles hold it already)
        int rc = dalloc(h, &h->d_demi, h->n);
        if (rc) return rc;
    }
    h->emi_q = emission_parts(h->N, h->C, h->num_cu);
    if (h->emi_q > 1 && !h->d_emi_part) {  // the part sums of column ranges
        int rc = dalloc(h, &h->d_emi_part, (int64_t)h->emi_q * h->n);
        if (rc) return rc;
    }
    int rc = upload(h, h->d_emi, blocks, count);
    if (rc) return rc;
    h->emission = true;
    return DTX_OK;
}

int dtx_set_emission_dense(dtx_handle* h, int64_t n, const double* M) {
    if (!h) return DTX_ERR_INVALID;
    if (n != h->n) return fail(h, DTX_ERR_INVALID, "M_Emi must be n x n with n = %lld",
                               (long long)h->n);
    if (!M) return fail(h, DTX_ERR_INVALID, "null M_Emi");
    const int N = h->N, C = h->C;
    std::vector<double> blocks((size_t)C * N * N);
    for (int64_t j = 0; j < n; ++j) {
        const int cj = (int)(j / N);
        for (int64_t i = 0; i < n; ++i) {
            const double v = M[i + n * j];
            const int ci = (int)(i / N);
            if (ci != cj) {
                if (v != 0.0)
                    return fail(h, DTX_ERR_UNSUPPORTED,
                                "M_Emi has a non-zero outside the per-cell diagonal blocks "
                                "(row %lld, column %lld)", (long long)i + 1, (long long)j + 1);
                continue;
            }
            blocks[(size_t)ci * N * N + (size_t)(i % N) + (size_t)N * (j % N)] = v;
        }
    }
    return dtx_set_emission_blocks(h, blocks.data());
}

int dtx_set_emission_csc(dtx_handle* h, int64_t n, const int64_t* colptr, const int64_t* rowval,
                         const double* nzval) {
    if (!h) return DTX_ERR_INVALID;
    if (n != h->n) return fail(h, DTX_ERR_INVALID, "M_Emi must be n x n with n = %lld",
                               (long long)h->n);
    if (!colptr || (colptr[n] > 1 && (!rowval || !nzval)))
        return fail(h, DTX_ERR_INVALID, "null CSC array");
    const int N = h->N, C = h->C;
    std::vector<double> blocks((size_t)C * N * N, 0.0);
    for (int64_t j = 0; j < n; ++j) {
        const int64_t k0 = colptr[j] - 1, k1 = colptr[j + 1] - 1;
        if (k0 < 0 || k1 < k0)
            return fail(h, DTX_ERR_INVALID, "colptr is not a valid 1-based CSC pointer");
        const int64_t cj = j / N;
        for (int64_t q = k0; q < k1; ++q) {
            const int64_t i = rowval[q] - 1;
            if (i < 0 || i >= n) return fail(h, DTX_ERR_INVALID, "rowval out of range");
            const double v = nzval[q];
            if (i / N != cj) {
                if (v != 0.0)
                    return fail(h, DTX_ERR_UNSUPPORTED,
                                "M_Emi has a non-zero outside the per-cell diagonal blocks "
                                "(row %lld, column %lld)", (long long)i + 1, (long long)j + 1);
                continue;
            }
            double& dst = blocks[(size_t)cj * N * N + (size_t)(i % N) + (size_t)N * (j % N)];
            dst = dst + v;
        }
    }
    return dtx_set_emission_blocks(h, blocks.data());
}

int dtx_set_vectors(dtx_handle* h, const double* Ap, const double* Vol, const double* f_init,
                    const double* df_inj) {
    if (!h) return DTX_ERR_INVALID;
    const bool keep_geometry = !Ap && !Vol && h->geometry_set;
    if ((!keep_geometry && (!Ap || !Vol)) || !f_init || !df_inj)
        return fail(h, DTX_ERR_INVALID, "null vector");
    const int64_t n = h->n;
    int rc = DTX_OK;
    if (!keep_geometry) {
        std::vector<double> inv((size_t)n);
        for (int64_t i = 0; i < n; ++i) inv[(size_t)i] = 1.0 / Ap[i];  // 1 ./ Ap_Flux
        rc = upload(h, h->d_invap, inv.data(), n);
        if (!rc) rc = upload(h, h->d_vol, Vol, h->C);
    }
    if (!rc) rc = upload(h, h->d_finit, f_init, n);
    if (!rc) rc = upload(h, h->d_dfinj, df_inj, n);
    if (rc) return rc;
    h->vectors_set = true;
    return dtx_reset(h);
}

int dtx_set_moment_weights(dtx_handle* h, const double* w) {
    if (!h) return DTX_ERR_INVALID;
    if (!w) return fail(h, DTX_ERR_INVALID, "null weights");
    if (!h->d_eweight) {
        int rc = dalloc(h, &h->d_eweight, h->N);
        if (rc) return rc;
    }
    h->eweight_host.assign(w, w + h->N);
    return upload(h, h->d_eweight, w, h->N);
}

int dtx_set_moment_correction(dtx_handle* h, int32_t enable) {
    if (!h) return DTX_ERR_INVALID;
    if (h->binary && !h->binary_dirty)
        return fail(h, DTX_ERR_STATE,
                    "the moment-rate correction must be chosen before the first step");
    h->moment_correction = enable == 2 ? 2 : (enable ? 1 : 0);
    return DTX_OK;
}

int dtx_synth_binary(dtx_handle* h, uint64_t seed, double scale) {
    if (!h) return DTX_ERR_INVALID;
    h->hch.clear();
    h->loss_host.clear();
    release_binary_device(h);
    dfree(h, &h->d_loss, (int64_t)h->N * h->N);
    int rc = dalloc(h, &h->d_loss, (int64_t)h->N * h->N);
    if (rc) return rc;
    HIP_CHECK(h, launch_fill_zero(h->d_loss, (int64_t)h->N * h->N, h->stream));
    uint64_t stream = 0;
    // interactions a + b -> a + b for every species pair a <= b (one gain block per output)
    for (int a = 0; a < h->S; ++a)
        for (int b = a; b < h->S; ++b) {
            const int outs[2] = {a, b};
            const int nout = (a == b) ? 1 : 2;
            for (int q = 0; q < nout; ++q) {
                const int so = outs[q];
                const uint64_t key = splitmix_key(seed, stream++);
                PackedChannel pc;
                pc.s_out = so;
                pc.s1 = a;
                pc.s2 = b;
                pc.kind = (a == b) ? CH_TRI : CH_RECT;
                pc.precision = h->precision;
                pc.No = h->sp_size[so];
                pc.N1 = h->sp_size[a];
                pc.N2 = h->sp_size[b];
                pc.npairs = (a == b) ? (int64_t)pc.N1 * (pc.N1 + 1) / 2 : (int64_t)pc.N1 * pc.N2;
                pc.nktiles = (pc.No + TK - 1) / TK;
                pc.synth = true;
                pc.synth_key = key;
                pc.synth_scale = scale;
                rc = alloc_channel(h, pc, chunk_table(h, pc.kind, a, b));
                h->chans.push_back(pc);
                if (rc) return rc;
                HIP_CHECK(h, launch_synth_gain_packed(h->precision, pc.A, pc.chunks, pc.nchunks,
                                                      pc.No, pc.nktiles, pc.kind, pc.N1, pc.N2,
                                                      h->sp_pad_off[a], h->sp_pad_off[b], key,
                                                      scale, h->stream));
                // the particle of the output species leaves: out a -> loss on a (rows i),
                // out b -> loss on b (rows j); a == b -> one loss over (i, j)
                const int loss_on_first = (a == b) || (so == a);
                HIP_CHECK(h, launch_synth_loss(h->d_loss, h->N, h->sp_off[a], pc.N1,
                                               h->sp_off[b], pc.N2, pc.No, key, scale,
                                               loss_on_first, h->stream));
            }
        }
    HIP_CHECK(h, stream_sync(h));
    h->synth_binary = true;
    h->binary = true;
    h->binary_dirty = true;
    return DTX_OK;
}

int dtx_reset(dtx_handle* h) {
    if (!h) return DTX_ERR_INVALID;
    if (!h->vectors_set) return fail(h, DTX_ERR_STATE, "dtx_set_vectors not called");
    HIP_CHECK(h, stream_sync(h));
    int rc = drain_status(h);
    if (rc) return rc;
    h->cur = 0;
    HIP_CHECK(h, hipMemcpyAsync(h->d_f[0], h->d_finit, (size_t)h->n * 8,
                                hipMemcpyDeviceToDevice, h->stream));
    rc = refresh_fT(h);
    if (rc) return rc;
    HIP_CHECK(h, stream_sync(h));
    h->step_counter = 0;
    h->reported = 0;
    return DTX_OK;
}

int dtx_get_f(dtx_handle* h, double* f) {
    if (!h || !f) return DTX_ERR_INVALID;
    HIP_CHECK(h, hipMemcpyAsync(f, h->d_f[h->cur], (size_t)h->n * 8, hipMemcpyDeviceToHost,
                                h->stream));
    HIP_CHECK(h, stream_sync(h));
    return drain_status(h);
}

int dtx_set_f(dtx_handle* h, const double* f) {
    if (!h || !f) return DTX_ERR_INVALID;
    int rc = upload(h, h->d_f[h->cur], f, h->n);
    if (rc) return rc;
    rc = refresh_fT(h);
    if (rc) return rc;
    HIP_CHECK(h, stream_sync(h));
    return DTX_OK;
}

int dtx_moments(dtx_handle* h, double* number, double* energy) {
    if (!h || !number || !energy) return DTX_ERR_INVALID;
    std::vector<double> m((size_t)h->C * h->S * 2);
    HIP_CHECK(h, stream_sync(h));   // the moments of the last step are complete
    HIP_CHECK(h, hipMemcpyAsync(m.data(), h->d_moments, m.size() * 8, hipMemcpyDeviceToHost,
                                h->stream));
    HIP_CHECK(h, stream_sync(h));
    for (size_t i = 0; i < (size_t)h->C * h->S; ++i) {
        number[i] = m[2 * i];
        energy[i] = m[2 * i + 1];
    }
    return drain_status(h);
}

int dtx_synchronize(dtx_handle* h) {
    if (!h) return DTX_ERR_INVALID;
    HIP_CHECK(h, stream_sync(h));
    return drain_status(h);
}

int dtx_step(dtx_handle* h, double t_start, double t_stop, double dt, int32_t verbose,
             double* dt_out, int32_t* save_out, dtx_step_stats* stats) {
    if (!h) return DTX_ERR_INVALID;
    int rc = ensure_ready(h);
    if (rc) return rc;
    int save;
    if (t_start + dt >= t_stop) {  // src/SteppingMethods.jl:18-23
        dt = t_stop - t_start;
        save = 1;
    } else {
        save = 0;
    }
    const int64_t idx = h->step_counter;
    rc = enqueue_step(h, t_start, t_stop, dt, verbose, save);
    if (rc) return rc;
    if (dt_out) *dt_out = dt;
    if (save_out) *save_out = save;
    if (stats) {
        HIP_CHECK(h, stream_sync(h));
        double st[STATUS_WORDS];
        HIP_CHECK(h, hipMemcpy(st, h->d_status + (idx % STATUS_RING) * STATUS_WORDS,
                               sizeof st, hipMemcpyDeviceToHost));
        fill_stats(st, stats);
        rc = drain_status(h);
        if (rc) return rc;
    }
    return DTX_OK;
}

int dtx_step_noalloc(dtx_handle* h, double t_start, double t_stop, double dt, int32_t verbose,
                     double* dt_out, int32_t* save_out, dtx_step_stats* stats) {
    if (!h) return DTX_ERR_INVALID;
    if (!h->vectors_set)
        return fail(h, DTX_ERR_STATE, "dtx_set_vectors must be called before stepping");
    if (h->binary && h->binary_dirty)
        return fail(h, DTX_ERR_STATE, "dtx_step_noalloc: the operators are not packed yet (a "
                    "dtx_step, dtx_solve or dtx_work call packs them)");
    const bool bin = h->binary && h->nitems > 0;
    if (bin && !h->scratch)
        return fail(h, DTX_ERR_HIP, "dtx_step_noalloc: the device scratch array is unavailable");
    if (bin && (size_t)h->scratch_need > SCRATCH_BYTES)
        return fail(h, DTX_ERR_UNSUPPORTED, "dtx_step_noalloc: the step needs %lld bytes of "
                    "scratch, more than the %zu available", (long long)h->scratch_need,
                    SCRATCH_BYTES);
    int save;
    if (t_start + dt >= t_stop) {  // src/SteppingMethods.jl:18-23
        dt = t_stop - t_start;
        save = 1;
    } else {
        save = 0;
    }
    const int64_t idx = h->step_counter;
    int rc = enqueue_step(h, t_start, t_stop, dt, verbose, save, /*use_scratch=*/true);
    if (rc) return rc;
    if (dt_out) *dt_out = dt;
    if (save_out) *save_out = save;
    if (stats) {
        HIP_CHECK(h, stream_sync(h));
        double st[STATUS_WORDS];
        HIP_CHECK(h, hipMemcpy(st, h->d_status + (idx % STATUS_RING) * STATUS_WORDS,
                               sizeof st, hipMemcpyDeviceToHost));
        fill_stats(st, stats);
        rc = drain_status(h);
        if (rc) return rc;
    }
    return DTX_OK;
}

int dtx_solve(dtx_handle* h, double dt_initial, int64_t n_save, const double* t_save,
              int32_t verbose, double* f_out, double* t_out, int64_t* n_steps) {
    if (!h) return DTX_ERR_INVALID;
    if (n_save < 1 || !t_save || !f_out || !t_out)
        return fail(h, DTX_ERR_INVALID, "invalid Solve arguments");
    int rc = ensure_ready(h);
    if (rc) return rc;
    rc = dtx_reset(h);  // src/Solver.jl:22
    if (rc) return rc;
    const int64_t n = h->n;
    double* staging = nullptr;
    HIP_CHECK(h, hipHostMalloc((void**)&staging, (size_t)n * n_save * 8, hipHostMallocDefault));
    HIP_CHECK(h, hipMemcpyAsync(staging, h->d_finit, (size_t)n * 8, hipMemcpyDeviceToHost,
                                h->stream));
    t_out[0] = t_save[0];
    double dt = dt_initial;
    int64_t steps = 0;
    if (!(dt > 0.0) || !std::isfinite(dt))
        rc = fail(h, DTX_ERR_INVALID, "dt_initial must be positive and finite");
    for (int64_t i = 1; i < n_save && rc == DTX_OK; ++i) {  // src/Solver.jl:27-44
        double t_start = t_save[i - 1];
        const double t_stop = t_save[i];
        int save = 0;
        // The clipped dt carries over between intervals (src/Solver.jl:35).  When rounding makes
        // it collapse, the reference would loop (practically) forever; refuse instead.
        if (!(dt > 0.0) || (t_stop - t_start) / dt > kMaxIntervalSteps) {
            rc = fail(h, DTX_ERR_INVALID,
                      "Solve: dt = %.17g carried into save interval %lld [%.17g, %.17g] would "
                      "need more than %.0e steps", dt, (long long)i, t_start, t_stop,
                      kMaxIntervalSteps);
            break;
        }
        while (!save) {
            if (t_start + dt >= t_stop) {
                dt = t_stop - t_start;
                save = 1;
            }
            rc = enqueue_step(h, t_start, t_stop, dt, verbose, save);
            if (rc) break;
            ++steps;
            t_start += dt;
        }
        if (rc) break;
        hipError_t e = hipMemcpyAsync(staging + i * n, h->d_f[h->cur], (size_t)n * 8,
                                      hipMemcpyDeviceToHost, h->stream);
        if (e != hipSuccess) rc = fail(h, DTX_ERR_HIP, "snapshot copy: %s", hipGetErrorString(e));
        t_out[i] = t_stop;
    }
    hipError_t e = stream_sync(h);
    if (rc == DTX_OK && e != hipSuccess)
        rc = fail(h, DTX_ERR_HIP, "solve: %s", hipGetErrorString(e));
    if (rc == DTX_OK) {
        memcpy(f_out, staging, (size_t)n * n_save * 8);
        rc = drain_status(h);
    }
    (void)hipHostFree(staging);
    if (n_steps) *n_steps = steps;
    return rc;
}

int dtx_recent_stats(dtx_handle* h, dtx_step_stats* out, int32_t max_steps, int32_t* count) {
    if (!h || !out || !count) return DTX_ERR_INVALID;
    HIP_CHECK(h, stream_sync(h));
    std::vector<double> ring((size_t)STATUS_RING * STATUS_WORDS);
    HIP_CHECK(h, hipMemcpy(ring.data(), h->d_status, ring.size() * 8, hipMemcpyDeviceToHost));
    int64_t avail = std::min<int64_t>(h->step_counter, STATUS_RING);
    int64_t take = std::min<int64_t>(avail, max_steps);
    for (int64_t q = 0; q < take; ++q) {
        int64_t s = h->step_counter - take + q;
        fill_stats(ring.data() + (s % STATUS_RING) * STATUS_WORDS, &out[q]);
    }
    *count = (int32_t)take;
    return drain_status(h);
}

int64_t dtx_device_bytes(const dtx_handle* h) { return h ? h->device_bytes : 0; }

int dtx_work(const dtx_handle* hc, dtx_work_info* out) {
    dtx_handle* h = const_cast<dtx_handle*>(hc);
    if (!h || !out) return DTX_ERR_INVALID;
    if (h->binary && h->binary_dirty && h->vectors_set) {
        int rc = ensure_ready(h);
        if (rc) return rc;
    }
    memset(out, 0, sizeof *out);
    const double C = h->C, N = h->N;
    out->pairs = h->pairs_total;
    out->bin_launch_items = h->nitems;
    out->fused_correction = h->correct_active && h->fused_update ? 1 : 0;
    out->gain_bytes = h->gain_bytes;
    if (h->binary) {
        out->bin_flops = h->bin_flops;  // gain contraction + FP64 loss channels
        // packed operator once + chunk tables + f reads (fT) + partial writes
        double pair_bytes = 0;
        for (auto& c : h->chans) pair_bytes += c.nchunks * 4.0;
        out->bin_bytes = h->gain_bytes + pair_bytes + C * N * 8.0 +
                         (double)h->max_slots * C * N * 8.0;
    }
    // update: f, f_out, fT, invAp, df_inj, ELL (col+val), partial reads, moments
    double ell_bytes = 0;
    for (int w = 0; w < 3; ++w) ell_bytes += (double)h->ell[w].width * h->ell[w].n * 12.0;
    out->update_bytes = h->n * 8.0 * 5 + ell_bytes + (double)h->max_slots * C * N * 8.0;
    out->emission_bytes = h->emission ? (double)C * N * N * 8.0 + h->n * 16.0 : 0.0;
    return DTX_OK;
}

int dtx_kernel_timing(dtx_handle* h, int32_t enable, int32_t reset, dtx_kernel_times* out) {
    if (!h) return DTX_ERR_INVALID;
    HIP_CHECK(h, stream_sync(h));
    for (auto& ev : h->events) {
        float ms = 0;
        HIP_CHECK(h, hipEventElapsedTime(&ms, ev.start, ev.stop));
        if (ev.cls == 0) {
            h->times.bin_ms += ms;
            h->times.bin_launches++;
        } else if (ev.cls == 2) {
            h->times.emission_ms += ms;
            h->times.emission_launches++;
        } else {
            h->times.update_ms += ms;
            h->times.update_launches++;
        }
        (void)hipEventDestroy(ev.start);
        (void)hipEventDestroy(ev.stop);
    }
    h->events.clear();
    if (out) *out = h->times;
    if (reset) h->times = dtx_kernel_times{};
    h->timing = enable > 0;
    h->timing_every = enable > 0 ? enable : 1;
    return drain_status(h);
}

}  // extern "C"
