// bf16_util.h -- round-to-nearest-even float <-> bfloat16 bit conversions, host and device.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define DTX_HD __host__ __device__
#else
#define DTX_HD
#endif

namespace dtx {

DTX_HD inline uint32_t f32_bits(float x) { return __builtin_bit_cast(uint32_t, x); }

DTX_HD inline float bits_f32(uint32_t u) { return __builtin_bit_cast(float, u); }

// float -> bf16 with round-to-nearest-even (finite inputs).
DTX_HD inline uint16_t bf16_rne(float x) {
    uint32_t u = f32_bits(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

DTX_HD inline float bf16_to_f32(uint16_t h) { return bits_f32((uint32_t)h << 16); }

// Split a double into hi + lo bf16 parts: hi = bf16(x), lo = bf16(x - hi).
DTX_HD inline void split_bf16(double x, uint16_t* hi, uint16_t* lo) {
    uint16_t h = bf16_rne((float)x);
    double r = x - (double)bf16_to_f32(h);
    *hi = h;
    *lo = bf16_rne((float)r);
}

// Position of operator element (row r, chunk offset t) inside one packed split-BF16 chunk half
// (64 rows x 32 j): the order in which the 16x16x32 MFMA consumes its A operand.  Rows come in
// blocks of 16; inside a block, lane l = (r mod 16) + 16 (t div 8) holds the 8 consecutive j of
// its row at 16 bytes.  Staging copies a chunk to LDS in this order and every wave reads its
// fragments as one linear 1 KB block per instruction, free of LDS bank conflicts.
DTX_HD inline int bf16_frag_index(int r, int t) {
    return (((r >> 4) * 64 + (r & 15) + 16 * (t >> 3)) << 3) + (t & 7);
}

// Position (in bf16 elements) of the split of f for cell c and padded state kp inside the
// fragment-ordered copy that the split-BF16 contraction reads as its B operand.  Per 32-state
// chunk and 16-cell block there is one 1 KB block in the order the 16x16x32 MFMA consumes B:
// lane (c mod 16) + 16 ((kp mod 32) div 8) holds the 8 consecutive states of its octet, so the
// B quad of any lane is 16 contiguous bytes and a wave loads a whole block with one
// instruction.  cblocks = 16-cell blocks per chunk (padded cells / 16).
DTX_HD inline int64_t fb_index(int c, int kp, int cblocks) {
    return ((((int64_t)(kp >> 5) * cblocks + (c >> 4)) * 64 + (c & 15) + 16 * ((kp >> 3) & 3))
            << 3) + (kp & 7);
}

}  // namespace dtx
