// lz4_block.hpp -- decoder of the LZ4 block format, as quipu files store
// compressed images (one block per image, written by LZ4_compress_default and
// read by LZ4_decompress_safe in quipu_io.ipp:176-235).
//
// The block format (LZ4 block format description, lz4 project documentation):
// a sequence of
//   token (high nibble: literal count, low nibble: match length - 4),
//   literal-count extension bytes while the nibble is 15 (each adds 0..255,
//     the last one is < 255),
//   the literals,
//   a 2-byte little-endian match offset (1..65535) back into the output,
//   match-length extension bytes while the nibble is 15,
// and the block ends right after the literals of its last sequence (which has
// no match).  A match may overlap its own output (offset < length), which
// repeats a pattern, so it is copied byte by byte in order.
//
// Every read and write is bounds-checked: a malformed or truncated block
// returns -1 and never touches memory outside [src, src+n) and [dst, dst+cap).
#pragma once

#include <cstddef>
#include <cstdint>

namespace vg {

// Decodes one block; returns the number of bytes written, or -1.
inline long long lz4_decode_block(const uint8_t* src, size_t n, uint8_t* dst, size_t cap) {
    size_t ip = 0, op = 0;
    if (n == 0) return -1;
    while (true) {
        if (ip >= n) return -1;
        const uint32_t token = src[ip++];
        // literals
        size_t lit = token >> 4;
        if (lit == 15) {
            uint32_t b;
            do {
                if (ip >= n) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > n - ip || lit > cap - op) return -1;
        for (size_t k = 0; k < lit; ++k) dst[op + k] = src[ip + k];
        ip += lit;
        op += lit;
        if (ip == n) return (long long)op;  // the last sequence has no match
        // match
        if (n - ip < 2) return -1;
        const size_t off = (size_t)src[ip] | ((size_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return -1;
        size_t len = (token & 15u) + 4;
        if ((token & 15u) == 15) {
            uint32_t b;
            do {
                if (ip >= n) return -1;
                b = src[ip++];
                len += b;
            } while (b == 255);
        }
        if (len > cap - op) return -1;
        const size_t from = op - off;
        for (size_t k = 0; k < len; ++k) dst[op + k] = dst[from + k];
        op += len;
    }
}

}  // namespace vg
