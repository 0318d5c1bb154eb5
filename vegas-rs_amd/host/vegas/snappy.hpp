// Snappy block compression for the parquet pages (the reference's
// ArrowWriter compresses with Compression::SNAPPY, src/output.rs:48,138).
//
// Written from the published format description (google/snappy,
// format_description.txt): a block is the uncompressed length as a varint,
// then elements -- literals (tag & 3 == 0: length - 1 in the tag's upper six
// bits when below 60, else in 1..4 little-endian bytes after it) and copies
// of earlier output (tag & 3 == 1: length 4..11 and an 11-bit offset; tag &
// 3 == 2: length 1..64 and a 16-bit little-endian offset).  A copy may
// overlap its own output (offset < length), which repeats a pattern.
//
// The compressor here is a greedy one: the input is cut into fragments of
// 64 KiB (so every offset fits 16 bits), and within a fragment a hash of the
// next four bytes looks up the last position with the same hash; a match of
// at least four bytes becomes a copy, extended as far as it goes.  After 32
// positions without a match the scan steps faster (one more byte per 32
// misses), so incompressible data costs little time.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace vegas {

namespace snappy_detail {

inline void varint(std::string& out, uint64_t v) {
    while (v >= 0x80) {
        out.push_back((char)((v & 0x7F) | 0x80));
        v >>= 7;
    }
    out.push_back((char)v);
}

inline void literal(std::string& out, const char* p, size_t n) {
    while (n > 0) {
        // one element holds at most 2^32 bytes; pages are far smaller, but
        // the split keeps the encoding valid for any input
        const size_t m = n > (size_t(1) << 32) ? (size_t(1) << 32) : n;
        const uint64_t k = m - 1;
        if (k < 60) {
            out.push_back((char)(k << 2));
        } else {
            int nb = 1;
            while (nb < 4 && (k >> (8 * nb)) != 0) ++nb;
            out.push_back((char)((59 + nb) << 2));
            for (int b = 0; b < nb; ++b) out.push_back((char)(k >> (8 * b)));
        }
        out.append(p, m);
        p += m;
        n -= m;
    }
}

// One copy element of length 1..64 and offset 1..65535.
inline void copy_one(std::string& out, uint32_t off, uint32_t len) {
    if (len >= 4 && len <= 11 && off < 2048) {
        out.push_back((char)(1u | ((len - 4) << 2) | ((off >> 8) << 5)));
        out.push_back((char)(off & 0xFF));
    } else {
        out.push_back((char)(2u | ((len - 1) << 2)));
        out.push_back((char)(off & 0xFF));
        out.push_back((char)(off >> 8));
    }
}

// A copy of any length, as elements of at most 64 bytes.
inline void copy(std::string& out, uint32_t off, size_t len) {
    while (len >= 68) {
        copy_one(out, off, 64);
        len -= 64;
    }
    if (len > 64) {  // 65..67: leave at least 4 for the last element
        copy_one(out, off, 60);
        len -= 60;
    }
    copy_one(out, off, (uint32_t)len);
}

inline uint32_t load32(const char* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

constexpr size_t kFragment = size_t(1) << 16;
constexpr int kHashBits = 14;

inline uint32_t hash4(uint32_t v) { return (v * 0x1E35A7BDu) >> (32 - kHashBits); }

// Elements for the fragment [p, p + n), n <= kFragment.
inline void fragment(std::string& out, const char* p, size_t n, std::vector<int32_t>& table) {
    std::fill(table.begin(), table.end(), -1);
    size_t lit = 0;  // start of the pending literal
    size_t i = 0;
    uint32_t misses = 0;
    while (n >= 4 && i + 4 <= n) {
        const uint32_t cur = load32(p + i);
        const uint32_t h = hash4(cur);
        const int32_t cand = table[h];
        table[h] = (int32_t)i;
        if (cand >= 0 && load32(p + cand) == cur) {
            size_t len = 4;
            while (i + len < n && p[cand + len] == p[i + len]) ++len;
            if (i > lit) literal(out, p + lit, i - lit);
            copy(out, (uint32_t)(i - (size_t)cand), len);
            // index the match's last positions so the next one can chain
            const size_t end = i + len;
            for (size_t k = end >= 3 ? end - 3 : 0; k < end && k + 4 <= n; ++k)
                if (k > i) table[hash4(load32(p + k))] = (int32_t)k;
            i = end;
            lit = i;
            misses = 0;
        } else {
            i += 1 + (misses++ >> 5);
        }
    }
    if (n > lit) literal(out, p + lit, n - lit);
}

}  // namespace snappy_detail

// The Snappy block (raw format, as parquet stores it) of n bytes at p.
inline std::string snappy_compress(const void* data, size_t n) {
    const char* p = static_cast<const char*>(data);
    std::string out;
    out.reserve(n + n / 6 + 16);
    snappy_detail::varint(out, n);
    std::vector<int32_t> table(size_t(1) << snappy_detail::kHashBits);
    for (size_t f = 0; f < n; f += snappy_detail::kFragment) {
        const size_t m = n - f < snappy_detail::kFragment ? n - f : snappy_detail::kFragment;
        snappy_detail::fragment(out, p + f, m, table);
    }
    return out;
}

}  // namespace vegas
