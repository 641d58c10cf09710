// md5.hpp -- MD5 message digest (RFC 1321), for runvx's image checksums (the reference runvx
// reports MD5 as 32 lowercase hex digits, utilities/runvx/vxUtils.cpp:347-369).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

namespace mvxe {

class Md5 {
public:
    Md5() { reset(); }

    void reset()
    {
        h_[0] = 0x67452301u; h_[1] = 0xefcdab89u; h_[2] = 0x98badcfeu; h_[3] = 0x10325476u;
        len_ = 0;
        fill_ = 0;
    }

    void update(const uint8_t *p, size_t n)
    {
        len_ += n;
        while (n) {
            const size_t take = n < 64 - fill_ ? n : 64 - fill_;
            memcpy(buf_ + fill_, p, take);
            fill_ += take; p += take; n -= take;
            if (fill_ == 64) { block(buf_); fill_ = 0; }
        }
    }

    std::string hex()
    {
        const uint64_t bits = len_ * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill_ != 56) update(&zero, 1);
        uint8_t lenLE[8];
        for (int i = 0; i < 8; i++) lenLE[i] = (uint8_t)(bits >> (8 * i));
        update(lenLE, 8);
        static const char digits[] = "0123456789abcdef";
        std::string s;
        for (int i = 0; i < 4; i++)
            for (int b = 0; b < 4; b++) {
                const uint8_t v = (uint8_t)(h_[i] >> (8 * b));
                s += digits[v >> 4];
                s += digits[v & 15];
            }
        return s;
    }

private:
    static uint32_t rotl(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }

    void block(const uint8_t *p)
    {
        // per-round shift amounts and the constants K[i] = floor(2^32 * |sin(i + 1)|) of RFC 1321 3.4
        static const int S[64] = {7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
                                  5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20,
                                  4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
                                  6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};
        static uint32_t K[64];
        static bool init = false;
        if (!init) {
            for (int i = 0; i < 64; i++) {
                // |sin(i+1)| * 2^32 computed in long double, truncated (the RFC's table)
                long double v = sinl((long double)(i + 1));
                if (v < 0) v = -v;
                K[i] = (uint32_t)(v * 4294967296.0L);
            }
            init = true;
        }
        uint32_t M[16];
        for (int i = 0; i < 16; i++) M[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) | ((uint32_t)p[4 * i + 3] << 24);
        uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3];
        for (int i = 0; i < 64; i++) {
            uint32_t f;
            int g;
            if (i < 16) { f = (b & c) | (~b & d); g = i; }
            else if (i < 32) { f = (d & b) | (~d & c); g = (5 * i + 1) & 15; }
            else if (i < 48) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
            else { f = c ^ (b | ~d); g = (7 * i) & 15; }
            const uint32_t t = d;
            d = c;
            c = b;
            b = b + rotl(a + f + K[i] + M[g], S[i]);
            a = t;
        }
        h_[0] += a; h_[1] += b; h_[2] += c; h_[3] += d;
    }

    uint32_t h_[4];
    uint64_t len_;
    uint8_t buf_[64];
    size_t fill_;
};

}  // namespace mvxe
