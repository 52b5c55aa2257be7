// Host-built tables of the EBCOT Tier-1 coder and decoder (ISO/IEC 15444-1 Tables C.2 and
// D.1), uploaded to the constant memory of the kernels that use them.
#pragma once
#include <algorithm>
#include <cstdint>

namespace gail_amd {

// Table D.1 for the four orientations (0 LL, 1 HL, 2 LH, 3 HH), indexed [orient * 128 +
// code] by a 7-bit neighbourhood code h | v << 2 | d << 4: h, v the significant horizontal
// and vertical neighbours (0..2), d the diagonal ones (0..4), each as a binary count.
// Codes no neighbourhood produces clamp to the largest counts.
inline void build_zc_table(uint8_t z[4 * 128]) {
    uint8_t t[4 * 45];
    for (int o = 0; o < 4; ++o)
        for (int h = 0; h < 3; ++h)
            for (int v = 0; v < 3; ++v)
                for (int d = 0; d < 5; ++d) {
                    int hh = h, vv = v, c;
                    if (o == 1) {  // HL: horizontal and vertical trade places
                        hh = v;
                        vv = h;
                    }
                    if (o == 3) {
                        const int hv = hh + vv;
                        c = d >= 3 ? 8 : d == 2 ? (hv ? 7 : 6) : d == 1 ? (hv >= 2 ? 5 : hv ? 4 : 3) : (hv >= 2 ? 2 : hv);
                    } else {
                        c = hh == 2 ? 8 : hh == 1 ? (vv ? 7 : d ? 6 : 5) : vv ? (vv == 2 ? 4 : 3) : (d >= 2 ? 2 : d);
                    }
                    t[o * 45 + h * 15 + v * 5 + d] = (uint8_t)c;
                }
    for (int o = 0; o < 4; ++o)
        for (int i = 0; i < 128; ++i) {
            const int h = std::min(2, i & 3), v = std::min(2, (i >> 2) & 3), d = std::min(4, (i >> 4) & 7);
            z[o * 128 + i] = t[o * 45 + h * 15 + v * 5 + d];
        }
}

// Table C.2 packed as Qe | NMPS << 16 | NLPS << 22 | SWITCH << 28. A context's state word is
// its entry of this table with its MPS in bit 31.
inline void build_mq_table(uint32_t tab[47]) {
    static const uint16_t qe[47] = {
        0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801, 0x3001, 0x2401,
        0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401, 0x3001, 0x2801, 0x2401, 0x2201,
        0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101, 0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1,
        0x0221, 0x0141, 0x0111, 0x0085, 0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
    static const uint8_t nm[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                                   17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                   33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
    static const uint8_t nl[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                                   15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                   30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
    for (int i = 0; i < 47; ++i)
        tab[i] = qe[i] | ((uint32_t)nm[i] << 16) | ((uint32_t)nl[i] << 22) | ((i == 0 || i == 6 || i == 14) ? 1u << 28 : 0u);
}

}  // namespace gail_amd
