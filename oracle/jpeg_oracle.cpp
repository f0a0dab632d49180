/*
 * jpeg_oracle.cpp — CPU restatement of the frame grab's JPEG encoding.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP JPEG encoder is compared
 * against and the CPU baseline bench.py times. Nothing in livre_amd/ links or
 * loads it.
 *
 * The reference grabs frames with libjpeg-turbo's tjCompress2
 * (livre/eq/FrameGrabber.cpp:49-84): TJPF_BGRA pixels, TJSAMP_444, quality
 * 100, flags = TJXOP_ROT180 (= 6, which tjCompress2 reads as TJFLAG_BOTTOMUP
 * plus an unused bit, so rows are taken bottom-up). libjpeg-turbo is a
 * third-party dependency (CMakeLists.txt:53, common_find_package(LibJpegTurbo),
 * no version pinned) and is not under the reference tree. This file restates
 * its published baseline-JPEG algorithm, ITU-T T.81 with the IJG choices:
 *   colour conversion   jccolor.c rgb_ycc_convert (16-bit fixed point)
 *   level shift + FDCT  jcdctmgr.c convsamp, jfdctint.c jpeg_fdct_islow
 *                       (tjCompress2 selects the accurate integer DCT for
 *                       quality >= 96, and the reference asks for 100)
 *   quantisation        jcdctmgr.c compute_reciprocal + quantize
 *   tables              jcparam.c jpeg_set_quality (standard tables of T.81
 *                       Annex K.1 scaled, force_baseline) and the standard
 *                       Huffman tables of Annex K.3
 *   entropy coding      jchuff.c encode_one_block, flush_bits (pad with 1s)
 *   markers             jcmarker.c: SOI, JFIF APP0 1.01, one DQT per table,
 *                       SOF0, one DHT per table, SOS, scan, EOI
 *   edges               jcsample.c expand_right_edge, jcprepct.c
 *                       expand_bottom_edge (replicate the last column / row)
 * Pinned by tests/golden/jpeg_*.jpg, written by libjpeg-turbo through Pillow
 * (tests/golden/make_jpeg_golden.py), byte for byte, at several sizes and
 * qualities, and by decoding its output with Pillow.
 */
#include "livre_oracle.h"

#include <cstring>
#include <vector>

namespace
{
const int kNatural[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                          12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                          35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                          58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// T.81 Annex K.1, natural order
const int kStdLuma[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                          14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                          18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                          49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const int kStdChroma[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                            24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                            99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                            99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

// T.81 Annex K.3: code counts per length 1..16, then the symbols
const unsigned char kDcLumaBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const unsigned char kDcChromaBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const unsigned char kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const unsigned char kAcLumaBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const unsigned char kAcLumaVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61,
    0x07, 0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52,
    0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25,
    0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45,
    0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64,
    0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99,
    0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6,
    0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3,
    0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8,
    0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const unsigned char kAcChromaBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const unsigned char kAcChromaVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61,
    0x71, 0x13, 0x22, 0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33,
    0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18,
    0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44,
    0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63,
    0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97,
    0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4,
    0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca,
    0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7,
    0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

struct Huff
{
    unsigned code[256];
    int size[256]; // 0 = symbol absent
};

// jchuff.c jpeg_make_c_derived_tbl: canonical codes, T.81 Annex C
void makeHuff(const unsigned char bits[16], const unsigned char* vals, Huff& h)
{
    std::memset(&h, 0, sizeof h);
    unsigned code = 0;
    int p = 0;
    for (int len = 1; len <= 16; ++len)
    {
        for (int i = 0; i < bits[len - 1]; ++i, ++p)
        {
            h.code[vals[p]] = code++;
            h.size[vals[p]] = len;
        }
        code <<= 1;
    }
}

// jcparam.c jpeg_quality_scaling + jpeg_add_quant_table(force_baseline)
void qualityTable(const int base[64], int quality, int out[64])
{
    if (quality <= 0)
        quality = 1;
    if (quality > 100)
        quality = 100;
    const int scale = quality < 50 ? 5000 / quality : 200 - quality * 2;
    for (int i = 0; i < 64; ++i)
    {
        long t = (long(base[i]) * scale + 50L) / 100L;
        if (t <= 0)
            t = 1;
        if (t > 255)
            t = 255;
        out[i] = int(t);
    }
}

// jcdctmgr.c compute_reciprocal for a divisor of quantval << 3 (the
// accurate DCT's output is scaled by 8). The quantiser divides by
// multiplying with a rounded reciprocal: q = ((|x| + corr) * recip) >> shift.
struct Divisor
{
    uint32_t recip, corr, shift;
};

Divisor reciprocal(uint32_t divisor)
{
    int b = 31 - __builtin_clz(divisor); // flss(divisor) - 1
    int r = 16 + b;
    uint32_t fq = (1u << r) / divisor;
    const uint32_t fr = (1u << r) % divisor;
    uint32_t c = divisor / 2;
    if (fr == 0)
    {
        fq >>= 1;
        --r;
    }
    else if (fr <= divisor / 2)
        ++c;
    else
        ++fq;
    return {fq, c, uint32_t(r)};
}

// jfdctint.c jpeg_fdct_islow: CONST_BITS 13, PASS1_BITS 2
const int32_t F0298 = 2446, F0390 = 3196, F0541 = 4433, F0765 = 6270, F0899 = 7373,
              F1175 = 9633, F1501 = 12299, F1847 = 15137, F1961 = 16069, F2053 = 16819,
              F2562 = 20995, F3072 = 25172;

inline int32_t descale(int32_t x, int n) { return (x + (1 << (n - 1))) >> n; }

void fdctIslow(int32_t d[64])
{
    for (int pass = 0; pass < 2; ++pass)
    {
        const int step = pass == 0 ? 1 : 8; // rows, then columns
        const int shiftOdd = pass == 0 ? 13 - 2 : 13 + 2;
        for (int line = 0; line < 8; ++line)
        {
            int32_t* p = pass == 0 ? d + 8 * line : d + line;
            const int32_t tmp0 = p[0] + p[7 * step], tmp7 = p[0] - p[7 * step];
            const int32_t tmp1 = p[step] + p[6 * step], tmp6 = p[step] - p[6 * step];
            const int32_t tmp2 = p[2 * step] + p[5 * step], tmp5 = p[2 * step] - p[5 * step];
            const int32_t tmp3 = p[3 * step] + p[4 * step], tmp4 = p[3 * step] - p[4 * step];
            const int32_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3;
            const int32_t tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
            if (pass == 0)
            {
                p[0] = (tmp10 + tmp11) * 4;
                p[4 * step] = (tmp10 - tmp11) * 4;
            }
            else
            {
                p[0] = descale(tmp10 + tmp11, 2);
                p[4 * step] = descale(tmp10 - tmp11, 2);
            }
            const int32_t z1e = (tmp12 + tmp13) * F0541;
            p[2 * step] = descale(z1e + tmp13 * F0765, shiftOdd);
            p[6 * step] = descale(z1e - tmp12 * F1847, shiftOdd);
            const int32_t z1 = tmp4 + tmp7, z2 = tmp5 + tmp6, z3 = tmp4 + tmp6, z4 = tmp5 + tmp7;
            const int32_t z5 = (z3 + z4) * F1175;
            const int32_t t4 = tmp4 * F0298, t5 = tmp5 * F2053, t6 = tmp6 * F3072,
                          t7 = tmp7 * F1501;
            const int32_t m1 = -z1 * F0899, m2 = -z2 * F2562;
            const int32_t m3 = -z3 * F1961 + z5, m4 = -z4 * F0390 + z5;
            p[7 * step] = descale(t4 + m1 + m3, shiftOdd);
            p[5 * step] = descale(t5 + m2 + m4, shiftOdd);
            p[3 * step] = descale(t6 + m2 + m3, shiftOdd);
            p[step] = descale(t7 + m1 + m4, shiftOdd);
        }
    }
}

struct BitWriter
{
    std::vector<uint8_t>& out;
    uint32_t acc = 0;
    int n = 0;
    void put(uint32_t code, int len)
    {
        for (int i = len - 1; i >= 0; --i)
        {
            acc = (acc << 1) | ((code >> i) & 1u);
            if (++n == 8)
            {
                out.push_back(uint8_t(acc));
                if (uint8_t(acc) == 0xFF)
                    out.push_back(0); // byte stuffing, T.81 F.1.2.3
                acc = 0;
                n = 0;
            }
        }
    }
    void flush() // jchuff.c flush_bits: fill the last byte with 1 bits
    {
        if (n > 0)
            put(0x7F, 8 - n);
    }
};

void marker(std::vector<uint8_t>& o, uint8_t m, int len)
{
    o.push_back(0xFF);
    o.push_back(m);
    if (len >= 0)
    {
        o.push_back(uint8_t(len >> 8));
        o.push_back(uint8_t(len));
    }
}

void dht(std::vector<uint8_t>& o, int cls, int id, const unsigned char bits[16],
         const unsigned char* vals)
{
    int n = 0;
    for (int i = 0; i < 16; ++i)
        n += bits[i];
    marker(o, 0xC4, 2 + 1 + 16 + n);
    o.push_back(uint8_t(cls << 4 | id));
    o.insert(o.end(), bits, bits + 16);
    o.insert(o.end(), vals, vals + n);
}

inline int nbitsOf(int v)
{
    int n = 0;
    for (unsigned t = unsigned(v < 0 ? -v : v); t; t >>= 1)
        ++n;
    return n;
}
} // namespace

extern "C" int orc_jpeg_encode(const uint8_t* pixels, int width, int height, int pitch,
                               int bgra, int bottom_up, int quality, uint8_t* out, size_t cap,
                               size_t* size)
{
    if (!pixels || width < 1 || height < 1 || width > 65535 || height > 65535 ||
        pitch < 4 * width || quality < 1 || quality > 100 || !size)
        return -1;
    int qLuma[64], qChroma[64];
    qualityTable(kStdLuma, quality, qLuma);
    qualityTable(kStdChroma, quality, qChroma);
    Divisor div[2][64];
    for (int i = 0; i < 64; ++i)
    {
        div[0][i] = reciprocal(uint32_t(qLuma[i]) << 3);
        div[1][i] = reciprocal(uint32_t(qChroma[i]) << 3);
    }
    Huff dc[2], ac[2];
    makeHuff(kDcLumaBits, kDcVals, dc[0]);
    makeHuff(kDcChromaBits, kDcVals, dc[1]);
    makeHuff(kAcLumaBits, kAcLumaVals, ac[0]);
    makeHuff(kAcChromaBits, kAcChromaVals, ac[1]);

    std::vector<uint8_t> o;
    marker(o, 0xD8, -1);
    const uint8_t jfif[14] = {'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    marker(o, 0xE0, 16);
    o.insert(o.end(), jfif, jfif + 14);
    for (int t = 0; t < 2; ++t)
    {
        marker(o, 0xDB, 67);
        o.push_back(uint8_t(t));
        for (int i = 0; i < 64; ++i)
            o.push_back(uint8_t((t ? qChroma : qLuma)[kNatural[i]]));
    }
    marker(o, 0xC0, 17);
    const uint8_t sof[15] = {8,    uint8_t(height >> 8), uint8_t(height), uint8_t(width >> 8),
                             uint8_t(width), 3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1};
    o.insert(o.end(), sof, sof + 15);
    dht(o, 0, 0, kDcLumaBits, kDcVals);
    dht(o, 1, 0, kAcLumaBits, kAcLumaVals);
    dht(o, 0, 1, kDcChromaBits, kDcVals);
    dht(o, 1, 1, kAcChromaBits, kAcChromaVals);
    marker(o, 0xDA, 12);
    const uint8_t sos[10] = {3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0};
    o.insert(o.end(), sos, sos + 10);

    BitWriter bw{o};
    const int mcuW = (width + 7) / 8, mcuH = (height + 7) / 8;
    const int ri = bgra ? 2 : 0, bi = bgra ? 0 : 2;
    int lastDc[3] = {0, 0, 0};
    for (int my = 0; my < mcuH; ++my)
        for (int mx = 0; mx < mcuW; ++mx)
        {
            int32_t blk[3][64];
            for (int y = 0; y < 8; ++y)
            {
                const int iy = my * 8 + y < height ? my * 8 + y : height - 1;
                const int sy = bottom_up ? height - 1 - iy : iy;
                for (int x = 0; x < 8; ++x)
                {
                    const int ix = mx * 8 + x < width ? mx * 8 + x : width - 1;
                    const uint8_t* px = pixels + size_t(sy) * size_t(pitch) + 4 * size_t(ix);
                    const int32_t r = px[ri], g = px[1], b = px[bi];
                    // jccolor.c: FIX(x) = round(x * 2^16); Cb/Cr round with 0.5 - eps
                    const int32_t Y = (19595 * r + 38470 * g + 7471 * b + 32768) >> 16;
                    const int32_t Cb =
                        (-11059 * r - 21709 * g + 32768 * b + (128 << 16) + 32767) >> 16;
                    const int32_t Cr =
                        (32768 * r - 27439 * g - 5329 * b + (128 << 16) + 32767) >> 16;
                    blk[0][8 * y + x] = Y - 128;
                    blk[1][8 * y + x] = Cb - 128;
                    blk[2][8 * y + x] = Cr - 128;
                }
            }
            for (int c = 0; c < 3; ++c)
            {
                fdctIslow(blk[c]);
                const Divisor* dv = div[c ? 1 : 0];
                int q[64];
                for (int i = 0; i < 64; ++i)
                {
                    const int32_t v = blk[c][i];
                    const uint32_t a = uint32_t(v < 0 ? -v : v);
                    const int32_t m =
                        int32_t((uint64_t(a + dv[i].corr) * dv[i].recip) >> dv[i].shift);
                    q[i] = v < 0 ? -m : m;
                }
                const Huff& hd = dc[c ? 1 : 0];
                const Huff& ha = ac[c ? 1 : 0];
                // jchuff.c encode_one_block
                const int diff = q[0] - lastDc[c];
                lastDc[c] = q[0];
                int nb = nbitsOf(diff);
                bw.put(hd.code[nb], hd.size[nb]);
                if (nb)
                    bw.put(uint32_t(diff < 0 ? diff - 1 : diff) & ((1u << nb) - 1), nb);
                int run = 0;
                for (int k = 1; k < 64; ++k)
                {
                    const int v = q[kNatural[k]];
                    if (v == 0)
                    {
                        ++run;
                        continue;
                    }
                    while (run > 15)
                    {
                        bw.put(ha.code[0xF0], ha.size[0xF0]);
                        run -= 16;
                    }
                    nb = nbitsOf(v);
                    const int sym = (run << 4) | nb;
                    bw.put(ha.code[sym], ha.size[sym]);
                    bw.put(uint32_t(v < 0 ? v - 1 : v) & ((1u << nb) - 1), nb);
                    run = 0;
                }
                if (run > 0)
                    bw.put(ha.code[0x00], ha.size[0x00]);
            }
        }
    bw.flush();
    marker(o, 0xD9, -1);
    *size = o.size();
    if (!out)
        return 0;
    if (o.size() > cap)
        return -2;
    std::memcpy(out, o.data(), o.size());
    return 0;
}
