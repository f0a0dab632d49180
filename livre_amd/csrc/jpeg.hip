// Frame grab as JPEG on the device: FrameGrabber::_encodeJpeg
// (livre/eq/FrameGrabber.cpp:49-84) hands the frame to libjpeg-turbo's
// tjCompress2 with BGRA pixels, 4:4:4, quality 100 and bottom-up rows. The
// bytes written here are those of that call: baseline sequential JPEG with the
// accurate integer DCT (which tjCompress2 selects for quality >= 96), the
// standard tables of ITU-T T.81 Annex K scaled to the quality, JFIF 1.01.
//
// Kernels, all on one stream:
//   dct      8 threads per 8x8 MCU: colour conversion, row and column DCT
//            through LDS, quantisation, zigzag order; coefficients to HBM and
//            the AC code length of each block
//   bits     DC code length (difference to the previous block of the
//            component) + block exclusive scan over chunks of 1024 blocks
//   scan     one workgroup: exclusive scan of the chunk totals (bit offsets)
//   clear    zero the words the scan will cover
//   emit     8 threads per MCU again, each writes the codes of 8 zigzag
//            positions of each block at its bit offset (atomic OR into
//            big-endian words); the last block pads with 1 bits
//   ffcount  0xFF bytes per 4096-byte chunk of the scan
//   ffscan   one workgroup: exclusive scan of those counts
//   stuff    copy every byte to its place, a 0x00 after each 0xFF
//            (T.81 F.1.2.3), then EOI
// The JFIF header is written by the host; the stuffed scan follows it.
#include "lvr_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace lvr
{
namespace
{
constexpr int kMcuPerWg = 32;          // 8 threads per MCU, 256 per workgroup
constexpr int kBlocksPerChunk = 1024;  // bit-length scan: 4 blocks per thread
constexpr int kStuffChunk = 256 * 16;  // byte stuffing: 16 bytes per thread
constexpr int kMaxBlockBits = 1664;    // 22 DC + 63 x 26 AC, rounded to 32
constexpr int kMaxDim = 16384;

// T.81 Annex K.1 quantisation tables, natural order
const int kStdLuma[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                          14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                          18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                          49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const int kStdChroma[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                            24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                            99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                            99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
// zigzag position -> natural index (T.81 Figure A.6)
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
// T.81 Annex K.3 Huffman tables: counts per code length, then symbols
const uint8_t kDcBits[2][16] = {{0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0},
                                {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0}};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcBits[2][16] = {{0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d},
                                {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77}};
const uint8_t kAcVals[2][162] = {
    {0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61,
     0x07, 0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52,
     0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25,
     0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45,
     0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64,
     0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
     0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99,
     0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6,
     0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3,
     0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8,
     0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa},
    {0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61,
     0x71, 0x13, 0x22, 0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33,
     0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18,
     0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44,
     0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63,
     0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
     0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97,
     0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4,
     0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca,
     0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7,
     0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa}};

// Per-quality device tables ([0] luma, [1] chroma).
struct JpegTables
{
    // quantiser: q = ((|x| + corr) * recip) >> shift, natural order
    uint32_t recip[2][64];
    uint16_t corr[2][64];
    uint8_t shift[2][64];
    uint8_t zigzagOf[64]; // natural index -> zigzag position
    uint8_t dcLen[2][16];
    uint16_t dcCode[2][16];
    uint8_t acLen[2][256];
    uint16_t acCode[2][256];
};

struct JpegMeta
{
    unsigned long long totalBits; // entropy-coded bits before padding
    unsigned long long nBytes;    // bytes after padding, before stuffing
    unsigned long long outBytes;  // stuffed scan + EOI
};

// jcdctmgr.c compute_reciprocal of libjpeg-turbo for a divisor quantval << 3
// (the accurate DCT's output carries a factor 8): the rounded reciprocal and
// its correction make the multiply-shift equal to the table's rounding.
void reciprocal(uint32_t divisor, uint32_t* recip, uint16_t* corr, uint8_t* shift)
{
    const int b = 31 - __builtin_clz(divisor);
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
    *recip = fq;
    *corr = uint16_t(c);
    *shift = uint8_t(r);
}

// jcparam.c jpeg_quality_scaling + jpeg_add_quant_table with force_baseline
// (tjCompress2 passes TRUE).
void qualityTable(const int base[64], int quality, int out[64])
{
    const int scale = quality < 50 ? 5000 / quality : 200 - quality * 2;
    for (int i = 0; i < 64; ++i)
    {
        long t = (long(base[i]) * scale + 50L) / 100L;
        out[i] = int(t < 1 ? 1 : (t > 255 ? 255 : t));
    }
}

// Canonical codes from the counts (T.81 Annex C).
void huffman(const uint8_t bits[16], const uint8_t* vals, uint16_t* code, uint8_t* len)
{
    uint32_t c = 0;
    int p = 0;
    for (int l = 1; l <= 16; ++l)
    {
        for (int i = 0; i < bits[l - 1]; ++i, ++p)
        {
            code[vals[p]] = uint16_t(c++);
            len[vals[p]] = uint8_t(l);
        }
        c <<= 1;
    }
}

void buildTables(int quality, JpegTables& t, int q[2][64])
{
    std::memset(&t, 0, sizeof t);
    qualityTable(kStdLuma, quality, q[0]);
    qualityTable(kStdChroma, quality, q[1]);
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < 64; ++i)
            reciprocal(uint32_t(q[k][i]) << 3, &t.recip[k][i], &t.corr[k][i], &t.shift[k][i]);
    for (int z = 0; z < 64; ++z)
        t.zigzagOf[kZigzag[z]] = uint8_t(z);
    for (int k = 0; k < 2; ++k)
    {
        huffman(kDcBits[k], kDcVals, t.dcCode[k], t.dcLen[k]);
        huffman(kAcBits[k], kAcVals[k], t.acCode[k], t.acLen[k]);
    }
}

void put16(std::vector<uint8_t>& o, int v)
{
    o.push_back(uint8_t(v >> 8));
    o.push_back(uint8_t(v));
}

// jcmarker.c write_file_header / write_frame_header / write_scan_header:
// SOI, JFIF APP0 1.01 (aspect 1:1), DQT per table, SOF0, DHT per table, SOS.
std::vector<uint8_t> header(int width, int height, const int q[2][64])
{
    std::vector<uint8_t> o = {0xFF, 0xD8, 0xFF, 0xE0};
    put16(o, 16);
    const uint8_t jfif[14] = {'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    o.insert(o.end(), jfif, jfif + 14);
    for (int k = 0; k < 2; ++k)
    {
        o.push_back(0xFF);
        o.push_back(0xDB);
        put16(o, 67);
        o.push_back(uint8_t(k));
        for (int z = 0; z < 64; ++z)
            o.push_back(uint8_t(q[k][kZigzag[z]]));
    }
    o.push_back(0xFF);
    o.push_back(0xC0);
    put16(o, 17);
    o.push_back(8);
    put16(o, height);
    put16(o, width);
    const uint8_t comps[10] = {3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1};
    o.insert(o.end(), comps, comps + 10);
    for (int k = 0; k < 2; ++k)
        for (int ac = 0; ac < 2; ++ac)
        {
            const uint8_t* bits = ac ? kAcBits[k] : kDcBits[k];
            const uint8_t* vals = ac ? kAcVals[k] : kDcVals;
            int n = 0;
            for (int i = 0; i < 16; ++i)
                n += bits[i];
            o.push_back(0xFF);
            o.push_back(0xC4);
            put16(o, 2 + 1 + 16 + n);
            o.push_back(uint8_t(ac << 4 | k));
            o.insert(o.end(), bits, bits + 16);
            o.insert(o.end(), vals, vals + n);
        }
    const uint8_t sos[14] = {0xFF, 0xDA, 0, 12, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0};
    o.insert(o.end(), sos, sos + 14);
    return o;
}

// ---- device side ----

// jfdctint.c jpeg_fdct_islow on one line of 8 (CONST_BITS 13, PASS1_BITS 2):
// the row pass keeps 2 extra bits, the column pass removes them.
template <bool ROWS>
__device__ inline void fdct8(int32_t v[8])
{
    constexpr int kOdd = ROWS ? 11 : 15;
    const int32_t tmp0 = v[0] + v[7], tmp7 = v[0] - v[7];
    const int32_t tmp1 = v[1] + v[6], tmp6 = v[1] - v[6];
    const int32_t tmp2 = v[2] + v[5], tmp5 = v[2] - v[5];
    const int32_t tmp3 = v[3] + v[4], tmp4 = v[3] - v[4];
    const int32_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3;
    const int32_t tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
    if (ROWS)
    {
        v[0] = (tmp10 + tmp11) * 4;
        v[4] = (tmp10 - tmp11) * 4;
    }
    else
    {
        v[0] = (tmp10 + tmp11 + 2) >> 2;
        v[4] = (tmp10 - tmp11 + 2) >> 2;
    }
    constexpr int32_t kRound = 1 << (kOdd - 1);
    const int32_t z1e = (tmp12 + tmp13) * 4433;
    v[2] = (z1e + tmp13 * 6270 + kRound) >> kOdd;
    v[6] = (z1e - tmp12 * 15137 + kRound) >> kOdd;
    const int32_t z1 = tmp4 + tmp7, z2 = tmp5 + tmp6, z3 = tmp4 + tmp6, z4 = tmp5 + tmp7;
    const int32_t z5 = (z3 + z4) * 9633;
    const int32_t m1 = -z1 * 7373, m2 = -z2 * 20995;
    const int32_t m3 = -z3 * 16069 + z5, m4 = -z4 * 3196 + z5;
    v[7] = (tmp4 * 2446 + m1 + m3 + kRound) >> kOdd;
    v[5] = (tmp5 * 16819 + m2 + m4 + kRound) >> kOdd;
    v[3] = (tmp6 * 25172 + m2 + m3 + kRound) >> kOdd;
    v[1] = (tmp7 * 12299 + m1 + m4 + kRound) >> kOdd;
}

__device__ inline int bitLength(int a) { return a ? 32 - __clz(a) : 0; }

// OR of the 8 lanes' bytes into the block's 64-bit nonzero mask (zigzag
// positions); lanes 8g..8g+7 hold one block.
__device__ inline uint64_t blockMask(uint32_t byteMask, int r)
{
    uint64_t m = uint64_t(byteMask) << (8 * r);
    m |= __shfl_xor(m, 1, 8);
    m |= __shfl_xor(m, 2, 8);
    m |= __shfl_xor(m, 4, 8);
    return m;
}

// Zero run before zigzag position k (> 0) of a block with nonzero mask `nz`.
__device__ inline int runBefore(uint64_t nz, int k)
{
    const uint64_t below = nz & ((1ull << k) - 1) & ~1ull;
    return k - (below ? 63 - __clzll(below) : 0) - 1;
}

// Code length of zigzag positions 8r..8r+7 (AC only; EOB on the last chunk),
// jchuff.c encode_one_block.
__device__ inline uint32_t acChunkBits(const int16_t v[8], int r, uint64_t nz,
                                       const uint8_t* acLen)
{
    uint32_t bits = 0;
    for (int j = 0; j < 8; ++j)
    {
        const int k = 8 * r + j;
        if (k == 0 || v[j] == 0)
            continue;
        const int run = runBefore(nz, k);
        const int nb = bitLength(abs(int(v[j])));
        bits += uint32_t(run >> 4) * acLen[0xF0] + acLen[((run & 15) << 4) | nb] + nb;
    }
    if (r == 7 && !(nz >> 63))
        bits += acLen[0x00];
    return bits;
}

__device__ inline void unpack8(const uint4 p, int16_t v[8])
{
    const uint32_t w[4] = {p.x, p.y, p.z, p.w};
    for (int i = 0; i < 4; ++i)
    {
        v[2 * i] = int16_t(w[i] & 0xFFFF);
        v[2 * i + 1] = int16_t(w[i] >> 16);
    }
}

__global__ __launch_bounds__(256) void jpegDctKernel(const uint8_t* __restrict__ px, int w,
                                                     int h, int pitch, int bgra, int bottomUp,
                                                     int nMcuX, int nMcu,
                                                     const JpegTables* __restrict__ tabs,
                                                     int16_t* __restrict__ coef,
                                                     uint32_t* __restrict__ blkBits)
{
    __shared__ int32_t rows[kMcuPerWg][3][64];
    __shared__ __align__(16) int16_t zz[kMcuPerWg][3][64];
    __shared__ uint32_t sRecip[2][64];
    __shared__ uint16_t sCorr[2][64];
    __shared__ uint8_t sShift[2][64];
    __shared__ uint8_t sZig[64];
    __shared__ uint8_t sAcLen[2][256];
    const int t = threadIdx.x;
    if (t < 128)
    {
        sRecip[t >> 6][t & 63] = tabs->recip[t >> 6][t & 63];
        sCorr[t >> 6][t & 63] = tabs->corr[t >> 6][t & 63];
        sShift[t >> 6][t & 63] = tabs->shift[t >> 6][t & 63];
    }
    if (t < 64)
        sZig[t] = tabs->zigzagOf[t];
    for (int i = t; i < 512; i += 256)
        sAcLen[i >> 8][i & 255] = tabs->acLen[i >> 8][i & 255];

    const int lm = t >> 3, r = t & 7;
    const int m = blockIdx.x * kMcuPerWg + lm;
    const bool live = m < nMcu;
    const int mcu = live ? m : nMcu - 1; // spare lanes redo the last MCU, store nothing
    const int mx = mcu % nMcuX, my = mcu / nMcuX;
    // jcprepct.c / jcsample.c: rows and columns past the edge repeat the last
    const int iy = min(my * 8 + r, h - 1);
    const uint8_t* row = px + size_t(bottomUp ? h - 1 - iy : iy) * size_t(pitch);
    int32_t y[8], cb[8], cr[8];
    // the row's 8 pixels: two 16-byte loads when they lie inside the image
    // and are 16-byte aligned, else one load per pixel with the edge repeated
    uint32_t px8[8];
    const uint8_t* seg = row + 32 * size_t(mx);
    if (mx * 8 + 8 <= w && (reinterpret_cast<uintptr_t>(seg) & 15) == 0)
    {
        const uint4 a = reinterpret_cast<const uint4*>(seg)[0];
        const uint4 b = reinterpret_cast<const uint4*>(seg)[1];
        px8[0] = a.x, px8[1] = a.y, px8[2] = a.z, px8[3] = a.w;
        px8[4] = b.x, px8[5] = b.y, px8[6] = b.z, px8[7] = b.w;
    }
    else
        for (int c = 0; c < 8; ++c)
            px8[c] = *reinterpret_cast<const uint32_t*>(row + 4 * size_t(min(mx * 8 + c, w - 1)));
    for (int c = 0; c < 8; ++c)
    {
        const uint32_t p = px8[c];
        const int32_t c0 = int32_t(p & 0xFF), g = int32_t((p >> 8) & 0xFF),
                      c2 = int32_t((p >> 16) & 0xFF);
        const int32_t R = bgra ? c2 : c0, B = bgra ? c0 : c2;
        // jccolor.c rgb_ycc_convert, 16-bit fixed point, then the level shift
        y[c] = ((19595 * R + 38470 * g + 7471 * B + 32768) >> 16) - 128;
        cb[c] = ((-11059 * R - 21709 * g + 32768 * B + (128 << 16) + 32767) >> 16) - 128;
        cr[c] = ((32768 * R - 27439 * g - 5329 * B + (128 << 16) + 32767) >> 16) - 128;
    }
    fdct8<true>(y);
    fdct8<true>(cb);
    fdct8<true>(cr);
    for (int k = 0; k < 8; ++k)
    {
        rows[lm][0][8 * r + k] = y[k];
        rows[lm][1][8 * r + k] = cb[k];
        rows[lm][2][8 * r + k] = cr[k];
    }
    __syncthreads();
    // column r of each component; quantise into zigzag order
    for (int comp = 0; comp < 3; ++comp)
    {
        int32_t v[8];
        for (int k = 0; k < 8; ++k)
            v[k] = rows[lm][comp][8 * k + r];
        fdct8<false>(v);
        const int tq = comp ? 1 : 0;
        for (int k = 0; k < 8; ++k)
        {
            const int n = 8 * k + r;
            const uint32_t a = uint32_t(abs(v[k]));
            const int32_t q = int32_t(((a + sCorr[tq][n]) * sRecip[tq][n]) >> sShift[tq][n]);
            zz[lm][comp][sZig[n]] = int16_t(v[k] < 0 ? -q : q);
        }
    }
    __syncthreads();
    // positions 8r..8r+7 of each block: store, and their code lengths
    for (int comp = 0; comp < 3; ++comp)
    {
        const uint4 p = *reinterpret_cast<const uint4*>(&zz[lm][comp][8 * r]);
        int16_t v[8];
        unpack8(p, v);
        uint32_t byteMask = 0;
        for (int j = 0; j < 8; ++j)
            byteMask |= uint32_t(v[j] != 0) << j;
        const uint64_t nz = blockMask(byteMask, r);
        uint32_t bits = acChunkBits(v, r, nz, sAcLen[comp ? 1 : 0]);
        bits += __shfl_xor(bits, 1, 8);
        bits += __shfl_xor(bits, 2, 8);
        bits += __shfl_xor(bits, 4, 8);
        if (live)
        {
            const size_t b = size_t(m) * 3 + comp;
            *reinterpret_cast<uint4*>(coef + b * 64 + 8 * r) = p;
            if (r == 0)
                blkBits[b] = bits;
        }
    }
}

// Exclusive scan of one value per thread over a 256-thread workgroup.
__device__ inline uint32_t wgExclusiveScan(uint32_t v, uint32_t* sWave, uint32_t* total)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t inc = v;
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d)
            inc += o;
    }
    if (lane == 63)
        sWave[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int i = 0; i < 4; ++i)
    {
        before += i < wave ? sWave[i] : 0;
        all += sWave[i];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

__global__ __launch_bounds__(256) void jpegBitsKernel(const int16_t* __restrict__ coef,
                                                      const uint32_t* __restrict__ blkBits,
                                                      int nBlocks,
                                                      const JpegTables* __restrict__ tabs,
                                                      uint32_t* __restrict__ blkOff,
                                                      uint32_t* __restrict__ chunkBits)
{
    __shared__ uint32_t sWave[4];
    const int base = blockIdx.x * kBlocksPerChunk + 4 * threadIdx.x;
    uint32_t bits[4], sum = 0;
    for (int i = 0; i < 4; ++i)
    {
        const int b = base + i;
        bits[i] = 0;
        if (b < nBlocks)
        {
            const int diff = int(coef[size_t(b) * 64]) - (b >= 3 ? int(coef[size_t(b - 3) * 64]) : 0);
            const int nb = bitLength(abs(diff));
            bits[i] = blkBits[b] + tabs->dcLen[b % 3 ? 1 : 0][nb] + nb;
        }
        sum += bits[i];
    }
    uint32_t total;
    uint32_t off = wgExclusiveScan(sum, sWave, &total);
    for (int i = 0; i < 4; ++i)
    {
        if (base + i < nBlocks)
            blkOff[base + i] = off;
        off += bits[i];
    }
    if (threadIdx.x == 0)
        chunkBits[blockIdx.x] = total;
}

// One workgroup: exclusive scan of n counts into 64-bit offsets.
__device__ inline unsigned long long scanCounts(const uint32_t* __restrict__ in, int n,
                                                unsigned long long* __restrict__ out,
                                                uint32_t* sWave)
{
    unsigned long long carry = 0;
    for (int base = 0; base < n; base += 256)
    {
        const int i = base + int(threadIdx.x);
        const uint32_t v = i < n ? in[i] : 0;
        uint32_t total;
        const uint32_t ex = wgExclusiveScan(v, sWave, &total);
        if (i < n)
            out[i] = carry + ex;
        carry += total;
    }
    return carry;
}

// nWords bounds every later index: a block never exceeds kMaxBlockBits, and
// the clamp keeps the buffers safe should that bound ever fail.
__global__ __launch_bounds__(256) void jpegChunkScanKernel(const uint32_t* __restrict__ chunkBits,
                                                           int nChunks,
                                                           unsigned long long* __restrict__ chunkOff,
                                                           JpegMeta* meta,
                                                           unsigned long long nWords)
{
    __shared__ uint32_t sWave[4];
    unsigned long long bits = scanCounts(chunkBits, nChunks, chunkOff, sWave);
    if (threadIdx.x == 0)
    {
        bits = bits < 32 * (nWords - 1) ? bits : 32 * (nWords - 1);
        meta->totalBits = bits;
        meta->nBytes = (bits + 7) / 8;
    }
}

__global__ __launch_bounds__(256) void jpegClearKernel(uint32_t* __restrict__ words,
                                                       const JpegMeta* __restrict__ meta)
{
    const unsigned long long n = (meta->totalBits + 31) / 32;
    for (unsigned long long i = blockIdx.x * 256ull + threadIdx.x; i < n;
         i += 256ull * gridDim.x)
        words[i] = 0;
}

// Big-endian bit writer aligned to the 32-bit word grid; the first word is
// shared with the previous writer, so words are ORed in.
struct BitSink
{
    uint32_t* words;
    unsigned long long wi;
    int fill;
    unsigned long long acc = 0;
    unsigned long long nWords;
    __device__ BitSink(uint32_t* w, unsigned long long pos, unsigned long long n)
        : words(w), wi(pos >> 5), fill(int(pos & 31)), nWords(n)
    {
    }
    __device__ void put(uint32_t code, int len) // len <= 16
    {
        acc = (acc << len) | code;
        fill += len;
        if (fill >= 32)
        {
            fill -= 32;
            if (wi < nWords)
                atomicOr(&words[wi], __builtin_bswap32(uint32_t(acc >> fill)));
            ++wi;
        }
    }
    __device__ void flush()
    {
        if (fill > 0 && wi < nWords)
            atomicOr(&words[wi], __builtin_bswap32(uint32_t(acc << (32 - fill))));
    }
};

__global__ __launch_bounds__(256) void jpegEmitKernel(const int16_t* __restrict__ coef, int nMcu,
                                                      const JpegTables* __restrict__ tabs,
                                                      const uint32_t* __restrict__ blkOff,
                                                      const unsigned long long* __restrict__ chunkOff,
                                                      uint32_t* __restrict__ words,
                                                      unsigned long long nWords)
{
    __shared__ uint8_t sDcLen[2][16];
    __shared__ uint16_t sDcCode[2][16];
    __shared__ uint8_t sAcLen[2][256];
    __shared__ uint16_t sAcCode[2][256];
    const int t = threadIdx.x;
    if (t < 32)
    {
        sDcLen[t >> 4][t & 15] = tabs->dcLen[t >> 4][t & 15];
        sDcCode[t >> 4][t & 15] = tabs->dcCode[t >> 4][t & 15];
    }
    for (int i = t; i < 512; i += 256)
    {
        sAcLen[i >> 8][i & 255] = tabs->acLen[i >> 8][i & 255];
        sAcCode[i >> 8][i & 255] = tabs->acCode[i >> 8][i & 255];
    }
    __syncthreads();
    const int lm = t >> 3, r = t & 7;
    const int m = blockIdx.x * kMcuPerWg + lm;
    const bool live = m < nMcu;
    const int nBlocks = 3 * nMcu;
    for (int comp = 0; comp < 3; ++comp)
    {
        const int tq = comp ? 1 : 0;
        const int b = (live ? m : nMcu - 1) * 3 + comp;
        const uint4 p = *reinterpret_cast<const uint4*>(coef + size_t(b) * 64 + 8 * r);
        int16_t v[8];
        unpack8(p, v);
        uint32_t byteMask = 0;
        for (int j = 0; j < 8; ++j)
            byteMask |= uint32_t(v[j] != 0) << j;
        const uint64_t nz = blockMask(byteMask, r);
        uint32_t bits = acChunkBits(v, r, nz, sAcLen[tq]);
        int diff = 0, dcBits = 0;
        if (r == 0)
        {
            diff = int(v[0]) - (b >= 3 ? int(coef[size_t(b - 3) * 64]) : 0);
            dcBits = bitLength(abs(diff));
            bits += sDcLen[tq][dcBits] + dcBits;
        }
        uint32_t inc = bits;
        for (int d = 1; d < 8; d <<= 1)
        {
            const uint32_t o = __shfl_up(inc, d, 8);
            if (r >= d)
                inc += o;
        }
        if (!live || bits == 0)
            continue;
        const unsigned long long start =
            chunkOff[b / kBlocksPerChunk] + blkOff[b] + (inc - bits);
        BitSink s(words, start, nWords);
        if (r == 0)
        {
            s.put(sDcCode[tq][dcBits], sDcLen[tq][dcBits]);
            if (dcBits)
                s.put(uint32_t(diff < 0 ? diff - 1 : diff) & ((1u << dcBits) - 1), dcBits);
        }
        for (int j = 0; j < 8; ++j)
        {
            const int k = 8 * r + j;
            if (k == 0 || v[j] == 0)
                continue;
            int run = runBefore(nz, k);
            for (; run > 15; run -= 16)
                s.put(sAcCode[tq][0xF0], sAcLen[tq][0xF0]);
            const int a = int(v[j]);
            const int nb = bitLength(abs(a));
            const int sym = (run << 4) | nb;
            s.put(sAcCode[tq][sym], sAcLen[tq][sym]);
            s.put(uint32_t(a < 0 ? a - 1 : a) & ((1u << nb) - 1), nb);
        }
        if (r == 7 && !(nz >> 63))
            s.put(sAcCode[tq][0x00], sAcLen[tq][0x00]);
        if (r == 7 && b == nBlocks - 1)
        {
            // jchuff.c flush_bits: the last byte is completed with 1 bits
            const int pad = int((8 - ((start + bits) & 7)) & 7);
            if (pad)
                s.put((1u << pad) - 1, pad);
        }
        s.flush();
    }
}

// 16 bytes of the scan from thread t of chunk c; bytes past nBytes are
// reported as absent (count -1 marks nothing to do).
__device__ inline int loadBytes(const uint32_t* words, unsigned long long pos,
                                unsigned long long nBytes, uint8_t bytes[16])
{
    if (pos >= nBytes)
        return 0;
    const uint4 p = *reinterpret_cast<const uint4*>(words + pos / 4);
    const uint32_t w[4] = {p.x, p.y, p.z, p.w};
    const int n = int(nBytes - pos < 16 ? nBytes - pos : 16);
    for (int i = 0; i < 16; ++i)
        bytes[i] = uint8_t(w[i >> 2] >> (8 * (i & 3)));
    return n;
}

__global__ __launch_bounds__(256) void jpegFfCountKernel(const uint32_t* __restrict__ words,
                                                         const JpegMeta* __restrict__ meta,
                                                         uint32_t* __restrict__ ffChunk)
{
    __shared__ uint32_t sWave[4];
    const unsigned long long nBytes = meta->nBytes;
    const unsigned long long nChunks = (nBytes + kStuffChunk - 1) / kStuffChunk;
    for (unsigned long long c = blockIdx.x; c < nChunks; c += gridDim.x)
    {
        uint8_t bytes[16];
        const int n = loadBytes(words, c * kStuffChunk + 16ull * threadIdx.x, nBytes, bytes);
        uint32_t ff = 0;
        for (int i = 0; i < 16; ++i)
            ff += (i < n && bytes[i] == 0xFF) ? 1u : 0u;
        uint32_t total;
        wgExclusiveScan(ff, sWave, &total);
        if (threadIdx.x == 0)
            ffChunk[c] = total;
    }
}

__global__ __launch_bounds__(256) void jpegFfScanKernel(const uint32_t* __restrict__ ffChunk,
                                                        unsigned long long* __restrict__ ffOff,
                                                        JpegMeta* meta,
                                                        unsigned long long* hostOutBytes)
{
    __shared__ uint32_t sWave[4];
    const unsigned long long nBytes = meta->nBytes;
    const int nChunks = int((nBytes + kStuffChunk - 1) / kStuffChunk);
    const unsigned long long ff = scanCounts(ffChunk, nChunks, ffOff, sWave);
    if (threadIdx.x == 0)
    {
        meta->outBytes = nBytes + ff + 2;
        *hostOutBytes = nBytes + ff + 2;
    }
}

__global__ __launch_bounds__(256) void jpegStuffKernel(const uint32_t* __restrict__ words,
                                                       const JpegMeta* __restrict__ meta,
                                                       const unsigned long long* __restrict__ ffOff,
                                                       uint8_t* __restrict__ out)
{
    __shared__ uint32_t sWave[4];
    const unsigned long long nBytes = meta->nBytes;
    const unsigned long long nChunks = (nBytes + kStuffChunk - 1) / kStuffChunk;
    for (unsigned long long c = blockIdx.x; c < nChunks; c += gridDim.x)
    {
        uint8_t bytes[16];
        const unsigned long long pos = c * kStuffChunk + 16ull * threadIdx.x;
        const int n = loadBytes(words, pos, nBytes, bytes);
        uint32_t ff = 0;
        for (int i = 0; i < 16; ++i)
            ff += (i < n && bytes[i] == 0xFF) ? 1u : 0u;
        uint32_t total;
        unsigned long long o = pos + ffOff[c] + wgExclusiveScan(ff, sWave, &total);
        for (int i = 0; i < n; ++i)
        {
            out[o++] = bytes[i];
            if (bytes[i] == 0xFF)
                out[o++] = 0x00;
        }
        if (n > 0 && pos + n == nBytes)
        {
            out[o] = 0xFF; // EOI
            out[o + 1] = 0xD9;
        }
    }
}
} // namespace
} // namespace lvr

struct lvr_jpeg_encoder
{
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    lvr::JpegTables* d_tables = nullptr;
    lvr::JpegTables* h_tables = nullptr; // pinned staging of the tables
    int quality = -1;
    int q[2][64] = {};
    int16_t* d_coef = nullptr;
    uint32_t* d_blkBits = nullptr;
    uint32_t* d_blkOff = nullptr;
    uint32_t* d_chunkBits = nullptr;
    unsigned long long* d_chunkOff = nullptr;
    uint32_t* d_words = nullptr;
    uint8_t* d_out = nullptr;
    uint32_t* d_ffChunk = nullptr;
    unsigned long long* d_ffOff = nullptr;
    lvr::JpegMeta* d_meta = nullptr;
    unsigned long long* h_outBytes = nullptr; // mapped pinned
    unsigned long long* d_hostOutBytes = nullptr;
    uint8_t* d_pixels = nullptr; // host images land here
    size_t pixelCap = 0;
    uint8_t* h_scan = nullptr; // pinned landing of the scan (grow-only)
    size_t hostCap = 0;
    size_t capBlocks = 0;
    hipEvent_t ev[5] = {};
    lvr_jpeg_stats stats = {};
    std::string error;
};

namespace
{
int jfail(lvr_jpeg_encoder* e, int code, const std::string& msg)
{
    e->error = msg;
    return code;
}

#define JPEG_HIP(enc, call)                                                                  \
    do                                                                                       \
    {                                                                                        \
        hipError_t _e = (call);                                                              \
        if (_e != hipSuccess)                                                                \
            return jfail((enc), LVR_EHIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

void freeBuffers(lvr_jpeg_encoder* e)
{
    void* bufs[] = {e->d_coef,  e->d_blkBits, e->d_blkOff,   e->d_chunkBits, e->d_chunkOff,
                    e->d_words, e->d_out,     e->d_ffChunk, e->d_ffOff};
    for (void* p : bufs)
        if (p)
            (void)hipFree(p);
    e->d_coef = nullptr;
    e->d_blkBits = e->d_blkOff = e->d_chunkBits = e->d_ffChunk = nullptr;
    e->d_chunkOff = e->d_ffOff = nullptr;
    e->d_words = nullptr;
    e->d_out = nullptr;
    e->capBlocks = 0;
}

size_t wordCount(size_t blocks)
{
    // whole 4096-byte chunks, so the stuffing kernels' 16-byte loads stay inside
    const size_t words = blocks * (lvr::kMaxBlockBits / 32) + 2;
    return (words + lvr::kStuffChunk / 4 - 1) / (lvr::kStuffChunk / 4) * (lvr::kStuffChunk / 4);
}

int ensureBuffers(lvr_jpeg_encoder* e, size_t blocks)
{
    if (blocks <= e->capBlocks)
        return LVR_OK;
    freeBuffers(e);
    const size_t chunks = (blocks + lvr::kBlocksPerChunk - 1) / lvr::kBlocksPerChunk;
    const size_t words = wordCount(blocks);
    const size_t ffChunks = words * 4 / lvr::kStuffChunk + 1;
    JPEG_HIP(e, hipMalloc((void**)&e->d_coef, blocks * 64 * sizeof(int16_t)));
    JPEG_HIP(e, hipMalloc((void**)&e->d_blkBits, blocks * sizeof(uint32_t)));
    JPEG_HIP(e, hipMalloc((void**)&e->d_blkOff, blocks * sizeof(uint32_t)));
    JPEG_HIP(e, hipMalloc((void**)&e->d_chunkBits, chunks * sizeof(uint32_t)));
    JPEG_HIP(e, hipMalloc((void**)&e->d_chunkOff, chunks * sizeof(unsigned long long)));
    JPEG_HIP(e, hipMalloc((void**)&e->d_words, words * sizeof(uint32_t)));
    JPEG_HIP(e, hipMalloc((void**)&e->d_out, words * 8 + 16));
    JPEG_HIP(e, hipMalloc((void**)&e->d_ffChunk, ffChunks * sizeof(uint32_t)));
    JPEG_HIP(e, hipMalloc((void**)&e->d_ffOff, ffChunks * sizeof(unsigned long long)));
    e->capBlocks = blocks;
    return LVR_OK;
}

int encoderInit(lvr_jpeg_encoder* e, int device, hipStream_t stream)
{
    e->device = device;
    JPEG_HIP(e, hipSetDevice(device));
    if (stream)
        e->stream = stream;
    else
    {
        JPEG_HIP(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        e->ownStream = true;
    }
    JPEG_HIP(e, hipMalloc((void**)&e->d_tables, sizeof(lvr::JpegTables)));
    JPEG_HIP(e, hipHostMalloc((void**)&e->h_tables, sizeof(lvr::JpegTables), hipHostMallocDefault));
    JPEG_HIP(e, hipMalloc((void**)&e->d_meta, sizeof(lvr::JpegMeta)));
    JPEG_HIP(e, hipHostMalloc((void**)&e->h_outBytes, sizeof(unsigned long long),
                              hipHostMallocMapped | hipHostMallocCoherent));
    JPEG_HIP(e, hipHostGetDevicePointer((void**)&e->d_hostOutBytes, e->h_outBytes, 0));
    for (hipEvent_t& ev : e->ev)
        JPEG_HIP(e, hipEventCreate(&ev));
    return LVR_OK;
}

int encode(lvr_jpeg_encoder* e, const void* pixels, int width, int height, int pitch,
           int format, int bottomUp, int quality, uint8_t* out, size_t cap, size_t* size)
{
    if (!pixels || !size || width < 1 || height < 1 || width > lvr::kMaxDim ||
        height > lvr::kMaxDim || pitch < 4 * width || pitch % 4 != 0 ||
        (reinterpret_cast<uintptr_t>(pixels) & 3) != 0 ||
        (format != LVR_PIXEL_RGBA && format != LVR_PIXEL_BGRA) || quality < 1 || quality > 100)
        return jfail(e, LVR_EINVAL,
                     "jpeg: need 1 <= width, height <= 16384, pitch >= 4 * width and a "
                     "multiple of 4, 4-byte aligned pixels, RGBA or BGRA, quality 1..100");
    JPEG_HIP(e, hipSetDevice(e->device));
    const int nMcuX = (width + 7) / 8, nMcu = nMcuX * ((height + 7) / 8);
    const int nBlocks = 3 * nMcu;
    int rc = ensureBuffers(e, size_t(nBlocks));
    if (rc)
        return rc;
    if (quality != e->quality)
    {
        // the previous upload may still be read by a queued encode
        JPEG_HIP(e, hipStreamSynchronize(e->stream));
        lvr::buildTables(quality, *e->h_tables, e->q);
        JPEG_HIP(e, hipMemcpyAsync(e->d_tables, e->h_tables, sizeof(lvr::JpegTables),
                                   hipMemcpyHostToDevice, e->stream));
        e->quality = quality;
    }
    const std::vector<uint8_t> hdr = lvr::header(width, height, e->q);
    const int nChunks = (nBlocks + lvr::kBlocksPerChunk - 1) / lvr::kBlocksPerChunk;
    const size_t words = wordCount(size_t(nBlocks));
    const int stuffGrid = int(std::min<size_t>(words * 4 / lvr::kStuffChunk + 1, 2048));
    const int clearGrid = int(std::min<size_t>((words + 255) / 256, 2048));
    const int mcuGrid = (nMcu + lvr::kMcuPerWg - 1) / lvr::kMcuPerWg;
    JPEG_HIP(e, hipEventRecord(e->ev[0], e->stream));
    hipLaunchKernelGGL(lvr::jpegDctKernel, dim3(mcuGrid), dim3(256), 0, e->stream,
                       static_cast<const uint8_t*>(pixels), width, height, pitch,
                       format == LVR_PIXEL_BGRA ? 1 : 0, bottomUp ? 1 : 0, nMcuX, nMcu,
                       e->d_tables, e->d_coef, e->d_blkBits);
    JPEG_HIP(e, hipEventRecord(e->ev[1], e->stream));
    hipLaunchKernelGGL(lvr::jpegBitsKernel, dim3(nChunks), dim3(256), 0, e->stream, e->d_coef,
                       e->d_blkBits, nBlocks, e->d_tables, e->d_blkOff, e->d_chunkBits);
    hipLaunchKernelGGL(lvr::jpegChunkScanKernel, dim3(1), dim3(256), 0, e->stream,
                       e->d_chunkBits, nChunks, e->d_chunkOff, e->d_meta,
                       (unsigned long long)words);
    hipLaunchKernelGGL(lvr::jpegClearKernel, dim3(clearGrid), dim3(256), 0, e->stream, e->d_words,
                       e->d_meta);
    JPEG_HIP(e, hipEventRecord(e->ev[2], e->stream));
    hipLaunchKernelGGL(lvr::jpegEmitKernel, dim3(mcuGrid), dim3(256), 0, e->stream, e->d_coef,
                       nMcu, e->d_tables, e->d_blkOff, e->d_chunkOff, e->d_words,
                       (unsigned long long)words);
    JPEG_HIP(e, hipEventRecord(e->ev[3], e->stream));
    hipLaunchKernelGGL(lvr::jpegFfCountKernel, dim3(stuffGrid), dim3(256), 0, e->stream,
                       e->d_words, e->d_meta, e->d_ffChunk);
    hipLaunchKernelGGL(lvr::jpegFfScanKernel, dim3(1), dim3(256), 0, e->stream, e->d_ffChunk,
                       e->d_ffOff, e->d_meta, e->d_hostOutBytes);
    hipLaunchKernelGGL(lvr::jpegStuffKernel, dim3(stuffGrid), dim3(256), 0, e->stream,
                       e->d_words, e->d_meta, e->d_ffOff, e->d_out);
    JPEG_HIP(e, hipGetLastError());
    JPEG_HIP(e, hipEventRecord(e->ev[4], e->stream));
    JPEG_HIP(e, hipStreamSynchronize(e->stream));
    const size_t scanBytes = size_t(*e->h_outBytes);
    if (scanBytes > words * 8 + 16)
        return jfail(e, LVR_EHIP, "jpeg: scan length out of range");
    const size_t total = hdr.size() + scanBytes;
    *size = total;
    float ms[4] = {};
    for (int i = 0; i < 4; ++i)
        JPEG_HIP(e, hipEventElapsedTime(&ms[i], e->ev[i], e->ev[i + 1]));
    e->stats.dct_ms = ms[0];
    e->stats.offsets_ms = ms[1];
    e->stats.emit_ms = ms[2];
    e->stats.stuff_ms = ms[3];
    e->stats.device_ms = ms[0] + ms[1] + ms[2] + ms[3];
    e->stats.bytes = total;
    e->stats.width = width;
    e->stats.height = height;
    if (!out)
        return LVR_OK;
    if (cap < total)
        return jfail(e, LVR_ERANGE, "jpeg: output buffer too small (" + std::to_string(total) +
                                        " bytes needed)");
    std::memcpy(out, hdr.data(), hdr.size());
    // the scan comes back by DMA into pinned memory, then one host copy
    // (a copy into pageable memory would be staged in small pieces)
    if (scanBytes > e->hostCap)
    {
        if (e->h_scan)
            (void)hipHostFree(e->h_scan);
        e->h_scan = nullptr;
        e->hostCap = 0;
        const size_t cap2 = std::max(scanBytes, size_t(1) << 20) * 2;
        JPEG_HIP(e, hipHostMalloc((void**)&e->h_scan, cap2, hipHostMallocDefault));
        e->hostCap = cap2;
    }
    JPEG_HIP(e, hipMemcpyAsync(e->h_scan, e->d_out, scanBytes, hipMemcpyDeviceToHost,
                               e->stream));
    JPEG_HIP(e, hipStreamSynchronize(e->stream));
    std::memcpy(out + hdr.size(), e->h_scan, scanBytes);
    return LVR_OK;
}
} // namespace

namespace
{
thread_local std::string gCreateError; // why the last create on this thread failed
}

extern "C" int lvr_jpeg_encoder_create(int device, lvr_jpeg_encoder** out)
{
    if (!out)
        return LVR_EINVAL;
    *out = nullptr;
    auto* e = new lvr_jpeg_encoder();
    const int rc = encoderInit(e, device, nullptr);
    if (rc)
    {
        gCreateError = e->error;
        lvr_jpeg_encoder_destroy(e);
        return rc;
    }
    *out = e;
    return LVR_OK;
}

extern "C" void lvr_jpeg_encoder_destroy(lvr_jpeg_encoder* e)
{
    if (!e)
        return;
    (void)hipSetDevice(e->device);
    if (e->stream)
        (void)hipStreamSynchronize(e->stream);
    freeBuffers(e);
    if (e->d_tables)
        (void)hipFree(e->d_tables);
    if (e->h_tables)
        (void)hipHostFree(e->h_tables);
    if (e->d_meta)
        (void)hipFree(e->d_meta);
    if (e->d_pixels)
        (void)hipFree(e->d_pixels);
    if (e->h_scan)
        (void)hipHostFree(e->h_scan);
    if (e->h_outBytes)
        (void)hipHostFree(e->h_outBytes);
    for (hipEvent_t ev : e->ev)
        if (ev)
            (void)hipEventDestroy(ev);
    if (e->ownStream && e->stream)
        (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int lvr_jpeg_encode(lvr_jpeg_encoder* e, const void* device_pixels, int width,
                               int height, int pitch_bytes, int pixel_format, int bottom_up,
                               int quality, uint8_t* out, size_t cap, size_t* size)
{
    if (!e)
        return LVR_EINVAL;
    return encode(e, device_pixels, width, height, pitch_bytes, pixel_format, bottom_up, quality,
                  out, cap, size);
}

extern "C" int lvr_jpeg_encode_host(lvr_jpeg_encoder* e, const void* host_pixels, int width,
                                    int height, int pitch_bytes, int pixel_format, int bottom_up,
                                    int quality, uint8_t* out, size_t cap, size_t* size)
{
    if (!e)
        return LVR_EINVAL;
    if (!host_pixels || width < 1 || height < 1 || width > lvr::kMaxDim ||
        height > lvr::kMaxDim || pitch_bytes < 4 * width || pitch_bytes % 4 != 0)
        return jfail(e, LVR_EINVAL,
                     "jpeg: need 1 <= width, height <= 16384 and pitch >= 4 * width, a "
                     "multiple of 4");
    JPEG_HIP(e, hipSetDevice(e->device));
    const size_t bytes = size_t(height - 1) * size_t(pitch_bytes) + 4 * size_t(width);
    if (bytes > e->pixelCap)
    {
        JPEG_HIP(e, hipStreamSynchronize(e->stream));
        if (e->d_pixels)
            (void)hipFree(e->d_pixels);
        e->d_pixels = nullptr;
        e->pixelCap = 0;
        JPEG_HIP(e, hipMalloc((void**)&e->d_pixels, bytes));
        e->pixelCap = bytes;
    }
    JPEG_HIP(e, hipMemcpyAsync(e->d_pixels, host_pixels, bytes, hipMemcpyHostToDevice, e->stream));
    return encode(e, e->d_pixels, width, height, pitch_bytes, pixel_format, bottom_up, quality,
                  out, cap, size);
}

extern "C" size_t lvr_jpeg_max_size(int width, int height)
{
    if (width < 1 || height < 1)
        return 0;
    const size_t blocks = size_t(3) * size_t((width + 7) / 8) * size_t((height + 7) / 8);
    return 1024 + blocks * (lvr::kMaxBlockBits / 8) * 2 + 2;
}

extern "C" const char* lvr_jpeg_error(const lvr_jpeg_encoder* e)
{
    return e ? e->error.c_str() : gCreateError.c_str();
}

extern "C" int lvr_jpeg_get_stats(const lvr_jpeg_encoder* e, lvr_jpeg_stats* out)
{
    if (!e || !out)
        return LVR_EINVAL;
    *out = e->stats;
    return LVR_OK;
}

namespace lvr
{
void destroyJpeg(lvr_ctx* ctx)
{
    if (ctx->jpeg)
        lvr_jpeg_encoder_destroy(ctx->jpeg);
    ctx->jpeg = nullptr;
}
} // namespace lvr

extern "C" int lvr_frame_grab_jpeg(lvr_ctx* ctx, int quality, uint8_t* out, size_t cap,
                                   size_t* size)
{
    if (!ctx || !size)
        return LVR_EINVAL;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    if (!ctx->d_frame || ctx->fbW < 1 || ctx->fbH < 1)
        return lvr::fail(ctx, LVR_EINVAL, "frame grab: no frame rendered");
    if (!ctx->jpeg)
    {
        ctx->jpeg = new lvr_jpeg_encoder();
        rc = encoderInit(ctx->jpeg, ctx->device, ctx->stream);
        if (rc)
        {
            const std::string msg = ctx->jpeg->error;
            lvr::destroyJpeg(ctx);
            return lvr::fail(ctx, rc, msg);
        }
    }
    // the framebuffer's row 0 is the bottom row: take rows bottom-up, as the
    // reference's TJFLAG_BOTTOMUP does with Equalizer's read-back image
    rc = encode(ctx->jpeg, ctx->d_frame, ctx->fbW, ctx->fbH, 4 * ctx->fbW, LVR_PIXEL_RGBA, 1,
                quality, out, cap, size);
    if (rc)
        return lvr::fail(ctx, rc, ctx->jpeg->error);
    return LVR_OK;
}

extern "C" int lvr_frame_grab_stats(lvr_ctx* ctx, lvr_jpeg_stats* out)
{
    if (!ctx || !out || !ctx->jpeg)
        return LVR_EINVAL;
    *out = ctx->jpeg->stats;
    return LVR_OK;
}
