// g5_codec.hip -- s4-bp128-d4 on the device, word-for-word the stream that
// CompositeCodec<SIMDBinaryPacking<SIMDIntegratedBlockPacker<CoarseDelta4SIMD,
// true>>, VariableByte<true>> writes (cpu_2d/compression/cpusimd/include/
// codecfactory.h:103-105) into a 16-byte aligned buffer.
//
// Stream (uint32 words):
//   [L = n/128*128, 123456, 123456, 123456]            simdbinarypacking.h:182-183
//   per group of <= 16 miniblocks of 128 values:
//     4 header words, 8-bit widths MSB first            simdbinarypacking.h:195-202
//     per miniblock 4*b words: 4-strided deltas d[k] = x[k] - x[k-4]
//     (x[-4..-1] = previous miniblock's last values, 0 at the start),
//     value k in lane k&3 at bit (k>>2)*b of that lane's word stream;
//     b = 32 stores the raw values, b = 0 stores nothing
//   varint tail of n - L values, deltas restarting at 0, 7-bit groups LSB
//   first, stop bit on the last byte, zero padded to a word (variablebyte.h).
//
// Encode: one wavefront per miniblock computes b with a wave OR-reduce; an
// exclusive scan of 4*b gives every miniblock's position; a second pass packs
// (each lane assembles whole output words from the deltas staged in LDS, so
// stores are coalesced and need no atomics); one wave writes the varint tail.
// Decode: the group headers' positions depend on all earlier widths.  Short
// streams (<= 4096 words) walk them in one workgroup through LDS-staged
// chunks; longer ones find them in parallel: every word index 4 + 4i is a
// candidate group start (groups are 4 + 4 * sum(widths) words), its next
// start follows from its own 16 widths, 5 rounds of pointer doubling give
// the 32-step jump, one thread walks the anchors (every 32nd start) and one
// thread per anchor fills its 32 starts.  Then one wave per miniblock
// unpacks; the 4-lane running sums across miniblocks come from an
// exclusive scan of per-miniblock (reset, sum) records, where b = 32
// miniblocks reset the running value to their raw last four values.
// A round trip of at most kSmallRoundTrip values runs all of these stages in
// one workgroup (k_rt_small).
#include "g5_codec.hpp"

namespace {

constexpr uint32_t kCookie = 123456u;

struct Carry {
    uint32_t r, v[4];
};

__device__ __forceinline__ Carry comb(const Carry &a, const Carry &b)
{
    if (b.r) return b;
    Carry c = a;
#pragma unroll
    for (int j = 0; j < 4; ++j) c.v[j] += b.v[j];
    return c;
}

__device__ __forceinline__ Carry load_carry(const uint32_t *p, int64_t i)
{
    Carry c;
    c.r = p[5 * i];
#pragma unroll
    for (int j = 0; j < 4; ++j) c.v[j] = p[5 * i + 1 + j];
    return c;
}

__device__ __forceinline__ void store_carry(uint32_t *p, int64_t i, const Carry &c)
{
    p[5 * i] = c.r;
#pragma unroll
    for (int j = 0; j < 4; ++j) p[5 * i + 1 + j] = c.v[j];
}

__device__ __forceinline__ uint32_t width_of(uint32_t acc) { return acc ? 32u - __clz(acc) : 0u; }

// ------------------------------------------------------------------ encode

// LDS written by some lanes of a wave, then read by others
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 4 * b of miniblock i (one wave, every lane calls it)
__device__ __forceinline__ int64_t enc_block_words(const uint32_t *__restrict__ in, int64_t i, int l)
{
    const int64_t base = i * 128;
    const uint32_t x0 = in[base + l], x1 = in[base + l + 64];
    const uint32_t r0 = (base + l >= 4) ? in[base + l - 4] : 0u;
    const uint32_t r1 = in[base + l + 60];
    const uint32_t acc = g5_wave_or32((x0 - r0) | (x1 - r1));
    return 4 * (int64_t)width_of(acc);
}

__global__ __launch_bounds__(256) void k_enc_widths(const uint32_t *__restrict__ in, int64_t nblk,
                                                    int64_t *__restrict__ words)
{
    const int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nblk) return;
    const int l = threadIdx.x & 63;
    const int64_t w = enc_block_words(in, i, l);
    if (l == 0) words[i] = w;
}

// miniblock i's packed words (and its group's header if it is a group's
// first): one wave, P = word-count prefix (global or LDS), lds = 128 words
// of the wave's own LDS
__device__ __forceinline__ void enc_pack_block(const uint32_t *__restrict__ in, int64_t nblk, int64_t i,
                                               const int64_t *P, uint32_t *__restrict__ out, uint32_t *lds, int l)
{
    const int64_t base = i * 128;
    const int64_t g = i >> 4;
    const uint32_t b = (uint32_t)((P[i + 1] - P[i]) >> 2);
    if ((i & 15) == 0 && l < 4) {  // group header: widths of miniblocks 16g .. 16g+15
        uint32_t w = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t j = i + 4 * l + t;
            const uint32_t bj = (j < nblk) ? (uint32_t)((P[j + 1] - P[j]) >> 2) : 0u;
            w |= bj << (24 - 8 * t);
        }
        out[4 + 4 * g + P[i] + l] = w;
    }
    uint32_t *dst = out + 8 + 4 * g + P[i];
    const uint32_t x0 = in[base + l], x1 = in[base + l + 64];
    if (b == 32) {
        dst[l] = x0;
        dst[l + 64] = x1;
        return;
    }
    if (b == 0) return;
    const uint32_t r0 = (base + l >= 4) ? in[base + l - 4] : 0u;
    const uint32_t r1 = in[base + l + 60];
    lds[l] = x0 - r0;
    lds[l + 64] = x1 - r1;
    wave_sync();
    for (uint32_t t = l; t < 4 * b; t += 64) {
        const uint32_t j = t & 3, w = t >> 2, lo = 32 * w;
        uint32_t word = 0;
        const uint32_t i0 = lo / b;
        uint32_t i1 = (lo + 31) / b;
        if (i1 > 31) i1 = 31;
        for (uint32_t idx = i0; idx <= i1; ++idx) {
            const int sh = (int)(idx * b) - (int)lo;
            const uint32_t v = lds[4 * idx + j];
            word |= sh >= 0 ? (v << sh) : (v >> (-sh));
        }
        dst[t] = word;
    }
    wave_sync();  // lds is reused by the wave's next miniblock
}

__global__ __launch_bounds__(256) void k_enc_pack(const uint32_t *__restrict__ in, int64_t nblk,
                                                  const int64_t *__restrict__ P, uint32_t *__restrict__ out)
{
    __shared__ uint32_t lds[4][128];
    const int wv = threadIdx.x >> 6;
    const int64_t i = blockIdx.x * 4 + wv;
    if (i >= nblk) return;
    enc_pack_block(in, nblk, i, P, out, lds[wv], threadIdx.x & 63);
}

__device__ __forceinline__ uint32_t varint_len(uint32_t v)
{
    return v < (1u << 7) ? 1u : v < (1u << 14) ? 2u : v < (1u << 21) ? 3u : v < (1u << 28) ? 4u : 5u;
}

constexpr int kTailBytes = 128 * 5 + 8;

// stream header + varint tail; one wave (every lane calls it), pn = the
// miniblocks' word total, bytes = kTailBytes of LDS; returns the stream's
// word count
__device__ int64_t enc_tail_wave(const uint32_t *__restrict__ in, int64_t n, int64_t L, int64_t nblk, int64_t pn,
                                 uint32_t *__restrict__ out, uint8_t *bytes, int l)
{
    if (l == 0) {
        out[0] = (uint32_t)L;
        out[1] = out[2] = out[3] = kCookie;
    }
    const int64_t tailpos = 4 + 4 * ((nblk + 15) / 16) + pn;
    const int64_t nt = n - L;
    for (int k = l; k < kTailBytes; k += 64) bytes[k] = 0;
    wave_sync();
    uint32_t run = 0;
    for (int r = 0; r < 2; ++r) {
        const int k = l + 64 * r;
        uint32_t v = 0, len = 0;
        if (k < nt) {
            v = in[L + k] - (k ? in[L + k - 1] : 0u);
            len = varint_len(v);
        }
        const int64_t inc = g5_wave_incl_scan((int64_t)len);
        uint32_t pos = run + (uint32_t)(inc - len);
        for (uint32_t q = 0; q < len; ++q) {
            const uint32_t last = (q + 1 == len);
            bytes[pos + q] = (uint8_t)(((v >> (7 * q)) & 127u) | (last ? 128u : 0u));
        }
        run += (uint32_t)__shfl(inc, 63, 64);
    }
    wave_sync();
    const uint32_t nw = (run + 3) / 4;
    for (uint32_t w = l; w < nw; w += 64)
        out[tailpos + w] = (uint32_t)bytes[4 * w] | ((uint32_t)bytes[4 * w + 1] << 8) |
                           ((uint32_t)bytes[4 * w + 2] << 16) | ((uint32_t)bytes[4 * w + 3] << 24);
    return tailpos + nw;
}

// one wave.  meta[0] = total words.
__global__ __launch_bounds__(64) void k_enc_tail(const uint32_t *__restrict__ in, int64_t n, int64_t L,
                                                 int64_t nblk, const int64_t *__restrict__ P,
                                                 uint32_t *__restrict__ out, int64_t *__restrict__ meta)
{
    __shared__ uint8_t bytes[kTailBytes];
    const int64_t nw = enc_tail_wave(in, n, L, nblk, P[nblk], out, bytes, threadIdx.x);
    if (threadIdx.x == 0) meta[0] = nw;
}

// ------------------------------------------------------------------ decode

enum { kMetaL = 0, kMetaGroups = 1, kMetaTailPos = 2, kMetaErr = 3, kMetaTail = 4, kMetaAnchors = 5, kMetaFlag = 6,
       kMetaNw = 7 };

constexpr int kWalkChunk = 8192;

// One workgroup: stream header check and the group-header chain.
__global__ __launch_bounds__(256) void k_dec_walk(const uint32_t *__restrict__ in, const int64_t *__restrict__ nwp,
                                                  int64_t cap, int64_t *__restrict__ grp,
                                                  int64_t *__restrict__ meta)
{
    __shared__ uint32_t lds[kWalkChunk];
    __shared__ int64_t s_pos, s_g, s_err;
    const int64_t nwords = *nwp;
    const int64_t L = in[0];
    const int64_t nblk = L / 128, ngroups = (nblk + 15) / 16;
    if (threadIdx.x == 0) {
        int64_t err = 0;
        if (in[1] != kCookie || in[2] != kCookie || in[3] != kCookie) err = 1;
        if (L % 128 != 0 || L > cap) err = 2;
        s_err = err;
        s_pos = 4;
        s_g = 0;
        meta[kMetaL] = L;
        meta[kMetaGroups] = ngroups;
    }
    __syncthreads();
    while (s_err == 0 && s_g < ngroups) {
        const int64_t pos = s_pos;
        const int64_t len = min((int64_t)kWalkChunk, nwords - pos);
        for (int64_t t = threadIdx.x; t < len; t += blockDim.x) lds[t] = in[pos + t];
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t local = 0, g = s_g, err = 0;
            const int64_t g_at_entry = g;
            while (g < ngroups && local + 4 <= len) {
                const int64_t cnt = min((int64_t)16, nblk - 16 * g);
                int64_t size = 4;
                for (int64_t j = 0; j < cnt; ++j) {
                    const uint32_t bj = (lds[local + (j >> 2)] >> (24 - 8 * (j & 3))) & 0xFF;
                    if (bj > 32) err = 3;
                    size += 4 * bj;
                }
                if (err) break;
                grp[g] = pos + local;
                local += size;
                ++g;
            }
            if (!err && g == g_at_entry) err = 4;  // truncated: no header fits
            if (!err && pos + local > nwords) err = 4;
            s_pos = pos + local;
            s_g = g;
            s_err = err;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        meta[kMetaTailPos] = s_pos;
        meta[kMetaErr] = s_err;
    }
}

// Miniblock i of the group whose header is at hp: its data offset (*off)
// and its (reset, lane sums) record; one wave, every lane calls it.
__device__ __forceinline__ Carry dec_block_carry(const uint32_t *__restrict__ in, int64_t hp, int64_t i, int l,
                                                 int64_t *off)
{
    const int64_t g = i >> 4;
    int64_t o = hp + 4;
    for (int64_t j = 16 * g; j < i; ++j) {
        const int64_t jj = j - 16 * g;
        o += 4 * ((in[hp + (jj >> 2)] >> (24 - 8 * (jj & 3))) & 0xFF);
    }
    const int64_t ii = i - 16 * g;
    const uint32_t b = (in[hp + (ii >> 2)] >> (24 - 8 * (ii & 3))) & 0xFF;
    Carry c;
    if (b == 32) {
        c.r = 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) c.v[j] = in[o + 124 + j];
    } else {
        c.r = 0;
        uint32_t s = 0;
        if (b > 0) {
            // lane l holds values k = l and k = l + 64, both of class l & 3
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t k = l + 64 * h, j = k & 3, idx = k >> 2;
                const uint32_t bit = idx * b, w = bit >> 5, sh = bit & 31;
                uint64_t v = in[o + 4 * w + j] >> sh;
                if (sh + b > 32) v |= (uint64_t)in[o + 4 * (w + 1) + j] << (32 - sh);
                s += (uint32_t)(v & ((1ull << b) - 1));
            }
        }
        // sum over lanes of the same class (lane distance multiples of 4)
        s += __shfl_xor(s, 4, 64);
        s += __shfl_xor(s, 8, 64);
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
#pragma unroll
        for (int j = 0; j < 4; ++j) c.v[j] = __shfl(s, j, 64);
    }
    *off = o;
    return c;
}

// One wave per miniblock: data offset and the (reset, lane sums) record.
__global__ __launch_bounds__(256) void k_dec_local(const uint32_t *__restrict__ in, int64_t nblk,
                                                   const int64_t *__restrict__ grp,
                                                   int64_t *__restrict__ off, uint32_t *__restrict__ carry,
                                                   const int64_t *__restrict__ meta)
{
    // on an error, or past the blocks the stream's header names, there are no
    // group positions to read
    const int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nblk || i >= meta[kMetaL] / 128 || meta[kMetaErr]) return;
    const int l = threadIdx.x & 63;
    int64_t o = 0;
    const Carry c = dec_block_carry(in, grp[i >> 4], i, l, &o);
    if (l == 0) {
        off[i] = o;
        store_carry(carry, i, c);
    }
}

// --- exclusive scan of Carry records (tile of 256, recursive over tiles)
__device__ __forceinline__ Carry block_scan_carry(Carry v, Carry *lds, Carry *total)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        Carry prev;
        const bool has = t >= off;
        if (has) prev = lds[t - off];
        __syncthreads();
        if (has) v = comb(prev, v);
        lds[t] = v;
        __syncthreads();
    }
    *total = lds[255];
    Carry ex;
    if (t > 0) ex = lds[t - 1];
    else ex = Carry{0, {0, 0, 0, 0}};
    __syncthreads();
    return ex;
}

__global__ __launch_bounds__(256) void k_carry_tiles(const uint32_t *__restrict__ c, int64_t n,
                                                     uint32_t *__restrict__ agg)
{
    __shared__ Carry lds[256];
    const int64_t i = blockIdx.x * 256 + threadIdx.x;
    Carry v = (i < n) ? load_carry(c, i) : Carry{0, {0, 0, 0, 0}};
    Carry tot;
    block_scan_carry(v, lds, &tot);
    if (threadIdx.x == 0) store_carry(agg, blockIdx.x, tot);
}

__global__ __launch_bounds__(256) void k_carry_apply(uint32_t *__restrict__ c, int64_t n,
                                                     const uint32_t *__restrict__ tile_prefix)
{
    __shared__ Carry lds[256];
    const int64_t i = blockIdx.x * 256 + threadIdx.x;
    Carry v = (i < n) ? load_carry(c, i) : Carry{0, {0, 0, 0, 0}};
    Carry tot;
    Carry ex = block_scan_carry(v, lds, &tot);
    if (tile_prefix) ex = comb(load_carry(tile_prefix, blockIdx.x), ex);
    if (i < n) store_carry(c, i, ex);
}

// ---- parallel group-header search ----
// g5_s4_decode takes the serial walk for streams of at most kSerialWords
// words (the word count is all the host knows before launching; the
// header's checks run on the device); an encode -> decode round trip knows
// the group count from the value count and walks at most kSerialGroups
// groups serially.
constexpr int64_t kSerialWords = 4096;
constexpr int64_t kSerialGroups = 64;
constexpr int kJumpLog = 5;  // anchors every 2^5 groups
constexpr int32_t kBadWidth = -1, kTruncated = -2;

// nxt[i] = index of the group start after candidate start p = 4 + 4i, or
// kBadWidth / kTruncated if a group header at p is impossible
__global__ void k_dec_next(const uint32_t *__restrict__ in, const int64_t *__restrict__ nwp, int64_t ncand,
                           int32_t *__restrict__ nxt)
{
    const int64_t nwords = *nwp;  // candidates past the stream get kTruncated
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ncand; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = 4 + 4 * i;
        int32_t r = kTruncated;
        if (p + 4 <= nwords) {
            int64_t size = 4;
            bool ok = true;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint32_t x = in[p + w];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t b = (x >> (8 * k)) & 0xFF;
                    ok = ok && b <= 32;
                    size += 4 * b;
                }
            }
            if (!ok) r = kBadWidth;
            else if (p + size <= nwords) r = (int32_t)((p + size - 4) / 4);
        }
        nxt[i] = r;
    }
}

// one doubling round: out[i] = in[in[i]] (negative entries propagate)
__global__ void k_dec_jump(const int32_t *__restrict__ in, int64_t ncand, int32_t *__restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ncand; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = in[i];
        out[i] = (a < 0 || a >= ncand) ? (a < 0 ? a : kTruncated) : in[a];
    }
}

// one thread: the stream header's checks (as the serial walk makes them),
// then anchor[k] = start index of group 32k, k = 0..(ngroups-1)/32; the
// jumps between anchors never cross the last group, whose unused widths are
// not read.  meta gets L, the group and anchor counts.
__global__ void k_dec_anchors(const uint32_t *__restrict__ in, int64_t cap, const int64_t *__restrict__ nwp,
                              const int32_t *__restrict__ jmp, int32_t *__restrict__ anchor,
                              int64_t *__restrict__ meta)
{
    const int64_t ncand = (*nwp - 4) / 4 + 1;
    const int64_t L = in[0], ngroups = (L / 128 + 15) / 16;
    meta[kMetaL] = L;
    meta[kMetaGroups] = ngroups;
    meta[kMetaAnchors] = 0;
    int64_t err = 0;
    if (in[1] != kCookie || in[2] != kCookie || in[3] != kCookie) err = 1;
    if (L % 128 != 0 || L > cap) err = 2;
    if (!err && ngroups > ncand) err = 4;  // more groups than possible starts: truncated
    if (err) {
        meta[kMetaErr] = err;
        return;
    }
    if (ngroups == 0) {
        meta[kMetaTailPos] = 4;
        return;
    }
    const int64_t nanch = ((ngroups - 1) >> kJumpLog) + 1;
    int32_t a = 0;
    for (int64_t k = 0; k < nanch; ++k) {
        anchor[k] = a;
        if (k + 1 == nanch) break;
        a = jmp[a];
        if (a < 0) {
            meta[kMetaErr] = a == kBadWidth ? 3 : 4;
            return;
        }
    }
    meta[kMetaAnchors] = nanch;
}

// thread k: group starts 32k .. 32k+31; the last group's own widths (only
// the miniblocks it holds) give the tail position
__global__ void k_dec_fill(const uint32_t *__restrict__ in, const int64_t *__restrict__ nwp,
                           const int32_t *__restrict__ nxt, const int32_t *__restrict__ anchor, int64_t ncand,
                           int64_t *__restrict__ grp, int64_t *__restrict__ meta)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= meta[kMetaAnchors] || meta[kMetaErr]) return;
    const int64_t nwords = *nwp;
    const int64_t ngroups = meta[kMetaGroups], nblk = meta[kMetaL] / 128;
    int32_t a = anchor[k];
    for (int j = 0; j < (1 << kJumpLog); ++j) {
        const int64_t g = (k << kJumpLog) + j;
        if (g >= ngroups) break;
        const int64_t p = 4 + 4 * (int64_t)a;
        grp[g] = p;
        if (g == ngroups - 1) {
            int64_t err = p + 4 > nwords ? 4 : 0, size = 4;
            const int64_t cnt = min((int64_t)16, nblk - 16 * g);
            for (int64_t b = 0; b < cnt && !err; ++b) {
                const uint32_t w = (in[p + (b >> 2)] >> (24 - 8 * (b & 3))) & 0xFF;
                if (w > 32) err = 3;
                size += 4 * w;
            }
            if (!err && p + size > nwords) err = 4;
            if (err) atomicMax((unsigned long long *)&meta[kMetaErr], (unsigned long long)err);
            else meta[kMetaTailPos] = p + size;
            break;
        }
        a = a < ncand ? nxt[a] : kTruncated;
        if (a < 0) {
            atomicMax((unsigned long long *)&meta[kMetaErr], (unsigned long long)(a == kBadWidth ? 3 : 4));
            return;
        }
    }
}

int64_t carry_scratch_records(int64_t n)
{
    int64_t w = 0;
    while (n > 256) {
        n = (n + 255) / 256;
        w += n;
    }
    return w + 1;
}

hipError_t carry_scan(uint32_t *c, int64_t n, uint32_t *scratch, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + 255) / 256;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_carry_apply, 1, 256, 0, s, c, n, (const uint32_t *)nullptr);
        return hipGetLastError();
    }
    uint32_t *agg = scratch;
    hipLaunchKernelGGL(k_carry_tiles, (unsigned)tiles, 256, 0, s, (const uint32_t *)c, n, agg);
    hipError_t e = carry_scan(agg, tiles, scratch + 5 * tiles, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_carry_apply, (unsigned)tiles, 256, 0, s, c, n, (const uint32_t *)agg);
    return hipGetLastError();
}

// Miniblock i's 128 values from its data at o, the group header at hp and
// the running values c before it; one wave, every lane calls it.
__device__ __forceinline__ void dec_write_block(const uint32_t *__restrict__ in, int64_t hp, int64_t i, int64_t o,
                                                const Carry &c, uint32_t *__restrict__ out, int l)
{
    const int64_t g = i >> 4, ii = i - 16 * g;
    const uint32_t b = (in[hp + (ii >> 2)] >> (24 - 8 * (ii & 3))) & 0xFF;
    if (b == 32) {
        out[128 * i + l] = in[o + l];
        out[128 * i + l + 64] = in[o + l + 64];
        return;
    }
    uint32_t d[2] = {0u, 0u};
    if (b > 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t k = l + 64 * h, j = k & 3, idx = k >> 2;
            const uint32_t bit = idx * b, w = bit >> 5, sh = bit & 31;
            uint64_t v = in[o + 4 * w + j] >> sh;
            if (sh + b > 32) v |= (uint64_t)in[o + 4 * (w + 1) + j] << (32 - sh);
            d[h] = (uint32_t)(v & ((1ull << b) - 1));
        }
    }
    // inclusive prefix over lanes of the same class (stride 4)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int dist = 4; dist < 64; dist <<= 1) {
            const uint32_t t = __shfl_up(d[h], dist, 64);
            if (l >= dist) d[h] += t;
        }
    }
    const uint32_t j = l & 3;
    const uint32_t first_half = __shfl(d[0], 60 + j, 64);
    out[128 * i + l] = c.v[j] + d[0];
    out[128 * i + l + 64] = c.v[j] + first_half + d[1];
}

__global__ __launch_bounds__(256) void k_dec_write(const uint32_t *__restrict__ in, int64_t nblk,
                                                   const int64_t *__restrict__ grp,
                                                   const int64_t *__restrict__ off,
                                                   const uint32_t *__restrict__ carry,
                                                   uint32_t *__restrict__ out, const int64_t *__restrict__ meta)
{
    const int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nblk || i >= meta[kMetaL] / 128 || meta[kMetaErr]) return;
    dec_write_block(in, grp[i >> 4], i, off[i], load_carry(carry, i), out, threadIdx.x & 63);
}

// varint tail, one thread; meta[kMetaTail] = values, or kMetaErr on overflow
// flag (round trip; may be null): set to 1 unless the stream held exactly
// n_expect values, L of them in miniblocks
__device__ void dec_tail_thread(const uint32_t *__restrict__ in, int64_t nwords, int64_t L, int64_t cap,
                                uint32_t *__restrict__ out, int64_t *__restrict__ meta, int64_t n_expect,
                                int64_t *__restrict__ flag)
{
    meta[kMetaNw] = nwords;
    if (meta[kMetaErr]) {  // the group positions are not valid
        if (flag) *flag = 1;
        return;
    }
    const int64_t tp = meta[kMetaTailPos];
    const uint8_t *bi = (const uint8_t *)(in + tp);
    const int64_t nb = (nwords - tp) * 4;
    uint32_t prev = 0, v = 0;
    uint32_t sh = 0;
    int64_t o = L;
    for (int64_t q = 0; q < nb; ++q) {
        const uint32_t c = bi[q];
        v += (c & 127u) << sh;
        sh += 7;
        if (c & 128u) {
            prev += v;
            if (o >= cap) {
                meta[kMetaErr] = 5;
                if (flag) *flag = 1;
                return;
            }
            out[o++] = prev;
            v = 0;
            sh = 0;
        }
    }
    meta[kMetaTail] = o - L;
    if (flag && (meta[kMetaL] != L || o != n_expect)) *flag = 1;
}

__global__ void k_dec_tail(const uint32_t *__restrict__ in, const int64_t *__restrict__ nwp, int64_t L, int64_t cap,
                           uint32_t *__restrict__ out, int64_t *__restrict__ meta, int64_t n_expect,
                           int64_t *__restrict__ flag)
{
    dec_tail_thread(in, *nwp, L, cap, out, meta, n_expect, flag);
}

// ---- round trip of a short queue in one workgroup ----
// The same stages as the multi-kernel round trip (widths, their prefix,
// packing, header and tail; then the header walk, per-miniblock records,
// their scan, unpacking and the tail with the count check), for at most
// kSmallRoundTrip values, so that a short level queue costs one launch
// instead of a dozen.  The decode reads the words the encode wrote to
// memory, as a receiver would.
constexpr int64_t kSmallRoundTrip = 8192;  // 32768: C2 53-56 -> 51 GTEPS (one workgroup is slower there)
constexpr int kSmallBlocks = kSmallRoundTrip / 128;  // 64 miniblocks, 4 groups

__global__ __launch_bounds__(256) void k_rt_small(const uint32_t *in, int64_t n, uint32_t *__restrict__ words,
                                                  int64_t *__restrict__ nw_out, uint32_t *out,
                                                  int64_t *__restrict__ meta, int64_t *__restrict__ flag)
{
    __shared__ int64_t s_P[kSmallBlocks + 1];
    __shared__ uint32_t s_lds[4][128];
    __shared__ uint8_t s_bytes[kTailBytes];
    __shared__ int64_t s_grp[kSmallBlocks / 16], s_off[kSmallBlocks];
    __shared__ Carry s_c[kSmallBlocks];
    __shared__ int64_t s_nw, s_err, s_tp, s_L;
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t L = n / 128 * 128, nblk = L / 128;
    if (threadIdx.x < 16) meta[threadIdx.x] = 0;
    // ---- encode
    for (int64_t i = wv; i < nblk; i += 4) {
        const int64_t w = enc_block_words(in, i, l);
        if (l == 0) s_P[i + 1] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_P[0] = 0;
        for (int64_t i = 0; i < nblk; ++i) s_P[i + 1] += s_P[i];
    }
    __syncthreads();
    for (int64_t i = wv; i < nblk; i += 4) enc_pack_block(in, nblk, i, s_P, words, s_lds[wv], l);
    if (wv == 0) {
        const int64_t nw = enc_tail_wave(in, n, L, nblk, s_P[nblk], words, s_bytes, l);
        if (l == 0) {
            s_nw = nw;
            *nw_out = nw;
        }
    }
    __threadfence();  // the words are read back below, by other waves
    __syncthreads();
    // ---- decode: header checks and the group-header walk
    const int64_t nwords = s_nw;
    if (threadIdx.x == 0) {
        const int64_t Ls = words[0];
        int64_t err = 0, pos = 4;
        if (words[1] != kCookie || words[2] != kCookie || words[3] != kCookie) err = 1;
        if (Ls % 128 != 0 || Ls > n) err = 2;
        const int64_t nb = err ? 0 : Ls / 128, ng = (nb + 15) / 16;
        for (int64_t g = 0; g < ng && !err; ++g) {
            if (pos + 4 > nwords) {
                err = 4;
                break;
            }
            const int64_t cnt = min((int64_t)16, nb - 16 * g);
            int64_t size = 4;
            for (int64_t j = 0; j < cnt; ++j) {
                const uint32_t bj = (words[pos + (j >> 2)] >> (24 - 8 * (j & 3))) & 0xFF;
                if (bj > 32) err = 3;
                size += 4 * bj;
            }
            s_grp[g] = pos;
            pos += size;
            if (pos > nwords) err = 4;
        }
        s_L = err ? 0 : Ls;
        s_err = err;
        s_tp = pos;
        meta[kMetaL] = Ls;
        meta[kMetaGroups] = ng;
        meta[kMetaTailPos] = pos;
        meta[kMetaErr] = err;
    }
    __syncthreads();
    const int64_t nbd = s_L / 128;  // 0 on an error
    for (int64_t i = wv; i < nbd; i += 4) {
        int64_t o = 0;
        const Carry c = dec_block_carry(words, s_grp[i >> 4], i, l, &o);
        if (l == 0) {
            s_off[i] = o;
            s_c[i] = c;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive scan of the records
        Carry run{0, {0, 0, 0, 0}};
        for (int64_t i = 0; i < nbd; ++i) {
            const Carry c = s_c[i];
            s_c[i] = run;
            run = comb(run, c);
        }
    }
    __syncthreads();
    for (int64_t i = wv; i < nbd; i += 4) dec_write_block(words, s_grp[i >> 4], i, s_off[i], s_c[i], out, l);
    if (threadIdx.x == 0) dec_tail_thread(words, nwords, s_L, n, out, meta, n, flag);
}

}  // namespace

namespace g5 {

size_t S4Codec::bind(uint8_t *base, int64_t n)
{
    const int64_t nblk = n / 128 + 2;
    const int64_t ngrp = nblk / 16 + 2;
    const int64_t ncand = (int64_t)g5_s4_max_words((size_t)n) / 4 + 2;
    size_t off = 0;
    // pieces start 256 B apart from each other
    auto take = [&](size_t bytes) -> uint8_t * {
        uint8_t *p = base ? base + off : nullptr;
        off += (bytes + 255) / 256 * 256;
        return p;
    };
    uint8_t *words = take((size_t)(nblk + 1) * 8);
    uint8_t *scan = take((scan_scratch_words(nblk + 1) + 8) * 8);
    uint8_t *offs = take((size_t)nblk * 8);
    uint8_t *carry = take((size_t)nblk * 20);
    uint8_t *carry_scan = take((size_t)(carry_scratch_records(nblk) + 2) * 20);
    uint8_t *grp = take((size_t)ngrp * 8);
    uint8_t *nxt = take((size_t)ncand * 4);
    uint8_t *jmp = take((size_t)ncand * 4);
    uint8_t *jtmp = take((size_t)ncand * 4);
    uint8_t *anchor = take((size_t)(ngrp / 32 + 2) * 4);
    uint8_t *meta = take(16 * 8);
    uint8_t *nw = take(2 * 8);
    if (base) {
        d_base = base;
        d_words = (int64_t *)words;
        d_scan = (int64_t *)scan;
        d_off = (int64_t *)offs;
        d_carry = (uint32_t *)carry;
        d_carry_scan = (uint32_t *)carry_scan;
        d_grp = (int64_t *)grp;
        d_nxt = (int32_t *)nxt;
        d_jmp = (int32_t *)jmp;
        d_jtmp = (int32_t *)jtmp;
        d_anchor = (int32_t *)anchor;
        d_meta = (int64_t *)meta;
        d_nw = (int64_t *)nw;
        ncand_cap = ncand;
        cap = n;
    }
    return off;
}

int S4Codec::reserve(int64_t n)
{
    if (n <= cap && d_meta) return G5_OK;
    release();
    uint8_t *base = nullptr;
    hipError_t e = hipMalloc(&base, bind(nullptr, n));
    if (e == hipSuccess) e = hipHostMalloc(&h_meta, 16 * 8, hipHostMallocDefault);
    if (e != hipSuccess) {
        if (base) (void)hipFree(base);
        h_meta = nullptr;
        return fail_hip(e, "S4Codec::reserve", __FILE__, __LINE__);
    }
    bind(base, n);
    owned = true;
    return G5_OK;
}

void S4Codec::release()
{
    if (owned) {
        if (d_base) (void)hipFree(d_base);
        if (h_meta) (void)hipHostFree(h_meta);
        h_meta = nullptr;
    }
    owned = false;
    d_base = nullptr;
    d_words = d_scan = d_off = d_grp = d_meta = nullptr;
    d_carry = d_carry_scan = nullptr;
    d_nxt = d_jmp = d_jtmp = d_anchor = nullptr;
    d_nw = nullptr;
    ncand_cap = 0;
    cap = 0;
}

// the encode kernels; the stream's word count goes to d_nw[0]
int S4Codec::encode_launch(const uint32_t *d_in, int64_t n, uint32_t *d_out, hipStream_t s)
{
    const int64_t L = n / 128 * 128, nblk = L / 128;
    if (nblk > 0) {
        hipLaunchKernelGGL(k_enc_widths, (unsigned)((nblk + 3) / 4), 256, 0, s, d_in, nblk, d_words);
        G5_HIP(hipGetLastError());
    }
    G5_HIP(exclusive_scan_i64(d_words, nblk, d_words + nblk, d_scan, s));
    if (nblk > 0) {
        hipLaunchKernelGGL(k_enc_pack, (unsigned)((nblk + 3) / 4), 256, 0, s, d_in, nblk,
                           (const int64_t *)d_words, d_out);
        G5_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_enc_tail, 1, 64, 0, s, d_in, n, L, nblk, (const int64_t *)d_words, d_out, d_nw);
    G5_HIP(hipGetLastError());
    return G5_OK;
}

int S4Codec::encode(const uint32_t *d_in, int64_t n, uint32_t *d_out, size_t cap_words, size_t *nwords,
                    hipStream_t s)
{
    G5_REQUIRE(n >= 0 && cap_words >= g5_s4_max_words((size_t)n), G5_E_OVERFLOW,
               "s4 encode: output capacity %zu below the bound for %lld values", cap_words, (long long)n);
    int rc = reserve(n);
    if (rc) return rc;
    if (acct) G5_HIP(acct->start());
    rc = encode_launch(d_in, n, d_out, s);
    if (rc) return rc;
    if (acct) G5_HIP(acct->stop());
    G5_HIP(hipMemcpyAsync(h_meta, d_nw, 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    *nwords = (size_t)h_meta[0];
    if (acct) G5_HIP(acct->commit(G5_UNIT_ENCODE, 4.0 * (double)n + 4.0 * (double)*nwords, 1));
    return G5_OK;
}

// the group-header search of a stream whose word count is *nwp on the device
// and at most words_bound: meta gets L, the groups, their positions in grp
// and the tail position, or an error
int S4Codec::search_launch(const uint32_t *d_in, const int64_t *nwp, int64_t words_bound, int64_t cap_values,
                           bool serial, hipStream_t s)
{
    G5_HIP(fill_i64(d_meta, 16, 0, s));
    const int64_t ncand = (words_bound - 4) / 4 + 1;  // bound on the 4-word group starts
    if (serial || ncand > ncand_cap) {
        hipLaunchKernelGGL(k_dec_walk, 1, 256, 0, s, d_in, nwp, cap_values, d_grp, d_meta);
        G5_HIP(hipGetLastError());
        return G5_OK;
    }
    // a stream holds at most ncand groups (k_dec_anchors rejects more)
    const int64_t nanch_max = ((ncand - 1) >> kJumpLog) + 1;
    const unsigned grid = (unsigned)grid_for(ncand, 256, 16384);
    hipLaunchKernelGGL(k_dec_next, grid, 256, 0, s, d_in, nwp, ncand, d_nxt);
    const int32_t *src = d_nxt;
    int32_t *bufs[2] = {d_jmp, d_jtmp};
    for (int r = 0; r < kJumpLog; ++r) {
        hipLaunchKernelGGL(k_dec_jump, grid, 256, 0, s, src, ncand, bufs[r & 1]);
        src = bufs[r & 1];
    }
    hipLaunchKernelGGL(k_dec_anchors, 1, 1, 0, s, d_in, cap_values, nwp, src, d_anchor, d_meta);
    hipLaunchKernelGGL(k_dec_fill, (unsigned)((nanch_max + 255) / 256), 256, 0, s, d_in, nwp,
                       (const int32_t *)d_nxt, (const int32_t *)d_anchor, ncand, d_grp, d_meta);
    G5_HIP(hipGetLastError());
    return G5_OK;
}

// miniblocks and the varint tail of a stream of L values whose header search
// has run (every kernel does nothing when meta holds an error)
int S4Codec::unpack_launch(const uint32_t *d_in, const int64_t *nwp, int64_t L, int64_t cap_values, uint32_t *d_out,
                           int64_t *d_flag, hipStream_t s)
{
    const int64_t nblk = L / 128;
    if (nblk > 0) {
        hipLaunchKernelGGL(k_dec_local, (unsigned)((nblk + 3) / 4), 256, 0, s, d_in, nblk, (const int64_t *)d_grp,
                           d_off, d_carry, (const int64_t *)d_meta);
        G5_HIP(hipGetLastError());
        G5_HIP(carry_scan(d_carry, nblk, d_carry_scan, s));
        hipLaunchKernelGGL(k_dec_write, (unsigned)((nblk + 3) / 4), 256, 0, s, d_in, nblk, (const int64_t *)d_grp,
                           (const int64_t *)d_off, (const uint32_t *)d_carry, d_out, (const int64_t *)d_meta);
        G5_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_dec_tail, 1, 1, 0, s, d_in, nwp, L, cap_values, d_out, d_meta, cap_values, d_flag);
    G5_HIP(hipGetLastError());
    return G5_OK;
}

int S4Codec::decode(const uint32_t *d_in, size_t nwords, uint32_t *d_out, size_t cap_values, size_t *n,
                    hipStream_t s)
{
    G5_REQUIRE(nwords >= 4, G5_E_STREAM, "s4 decode: stream of %zu words has no header", nwords);
    int rc = reserve((int64_t)cap_values);
    if (rc) return rc;
    G5_HIP(fill_i64(d_nw + 1, 1, (int64_t)nwords, s));
    if (acct) G5_HIP(acct->start());
    rc = search_launch(d_in, d_nw + 1, (int64_t)nwords, (int64_t)cap_values, (int64_t)nwords <= kSerialWords, s);
    if (rc) return rc;
    if (acct) G5_HIP(acct->stop());
    G5_HIP(hipMemcpyAsync(h_meta, d_meta, 8 * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    if (acct) G5_HIP(acct->commit(G5_UNIT_DECODE, 0, 0));
    G5_REQUIRE(h_meta[kMetaErr] == 0, G5_E_STREAM, "s4 decode: malformed stream (code %lld)",
               (long long)h_meta[kMetaErr]);
    const int64_t L = h_meta[kMetaL];
    if (acct) G5_HIP(acct->start());
    rc = unpack_launch(d_in, d_nw + 1, L, (int64_t)cap_values, d_out, nullptr, s);
    if (rc) return rc;
    if (acct) G5_HIP(acct->stop());
    G5_HIP(hipMemcpyAsync(h_meta, d_meta, 8 * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    G5_REQUIRE(h_meta[kMetaErr] == 0, G5_E_OVERFLOW, "s4 decode: output capacity %zu exceeded", cap_values);
    *n = (size_t)(L + h_meta[kMetaTail]);
    if (acct) G5_HIP(acct->commit(G5_UNIT_DECODE, 4.0 * (double)nwords + 4.0 * (double)*n, 1));
    return G5_OK;
}

int S4Codec::roundtrip(const uint32_t *d_in, int64_t n, uint32_t *d_words_out, size_t cap_words, uint32_t *d_out,
                       int64_t *d_flag, hipStream_t s)
{
    G5_REQUIRE(n >= 0 && cap_words >= g5_s4_max_words((size_t)n), G5_E_OVERFLOW,
               "s4 round trip: word capacity %zu below the bound for %lld values", cap_words, (long long)n);
    int rc = reserve(n);
    if (rc) return rc;
    if (n <= kSmallRoundTrip) {
        hipLaunchKernelGGL(k_rt_small, 1, 256, 0, s, d_in, n, d_words_out, d_nw, d_out, d_meta,
                           d_flag ? d_flag : d_meta + kMetaFlag);
        G5_HIP(hipGetLastError());
        return G5_OK;
    }
    rc = encode_launch(d_in, n, d_words_out, s);
    if (rc) return rc;
    // the stream's length stays on the device; sizes the host needs follow from n
    const int64_t L = n / 128 * 128;
    const int64_t ngroups = (L / 128 + 15) / 16;
    rc = search_launch(d_words_out, d_nw, (int64_t)g5_s4_max_words((size_t)n), n, ngroups <= kSerialGroups, s);
    if (rc) return rc;
    // cap n: the tail kernel's flag check is against exactly n values
    return unpack_launch(d_words_out, d_nw, L, n, d_out, d_flag ? d_flag : d_meta + kMetaFlag, s);
}

int S4Codec::roundtrip_result(size_t *nwords, const char *who)
{
    G5_HIP(hipMemcpy(h_meta, d_meta, 8 * 8, hipMemcpyDeviceToHost));
    G5_REQUIRE(h_meta[kMetaFlag] == 0, G5_E_STREAM, "%s: the decoded stream lost values", who);
    *nwords = (size_t)h_meta[kMetaNw];
    return G5_OK;
}

}  // namespace g5

namespace {
g5::S4Codec &abi_codec()
{
    static thread_local g5::S4Codec c;
    return c;
}
}  // namespace

extern "C" size_t g5_s4_max_words(size_t n)
{
    const size_t L = n / 128 * 128;
    return 4 + 4 * ((L + 2047) / 2048) + L + (5 * (n - L) + 3) / 4;
}

extern "C" int g5_s4_encode(const uint32_t *d_in, size_t n, uint32_t *d_out, size_t out_cap, size_t *nwords)
{
    G5_REQUIRE((d_in || n == 0) && d_out && nwords, G5_E_ARG, "g5_s4_encode: null argument");
    G5_REQUIRE(((uintptr_t)d_out & 15) == 0, G5_E_ARG, "g5_s4_encode: output must be 16-byte aligned");
    g5::ExchangeTimer t(nullptr);
    g5::S4Codec &c = abi_codec();
    c.acct = &t;
    const int rc = c.encode(d_in, (int64_t)n, d_out, out_cap, nwords, nullptr);
    c.acct = nullptr;
    return rc;
}

extern "C" int g5_s4_decode(const uint32_t *d_in, size_t nwords, uint32_t *d_out, size_t out_cap, size_t *n)
{
    G5_REQUIRE(d_in && (d_out || out_cap == 0) && n, G5_E_ARG, "g5_s4_decode: null argument");
    g5::ExchangeTimer t(nullptr);
    g5::S4Codec &c = abi_codec();
    c.acct = &t;
    const int rc = c.decode(d_in, nwords, d_out, out_cap, n, nullptr);
    c.acct = nullptr;
    return rc;
}

extern "C" int g5_s4_roundtrip(const uint32_t *d_in, size_t n, uint32_t *d_words, size_t cap_words, uint32_t *d_out,
                               size_t *nwords)
{
    G5_REQUIRE(d_in && d_words && d_out && nwords, G5_E_ARG, "g5_s4_roundtrip: null argument");
    g5::S4Codec &c = abi_codec();
    int rc = c.reserve((int64_t)n);
    if (rc) return rc;
    // the flag lives in the decoder's record, which the header search clears
    rc = c.roundtrip(d_in, (int64_t)n, d_words, cap_words, d_out, nullptr, nullptr);
    if (rc) return rc;
    return c.roundtrip_result(nwords, "g5_s4_roundtrip");
}
