// g5_compression.hh -- the frontier-queue codec interface of the reference
// (Compression<T, T_C>, cpu_2d/compression/compression.hh:11-28) with its
// "nocompression" schema and an s4-bp128-d4 schema whose encode and decode
// run on the GPU through the C ABI (g5_s4_encode / g5_s4_decode).
//
// Contract kept from the reference:
//   compress(in, n, &out, csize)    callee allocates *out (posix_memalign,
//                                   16-byte aligned), caller free()s it;
//                                   int64 ids are narrowed to uint32 (ids
//                                   above 2^32 - 1 are rejected,
//                                   cpusimd.hh:82-91); n <= threshold is sent
//                                   raw (cpusimd.hh:27,102-117)
//   decompress(in, csize, &out, n)  callee allocates, caller free()s; the
//                                   stream is decoded iff
//                                   isCompressed(n, csize) (cpusimd.hh:170-173)
//   isCompressed(n, csize)          n > threshold && csize != n
// Difference, on purpose: a coded stream exactly n words long would be read
// back as raw by the rule above (the reference decodes it wrongly); compress
// sends such a queue raw instead, so the receiver's rule always holds.
// Errors are thrown as std::runtime_error (the reference's codec throws or
// asserts).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "g5hip.h"

namespace g5host {

template <typename T, typename T_C>
class Compression
{
public:
    virtual void debugCompression(T *fq, const size_t size) const = 0;
    virtual void compress(T *__restrict__ fq_64, const size_t &size, T_C **__restrict__ compressed_fq_64,
                          size_t &compressedsize) const = 0;
    virtual void decompress(T_C *__restrict__ compressed_fq_64, const int size,
                            /*Out*/ T **__restrict__ uncompressed_fq_64, /*In Out*/ size_t &uncompressedsize) const = 0;
    virtual void verifyCompression(const T *fq, const T *uncompressed_fq_64, const size_t uncompressedsize) const = 0;
    virtual bool isCompressed(const size_t originalsize, const size_t compressedsize) const = 0;
    virtual void reconfigure(int compressionThreshold, std::string compressionExtraArgument) = 0;
    virtual ~Compression() {}
    virtual std::string name() const = 0;
    virtual void init() const = 0;
};

namespace detail {

constexpr size_t kAlign = 16;

template <typename X>
X *aligned_alloc_n(size_t n)
{
    void *p = nullptr;
    if (posix_memalign(&p, kAlign, (n ? n : 1) * sizeof(X)) != 0) throw std::bad_alloc();
    return static_cast<X *>(p);
}

[[noreturn]] inline void fail(const char *what, int rc)
{
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + g5_last_error());
}

inline void check(int rc, const char *what)
{
    if (rc != G5_OK) fail(what, rc);
}

// grow-only device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void reserve(size_t b)
    {
        if (b <= bytes) return;
        if (p) g5_dfree(p);
        p = nullptr;
        bytes = 0;
        check(g5_dmalloc(&p, b), "g5_dmalloc");
        bytes = b;
    }
    ~DevBuf()
    {
        if (p) g5_dfree(p);
    }
};

}  // namespace detail

// raw uint32 transfer (the reference's "nocompression" schema)
template <typename T, typename T_C>
class NoCompression : public Compression<T, T_C>
{
public:
    void debugCompression(T *, const size_t) const override {}
    void compress(T *__restrict__ fq_64, const size_t &size, T_C **__restrict__ out, size_t &csize) const override
    {
        *out = detail::aligned_alloc_n<T_C>(size);
        for (size_t i = 0; i < size; ++i) (*out)[i] = static_cast<T_C>(fq_64[i]);
        csize = size;
    }
    void decompress(T_C *__restrict__ in, const int size, T **__restrict__ out, size_t &n) const override
    {
        n = static_cast<size_t>(size);
        *out = detail::aligned_alloc_n<T>(n);
        for (size_t i = 0; i < n; ++i) (*out)[i] = static_cast<T>(in[i]);
    }
    void verifyCompression(const T *, const T *, const size_t) const override {}
    bool isCompressed(const size_t, const size_t) const override { return false; }
    void reconfigure(int, std::string) override {}
    std::string name() const override { return "nocompression"; }
    void init() const override {}
};

// s4-bp128-d4 on the GPU (the reference's "cpusimd" schema configured with
// -be s4-bp128-d4, cpu_2d/test3.sh:12): byte-identical word streams.
template <typename T, typename T_C>
class HipSimd : public Compression<T, T_C>
{
    static_assert(sizeof(T_C) == 4, "the codec moves 32-bit words");
    size_t threshold_ = 64;  // the reference's default (cpusimd.hh:44)
    mutable detail::DevBuf d_in_, d_out_;

    bool compressible(size_t n) const { return n > threshold_; }

public:
    void debugCompression(T *, const size_t) const override {}

    void compress(T *__restrict__ fq_64, const size_t &size, T_C **__restrict__ out, size_t &csize) const override
    {
        if (!compressible(size)) {
            *out = detail::aligned_alloc_n<T_C>(size);
            for (size_t i = 0; i < size; ++i) (*out)[i] = static_cast<T_C>(fq_64[i]);
            csize = size;
            return;
        }
        std::unique_ptr<uint32_t, decltype(&free)> q32(detail::aligned_alloc_n<uint32_t>(size), &free);
        for (size_t i = 0; i < size; ++i) {
            const T v = fq_64[i];
            if (static_cast<uint64_t>(v) > 0xFFFFFFFFull)
                throw std::runtime_error("HipSimd::compress: id above 2^32 - 1 cannot be narrowed");
            q32.get()[i] = static_cast<uint32_t>(v);
        }
        const size_t cap = g5_s4_max_words(size);
        d_in_.reserve(size * 4);
        d_out_.reserve(cap * 4);
        detail::check(g5_memcpy_h2d(d_in_.p, q32.get(), size * 4), "g5_memcpy_h2d");
        size_t words = 0;
        detail::check(g5_s4_encode((const uint32_t *)d_in_.p, size, (uint32_t *)d_out_.p, cap, &words), "g5_s4_encode");
        if (words == size) {
            // would be read back as raw by isCompressed: send raw
            *out = detail::aligned_alloc_n<T_C>(size);
            std::memcpy(*out, q32.get(), size * 4);
            csize = size;
            return;
        }
        *out = detail::aligned_alloc_n<T_C>(words);
        detail::check(g5_memcpy_d2h(*out, d_out_.p, words * 4), "g5_memcpy_d2h");
        csize = words;
    }

    void decompress(T_C *__restrict__ in, const int size, T **__restrict__ out, size_t &n) const override
    {
        const size_t csize = static_cast<size_t>(size);
        if (!isCompressed(n, csize)) {
            n = csize;
            *out = detail::aligned_alloc_n<T>(n);
            for (size_t i = 0; i < n; ++i) (*out)[i] = static_cast<T>(in[i]);
            return;
        }
        d_in_.reserve(csize * 4);
        d_out_.reserve(n * 4);
        detail::check(g5_memcpy_h2d(d_in_.p, in, csize * 4), "g5_memcpy_h2d");
        size_t got = 0;
        detail::check(g5_s4_decode((const uint32_t *)d_in_.p, csize, (uint32_t *)d_out_.p, n, &got), "g5_s4_decode");
        if (got != n) throw std::runtime_error("HipSimd::decompress: stream holds a different number of values");
        std::unique_ptr<uint32_t, decltype(&free)> q32(detail::aligned_alloc_n<uint32_t>(n), &free);
        detail::check(g5_memcpy_d2h(q32.get(), d_out_.p, n * 4), "g5_memcpy_d2h");
        *out = detail::aligned_alloc_n<T>(n);
        for (size_t i = 0; i < n; ++i) (*out)[i] = static_cast<T>(q32.get()[i]);
    }

    void verifyCompression(const T *fq, const T *back, const size_t n) const override
    {
        if (compressible(n) && std::memcmp(fq, back, n * sizeof(T)) != 0)
            throw std::runtime_error("HipSimd::verifyCompression: round trip differs");
    }

    bool isCompressed(const size_t originalsize, const size_t compressedsize) const override
    {
        return compressible(originalsize) && originalsize != compressedsize;
    }

    void reconfigure(int compressionThreshold, std::string codec) override
    {
        if (compressionThreshold <= 0) throw std::invalid_argument("HipSimd::reconfigure: threshold must be > 0");
        if (!codec.empty() && codec != "s4-bp128-d4")
            throw std::invalid_argument("HipSimd::reconfigure: only s4-bp128-d4 runs on the GPU, not " + codec);
        threshold_ = static_cast<size_t>(compressionThreshold);
    }

    std::string name() const override { return "hipsimd"; }
    void init() const override {}
    size_t threshold() const { return threshold_; }
};

// CompressionFactory<T, T_C>::getFromName (cpu_2d/compression/compressionfactory.hh):
// one shared instance per schema name.
template <typename T, typename T_C>
struct CompressionFactory {
    static Compression<T, T_C> *getFromName(const std::string &name)
    {
        static NoCompression<T, T_C> none;
        static HipSimd<T, T_C> hip;
        if (name == "nocompression") return &none;
        if (name == "hipsimd" || name == "cpusimd") return &hip;
        throw std::invalid_argument("CompressionFactory: unknown schema " + name);
    }
};

}  // namespace g5host
