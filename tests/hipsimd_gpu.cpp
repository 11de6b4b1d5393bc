// The HipSimd schema (host/g5_compression.hh) on the GPU, driven through
// the Compression<int64_t, uint32_t> interface exactly as GlobalBFS's fold
// and vreduce drive CpuSimd (cpu_2d/globalbfs.hh:864-905,
// cpu_2d/vreduce.hpp:263-265): compress at the sender, decompress at the
// receiver with the original length, verifyCompression on the result.
// Built and run by tests/test_gpu_codec_schema.py.
//
//   hipsimd_gpu <in> <out>
//   in : int32 threshold, int64 count, then per queue int64 n, int64 q[n]
//   out: per queue int64 csize, int32 isCompressed(n, csize), uint32 w[csize]
// Exit status 0 when every queue came back identical after the round trip.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "g5_compression.hh"

using namespace g5host;

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: hipsimd_gpu <in> <out>\n");
        return 2;
    }
    try {
        if (g5_init(0) != G5_OK) throw std::runtime_error(g5_last_error());
        FILE *fi = std::fopen(argv[1], "rb");
        FILE *fo = std::fopen(argv[2], "wb");
        if (!fi || !fo) throw std::runtime_error("cannot open the queue files");
        int32_t thr = 0;
        int64_t count = 0;
        if (std::fread(&thr, 4, 1, fi) != 1 || std::fread(&count, 8, 1, fi) != 1)
            throw std::runtime_error("short header");
        HipSimd<int64_t, uint32_t> hs;
        hs.reconfigure(thr, "s4-bp128-d4");
        Compression<int64_t, uint32_t> &schema = hs;
        int bad = 0;
        for (int64_t k = 0; k < count; ++k) {
            int64_t n = 0;
            if (std::fread(&n, 8, 1, fi) != 1) throw std::runtime_error("short queue header");
            std::vector<int64_t> q((size_t)n);
            if (n && std::fread(q.data(), 8, (size_t)n, fi) != (size_t)n) throw std::runtime_error("short queue");
            uint32_t *c = nullptr;
            size_t cs = 0;
            schema.compress(q.data(), (size_t)n, &c, cs);
            const int32_t coded = schema.isCompressed((size_t)n, cs) ? 1 : 0;
            const int64_t cs64 = (int64_t)cs;
            std::fwrite(&cs64, 8, 1, fo);
            std::fwrite(&coded, 4, 1, fo);
            if (cs) std::fwrite(c, 4, cs, fo);
            // the receiver knows the original length (globalbfs.hh:893-904)
            int64_t *back = nullptr;
            size_t nb = (size_t)n;
            schema.decompress(c, (int)cs, &back, nb);
            bool same = nb == (size_t)n;
            for (int64_t i = 0; same && i < n; ++i) same = back[i] == q[(size_t)i];
            if (same) schema.verifyCompression(q.data(), back, (size_t)n);
            if (!same) {
                std::fprintf(stderr, "queue %lld (n = %lld): the round trip differs\n", (long long)k, (long long)n);
                ++bad;
            }
            free(c);
            free(back);
        }
        std::fclose(fi);
        std::fclose(fo);
        if (bad) return 1;
        std::printf("hipsimd_gpu: %lld queues round-tripped\n", (long long)count);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "hipsimd_gpu: %s\n", e.what());
        return 3;
    }
}
