// g5_codec.hpp -- device s4-bp128-d4 codec state (see g5_codec.hip).
#pragma once

#include "g5_internal.hpp"

namespace g5 {

// Scratch for one stream at a time on one stream; sized for up to cap values.
// All device arrays live in one buffer: reserve() allocates it (owned), or
// bind() lays them out in a buffer the caller provides (borrowed: never
// freed here).
struct S4Codec {
    int64_t cap = 0;               // values
    uint8_t *d_base = nullptr;     // the buffer holding every device array below
    bool owned = false;            // d_base and h_meta were allocated by reserve()
    int64_t *d_words = nullptr;    // per-miniblock word counts -> prefix, nblk + 1
    int64_t *d_scan = nullptr;     // int64 scan scratch
    int64_t *d_off = nullptr;      // decode: per-miniblock data offset
    uint32_t *d_carry = nullptr;   // decode: per-miniblock carry records (5 words)
    uint32_t *d_carry_scan = nullptr;
    int64_t *d_grp = nullptr;      // decode: group header positions
    int32_t *d_nxt = nullptr;      // decode: next group start per candidate start (4 + 4i)
    int32_t *d_jmp = nullptr, *d_jtmp = nullptr;  // decode: 32-step jumps (pointer doubling)
    int32_t *d_anchor = nullptr;   // decode: every 32nd group start
    int64_t ncand_cap = 0;
    int64_t *d_meta = nullptr;     // small device record
    int64_t *d_nw = nullptr;       // [0] the last encode's stream length, [1] decode's input length
    int64_t *h_meta = nullptr;     // pinned mirror
    // accounting of encode() and decode() (g5_exchange_profile), when set
    ExchangeTimer *acct = nullptr;

    int reserve(int64_t n_values);
    void release();
    // Lays the device arrays for up to n_values out from base (nothing is
    // set when base is null) and returns the bytes they take; h_meta must be
    // set by the caller of a borrowed layout.
    size_t bind(uint8_t *base, int64_t n_values);
    static size_t scratch_bytes(int64_t n_values) { return S4Codec().bind(nullptr, n_values); }
    // d_in[n] -> d_out; requires cap_words >= g5_s4_max_words(n).
    int encode(const uint32_t *d_in, int64_t n, uint32_t *d_out, size_t cap_words, size_t *nwords,
               hipStream_t s);
    // d_in[nwords] -> d_out (capacity cap_values); *n = decoded count.
    int decode(const uint32_t *d_in, size_t nwords, uint32_t *d_out, size_t cap_values, size_t *n,
               hipStream_t s);
    // encode d_in[n] into d_words_out, then decode it into d_out (may alias
    // d_in), with no host synchronisation: *d_flag (device) is set to 1 if
    // the round trip did not give back exactly n values, and left alone
    // otherwise (d_flag null: the flag slot of the decoder's record, which
    // the round trip clears first).
    int roundtrip(const uint32_t *d_in, int64_t n, uint32_t *d_words_out, size_t cap_words, uint32_t *d_out,
                  int64_t *d_flag, hipStream_t s);
    // After a roundtrip with d_flag null, on the null stream: waits for it,
    // fails with G5_E_STREAM unless exactly its n values came back, and sets
    // *nwords to the stream's length (who names the caller in the error).
    int roundtrip_result(size_t *nwords, const char *who);

private:
    int encode_launch(const uint32_t *d_in, int64_t n, uint32_t *d_out, hipStream_t s);
    int search_launch(const uint32_t *d_in, const int64_t *nwp, int64_t words_bound, int64_t cap_values,
                      bool serial, hipStream_t s);
    int unpack_launch(const uint32_t *d_in, const int64_t *nwp, int64_t L, int64_t cap_values, uint32_t *d_out,
                      int64_t *d_flag, hipStream_t s);
};

}  // namespace g5
