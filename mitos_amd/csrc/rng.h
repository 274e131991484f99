// xoshiro256** (the generator of include/random_xorshift.h:15-29, seeded by
// splitmix64 as at :34-59) with GF(2) jump-ahead so that every z-row of the
// half-complex grid can start its own slice of the single serial stream.
//
// The state update s -> T s is linear over GF(2)^256.  T is built once as a
// 256x256 bit matrix (row i = which state bits feed output bit i); powers
// T^n come from repeated squaring on the host.  A device thread jumps by
// applying a few precomputed powers J^(2^b) to a base state.
#pragma once

#include <stdint.h>
#include <string.h>

#include <vector>

#include "mitos_common.h"

namespace mitos {

struct Xo {
    uint64_t s[4];
};

__host__ __device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
}

// one step of xoshiro256**; returns the output word
__host__ __device__ __forceinline__ uint64_t xo_next(Xo& st) {
    const uint64_t out = rotl64(st.s[1] * 5, 7) * 9;
    const uint64_t t = st.s[1] << 17;
    st.s[2] ^= st.s[0];
    st.s[3] ^= st.s[1];
    st.s[1] ^= st.s[2];
    st.s[0] ^= st.s[3];
    st.s[2] ^= t;
    st.s[3] = rotl64(st.s[3], 45);
    return out;
}

inline uint64_t splitmix64_next(uint64_t& x) {
    uint64_t r = x;
    x = r + 0x9E3779B97f4A7C15ULL;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
    return r ^ (r >> 31);
}

inline Xo xo_seed(uint64_t seed) {
    Xo st;
    uint64_t sm = seed;
    for (int i = 0; i < 4; i++) st.s[i] = splitmix64_next(sm);
    return st;
}

// 256x256 GF(2) matrix, row-major: rows[i][w] bit b = M[i][64*w + b]
struct BitMat {
    uint64_t rows[256][4];
};

inline void bm_identity(BitMat& m) {
    memset(&m, 0, sizeof(m));
    for (int i = 0; i < 256; i++) m.rows[i][i >> 6] = 1ULL << (i & 63);
}

// transition matrix of one xoshiro step
inline void bm_step(BitMat& m) {
    memset(&m, 0, sizeof(m));
    for (int j = 0; j < 256; j++) {
        Xo e = {{0, 0, 0, 0}};
        e.s[j >> 6] = 1ULL << (j & 63);
        xo_next(e);
        for (int i = 0; i < 256; i++)
            if ((e.s[i >> 6] >> (i & 63)) & 1ULL) m.rows[i][j >> 6] |= 1ULL << (j & 63);
    }
}

// c = a * b
inline void bm_mul(const BitMat& a, const BitMat& b, BitMat& c) {
    BitMat r;
    memset(&r, 0, sizeof(r));
    for (int i = 0; i < 256; i++)
        for (int j = 0; j < 256; j++)
            if ((a.rows[i][j >> 6] >> (j & 63)) & 1ULL)
                for (int w = 0; w < 4; w++) r.rows[i][w] ^= b.rows[j][w];
    c = r;
}

__host__ __device__ __forceinline__ Xo bm_apply_rows(const uint64_t* rows, const Xo& s) {
    Xo r = {{0, 0, 0, 0}};
    for (int w = 0; w < 4; w++) {
        uint64_t acc = 0;
        for (int i = 0; i < 64; i++) {
            const uint64_t* row = rows + (size_t)(w * 64 + i) * 4;
            const uint64_t p = (row[0] & s.s[0]) ^ (row[1] & s.s[1]) ^ (row[2] & s.s[2]) ^
                               (row[3] & s.s[3]);
            acc |= (uint64_t)(__builtin_popcountll(p) & 1) << i;
        }
        r.s[w] = acc;
    }
    return r;
}

// T^n
inline void bm_pow_step(uint64_t n, BitMat& out) {
    BitMat base, r;
    bm_step(base);
    bm_identity(r);
    while (n) {
        if (n & 1) bm_mul(r, base, r);
        n >>= 1;
        if (n) bm_mul(base, base, base);
    }
    out = r;
}

}  // namespace mitos
