// Per-term algebra of the propagate path, shared by host and device code.
//
// Row layout (u64 words per row, NW of them):
//   Pauli    (NW = 2L): words [0, L) = x-plane, [L, 2L) = z-plane, qubit q -> bit q & 63 of limb q >> 6.
//   Majorana (NW = W) : Majorana slot s -> bit s & 63 of word s >> 6.
// Every rule below restates the reference for this layout; the reference stores Majorana slot s at
// physical bit 2N-1-s of an MSb0 bitset (cpp/monoprop/algebra/AlgebraCommon.h:204-213).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MPA_HD __host__ __device__ __forceinline__
#else
#define MPA_HD inline
#endif

namespace mpa {

constexpr int kMaxWords = 8;
constexpr uint64_t kEvenSlots = 0x5555555555555555ULL;

MPA_HD int popc64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __popcll(x);
#else
    return __builtin_popcountll(x);
#endif
}

// Reference Monomial::count(): Pauli X,Y use one slot and Z two.
template <int NW, int PAULI>
MPA_HD int slot_pop(const uint64_t *m) {
    int p = 0;
    if (PAULI) {
        constexpr int L = NW / 2;
#pragma unroll
        for (int l = 0; l < L; ++l) p += popc64(m[l] ^ m[L + l]) + popc64(m[L + l]);
    } else {
#pragma unroll
        for (int w = 0; w < NW; ++w) p += popc64(m[w]);
    }
    return p;
}

// Structural keep of the partner (AlgebraCommon.h:285-342, fully paired exemption).
// Pauli: xor_sum = |x|, support = |x or z| (PauliAlgebra.h:17-21).
template <int NW, int PAULI>
MPA_HD bool keep_partner(const uint64_t *t, uint32_t cutoff, int cutoff_type) {
    int xs = 0, measure = 0;
    if (PAULI) {
        constexpr int L = NW / 2;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            xs += popc64(t[l]);
            measure += popc64(t[l] | t[L + l]);
        }
    } else {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint64_t a = t[w] & kEvenSlots, b = (t[w] >> 1) & kEvenSlots;
            xs += popc64(a ^ b);
            measure += cutoff_type == 0 ? popc64(t[w]) : popc64(a | b);
        }
    }
    return xs == 0 || static_cast<uint32_t>(measure) <= cutoff;
}

// Sine phase of the emitted partner t = m ^ g.
// Pauli (PauliAlgebra.h:116-140): e = y_G + y_M - y_T + 2|z_M & x_G| with y = |x & z|;
//   phase -1 iff e = 1 (mod 4).
// Majorana (Algebra.h:51-59, MajoranaAlgebra.h:82-122): (-1)^{#(m in M, g in G): g < m} times
//   REAL_PARTS[(C(|M|,2) + C(|G|,2) - C(|T|,2) + 3) mod 4].
template <int NW, int PAULI>
MPA_HD int emit_phase(const uint64_t *m, const uint64_t *g, const uint64_t *t) {
    if (PAULI) {
        constexpr int L = NW / 2;
        long e = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            e += popc64(g[l] & g[L + l]) + popc64(m[l] & m[L + l]) - popc64(t[l] & t[L + l]);
            e += 2 * popc64(m[L + l] & g[l]);
        }
        const long r = ((e % 4) + 4) % 4;
        return r == 1 ? -1 : 1;
    }
    uint64_t par = 0, carry = 0;
    uint64_t pm = 0, pg = 0, pt = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        uint64_t x = g[w];
        x ^= x << 1;
        x ^= x << 2;
        x ^= x << 4;
        x ^= x << 8;
        x ^= x << 16;
        x ^= x << 32;
        const uint64_t below = (x << 1) ^ (uint64_t{0} - carry);
        par ^= static_cast<uint64_t>(popc64(below & m[w]));
        carry ^= static_cast<uint64_t>(popc64(g[w]) & 1);
        pm += popc64(m[w]);
        pg += popc64(g[w]);
        pt += popc64(t[w]);
    }
    const int sign = (par & 1) ? -1 : 1;
    const uint64_t power = (pm * (pm - 1) / 2 + pg * (pg - 1) / 2 - pt * (pt - 1) / 2 + 3) % 4;
    const int real_part = power == 0 ? 1 : (power == 2 ? -1 : 0);
    return sign * real_part;
}

// Fully paired test and diagonal phase against the initial product state (MPOperator.h:235-254,
// PauliAlgebra.h:142-147, MajoranaAlgebra.h:52-60).  `occ`: Pauli occupied qubits in x-plane layout
// (L limbs); Majorana slot 2j of occupied mode j.
template <int NW, int PAULI>
MPA_HD bool is_paired(const uint64_t *m) {
    if (PAULI) {
        constexpr int L = NW / 2;
        uint64_t x = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) x |= m[l];
        return x == 0;
    }
    uint64_t d = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) d |= (m[w] & kEvenSlots) ^ ((m[w] >> 1) & kEvenSlots);
    return d == 0;
}

template <int NW, int PAULI>
MPA_HD double state_phase(const uint64_t *m, const uint64_t *occ) {
    int c = 0;
    if (PAULI) {
        constexpr int L = NW / 2;
#pragma unroll
        for (int l = 0; l < L; ++l) c += popc64(m[L + l] & occ[l]);
        return (c & 1) ? -1.0 : 1.0;
    }
    int pop = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        c += popc64(m[w] & occ[w]);
        pop += popc64(m[w]);
    }
    return ((c + pop / 2) & 1) ? -1.0 : 1.0;
}

// splitmix64 finaliser, mixed per word with the word position.
MPA_HD uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

template <int NW>
MPA_HD uint64_t hash_key(const uint64_t *k) {
    uint64_t h = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) h ^= mix64(k[w] + 0x9e3779b97f4a7c15ULL * static_cast<uint64_t>(w + 1));
    return h;
}

}  // namespace mpa
