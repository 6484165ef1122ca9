// grand_kernels.hip -- gfx950 kernels of the GRAND decoder.
//
// The search is GF(2) integer work on 64-wide wavefronts (no MFMA):
//
//  * triage_kernel     one wave per shot: packs the int32 input row with ballots,
//                      forms s = H e when errors are given, writes the result of
//                      zero-syndrome shots and appends the others to a worklist.
//  * fused_kernel      persistent: every wave claims worklist shots and runs the
//                      whole per-shot pipeline of the reference driver (column
//                      scores, up to three active sets, weight-ordered search,
//                      post-processing) with wave-local state in LDS.
//  * plain_*           the one-launch-per-weight-class engine the host drives shot
//                      by shot, built from the same device functions.
//
// Candidate test.  Every column c carries a 32-bit signature sig[c] = M h_c where
// M is a fixed random GF(2) matrix (32 x n); the map is linear, so a pattern
// {c_1..c_w} can only satisfy H e = s if sig(s) ^ sig[c_1] ^ ... ^ sig[c_w] == 0.
// A weight-w class is walked as its (w-1)-prefixes: lanes keep prefix XORs of
// signatures, and the last columns that complete a prefix are looked up in an
// LDS hash table of the active signatures (C(tam-1, w-1) probes instead of
// C(tam, w) pattern tests).  A signature hit is confirmed exactly against the
// bit-packed columns, so results are exact; the signature only decides which
// patterns get the full D-word comparison.  The plain engine keeps the
// pattern-per-lane walk of the reference's GPU kernels.
//
// Enumeration order and counts follow the reference CPU search
// (grand_decoder.cu:264-350, :534-666): weight classes in list order, ranks in
// lexicographic order, the first match wins and the count includes it.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include "grand_types.h"

namespace grand {

// Pointers into the workgroup's LDS carry the LDS address space, so every access
// through them is a ds_read / ds_write rather than a flat access.
#define LDS __attribute__((address_space(3)))

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_xor(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v ^= (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t t = (uint32_t)__shfl_xor((int)v, o, 64);
        v = t < v ? t : v;
    }
    return v;
}

__device__ __forceinline__ int bcast_i(int v, int src) { return __shfl(v, src, 64); }
// Value of the first active lane as a wave-uniform (scalar) value; used so that
// loop exits and branches after lane-0 atomics depend only on uniform values.
__device__ __forceinline__ unsigned uni(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ unsigned long long uni64(unsigned long long v) {
    const unsigned lo = uni((unsigned)v), hi = uni((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t bcast_u(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }
__device__ __forceinline__ float bcast_f(float v, int src) { return __shfl(v, src, 64); }

__device__ __forceinline__ unsigned long long now_ticks() { return wall_clock64(); }

// fp32 -> u32 key whose unsigned order is the float order; -0 is folded onto +0
// so that equal costs compare equal, as the reference's `cost < best` does.
__device__ __forceinline__ uint32_t order_key(float f) {
    uint32_t b = __float_as_uint(f == 0.0f ? 0.0f : f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ uint64_t ncr(const unsigned long long *t, int n, int k) {
    return (n < 0 || k < 0 || k > n) ? 0ull : t[n * kNcrStride + k];
}

// --------------------------------------------------------------- combinatorics
// Lexicographic unranking (rank 0 = {0..w-1}); the reference scans linearly
// (grand_decoder.cu:83-100), here each position is a binary search on the
// hockey-stick identity  sum_{y=p+1}^{x-1} C(t-1-y, k-1) = C(t-1-p, k) - C(t-x, k).
template <int W, class P>
__device__ __forceinline__ void unrank_t(uint64_t r, int tam, P nCr, int *c) {
    int prev = -1;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const int k = W - i;
        const uint64_t tot = nCr[(tam - 1 - prev) * kNcrStride + k];
        const uint64_t T = tot - r;
        int lo = prev + 1, hi = tam - k;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (nCr[(tam - mid) * kNcrStride + k] >= T) lo = mid; else hi = mid - 1;
        }
        r -= tot - nCr[(tam - lo) * kNcrStride + k];
        c[i] = lo;
        prev = lo;
    }
}

// Same with the weight known only at run time (per lane), at most MAXW.
template <int MAXW = kMaxW, class P = const unsigned long long *>
__device__ __forceinline__ void unrank_dyn(uint64_t r, int tam, int w, P nCr, int *c) {
    int prev = -1;
#pragma unroll
    for (int i = 0; i < MAXW; ++i) {
        if (i < w) {
            const int k = w - i;
            const uint64_t tot = nCr[(tam - 1 - prev) * kNcrStride + k];
            const uint64_t T = tot - r;
            int lo = prev + 1, hi = tam - k;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (nCr[(tam - mid) * kNcrStride + k] >= T) lo = mid; else hi = mid - 1;
            }
            r -= tot - nCr[(tam - lo) * kNcrStride + k];
            c[i] = lo;
            prev = lo;
        }
    }
}

// Unranking in a class small enough that every binomial it reads is below
// 2^32 (tam <= 128), from the wave's LDS copy of the table.
__device__ __forceinline__ void unrank_lds(uint32_t r, int tam, int w, const unsigned long long LDS *t,
                                           int *c) {
    int prev = -1;
#pragma unroll
    for (int i = 0; i < kMaxW; ++i) {
        if (i < w) {
            const int k = w - i;
            const uint32_t tot = (uint32_t)t[(tam - 1 - prev) * kNcrStride + k];
            const uint32_t T = tot - r;
            int lo = prev + 1, hi = tam - k;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((uint32_t)t[(tam - mid) * kNcrStride + k] >= T) lo = mid; else hi = mid - 1;
            }
            r -= tot - (uint32_t)t[(tam - lo) * kNcrStride + k];
            c[i] = lo;
            prev = lo;
        }
    }
}

// ------------------------------------------------------------- wave shot state
// Per-wave LDS layout (cap entries each): A (active columns) and their
// signatures; for ML also the llr of A and the LLR-sorted copies P / sigP /
// llrP; then the signature hash table (htcap = power of two >= 2 cap entries of
// (index, signature)).
struct WaveLds {
    int LDS *A;
    uint32_t LDS *sigA;
    float LDS *llrA;
    int LDS *P;
    uint32_t LDS *sigP;
    float LDS *llrP;
    unsigned long long LDS *ht;   // (signature << 32) | position, empty = kHtEmpty
    uint32_t LDS *stage;       // search kernel: the claimed shots, DevCall::stage_words each
    uint32_t LDS *pbm;         // pair table (search kernel, pair_cap > 0): key filter [kPairBitmapWords],
    uint32_t LDS *phead;       //   bucket chain heads [pair_buckets(kind)], keys and (y | x << 10 | next << 20)
    uint32_t LDS *pkey;
    uint32_t LDS *pval;
    uint32_t LDS *sw;  // syndrome words of the shot being searched, kMaxWords
    uint32_t LDS *sw2; // syndrome minus one column (learned pass), kMaxWords
    unsigned long long LDS *ncr; // C(n, k) for n < kNcrLdsRows, k < 8 (search kernel only)
};

extern __shared__ uint32_t lds_raw[];
__device__ __forceinline__ uint32_t LDS *lds_base() { return (uint32_t LDS *)lds_raw; }

constexpr int kNcrLdsRows = 129;   // n = 0..128

// Signature bitmap kept after the hash table: 256 bits per entry (at most 2^16
// bits), so a prefix whose target signature is absent is rejected by one LDS
// read in all but a fraction <= 1/256 of cases (1/64 at the largest lists),
// without walking a probe chain; a wave of 64 prefixes then rarely has any
// lane on the probe path.
__host__ __device__ __forceinline__ int bitmap_words(int cap) {
    int bits = 1024;
    while (bits < 256 * cap && bits < 65536) bits <<= 1;
    return bits / 32;
}

__host__ __device__ __forceinline__ int ht_capacity(int cap) {
    int h = 64;
    while (h < 2 * cap) h <<= 1;
    return h;
}

// Every wave's LDS, in order: syndrome words (kMaxWords), the learned pass's
// reduced syndrome (kMaxWords), C(n, k) for n < kNcrLdsRows (search kernel),
// the staged code (DevCode::lds_words, search kernel), then the per-class
// arrays sized by the active-list capacity and the claimed shots.
constexpr int kLdsNcrOffset = 2 * kMaxWords;
constexpr int kLdsCodeOffset = kLdsNcrOffset + 2 * kNcrLdsRows * kNcrStride;

__device__ __forceinline__ WaveLds carve_lds(int cap, bool ml, int code_words = 0, int batch_words = 0,
                                             int pair_cap = 0, int head_words = pair_head_words(false)) {
    uint32_t LDS *base = lds_base();
    WaveLds L;
    L.sw = base;
    L.sw2 = L.sw + kMaxWords;
    L.ncr = (unsigned long long LDS *)(base + kLdsNcrOffset);
    // the staged code is padded to 4 words, so the 64-bit hash table and the
    // pair table that follow are 16-byte aligned
    uint32_t LDS *p = base + kLdsCodeOffset + ((code_words + 3) & ~3);
    L.ht = (unsigned long long LDS *)p; p += 2 * ht_capacity(cap) + bitmap_words(cap);
    L.A = (int LDS *)p; p += cap;
    L.sigA = p; p += cap;
    if (ml) {
        L.llrA = (float LDS *)p; p += cap;
        L.P = (int LDS *)p; p += cap;
        L.sigP = p; p += cap;
        L.llrP = (float LDS *)p; p += cap;
    } else {
        L.llrA = nullptr; L.P = nullptr; L.sigP = nullptr; L.llrP = nullptr;
    }
    if (pair_cap > 0) {
        L.pbm = p; p += kPairBitmapWords;
        L.phead = p; p += head_words;
        L.pkey = p; p += pair_cap;
        L.pval = p; p += pair_cap;
    } else {
        L.pbm = nullptr; L.phead = nullptr; L.pkey = nullptr; L.pval = nullptr;
    }
    L.stage = batch_words > 0 ? p : nullptr;
    return L;
}

// The search kernel keeps C(n, k), n < kNcrLdsRows, in every wave's LDS at a
// fixed offset, filled when the kernel starts; lookups with small n read it
// instead of the global table.
__device__ __forceinline__ const unsigned long long LDS *ncr_lds() {
    return (const unsigned long long LDS *)(lds_base() + kLdsNcrOffset);
}

// The staged code (search kernel, code.lds_words > 0): column words [D][m],
// offsets [m], signatures [m], pivot rows [m] and RREF rows [rank][M32].
__device__ __forceinline__ const uint32_t LDS *code_lds() { return lds_base() + kLdsCodeOffset; }
__device__ __forceinline__ const uint32_t LDS *code_lds_off(const DevCode &code) {
    return code_lds() + (size_t)code.D * code.m;
}
__device__ __forceinline__ const uint32_t LDS *code_lds_sig(const DevCode &code) {
    return code_lds() + (size_t)(code.D + 1) * code.m;
}
__device__ __forceinline__ const int LDS *code_lds_pivrow(const DevCode &code) {
    return (const int LDS *)(code_lds() + (size_t)(code.D + 2) * code.m);
}
__device__ __forceinline__ const uint32_t LDS *code_lds_R(const DevCode &code) {
    return code_lds() + (size_t)(code.D + 3) * code.m;
}
__device__ __forceinline__ const uint32_t LDS *code_lds_dg(const DevCode &code) {
    return code_lds() + (size_t)(code.D + 2) * code.m;
}
// word d of column c
__device__ __forceinline__ uint32_t col_word(const DevCode &code, int c, int d) {
    return code.lds_words ? code_lds()[(size_t)d * code.m + c] : code.Hcm[(size_t)c * code.D + d];
}

__device__ __forceinline__ uint64_t ncr_w(const DevCode &code, int n, int k) {
    if (n < 0 || k < 0 || k > n) return 0ull;
    return (n < kNcrLdsRows) ? ncr_lds()[n * kNcrStride + k] : code.nCr[n * kNcrStride + k];
}

// Exact check of a candidate: s ^ XOR_l column(cols[l]) == 0 over all D words.
template <int MAXW = kMaxW>
__device__ __forceinline__ bool exact_match(const DevCode &code, const uint32_t LDS *s, const int *cols, int w) {
    bool ok = true;
#pragma unroll 1
    for (int d = 0; d < code.D; ++d) {
        uint32_t x = s[d];
#pragma unroll
        for (int l = 0; l < MAXW; ++l)
            if (l < w) x ^= col_word(code, cols[l], d);
        ok = ok && (x == 0u);
    }
    return ok;
}

// Active set for threshold thr: ascending columns with score >= thr, where
// score[c] = |supp(h_c) & supp(s)| + off[c]  (grand_decoder.cu:352-416, :495-504).
// Writes the first min(tam, cap) entries; returns the full tam.
// Active columns at threshold thr (score = |supp(h_c) & supp(s)| + off[c] >= thr)
// in column order, at most cap stored in A / sigA.  The columns are scored 64 per
// round, kBuildRounds rounds with their loads (offsets, signatures, column words)
// issued together.  Returns the number of active columns, or, once that passes
// cap, some number above cap: every caller treats a count above cap alike (an
// active set too large, or for the DEM family the reference's fallback to the
// previous list size), and only the first cap entries are ever kept.
constexpr int kBuildRounds = 6;
template <class PW, class PO, class PS>
__device__ __forceinline__ int build_active_t(PW Hcw, PO off, PS sig, int D, int m, const uint32_t LDS *s, int thr,
                                              int LDS *A, uint32_t LDS *sigA, int cap) {
    const int lane = lane_id();
    int tam = 0;
    for (int c0 = 0; c0 < m && tam <= cap; c0 += 64 * kBuildRounds) {
        int cc[kBuildRounds], sc[kBuildRounds];
        uint32_t sv[kBuildRounds];
#pragma unroll
        for (int r = 0; r < kBuildRounds; ++r) {
            const int c = c0 + 64 * r + lane;
            cc[r] = c < m ? c : m - 1;               // in-bounds column for the loads
            sc[r] = off[cc[r]];
            sv[r] = sig[cc[r]];
        }
        // column words four at a time (indices clamped, so the loads issue together)
        for (int d0 = 0; d0 < D; d0 += 4) {
            uint32_t h[kBuildRounds][4];
#pragma unroll
            for (int r = 0; r < kBuildRounds; ++r)
#pragma unroll
                for (int j = 0; j < 4; ++j) h[r][j] = Hcw[(size_t)((d0 + j < D) ? d0 + j : D - 1) * m + cc[r]];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (d0 + j < D) {
                    const uint32_t sw = s[d0 + j];
#pragma unroll
                    for (int r = 0; r < kBuildRounds; ++r) sc[r] += __popc(h[r][j] & sw);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kBuildRounds; ++r) {
            const int c = c0 + 64 * r + lane;
            const bool on = c < m && sc[r] >= thr;
            const uint64_t b = ballot(on);
            if (on) {
                const int p = tam + lanes_below(b);
                if (p < cap) { A[p] = c; sigA[p] = sv[r]; }
            }
            tam += __popcll(b);
        }
    }
    __syncthreads();
    return tam;
}
// Active list at threshold thr into A / sigA (the wave's LDS).  Out of line: its
// rounds of loads in flight would otherwise add to the registers of every caller.
__device__ __attribute__((noinline)) int build_active_into(const DevCode &code, const uint32_t LDS *s, int thr,
                                                           int LDS *A, uint32_t LDS *sigA, int cap) {
    if (code.lds_words)
        return build_active_t(code_lds(), (const int LDS *)code_lds_off(code), code_lds_sig(code), code.D, code.m, s,
                              thr, A, sigA, cap);
    return build_active_t(code.Hcw, code.off, code.sig, code.D, code.m, s, thr, A, sigA, cap);
}
__device__ __forceinline__ int build_active(const DevCode &code, const uint32_t LDS *s, int thr, const WaveLds &L,
                                            int cap) {
    return build_active_into(code, s, thr, L.A, L.sigA, cap);
}

// Stable sort of A[0..tam) by llr (ties keep active-list order), the order of
// the reference's insertion sort (grand_decoder.cu:598-610).
__device__ void sort_by_key(const WaveLds &L, int tam);
__device__ void sort_by_llr(const DevCode &code, const WaveLds &L, int tam) {
    const int lane = lane_id();
    for (int i = lane; i < tam; i += 64) L.llrA[i] = code.llr[L.A[i]];
    __syncthreads();
    sort_by_key(L, tam);
}

// Stable ascending sort of A[0..tam) by the keys in llrA[0..tam) into P / sigP / llrP.
__device__ void sort_by_key(const WaveLds &L, int tam) {
    const int lane = lane_id();
    for (int i0 = 0; i0 < tam; i0 += 64) {
        const int i = i0 + lane;
        const float li = (i < tam) ? L.llrA[i] : 0.0f;
        int rank = 0;
        for (int j = 0; j < tam; ++j) {
            const float lj = L.llrA[j];
            rank += (lj < li) || (lj == li && j < i);
        }
        if (i < tam) { L.P[rank] = L.A[i]; L.sigP[rank] = L.sigA[i]; L.llrP[rank] = li; }
    }
    __syncthreads();
}

// Open-addressing table of (signature, position) for sg[0..tam): slot =
// signature & mask, linear probing, at most half full.  Equal signatures
// (equal columns) occupy successive slots of one probe chain.
constexpr unsigned long long kHtEmpty = 0xFFFFFFFFull;
// Returns the table mask in the low 16 bits and the bitmap's bit mask in the
// high 16 bits; the bitmap follows the table's hsz entries.
__device__ unsigned build_ht(unsigned long long LDS *ht, const uint32_t LDS *sg, int tam) {
    const int lane = lane_id();
    const int hsz = ht_capacity(tam);
    const unsigned mask = (unsigned)hsz - 1u;
    uint32_t LDS *bm = (uint32_t LDS *)(ht + hsz);
    const int bw = bitmap_words(tam);
    const unsigned bmask = (unsigned)bw * 32u - 1u;
    for (int i = lane; i < hsz; i += 64) ht[i] = kHtEmpty;
    for (int i = lane; i < bw; i += 64) bm[i] = 0u;
    __syncthreads();
    for (int i = lane; i < tam; i += 64) {
        const uint32_t key = sg[i];
        __hip_atomic_fetch_or(&bm[(key & bmask) >> 5], 1u << (key & 31u), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_WORKGROUP);
        const unsigned long long ent = ((unsigned long long)key << 32) | (unsigned)i;
        unsigned h = key & mask;
        for (;;) {
            unsigned long long expect = kHtEmpty;
            if (__hip_atomic_compare_exchange_strong(&ht[h], &expect, ent, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                break;
            h = (h + 1u) & mask;
        }
    }
    __syncthreads();
    return mask | (bmask << 16);
}

// ------------------------------------------------------------ prefix scanners
// A weight-w class is walked as its (w-1)-prefixes, the (w-1)-subsets of
// [0, tam-1) in lexicographic order; for each prefix the positions x >
// prefix_last with sig[x] == t0 ^ sig(prefix) come from one hash probe chain,
// and each is confirmed exactly.  Lexicographic order of (prefix, x) is the
// order of the w-subsets, so ranks, first-hit and ML tie-breaks are those of a
// pattern-by-pattern walk.

// Rank of the w-subset (c[0..WP-1], x) of [0, tam), w = WP + 1.
template <int WP, class P>
__device__ __forceinline__ uint64_t rank_with_last(const int *c, int x, int tam, P nCr) {
    uint64_t r = 0;
    int prev = -1;
#pragma unroll
    for (int i = 0; i < WP; ++i) {
        const int k = WP + 1 - i;
        r += nCr[(tam - 1 - prev) * kNcrStride + k] - nCr[(tam - c[i]) * kNcrStride + k];
        prev = c[i];
    }
    return r + (uint64_t)(x - prev - 1);
}

struct PrefHit {
    uint64_t rank;   // rank of the hit in its weight class (~0 = none)
    float cost;
    int j;           // prefix offset of the hit within the lane's range
};

// Probe of a prefix c[0..WP) whose target signature t passed the bitmap: the
// smallest position x > last in t's probe chain whose pattern matches exactly,
// or -1.  Out of line, so that none of its work is scheduled into the scan.
template <int WP>
__device__ __attribute__((noinline)) int probe_prefix(const DevCode &code, const uint32_t LDS *s, uint32_t t,
                                                     int last, int tam, const int LDS *act,
                                                     const unsigned long long LDS *ht, unsigned hm, int p0, int p1,
                                                     int p2, int p3, int p4, int p5) {
    const int pc[6] = {p0, p1, p2, p3, p4, p5};
    int best = -1;
    for (unsigned pos = t & hm;; pos = (pos + 1u) & hm) {
        const unsigned long long e = ht[pos];
        if (e == kHtEmpty) break;
        const int x = (int)(uint32_t)e;
        if ((uint32_t)(e >> 32) == t && x > last && x < tam && (best < 0 || x < best)) {
            int cols[WP + 1];
#pragma unroll
            for (int l = 0; l < WP; ++l) cols[l] = act[pc[l]];
            cols[WP] = act[x];
            if (exact_match<WP + 1>(code, s, cols, WP + 1)) best = x;
        }
    }
    return best;
}

// First-hit walk of len prefixes from rank q0: one bitmap read per prefix on
// the common path, the next position's signature read while the bitmap read
// is in flight, candidates handed to probe_prefix.  Stops at the lane's first
// match (lexicographic order of (prefix, x) is rank order).
template <int WP>
__device__ __forceinline__ PrefHit scan_prefix_first(const DevCode &code, const uint32_t LDS *s, uint32_t t0,
                                                     uint64_t q0, int len, int tam, const int LDS *act,
                                                     const uint32_t LDS *sg, const unsigned long long LDS *ht,
                                                     unsigned hmask) {
    PrefHit h{~0ull, 0.0f, -1};
    if (len <= 0) return h;
    const int M = tam - 1;
    int c[WP];
    const bool small = tam < kNcrLdsRows;
    if (small) unrank_t<WP>(q0, M, ncr_lds(), c); else unrank_t<WP>(q0, M, code.nCr, c);
    uint32_t acc[WP];   // acc[l] = t0 ^ sg[c0] ^ ... ^ sg[c_l]
    uint32_t a = t0;
#pragma unroll
    for (int l = 0; l < WP; ++l) { a ^= sg[c[l]]; acc[l] = a; }
    uint32_t hi = (WP >= 2) ? acc[(WP >= 2) ? WP - 2 : 0] : t0;   // signature up to the second-last element
    uint32_t t = acc[WP - 1];
    const unsigned hm = hmask & 0xFFFFu, bmsk = hmask >> 16;
    const uint32_t LDS *bm = (const uint32_t LDS *)(ht + hm + 1u);
    for (int j = 0;;) {
        const int last = c[WP - 1];
        const uint32_t nxt = sg[last + 1];            // last + 1 <= M = tam - 1
        const unsigned bi = t & bmsk;
        if ((bm[bi >> 5] >> (bi & 31u)) & 1u) {
            const int x = probe_prefix<WP>(code, s, t, last, tam, act, ht, hm, c[0], WP > 1 ? c[WP > 1 ? 1 : 0] : 0,
                                           WP > 2 ? c[WP > 2 ? 2 : 0] : 0, WP > 3 ? c[WP > 3 ? 3 : 0] : 0,
                                           WP > 4 ? c[WP > 4 ? 4 : 0] : 0, WP > 5 ? c[WP > 5 ? 5 : 0] : 0);
            if (x >= 0) {
                h.rank = small ? rank_with_last<WP>(c, x, tam, ncr_lds()) : rank_with_last<WP>(c, x, tam, code.nCr);
                h.j = j;
                break;
            }
        }
        if (++j == len) break;
        // lexicographic successor among WP-subsets of [0, M)
        if (last + 1 < M) {
            c[WP - 1] = last + 1;
            t = hi ^ nxt;
        } else {
            // end of a run: increment the highest element that can move and
            // restart the ones after it; every signature is reread (reads
            // issued together, no branch per element)
            int i = -1;
#pragma unroll
            for (int l = 0; l < WP - 1; ++l) i = (c[l] < M - WP + l) ? l : i;
            int nc[WP];
#pragma unroll
            for (int l = 0; l < WP; ++l)
                nc[l] = (l < i) ? c[l] : (l == i) ? c[l] + 1 : (l == 0) ? c[0] : nc[(l == 0) ? 0 : l - 1] + 1;
            uint32_t v[WP];
#pragma unroll
            for (int l = 0; l < WP; ++l) v[l] = sg[nc[l]];
            uint32_t a2 = t0;
#pragma unroll
            for (int l = 0; l < WP; ++l) { a2 ^= v[l]; acc[l] = a2; c[l] = nc[l]; }
            hi = (WP >= 2) ? acc[(WP >= 2) ? WP - 2 : 0] : t0;
            t = acc[WP - 1];
        }
    }
    return h;
}

template <int WP, int MODE>
__device__ __forceinline__ PrefHit scan_prefix(const DevCode &code, const uint32_t LDS *s, uint32_t t0, uint64_t q0,
                                               int len, int tam, const int LDS *act, const uint32_t LDS *sg,
                                               const float LDS *llr, const unsigned long long LDS *ht, unsigned hmask) {
    if constexpr (MODE == 0) return scan_prefix_first<WP>(code, s, t0, q0, len, tam, act, sg, ht, hmask);
    PrefHit h{~0ull, 0.0f, -1};
    if (len <= 0) return h;
    const int M = tam - 1;
    int c[WP];
    const bool small = tam < kNcrLdsRows;
    if (small) unrank_t<WP>(q0, M, ncr_lds(), c); else unrank_t<WP>(q0, M, code.nCr, c);
    uint32_t acc[WP];   // acc[l] = t0 ^ sg[c0] ^ ... ^ sg[c_l]
    uint32_t a = t0;
#pragma unroll
    for (int l = 0; l < WP; ++l) { a ^= sg[c[l]]; acc[l] = a; }
    const unsigned hm = hmask & 0xFFFFu, bmsk = hmask >> 16;
    const uint32_t LDS *bm = (const uint32_t LDS *)(ht + hm + 1u);
    for (int j = 0; j < len; ++j) {
        const uint32_t t = acc[WP - 1];
        const int last = c[WP - 1];
        const uint32_t nxt = sg[last + 1 < M ? last + 1 : last];   // the successor's signature, read ahead
        const unsigned bi = t & bmsk;
        unsigned pos = ((bm[bi >> 5] >> (bi & 31u)) & 1u) ? (t & hm) : 0xFFFFFFFFu;
        for (; pos != 0xFFFFFFFFu;) {
            const unsigned long long e = ht[pos];
            if (e == kHtEmpty) break;
            const int x = (int)(uint32_t)e;
            if ((uint32_t)(e >> 32) == t && x > last && x < tam) {
                int cols[WP + 1];
#pragma unroll
                for (int l = 0; l <= WP; ++l) cols[l] = (l < WP) ? act[c[l]] : act[x];
                if (exact_match<WP + 1>(code, s, cols, WP + 1)) {
                    const uint64_t r = small ? rank_with_last<WP>(c, x, tam, ncr_lds())
                                             : rank_with_last<WP>(c, x, tam, code.nCr);
                    if (MODE == 0) {
                        if (r < h.rank) { h.rank = r; h.j = j; }
                    } else {
                        float cost = 0.0f;
#pragma unroll
                        for (int l = 0; l < WP; ++l) cost = __fadd_rn(cost, llr[c[l]]);
                        cost = __fadd_rn(cost, llr[x]);
                        if (h.j < 0 || cost < h.cost || (cost == h.cost && r < h.rank)) { h.rank = r; h.cost = cost; h.j = j; }
                    }
                }
            }
            pos = (pos + 1u) & hm;
        }
        if (MODE == 0 && h.j >= 0) break;
        if (j + 1 == len) break;
        // lexicographic successor among WP-subsets of [0, M)
        c[WP - 1] += 1;
        if (c[WP - 1] < M) {
            acc[WP - 1] = ((WP >= 2) ? acc[(WP >= 2) ? WP - 2 : 0] : t0) ^ nxt;
        } else {
            int i = -1;
#pragma unroll
            for (int l = 0; l < WP - 1; ++l) if (c[l] < M - WP + l) i = l;
#pragma unroll
            for (int l = 0; l < WP; ++l) {
                if (l == i) c[l] += 1;
                else if (l > 0 && l > i) c[l] = c[l - 1] + 1;
            }
#pragma unroll
            for (int l = 0; l < WP; ++l)
                if (l >= i) acc[l] = ((l == 0) ? t0 : acc[(l == 0) ? 0 : l - 1]) ^ sg[c[l]];
        }
    }
    return h;
}

template <int MODE>
__device__ __attribute__((noinline)) PrefHit scan_prefix_w(int w, const DevCode &code, const uint32_t LDS *s,
                                                           uint32_t t0, uint64_t q0, int len, int tam,
                                                           const int LDS *act, const uint32_t LDS *sg,
                                                           const float LDS *llr, const unsigned long long LDS *ht,
                                                           unsigned hmask) {
    switch (w) {
        case 2: return scan_prefix<1, MODE>(code, s, t0, q0, len, tam, act, sg, llr, ht, hmask);
        case 3: return scan_prefix<2, MODE>(code, s, t0, q0, len, tam, act, sg, llr, ht, hmask);
        case 4: return scan_prefix<3, MODE>(code, s, t0, q0, len, tam, act, sg, llr, ht, hmask);
        case 5: return scan_prefix<4, MODE>(code, s, t0, q0, len, tam, act, sg, llr, ht, hmask);
        default: return scan_prefix<5, MODE>(code, s, t0, q0, len, tam, act, sg, llr, ht, hmask);
    }
}

// First-hit scan for the learned pass, whose sub-classes go up to weight 7
// (patterns up to kMaxWin columns); kept apart so that the plain search's
// dispatcher does not carry the registers of the larger cases.
__device__ __attribute__((noinline)) PrefHit scan_prefix_learned(int w, const DevCode &code, const uint32_t LDS *s,
                                                                 uint32_t t0, uint64_t q0, int len, int tam,
                                                                 const int LDS *act, const uint32_t LDS *sg,
                                                                 const unsigned long long LDS *ht, unsigned hmask) {
    switch (w) {
        case 2: return scan_prefix<1, 0>(code, s, t0, q0, len, tam, act, sg, nullptr, ht, hmask);
        case 3: return scan_prefix<2, 0>(code, s, t0, q0, len, tam, act, sg, nullptr, ht, hmask);
        case 4: return scan_prefix<3, 0>(code, s, t0, q0, len, tam, act, sg, nullptr, ht, hmask);
        case 5: return scan_prefix<4, 0>(code, s, t0, q0, len, tam, act, sg, nullptr, ht, hmask);
        case 6: return scan_prefix<5, 0>(code, s, t0, q0, len, tam, act, sg, nullptr, ht, hmask);
        default: return scan_prefix<6, 0>(code, s, t0, q0, len, tam, act, sg, nullptr, ht, hmask);
    }
}

// ------------------------------------------------------------ pair table
// Meet in the middle for first-hit classes of weight w >= 3: a w-subset
// c_0 < ... < c_{w-1} of A[0..tam) is its (w-2)-stem c_0..c_{w-3} and the pair
// (y, x) = (c_{w-2}, c_{w-1}) after it.  The table holds every pair y < x of
// the list keyed by sg[y] ^ sg[x]; a stem with target T = sig(s) ^ sig(stem)
// looks up the pairs with key T and y past the stem.  Stems are walked in
// lexicographic order and, within a stem, the lexicographically smallest valid
// pair is taken, so the hit and its rank are those of the prefix walk (and of
// a pattern-by-pattern walk); C(tam-2, w-2) stems replace C(tam-1, w-1)
// prefixes.

// Pairs of A[0..tam) in chained buckets (bucket = key bits 16..26): phead[b] is
// the first entry of bucket b (kPairNone = empty); entry p holds the key in
// pkey[p] and y | x << 10 | next << 20 in pval[p] (tam <= 1024, at most
// kPairNone entries); pbm has the bit of every key's low 15 bits.  Each lane
// takes a contiguous block of the C(tam, 2) pairs in lexicographic order and
// pushes each on its bucket's chain with one atomic exchange.
template <int NB>
__device__ __attribute__((noinline)) void build_pairs_lds(uint32_t LDS *phead, uint32_t LDS *pbm, uint32_t LDS *pkey,
                                                        uint32_t LDS *pval, const uint32_t LDS *sg, int tam) {
    const int lane = lane_id();
    static_assert(NB % 256 == 0 && kPairBitmapWords % 256 == 0, "clearing in 4-word stores");
    const uint64_t none2 = ((uint64_t)kPairNone << 32) | kPairNone;
#pragma unroll
    for (int i = 0; i < NB / 256; ++i)
        *reinterpret_cast<uint64_t LDS *>(phead + 4 * (lane + 64 * i)) = none2,
        *reinterpret_cast<uint64_t LDS *>(phead + 4 * (lane + 64 * i) + 2) = none2;
#pragma unroll
    for (int i = 0; i < kPairBitmapWords / 256; ++i)
        *reinterpret_cast<uint64_t LDS *>(pbm + 4 * (lane + 64 * i)) = 0ull,
        *reinterpret_cast<uint64_t LDS *>(pbm + 4 * (lane + 64 * i) + 2) = 0ull;
    __syncthreads();
    const int P = tam * (tam - 1) / 2;
    const int per = (P + 63) / 64;
    const int k0 = lane * per < P ? lane * per : P;
    const int k1 = k0 + per < P ? k0 + per : P;
    // first pair of the block: rows y hold tam - 1 - y pairs and row y starts at
    // pair rs(y) = y (2 tam - 1 - y) / 2; y0 = the last row starting at or before k0
    // (a square root, then exact integer correction)
    auto rs = [tam](int y) { return y * (2 * tam - 1 - y) / 2; };
    const float bq = (float)(2 * tam - 1);
    int y0 = (int)((bq - sqrtf(fmaxf(bq * bq - 8.0f * (float)k0, 0.0f))) * 0.5f);
    y0 = y0 < 0 ? 0 : (y0 > tam - 2 ? tam - 2 : y0);
    while (y0 > 0 && rs(y0) > k0) --y0;
    while (y0 < tam - 2 && rs(y0 + 1) <= k0) ++y0;
    int y = y0, x = y0 + 1 + (k0 - rs(y0));
    constexpr int U = 8;                         // pairs per step, their LDS reads and atomics issued together
    for (int k = k0; k < k1; k += U) {
        int yy[U], xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            yy[u] = y; xx[u] = x;
            const bool wrap = x + 1 == tam;
            y = wrap ? y + 1 : y;
            x = wrap ? y + 1 : x + 1;
        }
        // unconditional reads (clamped indices) so that nothing waits inside a divergent branch
        uint32_t key[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = k + u < k1;
            key[u] = sg[ok ? yy[u] : 0] ^ sg[ok ? xx[u] : 1];
        }
        // Unconditional atomics (a conditional one becomes a branch that waits for
        // its result): past the block's end a lane ORs 0 into the filter and
        // exchanges on the spare head NB, which no lookup reads.
        uint32_t nxt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = k + u < k1;
            __hip_atomic_fetch_or(&pbm[(key[u] & 0x7FFFu) >> 5], ok ? 1u << (key[u] & 31u) : 0u, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
            uint32_t LDS *hp = &phead[ok ? (key[u] >> 16) & (unsigned)(NB - 1) : (unsigned)NB];
            nxt[u] = __hip_atomic_exchange(hp, (uint32_t)(k + u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k + u < k1) {
                pkey[k + u] = key[u];
                pval[k + u] = (uint32_t)yy[u] | ((uint32_t)xx[u] << 10) | (nxt[u] << 20);
            }
    }
    __syncthreads();
}

template <bool ML>
__device__ __forceinline__ void build_pairs(const WaveLds &L, const uint32_t LDS *sg, int tam) {
    build_pairs_lds<pair_buckets(ML)>(L.phead, L.pbm, L.pkey, L.pval, sg, tam);
}

// Lexicographically smallest pair (y, x), y > last, with key T whose pattern
// (stem p0.., y, x) matches exactly, packed y << 16 | x, or -1.  Out of line.
template <int WS>
__device__ __attribute__((noinline)) int probe_pairs(const DevCode &code, const uint32_t LDS *s, uint32_t T, int last,
                                                    const int LDS *act, const uint32_t LDS *pkey,
                                                    const uint32_t LDS *pval, const uint32_t LDS *phead, int p0,
                                                    int p1, int p2, int p3) {
    const int pc[4] = {p0, p1, p2, p3};
    unsigned e = phead[(T >> 16) & (unsigned)(kPairBucketsFirstHit - 1)];
    int best = -1;
    while (e != kPairNone) {
        const uint32_t v = pval[e];
        const uint32_t key = pkey[e];
        e = v >> 20;
        if (key != T) continue;
        const int y = (int)(v & 0x3FFu), x = (int)((v >> 10) & 0x3FFu);
        if (y <= last) continue;
        if (best >= 0) {
            const int by = best >> 16, bx = best & 0xFFFF;
            if (y > by || (y == by && x >= bx)) continue;
        }
        int cols[WS + 2];
#pragma unroll
        for (int l = 0; l < WS; ++l) cols[l] = act[pc[l]];
        cols[WS] = act[y];
        cols[WS + 1] = act[x];
        if (exact_match<WS + 2>(code, s, cols, WS + 2)) best = (y << 16) | x;
    }
    return best;
}

// First-hit walk of len stems from stem rank q0 (WS-subsets of [0, tam-2)),
// with the loop structure of scan_prefix_first.  The hit's j is its stem
// offset within the lane's range.
template <int WS>
__device__ __forceinline__ PrefHit scan_stems_first(const DevCode &code, const uint32_t LDS *s, uint32_t t0,
                                                    uint64_t q0, int len, int tam, const int LDS *act,
                                                    const uint32_t LDS *sg, const uint32_t LDS *pkey,
                                                    const uint32_t LDS *pval, const uint32_t LDS *phead,
                                                    const uint32_t LDS *pbm) {
    PrefHit h{~0ull, 0.0f, -1};
    if (len <= 0) return h;
    const int M = tam - 2;
    int c[WS];
    const bool small = tam < kNcrLdsRows;
    if (small) unrank_t<WS>(q0, M, ncr_lds(), c); else unrank_t<WS>(q0, M, code.nCr, c);
    uint32_t acc[WS];
    uint32_t a = t0;
#pragma unroll
    for (int l = 0; l < WS; ++l) { a ^= sg[c[l]]; acc[l] = a; }
    uint32_t hi = (WS >= 2) ? acc[(WS >= 2) ? WS - 2 : 0] : t0;
    uint32_t t = acc[WS - 1];
    for (int j = 0;;) {
        const int last = c[WS - 1];
        const uint32_t nxt = sg[last + 1];            // last + 1 <= M = tam - 2
        if ((pbm[(t & 0x7FFFu) >> 5] >> (t & 31u)) & 1u) {
            const int yx = probe_pairs<WS>(code, s, t, last, act, pkey, pval, phead, c[0],
                                           WS > 1 ? c[WS > 1 ? 1 : 0] : 0, WS > 2 ? c[WS > 2 ? 2 : 0] : 0,
                                           WS > 3 ? c[WS > 3 ? 3 : 0] : 0);
            if (yx >= 0) {
                int cc[WS + 1];
#pragma unroll
                for (int l = 0; l < WS; ++l) cc[l] = c[l];
                cc[WS] = yx >> 16;
                const int x = yx & 0xFFFF;
                h.rank = small ? rank_with_last<WS + 1>(cc, x, tam, ncr_lds())
                               : rank_with_last<WS + 1>(cc, x, tam, code.nCr);
                h.j = j;
                break;
            }
        }
        if (++j == len) break;
        if (last + 1 < M) {
            c[WS - 1] = last + 1;
            t = hi ^ nxt;
        } else {
            int i = -1;
#pragma unroll
            for (int l = 0; l < WS - 1; ++l) i = (c[l] < M - WS + l) ? l : i;
            int nc[WS];
#pragma unroll
            for (int l = 0; l < WS; ++l)
                nc[l] = (l < i) ? c[l] : (l == i) ? c[l] + 1 : (l == 0) ? c[0] : nc[(l == 0) ? 0 : l - 1] + 1;
            uint32_t v[WS];
#pragma unroll
            for (int l = 0; l < WS; ++l) v[l] = sg[nc[l]];
            uint32_t a2 = t0;
#pragma unroll
            for (int l = 0; l < WS; ++l) { a2 ^= v[l]; acc[l] = a2; c[l] = nc[l]; }
            hi = (WS >= 2) ? acc[(WS >= 2) ? WS - 2 : 0] : t0;
            t = acc[WS - 1];
        }
    }
    return h;
}

// ML probe of one (w-2)-stem of cost scost (its LLRs summed in order): among the
// pairs (y, x), last < y < x, that complete the syndrome exactly, the cheapest
// (scost + llr[y] + llr[x] in that order, fp32 round-to-nearest; equal costs:
// the lower (y, x), i.e. the lower rank within the stem).  Returns (cost bits <<
// 32) | y << 16 | x, or ~0 for none.
template <int WS>
__device__ __attribute__((noinline)) uint64_t probe_pairs_ml(const DevCode &code, const uint32_t LDS *s, uint32_t T,
                                                             int last, float scost, const int LDS *act,
                                                             const float LDS *llr, const uint32_t LDS *pkey,
                                                             const uint32_t LDS *pval, const uint32_t LDS *phead,
                                                             int p0, int p1, int p2, int p3) {
    const int pc[4] = {p0, p1, p2, p3};
    unsigned e = phead[(T >> 16) & (unsigned)(kPairBucketsMl - 1)];
    int by = -1, bx = -1;
    float bc = 0.0f;
    while (e != kPairNone) {
        const uint32_t v = pval[e];
        const uint32_t key = pkey[e];
        e = v >> 20;
        if (key != T) continue;
        const int y = (int)(v & 0x3FFu), x = (int)((v >> 10) & 0x3FFu);
        if (y <= last) continue;
        int cols[WS + 2];
#pragma unroll
        for (int l = 0; l < WS; ++l) cols[l] = act[pc[l]];
        cols[WS] = act[y];
        cols[WS + 1] = act[x];
        if (!exact_match<WS + 2>(code, s, cols, WS + 2)) continue;
        const float cost = __fadd_rn(__fadd_rn(scost, llr[y]), llr[x]);
        if (by < 0 || cost < bc || (cost == bc && (y < by || (y == by && x < bx)))) { by = y; bx = x; bc = cost; }
    }
    return by < 0 ? ~0ull : (((uint64_t)__float_as_uint(bc) << 32) | ((uint64_t)(uint32_t)by << 16) | (uint32_t)bx);
}

// ML walk of a weight-(WS + 2) class by its (WS)-stems over the pair table:
// every stem of the lane's range (positions q0 .. q0 + len - 1 of the
// lexicographic stems of [0, tam - 2)) is probed; the cheapest pattern wins,
// equal costs the lowest rank.  The stem's signature and cost are kept per
// element, the next element's signature and LLR read ahead.
template <int WS>
__device__ __forceinline__ PrefHit scan_stems_ml(const DevCode &code, const uint32_t LDS *s, uint32_t t0, uint64_t q0,
                                                 int len, int tam, const int LDS *act, const uint32_t LDS *sg,
                                                 const float LDS *llr, const uint32_t LDS *pkey,
                                                 const uint32_t LDS *pval, const uint32_t LDS *phead,
                                                 const uint32_t LDS *pbm) {
    PrefHit h{~0ull, 0.0f, -1};
    if (len <= 0) return h;
    const int M = tam - 2;
    int c[WS];
    const bool small = tam < kNcrLdsRows;
    if (small) unrank_t<WS>(q0, M, ncr_lds(), c); else unrank_t<WS>(q0, M, code.nCr, c);
    uint32_t acc[WS];
    float cacc[WS];
    uint32_t a = t0;
    float ca = 0.0f;
#pragma unroll
    for (int l = 0; l < WS; ++l) { a ^= sg[c[l]]; acc[l] = a; ca = __fadd_rn(ca, llr[c[l]]); cacc[l] = ca; }
    uint32_t hi = (WS >= 2) ? acc[(WS >= 2) ? WS - 2 : 0] : t0;
    float hic = (WS >= 2) ? cacc[(WS >= 2) ? WS - 2 : 0] : 0.0f;
    uint32_t t = acc[WS - 1];
    float sc = cacc[WS - 1];
    for (int j = 0;;) {
        const int last = c[WS - 1];
        const uint32_t nxt = sg[last + 1];            // last + 1 <= M = tam - 2
        const float nxl = llr[last + 1];
        if ((pbm[(t & 0x7FFFu) >> 5] >> (t & 31u)) & 1u) {
            const uint64_t v = probe_pairs_ml<WS>(code, s, t, last, sc, act, llr, pkey, pval, phead, c[0],
                                                  WS > 1 ? c[WS > 1 ? 1 : 0] : 0, WS > 2 ? c[WS > 2 ? 2 : 0] : 0,
                                                  WS > 3 ? c[WS > 3 ? 3 : 0] : 0);
            if (v != ~0ull) {
                int cc[WS + 1];
#pragma unroll
                for (int l = 0; l < WS; ++l) cc[l] = c[l];
                cc[WS] = (int)((v >> 16) & 0xFFFFu);
                const int x = (int)(v & 0xFFFFu);
                const float cost = __uint_as_float((uint32_t)(v >> 32));
                const uint64_t r = small ? rank_with_last<WS + 1>(cc, x, tam, ncr_lds())
                                         : rank_with_last<WS + 1>(cc, x, tam, code.nCr);
                if (h.j < 0 || cost < h.cost || (cost == h.cost && r < h.rank)) { h.rank = r; h.cost = cost; h.j = j; }
            }
        }
        if (++j == len) break;
        if (last + 1 < M) {
            c[WS - 1] = last + 1;
            t = hi ^ nxt;
            sc = __fadd_rn(hic, nxl);
        } else {
            int i = -1;
#pragma unroll
            for (int l = 0; l < WS - 1; ++l) i = (c[l] < M - WS + l) ? l : i;
            int nc[WS];
#pragma unroll
            for (int l = 0; l < WS; ++l)
                nc[l] = (l < i) ? c[l] : (l == i) ? c[l] + 1 : (l == 0) ? c[0] : nc[(l == 0) ? 0 : l - 1] + 1;
            uint32_t v[WS];
            float lv[WS];
#pragma unroll
            for (int l = 0; l < WS; ++l) { v[l] = sg[nc[l]]; lv[l] = llr[nc[l]]; }
            uint32_t a2 = t0;
            float c2 = 0.0f;
#pragma unroll
            for (int l = 0; l < WS; ++l) {
                a2 ^= v[l]; acc[l] = a2; c2 = __fadd_rn(c2, lv[l]); cacc[l] = c2; c[l] = nc[l];
            }
            hi = (WS >= 2) ? acc[(WS >= 2) ? WS - 2 : 0] : t0;
            hic = (WS >= 2) ? cacc[(WS >= 2) ? WS - 2 : 0] : 0.0f;
            t = acc[WS - 1];
            sc = cacc[WS - 1];
        }
    }
    return h;
}
__device__ __attribute__((noinline)) PrefHit scan_stems_ml_w(int w, const DevCode &code, const uint32_t LDS *s,
                                                            uint32_t t0, uint64_t q0, int len, int tam,
                                                            const int LDS *act, const uint32_t LDS *sg,
                                                            const float LDS *llr, const uint32_t LDS *pkey,
                                                            const uint32_t LDS *pval, const uint32_t LDS *phead,
                                                            const uint32_t LDS *pbm) {
    switch (w) {
        case 3: return scan_stems_ml<1>(code, s, t0, q0, len, tam, act, sg, llr, pkey, pval, phead, pbm);
        case 4: return scan_stems_ml<2>(code, s, t0, q0, len, tam, act, sg, llr, pkey, pval, phead, pbm);
        case 5: return scan_stems_ml<3>(code, s, t0, q0, len, tam, act, sg, llr, pkey, pval, phead, pbm);
        default: return scan_stems_ml<4>(code, s, t0, q0, len, tam, act, sg, llr, pkey, pval, phead, pbm);
    }
}
__device__ __attribute__((noinline)) PrefHit scan_stems_w(int w, const DevCode &code, const uint32_t LDS *s,
                                                         uint32_t t0, uint64_t q0, int len, int tam,
                                                         const int LDS *act, const uint32_t LDS *sg,
                                                         const uint32_t LDS *pkey, const uint32_t LDS *pval,
                                                         const uint32_t LDS *phead, const uint32_t LDS *pbm) {
    switch (w) {
        case 3: return scan_stems_first<1>(code, s, t0, q0, len, tam, act, sg, pkey, pval, phead, pbm);
        case 4: return scan_stems_first<2>(code, s, t0, q0, len, tam, act, sg, pkey, pval, phead, pbm);
        case 5: return scan_stems_first<3>(code, s, t0, q0, len, tam, act, sg, pkey, pval, phead, pbm);
        default: return scan_stems_first<4>(code, s, t0, q0, len, tam, act, sg, pkey, pval, phead, pbm);
    }
}

// Whether a first-hit class of weight w over tam positions is walked by stems:
// the table must fit and the class must have many more prefixes than the list
// has pairs (the table is built once per list and reused by its later classes).
__device__ __forceinline__ bool use_pairs(const DevCode &code, const DevCall &call, int w, int tam) {
    if (call.pair_cap <= 0 || w < 3 || w > 6 || tam < 4) return false;
    const uint64_t P = (uint64_t)tam * (uint64_t)(tam - 1) / 2u;
    return P <= (uint64_t)call.pair_cap && ncr_w(code, tam - 1, w - 1) > 2u * P;
}

// Weight 1: every position with sig == t0, scanned by the wave 64 at a time.
// MODE 0 returns the lowest matching position; MODE 1 the cheapest (ties: lowest).
template <int MODE>
__device__ PrefHit weight_one(const DevCode &code, const uint32_t LDS *s, uint32_t t0, int tam, const int LDS *act,
                              const uint32_t LDS *sg, const float LDS *llr) {
    const int lane = lane_id();
    PrefHit best{~0ull, 0.0f, -1};
    for (int i0 = 0; i0 < tam; i0 += 64) {
        const int i = i0 + lane;
        bool hit = false;
        if (i < tam && sg[i] == t0) {
            int cols[1] = {act[i]};
            hit = exact_match<1>(code, s, cols, 1);
        }
        const float cost = (hit && MODE == 1) ? __fadd_rn(0.0f, llr[i]) : 0.0f;
        const uint32_t key = hit ? order_key(cost) : 0xFFFFFFFFu;
        const uint32_t kmin = (MODE == 1) ? wave_min(key) : 0u;
        const uint64_t b = ballot(hit && (MODE == 0 || key == kmin));
        if (b) {
            const int bl = __ffsll((long long)b) - 1;
            const float cb = bcast_f(cost, bl);
            if (best.j < 0 || (MODE == 1 && cb < best.cost)) { best.rank = (uint64_t)(i0 + bl); best.cost = cb; best.j = 0; }
            if (MODE == 0) break;
        }
    }
    return best;
}

// Adds each lane's count to call.work[idx] (one atomic per wave), when counting.
__device__ __forceinline__ void count_work(unsigned long long *work, int idx, uint32_t n) {
    if (!work) return;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) n += (uint32_t)__shfl_xor((int)n, o, 64);
    if (lane_id() == 0 && n) atomicAdd(&work[idx], (unsigned long long)n);
}
// Steps a lane's walk of len units executed: up to and including its hit.
__device__ __forceinline__ uint32_t walked(const PrefHit &h, int len, bool ml) {
    return (uint32_t)((!ml && h.j >= 0) ? h.j + 1 : (len > 0 ? len : 0));
}

// ---------------------------------------------------------------- shared tasks
// ctl word: generation (20 bits) | nchunks (20 bits) | next chunk (24 bits).
// Chunks are claimed with one fetch-and-add on ctl: a returned `next` below
// `nchunks` is a claim of that chunk of generation `gen`; anything else claims
// nothing.  A wave adds only after it read next < nchunks, so the counter runs
// past nchunks by at most one step per resident wave (< 2^12), far from
// carrying into the nchunks field.  The generation counts the publications of
// the slot (it wraps after 2^20 of them).  (slot, generation), as returned by
// the claiming fetch-and-add itself, identifies the task a chunk belongs to.
constexpr unsigned kMaxChunks = (1u << 20) - 1;
constexpr int kStemCost = 4;    // prefix steps one stem step is counted as (split decision, chunk size)
constexpr int kMinStemR = 4;
__device__ __forceinline__ unsigned ctl_next(unsigned long long c) { return (unsigned)(c & 0xFFFFFFull); }
__device__ __forceinline__ unsigned ctl_nch(unsigned long long c) { return (unsigned)((c >> 24) & 0xFFFFFull); }
__device__ __forceinline__ unsigned ctl_gen(unsigned long long c) { return (unsigned)(c >> 44); }
__device__ __forceinline__ unsigned long long make_ctl(unsigned gen, unsigned nch) {
    return ((unsigned long long)(gen & 0xFFFFFu) << 44) | ((unsigned long long)nch << 24);
}
// best_q holds the task's generation in its top 20 bits and the unit index of
// the lowest hit so far in the low 44.  Every unit index of a task is below
// 2^44: its classes have weights 2..6 over at most 1023 positions, and the sum
// over w = 2..6 of C(1023, w - 1) prefixes is 9.3e12 < 2^44.  A claim reads it
// in the same round trip as the claiming fetch-and-add, so the value may
// predate the task's publication; one with another generation is ignored.
constexpr int kQBits = 44;
constexpr unsigned long long kQMask = (1ull << kQBits) - 1ull;
__device__ __forceinline__ unsigned long long make_bq(unsigned gen, unsigned long long q) {
    return ((unsigned long long)(gen & 0xFFFFFu) << kQBits) | (q & kQMask);
}

template <class T>
__device__ __forceinline__ T ld_relaxed(T *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <class T>
__device__ __forceinline__ void st_relaxed(T *p, T v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// Shared-task data is written with device-coherent (write-through) stores and
// read with device-coherent loads, so no cache-wide acquire/release is needed:
// a writer only waits for its own stores and atomics to complete before the
// store or atomic that publishes them.  (An agent-scope acquire or release
// fence would invalidate or write back the whole L2 of the XCD each time.)
// s_waitcnt 0 waits for every vector memory operation of the wave, stores
// included (gfx9 counts them in vmcnt).  The intrinsic alone is not a memory
// operation for the compiler, which could move memory accesses across it; the
// signal fences on both sides keep every access of the wave on its side.
__device__ __forceinline__ void wait_mem() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ float key_to_float(uint32_t key) {
    const uint32_t b = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
    return __uint_as_float(b);
}

// Claims the next chunk of slot T for this wave; returns it or ~0u, with the
// task's generation in gen.  Lane 0 issues the fetch-and-add and, behind it in
// the same round trip, the load of best_q; bq is its unit index when it
// carries the claimed task's generation, else ~0 (an older task's value, read
// before the claim was performed).  best_q only decreases during a task, so a
// value of the right generation at worst searches a chunk past a known hit.
// Every exit depends only on wave-uniform values.
__device__ __forceinline__ unsigned claim_chunk(TaskSlot *T, unsigned *open_count, unsigned &gen,
                                                unsigned long long &bq) {
    gen = 0;
    bq = ~0ull;
    const unsigned long long snap = uni64(ld_relaxed(&T->ctl));
    if (ctl_next(snap) >= ctl_nch(snap)) return ~0u;
    unsigned long long old = 0, b = ~0ull;
    if (lane_id() == 0) {
        old = __hip_atomic_fetch_add(&T->ctl, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b = ld_relaxed(&T->best_q);
    }
    old = uni64(old);
    b = uni64(b);
    if (ctl_next(old) >= ctl_nch(old)) return ~0u;
    const unsigned ch = ctl_next(old);
    if (lane_id() == 0 && ch + 1u == ctl_nch(old)) atomicSub(open_count, 1u);
    gen = ctl_gen(old);
    bq = ((unsigned)(b >> kQBits) == gen) ? (b & kQMask) : ~0ull;
    return ch;
}

// Units of one class of a task: (w-2)-stems with the pair table, else (w-1)-prefixes.
__device__ __forceinline__ uint64_t class_units(const DevCode &code, int tam, int w, bool pairs) {
    return pairs ? ncr_w(code, tam - 2, w - 2) : ncr_w(code, tam - 1, w - 1);
}

// Chunks of a task's class: units per lane per chunk, bounded so that the
// chunk count fits the 20-bit ctl field.  A stem step costs several prefix
// steps, so stem classes get a quarter of the units per lane (at least
// kMinStemR).
__device__ __forceinline__ int class_chunk_r(const DevCall &call, uint64_t units, bool pairs) {
    int R = pairs ? (call.rmax / kStemCost > kMinStemR ? call.rmax / kStemCost : kMinStemR) : call.rmax;
    const uint64_t per = 64ull * (uint64_t)kMaxChunks;
    if (units / per >= (uint64_t)R) R = (int)(units / per + 1);
    return R;
}

// One chunk (64 lanes x R units) of a shared task; the list's arrays and
// tables are in this wave's LDS.  A task holds consecutive weight classes of
// one list; chunk ch belongs to the class whose chunk range holds it.
// First-hit: atomic-min of the matching pattern index over the task's classes
// concatenated in list order (earlier classes' sizes plus the rank), and of
// its unit index (so that chunks past a known hit are skipped).  ML (one
// class): atomic-min of (cost key << 32 | rank).  Returns the chunk's result
// (~0 = none).
// The task's descriptor comes in registers: nclass, wpack, ppack, and per lane
// cch (lane i: first chunk of class i, lane nclass: the chunk count) and cR
// (lane i: units per lane per chunk of class i); bq is a recent best_q unit
// index of generation gen (~0: none known).
template <bool ML>
__device__ unsigned long long task_chunk(const DevCode &code, TaskSlot *T, const uint32_t LDS *s, uint32_t t0,
                                        int tam, unsigned ch, unsigned gen, unsigned long long bq, int nclass,
                                        unsigned wpack, unsigned ppack, unsigned cch, int cR, const int LDS *act,
                                        const uint32_t LDS *sg, const float LDS *llr,
                                        const unsigned long long LDS *ht, unsigned hmask,
                                        const uint32_t LDS *pkey, const uint32_t LDS *pval,
                                        const uint32_t LDS *phead, const uint32_t LDS *pbm,
                                        unsigned long long *work) {
    constexpr bool ml = ML;
    const int lane = lane_id();
    int w = 0, R = 1;
    bool pairs = false;
    uint64_t ubase = 0, pbase = 0, NP = 0;
    unsigned ch0 = 0;
    for (int i = 0; i < nclass; ++i) {
        const int wi = (int)((wpack >> (4 * i)) & 15u);
        const bool pi = ((ppack >> i) & 1u) != 0;
        const unsigned c1 = (unsigned)__builtin_amdgcn_readlane((int)cch, i + 1);
        if (ch < c1) {
            w = wi; pairs = pi;
            ch0 = (unsigned)__builtin_amdgcn_readlane((int)cch, i);
            R = __builtin_amdgcn_readlane(cR, i);
            NP = class_units(code, tam, wi, pi);
            break;
        }
        ubase += class_units(code, tam, wi, pi);
        pbase += ncr_w(code, tam, wi);
    }
    const uint64_t base = (uint64_t)(ch - ch0) * 64ull * (uint64_t)R;
    if (!ml && ubase + base > bq) return ~0ull - 1;
    const uint64_t q0 = base + (uint64_t)lane * (uint64_t)R;
    const int len = q0 >= NP ? 0 : (int)((NP - q0) < (uint64_t)R ? (NP - q0) : (uint64_t)R);
    const PrefHit h = ml ? (pairs ? scan_stems_ml_w(w, code, s, t0, q0, len, tam, act, sg, llr, pkey, pval, phead, pbm)
                                  : scan_prefix_w<1>(w, code, s, t0, q0, len, tam, act, sg, llr, ht, hmask))
                    : pairs ? scan_stems_w(w, code, s, t0, q0, len, tam, act, sg, pkey, pval, phead, pbm)
                            : scan_prefix_w<0>(w, code, s, t0, q0, len, tam, act, sg, llr, ht, hmask);
    count_work(work, pairs ? 1 : 0, walked(h, len, ml));
    unsigned long long out = ~0ull;
    if (!ml) {
        const uint64_t b = ballot(h.j >= 0);
        if (b) {
            const int hl = __ffsll((long long)b) - 1;
            const uint64_t r = pbase + (uint64_t)__shfl((long long)h.rank, hl, 64);
            const uint64_t q = ubase + q0 + (uint64_t)(h.j < 0 ? 0 : h.j);
            const uint64_t qh = (uint64_t)__shfl((long long)q, hl, 64);
            if (lane == 0) { atomicMin(&T->best, (unsigned long long)r); atomicMin(&T->best_q, make_bq(gen, qh)); }
            out = r;
        }
    } else {
        const uint32_t key = (h.j >= 0) ? order_key(h.cost) : 0xFFFFFFFFu;
        const uint32_t kmin = wave_min(key);
        // among lanes with the minimal key, the lowest rank (lanes hold ascending ranges)
        const uint64_t b = ballot(h.j >= 0 && key == kmin);
        if (b) {
            const int bl = __ffsll((long long)b) - 1;
            const uint64_t r = (uint64_t)__shfl((long long)h.rank, bl, 64);
            if (lane == 0) atomicMin(&T->best, ((unsigned long long)kmin << 32) | r);
            out = ((unsigned long long)kmin << 32) | r;
        }
    }
    return out;
}

// Chunk ranges of a task's classes (lane i < nclass: first chunk and units per
// lane per chunk of class i); returns the chunk count.
__device__ __forceinline__ unsigned task_ranges(const DevCode &code, const DevCall &call, int tam, int nclass,
                                                unsigned wpack, unsigned ppack, unsigned &myc, int &myR) {
    const int lane = lane_id();
    unsigned nch = 0;
    myc = 0;
    myR = 1;
    for (int i = 0; i < nclass; ++i) {
        const int wi = (int)((wpack >> (4 * i)) & 15u);
        const bool pi = ((ppack >> i) & 1u) != 0;
        const uint64_t u = class_units(code, tam, wi, pi);
        const int Ri = class_chunk_r(call, u, pi);
        if (lane == i) { myc = nch; myR = Ri; }
        nch += (unsigned)((u + 64ull * (uint64_t)Ri - 1) / (64ull * (uint64_t)Ri));
    }
    return nch;
}

// Waits until every chunk of T is done (bounded: past the limit the kernel
// raises kErrTaskWait instead of waiting on).
constexpr unsigned kTaskWaitLimit = 1u << 22;
__device__ __forceinline__ void wait_task_done(const DevCall &call, TaskSlot *T, unsigned nch) {
    for (unsigned spin = 0; uni(ld_relaxed(&T->done)) < nch; ++spin) {
        if (spin >= kTaskWaitLimit) {
            if (lane_id() == 0) atomicOr(&call.counters[2], kErrTaskWait);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// A free task slot, taken (busy = 1), or -1 when none is free.  64 candidates
// per step; lane 0 takes the first free one.  Out of line: publications are
// rare, and the scan's registers stay out of the callers.
__device__ __attribute__((noinline)) int find_free_slot(TaskSlot *tasks) {
    const int lane = lane_id();
    // find a free slot: 64 candidates per step, lane 0 takes the first free one
    int slot = -1;
    const int start = (int)((blockIdx.x * 131u) & (kTaskSlots - 1));
    // a slot is free when idle, or withdrawn with every chunk done (its last
    // chunk's result atomics completed before that chunk counted as done)
    for (int b0 = 0; b0 < kTaskSlots; b0 += 64) {
        const int k = (start + b0 + lane) & (kTaskSlots - 1);
        const unsigned bz = ld_relaxed(&tasks[k].busy);
        const unsigned dn = ld_relaxed(&tasks[k].done);
        const unsigned nc = ld_relaxed(&tasks[k].nchunks);
        const unsigned want = (bz == 0u || (bz == 2u && dn >= nc)) ? bz : 1u;
        uint64_t freem = ballot(want != 1u);
        while (freem) {
            const int bit = __ffsll((long long)freem) - 1;
            const int pick = (int)uni((unsigned)(start + b0 + bit) & (kTaskSlots - 1));
            const unsigned expect = (unsigned)__builtin_amdgcn_readlane((int)want, bit);
            unsigned old = 1u;
            if (lane == 0) old = atomicCAS(&tasks[pick].busy, expect, 1u);
            if (uni(old) == expect) {
                if (expect == 0u) { slot = pick; break; }
                // a withdrawn slot: its done count, read after the claim, must
                // show every chunk done (the loads above may predate it)
                unsigned ok = 0u;
                if (lane == 0) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    ok = ld_relaxed(&tasks[pick].done) >= ld_relaxed(&tasks[pick].nchunks) ? 1u : 0u;
                    if (!ok) st_relaxed(&tasks[pick].busy, 2u);
                }
                if (uni(ok)) { slot = pick; break; }
            }
            freem &= freem - 1;
        }
        if (slot >= 0) break;
    }
    return slot;
}

// Owner side: publishes nclass consecutive weight classes of the list act/sg/llr
// (weights in wpack, 4 bits each; bit i of ppack: class i by stems) as one
// shared task and returns its slot, or -1 when no slot is free.  The task is
// complete in memory before its ctl is published; helpers build their own
// tables from the copied list.
__device__ int publish_shared(const DevCode &code, const DevCall &call, int shot, const uint32_t LDS *s, uint32_t t0,
                              int tam, bool ml, int nclass, unsigned wpack, unsigned ppack, const int LDS *act,
                              const uint32_t LDS *sg, const float LDS *llr) {
    const int lane = lane_id();
    const int slot = find_free_slot(call.tasks);
    if (slot < 0) return -1;
    TaskSlot *T = &call.tasks[slot];
    uint32_t *arr = call.task_arrays + (size_t)slot * 3 * call.cap;
    for (int i = lane; i < tam; i += 64) {
        st_relaxed(&arr[i], (uint32_t)act[i]);
        st_relaxed(&arr[call.cap + i], (uint32_t)sg[i]);
        if (ml) st_relaxed(&arr[2 * call.cap + i], __float_as_uint(llr[i]));
    }
    unsigned myc;
    int myR;
    const unsigned nch = task_ranges(code, call, tam, nclass, wpack, ppack, myc, myR);
    const unsigned gen = (uni((unsigned)ctl_gen(ld_relaxed(&T->ctl))) + 1u) & 0xFFFFFu;
    if (lane == 0) {
        st_relaxed(&T->nchunks, nch); st_relaxed(&T->shot, shot);
        st_relaxed(&T->tam, tam); st_relaxed(&T->ml, ml ? 1 : 0); st_relaxed(&T->t0, t0);
        st_relaxed(&T->nclass, nclass); st_relaxed(&T->wpack, (int)wpack); st_relaxed(&T->ppack, (int)ppack);
        st_relaxed(&T->cch[nclass], nch);
        st_relaxed(&T->done, 0u); st_relaxed(&T->best, ~0ull); st_relaxed(&T->best_q, make_bq(gen, kQMask));
    }
    if (lane < nclass) { st_relaxed(&T->cch[lane], myc); st_relaxed(&T->cR[lane], myR); }
    if (lane < kMaxWords) st_relaxed(&T->s[lane], (uint32_t)s[lane]);
    wait_mem();                                   // the task is complete in memory before it is published
    if (lane == 0) {
        __hip_atomic_exchange(&T->ctl, make_ctl(gen, nch), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(&call.task_counters[0], 1u);
    }
    return slot;
}

// Owner side: works on its published task in `slot` (the list and tables in
// this wave's LDS), waits for every chunk, frees the slot and returns the
// result in best (~0 = no match): first-hit, the lowest matching pattern index
// over the classes concatenated in order; ML (one class), the lowest (cost
// key, rank).
template <bool ML>
__device__ void work_shared(const DevCode &code, const DevCall &call, int slot, const uint32_t LDS *s, uint32_t t0,
                            int tam, int nclass, unsigned wpack, unsigned ppack, const int LDS *act,
                            const uint32_t LDS *sg, const float LDS *llr, const unsigned long long LDS *ht,
                            unsigned hmask, const uint32_t LDS *pkey, const uint32_t LDS *pval,
                            const uint32_t LDS *phead, const uint32_t LDS *pbm, unsigned long long &best) {
    const int lane = lane_id();
    TaskSlot *T = &call.tasks[slot];
    unsigned myc;
    int myR;
    const unsigned nch = task_ranges(code, call, tam, nclass, wpack, ppack, myc, myR);
    const unsigned cchv = (lane < nclass) ? myc : (lane == nclass ? nch : 0u);
    for (;;) {
        unsigned g;
        unsigned long long bq;
        const unsigned ch = claim_chunk(T, &call.task_counters[0], g, bq);
        if (ch == ~0u) break;
        (void)task_chunk<ML>(code, T, s, t0, tam, ch, g, bq, nclass, wpack, ppack, cchv, myR, act, sg, llr, ht, hmask,
                         pkey, pval, phead, pbm, call.work);
        wait_mem();                               // the chunk's result atomics complete before it counts as done
        if (lane == 0) atomicAdd(&T->done, 1u);
    }
    wait_task_done(call, T, nch);
    best = uni64(ld_relaxed(&T->best));
    wait_mem();
    if (lane == 0) st_relaxed(&T->busy, 0u);
}

// Owner side: withdraws its published task in `slot` (its result is no longer
// needed), without waiting.  One exchange on ctl ends the claims; the chunks
// nobody claimed count as done (and the task leaves the open count, which its
// last claim would otherwise have done); chunks claimed later see best_q = 0
// and skip their search.  The slot is marked withdrawn: a publisher takes it
// once every chunk in flight is done.  (Only the owner changes the slot's
// generation and chunk count, so its own read of ctl holds them.)
__device__ void cancel_shared(const DevCall &call, int slot) {
    TaskSlot *T = &call.tasks[slot];
    if (lane_id() == 0) {
        const unsigned long long c = ld_relaxed(&T->ctl);
        const unsigned nch = ctl_nch(c);
        atomicMin(&T->best_q, make_bq(ctl_gen(c), 0ull));
        const unsigned long long prev = __hip_atomic_exchange(&T->ctl, (c & ~0xFFFFFFull) | (unsigned long long)nch,
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned nxt = ctl_next(prev);
        if (nxt < nch) {
            atomicAdd(&T->done, nch - nxt);
            atomicSub(&call.task_counters[0], 1u);
        }
    }
    wait_mem();                                   // the unclaimed chunks count as done before the slot is withdrawn
    if (lane_id() == 0) st_relaxed(&T->busy, 2u);
}

// Owner side: publish a task and work on it; false when no slot is free (the
// caller searches alone).
template <bool ML>
__device__ bool run_shared(const DevCode &code, const DevCall &call, int shot, const uint32_t LDS *s, uint32_t t0,
                           int tam, int nclass, unsigned wpack, unsigned ppack, const int LDS *act,
                           const uint32_t LDS *sg, const float LDS *llr, const unsigned long long LDS *ht, unsigned hmask,
                           const uint32_t LDS *pkey, const uint32_t LDS *pval,
                           const uint32_t LDS *phead, const uint32_t LDS *pbm, unsigned long long &best) {
    const int slot = publish_shared(code, call, shot, s, t0, tam, ML, nclass, wpack, ppack, act, sg, llr);
    if (slot < 0) return false;
    work_shared<ML>(code, call, slot, s, t0, tam, nclass, wpack, ppack, act, sg, llr, ht, hmask, pkey, pval, phead,
                pbm, best);
    return true;
}

// Helper side: claim one chunk of any open task and process it.  The task's
// arrays are copied into this wave's LDS and its tables (hash table and/or pair
// table, as its classes need) built once per task.
// The task a helper wave has in its LDS (arrays, syndrome words, tables) and
// its descriptor in registers; key = slot | generation << 7 of that task, as
// returned by the claim that loaded it (-1: none).
struct HelpCache {
    int key;
    unsigned hmask;
    int tam, nclass;
    unsigned wpack, ppack;
    unsigned cch;                     // lane i: first chunk of class i; lane nclass: chunk count
    int cR;                           // lane i: units per lane per chunk of class i
    uint32_t t0;
};

template <bool ML>
__device__ __forceinline__ bool help_task(const DevCode &code, const DevCall &call, const WaveLds &L, HelpCache &hc) {
    const int lane = lane_id();
    if (__builtin_amdgcn_readfirstlane(ld_relaxed(&call.task_counters[0])) == 0u) return false;
    static_assert(kTaskSlots == 128, "the open-slot scan reads two 64-slot halves");
    const int start = (int)((blockIdx.x * 37u) & (kTaskSlots - 1));
    const unsigned long long c0 = ld_relaxed(&call.tasks[(start + lane) & (kTaskSlots - 1)].ctl);
    const unsigned long long c1 = ld_relaxed(&call.tasks[(start + 64 + lane) & (kTaskSlots - 1)].ctl);
    const uint64_t b0 = ballot(ctl_next(c0) < ctl_nch(c0));
    const uint64_t b1 = ballot(ctl_next(c1) < ctl_nch(c1));
    int slot = -1;
    if (b0) slot = (int)uni((unsigned)(start + __ffsll((long long)b0) - 1) & (kTaskSlots - 1));
    else if (b1) slot = (int)uni((unsigned)(start + 64 + __ffsll((long long)b1) - 1) & (kTaskSlots - 1));
    if (slot < 0) return false;
    if (hc.key >= 0) {
        // stay on the task whose tables this wave holds while it has chunks
        const int cs = hc.key & (kTaskSlots - 1);
        const int rel = (cs - start) & (kTaskSlots - 1);
        if ((((rel < 64) ? b0 : b1) >> (rel & 63)) & 1ull) slot = cs;
    }
    TaskSlot *T = &call.tasks[slot];
    unsigned g;
    unsigned long long bq;
    const unsigned ch = claim_chunk(T, &call.task_counters[0], g, bq);
    if (ch == ~0u) return true;
    static_assert(kTaskSlots <= 128, "the cache key holds the slot in 7 bits");
    const int key = (int)((unsigned)slot | (g << 7));
    if (hc.key != key) {
        // The task was complete in memory before its ctl was published.  These
        // device-coherent loads are issued after the claim returned (the key
        // comes from its result), so they read the published task; the fences
        // keep the compiler from moving them above that branch.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        hc.tam = ld_relaxed(&T->tam);
        hc.nclass = ld_relaxed(&T->nclass);
        hc.wpack = (unsigned)ld_relaxed(&T->wpack);
        hc.ppack = (unsigned)ld_relaxed(&T->ppack);
        hc.t0 = ld_relaxed(&T->t0);
        hc.cch = (lane <= kTaskClasses) ? ld_relaxed(&T->cch[lane]) : 0u;
        hc.cR = (lane < kTaskClasses) ? ld_relaxed(&T->cR[lane]) : 1;
        const uint32_t sword = (lane < kMaxWords) ? ld_relaxed(&T->s[lane]) : 0u;
        hc.tam = uni((unsigned)hc.tam);
        hc.nclass = uni((unsigned)hc.nclass);
        hc.wpack = uni(hc.wpack);
        hc.ppack = uni(hc.ppack);
        hc.t0 = uni(hc.t0);
        if (lane < kMaxWords) L.sw[lane] = sword;
        constexpr bool ml = ML;
        int LDS *act = ml ? L.P : L.A;
        uint32_t LDS *sg = ml ? L.sigP : L.sigA;
        const uint32_t *arr = call.task_arrays + (size_t)slot * 3 * call.cap;
        for (int i = lane; i < hc.tam; i += 64) {
            act[i] = (int)ld_relaxed(&arr[i]);
            sg[i] = ld_relaxed(&arr[call.cap + i]);
            if (ml) L.llrP[i] = __uint_as_float(ld_relaxed(&arr[2 * call.cap + i]));
        }
        __syncthreads();
        const unsigned all = (hc.nclass >= 32) ? ~0u : ((1u << hc.nclass) - 1u);
        if ((hc.ppack & all) != 0u) build_pairs<ML>(L, sg, hc.tam);
        if ((hc.ppack & all) != all) hc.hmask = build_ht(L.ht, sg, hc.tam);
        hc.key = key;
    }
    constexpr bool ml = ML;
    const uint32_t LDS *s = L.sw;
    const int LDS *act = ml ? L.P : L.A;
    const uint32_t LDS *sg = ml ? L.sigP : L.sigA;
    (void)task_chunk<ML>(code, T, s, hc.t0, hc.tam, ch, g, bq, hc.nclass, hc.wpack, hc.ppack, hc.cch, hc.cR, act, sg,
                     ml ? L.llrP : nullptr, L.ht, hc.hmask, L.pkey, L.pval, L.phead, L.pbm, call.work);
    wait_mem();                                   // the chunk's result atomics complete before it counts as done
    if (lane == 0) atomicAdd(&T->done, 1u);
    return true;
}

// ------------------------------------------------------------ class searches
constexpr uint64_t kFlatMax = 128;   // classes up to this size go pattern by pattern, one per lane

// First matching rank in the weight-w class over act/sg (hash table ht, or the
// pair table with pairs), searched by this wave alone, or ~0.
__device__ uint64_t class_first_hit(const DevCode &code, const DevCall &call, int shot, const uint32_t LDS *s,
                                    uint32_t t0, int w, int tam, uint64_t Nw, const int LDS *act,
                                    const uint32_t LDS *sg, const unsigned long long LDS *ht, unsigned hmask,
                                    bool pairs, const uint32_t LDS *pkey, const uint32_t LDS *pval,
                                    const uint32_t LDS *phead, const uint32_t LDS *pbm) {
    const int lane = lane_id();
    if (w == 1) {
        const PrefHit h1 = weight_one<0>(code, s, t0, tam, act, sg, nullptr);
        count_work(call.work, 2, lane == 0 ? (uint32_t)(h1.j >= 0 ? h1.rank + 1 : tam) : 0u);
        return h1.rank;
    }
    // units: stems with the pair table, else prefixes
    const uint64_t NP = class_units(code, tam, w, pairs);
    uint64_t pos = 0;
    while (pos < NP) {
        const uint64_t rem = NP - pos;
        const uint64_t r64 = (rem + 63u) / 64u;
        const int R = (int)(r64 < (uint64_t)call.rmax ? r64 : (uint64_t)call.rmax);
        const uint64_t q0 = pos + (uint64_t)lane * (uint64_t)R;
        const int len = q0 >= NP ? 0 : (int)((NP - q0) < (uint64_t)R ? (NP - q0) : (uint64_t)R);
        const PrefHit h = pairs ? scan_stems_w(w, code, s, t0, q0, len, tam, act, sg, pkey, pval, phead, pbm)
                                : scan_prefix_w<0>(w, code, s, t0, q0, len, tam, act, sg, nullptr, ht, hmask);
        count_work(call.work, pairs ? 1 : 0, walked(h, len, false));
        const uint64_t b = ballot(h.j >= 0);
        if (b) return (uint64_t)__shfl((long long)h.rank, __ffsll((long long)b) - 1, 64);
        pos += 64ull * (uint64_t)R;
    }
    return ~0ull;
}

// Sum of (weight + 1) over the first `upto` patterns of the concatenated classes.
__device__ __forceinline__ uint64_t weighted_prefix(const uint64_t *N, const int *wl, int nw, uint64_t upto) {
    uint64_t acc = 0, cs = 0;
#pragma unroll
    for (int k = 0; k < kMaxW; ++k) {
        if (k < nw && upto > cs) {
            const uint64_t take = (upto - cs) < N[k] ? (upto - cs) : N[k];
            acc += take * (uint64_t)(wl[k] + 1);
        }
        if (k < nw) cs += N[k];
    }
    return acc;
}

// Work left in a run of classes (weights in wl4, 4 bits each, from the current
// class on; the run ends at the list's end or at weight 1), in prefix steps:
// the classes after the current one are searched anyway unless it hits, so the
// run is shared as one task when this exceeds the split threshold.
__device__ __attribute__((noinline)) uint64_t rest_work(const DevCode &code, const DevCall &call, uint64_t wl4, int tam) {
    uint64_t rest = 0;
    for (; (wl4 & 15u) >= 2u; wl4 >>= 4) {
        const int w = (int)(wl4 & 15u);
        if (w > tam) continue;
        const bool pk = use_pairs(code, call, w, tam);
        rest += class_units(code, tam, w, pk) * (pk ? (uint64_t)kStemCost : 1u);
    }
    return rest;
}

// ------------------------------------------------------------ first-hit search
// The run of classes that first_hit_wave shares for a weight list over tam
// positions, if it shares one (and reaches it without a hit): first_hit_wave's
// walk over the class sizes, without the searches.  The list comes packed,
// 4 bits per weight (first in the low bits, 0 ends it), without classes of
// weight above tam (empty: the walk passes over them).  nc classes, their
// weights in wpack, bit i of ppack set when class i is walked by stems.
__device__ __forceinline__ bool find_split(const DevCode &code, const DevCall &call, int tam, uint64_t wl4,
                                           int &nc, unsigned &wpack, unsigned &ppack) {
    while (wl4) {
        const int w = (int)(wl4 & 15u);
        const uint64_t N = ncr_w(code, tam, w);
        if (N <= kFlatMax) {                      // a run of small classes goes pattern by pattern
            wl4 >>= 4;
            continue;
        }
        if (w >= 2 && call.split_threshold && rest_work(code, call, wl4, tam) > call.split_threshold) {
            nc = 0; wpack = 0u; ppack = 0u;
            for (uint64_t r = wl4; r && nc < kTaskClasses && (r & 15u) >= 2u; r >>= 4) {
                const int wr = (int)(r & 15u);
                const bool pk = use_pairs(code, call, wr, tam);
                wpack |= (unsigned)wr << (4 * nc);
                ppack |= (pk ? 1u : 0u) << nc;
                ++nc;
            }
            return true;
        }
        wl4 >>= 4;                                // searched alone
    }
    return false;
}

// Weight list wl[0..nw) searched in order over A[0..tam) (grand_decoder.cu:295-350,
// :534-588); on a hit writes the decoded columns to win and returns true; count
// gets the patterns tested up to and including the hit.
// spec: slot of this list's shared run when it was published before the
// search (-1: none).  Returns bit 0: hit; bit 1: spec was used or withdrawn.
__device__ int first_hit_wave(const DevCode &code, const DevCall &call, int shot, const uint32_t LDS *s, uint32_t t0,
                              int tam, const int *wl, int nw, const WaveLds &L, int *win, int &fw,
                              unsigned long long &count, unsigned long long &wops, int spec) {
    int spec_done = 0;
    const int lane = lane_id();
    unsigned hmask = 0;          // hash table of A, built before the first class that needs it
    bool have_table = false;
    bool have_pairs = false;     // pair table of A, likewise
    uint64_t N[kMaxW];
    uint64_t total = 0;
#pragma unroll
    for (int k = 0; k < kMaxW; ++k) {
        N[k] = (k < nw) ? ncr_w(code, tam, wl[k]) : 0ull;
        total += N[k];
    }
    uint64_t wl_packed = 0;        // the weight list, 4 bits per class
#pragma unroll
    for (int k = 0; k < 2 * kMaxW; ++k) if (k < nw) wl_packed |= (uint64_t)wl[k] << (4 * k);
    uint64_t class_start = 0;
    int kc = 0;
    bool hit = false;
    uint64_t hit_idx = 0;
    int hit_w = 0;
    int c[kMaxW] = {0, 0, 0, 0, 0, 0};
    int hit_lane = 0;
    while (kc < nw && !hit) {
        if (N[kc] == 0) { ++kc; continue; }
        if (N[kc] <= kFlatMax) {
            // small classes: one pattern per lane over the flattened index of this
            // and the following small classes
            uint64_t flat_end = class_start + N[kc];
            int kend = kc + 1;
            while (kend < nw && N[kend] <= kFlatMax) flat_end += N[kend++];
            for (uint64_t pos = class_start; pos < flat_end && !hit; pos += 64) {
                const uint64_t idx = pos + (uint64_t)lane;
                bool cand = false;
                int w = 0;
                if (idx < flat_end) {
                    uint64_t cs = class_start;
                    int k = kc;
                    while (idx >= cs + N[k]) { cs += N[k]; ++k; }
                    w = wl[k];
                    unrank_lds((uint32_t)(idx - cs), tam, w, L.ncr, c);   // flat classes have tam <= 128
                    uint32_t x = t0;
#pragma unroll
                    for (int l = 0; l < kMaxW; ++l) if (l < w) x ^= L.sigA[c[l]];
                    if (x == 0u) {
                        int cols[kMaxW];
#pragma unroll
                        for (int l = 0; l < kMaxW; ++l) cols[l] = (l < w) ? L.A[c[l]] : 0;
                        cand = exact_match(code, s, cols, w);
                    }
                }
                const uint64_t b = ballot(cand);
                count_work(call.work, 2, idx < flat_end ? 1u : 0u);
                if (b) {
                    hit = true;
                    hit_lane = __ffsll((long long)b) - 1;
                    hit_idx = pos + (uint64_t)hit_lane;
                    hit_w = bcast_i(w, hit_lane);
                }
            }
            if (!hit) { class_start = flat_end; kc = kend; }
        } else {
            const int w = wl[kc];
            const bool pairs = use_pairs(code, call, w, tam);
            if (w >= 2 && call.split_threshold && rest_work(code, call, wl_packed >> (4 * kc), tam) > call.split_threshold) {
                // this class and the following ones of weight >= 2 as one shared task:
                // no barrier between the classes, and the task's lowest matching
                // pattern index is the first hit over them in list order
                int nc = 0, kend = kc;
                unsigned wpack = 0, ppack = 0;
                bool need_ht = false, need_pairs = false;
                while (kend < nw && nc < kTaskClasses && wl[kend] >= 2) {
                    if (N[kend] > 0) {
                        const bool pk = use_pairs(code, call, wl[kend], tam);
                        wpack |= (unsigned)wl[kend] << (4 * nc);
                        ppack |= (pk ? 1u : 0u) << nc;
                        need_pairs = need_pairs || pk;
                        need_ht = need_ht || !pk;
                        ++nc;
                    }
                    ++kend;
                }
                unsigned long long best;
                bool shared = false;
                if (spec >= 0) {
                    // the run was published before this iteration's search; it is
                    // used when it is this run (it always is: the same list and walk)
                    TaskSlot *T = &call.tasks[spec];
                    const bool same = uni((unsigned)ld_relaxed(&T->nclass)) == (unsigned)nc &&
                                      uni((unsigned)ld_relaxed(&T->wpack)) == wpack &&
                                      uni((unsigned)ld_relaxed(&T->ppack)) == ppack &&
                                      uni((unsigned)ld_relaxed(&T->tam)) == (unsigned)tam;
                    if (same) {
                        // tables only when chunks are left to claim (other waves may have
                        // claimed them all; the claims never come back)
                        const unsigned long long snap = uni64(ld_relaxed(&T->ctl));
                        const bool left = ctl_next(snap) < ctl_nch(snap);
                        if (left && need_pairs && !have_pairs) { build_pairs<false>(L, L.sigA, tam); have_pairs = true; }
                        if (left && need_ht && !have_table) { hmask = build_ht(L.ht, L.sigA, tam); have_table = true; }
                        work_shared<false>(code, call, spec, s, t0, tam, nc, wpack, ppack, L.A, L.sigA, nullptr, L.ht,
                                    hmask, L.pkey, L.pval, L.phead, L.pbm, best);
                        shared = true;
                    } else {
                        cancel_shared(call, spec);
                    }
                    spec = -1;
                    spec_done = 2;
                }
                if (!shared) {
                    // the list's tables, for this wave's own chunks of the run
                    if (need_pairs && !have_pairs) { build_pairs<false>(L, L.sigA, tam); have_pairs = true; }
                    if (need_ht && !have_table) { hmask = build_ht(L.ht, L.sigA, tam); have_table = true; }
                    shared = run_shared<false>(code, call, shot, s, t0, tam, nc, wpack, ppack, L.A, L.sigA, nullptr, L.ht,
                                        hmask, L.pkey, L.pval, L.phead, L.pbm, best);
                }
                if (shared) {
                    uint64_t span = 0;
                    for (int k = kc; k < kend; ++k) span += N[k];
                    if (best != ~0ull) {
                        // class of the hit, and its rank within the class
                        uint64_t r = best;
                        int k = kc;
                        while (r >= N[k]) { r -= N[k]; ++k; }
                        hit = true;
                        hit_lane = 0;
                        hit_w = wl[k];
                        hit_idx = class_start + best;
                        if (tam < kNcrLdsRows) unrank_dyn(r, tam, hit_w, ncr_lds(), c);
                        else unrank_dyn(r, tam, hit_w, code.nCr, c);
                    } else {
                        class_start += span;
                        kc = kend;
                    }
                    continue;
                }
            }
            if (pairs) {
                if (!have_pairs) { build_pairs<false>(L, L.sigA, tam); have_pairs = true; }
            } else if (!have_table && w > 1) {
                hmask = build_ht(L.ht, L.sigA, tam);
                have_table = true;
            }
            const uint64_t r =
                class_first_hit(code, call, shot, s, t0, w, tam, N[kc], L.A, L.sigA, L.ht, hmask, pairs, L.pkey,
                                L.pval, L.phead, L.pbm);
            if (r != ~0ull) {
                hit = true;
                hit_lane = 0;
                hit_w = w;
                hit_idx = class_start + r;
                if (tam < kNcrLdsRows) unrank_dyn(r, tam, w, ncr_lds(), c);   // same values in every lane
                else unrank_dyn(r, tam, w, code.nCr, c);
            } else {
                class_start += N[kc];
                ++kc;
            }
        }
    }
    if (hit) {
        fw = hit_w;
#pragma unroll
        for (int l = 0; l < kMaxW; ++l) {
            const int cl = bcast_i(c[l], hit_lane);
            win[l] = (l < hit_w) ? L.A[cl] : -1;
        }
        count += hit_idx + 1u;
        wops += weighted_prefix(N, wl, nw, hit_idx + 1u);
        return 1 | spec_done;
    }
    count += total;
    wops += weighted_prefix(N, wl, nw, total);
    return spec_done;
}

// -------------------------------------------------------------------- ML search
// grand_decoder.cu:590-666 over the LLR-sorted active list P: weights 1..max_w
// in order, stop before weight w when a best exists and prefix[w] >= best, keep
// the first strictly cheapest match (fp32 costs summed in pattern order).
__device__ bool ml_wave(const DevCode &code, const DevCall &call, int shot, const uint32_t LDS *s, uint32_t t0, int tam,
                        const WaveLds &L, int *win, int &fw, unsigned long long &count, unsigned long long &wops) {
    const int lane = lane_id();
    sort_by_llr(code, L, tam);
    const unsigned hmask = build_ht(L.ht, L.sigP, tam);
    bool have_pairs = false;
    const int max_w = call.W < tam ? call.W : tam;
    float prefix[kMaxW + 1];
    prefix[0] = 0.0f;
#pragma unroll
    for (int w = 1; w <= kMaxW; ++w) prefix[w] = (w <= max_w) ? __fadd_rn(prefix[w - 1], L.llrP[w - 1]) : 0.0f;

    bool have = false;
    float best = 0.0f;
    for (int w = 1; w <= max_w; ++w) {
        if (have && prefix[w] >= best) break;
        const uint64_t Nw = ncr_w(code, tam, w);
        count += Nw;
        wops += Nw * (uint64_t)(w + 1);
        bool chit = false;
        float cbest = 0.0f;
        uint64_t crank = 0;
        if (w == 1) {
            const PrefHit h = weight_one<1>(code, s, t0, tam, L.P, L.sigP, L.llrP);
            count_work(call.work, 2, lane == 0 ? (uint32_t)tam : 0u);
            if (h.j >= 0) { chit = true; cbest = h.cost; crank = h.rank; }
        } else {
            // deep classes by (w-2)-stems over the pair table of the sorted list
            // (built once per list), the others by (w-1)-prefixes
            const bool pairs = use_pairs(code, call, w, tam);
            if (pairs && !have_pairs) { build_pairs<true>(L, L.sigP, tam); have_pairs = true; }
            const uint64_t NP = class_units(code, tam, w, pairs);
            unsigned long long shared_best;
            if (call.split_threshold && NP * (pairs ? (uint64_t)kStemCost : 1u) > call.split_threshold &&
                Nw < (1ull << 32) &&
                run_shared<true>(code, call, shot, s, t0, tam, 1, (unsigned)w, pairs ? 1u : 0u, L.P, L.sigP, L.llrP,
                           L.ht, hmask, L.pkey, L.pval, L.phead, L.pbm, shared_best)) {
                if (shared_best != ~0ull) {
                    chit = true;
                    cbest = key_to_float((uint32_t)(shared_best >> 32));
                    crank = shared_best & 0xFFFFFFFFull;
                }
            } else {
                uint64_t pos = 0;
                while (pos < NP) {
                    const uint64_t rem = NP - pos;
                    const uint64_t r64 = (rem + 63u) / 64u;
                    const int R = (int)(r64 < (uint64_t)call.rmax ? r64 : (uint64_t)call.rmax);
                    const uint64_t q0 = pos + (uint64_t)lane * (uint64_t)R;
                    const int len = q0 >= NP ? 0 : (int)((NP - q0) < (uint64_t)R ? (NP - q0) : (uint64_t)R);
                    const PrefHit h = pairs ? scan_stems_ml_w(w, code, s, t0, q0, len, tam, L.P, L.sigP, L.llrP, L.pkey,
                                                              L.pval, L.phead, L.pbm)
                                            : scan_prefix_w<1>(w, code, s, t0, q0, len, tam, L.P, L.sigP, L.llrP, L.ht,
                                                               hmask);
                    count_work(call.work, pairs ? 1 : 0, walked(h, len, true));
                    const uint32_t key = (h.j >= 0) ? order_key(h.cost) : 0xFFFFFFFFu;
                    const uint32_t kmin = wave_min(key);
                    const uint64_t b = ballot(h.j >= 0 && key == kmin);
                    if (b) {
                        const int bl = __ffsll((long long)b) - 1;
                        const float cb = bcast_f(h.cost, bl);
                        if (!chit || cb < cbest) {   // chunks in rank order: strict keeps the earliest
                            chit = true;
                            cbest = cb;
                            crank = (uint64_t)__shfl((long long)h.rank, bl, 64);
                        }
                    }
                    pos += 64ull * (uint64_t)R;
                }
            }
        }
        if (chit && (!have || cbest < best)) {
            int c[kMaxW] = {0, 0, 0, 0, 0, 0};
            if (tam < kNcrLdsRows) unrank_dyn(crank, tam, w, ncr_lds(), c);
            else unrank_dyn(crank, tam, w, code.nCr, c);
            have = true;
            best = cbest;
            fw = w;
#pragma unroll
            for (int l = 0; l < kMaxW; ++l) win[l] = (l < w) ? L.P[c[l]] : -1;
        }
    }
    return have;
}

// --------------------------------------------------------- per-iteration logic
// Weight list of one iteration: 1..W (uniform) or the phase lists of the mask
// chosen for tam (grand_decoder.cu:754-786, :1001-1003).
__device__ __forceinline__ int weight_list(const DevCall &call, int tam, int *wl) {
    int nw = 0;
    if (call.kind == kUniform) {
        for (int w = 1; w <= call.W; ++w) wl[nw++] = w;
        return nw;
    }
    int mi = call.n_masks - 1;
    for (int i = 0; i < call.n_masks; ++i) if (tam <= call.mask_tam_upper[i]) { mi = i; break; }
    for (int k = 0; k < call.num_phases; ++k) {
        const int ph = call.mask_order[mi][k];
        if (ph == 0) { for (int w = 1; w <= 5 && w <= tam && w <= kMaxW; ++w) wl[nw++] = w; }
        else if (ph == 1) { if (6 <= tam) wl[nw++] = 6; }
    }
    return nw;
}

// Branch of the reference driver for an iteration with tam > 0:
// 1 = "GPU" branch, -1 = "CPU" branch, 0 = no search (grand_decoder.cu:1296, :1327).
__device__ __forceinline__ int branch_of(const DevCall &call, int tam) {
    if ((call.mode == 1 && tam > call.thr_h && tam <= call.max_size) || call.mode == 0) return 1;
    if ((call.mode == 1 && tam <= call.thr_h) || call.mode == -1) return -1;
    return 0;
}

// ----------------------------------------------------------------- result write
// The 72-byte record is written as 18 dwords by 18 lanes.
__device__ void write_result(GrandDecoderResult *dst, int found, int fw, int ow, int deg, int eq, int tam,
                             uint64_t count, int hib, int it, double t_sum, double t_search, double t_find) {
    const int lane = lane_id();
    uint32_t v = 0;
    const uint64_t tu = (uint64_t)__double_as_longlong(t_sum);
    const uint64_t ts = (uint64_t)__double_as_longlong(t_search);
    const uint64_t tf = (uint64_t)__double_as_longlong(t_find);
    switch (lane) {
        case 0: v = (uint32_t)found; break;
        case 1: v = (uint32_t)fw; break;
        case 2: v = (uint32_t)ow; break;
        case 3: v = (uint32_t)deg; break;
        case 4: v = (uint32_t)eq; break;
        case 5: v = (uint32_t)tam; break;
        case 6: v = (uint32_t)count; break;
        case 7: v = (uint32_t)(count >> 32); break;
        case 8: v = (uint32_t)hib; break;
        case 9: v = (uint32_t)it; break;
        case 10: v = (uint32_t)tu; break;
        case 11: v = (uint32_t)(tu >> 32); break;
        case 12: v = (uint32_t)ts; break;
        case 13: v = (uint32_t)(ts >> 32); break;
        case 16: v = (uint32_t)tf; break;
        case 17: v = (uint32_t)(tf >> 32); break;
        default: v = 0; break;
    }
    if (lane < 18) reinterpret_cast<uint32_t *>(dst)[lane] = v;
}

// e ^ e_hat (words dl0, dl1 of lanes lane and lane + 64) is in the row space of
// the other matrix iff it equals the XOR of the RREF rows whose pivot bit it
// has set: in reduced echelon form a pivot column is zero outside its own row,
// so adding rows never changes which other pivots are set.  The residue is
// sparse; its set bits are looked up in pivrow and the selected rows XORed
// (independent loads, no chain through the rank).
__device__ __attribute__((noinline)) bool rref_degenerate(const DevCode &code, uint32_t dl0, uint32_t dl1) {
    const int lane = lane_id();
    const int M32 = code.M32;
    uint32_t acc[2] = {0u, 0u};
    uint32_t rest[2] = {dl0, dl1};
    for (;;) {
        int c = -1;
        if (rest[0]) { c = lane * 32 + __ffs(rest[0]) - 1; rest[0] &= rest[0] - 1u; }
        else if (rest[1]) { c = (lane + 64) * 32 + __ffs(rest[1]) - 1; rest[1] &= rest[1] - 1u; }
        if (ballot(c >= 0) == 0ull) break;
        const int r = (c >= 0) ? (code.lds_post == 1 ? code_lds_pivrow(code)[c] : code.pivrow[c]) : -1;
        uint64_t sel = ballot(r >= 0);
        while (sel) {
            const int src = __ffsll((long long)sel) - 1;
            sel &= sel - 1;
            const int rr = bcast_i(r, src);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int w2 = lane + 64 * q;
                if (w2 < M32)
                    acc[q] ^= code.lds_post == 1 ? code_lds_R(code)[(size_t)rr * M32 + w2]
                                                 : code.R[(size_t)rr * M32 + w2];
            }
        }
    }
    return ballot(((dl0 ^ acc[0]) | (dl1 ^ acc[1])) != 0u) == 0ull;
}

// Post-processing and result record of a searched shot (grand_decoder.cu:1359-1370
// for the CSS family, :1623-1648 for the DEM family).  The shot's packed errors
// (cc) or observed logicals (cln) come from its staged copy in LDS (`slot`, laid
// out as DevCall::stage_words describes) or, with slot null, from global memory.
__device__ __attribute__((noinline)) void finalize_shot_v(const DevCode &code, const DevCall &call, int shot, int found,
                                                          int fw, int iteration, int tam, int hibrido, int w0, int w1,
                                                          int w2, int w3, int w4, int w5, int w6, int w7,
                                                          unsigned long long count, unsigned long long wops,
                                                          unsigned long long t0, unsigned long long t1,
                                                          unsigned long long ta, const uint32_t LDS *slot) {
    ShotOutcome o;
    o.found = found; o.fw = fw; o.iteration = iteration; o.tam = tam; o.hibrido = hibrido;
    o.win[0] = w0; o.win[1] = w1; o.win[2] = w2; o.win[3] = w3; o.win[4] = w4; o.win[5] = w5; o.win[6] = w6;
    o.win[7] = w7;
    o.count = count; o.wops = wops; o.t0 = t0; o.t1 = t1; o.ta = ta;
    const int lane = lane_id();
    const int M32 = code.M32;
    uint32_t dw[2] = {0u, 0u};   // decoded pattern words lane + 64 q, q < 2 (m <= 4096)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int wd = lane + 64 * q;
        uint32_t v = 0;
        if (o.found) {
#pragma unroll
            for (int l = 0; l < kMaxWin; ++l)
                if (l < o.fw && (o.win[l] >> 5) == wd) v |= 1u << (o.win[l] & 31);
        }
        dw[q] = v;
        if (wd < M32) call.dec[(size_t)shot * M32 + wd] = v;
    }
    int ow = 0, eq = 0, deg = 0;
    if (call.family == kCC && call.input_errors && slot) {
        // staged shot: original weight and degeneracy word from the triage
        const uint32_t LDS *ews = slot + 2 + code.D + 2;
        ow = (int)uni(slot[2 + code.D]);
        if (o.found) {
            // e_hat == e iff both have the same weight and every decoded column is set in e
            bool same = (ow == o.fw);
            uint32_t gh = 0;
#pragma unroll
            for (int l = 0; l < kMaxWin; ++l) {
                if (l < o.fw) {
                    const int c = o.win[l];
                    same = same && ((ews[c >> 5] >> (c & 31)) & 1u);
                    if (code.dg) gh ^= (code.lds_post == 2) ? code_lds_dg(code)[c] : code.dg[c];
                }
            }
            eq = same ? 1 : 0;
            if (code.dg) {
                deg = (uni(slot[2 + code.D + 1]) == uni(gh)) ? 1 : 0;
            } else {
                uint32_t ew[2] = {0u, 0u};
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int wd = lane + 64 * q;
                    if (wd < M32) ew[q] = ews[wd];
                }
                deg = rref_degenerate(code, ew[0] ^ dw[0], ew[1] ^ dw[1]) ? 1 : 0;
            }
        }
    } else if (call.family == kCC && call.input_errors) {
        uint32_t ew[2] = {0u, 0u};
        int pc = 0;
        bool same = true;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int wd = lane + 64 * q;
            if (wd < M32) ew[q] = call.e_packed[(size_t)shot * M32 + wd];
            pc += __popc(ew[q]);
            same = same && (ew[q] == dw[q]);
        }
        // original weight = popcount(e) summed across the wave
        for (int o2 = 32; o2 >= 1; o2 >>= 1) pc += __shfl_xor(pc, o2, 64);
        ow = pc;
        if (o.found) {
            eq = ballot(!same) == 0ull;
            deg = rref_degenerate(code, ew[0] ^ dw[0], ew[1] ^ dw[1]) ? 1 : 0;
        }
    } else if (call.family == kCLN) {
        const int NL32 = code.NL32;
        const int nlw = (code.NL + 31) / 32;     // words that hold logicals (0 when n_logicos = 0)
        bool same = true;
        for (int wd = lane; wd < nlw; wd += 64) {
            uint32_t le = 0;
            if (o.found) {
#pragma unroll
                for (int l = 0; l < kMaxWin; ++l)
                    if (l < o.fw) le ^= code.Lp[(size_t)o.win[l] * NL32 + wd];
            }
            call.declog[(size_t)shot * NL32 + wd] = le;
            const uint32_t lo = slot ? slot[2 + code.D + wd] : call.lg_packed[(size_t)shot * NL32 + wd];
            same = same && (le == lo);
        }
        if (o.found) deg = ballot(!same) == 0ull;
    }
    // time_sum_ms: the scores and the first active list; total_time_search_ms: the whole search
    const double t_sum = o.ta ? (double)(o.ta - o.t0) * call.ms_per_tick : 0.0;
    const double t_search = (double)(o.t1 - o.t0) * call.ms_per_tick;
    write_result(&call.results[shot], o.found, o.found ? o.fw : 0, ow, deg, eq, o.tam, o.count, o.hibrido,
                 o.iteration, t_sum, t_search, o.found ? t_search : 0.0);
    if (call.stats && lane == 0) {
        call.stats[3 * (size_t)shot] = o.t0;
        call.stats[3 * (size_t)shot + 1] = o.t1;
        call.stats[3 * (size_t)shot + 2] = o.wops;
    }
}

// The outcome goes to the out-of-line finalizer as scalars (registers), so the
// shot's outcome never needs stack memory whatever the inliner decides.
static_assert(kMaxWin == 8, "finalize_shot_v takes eight decoded columns");
__device__ __forceinline__ void finalize_shot(const DevCode &code, const DevCall &call, int shot, const ShotOutcome &o,
                                              const uint32_t LDS *slot) {
    finalize_shot_v(code, call, shot, o.found, o.fw, o.iteration, o.tam, o.hibrido, o.win[0], o.win[1], o.win[2],
                    o.win[3], o.win[4], o.win[5], o.win[6], o.win[7], o.count, o.wops, o.t0, o.t1, o.ta, slot);
}

// ------------------------------------------------------------ learned pass
// Learned search order (no counterpart in the reference).  The positions of the
// first active list are sorted by descending score (ties: list order) into P,
// and patterns of weight 1..lmaxw over P are tried stage by stage: stage k
// holds the patterns whose last position in P is k, by weight, each weight
// lexicographic over P[0..k).  So a pattern comes as soon as its worst-scored
// position does.  Every candidate is checked exactly against the syndrome; the
// pass stops at the first match or after lbudget patterns.  Returns the match
// (count = patterns tried up to and including it) or false with count = the
// patterns tried.

// First match among the w'-subsets S of P[0..k) with H(S) = s2 (signature t2),
// rank < lim in lexicographic order; returns its rank or ~0.
__device__ uint64_t learned_class(const DevCode &code, const DevCall &call, const uint32_t LDS *s2, uint32_t t2,
                                  int wp, int k, uint64_t lim, const WaveLds &L, unsigned hmask) {
    const int lane = lane_id();
    if (wp == 1) {
        for (int i0 = 0; i0 < k && (uint64_t)i0 < lim; i0 += 64) {
            const int i = i0 + lane;
            bool hit = false;
            if (i < k && (uint64_t)i < lim && L.sigP[i] == t2) {
                int cols[1] = {L.P[i]};
                hit = exact_match<1>(code, s2, cols, 1);
            }
            const uint64_t b = ballot(hit);
            if (b) return (uint64_t)(i0 + __ffsll((long long)b) - 1);
        }
        return ~0ull;
    }
    const uint64_t NP = ncr_w(code, k - 1, wp - 1);
    uint64_t pos = 0;
    while (pos < NP) {
        const uint64_t rem = NP - pos;
        const uint64_t r64 = (rem + 63u) / 64u;
        const int R = (int)(r64 < (uint64_t)call.rmax ? r64 : (uint64_t)call.rmax);
        const uint64_t q0 = pos + (uint64_t)lane * (uint64_t)R;
        const int len = q0 >= NP ? 0 : (int)((NP - q0) < (uint64_t)R ? (NP - q0) : (uint64_t)R);
        const PrefHit h = scan_prefix_learned(wp, code, s2, t2, q0, len, k, L.P, L.sigP, L.ht, hmask);
        const uint64_t b = ballot(h.j >= 0);
        if (b) {
            const uint64_t r = (uint64_t)__shfl((long long)h.rank, __ffsll((long long)b) - 1, 64);
            return r < lim ? r : ~0ull;   // later prefixes only hold higher ranks
        }
        pos += 64ull * (uint64_t)R;
    }
    return ~0ull;
}

__device__ bool learned_pass(const DevCode &code, const DevCall &call, int shot, const uint32_t LDS *s, uint32_t t0,
                             const WaveLds &L, ShotOutcome &o) {
    const int lane = lane_id();
    const int thr = code.max_soma;
    int tam = build_active(code, s, thr, L, call.cap);
    o.ta = now_ticks();
    if (call.family == kCLN && tam > call.max_size) tam = 0;      // the first iteration has no list
    if (tam <= 0 || tam > call.cap || (!call.lper_column && tam > call.lstride) || branch_of(call, tam) == 0)
        return false;
    const float *sc = call.lscores + (size_t)shot * call.lstride;
    for (int i = lane; i < tam; i += 64) {
        const float v = sc[call.lper_column ? L.A[i] : i];
        L.llrA[i] = (v == v) ? -v : __builtin_inff();             // descending score; NaN last
    }
    __syncthreads();
    sort_by_key(L, tam);
    const unsigned hmask = build_ht(L.ht, L.sigP, tam);
    const uint64_t budget = call.lbudget;
    const int maxw = call.lmaxw < kMaxWin ? call.lmaxw : kMaxWin;
    uint64_t G = 0;
    bool stop = false;
    for (int k = 0; k < tam && !stop; ++k) {
        const int ck = L.P[k];
        if (lane < kMaxWords) L.sw2[lane] = (lane < code.D) ? (s[lane] ^ col_word(code, ck, lane)) : 0u;
        __syncthreads();
        const uint32_t t2 = t0 ^ L.sigP[k];
        for (int wp = 0; wp < maxw; ++wp) {
            const uint64_t N = (wp == 0) ? 1ull : ncr_w(code, k, wp);
            if (N == 0) break;                      // larger weights do not fit in P[0..k) either
            if (G >= budget) { stop = true; break; }
            const uint64_t lim = (budget - G) < N ? (budget - G) : N;
            uint64_t r = ~0ull;
            if (wp == 0) {
                bool z = true;
                if (t2 == 0u)
                    for (int d = 0; d < code.D; ++d) z = z && (L.sw2[d] == 0u);
                r = (t2 == 0u && z) ? 0ull : ~0ull;
            } else {
                r = learned_class(code, call, L.sw2, t2, wp, k, lim, L, hmask);
            }
            if (r != ~0ull) {
                int c[kMaxWin] = {0, 0, 0, 0, 0, 0, 0, 0};
                if (wp > 0) {
                    if (k < kNcrLdsRows) unrank_dyn<kMaxWin>(r, k, wp, ncr_lds(), c);
                    else unrank_dyn<kMaxWin>(r, k, wp, code.nCr, c);
                }
                o.fw = wp + 1;
#pragma unroll
                for (int l = 0; l < kMaxWin; ++l) o.win[l] = (l < wp) ? L.P[c[l]] : ((l == wp) ? ck : -1);
                o.count += G + r + 1;
                o.wops += (r + 1) * (uint64_t)(wp + 2);   // the earlier classes' share is already in wops
                o.found = 1;
                o.iteration = 1;
                o.tam = tam;
                o.hibrido = (branch_of(call, tam) > 0) ? 0 : -1;
                return true;
            }
            G += lim;
            o.wops += lim * (uint64_t)(wp + 2);
            if (lim < N) { stop = true; break; }
        }
    }
    o.count += G;
    return false;
}

// The whole per-shot pipeline in one wave (grand_decoder.cu:1284-1343 and the
// equivalent loops of the other four drivers).
// One claimed shot, staged in LDS by stage_claim: slot = [shot, signature,
// syndrome words (D), then original weight, degeneracy word and error words
// (M32) for cc with errors, or the observed logicals (NL32) for cln].
// Publishes the run of classes that the third iteration (threshold
// max_soma - 2) would share, before the second iteration is searched: a shot
// whose second list is long (kSpecTam positions or more) is likely to need the
// third, whose run then proceeds on the other waves while this wave searches
// the second list.  When the third iteration is searched it works on this task
// instead of publishing one; when it is not, the task is withdrawn (a result of
// an earlier iteration is never changed by it).  Returns the slot or -1.
// Overwrites the wave's list (A, sigA).
// Out of line: the wave's LDS layout is carved again from the call (the
// search kernel's own layout).
__device__ uint64_t lane_weight_list(const DevCall &call, int tam);
constexpr int kSpecTam = 16;
__device__ __attribute__((noinline)) int publish_iteration3(const DevCode &code, const DevCall &call, int shot,
                                                             const uint32_t LDS *s, uint32_t t0) {
    const WaveLds L = carve_lds(call.cap, call.kind == kML || call.lscores != nullptr, code.lds_words,
                                call.claim_batch * call.stage_words, call.pair_cap);
    const int tam3 = build_active_into(code, s, code.max_soma - 2, L.A, L.sigA, call.cap);
    if (tam3 <= 0 || tam3 > call.cap || branch_of(call, tam3) == 0) return -1;
    int nc;
    unsigned wpack, ppack;
    if (!find_split(code, call, tam3, lane_weight_list(call, tam3), nc, wpack, ppack)) return -1;
    return publish_shared(code, call, shot, s, t0, tam3, false, nc, wpack, ppack, L.A, L.sigA, nullptr);
}

template <bool ML>
__device__ __forceinline__ void decode_shot_wave(const DevCode &code, const DevCall &call, const uint32_t LDS *slot,
                                                 const WaveLds &L) {
    ShotOutcome o;
    o.t0 = now_ticks();
    const int shot = (int)uni(slot[0]);
    if (lane_id() < kMaxWords) L.sw[lane_id()] = (lane_id() < code.D) ? slot[2 + lane_id()] : 0u;
    __syncthreads();
    const uint32_t LDS *s = L.sw;
    const uint32_t t0 = uni(slot[1]);

    o.found = 0; o.fw = 0; o.count = 0; o.wops = 0; o.hibrido = -2; o.ta = 0;
#pragma unroll
    for (int l = 0; l < kMaxWin; ++l) o.win[l] = -1;
    int it = 0, tam = 0, tam_old = 0;
    int spec = -1;                                 // the third iteration's run, published early (slot)
    const bool spec_ok = !ML && call.family == kCC && call.lscores == nullptr && call.split_threshold &&
                         code.max_soma >= 3;
    const bool learned_hit = !ML && call.lscores && learned_pass(code, call, shot, s, t0, L, o);
    while (!learned_hit && it < code.max_soma && it < 3) {
        const int thr = code.max_soma - it;
        ++it;
        tam = build_active(code, s, thr, L, call.cap);
        if (it == 1 && o.ta == 0) o.ta = now_ticks();
        if (call.family == kCLN) { if (tam > call.max_size) tam = tam_old; else tam_old = tam; }
        if (tam <= 0) continue;
        if (tam > call.cap) {                      // beyond the supported active-set size
            if (lane_id() == 0) atomicOr(&call.counters[2], kErrTamOverflow);
            break;
        }
        const int br = branch_of(call, tam);
        if (br == 0) continue;
        if (spec_ok && it == 2 && tam >= kSpecTam) {
            // publish_iteration3 builds the third list in A / sigA before it
            // decides whether to publish, so the second list is rebuilt whatever
            // it returned
            spec = publish_iteration3(code, call, shot, s, t0);
            tam = build_active(code, s, thr, L, call.cap);
        }
        o.hibrido = (br > 0) ? (o.hibrido == -1 ? 1 : 0) : -1;
        bool hit;
        if constexpr (ML) {
            hit = ml_wave(code, call, shot, s, t0, tam, L, o.win, o.fw, o.count, o.wops);
        } else {
            int wl[2 * kMaxW];
            const int nw = weight_list(call, tam, wl);
            const int r = first_hit_wave(code, call, shot, s, t0, tam, wl, nw, L, o.win, o.fw, o.count, o.wops,
                                         it == 3 ? spec : -1);
            hit = (r & 1) != 0;
            if (r & 2) spec = -1;
        }
        if (hit) { o.found = 1; break; }
    }
    if (spec >= 0) cancel_shared(call, spec);       // the third iteration was not searched
    if (!learned_hit) {
        o.iteration = it;
        o.tam = tam;
    }
    // the decoded columns are reported sorted (the ML and learned winners are in score order)
    if (o.found) {
#pragma unroll
        for (int a = 0; a < kMaxWin; ++a)
#pragma unroll
            for (int b = 0; b < kMaxWin - 1 - a; ++b)
                if (b + 1 < o.fw && o.win[b] > o.win[b + 1]) { const int t = o.win[b]; o.win[b] = o.win[b + 1]; o.win[b + 1] = t; }
    }
    o.t1 = now_ticks();
    finalize_shot(code, call, shot, o, slot);
}

// First worklist slot of partition k and its number of slots (64 per group) for
// G triage groups dealt over P partitions (group g to partition g % P).
__device__ __forceinline__ unsigned wl_part_start(int k, int G, int P) {
    const int q = G / P, r = G % P;
    return 64u * (unsigned)(k * q + (k < r ? k : r));
}
__device__ __forceinline__ unsigned wl_part_slots(int k, int G, int P) {
    return 64u * (unsigned)(G / P + (k < G % P ? 1 : 0));
}

// Worklist slot of position p: positions below `heavy` are the front entries of
// all partitions in partition order, the others their back entries.  Lane j < P
// holds the exclusive prefix sums wfx (fronts) and wbx (backs) of partition j;
// the partition of p is the last whose prefix is <= p (empty partitions share
// the next one's prefix).  Call with every lane active: it reads other lanes.
__device__ __forceinline__ unsigned wl_slot(unsigned p, unsigned heavy, unsigned wfx, unsigned wbx, int P, int G) {
    const bool fr = p < heavy;
    const unsigned q = fr ? p : p - heavy;
    int k = 0;
    unsigned base = 0u;
    for (int j = 1; j < P; ++j) {
        const unsigned fj = (unsigned)__builtin_amdgcn_readlane((int)wfx, j);
        const unsigned bj = (unsigned)__builtin_amdgcn_readlane((int)wbx, j);
        const unsigned pj = fr ? fj : bj;
        if (pj <= q) { k = j; base = pj; }
    }
    const unsigned r0 = wl_part_start(k, G, P);
    return fr ? r0 + (q - base) : r0 + wl_part_slots(k, G, P) - 1u - (q - base);
}

// Stages the shots of one claim (worklist positions p0, p0 + 8, ...; nb of them)
// in LDS: the worklist entries first, then every shot's words in one round of
// loads, so the shots decode without waiting on global memory between them.
__device__ void stage_claim(const DevCode &code, const DevCall &call, unsigned p0, int nb, unsigned heavy,
                            unsigned wfx, unsigned wbx, const WaveLds &L) {
    const int lane = lane_id();
    const int SW = call.stage_words;
    const int per = SW - 1;                        // loaded words per shot: signature, syndrome, errors or logicals
    int myshot = 0;
    {
        const unsigned p = p0 + (unsigned)kClaimWords * (unsigned)lane;
        const unsigned slot = wl_slot(p, heavy, wfx, wbx, call.wl_parts, (call.num_shots + 63) / 64);
        if (lane < nb) myshot = call.worklist[slot];
    }
    if (lane < nb) L.stage[lane * SW] = (uint32_t)myshot;
    const int E = call.input_errors ? 2 + code.M32 : 0;
    const int tot = nb * per;
    for (int k0 = 0; k0 < tot; k0 += 256) {
        uint32_t v[4];
        int dst[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + lane + 64 * q;
            const int j = (k < tot) ? k / per : 0;
            const int f = k - j * per;
            const int shot = __shfl(myshot, j, 64);
            uint32_t x = 0u;
            if (k < tot) {
                if (f == 0) x = call.s_sig[shot];
                else if (f <= code.D) x = call.s_packed[(size_t)shot * code.D + (f - 1)];
                else if (f <= code.D + 2 && E > 0) x = call.e_aux[2 * (size_t)shot + (f - 1 - code.D)];
                else if (f <= code.D + E) x = call.e_packed[(size_t)shot * code.M32 + (f - 3 - code.D)];
                else x = call.lg_packed[(size_t)shot * code.NL32 + (f - 1 - code.D - E)];
            }
            v[q] = x;
            dst[q] = (k < tot) ? j * SW + 1 + f : -1;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (dst[q] >= 0) L.stage[dst[q]] = v[q];
    }
    __syncthreads();
}

// ================================================================== kernels

// Record of a shot whose syndrome is zero, written by its own lane
// (grand_decoder.cu:1223-1241, :1487-1499).
__device__ __forceinline__ void write_zero_result(GrandDecoderResult *dst, int ow) {
    uint32_t *p = reinterpret_cast<uint32_t *>(dst);
#pragma unroll
    for (int i = 0; i < 18; ++i) p[i] = 0u;
    p[2] = (uint32_t)ow;            // original_weight
    p[8] = (uint32_t)-2;            // hibrido: no search
}

// ------------------------------------------------------------ lane search
// A shot searched by its own lane in the triage kernel: the first iteration of
// decode_shot_wave (threshold max_soma), and for the CSS family the second
// (threshold max_soma - 1) after a first searched whole without a hit, when
// the iteration's active list has at most kLaneMax positions and the leading
// weight classes of its list are of weight <= 3 with at most kLaneBudget
// patterns in all.  At threshold max_soma - j a column is active iff at most
// j of its checks lie outside supp(s) (score = |supp(h_c) & supp(s)| +
// max_soma - colw[c], grand_decoder.cu:406-416, :495-504): its count of s's
// checks is within j of colw[c].  Classes are
// enumerated in the order of the wave path (lexicographic within a class,
// classes in list order), each candidate filtered by its signature and checked
// exactly, so the hit, its count, its iteration and its columns are the wave
// path's.  Returns false, having written nothing, when the shot needs the
// search kernel (list too long, a class too large, a cap of the DEM family, or
// no match in the iterations searched here).
// Weight list of the first iteration for an active list of `tam` positions:
// the weights of decode_shot_wave's list in order, 4 bits each, first in the
// low bits, without the slots whose class is empty (weight above tam); 0 ends
// the list (at most 15 slots: three phases of five).
__device__ uint64_t lane_weight_list(const DevCall &call, int tam) {
    if (call.kind != kUniform && call.n_masks <= 0) return 0ull;
    int mi = call.n_masks - 1;
    if (call.kind != kUniform)
        for (int i = 0; i < call.n_masks; ++i) if (tam <= call.mask_tam_upper[i]) { mi = i; break; }
    const int nslots = (call.kind == kUniform) ? call.W : call.num_phases * 5;
    uint64_t wl = 0ull;
    int n = 0;
    for (int q = 0; q < nslots; ++q) {
        int w;
        if (call.kind == kUniform) {
            w = q + 1;
        } else {
            const int ph = call.mask_order[mi][q / 5], j = q % 5;
            w = (ph == 0) ? j + 1 : (ph == 1 && j == 0) ? 6 : 0;
        }
        if (w == 0 || w > tam) continue;
        wl |= (uint64_t)w << (4 * n);
        ++n;
    }
    return wl;
}

// The code tables of the triage and lane kernels (the image DevCode::timg):
// column words (word-major, hcw[d * m + c]), signatures, degeneracy words
// (has_dg), the rows of H (rows[r * rw + wd], column bits of check r) and the
// column-weight planes pl[7][M32] (nonempty columns, colw's three bit planes,
// colw - 1's three bit planes).  In the wave's LDS when staged there, else read
// from global memory.
template <class U>
struct TriTables {
    U hcw;
    U sig;
    U dg;
    U rows;
    U pl;
    int rw;
    int D;
    bool has_dg;
};
using TriLds = TriTables<const uint32_t LDS *>;
using TriGlobal = TriTables<const uint32_t *>;
template <class U>
__device__ __forceinline__ TriTables<U> tri_tables(const DevCode &code, U img) {
    return TriTables<U>{img, img + code.t_sig, img + code.t_dg, img + code.t_rows, img + code.t_pl, code.t_rw,
                        code.D, code.dg != nullptr};
}

// s == h_a (+ h_b for w > 1, + h_c for w > 2); a, b and c are columns of the
// code whatever w is (the words of all three are loaded together).
template <int D, class T>
__device__ __forceinline__ bool lane_exact(const T &tb, int m, const uint32_t (&s)[kMaxWords], int a, int b, int c,
                                           int w) {
    uint32_t xa[D], xb[D], xc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xa[d] = tb.hcw[d * m + a];
        xb[d] = tb.hcw[d * m + b];
        xc[d] = tb.hcw[d * m + c];
    }
    uint32_t bad = 0u;
#pragma unroll
    for (int d = 0; d < D; ++d) bad |= s[d] ^ xa[d] ^ (w > 1 ? xb[d] : 0u) ^ (w > 2 ? xc[d] : 0u);
    return bad == 0u;
}

// Words per row the lane search reads: M32 rounded up to 4, 6 or kLaneMaxM32.
// The image has kLaneMaxM32 + 2 zero words after its last table, so reading NW
// words from any row or plane start stays inside it; the words past M32 are
// never used.
__device__ __forceinline__ int lane_words(int M32) { return M32 <= 4 ? 4 : M32 <= 6 ? 6 : kLaneMaxM32; }

// Bit-sliced counts q2 q1 q0 (0..7) of s's checks in every column: the sum of
// the rows of H at the set bits of s.  One loop over all set bits of the D
// words; a row's NW words are loaded together.
template <int NW, int D, class T>
__device__ __forceinline__ void lane_counts(const T &tb, const uint32_t (&s)[kMaxWords], uint32_t (&q0)[kLaneMaxM32],
                                            uint32_t (&q1)[kLaneMaxM32], uint32_t (&q2)[kLaneMaxM32]) {
#pragma unroll
    for (int wd = 0; wd < kLaneMaxM32; ++wd) { q0[wd] = 0u; q1[wd] = 0u; q2[wd] = 0u; }
    int d = 0;
    uint32_t bs = s[0];
    for (;;) {
        while (bs == 0u && d + 1 < D) {
            ++d;
            uint32_t nx = s[0];
#pragma unroll
            for (int e = 1; e < D; ++e) nx = (d == e) ? s[e] : nx;
            bs = nx;
        }
        if (bs == 0u) break;
        const int r = 32 * d + __ffs(bs) - 1;
        bs &= bs - 1u;
        const auto row = tb.rows + r * tb.rw;
        uint32_t x[NW];
#pragma unroll
        for (int wd = 0; wd < NW; ++wd) x[wd] = row[wd];
#pragma unroll
        for (int wd = 0; wd < NW; ++wd) {
            const uint32_t c0 = q0[wd] & x[wd];
            q0[wd] ^= x[wd];
            const uint32_t c1 = q1[wd] & c0;
            q1[wd] ^= c0;
            q2[wd] ^= c1;
        }
    }
}

// Active columns at iteration it (count == colw, or for it = 2 also colw - 1;
// nonempty columns only) as words act[0 .. M32); returns how many.
template <int NW, class T>
__device__ __forceinline__ int lane_active(const T &tb, int M32, int it, const uint32_t (&q0)[kLaneMaxM32],
                                           const uint32_t (&q1)[kLaneMaxM32], const uint32_t (&q2)[kLaneMaxM32],
                                           uint32_t (&act)[kLaneMaxM32]) {
    uint32_t pl[7][NW];
#pragma unroll
    for (int k = 0; k < 7; ++k)
#pragma unroll
        for (int wd = 0; wd < NW; ++wd) pl[k][wd] = tb.pl[k * M32 + wd];
    int tam = 0;
#pragma unroll
    for (int wd = 0; wd < kLaneMaxM32; ++wd) {
        act[wd] = 0u;
        if (wd < NW) {
            uint32_t a = ~((q0[wd] ^ pl[1][wd]) | (q1[wd] ^ pl[2][wd]) | (q2[wd] ^ pl[3][wd]));
            if (it == 2) a |= ~((q0[wd] ^ pl[4][wd]) | (q1[wd] ^ pl[5][wd]) | (q2[wd] ^ pl[6][wd]));
            act[wd] = (wd < M32) ? (a & pl[0][wd]) : 0u;
            tam += __popc(act[wd]);
        }
    }
    return tam;
}

template <int D, class T>
__device__ __forceinline__ bool lane_search(const DevCode &code, const DevCall &call, const T &tb, int shot,
                            const uint32_t (&s)[kMaxWords], uint32_t sg, int ow, uint32_t dgv, int LDS *LA,
                            uint32_t LDS *LS, const unsigned long long LDS *WL, unsigned long long t0) {
    const int lane = lane_id();
    const int m = code.m;
    const int M32 = code.M32;
    auto nc2 = [](int n) { return n * (n - 1) / 2; };
    auto nc3 = [](int n) { return n * (n - 1) * (n - 2) / 6; };
    // Iterations of decode_shot_wave searched here: the first, and for the CSS
    // family the second (when max_soma >= 2), when the first is searched whole
    // without a hit (or has no list).  done / wdone: patterns (weighted
    // patterns) of the iterations searched without a hit; hib: the hibrido
    // code as decode_shot_wave updates it.
    const int itmax = (call.family == kCC && code.max_soma >= 2) ? 2 : 1;
    unsigned long long ta = 0;
    unsigned done = 0u, wdone = 0u, cum = 0u, wsum = 0u, hidx = 0u;
    int hib = -2, it = 1, tam = 0, hw = 0, ha = 0, hb = 0, hc = 0;
    // The score of column c is |supp(h_c) & supp(s)| + max_soma - colw[c]
    // (nonempty columns), so at threshold max_soma - (it - 1) the column is
    // active iff its count of s's checks is colw[c] (it = 1) or colw[c] - 1 or
    // colw[c] (it = 2).  The counts of all columns, bit-sliced over the column
    // words (planes q0, q1, q2: counts 0..7), are the sum of the rows of H at
    // s's checks; both iterations' active sets are read off them.
    uint32_t q0[kLaneMaxM32], q1[kLaneMaxM32], q2[kLaneMaxM32];
    const int nw = lane_words(M32);
    if (nw == 4) lane_counts<4, D>(tb, s, q0, q1, q2);
    else if (nw == 6) lane_counts<6, D>(tb, s, q0, q1, q2);
    else lane_counts<kLaneMaxM32, D>(tb, s, q0, q1, q2);
    for (; it <= itmax; ++it) {
        // the active columns, in column order (the order of the wave path's
        // active list); a list longer than kLaneMax is not searched here
        uint32_t act[kLaneMaxM32];
        if (nw == 4) tam = lane_active<4>(tb, M32, it, q0, q1, q2, act);
        else if (nw == 6) tam = lane_active<6>(tb, M32, it, q0, q1, q2, act);
        else tam = lane_active<kLaneMaxM32>(tb, M32, it, q0, q1, q2, act);
        if (tam > 0 && tam <= kLaneMax) {
            int k = 0;
#pragma unroll
            for (int wd = 0; wd < kLaneMaxM32; ++wd) {
                uint32_t w = act[wd];
                while (w) {
                    LA[k * 64 + lane] = 32 * wd + __ffs(w) - 1;
                    w &= w - 1u;
                    ++k;
                }
            }
        }
        if (it == 1) ta = now_ticks();             // the first active list is built (time_sum_ms)
        if (tam == 0) continue;                     // no list: the next iteration
        if (tam > kLaneMax || tam > call.cap) return false;
        if (call.family == kCLN && tam > call.max_size) return false;   // tam falls back to 0: second iteration
        const int br = branch_of(call, tam);
        if (br == 0) continue;                      // not searched in this mode: the next iteration
        hib = (br > 0) ? (hib == -1 ? 1 : 0) : -1;
        // the weight list of decode_shot_wave for this active-list size
        const uint64_t wl = WL[tam];
        // the weights the walk reaches before it would give up (a class heavier
        // than 3, or the budget exceeded); `whole`: the walk reaches every class of
        // the list, of tot patterns (wtot weighted by w + 1) in all
        unsigned need = 0u, tot = 0u, wtot = 0u;
        bool whole = true;
        for (uint64_t l = wl; l; l >>= 4) {
            const int w = (int)(l & 15u);
            const unsigned N = (w > 3) ? 0u : (unsigned)((w == 1) ? tam : (w == 2) ? nc2(tam) : nc3(tam));
            if (w > 3 || tot + N > (unsigned)kLaneBudget) { whole = false; break; }
            need |= 1u << w;
            tot += N;
            wtot += N * (unsigned)(w + 1);
        }
        // The list's columns (C) and signatures (S) in registers, loaded
        // together (entries past tam read column 0), and the signatures in LS
        // for the pattern-by-pattern walk after a collision.  First signature
        // match of each needed class, lexicographic order: the fully unrolled
        // loops run backwards and the last match written is the first in
        // order.  Packed as a | b << 4 | c << 8, -1 = none.  tmax: the longest
        // list among the wave's searching lanes (bounds the loops).
        int C[kLaneMax];
        uint32_t S[kLaneMax];
#pragma unroll
        for (int i = 0; i < kLaneMax; ++i) C[i] = LA[i * 64 + lane];
#pragma unroll
        for (int i = 0; i < kLaneMax; ++i) C[i] = (i < tam) ? C[i] : 0;
#pragma unroll
        for (int i = 0; i < kLaneMax; ++i) S[i] = tb.sig[C[i]];
#pragma unroll
        for (int i = 0; i < kLaneMax; ++i)
            if (i < tam) LS[i * 64 + lane] = S[i];
        int tmax = 1;
        for (int t = kLaneMax; t > 1; --t)
            if (ballot(tam >= t)) { tmax = t; break; }
        int f1 = -1, f2 = -1, f3 = -1;
        if (ballot(need & 2u)) {
#pragma unroll
            for (int x = kLaneMax - 1; x >= 0; --x)
                if (x < tam && S[x] == sg) f1 = x;
        }
        if (ballot(need & 4u)) {
#pragma unroll
            for (int x = kLaneMax - 2; x >= 0; --x) {
                if (x + 2 > tmax) continue;
                const uint32_t t = sg ^ S[x];
#pragma unroll
                for (int y = kLaneMax - 1; y > x; --y)
                    if (y < tam && S[y] == t) f2 = x | (y << 4);
            }
        }
        if (ballot(need & 8u)) {
#pragma unroll
            for (int x = kLaneMax - 3; x >= 0; --x) {
                if (x + 3 > tmax) continue;
                const uint32_t tx = sg ^ S[x];
#pragma unroll
                for (int y = kLaneMax - 2; y > x; --y) {
                    if (y + 2 > tmax) continue;
                    const uint32_t t = tx ^ S[y];
#pragma unroll
                    for (int z = kLaneMax - 1; z > y; --z)
                        if (z < tam && S[z] == t) f3 = x | (y << 4) | (z << 8);
                }
            }
        }
        // walk the classes: the first class with a signature match holds the
        // candidate; its rank within the class is in closed form
        cum = 0; wsum = 0;
        for (uint64_t l = wl; l; l >>= 4) {
            const int w = (int)(l & 15u);
            if (w > 3) break;
            const unsigned N = (unsigned)((w == 1) ? tam : (w == 2) ? nc2(tam) : nc3(tam));
            if (cum + N > (unsigned)kLaneBudget) break;
            const int f = (w == 1) ? f1 : (w == 2) ? f2 : f3;
            if (f >= 0) {
                const int x = f & 15, y = (f >> 4) & 15, z = f >> 8;
                const int rank = (w == 1) ? x
                               : (w == 2) ? nc2(tam) - nc2(tam - x) + (y - x - 1)
                                          : nc3(tam) - nc3(tam - x) + nc2(tam - x - 1) - nc2(tam - y) + (z - y - 1);
                hw = w; ha = x; hb = y; hc = z; hidx = cum + (unsigned)rank;
                break;
            }
            cum += N;
            wsum += N * (unsigned)(w + 1);
        }
        // The candidate is confirmed exactly.  A signature collision (an earlier
        // pattern matching the signature but not the syndrome) redoes the walk
        // pattern by pattern with the exact check on every signature match.
        int ca = C[0], cb = C[0], cc = C[0];
#pragma unroll
        for (int i = 1; i < kLaneMax; ++i) {
            if (ha == i) ca = C[i];
            if (hb == i) cb = C[i];
            if (hc == i) cc = C[i];
        }
        if (hw != 0 && !lane_exact<D>(tb, m, s, ca, cb, cc, hw)) {
            cum = 0; wsum = 0; hw = 0;
            for (uint64_t l = wl; l && hw == 0; l >>= 4) {
                const int w = (int)(l & 15u);
                if (w > 3) break;
                const unsigned N = (unsigned)((w == 1) ? tam : (w == 2) ? nc2(tam) : nc3(tam));
                if (cum + N > (unsigned)kLaneBudget) break;
                unsigned r = 0;
                if (w == 1) {
                    for (int x = 0; x < tam && hw == 0; ++x, ++r)
                        if (LS[x * 64 + lane] == sg && lane_exact<D>(tb, m, s, LA[x * 64 + lane], 0, 0, 1)) {
                            hw = 1; ha = x; hidx = cum + r;
                        }
                } else if (w == 2) {
                    for (int x = 0; x + 1 < tam && hw == 0; ++x) {
                        const uint32_t tx = sg ^ LS[x * 64 + lane];
                        for (int y = x + 1; y < tam && hw == 0; ++y, ++r)
                            if (LS[y * 64 + lane] == tx &&
                                lane_exact<D>(tb, m, s, LA[x * 64 + lane], LA[y * 64 + lane], 0, 2)) {
                                hw = 2; ha = x; hb = y; hidx = cum + r;
                            }
                    }
                } else {
                    for (int x = 0; x + 2 < tam && hw == 0; ++x) {
                        const uint32_t tx = sg ^ LS[x * 64 + lane];
                        for (int y = x + 1; y + 1 < tam && hw == 0; ++y) {
                            const uint32_t txy = tx ^ LS[y * 64 + lane];
                            for (int z = y + 1; z < tam && hw == 0; ++z, ++r)
                                if (LS[z * 64 + lane] == txy &&
                                    lane_exact<D>(tb, m, s, LA[x * 64 + lane], LA[y * 64 + lane], LA[z * 64 + lane], 3)) {
                                    hw = 3; ha = x; hb = y; hc = z; hidx = cum + r;
                                }
                        }
                    }
                }
                if (hw == 0) { cum += N; wsum += N * (unsigned)(w + 1); }
            }
        }
        if (hw != 0) break;
        if (!whole) return false;                   // classes the lane does not search
        done += tot;
        wdone += wtot;
    }
    if (hw == 0) return false;
    const unsigned long long t1 = now_ticks();
    const int c0 = LA[ha * 64 + lane];
    const int c1 = hw > 1 ? LA[hb * 64 + lane] : c0;
    const int c2 = hw > 2 ? LA[hc * 64 + lane] : c0;
    const unsigned count = done + hidx + 1u;
    wsum = wdone + wsum + (hidx + 1u - cum) * (unsigned)(hw + 1);
    for (int wd = 0; wd < M32; ++wd) {
        uint32_t v = 0u;
        if ((c0 >> 5) == wd) v |= 1u << (c0 & 31);
        if (hw > 1 && (c1 >> 5) == wd) v |= 1u << (c1 & 31);
        if (hw > 2 && (c2 >> 5) == wd) v |= 1u << (c2 & 31);
        call.dec[(size_t)shot * M32 + wd] = v;
    }
    int eq = 0, deg = 0;
    if (call.family == kCC && call.input_errors) {
        const uint32_t *e = call.e_packed + (size_t)shot * M32;
        bool all = (ow == hw) && ((e[c0 >> 5] >> (c0 & 31)) & 1u);
        if (hw > 1) all = all && ((e[c1 >> 5] >> (c1 & 31)) & 1u);
        if (hw > 2) all = all && ((e[c2 >> 5] >> (c2 & 31)) & 1u);
        eq = all ? 1 : 0;
        uint32_t gh = tb.dg[c0];
        if (hw > 1) gh ^= tb.dg[c1];
        if (hw > 2) gh ^= tb.dg[c2];
        deg = (gh == dgv) ? 1 : 0;
    } else if (call.family == kCLN) {
        const int NL32 = code.NL32;
        const int nlw = (code.NL + 31) / 32;
        bool same = true;
        for (int wd = 0; wd < nlw; ++wd) {
            uint32_t le = code.Lp[(size_t)c0 * NL32 + wd];
            if (hw > 1) le ^= code.Lp[(size_t)c1 * NL32 + wd];
            if (hw > 2) le ^= code.Lp[(size_t)c2 * NL32 + wd];
            call.declog[(size_t)shot * NL32 + wd] = le;
            same = same && (le == call.lg_packed[(size_t)shot * NL32 + wd]);
        }
        deg = same ? 1 : 0;
    }
    const double ts = (double)(t1 - t0) * call.ms_per_tick;
    const uint64_t tsb = (uint64_t)__double_as_longlong(ts);
    const uint64_t tub = (uint64_t)__double_as_longlong((double)(ta - t0) * call.ms_per_tick);
    uint32_t *p = reinterpret_cast<uint32_t *>(&call.results[shot]);
    p[0] = 1u; p[1] = (uint32_t)hw; p[2] = (uint32_t)ow; p[3] = (uint32_t)deg; p[4] = (uint32_t)eq;
    p[5] = (uint32_t)tam; p[6] = count; p[7] = 0u; p[8] = (uint32_t)hib; p[9] = (uint32_t)it;
    p[10] = (uint32_t)tub; p[11] = (uint32_t)(tub >> 32); p[12] = (uint32_t)tsb; p[13] = (uint32_t)(tsb >> 32);
    p[14] = 0u; p[15] = 0u;
    p[16] = (uint32_t)tsb; p[17] = (uint32_t)(tsb >> 32);
    if (call.stats) {
        call.stats[3 * (size_t)shot] = t0;
        call.stats[3 * (size_t)shot + 1] = t1;
        call.stats[3 * (size_t)shot + 2] = wsum;
    }
    return true;
}

template <class T>
__device__ __forceinline__ bool lane_search_any(const DevCode &code, const DevCall &call, const T &tb, int shot,
                                                const uint32_t (&s)[kMaxWords], uint32_t sg, int ow, uint32_t dgv,
                                                int LDS *LA, uint32_t LDS *LS, const unsigned long long LDS *WL,
                                                unsigned long long t0) {
    switch (code.D) {
        case 1: return lane_search<1>(code, call, tb, shot, s, sg, ow, dgv, LA, LS, WL, t0);
        case 2: return lane_search<2>(code, call, tb, shot, s, sg, ow, dgv, LA, LS, WL, t0);
        case 3: return lane_search<3>(code, call, tb, shot, s, sg, ow, dgv, LA, LS, WL, t0);
        case 4: return lane_search<4>(code, call, tb, shot, s, sg, ow, dgv, LA, LS, WL, t0);
        case 5: return lane_search<5>(code, call, tb, shot, s, sg, ow, dgv, LA, LS, WL, t0);
        case 6: return lane_search<6>(code, call, tb, shot, s, sg, ow, dgv, LA, LS, WL, t0);
        default: return false;
    }
}

// Row and column of entry e of a row-major block with `ncols` columns, without
// an integer division: e < 2^18, so e * (1/ncols) in float is off by at most one.
__device__ __forceinline__ void entry_rc(int e, int ncols, float inv, int &r, int &c) {
    r = (int)((float)e * inv);
    c = e - r * ncols;
    if (c < 0) { --r; c += ncols; }
    if (c >= ncols) { ++r; c -= ncols; }
}

// Packs `cnt` consecutive int32 rows of `ncols` entries (row-major, contiguous,
// starting at row `base`) into bits[64][W32]: bit c of word c / 32 of row r is
// set iff entry (r, c) is odd; `bad` collects any value other than 0/1.  The
// block is read front to back as 16-byte vectors (fully coalesced), a batch of
// kPackBatch per lane issued together.  Set entries are rare, so only a vector
// holding one works out its row and column and sets its bits with one LDS OR
// (the four entries of a vector share a row and a word when ncols % 4 == 0).
constexpr int kPackBatch = 12;
__device__ __forceinline__ void pack_rows(const int *rows, int ncols, int base, int cnt, int W32, uint32_t LDS *bits,
                                          unsigned &bad) {
    const int lane = lane_id();
    __syncthreads();                              // earlier readers of `bits` are done
    for (int i = lane; i < 64 * W32; i += 64) bits[i] = 0u;
    __syncthreads();
    const int *blk = rows + (size_t)base * ncols;
    const int total = cnt * ncols;                // <= 64 * kMaxM
    const float inv = 1.0f / (float)ncols;
    if ((ncols & 3) == 0 && (((uintptr_t)blk) & 15) == 0) {
        const int4 *v4 = reinterpret_cast<const int4 *>(blk);
        const int nv = total >> 2;
        for (int q0 = 0; q0 < nv; q0 += 64 * kPackBatch) {
            int4 v[kPackBatch];
#pragma unroll
            for (int k = 0; k < kPackBatch; ++k) {
                const int q = q0 + 64 * k + lane;
                v[k] = (q < nv) ? v4[q] : make_int4(0, 0, 0, 0);
            }
#pragma unroll
            for (int k = 0; k < kPackBatch; ++k) {
                const uint32_t o = (uint32_t)(v[k].x | v[k].y | v[k].z | v[k].w);
                bad |= o & ~1u;
                if (o & 1u) {
                    int r, c;
                    entry_rc(4 * (q0 + 64 * k + lane), ncols, inv, r, c);
                    const uint32_t nib = ((uint32_t)v[k].x & 1u) | (((uint32_t)v[k].y & 1u) << 1) |
                                         (((uint32_t)v[k].z & 1u) << 2) | (((uint32_t)v[k].w & 1u) << 3);
                    __hip_atomic_fetch_or(&bits[r * W32 + (c >> 5)], nib << (c & 31), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    } else {
        for (int q0 = 0; q0 < total; q0 += 64 * kPackBatch) {
            int v[kPackBatch];
#pragma unroll
            for (int k = 0; k < kPackBatch; ++k) {
                const int q = q0 + 64 * k + lane;
                v[k] = (q < total) ? blk[q] : 0;
            }
#pragma unroll
            for (int k = 0; k < kPackBatch; ++k) {
                bad |= (uint32_t)v[k] & ~1u;
                if (v[k] & 1) {
                    int r, c;
                    entry_rc(q0 + 64 * k + lane, ncols, inv, r, c);
                    __hip_atomic_fetch_or(&bits[r * W32 + (c >> 5)], 1u << (c & 31), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }
    __syncthreads();
}

// Writes the table image held in registers (sv_vecs 16-byte vectors, vector
// lane + 64 i in sv[i]) to the wave's LDS.
constexpr int kTriageStageVec = 8;
__device__ __forceinline__ void flush_stage(const uint4 (&sv)[kTriageStageVec], int sv_vecs, uint32_t LDS *tab) {
    const int lane = lane_id();
#pragma unroll
    for (int i = 0; i < kTriageStageVec; ++i) {
        const int q = lane + 64 * i;
        if (q < sv_vecs) {
            tab[4 * q] = sv[i].x; tab[4 * q + 1] = sv[i].y; tab[4 * q + 2] = sv[i].z; tab[4 * q + 3] = sv[i].w;
        }
    }
    __syncthreads();
}

// s = H e, sig(s), the original weight and the XOR of the degeneracy words over
// the lane's set error bits (its packed row in `bits`), for codes of at most D
// syndrome words: a set bit's D column words, signature and degeneracy word are
// loaded together.  D = kMaxWords covers every code, with the words past code.D
// guarded.
template <int D, class T>
__device__ __forceinline__ void syndrome_of(const T &tb, int m, const uint32_t LDS *bits, int M32,
                                            uint32_t (&s)[kMaxWords], uint32_t &sg, uint32_t &dgv, int &ow) {
    const int lane = lane_id();
    const int Dc = tb.D;
    for (int wd = 0; wd < M32; ++wd) {
        uint32_t b = bits[lane * M32 + wd];
        ow += __popc(b);
        while (b) {
            const int c = 32 * wd + __ffs(b) - 1;
            b &= b - 1u;
            uint32_t x[D];
#pragma unroll
            for (int d = 0; d < D; ++d) x[d] = (D < kMaxWords || d < Dc) ? tb.hcw[d * m + c] : 0u;
            const uint32_t g = tb.sig[c];
            const uint32_t h = tb.dg[tb.has_dg ? c : 0];
#pragma unroll
            for (int d = 0; d < D; ++d) s[d] ^= x[d];
            sg ^= g;
            if (tb.has_dg) dgv ^= h;
        }
    }
}

// One lane per shot, 64 shots per wave: packs the input rows (errors or
// syndromes, and logicals), forms s = H e and sig(s) as XORs over the set error
// bits (both maps are GF(2)-linear), searches small first iterations by lane,
// writes the result of zero-syndrome and lane-solved shots and appends the
// others to the worklist with one atomic per 64 shots.
template <class T>
__device__ __forceinline__ void triage_groups(const DevCode &code, const DevCall &call, const T &tb, int LDS *LA,
                                              uint32_t LDS *LS, const unsigned long long LDS *WL,
                                              uint32_t LDS *bits, const uint4 (&sv)[kTriageStageVec], int sv_vecs,
                                              uint32_t LDS *tab) {
    const int lane = lane_id();
    const int m = code.m;
    const int groups = (call.num_shots + 63) / 64;
    unsigned bad = 0;
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int base = g * 64;
        const int cnt = (call.num_shots - base) < 64 ? (call.num_shots - base) : 64;
        const bool valid = lane < cnt;
        const int shot = base + lane;
        uint32_t s[kMaxWords];
#pragma unroll
        for (int d = 0; d < kMaxWords; ++d) s[d] = 0u;
        uint32_t sg = 0, dgv = 0;
        int ow = 0;
        if (call.input_errors) {
            const int M32 = code.M32;
            pack_rows(call.input, m, base, cnt, M32, bits, bad);
            if (g == (int)blockIdx.x) flush_stage(sv, sv_vecs, tab);
            // the block's packed rows are consecutive in e_packed
            for (int i = lane; i < cnt * M32; i += 64) call.e_packed[(size_t)base * M32 + i] = bits[i];
            switch (code.D) {
                case 1: syndrome_of<1>(tb, m, bits, M32, s, sg, dgv, ow); break;
                case 2: syndrome_of<2>(tb, m, bits, M32, s, sg, dgv, ow); break;
                case 3: syndrome_of<3>(tb, m, bits, M32, s, sg, dgv, ow); break;
                case 4: syndrome_of<4>(tb, m, bits, M32, s, sg, dgv, ow); break;
                case 5: syndrome_of<5>(tb, m, bits, M32, s, sg, dgv, ow); break;
                case 6: syndrome_of<6>(tb, m, bits, M32, s, sg, dgv, ow); break;
                case 7: syndrome_of<7>(tb, m, bits, M32, s, sg, dgv, ow); break;
                case 8: syndrome_of<8>(tb, m, bits, M32, s, sg, dgv, ow); break;
                default: syndrome_of<kMaxWords>(tb, m, bits, M32, s, sg, dgv, ow); break;
            }
        } else {
            const int D = code.D;
            pack_rows(call.input, code.n, base, cnt, D, bits, bad);
            if (g == (int)blockIdx.x) flush_stage(sv, sv_vecs, tab);
#pragma unroll
            for (int d = 0; d < kMaxWords; ++d) {
                if (d < D) {
                    uint32_t b = bits[lane * D + d];
                    s[d] = b;
                    while (b) {
                        sg ^= code.rowsig[32 * d + __ffs(b) - 1];
                        b &= b - 1u;
                    }
                }
            }
        }
        if (call.family == kCLN) {
            const int NL32 = code.NL32;
            pack_rows(call.logicals, code.NL, base, cnt, NL32, bits, bad);
            for (int i = lane; i < cnt * NL32; i += 64) call.lg_packed[(size_t)base * NL32 + i] = bits[i];
        }
        uint32_t any = 0;
#pragma unroll
        for (int d = 0; d < kMaxWords; ++d) any |= s[d];
        const bool nz = valid && any != 0u;
        bool solved = false;
        if (call.triage_search && ballot(nz)) {
            const unsigned long long t0 = now_ticks();
            if (nz) solved = lane_search_any(code, call, tb, shot, s, sg, ow, dgv, LA, LS, WL, t0);
        }
        const bool queued = nz && !solved;
        if (valid) {
#pragma unroll
            for (int d = 0; d < kMaxWords; ++d)
                if (d < code.D) call.s_packed[(size_t)shot * code.D + d] = s[d];
            call.s_sig[shot] = sg;
            if (call.input_errors && queued) {
                call.e_aux[2 * (size_t)shot] = (uint32_t)ow;
                call.e_aux[2 * (size_t)shot + 1] = dgv;
            }
        }
        if (valid && !nz) {
            write_zero_result(&call.results[shot], (call.family == kCC) ? ow : 0);
            for (int wd = 0; wd < code.M32; ++wd) call.dec[(size_t)shot * code.M32 + wd] = 0u;
            if (call.family == kCLN)
                for (int wd = 0; wd < code.NL32; ++wd) call.declog[(size_t)shot * code.NL32 + wd] = 0u;
            if (call.stats) {
                call.stats[3 * (size_t)shot] = 0;
                call.stats[3 * (size_t)shot + 1] = 0;
                call.stats[3 * (size_t)shot + 2] = 0;
            }
        }
        // search kernel: shots whose syndrome weight is at least that of two
        // disjoint columns go to the front of the worklist (their deep classes then
        // overlap with the rest of the batch), the others to the back; the plain
        // engine reads one list.  The group's partition (g % wl_parts) spreads the
        // counter atomics over wl_parts cache lines.
        int sw = 0;
#pragma unroll
        for (int d = 0; d < kMaxWords; ++d) sw += __popc(s[d]);
        const bool front = queued && (!call.split_worklist || sw >= 2 * code.max_soma);
        const uint64_t mf = ballot(front), mb = ballot(queued && !front);
        const int part = g % call.wl_parts;
        unsigned *pc = call.counters + kWlBase + kWlStride * part;
        const unsigned r0 = wl_part_start(part, groups, call.wl_parts);
        if (mf) {
            unsigned pos = 0;
            if (lane == 0) pos = atomicAdd(&pc[0], (unsigned)__popcll(mf));
            pos = uni(pos);
            if (front) call.worklist[r0 + pos + lanes_below(mf)] = shot;
        }
        if (mb) {
            unsigned pos = 0;
            if (lane == 0) pos = atomicAdd(&pc[1], (unsigned)__popcll(mb));
            pos = uni(pos);
            const unsigned r1 = r0 + wl_part_slots(part, groups, call.wl_parts) - 1u;
            if (queued && !front) call.worklist[r1 - (pos + lanes_below(mb))] = shot;
        }
    }
    if (ballot(bad != 0u) && lane == 0) atomicOr(&call.counters[2], kErrNonBinary);
}

// LDS: the lane search's [kLaneMax + 1][64] positions and signatures and its
// weight lists by active-list size [kTriageWlWords / 2]; the code tables when
// call.triage_stage_words > 0 (the image DevCode::timg); then the packed
// block, 64 rows of the larger of the input row's words (M32 for errors, D for
// syndromes) and, for the DEM family, the NL32 words of the logicals.
__global__ __launch_bounds__(64) void triage_kernel(const DevCode *__restrict__ codep, const DevCall *__restrict__ callp) {
    const DevCode &code = *codep;
    const DevCall &call = *callp;
    const int lane = lane_id();
    int LDS *LA = (int LDS *)lds_base();
    uint32_t LDS *LS = lds_base() + (kLaneMax + 1) * 64;
    unsigned long long LDS *WL = (unsigned long long LDS *)(lds_base() + 2 * (kLaneMax + 1) * 64);
    uint32_t LDS *tab = lds_base() + 2 * (kLaneMax + 1) * 64 + kTriageWlWords;
    if (lane <= kLaneMax) WL[lane] = (lane == 0) ? 0ull : lane_weight_list(call, lane);
    const int tsw = call.triage_stage_words;          // a multiple of 4
    uint4 sv[kTriageStageVec];
    if (tsw > 0) {
        // An image of up to 256 * kTriageStageVec words is loaded here and
        // written to LDS after the first block is packed (its loads overlap the
        // block's); a larger one is staged now.
        const uint4 *src = reinterpret_cast<const uint4 *>(code.timg);
        const int nv = tsw >> 2;
        if (nv <= 64 * kTriageStageVec) {
#pragma unroll
            for (int i = 0; i < kTriageStageVec; ++i) {
                const int q = lane + 64 * i;
                sv[i] = (q < nv) ? src[q] : make_uint4(0u, 0u, 0u, 0u);
            }
        } else {
            for (int q = lane; q < nv; q += 64) {
                const uint4 v = src[q];
                tab[4 * q] = v.x; tab[4 * q + 1] = v.y; tab[4 * q + 2] = v.z; tab[4 * q + 3] = v.w;
            }
        }
        triage_groups(code, call, tri_tables(code, (const uint32_t LDS *)tab), LA, LS, WL, tab + tsw, sv,
                      nv <= 64 * kTriageStageVec ? nv : 0, tab);
    } else {
        triage_groups(code, call, tri_tables(code, code.timg), LA, LS, WL, tab, sv, 0, tab);
    }
}

// Persistent search kernel: waves claim worklist entries in batches until none remain.
template <bool ML>
__global__ __launch_bounds__(64) void fused_kernel(const DevCode *__restrict__ codep, const DevCall *__restrict__ callp) {
    const DevCode &code = *codep;
    const DevCall &call = *callp;
    const WaveLds L = carve_lds(call.cap, ML || call.lscores != nullptr, code.lds_words,
                                call.claim_batch * call.stage_words, call.pair_cap, pair_head_words(ML));
    if (code.lds_words > 0) {
        uint32_t LDS *dst = lds_base() + kLdsCodeOffset;
        const int nw = code.D * code.m;
#pragma unroll 8
        for (int i = lane_id(); i < nw; i += 64) dst[i] = code.Hcw[i];
#pragma unroll 4
        for (int i = lane_id(); i < code.m; i += 64) {
            dst[nw + i] = (uint32_t)code.off[i];
            dst[nw + code.m + i] = code.sig[i];
        }
        if (code.lds_post == 2) {
            uint32_t LDS *post = dst + nw + 2 * code.m;
#pragma unroll 4
            for (int i = lane_id(); i < code.m; i += 64) post[i] = code.dg[i];
        } else if (code.lds_post == 1) {
            uint32_t LDS *post = dst + nw + 2 * code.m;
            const int nr = code.rank * code.M32;
#pragma unroll 4
            for (int i = lane_id(); i < code.m; i += 64) post[i] = (uint32_t)code.pivrow[i];
#pragma unroll 4
            for (int i = lane_id(); i < nr; i += 64) post[code.m + i] = code.R[i];
        }
    }
    // worklist: shots likely to search deep first (the front entries of every
    // partition, `heavy` of them), then the others (the partitions' back
    // entries); lane k < wl_parts gets the exclusive prefix sums of partition k
    unsigned wf = 0u, wb = 0u;
    if (lane_id() < call.wl_parts) {
        wf = call.counters[kWlBase + kWlStride * lane_id()];
        wb = call.counters[kWlBase + kWlStride * lane_id() + 1];
    }
    unsigned wfx = wf, wbx = wb;
#pragma unroll
    for (int o = 1; o < kWlParts; o <<= 1) {
        const unsigned a = (unsigned)__shfl_up((int)wfx, o, 64), b = (unsigned)__shfl_up((int)wbx, o, 64);
        if (lane_id() >= o) { wfx += a; wbx += b; }
    }
    const unsigned heavy = uni((unsigned)__shfl((int)wfx, call.wl_parts - 1, 64));
    const unsigned total = heavy + uni((unsigned)__shfl((int)wbx, call.wl_parts - 1, 64));
    wfx -= wf;
    wbx -= wb;
    // shots per claim: the tuned batch when the worklist is long, fewer when
    // there are not several claims per wave (small batches finish sooner)
    unsigned B = total / (4u * gridDim.x);
    B = B < 1u ? 1u : (B > (unsigned)call.claim_batch ? (unsigned)call.claim_batch : B);
    HelpCache hc;          // task whose arrays and tables this wave's LDS holds
    hc.key = -1;
    hc.hmask = 0;
    {
        // 17 loads per lane, issued together
        constexpr int kN = kNcrLdsRows * kNcrStride;
        unsigned long long v[(kN + 63) / 64];
#pragma unroll
        for (int q = 0; q < (kN + 63) / 64; ++q) {
            const int i = lane_id() + 64 * q;
            v[q] = (i < kN) ? code.nCr[i] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < (kN + 63) / 64; ++q) {
            const int i = lane_id() + 64 * q;
            if (i < kN) L.ncr[i] = v[q];
        }
    }
    __syncthreads();
    // Claims.  The worklist positions are dealt over kClaimWords claim words
    // (word x hands out positions x, x + 8, x + 16, ...), each in its own cache
    // line; a wave starts at the word of its XCD slot (workgroups are dispatched
    // to the 8 XCDs round-robin) and moves to the next word when one runs dry, so
    // the hot atomics are spread 8 ways.  A claim word is (registrations << 32) |
    // cursor: every claim RMW adds (1 << 32) + B, so a wave that got shots is
    // registered in that word before its cursor passed the end (RMWs on one
    // location are totally ordered), and each wave adds its number of claim RMWs
    // to the finished count once it has found every word dry.  Only a wave still
    // decoding can publish a shared class, so helpers leave when finished equals
    // the registrations of all words and no class is open; waves that are not
    // resident yet hold no work.  The open-class poll rides along with the claim.
    unsigned long long *cw = reinterpret_cast<unsigned long long *>(call.counters) + kClaimBase;
    int cur = (int)(blockIdx.x & (kClaimWords - 1));
    unsigned claims = 0;
    bool drained = false;
    for (;;) {
        unsigned got = 0xFFFFFFFFu, open = 0u;
        if (lane_id() == 0) {
            if (call.split_threshold) open = ld_relaxed(&call.task_counters[0]);
            if (!drained) got = (unsigned)atomicAdd(cw + (size_t)cur * kClaimStride, (1ull << 32) + B);
        }
        open = uni(open);
        got = uni(got);
        // the claimed shots are decoded first; a wave serves the open classes of
        // deep shots only while it holds no claim (claims held during helping
        // left light shots to the end of the batch)
        bool helped = false;
        const bool have_claim = !drained && got < total && (unsigned)cur + kClaimWords * got < total;
        if (open != 0u && !have_claim)
            while (help_task<ML>(code, call, L, hc)) helped = true;
        if (!drained) {
            ++claims;
            const unsigned p0 = (unsigned)cur + kClaimWords * got;
            if (got < total && p0 < total) {
                const unsigned left = (total - 1u - p0) / kClaimWords + 1u;   // positions p0, p0 + 8, ... below total
                const int nb = (int)(left < B ? left : B);
                stage_claim(code, call, p0, nb, heavy, wfx, wbx, L);
                for (int j = 0; j < nb; ++j) decode_shot_wave<ML>(code, call, L.stage + j * call.stage_words, L);
                hc.key = -1;
                continue;
            }
            // this word is dry: read every word's cursor at once (they only grow)
            // and move to the next word that still holds positions, or drain
            unsigned live = 0u;
            if (lane_id() < kClaimWords) {
                const unsigned c = (unsigned)ld_relaxed(cw + (size_t)lane_id() * kClaimStride);
                live = ((unsigned)lane_id() + kClaimWords * c < total) ? 1u : 0u;
            }
            const uint64_t lv = ballot(live != 0u) & ~(1ull << cur);
            if (lv) {
                const uint64_t after = lv & ~((2ull << cur) - 1ull);   // words after cur, else the first one
                cur = (int)uni((unsigned)(__ffsll((long long)(after ? after : lv)) - 1));
                continue;
            }
            drained = true;
            if (lane_id() == 0) atomicAdd(&call.task_counters[1], claims);
        }
        if (!call.split_threshold) break;
        if (helped) continue;
        const unsigned fin = uni(ld_relaxed(&call.task_counters[1]));     // finished first, then registrations
        unsigned reg = (lane_id() < kClaimWords)
                           ? (unsigned)(ld_relaxed(cw + (size_t)lane_id() * kClaimStride) >> 32) : 0u;
#pragma unroll
        for (int o = 4; o >= 1; o >>= 1) reg += (unsigned)__shfl_xor((int)reg, o, 64);
        reg = uni(reg);
        const unsigned op = uni(ld_relaxed(&call.task_counters[0]));
        if (fin == reg && op == 0u) break;
        __builtin_amdgcn_s_sleep(64);
    }
}

// --------------------------------------------------------------- lane scanners
// A lane walks `len` consecutive ranks of one weight-W class starting at rank0.
// MODE 0: stop at the first exact match.  MODE 1 (ML): visit all, keep the
// first strictly cheapest match.  Returns the offset of the hit within the
// lane's range (-1 = none) and its fp32 cost.
struct ScanHit {
    int j;
    float cost;
};

template <int W, int MODE>
__device__ __forceinline__ ScanHit scan_range(const DevCode &code, const uint32_t LDS *s, uint32_t t0, uint64_t rank0,
                                              int len, int tam, const int LDS *act, const uint32_t LDS *sg,
                                              const float LDS *llr) {
    ScanHit h{-1, 0.0f};
    if (len <= 0) return h;
    int c[W];
    unrank_t<W>(rank0, tam, code.nCr, c);
    uint32_t acc[W];    // acc[l] = t0 ^ sg[c0] ^ ... ^ sg[c_l]
    uint32_t a = t0;
#pragma unroll
    for (int l = 0; l < W - 1; ++l) { a ^= sg[c[l]]; acc[l] = a; }
    uint32_t pre = a;   // prefix of the first W-1 elements
    for (int j = 0; j < len; ++j) {
        const uint32_t x = pre ^ sg[c[W - 1]];
        if (x == 0u) {
            int cols[kMaxW];
#pragma unroll
            for (int l = 0; l < kMaxW; ++l) cols[l] = (l < W) ? act[c[l]] : 0;
            if (exact_match(code, s, cols, W)) {
                if (MODE == 0) { h.j = j; break; }
                float cost = 0.0f;
#pragma unroll
                for (int l = 0; l < W; ++l) cost = __fadd_rn(cost, llr[c[l]]);
                if (h.j < 0 || cost < h.cost) { h.j = j; h.cost = cost; }
            }
        }
        if (j + 1 == len) break;
        // lexicographic successor
        c[W - 1] += 1;
        if (c[W - 1] >= tam) {
            int i = -1;
#pragma unroll
            for (int l = 0; l < W - 1; ++l) if (c[l] < tam - W + l) i = l;
#pragma unroll
            for (int l = 0; l < W; ++l) {
                if (l == i) c[l] += 1;
                else if (l > 0 && l > i) c[l] = c[l - 1] + 1;
            }
#pragma unroll
            for (int l = 0; l < W - 1; ++l)
                if (l >= i) acc[l] = ((l == 0) ? t0 : acc[l - 1]) ^ sg[c[l]];
            pre = (W >= 2) ? acc[(W >= 2) ? W - 2 : 0] : t0;
        }
    }
    return h;
}

template <int MODE>
__device__ __attribute__((noinline)) ScanHit scan_range_w(int w, const DevCode &code, const uint32_t LDS *s,
                                                          uint32_t t0, uint64_t rank0, int len, int tam,
                                                          const int LDS *act, const uint32_t LDS *sg,
                                                          const float LDS *llr) {
    switch (w) {
        case 1: return scan_range<1, MODE>(code, s, t0, rank0, len, tam, act, sg, llr);
        case 2: return scan_range<2, MODE>(code, s, t0, rank0, len, tam, act, sg, llr);
        case 3: return scan_range<3, MODE>(code, s, t0, rank0, len, tam, act, sg, llr);
        case 4: return scan_range<4, MODE>(code, s, t0, rank0, len, tam, act, sg, llr);
        case 5: return scan_range<5, MODE>(code, s, t0, rank0, len, tam, act, sg, llr);
        default: return scan_range<6, MODE>(code, s, t0, rank0, len, tam, act, sg, llr);
    }
}

// ------------------------------------------------------------ plain engine
// Per-shot state the host loop reads back after every launch.
struct PlainState {
    int tam_new;
    int pad;
    unsigned long long best;      // first-hit: lowest matching rank in the class
    float prefix[kMaxW + 1];
};

// Builds A for threshold thr into the shot's global scratch (and, for ML with
// sort_tam > 0, the LLR-sorted copy with its prefix sums).
__global__ __launch_bounds__(64) void plain_active_kernel(const DevCode *__restrict__ codep, const DevCall *__restrict__ callp, int shot, int thr, int sort_tam,
                                                          int *gA, uint32_t *gsig, float *gllr, int *gP, uint32_t *gsigP,
                                                          float *gllrP, PlainState *st) {
    const DevCode &code = *codep;
    const DevCall &call = *callp;
    const int lane = lane_id();
    const WaveLds L = carve_lds(call.cap, true);
    if (sort_tam > 0) {
        for (int i = lane; i < sort_tam; i += 64) { L.A[i] = gA[i]; L.sigA[i] = gsig[i]; }
        __syncthreads();
        sort_by_llr(code, L, sort_tam);
        for (int i = lane; i < sort_tam; i += 64) { gP[i] = L.P[i]; gsigP[i] = L.sigP[i]; gllrP[i] = L.llrP[i]; }
        if (lane == 0) {
            float p = 0.0f;
            st->prefix[0] = 0.0f;
            for (int w = 1; w <= kMaxW; ++w) { if (w <= sort_tam) p = __fadd_rn(p, L.llrP[w - 1]); st->prefix[w] = p; }
        }
        return;
    }
    if (lane < kMaxWords) L.sw[lane] = (lane < code.D) ? call.s_packed[(size_t)shot * code.D + lane] : 0u;
    __syncthreads();
    const int tam = build_active(code, L.sw, thr, L, call.cap);
    const int keep = tam < call.cap ? tam : call.cap;
    for (int i = lane; i < keep; i += 64) { gA[i] = L.A[i]; gsig[i] = L.sigA[i]; }
    if (lane == 0) { st->tam_new = tam; st->best = ~0ull; }
}

// One wave per chunk of 64 * R ranks of the weight-w class (grid-stride).
// First-hit: atomicMin of the matching rank.  ML: per-wave best record.
struct MlRecord {
    unsigned key;          // order_key(cost), 0xFFFFFFFF = none
    int pad;
    unsigned long long rank;
    int c[kMaxW];
};

__global__ __launch_bounds__(64) void plain_class_kernel(const DevCode *__restrict__ codep, const DevCall *__restrict__ callp, int shot, int w, int tam, int R,
                                                         unsigned long long nchunks, int ml, const int *gA,
                                                         const uint32_t *gsig, const float *gllr, PlainState *st,
                                                         MlRecord *rec) {
    const DevCode &code = *codep;
    const DevCall &call = *callp;
    const int lane = lane_id();
    // stage the class's arrays and the syndrome in LDS
    const WaveLds L = carve_lds(call.cap, true);
    if (lane < kMaxWords) L.sw[lane] = (lane < code.D) ? call.s_packed[(size_t)shot * code.D + lane] : 0u;
    for (int i = lane; i < tam; i += 64) {
        L.P[i] = gA[i];
        L.sigP[i] = gsig[i];
        if (ml) L.llrP[i] = gllr[i];
    }
    __syncthreads();
    const uint32_t LDS *s = L.sw;
    const uint32_t t0 = call.s_sig[shot];
    const uint64_t Nw = ncr(code.nCr, tam, w);
    uint32_t bkey = 0xFFFFFFFFu;
    uint64_t brank = 0;
    for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint64_t base = ch * 64ull * (uint64_t)R;
        if (!ml && base > __hip_atomic_load(&st->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        const uint64_t lstart = base + (uint64_t)lane * (uint64_t)R;
        const int len = lstart >= Nw ? 0 : (int)((Nw - lstart) < (uint64_t)R ? (Nw - lstart) : (uint64_t)R);
        if (!ml) {
            const int j = scan_range_w<0>(w, code, s, t0, lstart, len, tam, L.P, L.sigP, nullptr).j;
            const uint64_t b = ballot(j >= 0);
            if (b) {
                const int hl = __ffsll((long long)b) - 1;
                const uint64_t r = base + (uint64_t)hl * (uint64_t)R + (uint64_t)bcast_i(j, hl);
                if (lane == 0) atomicMin(&st->best, (unsigned long long)r);
                break;
            }
        } else {
            const ScanHit h = scan_range_w<1>(w, code, s, t0, lstart, len, tam, L.P, L.sigP, L.llrP);
            const uint32_t key = (h.j >= 0) ? order_key(h.cost) : 0xFFFFFFFFu;
            const uint32_t kmin = wave_min(key);
            const uint64_t b = ballot(h.j >= 0 && key == kmin);
            if (b && kmin < bkey) {   // chunks visited in rank order: strict keeps the earliest
                const int bl = __ffsll((long long)b) - 1;
                bkey = kmin;
                brank = (uint64_t)__shfl((long long)(lstart + (uint64_t)(h.j < 0 ? 0 : h.j)), bl, 64);
            }
        }
    }
    if (ml) {
        int bc[kMaxW] = {0, 0, 0, 0, 0, 0};
        if (bkey != 0xFFFFFFFFu) unrank_dyn(brank, tam, w, code.nCr, bc);
        if (lane == 0) {
            MlRecord r;
            r.key = bkey; r.pad = 0; r.rank = brank;
#pragma unroll
            for (int l = 0; l < kMaxW; ++l) r.c[l] = bc[l];
            rec[blockIdx.x] = r;
        }
    }
}

// Reduces per-wave ML records: smallest (key, rank).
__global__ __launch_bounds__(64) void plain_ml_reduce_kernel(const MlRecord *rec, int nrec, MlRecord *out) {
    if (lane_id() != 0) return;
    MlRecord b = rec[0];
    for (int i = 1; i < nrec; ++i) {
        const MlRecord r = rec[i];
        if (r.key < b.key || (r.key == b.key && r.rank < b.rank)) b = r;
    }
    *out = b;
}

// Rebuilds the hit of a first-hit class (rank -> columns) for the finalizer.
__global__ __launch_bounds__(64) void plain_finalize_kernel(const DevCode *__restrict__ codep, const DevCall *__restrict__ callp, int shot,
                                                            ShotOutcome o, const int *gA, unsigned long long hit_rank) {
    const DevCode &code = *codep;
    const DevCall &call = *callp;
    if (o.found && o.win[0] < 0) {
        int c[kMaxW];
        unrank_dyn(hit_rank, o.tam, o.fw, code.nCr, c);
#pragma unroll
        for (int l = 0; l < kMaxW; ++l) o.win[l] = (l < o.fw) ? gA[c[l]] : -1;
    }
    if (o.found) {
#pragma unroll
        for (int a = 0; a < kMaxWin; ++a)
#pragma unroll
            for (int b = 0; b < kMaxWin - 1 - a; ++b)
                if (b + 1 < o.fw && o.win[b] > o.win[b + 1]) { const int t = o.win[b]; o.win[b] = o.win[b + 1]; o.win[b + 1] = t; }
    }
    finalize_shot(code, call, shot, o, nullptr);
}

// Expands packed words to the ABI's int32 0/1 rows.
__global__ void expand_bits_kernel(const uint32_t *packed, int words, int cols, int rows, int *out) {
    const size_t total = (size_t)rows * cols;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / cols;
        const int c = (int)(i - r * cols);
        out[i] = (int)((packed[r * words + (c >> 5)] >> (c & 31)) & 1u);
    }
}

// Installs a batch's descriptor (passed by value in the kernel arguments) in
// device memory and clears the batch counters, in one launch.
static_assert(sizeof(DevCall) % 4 == 0 && sizeof(DevCall) <= 3072, "DevCall must fit the kernel arguments");
__global__ void prep_kernel(const DevCall c, DevCall *dst, unsigned *counters) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&c);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    for (int i = threadIdx.x; i < (int)(sizeof(DevCall) / 4); i += blockDim.x) d[i] = src[i];
    if (threadIdx.x < 16) counters[threadIdx.x] = 0u;
    unsigned long long *claim_words = reinterpret_cast<unsigned long long *>(counters) + kClaimBase;
    if (threadIdx.x < kClaimWords) claim_words[threadIdx.x * kClaimStride] = 0ull;
    if (threadIdx.x < 2 * kWlParts) counters[kWlBase + kWlStride * (threadIdx.x >> 1) + (threadIdx.x & 1)] = 0u;
}

__global__ void fill_u32_kernel(unsigned *p, unsigned v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace grand

// ----------------------------------------------------------- launch wrappers
// `code`/`call` are device copies of the descriptors; `hcall` is the host copy
// (launch geometry only).
extern "C" {
using grand::DevCall;
using grand::DevCode;

// LDS of one wave without the staged code and claimed shots (carve_lds).
GRAND_INTERNAL size_t grand_fused_lds_bytes(int cap, int ml) {
    return sizeof(uint32_t) * ((size_t)grand::kLdsCodeOffset + 2 * (size_t)grand::ht_capacity(cap) +
                               (size_t)grand::bitmap_words(cap) + (size_t)cap * (ml ? 6 : 2));
}

GRAND_INTERNAL hipError_t grand_launch_prep(const DevCall *hcall, DevCall *dcall, unsigned *counters, hipStream_t st) {
    hipLaunchKernelGGL(grand::prep_kernel, dim3(1), dim3(256), 0, st, *hcall, dcall, counters);
    return hipGetLastError();
}

GRAND_INTERNAL hipError_t grand_launch_fill(unsigned *p, unsigned v, int n, hipStream_t st) {
    hipLaunchKernelGGL(grand::fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// `block_words`: words per shot of the triage kernel's packed block, the larger
// of the input row (M32 words of errors or D words of syndrome) and, for the DEM
// family, the NL32 words of observed logicals; `stage_words`: the code tables it
// stages in LDS (DevCall::triage_stage_words).
GRAND_INTERNAL hipError_t grand_launch_triage(const DevCode *code, const DevCall *call, int grid, int block_words,
                                              int stage_words, hipStream_t st) {
    hipLaunchKernelGGL(grand::triage_kernel, dim3(grid), dim3(64),
                       sizeof(uint32_t) * (2 * (grand::kLaneMax + 1) * 64 + grand::kTriageWlWords + (size_t)stage_words +
                                           64 * (size_t)block_words),
                       st, code, call);
    return hipGetLastError();
}

GRAND_INTERNAL hipError_t grand_launch_fused(const DevCode *code, const DevCall *call, const DevCall *hcall, int code_lds_words,
                              int grid, hipStream_t st) {
    const size_t lds = grand_fused_lds_bytes(hcall->cap, hcall->kind == grand::kML || hcall->lscores != nullptr) +
                       sizeof(uint32_t) * ((size_t)((code_lds_words + 3) & ~3) + (size_t)hcall->claim_batch * hcall->stage_words +
                                           (hcall->pair_cap > 0 ? (size_t)grand::kPairBitmapWords + grand::pair_head_words(hcall->kind == grand::kML) +
                                                                      2 * (size_t)hcall->pair_cap
                                                                : 0));
    if (hcall->kind == grand::kML)
        hipLaunchKernelGGL(grand::fused_kernel<true>, dim3(grid), dim3(64), lds, st, code, call);
    else
        hipLaunchKernelGGL(grand::fused_kernel<false>, dim3(grid), dim3(64), lds, st, code, call);
    return hipGetLastError();
}

GRAND_INTERNAL hipError_t grand_launch_expand(const uint32_t *packed, int words, int cols, int rows, int *out, hipStream_t st) {
    const size_t total = (size_t)rows * cols;
    int grid = (int)((total + 255) / 256);
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(grand::expand_bits_kernel, dim3(grid), dim3(256), 0, st, packed, words, cols, rows, out);
    return hipGetLastError();
}

// Fixed device scratch of grand_session_decode_device_noalloc: a file-scope
// array every GPU holds (its own copy, at its own address) until the process
// ends.  Host code never uses the array's name as a device address: its
// address on the current device comes from hipGetSymbolAddress, resolved once
// per device and kept, and only that address is passed to kernels and fills.
// The task table sits at offset 0 and is cleared here, once per device, with
// a fill kernel on `st`; every decode leaves it idle.  Every call on one GPU
// shares the array, so calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[grand::kScratchBytes];

GRAND_INTERNAL hipError_t grand_scratch_address(void **ptr, hipStream_t st) {
    static std::mutex mu;
    static void *addr[grand::kScratchDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= grand::kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        const int words = (int)(sizeof(grand::TaskSlot) * grand::kTaskSlots / 4);
        hipLaunchKernelGGL(grand::fill_u32_kernel, dim3((words + 255) / 256), dim3(256), 0, st, (unsigned *)p, 0u, words);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

GRAND_INTERNAL size_t grand_plain_state_bytes() { return sizeof(grand::PlainState); }
GRAND_INTERNAL size_t grand_plain_record_bytes() { return sizeof(grand::MlRecord); }

GRAND_INTERNAL hipError_t grand_launch_plain_active(const DevCode *code, const DevCall *call, int cap, int shot, int thr,
                                     int sort_tam, void *scratch, void *state, hipStream_t st) {
    int *gA = (int *)scratch;
    uint32_t *gsig = (uint32_t *)(gA + cap);
    float *gllr = (float *)(gsig + cap);
    int *gP = (int *)(gllr + cap);
    uint32_t *gsigP = (uint32_t *)(gP + cap);
    float *gllrP = (float *)(gsigP + cap);
    hipLaunchKernelGGL(grand::plain_active_kernel, dim3(1), dim3(64), grand_fused_lds_bytes(cap, 1), st, code, call,
                       shot, thr, sort_tam, gA, gsig, gllr, gP, gsigP, gllrP, (grand::PlainState *)state);
    return hipGetLastError();
}

GRAND_INTERNAL hipError_t grand_launch_plain_class(const DevCode *code, const DevCall *call, int cap, int shot, int w, int tam, int R,
                                    unsigned long long nchunks, int grid, int ml, void *scratch, void *state,
                                    void *records, hipStream_t st) {
    int *gA = (int *)scratch;
    uint32_t *gsig = (uint32_t *)(gA + cap);
    float *gllr = (float *)(gsig + cap);
    int *gP = (int *)(gllr + cap);
    uint32_t *gsigP = (uint32_t *)(gP + cap);
    float *gllrP = (float *)(gsigP + cap);
    hipLaunchKernelGGL(grand::plain_class_kernel, dim3(grid), dim3(64), grand_fused_lds_bytes(cap, 1), st, code, call,
                       shot, w, tam, R, nchunks, ml, ml ? gP : gA, ml ? gsigP : gsig, gllrP,
                       (grand::PlainState *)state, (grand::MlRecord *)records);
    return hipGetLastError();
}

GRAND_INTERNAL hipError_t grand_launch_plain_ml_reduce(void *records, int nrec, void *out, hipStream_t st) {
    hipLaunchKernelGGL(grand::plain_ml_reduce_kernel, dim3(1), dim3(64), 0, st, (const grand::MlRecord *)records, nrec,
                       (grand::MlRecord *)out);
    return hipGetLastError();
}

GRAND_INTERNAL hipError_t grand_launch_plain_finalize(const DevCode *code, const DevCall *call, int shot,
                                       const grand::ShotOutcome *o, void *scratch, unsigned long long hit_rank,
                                       hipStream_t st) {
    hipLaunchKernelGGL(grand::plain_finalize_kernel, dim3(1), dim3(64), 0, st, code, call, shot, *o,
                       (const int *)scratch, hit_rank);
    return hipGetLastError();
}
}
