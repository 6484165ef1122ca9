// grand_host.cpp -- host side of libgranddecoder.so: the reference C ABI, the
// device-resident session API, code packing/caching and the two engines
// (fused persistent kernel, plain one-launch-per-weight-class loop).
//
// Validation order and messages follow each reference driver:
//   grand_decode_prob_cc  grand_decoder.cu:833-869
//   grand_decode_cc       grand_decoder.cu:1155-1180
//   grand_decode_cln      grand_decoder.cu:1440-1446
//   grand_decode_prob_cln grand_decoder.cu:1713-1746
//   grand_decode_ML_cln   grand_decoder.cu:2039-2053
// followed by this library's own limits (n <= 512, m <= 4096, 0/1 matrices,
// finite LLRs), each reported on stderr with a -1 return.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "grand_types.h"
#include "../../include/grand_session.h"

extern "C" {
GRAND_INTERNAL size_t grand_fused_lds_bytes(int cap, int ml);
GRAND_INTERNAL hipError_t grand_launch_fill(unsigned *p, unsigned v, int n, hipStream_t st);
GRAND_INTERNAL hipError_t grand_launch_prep(const grand::DevCall *hcall, grand::DevCall *dcall, unsigned *counters, hipStream_t st);
GRAND_INTERNAL hipError_t grand_launch_triage(const grand::DevCode *code, const grand::DevCall *call, int grid,
                                              int block_words, int stage_words, hipStream_t st);
GRAND_INTERNAL hipError_t grand_launch_fused(const grand::DevCode *code, const grand::DevCall *call, const grand::DevCall *hcall,
                              int code_lds_words, int grid, hipStream_t st);
GRAND_INTERNAL hipError_t grand_launch_expand(const uint32_t *packed, int words, int cols, int rows, int *out, hipStream_t st);
GRAND_INTERNAL hipError_t grand_scratch_address(void **ptr, hipStream_t st);
GRAND_INTERNAL size_t grand_plain_state_bytes();
GRAND_INTERNAL size_t grand_plain_record_bytes();
GRAND_INTERNAL hipError_t grand_launch_plain_active(const grand::DevCode *code, const grand::DevCall *call, int cap, int shot,
                                     int thr, int sort_tam, void *scratch, void *state, hipStream_t st);
GRAND_INTERNAL hipError_t grand_launch_plain_class(const grand::DevCode *code, const grand::DevCall *call, int cap, int shot, int w,
                                    int tam, int R, unsigned long long nchunks, int grid, int ml, void *scratch,
                                    void *state, void *records, hipStream_t st);
GRAND_INTERNAL hipError_t grand_launch_plain_ml_reduce(void *records, int nrec, void *out, hipStream_t st);
GRAND_INTERNAL hipError_t grand_launch_plain_finalize(const grand::DevCode *code, const grand::DevCall *call, int shot,
                                       const grand::ShotOutcome *o, void *scratch, unsigned long long hit_rank,
                                       hipStream_t st);
}

namespace grand {
namespace {

bool hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "grand: HIP error in %s: %s\n", what, hipGetErrorString(e));
    return false;
}
#define HIPX(expr) do { if (!hip_check((expr), #expr)) return -1; } while (0)

// ------------------------------------------------------------------ runtime
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem() { if (p) (void)hipFree(p); }
    bool ensure(size_t b) {
        if (b <= bytes) return true;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        size_t want = std::max<size_t>(b, 256);
        if (!hip_check(hipMalloc(&p, want), "hipMalloc")) return false;
        bytes = want;
        return true;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Runtime {
    bool ok = false;
    int device = 0;
    int cu_count = 256;
    double ms_per_tick = 1e-5;
    hipStream_t stream = nullptr;
    DevMem nCr;
    std::vector<unsigned long long> nCr_host;
    std::recursive_mutex mu;
};

Runtime &rt() {
    static Runtime r;
    static std::once_flag once;
    std::call_once(once, [] {
        r.nCr_host.assign((size_t)(kMaxTam + 1) * kNcrStride, 0ull);
        for (int n = 0; n <= kMaxTam; ++n)
            for (int k = 0; k < kNcrStride; ++k) {
                if (k > n) continue;
                unsigned long long v = 1;
                for (int i = 1; i <= k; ++i) v = v * (unsigned long long)(n - i + 1) / (unsigned long long)i;
                r.nCr_host[(size_t)n * kNcrStride + k] = v;
            }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
            fprintf(stderr, "grand: no HIP device available\n");
            return;
        }
        if (!hip_check(hipGetDevice(&r.device), "hipGetDevice")) return;
        hipDeviceProp_t prop;
        if (!hip_check(hipGetDeviceProperties(&prop, r.device), "hipGetDeviceProperties")) return;
        r.cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, r.device) == hipSuccess && khz > 0)
            r.ms_per_tick = 1.0 / (double)khz;
        if (!hip_check(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking), "hipStreamCreate")) return;
        if (!r.nCr.ensure(r.nCr_host.size() * sizeof(unsigned long long))) return;
        if (!hip_check(hipMemcpy(r.nCr.p, r.nCr_host.data(), r.nCr_host.size() * sizeof(unsigned long long),
                                 hipMemcpyHostToDevice), "hipMemcpy nCr")) return;
        r.ok = true;
    });
    return r;
}

uint64_t ncr_host(int n, int k) {
    if (n < 0 || k < 0 || k > n || n > kMaxTam || k >= kNcrStride) return 0;
    return rt().nCr_host[(size_t)n * kNcrStride + k];
}

// ----------------------------------------------------------------- file readers
struct Masks {
    int n = 0;
    int tam_upper[kMaxMasks];
    int order[kMaxMasks][3];
};

// grand_decoder.cu:728-752: up to 20 lines of "%d %d %d %d %f %f %f".
int read_masks(const char *path, Masks &mk) {
    mk.n = 0;
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) { fprintf(stderr, "Error opening probability file: %s\n", path ? path : "(null)"); return 0; }
    while (mk.n < kMaxMasks) {
        int t, a, b, c;
        float p0, p1, p2;
        if (fscanf(f, "%d %d %d %d %f %f %f", &t, &a, &b, &c, &p0, &p1, &p2) != 7) break;
        mk.tam_upper[mk.n] = t;
        mk.order[mk.n][0] = a; mk.order[mk.n][1] = b; mk.order[mk.n][2] = c;
        ++mk.n;
    }
    fclose(f);
    return mk.n;
}

// grand_decoder.cu:695-713: m floats.
bool read_llr(const char *path, std::vector<float> &llr, int m) {
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) { fprintf(stderr, "Error opening LLR file: %s\n", path ? path : "(null)"); return false; }
    llr.assign(m, 0.0f);
    for (int i = 0; i < m; ++i)
        if (fscanf(f, "%f", &llr[i]) != 1) {
            fprintf(stderr, "Error reading LLR file (expected %d values): %s\n", m, path);
            fclose(f);
            return false;
        }
    fclose(f);
    return true;
}

// ------------------------------------------------------------------- hashing
struct Hasher {
    uint64_t h = 0x6a09e667f3bcc908ull;
    void mix(uint64_t v) {
        h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 33;
    }
    void ints(const int *p, size_t n) {
        size_t i = 0;
        for (; i + 2 <= n; i += 2) mix((uint64_t)(uint32_t)p[i] | ((uint64_t)(uint32_t)p[i + 1] << 32));
        if (i < n) mix((uint64_t)(uint32_t)p[i]);
        mix(n);
    }
};

bool all_binary(const int *p, size_t n) {
    int acc = 0;
    for (size_t i = 0; i < n; ++i) acc |= p[i] & ~1;
    return acc == 0;
}

// ------------------------------------------------------------------ code upload
uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Degeneracy words of a CSS pair.  For a found shot, v = e ^ e_hat lies in
// ker(H) (both have syndrome H e), and v is in rowspace(other) iff v is
// orthogonal to ker(other).  A vector of ker(other) that lies in rowspace(H) +
// span(earlier kept ones) adds nothing for v in ker(H), so only the ones
// independent of rowspace(H) are kept: as many as the code's logical qubits for
// a CSS code.  With at most 32 of them, dg[c] holds their bits at column c and
// v is in rowspace(other) iff the XOR of dg over supp(e) equals the XOR over
// supp(e_hat).  Hrow: rows of H packed [n][M32]; R, piv: RREF of the other
// matrix (rows [rank][M32] and pivot columns).  Returns the number of kept
// vectors and fills dg, or returns -1 (dg empty) when more than 32 are needed.
int degeneracy_words(const std::vector<uint32_t> &Hrow, const std::vector<uint32_t> &R, const std::vector<int> &piv,
                     int n, int m, std::vector<uint32_t> &dg) {
    const int M32 = (m + 31) / 32;
    const int rank = (int)piv.size();
    dg.clear();
    std::vector<std::vector<uint32_t>> basis;     // reduced: each pivot column set in its own vector only
    std::vector<int> bpiv;
    auto bit = [](const std::vector<uint32_t> &v, int c) { return (v[c >> 5] >> (c & 31)) & 1u; };
    auto insert = [&](std::vector<uint32_t> x) {
        for (size_t i = 0; i < basis.size(); ++i)
            if (bit(x, bpiv[i])) for (int j = 0; j < M32; ++j) x[j] ^= basis[i][j];
        int p = -1;
        for (int j = 0; j < M32 && p < 0; ++j) if (x[j]) p = j * 32 + __builtin_ctz(x[j]);
        if (p < 0) return false;
        for (size_t i = 0; i < basis.size(); ++i)
            if (bit(basis[i], p)) for (int j = 0; j < M32; ++j) basis[i][j] ^= x[j];
        basis.push_back(std::move(x));
        bpiv.push_back(p);
        return true;
    };
    for (int r = 0; r < n; ++r)
        (void)insert(std::vector<uint32_t>(Hrow.begin() + (size_t)r * M32, Hrow.begin() + (size_t)(r + 1) * M32));
    std::vector<char> is_piv(m, 0);
    for (int r = 0; r < rank; ++r) is_piv[piv[r]] = 1;
    std::vector<std::vector<uint32_t>> kept;
    for (int f = 0; f < m; ++f) {
        if (is_piv[f]) continue;
        std::vector<uint32_t> k(M32, 0);          // null vector of the RREF: e_f plus the pivots of rows with bit f
        k[f >> 5] |= 1u << (f & 31);
        for (int r = 0; r < rank; ++r)
            if ((R[(size_t)r * M32 + (f >> 5)] >> (f & 31)) & 1u) k[piv[r] >> 5] |= 1u << (piv[r] & 31);
        if (insert(k)) {
            if (kept.size() == 32) return -1;
            kept.push_back(std::move(k));
        }
    }
    dg.assign(m, 0u);
    for (size_t j = 0; j < kept.size(); ++j)
        for (int c = 0; c < m; ++c) if (bit(kept[j], c)) dg[c] |= 1u << j;
    return (int)kept.size();
}

// RREF of rows [n][M32] over GF(2) (grand_decoder.cu:418-454): the reduced rows
// [rank][M32] and their pivot columns.
void gf2_rref(std::vector<uint32_t> rows, int n, int m, std::vector<uint32_t> &R, std::vector<int> &piv) {
    const int M32 = (m + 31) / 32;
    piv.clear();
    int rr = 0;
    for (int c = 0; c < m && rr < n; ++c) {
        const int wd = c >> 5;
        const uint32_t bit = 1u << (c & 31);
        int p = -1;
        for (int i = rr; i < n; ++i) if (rows[(size_t)i * M32 + wd] & bit) { p = i; break; }
        if (p < 0) continue;
        if (p != rr) for (int j = 0; j < M32; ++j) std::swap(rows[(size_t)rr * M32 + j], rows[(size_t)p * M32 + j]);
        for (int i = 0; i < n; ++i)
            if (i != rr && (rows[(size_t)i * M32 + wd] & bit))
                for (int j = 0; j < M32; ++j) rows[(size_t)i * M32 + j] ^= rows[(size_t)rr * M32 + j];
        piv.push_back(c);
        ++rr;
    }
    R.assign(rows.begin(), rows.begin() + (size_t)rr * M32);
}

struct CodeEntry {
    uint64_t key = 0;
    DevCode dev{};
    DevMem mem;
    DevMem dcode;                     // device copy of `dev` (lds_words = 0)
    DevMem dcode_search;              // the search kernel's copy (lds_words > 0 when the code is staged in LDS)
    int family = kCC;
    int n = 0, m = 0, NL = 0;
    int search_lds_words = 0;         // dcode_search's lds_words
    size_t bytes = 0;
    // the uploaded content, compared on every cache hit (the key is a 64-bit
    // hash): the searched matrix, the other CSS matrix (cc) or L (cln), LLRs
    std::vector<int> src_h, src_other;
    std::vector<float> src_llr;
    bool same(int family_, int n_, int m_, int NL_, const int *H, const int *other, size_t other_len,
              const std::vector<float> &llr) const {
        if (family_ != family || n_ != n || m_ != m || NL_ != NL) return false;
        if (src_other.size() != other_len || src_llr.size() != llr.size()) return false;
        if (memcmp(src_h.data(), H, src_h.size() * sizeof(int)) != 0) return false;
        if (other_len && memcmp(src_other.data(), other, other_len * sizeof(int)) != 0) return false;
        return llr.empty() || memcmp(src_llr.data(), llr.data(), llr.size() * sizeof(float)) == 0;
    }
};

// Packs H (n x m), the RREF of the other CSS matrix (cc) or L (cln), and LLRs.
std::shared_ptr<CodeEntry> upload_code(uint64_t key, int family, const int *H, const int *other, const int *L, int n,
                                       int m, int NL, const std::vector<float> *llr) {
    auto ce = std::make_shared<CodeEntry>();
    ce->key = key; ce->family = family; ce->n = n; ce->m = m; ce->NL = NL;
    const int D = (n + 31) / 32, M32 = (m + 31) / 32, NL32 = std::max(1, (NL + 31) / 32);

    std::vector<uint32_t> Hcw((size_t)D * m, 0), Hcm((size_t)m * D, 0), Hrow((size_t)n * M32, 0);
    std::vector<int> colw(m, 0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < m; ++c)
            if (H[(size_t)r * m + c]) {
                Hcw[(size_t)(r >> 5) * m + c] |= 1u << (r & 31);
                Hcm[(size_t)c * D + (r >> 5)] |= 1u << (r & 31);
                Hrow[(size_t)r * M32 + (c >> 5)] |= 1u << (c & 31);
                ++colw[c];
            }
    // grand_decoder.cu:376-416: max column weight, regularity, score offset
    int max_soma = 0;
    for (int c = 0; c < m; ++c) max_soma = std::max(max_soma, colw[c]);
    bool regular = true;
    for (int c = 1; c < m; ++c) if (colw[c] != colw[0]) { regular = false; break; }
    std::vector<int> off(m, 0);
    for (int c = 0; c < m; ++c)
        off[c] = (colw[c] == 0) ? kNegScore : (regular ? 0 : max_soma - colw[c]);

    // Signatures: distinct nonzero random words per row, XORed per column.
    std::vector<uint32_t> rowsig(n), sig(m, 0);
    uint64_t seed = 0x243f6a8885a308d3ull;
    for (int r = 0; r < n; ++r) {
        for (;;) {
            const uint32_t v = (uint32_t)(splitmix64(seed) >> 32);
            bool dup = (v == 0);
            for (int q = 0; q < r && !dup; ++q) dup = (rowsig[q] == v);
            if (!dup) { rowsig[r] = v; break; }
        }
    }
    for (int c = 0; c < m; ++c) {
        uint32_t v = 0;
        for (int d = 0; d < D; ++d) {
            uint32_t w = Hcm[(size_t)c * D + d];
            while (w) { const int b = __builtin_ctz(w); v ^= rowsig[d * 32 + b]; w &= w - 1; }
        }
        sig[c] = v;
    }

    // RREF of the other CSS matrix over GF(2) (grand_decoder.cu:418-454); the
    // degenerate test only depends on its row space.
    std::vector<uint32_t> R;
    std::vector<int> piv;
    int rank = 0;
    if (family == kCC) {
        std::vector<uint32_t> rows((size_t)n * M32, 0);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < m; ++c)
                if (other[(size_t)r * m + c]) rows[(size_t)r * M32 + (c >> 5)] |= 1u << (c & 31);
        gf2_rref(std::move(rows), n, m, R, piv);
        rank = (int)piv.size();
    }
    std::vector<int> pivrow(m, -1);
    for (int r = 0; r < rank; ++r) pivrow[piv[r]] = r;
    std::vector<uint32_t> dg;
    if (family == kCC) (void)degeneracy_words(Hrow, R, piv, n, m, dg);
    std::vector<uint32_t> Lp;
    if (family == kCLN) {
        Lp.assign((size_t)m * NL32, 0);
        for (int c = 0; c < m; ++c)
            for (int q = 0; q < NL; ++q)
                if (L[(size_t)c * NL + q]) Lp[(size_t)c * NL32 + (q >> 5)] |= 1u << (q & 31);
    }

    // The triage kernel's tables as one image (staged in its LDS when small): column words [D][m], signatures [m], degeneracy words [m]
    // (when there are any); then, 16-byte aligned, the rows of H [n][t_rw] with
    // the odd stride t_rw = M32 | 1 (lanes reading the rows of different checks
    // then spread over the LDS banks), and the column-weight planes
    // [7][M32]: the nonempty columns, the three bit planes of colw, the three
    // bit planes of colw - 1.  s = H e uses the first part; the lane search
    // counts the checks of s in every column as the bit-sliced
    // sum of the rows of H at s's checks, and a column is active at threshold
    // max_soma - j iff its count is within j of colw.
    std::vector<uint32_t> timg;
    int t_sig, t_dg, t_rows, t_rw, t_pl, lane_ok;
    {
        int maxw = 0;
        for (int c = 0; c < m; ++c) maxw = std::max(maxw, colw[c]);
        lane_ok = (D <= kLaneMaxD && M32 <= kLaneMaxM32 && maxw <= kLaneMaxColw) ? 1 : 0;
        t_rw = M32 | 1;
        timg.assign(Hcw.begin(), Hcw.end());
        t_sig = (int)timg.size();
        timg.insert(timg.end(), sig.begin(), sig.end());
        t_dg = (int)timg.size();
        timg.insert(timg.end(), dg.begin(), dg.end());
        timg.resize((timg.size() + 3) & ~(size_t)3, 0u);
        t_rows = (int)timg.size();
        timg.resize(timg.size() + (size_t)n * t_rw, 0u);
        for (int r = 0; r < n; ++r)
            for (int wd = 0; wd < M32; ++wd) timg[(size_t)t_rows + (size_t)r * t_rw + wd] = Hrow[(size_t)r * M32 + wd];
        t_pl = (int)timg.size();
        timg.resize(timg.size() + (size_t)7 * M32, 0u);
        uint32_t *pl = timg.data() + t_pl;
        for (int c = 0; c < m; ++c) {
            if (colw[c] == 0) continue;
            const uint32_t bit = 1u << (c & 31);
            const int wd = c >> 5;
            pl[wd] |= bit;
            for (int b = 0; b < 3; ++b) {
                if ((colw[c] >> b) & 1) pl[(1 + b) * M32 + wd] |= bit;
                if (((colw[c] - 1) >> b) & 1) pl[(4 + b) * M32 + wd] |= bit;
            }
        }
        // zero words after the last table: the lane search reads rows and
        // planes kLaneMaxM32 words wide (the words past M32 are not used)
        timg.resize((timg.size() + kLaneMaxM32 + 2 + 3) & ~(size_t)3, 0u);
    }

    // One allocation, 256-byte aligned segments.
    struct Seg { const void *src; size_t bytes; size_t off; };
    std::vector<Seg> segs;
    size_t total = 0;
    auto add = [&](const void *src, size_t bytes) {
        segs.push_back({src, bytes, total});
        total += (bytes + 255) & ~(size_t)255;
        return segs.size() - 1;
    };
    const size_t iHcw = add(Hcw.data(), Hcw.size() * 4), iHcm = add(Hcm.data(), Hcm.size() * 4),
                 iHrow = add(Hrow.data(), Hrow.size() * 4), iSig = add(sig.data(), sig.size() * 4),
                 iRs = add(rowsig.data(), rowsig.size() * 4), iOff = add(off.data(), off.size() * 4),
                 iR = add(R.data(), R.size() * 4), iPiv = add(pivrow.data(), pivrow.size() * 4),
                 iL = add(Lp.data(), Lp.size() * 4), iDg = add(dg.data(), dg.size() * 4),
                 iTimg = add(timg.data(), timg.size() * 4),

                 iLlr = add(llr ? llr->data() : nullptr, llr ? llr->size() * 4 : 0);
    if (!ce->mem.ensure(total)) return nullptr;
    char *base = ce->mem.as<char>();
    for (const Seg &s : segs)
        if (s.bytes && !hip_check(hipMemcpy(base + s.off, s.src, s.bytes, hipMemcpyHostToDevice), "hipMemcpy code"))
            return nullptr;
    ce->bytes = total;
    DevCode &dc = ce->dev;
    dc.n = n; dc.m = m; dc.D = D; dc.M32 = M32; dc.NL = NL; dc.NL32 = NL32;
    dc.max_soma = max_soma;
    dc.rank = rank;
    dc.Hcw = (const uint32_t *)(base + segs[iHcw].off);
    dc.Hcm = (const uint32_t *)(base + segs[iHcm].off);
    dc.Hrow = (const uint32_t *)(base + segs[iHrow].off);
    dc.sig = (const uint32_t *)(base + segs[iSig].off);
    dc.rowsig = (const uint32_t *)(base + segs[iRs].off);
    dc.off = (const int *)(base + segs[iOff].off);
    dc.R = (const uint32_t *)(base + segs[iR].off);
    dc.pivrow = (const int *)(base + segs[iPiv].off);
    dc.Lp = (const uint32_t *)(base + segs[iL].off);
    dc.dg = dg.empty() ? nullptr : (const uint32_t *)(base + segs[iDg].off);
    dc.timg = (const uint32_t *)(base + segs[iTimg].off);
    dc.timg_words = (int)timg.size();
    dc.t_sig = t_sig; dc.t_dg = t_dg; dc.t_rows = t_rows; dc.t_rw = t_rw; dc.t_pl = t_pl;
    dc.lane_ok = lane_ok;

    dc.llr = (const float *)(base + segs[iLlr].off);
    dc.nCr = rt().nCr.as<unsigned long long>();
    dc.lds_words = 0;
    dc.lds_post = 0;
    if (!ce->dcode.ensure(sizeof(DevCode)) ||
        !hip_check(hipMemcpy(ce->dcode.p, &dc, sizeof(DevCode), hipMemcpyHostToDevice), "hipMemcpy DevCode"))
        return nullptr;
    // the search kernel stages the column words, offsets and signatures of a
    // small code in LDS, and the degeneracy data (degeneracy words, or pivot
    // rows and RREF rows) too when they also fit
    DevCode ds = dc;
    const long long cw = (long long)(D + 2) * m;
    const long long post = (family != kCC) ? 0 : dc.dg ? (long long)m : (long long)m + (long long)rank * M32;
    if (cw + post <= kCodeLdsMaxWords) {
        ds.lds_words = (int)(cw + post);
        ds.lds_post = post == 0 ? 0 : dc.dg ? 2 : 1;
    } else if (cw <= kCodeLdsMaxWords) {
        ds.lds_words = (int)cw;
    }
    if (!ce->dcode_search.ensure(sizeof(DevCode)) ||
        !hip_check(hipMemcpy(ce->dcode_search.p, &ds, sizeof(DevCode), hipMemcpyHostToDevice), "hipMemcpy DevCode"))
        return nullptr;
    ce->search_lds_words = ds.lds_words;
    return ce;
}

// Small LRU of uploaded codes, keyed by content hash.
std::list<std::shared_ptr<CodeEntry>> g_codes;

std::shared_ptr<CodeEntry> find_code(uint64_t key, int family, int n, int m, int NL, const int *H, const int *other,
                                     size_t other_len, const std::vector<float> &llr) {
    for (auto it = g_codes.begin(); it != g_codes.end(); ++it)
        if ((*it)->key == key && (*it)->same(family, n, m, NL, H, other, other_len, llr)) {
            auto e = *it;
            g_codes.erase(it);
            g_codes.push_front(e);
            return e;
        }
    return nullptr;
}

void remember_code(const std::shared_ptr<CodeEntry> &e) {
    g_codes.push_front(e);
    while (g_codes.size() > 16) g_codes.pop_back();
}

}  // namespace

// ================================================================= sessions
struct Workspace {
    DevMem s_packed, e_packed, e_aux, lg_packed, s_sig, worklist, counters, dec, declog, tasks, task_arrays;
    int shots = 0;
    bool tasks_fresh = false;   // the task table was just allocated and must be cleared once
    bool ensure(const DevCode &c, int shots_, int cap) {
        const size_t S = (size_t)std::max(shots_, 1);
        if (tasks.bytes == 0) tasks_fresh = true;
        if (!s_packed.ensure(S * c.D * 4) || !e_packed.ensure(S * c.M32 * 4) || !e_aux.ensure(S * 8) ||
            !lg_packed.ensure(S * c.NL32 * 4) ||
            !s_sig.ensure(S * 4) || !worklist.ensure((S + 63) / 64 * 64 * 4) || !counters.ensure(kCounterBytes) ||
            !dec.ensure(S * c.M32 * 4) ||
            !declog.ensure(S * c.NL32 * 4) || !tasks.ensure(sizeof(TaskSlot) * kTaskSlots) ||
            !task_arrays.ensure((size_t)kTaskSlots * 3 * std::max(cap, 1) * 4))
            return false;
        shots = shots_;
        return true;
    }
    size_t bytes_for(const DevCode &c, int shots_, int cap) const {
        const size_t S = (size_t)std::max(shots_, 1);
        return S * 4 * ((size_t)c.D + 2 * c.M32 + 2 * c.NL32 + 4) + 256 + kCounterBytes + sizeof(TaskSlot) * kTaskSlots +
               (size_t)kTaskSlots * 3 * std::max(cap, 1) * 4;
    }
};

}  // namespace grand

struct GrandSession {
    int kind = 0;
    int family = grand::kCC;
    int search = grand::kUniform;
    std::shared_ptr<grand::CodeEntry> code;
    grand::DevCall call{};
    GrandDecoderExConfig ex{};
    grand::Workspace ws;
    grand::DevMem plain_scratch, plain_state, plain_records, plain_out;
    grand::DevMem dcall;              // device copy of `call`, refreshed before every batch
    // HIP-event timing: a ring of kTimingSlots event triples (before triage,
    // after triage, after search), one per decode.  A slot's times are read
    // (waiting for its events) when the slot is reused or the times are asked
    // for, so decodes with timing on still run back to back.
    static constexpr int kTimingSlots = 64;
    bool timing = false;
    std::vector<hipEvent_t> ev;       // [kTimingSlots * 3]
    int tslot = -1;                   // slot of the last timed decode
    unsigned long long tpending = 0;  // bit per slot whose times are not read yet
    float slot_ms[kTimingSlots][2] = {};
    double sum_ms[2] = {0.0, 0.0};
    long long sum_n = 0;
    ~GrandSession() {
        for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
    }
    // reads slot k's times into slot_ms and the sums
    bool harvest(int k) {
        if (!((tpending >> k) & 1ull)) return true;
        if (hipEventSynchronize(ev[3 * k + 2]) != hipSuccess) return false;
        float a = 0.0f, b = 0.0f;
        if (hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]) != hipSuccess ||
            hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]) != hipSuccess)
            return false;
        slot_ms[k][0] = a; slot_ms[k][1] = b;
        sum_ms[0] += a; sum_ms[1] += b; ++sum_n;
        tpending &= ~(1ull << k);
        return true;
    }
};

namespace grand {
namespace {

int kind_family(int kind) { return (kind <= GRAND_KIND_PROB_CC) ? kCC : kCLN; }
int kind_search(int kind) {
    switch (kind) {
        case GRAND_KIND_PROB_CC: case GRAND_KIND_PROB_CLN: return kPhase;
        case GRAND_KIND_ML_CLN: return kML;
        default: return kUniform;
    }
}

// Tuning knobs: built-in defaults per search kind, replaced for every kind by
// grand_set_tuning (the autotuner in the Python package), overridable per
// session.  First-hit kinds: 1 wave per CU.  A wave's LDS share holds the
// pair table of lists up to 81 columns, and with the first two iterations of
// most shots searched by lane in the triage kernel, the search kernel's batch
// is a few hundred shots: one wave per CU decodes them as fast as two and
// leaves the CUs' other SIMDs and LDS to the next batch's triage on another
// stream (BB144 prob_cc, 100k shots, per decode: 146.6 us on one stream and
// 84.7 us on two, against 150.9 and 120.0 us at 2 waves per CU).  Runs are
// shared above 65536 prefix steps of remaining work (144.4 and 82.3 us; 32768:
// 146.6 and 85.5 us; 131072: 171.7 and 92.5 us).  ML: 4 waves
// per CU, whose LDS share holds the pair table the deep classes are walked
// by, one shot per claim, runs shared above 65536 prefix steps (the DEM ML
// workload, 1M shots, with 2048 pair bucket heads: 15.5 M shots/s; claim batch
// 2 and split 32768: 14.5 M;
// 8 waves per CU, where the table does not fit: 13.2 M).
// The pair-table LDS (pair_table_kb: 0 = the wave's whole remaining share,
// -1 = none, k = at most k KiB) applies to both.
struct Tuning { int waves_per_cu; int claim_batch; int rmax; int split_threshold; int pair_table_kb; };
constexpr Tuning kTuningFirstHit{1, 2, 64, 65536, 0};
constexpr Tuning kTuningMl{4, 1, 64, 65536, 0};
Tuning g_tuning = kTuningFirstHit;
Tuning g_tuning_ml = kTuningMl;
const Tuning &tuning_for(int search) { return search == kML ? g_tuning_ml : g_tuning; }

// Validates exactly as the reference driver for `kind` would, then applies the
// library's own limits, and creates a session (code upload cached by content).
GrandSession *create_session(int kind, const int *H, const int *H2, int n, int m, int NL, char type,
                             const GrandDecoderConfig *cfg, const char *file, const GrandDecoderExConfig *ex) {
    if (!cfg) { fprintf(stderr, "grand: config is NULL\n"); return nullptr; }
    const int family = kind_family(kind), search = kind_search(kind);
    const int W = cfg->max_error_weight;
    Masks mk;
    int num_phases = 1;
    std::vector<float> llr;
    const int *Hs = nullptr, *other = nullptr;
    if (family == kCC) {
        if (type == 'x' || type == 'X') { Hs = H2; other = H; }          // x errors: Hz searched, Hx other
        else if (type == 'z' || type == 'Z') { Hs = H; other = H2; }
        else { fprintf(stderr, "Invalid type_of_error: %c. Must be 'x' or 'z'.\n", type); return nullptr; }
        if (search == kPhase) {
            read_masks(file, mk);
            if (W == 6) num_phases = 2; else if (W == 7) num_phases = 3;
            else if (W > 0 && W < 6) num_phases = 1;
            else { fprintf(stderr, "Invalid max_error_weight: %d. Must be 6 or 7.\n", W); return nullptr; }
            if (mk.n == 0) {   // the reference indexes masks[-1] here (grand_decoder.cu:761)
                fprintf(stderr, "No phase order masks read from: %s\n", file ? file : "(null)");
                return nullptr;
            }
        } else if (W <= 0 || W > 7) {
            fprintf(stderr, "Invalid max_error_weight: %d. Must be between 1 and 7.\n", W);
            return nullptr;
        }
    } else {
        Hs = H;
        if (W <= 0 || W > 7) {
            fprintf(stderr, "Invalid max_error_weight: %d. Must be between 1 and 7.\n", W);
            return nullptr;
        }
        if (search == kML && m > 0 && !read_llr(file, llr, m)) return nullptr;
        if (search == kPhase) {
            if (read_masks(file, mk) == 0) {
                fprintf(stderr, "No phase order masks read from: %s\n", file ? file : "(null)");
                return nullptr;
            }
            num_phases = (W == 6) ? 2 : (W == 7) ? 3 : 1;
        }
    }
    // library limits
    if (n < 1 || n > kMaxWords * 32) { fprintf(stderr, "grand: n = %d outside [1, %d]\n", n, kMaxWords * 32); return nullptr; }
    if (m < 1 || m > kMaxM) { fprintf(stderr, "grand: m = %d outside [1, %d]\n", m, kMaxM); return nullptr; }
    if (family == kCLN && (NL < 0 || NL > 2048)) { fprintf(stderr, "grand: n_logicos = %d outside [0, 2048]\n", NL); return nullptr; }
    if (!Hs || (family == kCC && !other) || (family == kCLN && NL > 0 && !H2)) {
        fprintf(stderr, "grand: matrix pointer is NULL\n");
        return nullptr;
    }
    for (float v : llr)
        if (!std::isfinite(v)) { fprintf(stderr, "grand: LLR file holds a non-finite value\n"); return nullptr; }
    Runtime &R = rt();
    if (!R.ok) { fprintf(stderr, "grand: no usable GPU\n"); return nullptr; }

    Hasher h;
    h.mix((uint64_t)family); h.mix((uint64_t)n); h.mix((uint64_t)m); h.mix((uint64_t)NL);
    h.ints(Hs, (size_t)n * m);
    bool bin = all_binary(Hs, (size_t)n * m);
    if (family == kCC) { h.ints(other, (size_t)n * m); bin = bin && all_binary(other, (size_t)n * m); }
    else if (NL > 0) { h.ints(H2, (size_t)m * NL); bin = bin && all_binary(H2, (size_t)m * NL); }
    if (!llr.empty()) { h.mix(0x4c4c52ull); h.ints(reinterpret_cast<const int *>(llr.data()), llr.size()); }
    if (!bin) { fprintf(stderr, "grand: check/logical matrices must hold only 0 and 1\n"); return nullptr; }

    const int *second = (family == kCC) ? other : (NL > 0 ? H2 : nullptr);
    const size_t second_len = (family == kCC) ? (size_t)n * m : (NL > 0 ? (size_t)m * NL : 0);
    std::shared_ptr<CodeEntry> ce = find_code(h.h, family, n, m, NL, Hs, second, second_len, llr);
    if (!ce) {
        ce = upload_code(h.h, family, Hs, other, family == kCLN ? H2 : nullptr, n, m, NL, llr.empty() ? nullptr : &llr);
        if (!ce) return nullptr;
        ce->src_h.assign(Hs, Hs + (size_t)n * m);
        if (second) ce->src_other.assign(second, second + second_len);
        ce->src_llr = llr;
        remember_code(ce);
    }
    auto *s = new GrandSession();
    s->kind = kind; s->family = family; s->search = search; s->code = ce;
    if (ex) s->ex = *ex;
    DevCall &c = s->call;
    c.family = family; c.kind = search;
    c.W = std::min(W, kMaxW);
    c.mode = cfg->mode_of_search; c.thr_h = cfg->threshold_for_hybrid; c.max_size = cfg->max_size_active_positions;
    c.num_phases = num_phases;
    c.n_masks = mk.n;
    for (int i = 0; i < mk.n; ++i) {
        c.mask_tam_upper[i] = mk.tam_upper[i];
        for (int k = 0; k < 3; ++k) c.mask_order[i][k] = mk.order[i][k];
    }
    int cap = std::min(m, kMaxTam);
    if (family == kCLN) cap = std::min(cap, std::max(0, cfg->max_size_active_positions));
    c.cap = cap;
    c.ms_per_tick = R.ms_per_tick;
    return s;
}

// ------------------------------------------------------------- plain engine
// Host loop over nonzero shots, one launch per (shot, iteration, weight class),
// a synchronous read-back after each: the reference driver's structure
// (grand_decoder.cu:1284-1343) with the deterministic CPU-order search.
int run_plain(GrandSession *s, hipStream_t st) {
    const DevCode &code = s->code->dev;
    DevCall &call = s->call;
    const DevCode *dcode = s->code->dcode.as<DevCode>();
    const DevCall *dcall = s->dcall.as<DevCall>();
    unsigned nwork = 0;
    // one worklist partition (wl_parts = 1): its fronts are worklist[0 .. nwork)
    HIPX(hipMemcpyAsync(&nwork, call.counters + kWlBase, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    std::vector<int> work(nwork);
    if (nwork) HIPX(hipMemcpyAsync(work.data(), call.worklist, sizeof(int) * nwork, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    const int cap = std::max(call.cap, 1);
    const int kRecords = 4096;
    if (!s->plain_scratch.ensure((size_t)cap * 6 * 4) || !s->plain_state.ensure(grand_plain_state_bytes()) ||
        !s->plain_records.ensure((size_t)kRecords * grand_plain_record_bytes()) ||
        !s->plain_out.ensure(grand_plain_record_bytes()))
        return -1;
    struct StateHost { int tam_new; int pad; unsigned long long best; float prefix[kMaxW + 1]; } sh;
    struct RecHost { unsigned key; int pad; unsigned long long rank; int c[kMaxW]; } rec;
    std::vector<int> P;
    const int rmax = call.rmax;

    for (unsigned wi = 0; wi < nwork; ++wi) {
        const int shot = work[wi];
        ShotOutcome o{};
        o.hibrido = -2;
        for (int l = 0; l < kMaxWin; ++l) o.win[l] = -1;
        unsigned long long hit_rank = 0;
        int it = 0, tam = 0, tam_old = 0;
        bool found = false;
        while (it < code.max_soma && it < 3) {
            const int thr = code.max_soma - it;
            ++it;
            HIPX(grand_launch_plain_active(dcode, dcall, cap, shot, thr, 0, s->plain_scratch.p, s->plain_state.p, st));
            HIPX(hipMemcpyAsync(&sh, s->plain_state.p, sizeof(sh), hipMemcpyDeviceToHost, st));
            HIPX(hipStreamSynchronize(st));
            tam = sh.tam_new;
            if (call.family == kCLN) { if (tam > call.max_size) tam = tam_old; else tam_old = tam; }
            if (tam <= 0) continue;
            if (tam > call.cap) { fprintf(stderr, "grand: active set of %d positions exceeds %d\n", tam, kMaxTam); return -2; }
            int br = 0;
            if ((call.mode == 1 && tam > call.thr_h && tam <= call.max_size) || call.mode == 0) br = 1;
            else if ((call.mode == 1 && tam <= call.thr_h) || call.mode == -1) br = -1;
            if (br == 0) continue;
            o.hibrido = (br > 0) ? (o.hibrido == -1 ? 1 : 0) : -1;

            if (call.kind == kML) {
                HIPX(grand_launch_plain_active(dcode, dcall, cap, shot, thr, tam, s->plain_scratch.p, s->plain_state.p, st));
                HIPX(hipMemcpyAsync(&sh, s->plain_state.p, sizeof(sh), hipMemcpyDeviceToHost, st));
                P.resize(tam);
                HIPX(hipMemcpyAsync(P.data(), s->plain_scratch.as<int>() + 3 * cap, sizeof(int) * tam,
                                    hipMemcpyDeviceToHost, st));
                HIPX(hipStreamSynchronize(st));
                bool have = false;
                float best = 0.0f;
                const int max_w = std::min(call.W, tam);
                for (int w = 1; w <= max_w; ++w) {
                    if (have && sh.prefix[w] >= best) break;
                    const uint64_t Nw = ncr_host(tam, w);
                    o.count += Nw;
                    const uint64_t per = (Nw + 63) / 64;
                    const int R = (int)std::min<uint64_t>(std::max<uint64_t>(per, 1), (uint64_t)rmax);
                    const uint64_t nchunks = (Nw + 64ull * R - 1) / (64ull * R);
                    const int grid = (int)std::min<uint64_t>(nchunks, (uint64_t)kRecords);
                    HIPX(grand_launch_plain_class(dcode, dcall, cap, shot, w, tam, R, nchunks, grid, 1,
                                                  s->plain_scratch.p, s->plain_state.p, s->plain_records.p, st));
                    HIPX(grand_launch_plain_ml_reduce(s->plain_records.p, grid, s->plain_out.p, st));
                    HIPX(hipMemcpyAsync(&rec, s->plain_out.p, sizeof(rec), hipMemcpyDeviceToHost, st));
                    HIPX(hipStreamSynchronize(st));
                    o.wops += Nw * (uint64_t)(w + 1);
                    if (rec.key == 0xFFFFFFFFu) continue;
                    const uint32_t b = (rec.key & 0x80000000u) ? (rec.key & 0x7FFFFFFFu) : ~rec.key;
                    float cost;
                    std::memcpy(&cost, &b, 4);
                    if (!have || cost < best) {
                        have = true; best = cost; o.fw = w;
                        for (int l = 0; l < kMaxW; ++l) o.win[l] = (l < w) ? P[rec.c[l]] : -1;
                    }
                }
                found = have;
            } else {
                // weight list as in the fused kernel
                int wl[2 * kMaxW], nw = 0;
                if (call.kind == kUniform) { for (int w = 1; w <= call.W; ++w) wl[nw++] = w; }
                else {
                    int mi = call.n_masks - 1;
                    for (int i = 0; i < call.n_masks; ++i) if (tam <= call.mask_tam_upper[i]) { mi = i; break; }
                    for (int k = 0; k < call.num_phases; ++k) {
                        const int ph = call.mask_order[mi][k];
                        if (ph == 0) { for (int w = 1; w <= 5 && w <= tam && w <= kMaxW; ++w) wl[nw++] = w; }
                        else if (ph == 1) { if (6 <= tam) wl[nw++] = 6; }
                    }
                }
                for (int k = 0; k < nw && !found; ++k) {
                    const int w = wl[k];
                    const uint64_t Nw = ncr_host(tam, w);
                    if (Nw == 0) continue;
                    const uint64_t per = (Nw + 63) / 64;
                    const int R = (int)std::min<uint64_t>(std::max<uint64_t>(per, 1), (uint64_t)rmax);
                    const uint64_t nchunks = (Nw + 64ull * R - 1) / (64ull * R);
                    const int grid = (int)std::min<uint64_t>(nchunks, (uint64_t)kRecords);
                    HIPX(grand_launch_plain_class(dcode, dcall, cap, shot, w, tam, R, nchunks, grid, 0,
                                                  s->plain_scratch.p, s->plain_state.p, s->plain_records.p, st));
                    HIPX(hipMemcpyAsync(&sh, s->plain_state.p, sizeof(sh), hipMemcpyDeviceToHost, st));
                    HIPX(hipStreamSynchronize(st));
                    if (sh.best != ~0ull) {
                        found = true;
                        o.fw = w;
                        hit_rank = sh.best;
                        o.count += sh.best + 1;
                        o.wops += (sh.best + 1) * (uint64_t)(w + 1);
                    } else {
                        o.count += Nw;
                        o.wops += Nw * (uint64_t)(w + 1);
                    }
                }
            }
            if (found) break;
        }
        o.found = found ? 1 : 0;
        o.iteration = it;
        o.tam = tam;
        HIPX(grand_launch_plain_finalize(dcode, dcall, shot, &o, s->plain_scratch.p, hit_rank, st));
    }
    return 0;
}

// Device scratch of one decode: the session's Workspace, or regions of the
// library's fixed scratch array (decode without allocation).
struct WsView {
    uint32_t *s_packed, *e_packed, *e_aux, *lg_packed, *s_sig, *dec, *declog, *task_arrays;
    int *worklist;
    unsigned *counters;
    TaskSlot *tasks;
    DevCall *dcall;
};

// Queues one decode with all device scratch given by `w`; allocates nothing.
int decode_view(GrandSession *s, const WsView &w, const int *d_input, int input_errors, const int *d_logicals,
                int num_shots, GrandDecoderResult *d_results, unsigned *d_dec, unsigned *d_declog,
                unsigned long long *d_stats, hipStream_t st) {
    Runtime &R = rt();
    const DevCode &code = s->code->dev;
    DevCall &c = s->call;
    c.num_shots = num_shots;
    c.input_errors = (s->family == kCC) ? (input_errors ? 1 : 0) : 0;
    c.input = d_input;
    c.logicals = d_logicals;
    c.s_packed = w.s_packed;
    c.e_packed = w.e_packed;
    c.e_aux = w.e_aux;
    c.lg_packed = w.lg_packed;
    c.s_sig = w.s_sig;
    c.worklist = w.worklist;
    c.counters = w.counters;
    c.results = d_results;
    c.dec = d_dec ? d_dec : w.dec;
    c.declog = d_declog ? d_declog : w.declog;
    c.stats = d_stats;
    c.tasks = w.tasks;
    c.task_arrays = w.task_arrays;
    c.task_counters = w.counters + 4;
    {
        const int st_ex = s->ex.split_threshold;
        c.split_threshold = st_ex < 0 ? 0ull
                          : st_ex > 0 ? (unsigned long long)st_ex
                          : (unsigned long long)tuning_for(s->search).split_threshold;
    }
    // tuning knobs are read at every decode, so the autotuner can vary them on a live session
    c.rmax = tuning_for(s->search).rmax;
    c.claim_batch = std::max(1, std::min((s->ex.claim_batch > 0) ? s->ex.claim_batch : tuning_for(s->search).claim_batch,
                                         kMaxBatch));
    // shots whose first iteration is small are searched one per lane by the
    // triage kernel (fused engine, first-hit kinds, small codes); with errors
    // as input the degeneracy words must exist
    c.triage_search = (s->ex.engine != GRAND_ENGINE_PLAIN && s->ex.triage_search >= 0 && s->search != kML &&
                       c.lscores == nullptr && code.lane_ok &&
                       !(s->family == kCC && c.input_errors && code.dg == nullptr)) ? 1 : 0;
    c.split_worklist = (s->ex.engine != GRAND_ENGINE_PLAIN) ? 1 : 0;
    c.wl_parts = (s->ex.engine != GRAND_ENGINE_PLAIN) ? kWlParts : 1;
    // the triage kernel reads the column words, offsets, signatures and
    // degeneracy words of a small code from its LDS (lane search, s = H e)
    {
        const long long tsw = code.timg_words;
        c.triage_stage_words = ((c.triage_search || c.input_errors) && tsw <= kCodeLdsMaxWords) ? (int)tsw : 0;
    }
    // one claimed shot in the search kernel's LDS: id, signature, syndrome words,
    // then the packed errors (cc with errors) or the observed logicals (cln)
    c.stage_words = 2 + code.D + (c.input_errors ? 2 + code.M32 : 0) + (s->family == kCLN ? code.NL32 : 0);
    // search-kernel pair table (deep first-hit and ML classes): what is left of the wave's share
    // of the CU's LDS (at most 64 KB per workgroup) after everything else
    const int wpc = (s->ex.waves_per_cu > 0) ? s->ex.waves_per_cu : tuning_for(s->search).waves_per_cu;
    c.pair_cap = 0;
    if (s->ex.engine != GRAND_ENGINE_PLAIN) {
        const size_t base = grand_fused_lds_bytes(c.cap, s->search == kML || c.lscores != nullptr) +
                            4 * ((size_t)((s->code->search_lds_words + 3) & ~3) + (size_t)c.claim_batch * c.stage_words) +
                            4 * ((size_t)kPairBitmapWords + pair_head_words(s->search == kML));
        const size_t budget = std::min<size_t>(65536, kLdsPerCu / (size_t)std::max(wpc, 1));
        const long long all = (long long)c.cap * (c.cap - 1) / 2;
        const int kb = (s->ex.pair_table_kb != 0) ? s->ex.pair_table_kb : tuning_for(s->search).pair_table_kb;
        if (kb >= 0 && budget > base) {
            long long room = (long long)((budget - base) / 8);
            if (kb > 0) room = std::min(room, (long long)kb * 1024 / 8);
            const long long pc = std::min(room, std::min(all, (long long)kPairNone));
            if (pc >= 256) c.pair_cap = (int)pc;
        }
    }

    // The worklist length, claim cursor, error bits and task counters restart
    // with every batch, in the launch that installs the batch descriptor.
    const int groups = (num_shots + 63) / 64;
    const int tgrid = std::max(1, std::min(groups, R.cu_count * 16));
    HIPX(grand_launch_prep(&c, w.dcall, c.counters, st));
    const DevCode *dcode = s->code->dcode.as<DevCode>();
    const DevCall *dcall = w.dcall;
    hipEvent_t *tev = nullptr;
    if (s->timing) {
        const int k = (s->tslot + 1) % GrandSession::kTimingSlots;
        if (!s->harvest(k)) return -1;
        s->tslot = k;
        tev = s->ev.data() + 3 * k;
        HIPX(hipEventRecord(tev[0], st));
    }
    const int block_words = std::max(c.input_errors ? code.M32 : code.D, s->family == kCLN ? code.NL32 : 0);
    HIPX(grand_launch_triage(dcode, dcall, tgrid, block_words, c.triage_stage_words, st));
    if (tev) HIPX(hipEventRecord(tev[1], st));
    if (s->ex.engine == GRAND_ENGINE_PLAIN) {
        int rc = run_plain(s, st);
        if (rc == -2) { (void)grand_launch_fill(c.counters + 2, kErrTamOverflow, 1, st); rc = 0; }
        if (tev) { HIPX(hipEventRecord(tev[2], st)); s->tpending |= 1ull << s->tslot; }
        return rc;
    }
    // a deep shot of a small batch shares its classes with the other waves, so
    // with shared tasks on the grid keeps at least one wave per CU
    const int want = c.split_threshold ? std::max(num_shots, R.cu_count) : num_shots;
    const int grid = std::max(1, std::min(R.cu_count * wpc, want));
    HIPX(grand_launch_fused(s->code->dcode_search.as<DevCode>(), dcall, &c, s->code->search_lds_words, grid, st));
    if (tev) { HIPX(hipEventRecord(tev[2], st)); s->tpending |= 1ull << s->tslot; }
    return 0;
}

// The regions of the fixed scratch array for a batch of num_shots shots, each
// 256-byte aligned: the task table first (it stays valid from call to call),
// then counters, the device descriptor, the task arrays and the per-shot
// buffers.  Returns the bytes used; with base null only sizes.
size_t scratch_layout(const DevCode &c, int num_shots, int cap, char *base, WsView &w) {
    const size_t S = (size_t)std::max(num_shots, 1);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    w.tasks = reinterpret_cast<TaskSlot *>(take(sizeof(TaskSlot) * kTaskSlots));
    w.counters = reinterpret_cast<unsigned *>(take(kCounterBytes));
    w.dcall = reinterpret_cast<DevCall *>(take(sizeof(DevCall)));
    w.task_arrays = reinterpret_cast<uint32_t *>(take((size_t)kTaskSlots * 3 * std::max(cap, 1) * 4));
    w.s_packed = reinterpret_cast<uint32_t *>(take(S * c.D * 4));
    w.e_packed = reinterpret_cast<uint32_t *>(take(S * c.M32 * 4));
    w.e_aux = reinterpret_cast<uint32_t *>(take(S * 8));
    w.lg_packed = reinterpret_cast<uint32_t *>(take(S * c.NL32 * 4));
    w.s_sig = reinterpret_cast<uint32_t *>(take(S * 4));
    w.worklist = reinterpret_cast<int *>(take((S + 63) / 64 * 64 * 4));
    w.dec = reinterpret_cast<uint32_t *>(take(S * c.M32 * 4));
    w.declog = reinterpret_cast<uint32_t *>(take(S * c.NL32 * 4));
    return off;
}

// Largest batch whose scratch layout fits the fixed array.
int noalloc_max_shots(GrandSession *s) {
    WsView w{};
    int lo = 0, hi = 1 << 26;
    while (lo < hi) {                              // largest batch whose layout fits the array
        const int mid = lo + (hi - lo + 1) / 2;
        if (scratch_layout(s->code->dev, mid, s->call.cap, nullptr, w) <= kScratchBytes) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Queues one decode with the session's own scratch, grown on demand.
int decode_device(GrandSession *s, const int *d_input, int input_errors, const int *d_logicals, int num_shots,
                  GrandDecoderResult *d_results, unsigned *d_dec, unsigned *d_declog, unsigned long long *d_stats,
                  hipStream_t st) {
    if (num_shots <= 0) return 0;
    const DevCode &code = s->code->dev;
    if (!s->ws.ensure(code, num_shots, s->call.cap) || !s->dcall.ensure(sizeof(DevCall))) return -1;
    // The task table is cleared once when allocated: every launch leaves all
    // slots idle (each owner waits for its chunks and frees its slot).
    if (s->ws.tasks_fresh) {
        HIPX(grand_launch_fill(reinterpret_cast<unsigned *>(s->ws.tasks.p), 0u,
                               (int)(sizeof(TaskSlot) * kTaskSlots / 4), st));
        s->ws.tasks_fresh = false;
    }
    WsView w;
    w.s_packed = s->ws.s_packed.as<uint32_t>();
    w.e_packed = s->ws.e_packed.as<uint32_t>();
    w.e_aux = s->ws.e_aux.as<uint32_t>();
    w.lg_packed = s->ws.lg_packed.as<uint32_t>();
    w.s_sig = s->ws.s_sig.as<uint32_t>();
    w.dec = s->ws.dec.as<uint32_t>();
    w.declog = s->ws.declog.as<uint32_t>();
    w.task_arrays = s->ws.task_arrays.as<uint32_t>();
    w.worklist = s->ws.worklist.as<int>();
    w.counters = s->ws.counters.as<unsigned>();
    w.tasks = s->ws.tasks.as<TaskSlot>();
    w.dcall = s->dcall.as<DevCall>();
    return decode_view(s, w, d_input, input_errors, d_logicals, num_shots, d_results, d_dec, d_declog, d_stats, st);
}

// ----------------------------------------------------------------- ABI path
struct AbiBuffers {
    DevMem input, logicals, results, dec_int, declog_int;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // H2D start/end, D2H start/end
    bool events() {
        if (ev[0]) return true;
        for (auto &e : ev)
            if (hipEventCreate(&e) != hipSuccess) { e = nullptr; return false; }
        return true;
    }
};
AbiBuffers g_abi;

bool print_enabled() {
    const char *v = getenv("GRAND_PRINT_SHOTS");
    return v && v[0] == '1';
}

void print_shot(int family, bool with_errors, int shot, const GrandDecoderResult &r) {
    if (family == kCC && with_errors)
        printf("shot %d: found=%d, original_weight=%d, found_weight=%d, degenerate=%d, is_equal=%d, iteration=%d, tam=%d, total_tested_combinations=%llu, hybrid=%d, time_sum_ms=%.5f, total_time_search_ms=%.5f, transfer_time_ms=%.5f, total_to_find_ms=%.5f\n",
               shot + 1, r.found, r.original_weight, r.found_weight, r.degenerate, r.is_equal, r.iteration, r.tam,
               r.total_tested_combinations, r.hibrido, r.time_sum_ms, r.total_time_search_ms, r.transfer_time_ms,
               r.total_to_find_ms);
    else if (family == kCC)
        printf("shot %d: found=%d, found_weight=%d, iteration=%d, tam=%d, total_tested_combinations=%llu, hybrid=%d, time_sum_ms=%.5f, total_time_search_ms=%.5f, transfer_time_ms=%.5f, total_to_find_ms=%.5f\n",
               shot + 1, r.found, r.found_weight, r.iteration, r.tam, r.total_tested_combinations, r.hibrido,
               r.time_sum_ms, r.total_time_search_ms, r.transfer_time_ms, r.total_to_find_ms);
    else
        printf("shot %d: found=%d, found_weight=%d, degenerate=%d, iteration=%d, tam=%d, total_tested_combinations=%llu, hybrid=%d, time_sum_ms=%.5f, total_time_search_ms=%.5f, transfer_time_ms=%.5f, total_to_find_ms=%.5f\n",
               shot + 1, r.found, r.found_weight, r.degenerate, r.iteration, r.tam, r.total_tested_combinations,
               r.hibrido, r.time_sum_ms, r.total_time_search_ms, r.transfer_time_ms, r.total_to_find_ms);
}

int report_error_bits(unsigned bits) {
    if (bits & kErrNonBinary) fprintf(stderr, "grand: input rows must hold only 0 and 1\n");
    if (bits & kErrTamOverflow) fprintf(stderr, "grand: an active set exceeds the supported %d positions\n", kMaxTam);
    if (bits & kErrTaskWait) fprintf(stderr, "grand: a shared search task did not complete\n");
    return bits ? -1 : 0;
}

int run_abi(int kind, const int *H, const int *H2, int n, int m, int NL, char type, const GrandDecoderConfig *cfg,
            const char *file, const GrandDecoderExConfig *ex, const int *input, int input_errors, const int *logicals,
            int num_shots, GrandDecoderResult *results, int *decoded, int *decoded_logicals) {
    std::lock_guard<std::recursive_mutex> lock(rt().mu);
    GrandSession *s = create_session(kind, H, H2, n, m, NL, type, cfg, file, ex);
    if (!s) return -1;
    std::unique_ptr<GrandSession> guard(s);
    if (num_shots <= 0) return 0;
    const int family = s->family;
    if (!input || !results || (family == kCLN && NL > 0 && !logicals)) {
        fprintf(stderr, "grand: input/results/logicals pointer is NULL\n");
        return -1;
    }
    Runtime &R = rt();
    hipStream_t st = R.stream;
    const DevCode &code = s->code->dev;
    const int cols = (family == kCC && input_errors) ? m : n;
    const size_t S = (size_t)num_shots;
    AbiBuffers &B = g_abi;
    if (!B.input.ensure(S * cols * 4) || !B.results.ensure(S * sizeof(GrandDecoderResult)) ||
        (family == kCLN && !B.logicals.ensure(S * std::max(NL, 1) * 4)))
        return -1;
    const bool timed = B.events();
    if (timed) HIPX(hipEventRecord(B.ev[0], st));
    HIPX(hipMemcpyAsync(B.input.p, input, S * cols * 4, hipMemcpyHostToDevice, st));
    if (family == kCLN && NL > 0) HIPX(hipMemcpyAsync(B.logicals.p, logicals, S * NL * 4, hipMemcpyHostToDevice, st));
    if (timed) HIPX(hipEventRecord(B.ev[1], st));
    if (decode_device(s, B.input.as<int>(), input_errors, B.logicals.as<int>(), num_shots,
                      B.results.as<GrandDecoderResult>(), nullptr, nullptr, nullptr, st) != 0)
        return -1;
    if (timed) HIPX(hipEventRecord(B.ev[2], st));
    HIPX(hipMemcpyAsync(results, B.results.p, S * sizeof(GrandDecoderResult), hipMemcpyDeviceToHost, st));
    if (decoded) {
        if (!B.dec_int.ensure(S * m * 4)) return -1;
        HIPX(grand_launch_expand(s->call.dec, code.M32, m, num_shots, B.dec_int.as<int>(), st));
        HIPX(hipMemcpyAsync(decoded, B.dec_int.p, S * m * 4, hipMemcpyDeviceToHost, st));
    }
    if (decoded_logicals && family == kCLN && NL > 0) {
        if (!B.declog_int.ensure(S * NL * 4)) return -1;
        HIPX(grand_launch_expand(s->call.declog, code.NL32, NL, num_shots, B.declog_int.as<int>(), st));
        HIPX(hipMemcpyAsync(decoded_logicals, B.declog_int.p, S * NL * 4, hipMemcpyDeviceToHost, st));
    }
    unsigned bits = 0;
    HIPX(hipMemcpyAsync(&bits, s->call.counters + 2, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    if (B.ev[0]) HIPX(hipEventRecord(B.ev[3], st));
    HIPX(hipStreamSynchronize(st));
    if (B.ev[0]) {
        // transfer_time_ms: the call's host<->device copies, an equal share per shot
        float h2d = 0.0f, d2h = 0.0f;
        if (hipEventElapsedTime(&h2d, B.ev[0], B.ev[1]) == hipSuccess &&
            hipEventElapsedTime(&d2h, B.ev[2], B.ev[3]) == hipSuccess) {
            const double share = ((double)h2d + (double)d2h) / (double)num_shots;
            for (int i = 0; i < num_shots; ++i) results[i].transfer_time_ms = share;
        }
    }
    if (report_error_bits(bits)) return -1;
    if (print_enabled())
        for (int i = 0; i < num_shots; ++i) print_shot(family, family == kCC && input_errors, i, results[i]);
    return 0;
}

}  // namespace
}  // namespace grand

// ====================================================================== C ABI
extern "C" {

using namespace grand;

int grand_decode_cc_ex(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                       const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                       const GrandDecoderExConfig *ex, int *decoded) {
    return run_abi(GRAND_KIND_CC, Hx, Hz, n, m, 0, type_of_error, config, nullptr, ex, errors ? errors : syndromes,
                   errors != nullptr, nullptr, num_shots, results, decoded, nullptr);
}

int grand_decode_prob_cc_ex(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                            const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                            const char *probability_file, const GrandDecoderExConfig *ex, int *decoded) {
    return run_abi(GRAND_KIND_PROB_CC, Hx, Hz, n, m, 0, type_of_error, config, probability_file, ex,
                   errors ? errors : syndromes, errors != nullptr, nullptr, num_shots, results, decoded, nullptr);
}

int grand_decode_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                        int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
                        const GrandDecoderExConfig *ex, int *decoded, int *decoded_logicals) {
    return run_abi(GRAND_KIND_CLN, H_dem, L_dem, n, m, n_logicos, 'x', config, nullptr, ex, detectors, 0, logicals,
                   num_shots, results, decoded, decoded_logicals);
}

int grand_decode_prob_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors,
                             int *logicals, int num_shots, const GrandDecoderConfig *config,
                             GrandDecoderResult *results, const char *probability_file,
                             const GrandDecoderExConfig *ex, int *decoded, int *decoded_logicals) {
    return run_abi(GRAND_KIND_PROB_CLN, H_dem, L_dem, n, m, n_logicos, 'x', config, probability_file, ex, detectors,
                   0, logicals, num_shots, results, decoded, decoded_logicals);
}

int grand_decode_ML_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors,
                           int *logicals, int num_shots, const GrandDecoderConfig *config,
                           GrandDecoderResult *results, const char *llr_file, const GrandDecoderExConfig *ex,
                           int *decoded, int *decoded_logicals) {
    return run_abi(GRAND_KIND_ML_CLN, H_dem, L_dem, n, m, n_logicos, 'x', config, llr_file, ex, detectors, 0,
                   logicals, num_shots, results, decoded, decoded_logicals);
}

int grand_decode_cc(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                    const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error) {
    return grand_decode_cc_ex(Hx, Hz, n, m, errors, syndromes, num_shots, config, results, type_of_error, nullptr,
                              nullptr);
}

int grand_decode_prob_cc(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                         const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                         const char *probability_file) {
    return grand_decode_prob_cc_ex(Hx, Hz, n, m, errors, syndromes, num_shots, config, results, type_of_error,
                                   probability_file, nullptr, nullptr);
}

int grand_decode_cln(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                     int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results) {
    return grand_decode_cln_ex(H_dem, L_dem, n, m, n_logicos, detectors, logicals, num_shots, config, results,
                               nullptr, nullptr, nullptr);
}

int grand_decode_prob_cln(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                          int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
                          const char *probability_file) {
    return grand_decode_prob_cln_ex(H_dem, L_dem, n, m, n_logicos, detectors, logicals, num_shots, config, results,
                                    probability_file, nullptr, nullptr, nullptr);
}

int grand_decode_ML_cln(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                        int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
                        const char *llr_file) {
    return grand_decode_ML_cln_ex(H_dem, L_dem, n, m, n_logicos, detectors, logicals, num_shots, config, results,
                                  llr_file, nullptr, nullptr, nullptr);
}

// ----------------------------------------------------------------- sessions
GrandSession *grand_session_create(int kind, const int *H, const int *H2, int n, int m, int n_logicos,
                                   char type_of_error, const GrandDecoderConfig *config, const char *file,
                                   const GrandDecoderExConfig *ex) {
    if (kind < GRAND_KIND_CC || kind > GRAND_KIND_ML_CLN) { fprintf(stderr, "grand: unknown kind %d\n", kind); return nullptr; }
    std::lock_guard<std::recursive_mutex> lock(rt().mu);
    return create_session(kind, H, H2, n, m, n_logicos, type_of_error, config, file, ex);
}

int grand_session_decode_device(GrandSession *s, const int *d_input, int input_is_errors, const int *d_logicals,
                                int num_shots, GrandDecoderResult *d_results, unsigned *d_decoded,
                                unsigned *d_decoded_logicals, unsigned long long *d_stats, void *stream) {
    if (!s || (num_shots > 0 && (!d_input || !d_results))) return -1;
    if (s->family == kCLN && s->code->NL > 0 && num_shots > 0 && !d_logicals) return -1;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream, ordered with other default-stream work
    return decode_device(s, d_input, input_is_errors, d_logicals, num_shots, d_results, d_decoded, d_decoded_logicals,
                         d_stats, st);
}

int grand_session_decode_device_noalloc(GrandSession *s, const int *d_input, int input_is_errors,
                                        const int *d_logicals, int num_shots, GrandDecoderResult *d_results,
                                        unsigned *d_decoded, unsigned *d_decoded_logicals,
                                        unsigned long long *d_stats, void *stream) {
    if (!s || num_shots < 0 || (num_shots > 0 && (!d_input || !d_results))) return -1;
    if (s->family == kCLN && s->code->NL > 0 && num_shots > 0 && !d_logicals) return -1;
    if (num_shots == 0) return 0;
    if (s->ex.engine == GRAND_ENGINE_PLAIN) {
        fprintf(stderr, "grand: decode without allocation is not available for the plain engine\n");
        return -1;
    }
    WsView w{};
    const size_t need = scratch_layout(s->code->dev, num_shots, s->call.cap, nullptr, w);
    if (need > kScratchBytes) {
        fprintf(stderr, "grand: %d shots need %zu bytes of device scratch; decode without allocation holds %zu "
                        "(at most %d shots of this code)\n",
                num_shots, need, kScratchBytes, noalloc_max_shots(s));
        return -1;
    }
    std::lock_guard<std::recursive_mutex> lock(rt().mu);
    hipStream_t st = static_cast<hipStream_t>(stream);
    void *base = nullptr;
    if (!hip_check(grand_scratch_address(&base, st), "device scratch address")) return -1;
    scratch_layout(s->code->dev, num_shots, s->call.cap, static_cast<char *>(base), w);
    return decode_view(s, w, d_input, input_is_errors, d_logicals, num_shots, d_results, d_decoded,
                       d_decoded_logicals, d_stats, st);
}


int grand_session_decode_learned(GrandSession *s, const int *d_input, int input_is_errors, const int *d_logicals,
                                 int num_shots, const float *d_scores, int stride, int scores_per_column,
                                 int max_weight, unsigned long long budget, GrandDecoderResult *d_results,
                                 unsigned *d_decoded, unsigned *d_decoded_logicals, unsigned long long *d_stats,
                                 void *stream) {
    if (!s || s->search == kML || s->ex.engine == GRAND_ENGINE_PLAIN) return -1;
    if (num_shots > 0 && (!d_input || !d_results || !d_scores)) return -1;
    if (s->family == kCLN && s->code->NL > 0 && num_shots > 0 && !d_logicals) return -1;
    if (max_weight < 1 || max_weight > kMaxWin || stride < 1 || (scores_per_column && stride < s->code->dev.m)) return -1;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream, ordered with other default-stream work
    DevCall &c = s->call;
    c.lscores = d_scores;
    c.lstride = stride;
    c.lper_column = scores_per_column ? 1 : 0;
    c.lmaxw = max_weight;
    c.lbudget = budget;
    const int rc = decode_device(s, d_input, input_is_errors, d_logicals, num_shots, d_results, d_decoded,
                                 d_decoded_logicals, d_stats, st);
    c.lscores = nullptr;   // the descriptor was copied into the launch; later decodes are plain
    return rc;
}

struct GrandGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GrandGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

GrandGraph *grand_session_capture(GrandSession *s, const int *d_input, int input_is_errors, const int *d_logicals,
                                  int num_shots, GrandDecoderResult *d_results, unsigned *d_decoded,
                                  unsigned *d_decoded_logicals, unsigned long long *d_stats, void *stream) {
    if (!s || !stream || num_shots <= 0 || !d_input || !d_results || s->ex.engine == GRAND_ENGINE_PLAIN) return nullptr;
    if (s->family == kCLN && s->code->NL > 0 && !d_logicals) return nullptr;
    hipStream_t st = (hipStream_t)stream;
    // everything that allocates or clears memory happens outside the capture
    if (grand_session_reserve(s, num_shots) != 0) return nullptr;
    if (s->ws.tasks_fresh) {
        if (!hip_check(grand_launch_fill(reinterpret_cast<unsigned *>(s->ws.tasks.p), 0u,
                                         (int)(sizeof(TaskSlot) * kTaskSlots / 4), st), "fill"))
            return nullptr;
        s->ws.tasks_fresh = false;
    }
    auto g = std::make_unique<GrandGraph>();
    const bool timing = s->timing;
    s->timing = false;
    if (!hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) {
        s->timing = timing;
        return nullptr;
    }
    const int rc = decode_device(s, d_input, input_is_errors, d_logicals, num_shots, d_results, d_decoded,
                                 d_decoded_logicals, d_stats, st);
    const hipError_t ec = hipStreamEndCapture(st, &g->graph);
    s->timing = timing;
    if (rc != 0 || !hip_check(ec, "hipStreamEndCapture")) return nullptr;
    if (!hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate")) return nullptr;
    return g.release();
}

int grand_graph_launch(GrandGraph *g, void *stream) {
    if (!g || !g->exec) return -1;
    return hip_check(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch") ? 0 : -1;
}

void grand_graph_destroy(GrandGraph *g) { delete g; }

int grand_session_error_bits(GrandSession *s, void *stream) {
    if (!s || !s->call.counters) return 0;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream, ordered with other default-stream work
    unsigned bits = 0;
    if (hipMemcpyAsync(&bits, s->call.counters + 2, sizeof(unsigned), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    return (int)bits;
}

int grand_session_error_bits_async(GrandSession *s, unsigned *host_dst, void *stream) {
    if (!s || !host_dst || !s->call.counters) return -1;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream, ordered with other default-stream work
    return hipMemcpyAsync(host_dst, s->call.counters + 2, sizeof(unsigned), hipMemcpyDeviceToHost, st) == hipSuccess
               ? 0 : -1;
}

int grand_session_reserve(GrandSession *s, int num_shots) {
    if (!s) return -1;
    if (!s->ws.ensure(s->code->dev, std::max(num_shots, 1), s->call.cap)) return -1;
    if (!s->dcall.ensure(sizeof(DevCall))) return -1;
    s->call.counters = s->ws.counters.as<unsigned>();
    return 0;
}

int grand_session_enable_timing(GrandSession *s, int on) {
    if (!s) return -1;
    if (on && s->ev.empty()) {
        // timing-only events: no system-scope fence (cache writeback and
        // invalidate) when they complete, which otherwise stalls the stream
        // between the triage and search kernels
        s->ev.assign(3 * GrandSession::kTimingSlots, nullptr);
        for (hipEvent_t &e : s->ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return -1;
    }
    s->timing = on != 0;
    return 0;
}

int grand_session_last_timing(GrandSession *s, float *triage_ms, float *search_ms) {
    if (!s || !s->timing || s->tslot < 0) return -1;
    if (!s->harvest(s->tslot)) return -1;
    if (triage_ms) *triage_ms = s->slot_ms[s->tslot][0];
    if (search_ms) *search_ms = s->slot_ms[s->tslot][1];
    return 0;
}

int grand_session_timing_totals(GrandSession *s, double *triage_ms, double *search_ms, long long *decodes) {
    if (!s || s->ev.empty()) return -1;
    for (int k = 0; k < GrandSession::kTimingSlots; ++k)
        if (!s->harvest(k)) return -1;
    if (triage_ms) *triage_ms = s->sum_ms[0];
    if (search_ms) *search_ms = s->sum_ms[1];
    if (decodes) *decodes = s->sum_n;
    return 0;
}

int grand_session_timing_reset(GrandSession *s) {
    if (!s) return -1;
    if (!s->ev.empty())
        for (int k = 0; k < GrandSession::kTimingSlots; ++k)
            if (!s->harvest(k)) return -1;
    s->sum_ms[0] = s->sum_ms[1] = 0.0;
    s->sum_n = 0;
    return 0;
}

int grand_session_set_work_counters(GrandSession *s, unsigned long long *d_work) {
    if (!s) return -1;
    s->call.work = d_work;
    return 0;
}

double grand_session_ms_per_tick(GrandSession *s) { return s ? s->call.ms_per_tick : 0.0; }
int grand_session_pattern_words(GrandSession *s) { return s ? s->code->dev.M32 : 0; }
int grand_session_logical_words(GrandSession *s) { return s ? s->code->dev.NL32 : 0; }

size_t grand_session_footprint(GrandSession *s, int num_shots) {
    if (!s) return 0;
    return s->code->bytes + s->ws.bytes_for(s->code->dev, num_shots, s->call.cap);
}

void grand_session_destroy(GrandSession *s) {
    std::lock_guard<std::recursive_mutex> lock(rt().mu);
    delete s;
}

// Degeneracy words of a CSS pair (host only; the same words the decoder uses).
int grand_degeneracy_words(const int *H, const int *other, int n, int m, unsigned *words) {
    if (!H || !other || n < 1 || m < 1) return -1;
    const int M32 = (m + 31) / 32;
    std::vector<uint32_t> Hrow((size_t)n * M32, 0), rows((size_t)n * M32, 0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < m; ++c) {
            if (H[(size_t)r * m + c]) Hrow[(size_t)r * M32 + (c >> 5)] |= 1u << (c & 31);
            if (other[(size_t)r * m + c]) rows[(size_t)r * M32 + (c >> 5)] |= 1u << (c & 31);
        }
    std::vector<uint32_t> R, dg;
    std::vector<int> piv;
    grand::gf2_rref(std::move(rows), n, m, R, piv);
    const int t = grand::degeneracy_words(Hrow, R, piv, n, m, dg);
    if (t >= 0 && words) for (int c = 0; c < m; ++c) words[c] = dg[c];
    return t;
}

// Compute units and peak engine clock of the current device (roofline inputs).
int grand_device_info(int *cus, int *clock_khz, int *wall_clock_khz) {
    Runtime &R = rt();
    if (!R.ok) return -1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, R.device) != hipSuccess) return -1;
    if (cus) *cus = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (wall_clock_khz) *wall_clock_khz = (int)(1.0 / R.ms_per_tick + 0.5);
    return 0;
}

// Tuning knobs (written by the autotuner): set for every search kind; all four
// negative restores the built-in defaults.
void grand_set_tuning(int waves_per_cu, int claim_batch, int rmax, int split_threshold) {
    if (waves_per_cu < 0 && claim_batch < 0 && rmax < 0 && split_threshold < 0) {
        g_tuning = kTuningFirstHit;
        g_tuning_ml = kTuningMl;
        return;
    }
    for (Tuning *t : {&g_tuning, &g_tuning_ml}) {
        if (waves_per_cu > 0) t->waves_per_cu = waves_per_cu;
        if (claim_batch > 0) t->claim_batch = std::min(claim_batch, kMaxBatch);
        if (rmax > 0) t->rmax = rmax;
        if (split_threshold > 0) t->split_threshold = split_threshold;
    }
}

void grand_set_pair_table_kb(int kb) {
    g_tuning.pair_table_kb = kb;
    g_tuning_ml.pair_table_kb = kb;
}

int grand_get_pair_table_kb(void) { return g_tuning.pair_table_kb; }

void grand_get_tuning(int *waves_per_cu, int *claim_batch, int *rmax, int *split_threshold) {
    if (waves_per_cu) *waves_per_cu = g_tuning.waves_per_cu;
    if (claim_batch) *claim_batch = g_tuning.claim_batch;
    if (rmax) *rmax = g_tuning.rmax;
    if (split_threshold) *split_threshold = g_tuning.split_threshold;
}

}  // extern "C"
