// grand_oracle.cpp -- CPU restatement of the reference GRAND decoder's CPU search.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
// (library_s_grand_amd/libgranddecoder.so): only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it.  The product never links or calls it.
//
// It restates, function by function, the reference's CPU-mode semantics
// (mode_of_search = -1) from grand_decoder.cu, with every search branch of the
// reference driver (its "GPU" and "CPU" branches alike) evaluated with the CPU
// search semantics: lexicographically first match at the lowest weight position,
// patterns counted up to and including the hit (grand_decoder.cu:534-588,
// :295-350), and the CPU maximum-likelihood search (grand_decoder.cu:590-666).
// The branch decision itself (which iterations are searched, the `hibrido` code)
// follows the reference driver exactly (grand_decoder.cu:981-1063, :1296-1335).
//
// Exports the reference C ABI (grand_decoder.h:92-176) plus `_ex` variants that
// also return the decoded pattern and the decoded logicals, and
// oracle_set_threads() for an OpenMP shot-sharded timing run.
//
// Differences from the reference that do not change any result on inputs the
// reference handles: counts are 64-bit (the reference's per-search counter is a
// 32-bit int, grand_decoder.cu:534, which overflows past 2^31 patterns); the
// binomial table covers active sets up to 1024 positions (the reference's covers
// 300, grand_decoder.cu:13, and indexes past it otherwise); per-shot printing is
// off unless GRAND_PRINT_SHOTS=1.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/grand_decoder.h"

namespace {

constexpr int kMaxW = GRAND_MAX_W;          // grand_decoder.h:28
constexpr int kMaxMasks = 20;               // grand_decoder.cu:11
constexpr int kMaxTam = 1024;               // reference: 300 (grand_decoder.cu:13)

int g_threads = 1;

// ---- binomial table: grand_decoder.cu:42-63 ---------------------------------
struct Binom {
    uint64_t t[kMaxTam + 1][kMaxW + 1];
    Binom() {
        for (int n = 0; n <= kMaxTam; ++n)
            for (int k = 0; k <= kMaxW; ++k) {
                if (k > n) { t[n][k] = 0; continue; }
                uint64_t r = 1;                       // r = C(n, i) after step i
                for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - i + 1) / (uint64_t)i;
                t[n][k] = r;
            }
    }
    uint64_t operator()(int n, int k) const {
        if (n < 0 || k < 0 || k > kMaxW || n > kMaxTam || k > n) return 0;
        return t[n][k];
    }
};
const Binom C;

// Lexicographic unranking by linear scan: grand_decoder.cu:264-280.
void unrank_lex(uint64_t rank, int n, int k, int *out) {
    int cand = 0;
    for (int i = 0; i < k; ++i) {
        const int rest = k - 1 - i;
        uint64_t cnt = C(n - 1 - cand, rest);
        while (rank >= cnt && (n - 1 - cand) > rest) {
            rank -= cnt;
            ++cand;
            cnt = C(n - 1 - cand, rest);
        }
        out[i] = cand++;
    }
}

// Row-by-row parity test of a candidate: grand_decoder.cu:282-293.
bool parity_matches(const int *H, const int *s, const int *pos, const int *combo, int w, int n, int m) {
    for (int r = 0; r < n; ++r) {
        int p = 0;
        for (int j = 0; j < w; ++j) p ^= H[r * m + pos[combo[j]]];
        if (p != s[r]) return false;
    }
    return true;
}

// ---- phase-order masks: grand_decoder.cu:15-19, :728-786 -----------------------
struct PhaseMask { int tam_upper; int order[3]; float prob[3]; };

int read_masks(const char *path, PhaseMask *masks) {
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) { fprintf(stderr, "Error opening probability file: %s\n", path ? path : "(null)"); return 0; }
    int cnt = 0;
    while (cnt < kMaxMasks) {
        PhaseMask t;
        if (fscanf(f, "%d %d %d %d %f %f %f", &t.tam_upper, &t.order[0], &t.order[1], &t.order[2],
                   &t.prob[0], &t.prob[1], &t.prob[2]) != 7) break;
        masks[cnt++] = t;
    }
    fclose(f);
    return cnt;
}

const PhaseMask &mask_for(int tam, const PhaseMask *masks, int n) {
    for (int i = 0; i < n; ++i) if (tam <= masks[i].tam_upper) return masks[i];
    return masks[n - 1];
}

// Phase 0 -> weights 1..5, phase 1 -> {6}, phase 2 -> {7} (never reachable with
// GRAND_MAX_W = 6); each capped by tam.  grand_decoder.cu:764-786.
int phase_weights(int phase, int tam, int *w) {
    int len = 0;
    if (phase == 0) { for (int x = 1; x <= 5 && x <= tam && x <= kMaxW; ++x) w[len++] = x; }
    else if (phase == 1) { if (6 <= tam && 6 <= kMaxW) w[len++] = 6; }
    else if (phase == 2) { if (7 <= tam && 7 <= kMaxW) w[len++] = 7; }
    return len;
}

bool read_llr(const char *path, float *llr, int m) {
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) { fprintf(stderr, "Error opening LLR file: %s\n", path ? path : "(null)"); return false; }
    for (int i = 0; i < m; ++i)
        if (fscanf(f, "%f", &llr[i]) != 1) {
            fprintf(stderr, "Error reading LLR file (expected %d values): %s\n", m, path);
            fclose(f);
            return false;
        }
    fclose(f);
    return true;
}

// ---- GF(2) row reduction and row-space membership: grand_decoder.cu:418-474 ---
int rref(std::vector<int> &A, int rows, int cols, std::vector<int> &piv) {
    int r = 0;
    for (int c = 0; c < cols && r < rows; ++c) {
        int p = -1;
        for (int i = r; i < rows; ++i) if (A[(size_t)i * cols + c] & 1) { p = i; break; }
        if (p < 0) continue;
        if (p != r) for (int j = 0; j < cols; ++j) std::swap(A[(size_t)r * cols + j], A[(size_t)p * cols + j]);
        for (int i = 0; i < rows; ++i)
            if (i != r && (A[(size_t)i * cols + c] & 1))
                for (int j = 0; j < cols; ++j) A[(size_t)i * cols + j] ^= A[(size_t)r * cols + j];
        piv[r++] = c;
    }
    return r;
}

bool in_rowspace(const std::vector<int> &R, int cols, const std::vector<int> &piv, int rank, std::vector<int> &e) {
    for (int r = 0; r < rank; ++r) {
        const int c = piv[r];
        if (e[c] & 1) for (int j = c; j < cols; ++j) e[j] ^= R[(size_t)r * cols + j];
    }
    for (int j = 0; j < cols; ++j) if (e[j] & 1) return false;
    return true;
}

// ---- the call-level setup shared by all five entry points ----------------------
enum Family { kCC, kCLN };
enum Search { kUniform, kPhase, kML };

struct Call {
    Family fam;
    Search kind;
    int n, m, n_log;
    const int *H;              // the searched matrix (n x m)
    std::vector<int> R;        // RREF of the other CSS matrix (cc only)
    std::vector<int> piv;
    int rank = 0;
    const int *L = nullptr;    // m x n_log (cln)
    std::vector<int> colw;
    int max_soma = 0;
    bool regular = true;
    PhaseMask masks[kMaxMasks];
    int n_masks = 0;
    int num_phases = 1;
    std::vector<float> llr;
    GrandDecoderConfig cfg;
};

// grand_decoder.cu:376-404
void column_stats(Call &c) {
    c.colw.assign(c.m, 0);
    c.max_soma = 0;
    for (int col = 0; col < c.m; ++col) {
        int s = 0;
        for (int r = 0; r < c.n; ++r) s += c.H[(size_t)r * c.m + col];
        c.colw[col] = s;
        c.max_soma = std::max(c.max_soma, s);
    }
    c.regular = true;
    for (int col = 1; col < c.m; ++col) if (c.colw[col] != c.colw[0]) { c.regular = false; break; }
}

struct ShotOut {
    GrandDecoderResult res;
    std::vector<int> e_est;     // m
    std::vector<int> l_est;     // n_log
};

// First-hit enumeration over an ordered weight list: every weight class in
// lexicographic rank order, first parity match wins, count includes the hit.
// grand_decoder.cu:295-350 (phase lists) and :534-588 (weights 1..W).
bool first_hit(const Call &c, const int *s, const int *pos, int tam, const int *wl, int nw,
               uint64_t &count, std::vector<int> &e_est) {
    for (int k = 0; k < nw; ++k) {
        const int w = wl[k];
        if (w <= 0 || w > tam) continue;
        const uint64_t total = C(tam, w);
        int combo[kMaxW];
        for (uint64_t r = 0; r < total; ++r) {
            ++count;
            unrank_lex(r, tam, w, combo);
            if (parity_matches(c.H, s, pos, combo, w, c.n, c.m)) {
                std::fill(e_est.begin(), e_est.end(), 0);
                for (int j = 0; j < w; ++j) e_est[pos[combo[j]]] = 1;
                return true;
            }
        }
    }
    return false;
}

// Maximum-likelihood search: grand_decoder.cu:590-666.
bool ml_search(const Call &c, const int *s, const int *pos, int tam, int W, uint64_t &count, std::vector<int> &e_est) {
    if (W > kMaxW) W = kMaxW;
    std::vector<int> P(pos, pos + tam);
    for (int i = 1; i < tam; ++i) {                  // stable insertion sort by llr
        const int key = P[i];
        const float kl = c.llr[key];
        int j = i - 1;
        while (j >= 0 && c.llr[P[j]] > kl) { P[j + 1] = P[j]; --j; }
        P[j + 1] = key;
    }
    const int max_w = std::min(W, tam);
    float prefix[kMaxW + 1];
    prefix[0] = 0.0f;
    for (int w = 1; w <= max_w; ++w) prefix[w] = prefix[w - 1] + c.llr[P[w - 1]];

    bool have = false;
    float best = 0.0f;
    int best_w = 0, best_combo[kMaxW];
    for (int w = 1; w <= max_w; ++w) {
        if (have && prefix[w] >= best) break;
        const uint64_t total = C(tam, w);
        int combo[kMaxW];
        for (uint64_t r = 0; r < total; ++r) {
            ++count;
            unrank_lex(r, tam, w, combo);
            if (!parity_matches(c.H, s, P.data(), combo, w, c.n, c.m)) continue;
            float cost = 0.0f;
            for (int j = 0; j < w; ++j) cost += c.llr[P[combo[j]]];
            if (!have || cost < best) {
                have = true; best = cost; best_w = w;
                for (int j = 0; j < w; ++j) best_combo[j] = P[combo[j]];
            }
        }
    }
    if (!have) return false;
    std::fill(e_est.begin(), e_est.end(), 0);
    for (int j = 0; j < best_w; ++j) e_est[best_combo[j]] = 1;
    return true;
}

void print_shot(const Call &c, int shot, const GrandDecoderResult &r, bool with_errors) {
    if (c.fam == kCC && with_errors)
        printf("shot %d: found=%d, original_weight=%d, found_weight=%d, degenerate=%d, is_equal=%d, iteration=%d, tam=%d, total_tested_combinations=%llu, hybrid=%d, time_sum_ms=%.5f, total_time_search_ms=%.5f, transfer_time_ms=%.5f, total_to_find_ms=%.5f\n",
               shot + 1, r.found, r.original_weight, r.found_weight, r.degenerate, r.is_equal, r.iteration, r.tam,
               r.total_tested_combinations, r.hibrido, r.time_sum_ms, r.total_time_search_ms, r.transfer_time_ms, r.total_to_find_ms);
    else if (c.fam == kCC)
        printf("shot %d: found=%d, found_weight=%d, iteration=%d, tam=%d, total_tested_combinations=%llu, hybrid=%d, time_sum_ms=%.5f, total_time_search_ms=%.5f, transfer_time_ms=%.5f, total_to_find_ms=%.5f\n",
               shot + 1, r.found, r.found_weight, r.iteration, r.tam, r.total_tested_combinations, r.hibrido,
               r.time_sum_ms, r.total_time_search_ms, r.transfer_time_ms, r.total_to_find_ms);
    else
        printf("shot %d: found=%d, found_weight=%d, degenerate=%d, iteration=%d, tam=%d, total_tested_combinations=%llu, hybrid=%d, time_sum_ms=%.5f, total_time_search_ms=%.5f, transfer_time_ms=%.5f, total_to_find_ms=%.5f\n",
               shot + 1, r.found, r.found_weight, r.degenerate, r.iteration, r.tam, r.total_tested_combinations,
               r.hibrido, r.time_sum_ms, r.total_time_search_ms, r.transfer_time_ms, r.total_to_find_ms);
}

// One shot through the reference driver body: syndrome, column scores, offset,
// up to three active-set iterations with the mode/branch decision, search,
// post-processing.  cc: grand_decoder.cu:1199-1381 / :885-1102;
// cln: :1466-1654, :1762-1972, :2073-2255.
void decode_shot(const Call &c, const int *e_real, const int *syn_in, const int *logical, ShotOut &o) {
    const int n = c.n, m = c.m;
    GrandDecoderResult &res = o.res;
    memset(&res, 0, sizeof(res));
    o.e_est.assign(m, 0);
    o.l_est.assign(c.n_log, 0);

    std::vector<int> s(n, 0);
    int original_weight = 0;
    if (e_real) {
        for (int j = 0; j < m; ++j) if (e_real[j] >= 1) ++original_weight;     // :495-504
        for (int r = 0; r < n; ++r) {                                          // :476-486
            int v = 0;
            for (int j = 0; j < m; ++j) v ^= (c.H[(size_t)r * m + j] & e_real[j]);
            s[r] = v;
        }
    } else {
        for (int r = 0; r < n; ++r) s[r] = syn_in[r];
    }
    int sw = 0;
    for (int r = 0; r < n; ++r) if (s[r] >= 1) ++sw;

    res.original_weight = (c.fam == kCC) ? original_weight : 0;
    res.hibrido = -2;
    if (sw == 0) return;                                                       // :1223-1241

    // Column scores over the flipped checks (:352-374) and the irregular offset (:406-416).
    std::vector<int> score(m, 0);
    for (int r = 0; r < n; ++r)
        if (s[r] == 1)  // the reference copies rows with s == 1 into H_red (:357)
            for (int j = 0; j < m; ++j) score[j] += c.H[(size_t)r * m + j];
    if (!c.regular)
        for (int j = 0; j < m; ++j) score[j] = (c.colw[j] == 0) ? 0 : score[j] + (c.max_soma - c.colw[j]);

    const GrandDecoderConfig &cfg = c.cfg;
    std::vector<int> A(m);
    int it = 0, tam = 0, tam_old = 0, hib = -2;
    bool found = false;
    uint64_t count = 0;
    while (it < c.max_soma && it < 3) {
        const int thr = c.max_soma - it;
        ++it;
        tam = 0;
        for (int j = 0; j < m; ++j) if (score[j] >= thr) A[tam++] = j;
        if (c.fam == kCLN) {                                                   // :1550-1554
            if (tam > cfg.max_size_active_positions) tam = tam_old; else tam_old = tam;
        }
        if (tam <= 0) continue;

        const bool gpu_branch = (tam > cfg.threshold_for_hybrid && tam <= cfg.max_size_active_positions &&
                                 cfg.mode_of_search == 1) || cfg.mode_of_search == 0;
        const bool cpu_branch = !gpu_branch &&
            ((tam <= cfg.threshold_for_hybrid && cfg.mode_of_search == 1) || cfg.mode_of_search == -1);
        if (!gpu_branch && !cpu_branch) continue;                              // neither branch runs
        hib = gpu_branch ? (hib == -1 ? 1 : 0) : -1;

        if (c.kind == kML) {
            found = ml_search(c, s.data(), A.data(), tam, cfg.max_error_weight, count, o.e_est);
        } else if (c.kind == kUniform) {
            int wl[kMaxW], nw = 0;
            const int W = std::min(cfg.max_error_weight, kMaxW);
            for (int w = 1; w <= W; ++w) wl[nw++] = w;
            found = first_hit(c, s.data(), A.data(), tam, wl, nw, count, o.e_est);
        } else {
            const PhaseMask &pm = mask_for(tam, c.masks, c.n_masks);
            for (int k = 0; k < c.num_phases && !found; ++k) {
                int wl[kMaxW];
                const int nw = phase_weights(pm.order[k], tam, wl);
                if (nw == 0) continue;
                found = first_hit(c, s.data(), A.data(), tam, wl, nw, count, o.e_est);
            }
        }
        if (found) break;
    }

    res.iteration = it;
    res.tam = tam;
    res.total_tested_combinations = count;
    res.hibrido = hib;
    if (!found) return;

    res.found = 1;
    int fw = 0;
    for (int j = 0; j < m; ++j) fw += (o.e_est[j] >= 1);
    res.found_weight = fw;
    if (c.fam == kCC) {
        if (e_real) {                                                           // :1080-1087
            bool eq = true;
            for (int j = 0; j < m; ++j) if (e_real[j] != o.e_est[j]) { eq = false; break; }
            res.is_equal = eq;
            std::vector<int> delta(m);
            for (int j = 0; j < m; ++j) delta[j] = e_real[j] ^ o.e_est[j];
            res.degenerate = in_rowspace(c.R, m, c.piv, c.rank, delta);
        }
    } else {                                                                    // :1623-1644
        for (int j = 0; j < m; ++j)
            if (o.e_est[j] >= 1)
                for (int q = 0; q < c.n_log; ++q) o.l_est[q] ^= c.L[(size_t)j * c.n_log + q];
        bool same = true;
        for (int q = 0; q < c.n_log; ++q) if ((o.l_est[q] ^ logical[q]) == 1) same = false;
        res.degenerate = same;
    }
}

bool print_enabled() {
    const char *v = getenv("GRAND_PRINT_SHOTS");
    return v && v[0] == '1';
}

void run_shots(const Call &c, const int *errors, const int *syn, const int *logicals, int num_shots,
               GrandDecoderResult *results, int *decoded, int *decoded_logicals) {
    const bool pr = print_enabled();
    #pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1 && !pr)
    for (int shot = 0; shot < num_shots; ++shot) {
        ShotOut o;
        const int *e = errors ? errors + (size_t)shot * c.m : nullptr;
        const int *sy = (!errors && syn) ? syn + (size_t)shot * c.n : nullptr;
        const int *lg = logicals ? logicals + (size_t)shot * c.n_log : nullptr;
        decode_shot(c, e, sy, lg, o);
        results[shot] = o.res;
        if (decoded) memcpy(decoded + (size_t)shot * c.m, o.e_est.data(), sizeof(int) * c.m);
        if (decoded_logicals && c.n_log > 0)
            memcpy(decoded_logicals + (size_t)shot * c.n_log, o.l_est.data(), sizeof(int) * c.n_log);
        if (pr) print_shot(c, shot, o.res, errors != nullptr);
    }
}

int cc_entry(Search kind, const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
             const GrandDecoderConfig *config, GrandDecoderResult *results, char type, const char *prob_file,
             int *decoded) {
    Call c;
    c.fam = kCC; c.kind = kind; c.n = n; c.m = m; c.n_log = 0; c.cfg = *config;
    const int *other;
    if (type == 'x' || type == 'X') { c.H = Hz; other = Hx; }                   // :833-846
    else if (type == 'z' || type == 'Z') { c.H = Hx; other = Hz; }
    else { fprintf(stderr, "Invalid type_of_error: %c. Must be 'x' or 'z'.\n", type); return -1; }
    const int W = config->max_error_weight;
    if (kind == kPhase) {
        c.n_masks = read_masks(prob_file, c.masks);                            // :849
        if (W == 6) c.num_phases = 2; else if (W == 7) c.num_phases = 3;
        else if (W > 0 && W < 6) c.num_phases = 1;
        else { fprintf(stderr, "Invalid max_error_weight: %d. Must be 6 or 7.\n", W); return -1; }
        if (c.n_masks == 0) {   // the reference reads masks[-1] here (grand_decoder.cu:761)
            fprintf(stderr, "No phase order masks read from: %s\n", prob_file ? prob_file : "(null)");
            return -1;
        }
    } else if (W <= 0 || W > 7) {
        fprintf(stderr, "Invalid max_error_weight: %d. Must be between 1 and 7.\n", W);
        return -1;
    }
    column_stats(c);
    c.R.assign(other, other + (size_t)n * m);
    c.piv.assign(std::max(n, 1), 0);
    c.rank = rref(c.R, n, m, c.piv);
    run_shots(c, errors, syndromes, nullptr, num_shots, results, decoded, nullptr);
    return 0;
}

int cln_entry(Search kind, const int *H_dem, const int *L_dem, int n, int m, int n_log, const int *detectors,
              const int *logicals, int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
              const char *file, int *decoded, int *decoded_logicals) {
    Call c;
    c.fam = kCLN; c.kind = kind; c.n = n; c.m = m; c.n_log = n_log; c.cfg = *config;
    c.H = H_dem; c.L = L_dem;
    const int W = config->max_error_weight;
    if (W <= 0 || W > 7) {
        fprintf(stderr, "Invalid max_error_weight: %d. Must be between 1 and 7.\n", W);
        return -1;
    }
    if (kind == kML) {
        c.llr.assign(m, 0.0f);
        if (!read_llr(file, c.llr.data(), m)) return -1;                       // :2048
    }
    if (kind == kPhase) {
        c.n_masks = read_masks(file, c.masks);                                 // :1723-1731
        if (c.n_masks == 0) {
            fprintf(stderr, "No phase order masks read from: %s\n", file ? file : "(null)");
            return -1;
        }
        c.num_phases = (W == 6) ? 2 : (W == 7) ? 3 : 1;
    }
    column_stats(c);
    run_shots(c, nullptr, detectors, logicals, num_shots, results, decoded, decoded_logicals);
    return 0;
}

}  // namespace

extern "C" {

void oracle_set_threads(int t) { g_threads = t < 1 ? 1 : t; }

// Reference binomial and unranking, exported so tests can pin them directly.
unsigned long long oracle_nCr(int n, int k) { return C(n, k); }
void oracle_unrank(unsigned long long rank, int n, int k, int *out) { unrank_lex(rank, n, k, out); }

int grand_decode_cc_ex(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                       const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                       const GrandDecoderExConfig *, int *decoded) {
    return cc_entry(kUniform, Hx, Hz, n, m, errors, syndromes, num_shots, config, results, type_of_error, nullptr, decoded);
}
int grand_decode_prob_cc_ex(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                            const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                            const char *probability_file, const GrandDecoderExConfig *, int *decoded) {
    return cc_entry(kPhase, Hx, Hz, n, m, errors, syndromes, num_shots, config, results, type_of_error,
                    probability_file, decoded);
}
int grand_decode_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                        int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
                        const GrandDecoderExConfig *, int *decoded, int *decoded_logicals) {
    return cln_entry(kUniform, H_dem, L_dem, n, m, n_logicos, detectors, logicals, num_shots, config, results,
                     nullptr, decoded, decoded_logicals);
}
int grand_decode_prob_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors,
                             int *logicals, int num_shots, const GrandDecoderConfig *config,
                             GrandDecoderResult *results, const char *probability_file,
                             const GrandDecoderExConfig *, int *decoded, int *decoded_logicals) {
    return cln_entry(kPhase, H_dem, L_dem, n, m, n_logicos, detectors, logicals, num_shots, config, results,
                     probability_file, decoded, decoded_logicals);
}
int grand_decode_ML_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors,
                           int *logicals, int num_shots, const GrandDecoderConfig *config,
                           GrandDecoderResult *results, const char *llr_file, const GrandDecoderExConfig *,
                           int *decoded, int *decoded_logicals) {
    return cln_entry(kML, H_dem, L_dem, n, m, n_logicos, detectors, logicals, num_shots, config, results,
                     llr_file, decoded, decoded_logicals);
}

int grand_decode_cc(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                    const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error) {
    return grand_decode_cc_ex(Hx, Hz, n, m, errors, syndromes, num_shots, config, results, type_of_error, nullptr, nullptr);
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

}  // extern "C"
