// grand_types.h -- device-resident code/call descriptors shared by the host
// driver (grand_host.cpp) and the gfx950 kernels (grand_kernels.hip).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include "../../include/grand_decoder.h"

// Symbols shared between the library's own translation units only (kernel
// launchers); they stay out of the shared library's exported interface.
#define GRAND_INTERNAL __attribute__((visibility("hidden")))

namespace grand {

constexpr int kMaxW = GRAND_MAX_W;   // largest enumerated weight of the plain search
constexpr int kMaxWin = 8;            // most columns of a decoded pattern (the learned pass goes up to 8)
constexpr int kMaxTam = 1024;        // largest active set (reference: 300, grand_decoder.cu:13)
constexpr int kMaxWords = 16;        // 32-bit words per packed column -> n <= 512 checks
constexpr int kMaxM = 4096;          // columns -> <= 128 packed words per pattern
constexpr int kMaxMasks = 20;        // phase-order masks (grand_decoder.cu:11)
constexpr int kNcrStride = 8;        // nCr table row stride (k = 0..7)
constexpr int kNegScore = -(1 << 20);

enum Family : int { kCC = 0, kCLN = 1 };
enum Kind : int { kUniform = 0, kPhase = 1, kML = 2 };

// Error bits raised by kernels and turned into a -1 return by the host.
enum ErrBits : unsigned { kErrNonBinary = 1u, kErrTamOverflow = 2u, kErrTaskWait = 4u };

// Immutable per-code device data.  Column bits are kept twice: word-major
// (Hcw[d * m + c]) for coalesced score computation, column-major
// (Hcm[c * D + d]) for the exact check of a candidate.
struct DevCode {
    int n, m, D, M32, NL, NL32;
    int max_soma;
    int rank;                         // rank of the other CSS matrix (cc)
    const uint32_t *Hcw;              // [D][m]
    const uint32_t *Hcm;              // [m][D]
    const uint32_t *Hrow;             // [n][M32]  rows of H, for s = H e
    const uint32_t *sig;              // [m] 32-bit GF(2)-linear signature of each column
    const uint32_t *rowsig;           // [n] signature of each unit row vector
    const int *off;                   // [m] score offset (irregular codes), kNegScore for empty columns
    const uint32_t *R;                // [rank][M32] RREF of the other CSS matrix (cc)
    const int *pivrow;                // [m] RREF row whose pivot is this column, -1 for other columns
    const uint32_t *Lp;               // [m][NL32] logical flips per mechanism (cln)
    const uint32_t *dg;               // [m] degeneracy words (cc), or null when more than 32 are needed
    // the triage kernel's tables as one image of timg_words words: column
    // words [D][m] at 0, signatures [m] at t_sig, degeneracy words [m] at t_dg
    // (when dg is set); then, 16-byte aligned at t_rows, the rows of H
    // [n][t_rw] (the odd stride t_rw = M32 | 1: no bank conflicts between lanes
    // reading different rows), and at t_pl the column-weight planes [7][M32]:
    // the nonempty columns, the three bit planes of colw and the three bit
    // planes of colw - 1, then kLaneMaxM32 + 2 zero words.  lane_ok: the code
    // fits the lane search (D <= kLaneMaxD, M32 <= kLaneMaxM32, colw <=
    // kLaneMaxColw).
    const uint32_t *timg;
    int timg_words, t_sig, t_dg, t_rows, t_rw, t_pl, lane_ok;
    const float *llr;                 // [m] (ML)
    const unsigned long long *nCr;    // [(kMaxTam + 1) * kNcrStride]
    // Small codes are staged in the search kernel's LDS.  The search kernel gets
    // its own copy of this descriptor with lds_words > 0; every other kernel's
    // copy has 0 and reads global memory.  Staged words, in order: column words
    // [D][m], score offsets [m], signatures [m], then for cc either (lds_post = 2)
    // the degeneracy words [m] or (lds_post = 1) the RREF row of each pivot
    // column [m] and the RREF rows [rank][M32].
    int lds_words;
    int lds_post;
};

// A weight class too large for one wave, shared through the task table.  The
// owner wave publishes it; any wave claims chunks of 64 * R prefixes with one
// fetch-and-add on ctl = (generation << 44) | (nchunks << 24) | next_chunk.
constexpr int kTaskSlots = 128;
constexpr int kCodeLdsMaxWords = 4096;   // codes up to 16 KB of staged words (DevCode::lds_words) go to LDS
constexpr int kMaxBatch = 8;             // most shots per worklist claim
// Pair table of the search kernel (first-hit classes of weight >= 3)
// Bucket heads by signature bits 16..: 2048 for the first-hit kinds (lists of up to 4095
// pairs, short chains), 1024 for ML, whose waves hold a quarter of a CU's LDS each, so that a
// list of 64 columns (2016 pairs) fits with its heads.
constexpr int kPairBucketsFirstHit = 2048;
constexpr int kPairBucketsMl = 1024;
constexpr int pair_buckets(bool ml) { return ml ? kPairBucketsMl : kPairBucketsFirstHit; }
constexpr int kPairBitmapWords = 1024;   // 2^15-bit filter, by signature bits 0..14
constexpr unsigned kPairNone = 0xFFFu;   // end of a bucket chain; also the most pairs a table holds
// bucket heads, a spare head (never looked up), padding
constexpr int pair_head_words(bool ml) { return pair_buckets(ml) + 4; }
constexpr size_t kLdsPerCu = 160 * 1024; // gfx950
// Search of a shot by its own lane in the triage kernel (first two iterations).
// Active-list entries a lane keeps: a longer list makes every lane of its wave
// run the longer register compares, which costs the triage kernel more than the
// search kernel spends on those shots (BB144 prob_cc, 100k shots, per decode on
// one / two streams: 138.0 / 78.7 us at 7, 139.2 / 80.1 at 8, 144.7 / 83.1 at 12,
// 147.8 / 82.6 at 5)
constexpr int kLaneMax = 7;
constexpr int kLaneBudget = 300;         // patterns a lane may test
constexpr int kLaneMaxD = 6;             // codes with at most 6 syndrome words ...
constexpr int kLaneMaxM32 = 10;          // ... at most 320 columns (bit-sliced column counts in registers) ...
constexpr int kLaneMaxColw = 7;          // ... and column weights of at most 7 (three count planes)
constexpr int kTriageWlWords = 32;       // LDS words of the lane search's weight lists (one u64 per list size)

// Claim words of the persistent search kernel: kClaimWords 64-bit words, one
// per 128-byte line, starting at 32-bit counter index 2 * kClaimBase.
constexpr int kClaimWords = 8;
constexpr int kClaimBase = 16;        // in 64-bit words: byte offset 128
constexpr int kClaimStride = 16;      // in 64-bit words: 128 bytes
// Worklist partitions: group g of the triage kernel (shots 64g .. 64g + 63)
// appends the shots it queues to partition g % wl_parts, whose front and back
// counters (32-bit counter indices kWlBase + kWlStride * k and one after) sit on
// a 128-byte line of their own after the claim words.  Partition k owns the
// worklist slots of its groups' count of 64-slot rows, starting at
// 64 * (k * (G / P) + min(k, G % P)) for G groups and P partitions; fronts fill
// it from the start, backs from the end.
constexpr int kWlParts = 16;
constexpr int kWlBase = 2 * (kClaimBase + kClaimWords * kClaimStride);
constexpr int kWlStride = 32;
constexpr int kCounterBytes = 4 * (kWlBase + kWlParts * kWlStride);
constexpr int kTaskClasses = 6;          // weight classes one shared task can hold
// Fixed device scratch of the decode that allocates nothing: its size, and the
// most devices whose address of it the library keeps.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kScratchDevices = 64;
struct TaskSlot {
    unsigned long long ctl;           // generation | nchunks | next chunk to claim
    unsigned long long best;          // first-hit: lowest matching pattern index over the task's classes
                                      // (classes concatenated in list order); ML: (cost key << 32) | rank
    unsigned long long best_q;        // first-hit: generation << 44 | unit index of that hit over the task's
                                      // classes (later chunks are skipped)
    unsigned done;                    // chunks finished
    unsigned busy;                    // 0 free, 1 owned, 2 withdrawn (free once done == nchunks)
    unsigned nchunks;
    int shot, tam, ml;
    int nclass;                       // weight classes of the task, consecutive in the list (ML: 1)
    int wpack;                        // their weights, 4 bits each, first class in the low bits
    int ppack;                        // bit i: class i is walked by (w-2)-stems and the pair table
    unsigned cch[kTaskClasses + 1];   // first chunk of each class; cch[nclass] = nchunks
    int cR[kTaskClasses];             // units (prefixes or stems) per lane per chunk of each class
    uint32_t t0;                      // signature of the syndrome
    uint32_t s[kMaxWords];            // syndrome words
};

// Byte offsets the compiled-code check of the publication order relies on.
static_assert(offsetof(TaskSlot, ctl) == 0 && offsetof(TaskSlot, best_q) == 16 && offsetof(TaskSlot, done) == 24 &&
                  offsetof(TaskSlot, busy) == 28,
              "TaskSlot layout");

// Per-call parameters and per-shot buffers.
struct DevCall {
    int family, kind;
    int W;                            // max_error_weight clamped to kMaxW
    int mode, thr_h, max_size;
    int num_phases, n_masks;
    int mask_tam_upper[kMaxMasks];
    int mask_order[kMaxMasks][3];
    int cap;                          // active-list capacity held in LDS (<= kMaxTam)
    int rmax;                         // patterns per lane per chunk in the range path
    int claim_batch;                  // shots taken per worklist claim (<= kMaxBatch)
    int triage_search;                // 1: the triage kernel searches small first iterations one shot per lane
    int split_worklist;               // 1 (search kernel): worklist in two segments, likely-deep shots first
    int wl_parts;                     // worklist partitions (1 for the plain engine, else kWlParts)
    int pair_cap;                     // search kernel: pairs the LDS pair table holds (0 = no pair table)
    int triage_stage_words;           // triage kernel: words of its table image staged in LDS
                                      // (DevCode::timg), 0 = read from global memory
    int stage_words;                  // search kernel: LDS words per claimed shot: id, signature, syndrome,
                                      // then original weight, degeneracy word and error words (cc with
                                      // errors) or the observed logicals (cln)
    int num_shots;
    int input_errors;                 // 1: input rows are error patterns (cc), 0: syndromes/detectors
    double ms_per_tick;               // device wall clock period
    const int *input;                 // [shots][m] or [shots][n]
    const int *logicals;              // [shots][NL] (cln)
    uint32_t *s_packed;               // [shots][D]
    uint32_t *e_packed;               // [shots][M32] (cc with errors)
    uint32_t *e_aux;                  // [shots][2] original weight and XOR of dg over supp(e) (cc with errors)
    uint32_t *lg_packed;              // [shots][NL32] (cln)
    uint32_t *s_sig;                  // [shots]
    int *worklist;                    // nonzero-syndrome shots, 64 slots per triage group
    unsigned *counters;               // [2] error bits, [4..6] task counters, claim words, then the
                                      // front and back lengths of each worklist partition
    GrandDecoderResult *results;      // [shots]
    uint32_t *dec;                    // [shots][M32] decoded pattern
    uint32_t *declog;                 // [shots][NL32] decoded logicals (cln)
    unsigned long long *stats;        // [shots][3] claim tick, completion tick, weighted pattern count; or null
    TaskSlot *tasks;                  // [kTaskSlots]
    uint32_t *task_arrays;            // [kTaskSlots][3][cap]: active list, signatures, sorted LLRs
    unsigned *task_counters;          // [0] open tasks, [1] waves finished with their own shots
    unsigned long long split_threshold;  // classes larger than this are shared (0 = never)
    // optional work counts of the search kernel (null: not counted), summed over
    // the launch: [0] (w-1)-prefix steps (one hash-table probe each, hit or
    // not), [1] (w-2)-stem steps (one pair-table lookup each), [2] patterns
    // tested one per lane (small classes and weight 1)
    unsigned long long *work;
    // learned search order, tried before the plain search (lscores == null: off)
    const float *lscores;             // per shot: [lstride] scores, by slot of the first active list or by column
    int lstride;
    int lper_column;                  // 1: lscores indexed by column, 0: by position in the first active list
    int lmaxw;                        // pattern weights 1..lmaxw are tried (<= kMaxWin)
    unsigned long long lbudget;       // patterns the learned pass may test before plain search takes over
};

// Search outcome of one shot, handed from a search engine to the finalizer.
struct ShotOutcome {
    int found, fw, iteration, tam, hibrido;
    int win[kMaxWin];                 // decoded columns (ascending within the pattern order)
    unsigned long long count;
    unsigned long long wops;          // sum over tested patterns of (weight + 1)
    unsigned long long t0, t1;        // device timestamps: search start and end
    unsigned long long ta;            // device timestamp when the first active list was built (0: none)
};

}  // namespace grand
