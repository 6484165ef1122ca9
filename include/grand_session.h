/*
 * grand_session.h -- device-resident decoding sessions (libgranddecoder.so).
 *
 * No counterpart in the reference, whose five entry points take host buffers
 * and rebuild every device structure per call (grand_decoder.cu:790-2277).  A
 * session uploads one code (packed check matrix, signatures, RREF, L_dem,
 * LLRs, phase masks) once and then decodes batches whose inputs and outputs
 * already live in HBM, asynchronously on a caller's HIP stream.  The five ABI
 * entry points are built on it; the benchmark, the serving loop and graph
 * replay use it directly.
 *
 * Kinds: GRAND_KIND_CC (grand_decode_cc), GRAND_KIND_PROB_CC
 * (grand_decode_prob_cc), GRAND_KIND_CLN, GRAND_KIND_PROB_CLN,
 * GRAND_KIND_ML_CLN.  For the CSS kinds `H` is Hx and `H2` is Hz, selected by
 * type_of_error as in the ABI; for the DEM kinds `H` is H_dem and `H2` is L_dem.
 * `file` is the probability (phase-mask) file or the LLR file.
 */
#ifndef GRAND_SESSION_H
#define GRAND_SESSION_H

#include "grand_decoder.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GRAND_KIND_CC 0
#define GRAND_KIND_PROB_CC 1
#define GRAND_KIND_CLN 2
#define GRAND_KIND_PROB_CLN 3
#define GRAND_KIND_ML_CLN 4

typedef struct GrandSession GrandSession;

/* Validates like the matching ABI entry point (same messages, same order) and
 * uploads the code.  Returns NULL on invalid input. */
GrandSession *grand_session_create(int kind, const int *H, const int *H2, int n, int m, int n_logicos,
                                   char type_of_error, const GrandDecoderConfig *config, const char *file,
                                   const GrandDecoderExConfig *ex);

/* Decodes num_shots shots.  d_input: num_shots x m error rows (CSS kinds with
 * input_is_errors = 1) or num_shots x n syndrome/detector rows; d_logicals:
 * num_shots x n_logicos (DEM kinds).  d_results: num_shots records.
 * d_decoded: optional num_shots x ceil(m/32) uint32 words of the decoded
 * pattern; d_decoded_logicals: optional num_shots x ceil(n_logicos/32) words.
 * d_stats: optional num_shots x 3 uint64: device tick when the shot's search
 * started, tick when it completed (0, 0 for zero-syndrome shots), and the
 * weighted pattern count  sum over tested patterns of (weight + 1).
 * All pointers are device memory; work is queued on `stream` (a hipStream_t,
 * NULL = the default stream, so work the caller queued there comes first) and
 * the call returns without waiting.
 * Returns 0, or -1 for invalid arguments. */
int grand_session_decode_device(GrandSession *s, const int *d_input, int input_is_errors, const int *d_logicals,
                                int num_shots, GrandDecoderResult *d_results, unsigned *d_decoded,
                                unsigned *d_decoded_logicals, unsigned long long *d_stats, void *stream);

/* Decode without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  As
 * grand_session_decode_device, same arguments and results, but the decode's
 * device scratch (packed syndromes, worklist, counters, shared-task table,
 * the launch descriptor) comes from a fixed array of 32 MiB that the library
 * holds on every GPU it runs on until the process ends; it calls no hipMalloc
 * or hipFree.  The session itself (the uploaded code) is created beforehand
 * with grand_session_create.  A batch whose scratch exceeds the array is
 * rejected with -1 and a message on stderr that names the most shots of the
 * session's code that fit (about 440,000 at BB144, 41,000
 * on a 216 x 3000 DEM); so is the plain engine.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (concurrent streams or threads) race. */
int grand_session_decode_device_noalloc(GrandSession *s, const int *d_input, int input_is_errors,
                                        const int *d_logicals, int num_shots, GrandDecoderResult *d_results,
                                        unsigned *d_decoded, unsigned *d_decoded_logicals,
                                        unsigned long long *d_stats, void *stream);

/* Learned search order.  As grand_session_decode_device, but every shot first
 * runs a learned pass over its first active list: positions sorted by
 * descending score, patterns of weight 1..max_weight tried so that a pattern
 * comes as soon as its worst-scored position does, each checked exactly
 * against the syndrome.  The first match is the result; after `budget` patterns
 * without one, the plain search runs as usual.  total_tested_combinations
 * counts the patterns of both.  d_scores: device float [num_shots][stride],
 * indexed by position in the first active list (scores_per_column = 0, stride
 * >= the list length, longer lists skip the pass) or by column
 * (scores_per_column = 1, stride >= m).  Not for GRAND_KIND_ML_CLN or the
 * plain engine.  budget = 0 gives the plain results.  Returns 0 or -1. */
int grand_session_decode_learned(GrandSession *s, const int *d_input, int input_is_errors, const int *d_logicals,
                                 int num_shots, const float *d_scores, int stride, int scores_per_column,
                                 int max_weight, unsigned long long budget, GrandDecoderResult *d_results,
                                 unsigned *d_decoded, unsigned *d_decoded_logicals, unsigned long long *d_stats,
                                 void *stream);

/* HIP graph of one decode, for callers that repeat the same call (same
 * buffers and shot count, new contents).  grand_session_capture records the
 * kernels grand_session_decode_device would queue (descriptor install,
 * triage, search; no memory set or copy nodes) on `stream`, which must be a
 * non-default stream, and instantiates them; every grand_graph_launch decodes
 * the current contents of the buffers.  Allocation happens before capture.
 * Launches of one graph, and decodes on its session, must not overlap.  Not
 * for the plain engine.  Returns NULL on failure. */
typedef struct GrandGraph GrandGraph;
GrandGraph *grand_session_capture(GrandSession *s, const int *d_input, int input_is_errors, const int *d_logicals,
                                  int num_shots, GrandDecoderResult *d_results, unsigned *d_decoded,
                                  unsigned *d_decoded_logicals, unsigned long long *d_stats, void *stream);
int grand_graph_launch(GrandGraph *g, void *stream);
void grand_graph_destroy(GrandGraph *g);

/* Error bits raised by the kernels of the last decode on this session
 * (1 = an input value other than 0/1, 2 = an active set larger than the
 * supported 1024 positions); synchronizes `stream`. */
int grand_session_error_bits(GrandSession *s, void *stream);

/* Queues a copy of the error-bit word of the last decode into *host_dst
 * (pinned host memory) on `stream`, without waiting. */
int grand_session_error_bits_async(GrandSession *s, unsigned *host_dst, void *stream);

/* Allocates the scratch of a num_shots batch now, so that later decodes of up
 * to that many shots allocate nothing.  Returns 0, or -1 when out of memory. */
int grand_session_reserve(GrandSession *s, int num_shots);

/* HIP-event timing of the kernels of each decode, recorded on the decode's
 * stream: triage (pack + syndrome + worklist) and search (the persistent
 * kernel, or the whole plain loop).  The session keeps the events of its last
 * 64 timed decodes, so timed decodes still run back to back; last_timing waits
 * for the last decode, timing_totals for all of them and returns the sums of
 * the times since the last timing_reset and the number of decodes summed. */
int grand_session_enable_timing(GrandSession *s, int on);
int grand_session_last_timing(GrandSession *s, float *triage_ms, float *search_ms);
int grand_session_timing_totals(GrandSession *s, double *triage_ms, double *search_ms, long long *decodes);
int grand_session_timing_reset(GrandSession *s);

/* Work counters of the search kernel for later decodes on this session (no
 * counterpart in the reference): d_work is device memory of 3 uint64 that the
 * kernels add to (the caller zeroes it): [0] (w-1)-prefix steps walked, one
 * signature-table probe each; [1] (w-2)-stem steps walked, one pair-table
 * lookup each; [2] patterns tested one per lane (small classes, weight 1).
 * Steps walked past a lane's hit are not counted; every other step is, so
 * these are the units the hardware executed.  NULL stops counting.
 * Returns 0, or -1 for a NULL session. */
int grand_session_set_work_counters(GrandSession *s, unsigned long long *d_work);

/* Device wall-clock period in milliseconds per tick (for d_stats). */
double grand_session_ms_per_tick(GrandSession *s);

/* Words per decoded pattern (ceil(m/32)) and per decoded logical row. */
int grand_session_pattern_words(GrandSession *s);
int grand_session_logical_words(GrandSession *s);

/* Bytes of device memory the session holds for its code plus the scratch a
 * batch of num_shots needs (the serving loop admits requests by this). */
size_t grand_session_footprint(GrandSession *s, int num_shots);

void grand_session_destroy(GrandSession *s);

/* Compute units, peak engine clock and device wall-clock rate (kHz) of the
 * current device.  Returns 0, or -1 without a usable GPU. */
int grand_device_info(int *cus, int *clock_khz, int *wall_clock_khz);

/* Process-wide knobs of the persistent search kernel, read by every decode
 * (values <= 0 keep the current setting): workgroups per CU, shots per
 * worklist claim (at most 8), prefixes per lane per chunk, and the class size
 * in prefixes above which a weight class is shared by all waves.  Built-in
 * defaults: 1, 2, 64, 65536 for the first-hit kinds and 8, 2, 64, 32768 for the
 * ML kind; a set replaces both, a get reads the first-hit set, and a set with
 * all four values negative restores the built-in defaults. */
void grand_set_tuning(int waves_per_cu, int claim_batch, int rmax, int split_threshold);
void grand_get_tuning(int *waves_per_cu, int *claim_batch, int *rmax, int *split_threshold);

/* LDS packing of the search kernel (first-hit kinds), process-wide like the
 * knobs above: the pair table of deep weight classes takes the LDS a wave's
 * share of the CU has left after its hash table, bitmap, code and claimed
 * shots.  kb = 0 gives it all of that (built-in default), kb > 0 at most kb
 * KiB, kb < 0 none (deep classes are then walked prefix by prefix); classes
 * whose pairs do not fit are walked too, so every value decodes the same.
 * grand_set_tuning with all four values negative also restores 0. */
void grand_set_pair_table_kb(int kb);
int grand_get_pair_table_kb(void);

/* Degeneracy words of a CSS pair (host only, no GPU needed): H is the matrix
 * the syndromes come from and other the one whose row space decides
 * `degenerate` (n x m, row-major int 0/1 each).  For v in ker(H), v is in
 * rowspace(other) iff the XOR of words[c] over c in supp(v) is 0.  Returns the
 * number of functionals (at most 32) and fills words[0..m), or -1 when more
 * than 32 would be needed (the decoder then walks the RREF rows instead). */
int grand_degeneracy_words(const int *H, const int *other, int n, int m, unsigned *words);

#ifdef __cplusplus
}
#endif

#endif /* GRAND_SESSION_H */
