/*
 * grand_decoder.h -- C ABI of the MI355X GRAND decoder (libgranddecoder.so).
 *
 * Drop-in for the reference library's ABI (CarolinaBoavida/Library_S_GRAND,
 * grand_decoder.h): the five entry points keep their names, argument order,
 * argument meaning, return convention (0 = ok, -1 = invalid input with a
 * message on stderr) and the byte layout of both structs, so existing ctypes
 * callers (Examples/auxilary_functions.py:172-253) bind it unchanged.
 *
 * Unlike the reference header this one is valid C as well as C++.
 *
 * Layout contract (checked by tests/test_abi_cpu.py):
 *   GrandDecoderConfig  24 bytes, 6 x int32.
 *   GrandDecoderResult  72 bytes; total_tested_combinations at offset 24,
 *                       hibrido at 32, iteration at 36, time_sum_ms at 40.
 *
 * Matrices are row-major int32 holding 0/1: Hx, Hz, H_dem are n x m; errors
 * are num_shots x m; syndromes and detectors num_shots x n; L_dem is
 * m x n_logicos; logicals (an input: the observed logical flips) is
 * num_shots x n_logicos.  Every pointer is caller-owned host memory.
 */
#ifndef GRAND_DECODER_H
#define GRAND_DECODER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest error weight any search enumerates (reference grand_decoder.h:28).
 * max_error_weight 7 is accepted and searched as 6, as in the reference. */
#define GRAND_MAX_W 6

/* Reference grand_decoder.h:37-44.  threads_per_block and use_gpu_sum are
 * accepted and ignored: launch shapes come from the autotuner and the column
 * scores are always computed on the device.  mode_of_search keeps the
 * reference's meaning for which active-set iterations are searched and for the
 * `hibrido` code (-1 = "CPU" branch, 0 = "GPU" branch, 1 = hybrid with the
 * threshold_for_hybrid / max_size_active_positions window, any other value
 * searches nothing); every searched iteration runs on the GPU with the
 * reference CPU search's semantics. */
typedef struct {
    int max_error_weight;
    int threads_per_block;
    int use_gpu_sum;
    int max_size_active_positions;
    int threshold_for_hybrid;
    int mode_of_search;
} GrandDecoderConfig;

/* Reference grand_decoder.h:53-68.  One record per shot.
 * total_tested_combinations counts patterns as the reference CPU search does:
 * up to and including the hit.  The four timing doubles (not parity fields):
 *   time_sum_ms          device time from the start of the shot's search to
 *                        its first active list (column scores included);
 *   total_time_search_ms device time of the shot's search;
 *   transfer_time_ms     host<->device copies of the whole call (inputs in,
 *                        results and decoded patterns out, from HIP events),
 *                        an equal share per shot: the library copies all
 *                        shots at once, not one shot at a time;
 *   total_to_find_ms     the search time when found, else 0.
 * Shots with an empty syndrome have 0 in every timing field except
 * transfer_time_ms.  The _ex entry points fill the same fields. */
typedef struct {
    int found;
    int found_weight;
    int original_weight;
    int degenerate;
    int is_equal;
    int tam;
    unsigned long long total_tested_combinations;
    int hibrido;
    int iteration;
    double time_sum_ms;
    double total_time_search_ms;
    double transfer_time_ms;
    double total_to_find_ms;
} GrandDecoderResult;

/* ---- the reference ABI ---------------------------------------------------- */

/* Replaces grand_decode_prob_cc (reference grand_decoder.h:92,
 * grand_decoder.cu:790-1115).  Phase-order search: the probability file holds
 * up to 20 lines "tam_upper order0 order1 order2 p0 p1 p2"; phase 0 searches
 * weights 1..5, phase 1 weight 6. */
int grand_decode_prob_cc(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                         const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                         const char *probability_file);

/* Replaces grand_decode_cc (reference grand_decoder.h:112, grand_decoder.cu:1117-1396).
 * Uniform order: weights 1..min(max_error_weight, 6). errors may be NULL, in
 * which case syndromes are decoded; with errors, is_equal and degenerate are
 * filled. */
int grand_decode_cc(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                    const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error);

/* Replaces grand_decode_cln (reference grand_decoder.h:133, grand_decoder.cu:1399-1670). */
int grand_decode_cln(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                     int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results);

/* Replaces grand_decode_prob_cln (reference grand_decoder.h:154, grand_decoder.cu:1672-1994). */
int grand_decode_prob_cln(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                          int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
                          const char *probability_file);

/* Replaces grand_decode_ML_cln (reference grand_decoder.h:176, grand_decoder.cu:1996-2277).
 * The LLR file holds m floats.  Selection is the reference CPU ML search
 * (grand_decoder.cu:590-666): stable LLR sort of the active set, fp32
 * sequential costs, first strictly smaller cost wins, prefix prune. */
int grand_decode_ML_cln(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                        int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
                        const char *llr_file);

/* ---- extended entry points ---------------------------------------------------
 * The reference ABI never returns the decoded pattern.  The _ex variants take
 * the same arguments plus an optional GrandDecoderExConfig and optional
 * outputs: decoded (num_shots x m int32, the decoded error pattern, zero when
 * not found) and, for the DEM family, decoded_logicals (num_shots x n_logicos,
 * the XOR of L_dem rows over the decoded pattern, zero when not found).
 * They have no counterpart in the reference; the reference entry points are
 * the _ex ones with ex = NULL and no outputs. */

#define GRAND_ENGINE_FUSED 0  /* one persistent kernel for the whole batch (default) */
#define GRAND_ENGINE_PLAIN 1  /* host loop over shots, one launch per weight class    */

typedef struct {
    int engine;          /* GRAND_ENGINE_FUSED or GRAND_ENGINE_PLAIN */
    int waves_per_cu;    /* persistent-grid occupancy; 0 = autotuned */
    int claim_batch;     /* shots claimed per queue access; 0 = autotuned */
    int split_threshold; /* weight classes larger than this are shared across waves; 0 = autotuned, -1 = never */
    int triage_search;   /* fused engine: shots whose first iteration is small are searched one per lane in the
                            triage kernel; 0 = on where it applies, -1 = off (every shot goes to the search kernel) */
    int pair_table_kb;   /* first-hit kinds: LDS per search wave for the pair table of deep classes; 0 = autotuned,
                            -1 = no pair table, k > 0 = at most k KiB (never more than the wave's share of the CU) */
    int reserved[2];
} GrandDecoderExConfig;

int grand_decode_cc_ex(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                       const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                       const GrandDecoderExConfig *ex, int *decoded);
int grand_decode_prob_cc_ex(const int *Hx, const int *Hz, int n, int m, int *errors, int *syndromes, int num_shots,
                            const GrandDecoderConfig *config, GrandDecoderResult *results, char type_of_error,
                            const char *probability_file, const GrandDecoderExConfig *ex, int *decoded);
int grand_decode_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors, int *logicals,
                        int num_shots, const GrandDecoderConfig *config, GrandDecoderResult *results,
                        const GrandDecoderExConfig *ex, int *decoded, int *decoded_logicals);
int grand_decode_prob_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors,
                             int *logicals, int num_shots, const GrandDecoderConfig *config,
                             GrandDecoderResult *results, const char *probability_file,
                             const GrandDecoderExConfig *ex, int *decoded, int *decoded_logicals);
int grand_decode_ML_cln_ex(int *H_dem, const int *L_dem, int n, int m, int n_logicos, int *detectors,
                           int *logicals, int num_shots, const GrandDecoderConfig *config,
                           GrandDecoderResult *results, const char *llr_file, const GrandDecoderExConfig *ex,
                           int *decoded, int *decoded_logicals);

#ifdef __cplusplus
}
#endif

#endif /* GRAND_DECODER_H */
