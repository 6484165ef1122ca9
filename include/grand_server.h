/*
 * grand_server.h -- continuous-batching decode server (libgranddecoder.so).
 *
 * No counterpart in the reference, which decodes one synchronous batch per
 * call (grand_decoder.cu:790-2277).  A server keeps several codes resident on
 * the GPU (one session per lane, workspaces and pinned staging allocated up
 * front) and decodes requests that arrive at any time:
 *
 *   - grand_server_add_code registers a code; it is admitted only if its HBM
 *     footprint (code + lanes x the scratch of a full batch) fits the budget;
 *   - grand_server_submit queues a request (host rows in, host results out);
 *     it is refused when the queue already holds max_queued_shots shots;
 *   - a worker thread keeps every lane busy: as soon as a lane is free it takes
 *     all pending requests of its code (up to max_batch_shots shots, requests
 *     whole and in arrival order), stages them, and launches one decode on the
 *     lane's own stream, so batches of different codes and successive batches
 *     of one code run concurrently;
 *   - grand_server_wait / poll return a request's status and latency (submit
 *     to results copied out).
 *
 * Results of a request are those of decoding its rows alone: shots are
 * independent, and the batch a request lands in changes nothing but latency.
 */
#ifndef GRAND_SERVER_H
#define GRAND_SERVER_H

#include "grand_decoder.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GrandServer GrandServer;

typedef struct {
    int max_batch_shots;         /* shots per launched batch and per request (default 16384) */
    int lanes_per_code;          /* concurrent batches per code (default 2) */
    int max_queued_shots;        /* queued shots beyond which submit refuses (default 1 << 20) */
    int reserved;
    long long hbm_budget_bytes;  /* device memory the server may hold (default: half of free memory) */
} GrandServerConfig;

/* Request status codes returned by poll / wait. */
#define GRAND_REQ_PENDING 1
#define GRAND_REQ_DONE 0
#define GRAND_REQ_FAILED (-1)

/* cfg may be NULL (all defaults); zero fields take their defaults. */
GrandServer *grand_server_create(const GrandServerConfig *cfg);

/* Registers a code with the arguments of grand_session_create (kind, matrices,
 * config, phase or LLR file).  Returns the code id (>= 0), -1 for invalid
 * input, or -2 when its footprint does not fit the remaining HBM budget. */
int grand_server_add_code(GrandServer *srv, int kind, const int *H, const int *H2, int n, int m, int n_logicos,
                          char type_of_error, const GrandDecoderConfig *config, const char *file);

/* Bytes the server holds on the device for code `code_id` (-1: all codes). */
long long grand_server_footprint(GrandServer *srv, int code_id);

/* Queues num_shots rows of `input` (errors, m ints per row, when
 * input_is_errors = 1 for the CSS kinds; else syndromes/detectors, n ints) and,
 * for the DEM kinds, `logicals` (n_logicos ints per row).  Results go to
 * `results` (num_shots records) and, if not NULL, the decoded patterns to
 * `decoded` (num_shots x m ints, 0/1).  The buffers must stay valid until the
 * request completes.  Returns a ticket (>= 0), -1 for invalid arguments, or -2
 * when the queue is full. */
long long grand_server_submit(GrandServer *srv, int code_id, const int *input, int input_is_errors,
                              const int *logicals, int num_shots, GrandDecoderResult *results, int *decoded);

/* Status of a request without waiting; latency_ms (optional) gets submit-to-
 * completion time of a finished request. */
int grand_server_poll(GrandServer *srv, long long ticket, double *latency_ms);

/* Waits for a request; returns GRAND_REQ_DONE or GRAND_REQ_FAILED. */
int grand_server_wait(GrandServer *srv, long long ticket, double *latency_ms);

/* Totals since creation: requests completed, shots completed, batches launched. */
void grand_server_counters(GrandServer *srv, long long *requests, long long *shots, long long *batches);

/* Drains outstanding work, stops the worker and frees everything. */
void grand_server_destroy(GrandServer *srv);

#ifdef __cplusplus
}
#endif

#endif /* GRAND_SERVER_H */
