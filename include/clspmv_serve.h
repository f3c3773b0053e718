/*
 * clspmv_serve.h -- request-serving loop over matrices resident in HBM.
 *
 * No counterpart in the reference: its iteration API (src/eval.cpp:19-72,
 * cpy_vector_from_gpu, src/spmv_cocktail.cpp:1510) serves one matrix and
 * one caller, blocking on every call.  Here many device matrices stay
 * resident and a scheduler thread runs requests from many callers:
 *
 *   - continuous batching: at every pass the scheduler admits every waiting
 *     request whose matrix is idle and whose buffers fit the HBM budget,
 *     while earlier requests are still running; requests run concurrently
 *     on several HIP streams;
 *   - each (matrix, k) pass is captured once into a HIP graph (zero y, then
 *     one multi-vector pass, one SpMM, or k SpMVs) and replayed for every
 *     later request;
 *   - results do not depend on the mix: every kernel has a fixed summation
 *     order and a request only touches its matrix's own buffers.
 *
 * A request computes y_j = A x_j for j < k, with x_j generated on the
 * device from (seed, j): uniform in [-1, 1) (clspmv_request_x gives the same
 * values on the host).  Latency is completion minus arrival on the server's
 * clock (seconds since clspmv_server_create).
 */
#ifndef CLSPMV_SERVE_H
#define CLSPMV_SERVE_H

#include <stddef.h>
#include "clspmv_device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clspmv_server clspmv_server;

typedef struct {
    int    device_ordinal;
    int    streams;      /* HIP streams the requests are spread over (>= 1)  */
    int    max_batch;    /* most requests admitted in one scheduler pass     */
    size_t hbm_budget;   /* bytes of x / y buffers of requests in flight     */
    int    use_graphs;   /* 1: replay captured graphs; 0: plain launches     */
} clspmv_server_config;

int clspmv_server_create(const clspmv_server_config* cfg, clspmv_server** out);
/* Finishes the requests in flight, then frees the server (not the
 * matrices). */
int clspmv_server_destroy(clspmv_server* s);

/* Registers a resident device matrix (not owned; it must outlive the
 * server) and allocates its request buffers for up to kmax right-hand
 * sides.  FP32 matrices run `method` (clspmv_spmv numbering) once per
 * right-hand side; CLSPMV_FMT_BSPMM matrices run clspmv_spmm with x rounded
 * to the matrix's 16-bit type.  DIA and CSR matrices (CSR with its plain
 * encoding) run a request with k > 1 as one multi-vector pass
 * (clspmv_spmv_multi) instead wherever that was faster: the first
 * registration of a device matrix times both for each k up to kmax, and the
 * choice stays with the matrix (every server then runs the same kernels;
 * clspmv_dmat_set_panel clears it).  *id receives the matrix id. */
int clspmv_server_add_matrix(clspmv_server* s, const clspmv_dmat* A, int method, int kmax,
                             int* id);

typedef struct {
    int    matrix;
    int    k;                 /* right-hand sides, 1..kmax                  */
    unsigned long long seed;  /* x_j = uniform(seed, j)                     */
    float* host_y;            /* NULL, or rows * k floats (column-major)    */
    double arrival;           /* server clock; <= 0 means "now"             */
} clspmv_request;

/* Queues a request (thread-safe); *ticket identifies it. */
int clspmv_server_submit(clspmv_server* s, const clspmv_request* r, long long* ticket);
/* Blocks until the request is done; *latency_s = completion - arrival. */
int clspmv_server_wait(clspmv_server* s, long long ticket, double* latency_s);
/* Seconds on the server clock. */
double clspmv_server_clock(clspmv_server* s);

/* Replays n requests: each is submitted at its arrival time (seconds on the
 * server clock, relative to the start of the replay) and the call returns
 * when all are done, with each request's latency and the wall time.
 * mode 0: the serving loop.  mode 1: plain per-request launches, the
 * baseline -- the calling thread runs each request alone in arrival order:
 * generate x, zero y, one SpMV (or SpMM) launch per right-hand side as a
 * caller of clspmv_spmv would, copy y out if asked, synchronize. */
int clspmv_server_replay(clspmv_server* s, int n, const clspmv_request* reqs, int mode,
                         double* latencies, double* wall_s);

/* x_j of a request on the host: n floats uniform in [-1, 1) from (seed, j),
 * equal to what the device generates. */
void clspmv_request_x(unsigned long long seed, int j, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif /* CLSPMV_SERVE_H */
