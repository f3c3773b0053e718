/*
 * clspmv_serve.h -- request-serving loop over matrices resident in HBM.
 *
 * No counterpart in the reference: its iteration API (src/eval.cpp:19-72,
 * cpy_vector_from_gpu, src/spmv_cocktail.cpp:1510) serves one matrix and
 * one caller, blocking on every call.  Here many device matrices stay
 * resident and a scheduler thread runs requests from many callers:
 *
 *   - continuous batching: at every pass the scheduler retires finished
 *     passes and, for every matrix with waiting requests that is idle,
 *     coalesces its waiting requests (in arrival order, up to max_batch
 *     requests and kmax right-hand sides) into ONE pass over the matrix:
 *     one multi-vector pass (DIA, CSR), one MFMA SpMM (4x4-block BSPMM), or
 *     one SpMV per right-hand side; passes of different matrices run
 *     concurrently on several HIP streams while new requests keep arriving;
 *   - admission by memory: each matrix's working buffers (x and y for kmax
 *     right-hand sides) are carved out of one device pool of hbm_budget
 *     bytes allocated when the server is created; a matrix whose buffers do
 *     not fit waits until idle matrices' buffers are evicted (least
 *     recently used first).  Copies to host_y go through a pinned staging
 *     area of host_staging bytes, taken per request;
 *   - each (matrix, right-hand-side count) pass is captured once into a HIP
 *     graph against the matrix's buffers (zero y, then the products) and
 *     replayed for every later batch of that size, until the buffers are
 *     evicted;
 *   - results do not depend on the mix: the kernels a matrix runs are fixed
 *     when it is registered, each right-hand side's arithmetic in them does
 *     not depend on the other columns of the pass, and every kernel has a
 *     fixed summation order, so a request gives the same y bit for bit
 *     alone or batched with others.
 *
 * A request computes y_j = A x_j for j < k.  x_j comes from the caller (x,
 * column-major, cols floats per right-hand side, on the host or on the
 * server's device) or, with x = NULL, is generated on the device from
 * (seed, j): uniform in [-1, 1) (clspmv_request_x gives the same values on
 * the host).  Latency is completion minus arrival on the server's clock
 * (seconds since clspmv_server_create).
 *
 * A device matrix registered with a server is driven by the servers only:
 * one pass at a time per device matrix, across every server that holds it
 * (passes write the matrix's own scratch, e.g. the x-panel segment sums),
 * and the caller must not launch on it concurrently through clspmv_spmv.
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
    int    streams;       /* HIP streams passes are spread over (>= 1)        */
    int    max_batch;     /* most requests coalesced into one pass (>= 1)     */
    size_t hbm_budget;    /* device pool for the matrices' working buffers    */
    int    use_graphs;    /* 1: replay captured graphs; 0: plain launches     */
    size_t host_staging;  /* pinned bytes for copies to host_y; 0: 256 MiB    */
} clspmv_server_config;

int clspmv_server_create(const clspmv_server_config* cfg, clspmv_server** out);
/* Finishes the requests in flight, then frees the server (not the
 * matrices). */
int clspmv_server_destroy(clspmv_server* s);

/* Registers a resident device matrix (not owned; it must outlive the
 * server).  A pass carries up to kmax right-hand sides; the matrix's
 * working buffers for kmax of them must fit hbm_budget (else
 * CLSPMV_ERR_TOO_LARGE).  The kernels are fixed here, once per device
 * matrix (every server then runs the same ones; clspmv_dmat_set_panel
 * clears the choice):
 *   - CLSPMV_FMT_BSPMM: the MFMA SpMM, x rounded to the matrix's 16-bit
 *     type;
 *   - DIA and CSR with its own arrays: one multi-vector pass per 16
 *     right-hand sides where that measured faster than separate SpMVs at
 *     min(kmax, 8) right-hand sides, else as below;
 *   - every other matrix: one SpMV per right-hand side with `method`
 *     (clspmv_spmv numbering; CLSPMV_METHOD_AUTO picks by shape).
 * The measurement runs on a stream and buffers of its own (allocated for
 * it and freed after), while no pass of any server uses the matrix.
 * *id receives the matrix id. */
int clspmv_server_add_matrix(clspmv_server* s, const clspmv_dmat* A, int method, int kmax,
                             int* id);

typedef struct {
    int    matrix;
    int    k;                 /* right-hand sides, 1..kmax                    */
    unsigned long long seed;  /* x_j = uniform(seed, j) when x is NULL        */
    const float* x;           /* NULL, or cols * k floats, column-major       */
    int    x_on_device;       /* x is a device pointer (server's GPU)         */
    float* host_y;            /* NULL, or rows * k floats (column-major)      */
    float* device_y;          /* NULL, or rows * k floats on the device       */
    double arrival;           /* server clock; <= 0 means "now"               */
} clspmv_request;

/* Queues a request (thread-safe); *ticket identifies it.  x, host_y and
 * device_y must stay valid until the request is waited for. */
int clspmv_server_submit(clspmv_server* s, const clspmv_request* r, long long* ticket);
/* Blocks until the request is done (host_y filled); *latency_s =
 * completion - arrival. */
int clspmv_server_wait(clspmv_server* s, long long ticket, double* latency_s);
/* Seconds on the server clock. */
double clspmv_server_clock(clspmv_server* s);

/* Counters of the serving loop since the server was created. */
typedef struct {
    long long requests;   /* requests finished                           */
    long long passes;     /* batched passes launched                     */
    long long columns;    /* right-hand sides computed                   */
    long long evictions;  /* working buffers evicted from the pool       */
    size_t    pool_bytes; /* device pool size (hbm_budget)               */
    size_t    pool_used;  /* bytes of resident working buffers           */
} clspmv_server_stats;
int clspmv_server_get_stats(clspmv_server* s, clspmv_server_stats* out);

/* Replays n requests: each is submitted at its arrival time (seconds on the
 * server clock, relative to the start of the replay) and the call returns
 * when all are done, with each request's latency and the wall time.
 * mode 0: the serving loop.  mode 1: plain per-request launches, the
 * baseline -- the calling thread runs each request alone in arrival order
 * on the matrix's buffers: stage x, zero y, one clspmv_spmv launch per
 * right-hand side with the matrix's single-vector method (clspmv_spmm once
 * for BSPMM), copy y out if asked, synchronize. */
int clspmv_server_replay(clspmv_server* s, int n, const clspmv_request* reqs, int mode,
                         double* latencies, double* wall_s);

/* x_j of a request on the host: n floats uniform in [-1, 1) from (seed, j),
 * equal to what the device generates. */
void clspmv_request_x(unsigned long long seed, int j, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif /* CLSPMV_SERVE_H */
