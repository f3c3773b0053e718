// Request-serving loop over resident device matrices (clspmv_serve.h).
//
// One scheduler thread owns every HIP call of the serving path.  Each pass
// it (1) retires finished requests (an event per request), (2) admits the
// waiting requests whose matrix is idle and whose x / y buffers fit the
// in-flight HBM budget -- skipping busy matrices rather than blocking on
// them -- and (3) launches each admitted request on its matrix's stream:
// generate x from the seed, then replay the captured (matrix, k) graph
// (zero y, then one multi-vector pass, one SpMM, or k SpMVs).  Callers only
// touch the queue.  Whether a k runs as one multi-vector pass or as k SpMVs
// is measured once per device matrix when it is registered.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "clspmv_serve.h"
#include "kernels.hpp"
#include "rng.hpp"
#include "runtime.hpp"

using namespace clspmv;

namespace {

__global__ void fill_x_kernel(unsigned long long seed, int k, int cols, int ldx, float* X) {
    const long long n = (long long)k * cols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(t / cols), i = (int)(t % cols);
        X[(size_t)j * ldx + i] = uniform_at(seed, j, i);
    }
}

// Row-major x for the multi-vector kernels: X[i*k + j] = x_j[i].
__global__ void fill_x_rows_kernel(unsigned long long seed, int k, int cols, float* X) {
    const long long n = (long long)k * cols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t / k), j = (int)(t % k);
        X[t] = uniform_at(seed, j, i);
    }
}

__global__ void fill_x16_kernel(unsigned long long seed, int k, int cols, int ldx, int dtype,
                                unsigned short* X) {
    const long long n = (long long)k * cols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(t / cols), i = (int)(t % cols);
        const float v = uniform_at(seed, j, i);
        X[(size_t)j * ldx + i] = dtype == CLSPMV_DT_BF16 ? to_bf16(v) : to_f16(v);
    }
}

__global__ void zero_kernel(long long n, float* y) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        y[t] = 0.0f;
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

struct SMatrix {
    const clspmv_dmat* A = nullptr;
    int method = 0, rows = 0, cols = 0, kmax = 0, ldx = 0, ldy = 0, stream = 0;
    bool lowp = false;
    bool multi = false;       // the matrix has a multi-vector kernel
    std::vector<char> pick;   // per k: run one multi-vector pass (X row-major)
    void* X = nullptr;
    float* Y = nullptr;
    float* pinned = nullptr;  // rows * kmax staging for host copies
    std::vector<hipGraph_t> graph;
    std::vector<hipGraphExec_t> exec;
    bool busy = false;
};

struct SRequest {
    clspmv_request r{};
    double arrival = 0.0, done = -1.0;
    size_t bytes = 0;
    int state = 0;  // 0 waiting, 1 running, 2 done, < 0 failed (status)
    hipEvent_t ev = nullptr;
};

}  // namespace

struct clspmv_server {
    clspmv_server_config cfg{};
    std::vector<hipStream_t> streams;
    std::vector<SMatrix> mats;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<long long> pending;  // arrival order
    std::unordered_map<long long, SRequest> reqs;
    std::vector<long long> running;
    std::vector<hipEvent_t> free_events;
    size_t inflight = 0;
    long long next_ticket = 1;
    bool stop = false;
    std::thread sched;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

size_t request_bytes(const SMatrix& M, int k) {
    return (size_t)k * ((size_t)M.ldx * (M.lowp ? 2 : 4) + (size_t)M.ldy * 4);
}

bool use_multi(const SMatrix& M, int k) { return M.multi && k > 1 && M.pick[k]; }

// Zero y, then the k products (one multi-vector pass when `multi`, else one
// SpMV per right-hand side).  Runs on `s` (captured or not).
int enqueue_pass(hipStream_t s, const SMatrix& M, int k, bool multi) {
    const long long ny = (long long)M.ldy * k;
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(ny)), dim3(256), 0, s, ny, M.Y);
    CLSPMV_HIP(hipGetLastError());
    if (M.lowp) {
        CLSPMV_HIP(launch::bspmm4(s, M.A->dtype == CLSPMV_DT_BF16, M.A->long_brows, M.A->brows, M.rows, M.cols, k,
                                  M.A->rp, M.A->col, M.A->lpval,
                                  static_cast<const unsigned short*>(M.X), M.ldx, M.Y, M.ldy));
        return CLSPMV_OK;
    }
    if (multi) return multi_launch(s, M.A, k, static_cast<const float*>(M.X), M.Y, M.ldy);
    for (int j = 0; j < k; ++j) {
        int st = dmat_launch(s, M.A, M.method, static_cast<const float*>(M.X) + (size_t)j * M.ldx,
                             M.Y + (size_t)j * M.ldy);
        if (st) return st;
    }
    return CLSPMV_OK;
}

int enqueue_x(hipStream_t s, const SMatrix& M, int k, unsigned long long seed, bool multi) {
    const long long n = (long long)k * M.cols;
    if (M.lowp)
        hipLaunchKernelGGL(fill_x16_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, k, M.cols, M.ldx,
                           M.A->dtype, static_cast<unsigned short*>(M.X));
    else if (multi)
        hipLaunchKernelGGL(fill_x_rows_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, k, M.cols,
                           static_cast<float*>(M.X));
    else
        hipLaunchKernelGGL(fill_x_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, k, M.cols, M.ldx,
                           static_cast<float*>(M.X));
    CLSPMV_HIP(hipGetLastError());
    return CLSPMV_OK;
}

int graph_for(clspmv_server* S, SMatrix& M, int k, hipGraphExec_t* out) {
    if (!M.exec[k]) {
        hipStream_t s = S->streams[M.stream];
        CLSPMV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int st = enqueue_pass(s, M, k, use_multi(M, k));
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "capture failed: %s", hipGetErrorString(e));
        M.graph[k] = g;
        CLSPMV_HIP(hipGraphInstantiate(&M.exec[k], g, nullptr, nullptr, 0));
    }
    *out = M.exec[k];
    return CLSPMV_OK;
}

// Mean time of `reps` runs of `fn` on `s` after one untimed run, in ms.
template <class F>
int time_pass(hipStream_t s, hipEvent_t e0, hipEvent_t e1, F fn, float* ms) {
    const int reps = 3;
    int st = fn();
    if (st) return st;
    CLSPMV_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps && !st; ++i) st = fn();
    if (st) return st;
    CLSPMV_HIP(hipEventRecord(e1, s));
    CLSPMV_HIP(hipEventSynchronize(e1));
    CLSPMV_HIP(hipEventElapsedTime(ms, e0, e1));
    *ms /= reps;
    return CLSPMV_OK;
}

// Fills M.pick: for each k in 2..kmax, one multi-vector pass where it
// measured faster than k SpMVs of the chosen method.  Measured on `s` (a
// stream of its own, at registration, under the server lock) once per
// device matrix -- the choice is kept on it -- so servers sharing a matrix
// run the same kernels and give the same results.
int calibrate(SMatrix& M, hipStream_t s) {
    M.pick.assign(M.kmax + 1, 0);
    if (!M.multi) return CLSPMV_OK;
    std::vector<signed char>& cache = M.A->multi_pick;
    if ((int)cache.size() < M.kmax + 1) cache.resize(M.kmax + 1, -1);
    bool need = false;
    for (int k = 2; k <= M.kmax; ++k) need = need || cache[k] < 0;
    if (need) {
        const long long n = (long long)M.ldx * M.kmax;
        hipLaunchKernelGGL(fill_x_kernel, dim3(grid_for(n)), dim3(256), 0, s, 12345ull, M.kmax, M.cols,
                           M.ldx, static_cast<float*>(M.X));
        CLSPMV_HIP(hipGetLastError());
        hipEvent_t e0 = nullptr, e1 = nullptr;
        CLSPMV_HIP(hipEventCreate(&e0));
        int st = CLSPMV_OK;
        if (hipEventCreate(&e1) != hipSuccess) st = fail(CLSPMV_ERR_HIP, "hipEventCreate failed");
        float t1 = 0.0f;
        if (!st) st = time_pass(s, e0, e1, [&] { return enqueue_pass(s, M, 1, false); }, &t1);
        for (int k = 2; k <= M.kmax && !st; ++k) {
            if (cache[k] >= 0) continue;
            float tm = 0.0f;
            st = time_pass(s, e0, e1, [&] { return enqueue_pass(s, M, k, true); }, &tm);
            if (!st) cache[k] = tm < k * t1 ? 1 : 0;
        }
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) return st;
    }
    for (int k = 2; k <= M.kmax; ++k) M.pick[k] = cache[k] == 1;
    return CLSPMV_OK;
}

int copy_out(hipStream_t s, const SMatrix& M, int k, float* dst) {
    CLSPMV_HIP(hipMemcpy2DAsync(dst, sizeof(float) * M.rows, M.Y, sizeof(float) * M.ldy,
                                sizeof(float) * M.rows, k, hipMemcpyDeviceToHost, s));
    return CLSPMV_OK;
}

// Launches an admitted request (scheduler thread, lock held).
int launch_request(clspmv_server* S, SRequest& q) {
    SMatrix& M = S->mats[q.r.matrix];
    hipStream_t s = S->streams[M.stream];
    int st = enqueue_x(s, M, q.r.k, q.r.seed, use_multi(M, q.r.k));
    if (st) return st;
    if (S->cfg.use_graphs) {
        hipGraphExec_t ex = nullptr;
        if ((st = graph_for(S, M, q.r.k, &ex))) return st;
        CLSPMV_HIP(hipGraphLaunch(ex, s));
    } else if ((st = enqueue_pass(s, M, q.r.k, use_multi(M, q.r.k)))) {
        return st;
    }
    if (q.r.host_y && (st = copy_out(s, M, q.r.k, M.pinned))) return st;
    if (S->free_events.empty()) {
        hipEvent_t e = nullptr;
        CLSPMV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        S->free_events.push_back(e);
    }
    q.ev = S->free_events.back();
    S->free_events.pop_back();
    CLSPMV_HIP(hipEventRecord(q.ev, s));
    return CLSPMV_OK;
}

void finish_request(clspmv_server* S, SRequest& q, int status, double t) {
    SMatrix& M = S->mats[q.r.matrix];
    if (!status && q.r.host_y)
        std::copy(M.pinned, M.pinned + (size_t)M.rows * q.r.k, q.r.host_y);
    M.busy = false;
    S->inflight -= q.bytes;
    if (q.ev) S->free_events.push_back(q.ev);
    q.ev = nullptr;
    q.done = t;
    q.state = status ? status : 2;
}

void scheduler(clspmv_server* S) {
    (void)hipSetDevice(S->cfg.device_ordinal);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        if (S->stop && S->pending.empty() && S->running.empty()) break;
        bool progress = false;
        const double t = S->now();
        // Retire.
        for (size_t i = 0; i < S->running.size();) {
            SRequest& q = S->reqs[S->running[i]];
            const hipError_t e = hipEventQuery(q.ev);
            if (e == hipErrorNotReady) {
                ++i;
                continue;
            }
            finish_request(S, q, e == hipSuccess ? 0 : CLSPMV_ERR_HIP, S->now());
            S->running[i] = S->running.back();
            S->running.pop_back();
            progress = true;
        }
        // Admit.
        int admitted = 0;
        for (auto it = S->pending.begin(); it != S->pending.end() && admitted < S->cfg.max_batch;) {
            SRequest& q = S->reqs[*it];
            if (q.arrival > t) break;  // arrival order: the rest are later
            SMatrix& M = S->mats[q.r.matrix];
            if (M.busy || S->inflight + q.bytes > S->cfg.hbm_budget) {
                ++it;
                continue;
            }
            M.busy = true;
            S->inflight += q.bytes;
            q.state = 1;
            const int st = launch_request(S, q);
            if (st) {
                finish_request(S, q, st, S->now());
            } else {
                S->running.push_back(*it);
            }
            it = S->pending.erase(it);
            ++admitted;
            progress = true;
        }
        if (progress) S->cv.notify_all();
        if (!progress) {
            // Nothing to do right now: wait briefly for a submission (the
            // device keeps running; completions are polled).
            if (S->running.empty() && S->pending.empty())
                S->cv.wait_for(lk, std::chrono::milliseconds(1));
            else {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }
}

// The baseline: one request at a time, one SpMV launch per right-hand side
// (as a caller of clspmv_spmv would issue them), synchronised at the end.
int run_plain(clspmv_server* S, const clspmv_request& r) {
    SMatrix& M = S->mats[r.matrix];
    hipStream_t s = S->streams[0];
    int st = enqueue_x(s, M, r.k, r.seed, false);
    if (!st) st = enqueue_pass(s, M, r.k, false);
    if (!st && r.host_y) st = copy_out(s, M, r.k, M.pinned);
    if (!st) CLSPMV_HIP(hipStreamSynchronize(s));
    if (!st && r.host_y) std::copy(M.pinned, M.pinned + (size_t)M.rows * r.k, r.host_y);
    return st;
}

void sleep_until(clspmv_server* S, double t) {
    for (;;) {
        const double d = t - S->now();
        if (d <= 0) return;
        if (d > 200e-6)
            std::this_thread::sleep_for(std::chrono::duration<double>(d - 100e-6));
        else
            std::this_thread::yield();
    }
}

}  // namespace

extern "C" {

int clspmv_server_create(const clspmv_server_config* cfg, clspmv_server** out) {
    if (!cfg || !out) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (cfg->streams < 1 || cfg->max_batch < 1)
        return fail(CLSPMV_ERR_INVALID, "streams and max_batch must be >= 1");
    int n = 0;
    CLSPMV_HIP(hipGetDeviceCount(&n));
    if (cfg->device_ordinal < 0 || cfg->device_ordinal >= n)
        return fail(CLSPMV_ERR_INVALID, "device %d not present", cfg->device_ordinal);
    CLSPMV_HIP(hipSetDevice(cfg->device_ordinal));
    clspmv_server* S = new clspmv_server();
    S->cfg = *cfg;
    S->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < cfg->streams; ++i) {
        hipStream_t s = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (hipStream_t t : S->streams) (void)hipStreamDestroy(t);
            delete S;
            return fail(CLSPMV_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
        }
        S->streams.push_back(s);
    }
    S->sched = std::thread(scheduler, S);
    *out = S;
    return CLSPMV_OK;
}

int clspmv_server_destroy(clspmv_server* S) {
    if (!S) return CLSPMV_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    S->sched.join();
    (void)hipSetDevice(S->cfg.device_ordinal);
    for (SMatrix& M : S->mats) {
        for (auto e : M.exec)
            if (e) (void)hipGraphExecDestroy(e);
        for (auto g : M.graph)
            if (g) (void)hipGraphDestroy(g);
        if (M.X) (void)hipFree(M.X);
        if (M.Y) (void)hipFree(M.Y);
        if (M.pinned) (void)hipHostFree(M.pinned);
    }
    for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
    for (auto& kv : S->reqs)
        if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
    for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
    delete S;
    return CLSPMV_OK;
}

int clspmv_server_add_matrix(clspmv_server* S, const clspmv_dmat* A, int method, int kmax, int* id) {
    if (!S || !A || !id) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (kmax < 1) return fail(CLSPMV_ERR_INVALID, "kmax must be >= 1");
    SMatrix M;
    M.A = A;
    M.lowp = A->info.format == CLSPMV_FMT_BSPMM;
    M.multi = !M.lowp && has_multi(A);
    M.method = M.lowp ? 0 : (method == CLSPMV_METHOD_AUTO ? auto_method(A) : method);
    if (!M.lowp && (M.method < 0 || M.method >= clspmv_method_count(A->info.format)))
        return fail(CLSPMV_ERR_INVALID, "method %d not available for format %d", method, A->info.format);
    M.rows = A->info.rows;
    M.cols = A->info.cols;
    M.kmax = kmax;
    M.ldx = round_up(std::max(M.cols, 1), 4);
    M.ldy = round_up(std::max(M.rows, 1), 4);
    M.graph.assign(kmax + 1, nullptr);
    M.exec.assign(kmax + 1, nullptr);
    std::lock_guard<std::mutex> lk(S->mu);
    if (request_bytes(M, 1) > S->cfg.hbm_budget)
        return fail(CLSPMV_ERR_TOO_LARGE, "one request of this matrix exceeds the HBM budget");
    CLSPMV_HIP(hipSetDevice(S->cfg.device_ordinal));
    hipError_t e = hipMalloc(&M.X, (size_t)M.ldx * kmax * (M.lowp ? 2 : 4));
    if (e == hipSuccess) e = hipMalloc(&M.Y, sizeof(float) * (size_t)M.ldy * kmax);
    if (e == hipSuccess) e = hipHostMalloc(&M.pinned, sizeof(float) * (size_t)M.rows * kmax + 16);
    if (e != hipSuccess) {
        if (M.X) (void)hipFree(M.X);
        if (M.Y) (void)hipFree(M.Y);
        return fail(CLSPMV_ERR_ALLOC, "request buffers: %s", hipGetErrorString(e));
    }
    M.stream = (int)(S->mats.size() % S->streams.size());
    hipStream_t cs = nullptr;
    int st = hipStreamCreate(&cs) == hipSuccess ? CLSPMV_OK : fail(CLSPMV_ERR_HIP, "hipStreamCreate failed");
    if (!st) st = calibrate(M, cs);
    if (cs) (void)hipStreamDestroy(cs);
    if (st) {
        (void)hipFree(M.X);
        (void)hipFree(M.Y);
        (void)hipHostFree(M.pinned);
        return st;
    }
    *id = (int)S->mats.size();
    S->mats.push_back(std::move(M));
    return CLSPMV_OK;
}

double clspmv_server_clock(clspmv_server* S) { return S ? S->now() : 0.0; }

int clspmv_server_submit(clspmv_server* S, const clspmv_request* r, long long* ticket) {
    if (!S || !r || !ticket) return fail(CLSPMV_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (r->matrix < 0 || r->matrix >= (int)S->mats.size())
        return fail(CLSPMV_ERR_INVALID, "matrix %d not registered", r->matrix);
    const SMatrix& M = S->mats[r->matrix];
    if (r->k < 1 || r->k > M.kmax) return fail(CLSPMV_ERR_INVALID, "k = %d outside 1..%d", r->k, M.kmax);
    SRequest q;
    q.r = *r;
    q.arrival = r->arrival > 0 ? r->arrival : S->now();
    q.bytes = request_bytes(M, r->k);
    if (q.bytes > S->cfg.hbm_budget) return fail(CLSPMV_ERR_TOO_LARGE, "request exceeds the HBM budget");
    const long long t = S->next_ticket++;
    S->reqs.emplace(t, q);
    // Keep the queue in arrival order (submissions usually arrive in order).
    auto pos = S->pending.end();
    while (pos != S->pending.begin() && S->reqs[*(pos - 1)].arrival > q.arrival) --pos;
    S->pending.insert(pos, t);
    *ticket = t;
    S->cv.notify_all();
    return CLSPMV_OK;
}

int clspmv_server_wait(clspmv_server* S, long long ticket, double* latency_s) {
    if (!S) return fail(CLSPMV_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return fail(CLSPMV_ERR_INVALID, "unknown ticket %lld", ticket);
    S->cv.wait(lk, [&] { return S->reqs[ticket].state >= 2 || S->reqs[ticket].state < 0; });
    SRequest q = S->reqs[ticket];
    S->reqs.erase(ticket);
    if (latency_s) *latency_s = q.done - q.arrival;
    if (q.state < 0) return fail(q.state, "request %lld failed", ticket);
    return CLSPMV_OK;
}

int clspmv_server_replay(clspmv_server* S, int n, const clspmv_request* reqs, int mode,
                         double* latencies, double* wall_s) {
    if (!S || (n > 0 && !reqs)) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (mode != 0 && mode != 1) return fail(CLSPMV_ERR_INVALID, "mode %d", mode);
    const double start = S->now();
    double last = start;
    if (mode == 1) {
        CLSPMV_HIP(hipSetDevice(S->cfg.device_ordinal));
        for (int i = 0; i < n; ++i) {
            const double arr = start + std::max(0.0, reqs[i].arrival);
            sleep_until(S, arr);
            int st = run_plain(S, reqs[i]);
            if (st) return st;
            last = S->now();
            if (latencies) latencies[i] = last - arr;
        }
    } else {
        std::vector<long long> tickets(n);
        for (int i = 0; i < n; ++i) {
            clspmv_request r = reqs[i];
            r.arrival = start + std::max(0.0, reqs[i].arrival);
            if (r.arrival < 1e-9) r.arrival = 1e-9;
            sleep_until(S, r.arrival);
            int st = clspmv_server_submit(S, &r, &tickets[i]);
            if (st) return st;
        }
        for (int i = 0; i < n; ++i) {
            double lat = 0.0;
            int st = clspmv_server_wait(S, tickets[i], &lat);
            if (st) return st;
            if (latencies) latencies[i] = lat;
        }
        last = S->now();
    }
    if (wall_s) *wall_s = last - start;
    return CLSPMV_OK;
}

void clspmv_request_x(unsigned long long seed, int j, int n, float* out) {
    for (int i = 0; i < n; ++i) out[i] = uniform_at(seed, j, i);
}

}  // extern "C"
