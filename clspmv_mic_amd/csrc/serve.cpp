// Request-serving loop over resident device matrices (clspmv_serve.h).
//
// One scheduler thread owns every HIP call of the serving path.  Each pass
// it (1) retires finished batches (an event per batch), (2) for every idle
// matrix with waiting requests, oldest first, makes sure the matrix's
// working buffers are resident in the server's device pool (evicting idle
// matrices' buffers, least recently used first, when the pool is full),
// takes the matrix's claim (one pass per device matrix across servers),
// coalesces its waiting requests into one batch, and (3) launches the batch
// on the matrix's stream: stage each request's x into its columns of the
// matrix's X, replay the captured (matrix, K) graph (one MFMA SpMM writing
// y, or zero y, then one multi-vector pass per 16 columns or one SpMV per
// column),
// copy each request's y out.  Callers only touch the queue.
//
// Bit-for-bit independence from the mix: the kernels of a matrix are fixed
// at registration, and in each of them a column's arithmetic does not
// depend on the other columns: DIA's multi-vector kernel sums each column
// alone in diagonal order; CSR's runs at the fixed 16-column width
// (launch::csr_multi with kb = 16); the SpMM's MFMA output columns depend
// only on their own X column, in block order; per-column SpMVs are
// independent launches.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "clspmv_serve.h"
#include "kernels.hpp"
#include "rng.hpp"
#include "runtime.hpp"

using namespace clspmv;

namespace {

constexpr size_t kNone = ~(size_t)0;
constexpr int kMultiCols = launch::kMultiMax;  // columns per multi-vector launch
constexpr size_t kPoolAlign = 256;

// How a matrix computes a batch of K right-hand sides.
enum Path { kSingle = 0, kMulti = 1, kSpmm = 2 };

// Layouts the staging kernel writes x into.
enum Layout {
    kColF32 = 0,    // X[g * ldx + i]: column-major FP32 (per-column SpMV)
    kChunkF32 = 1,  // chunk c = g / 16 of K columns: X[c*16*ldx + i*kc + g%16] (multi-vector)
    kTile16 = 2,    // 4-row tiles of 16-bit values, stride K: X[((i/4)*K + g)*4 + i%4] (SpMM)
    kCol16 = 3      // X[g * ldx + i]: column-major 16-bit (plain clspmv_spmm baseline)
};

// x of one request (k columns from `src`, column-major with leading
// dimension src_ld, or generated from the seed when src is null) into
// columns [g0, g0 + k) of a batch of K columns in `layout`.  Rows from cols
// up to the next multiple of 4 are zero-filled for the 16-bit layouts (the
// SpMM reads whole 4-row pieces).
__global__ void stage_x_kernel(unsigned long long seed, const float* __restrict__ src, int src_ld, int k,
                               int cols, int layout, int dtype, int K, int g0, int ldx, void* __restrict__ X) {
    const int cpad = layout >= kTile16 ? (cols + 3) / 4 * 4 : cols;
    const long long n = (long long)k * cpad;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(t / cpad), i = (int)(t % cpad);
        float v = 0.0f;
        if (i < cols) v = src ? src[(size_t)j * src_ld + i] : uniform_at(seed, j, i);
        const int g = g0 + j;
        switch (layout) {
            case kColF32:
                static_cast<float*>(X)[(size_t)g * ldx + i] = v;
                break;
            case kChunkF32: {
                const int c = g / kMultiCols, kc = min(kMultiCols, K - c * kMultiCols);
                static_cast<float*>(X)[(size_t)c * kMultiCols * ldx + (size_t)i * kc + g % kMultiCols] = v;
                break;
            }
            case kTile16:
                static_cast<unsigned short*>(X)[((size_t)(i >> 2) * K + g) * 4 + (i & 3)] =
                    dtype == CLSPMV_DT_BF16 ? to_bf16(v) : to_f16(v);
                break;
            default:
                static_cast<unsigned short*>(X)[(size_t)g * ldx + i] =
                    dtype == CLSPMV_DT_BF16 ? to_bf16(v) : to_f16(v);
                break;
        }
    }
}

__global__ void zero_kernel(long long n, float* y) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        y[t] = 0.0f;
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

// ---- claims: one pass at a time per device matrix, across servers --------
std::mutex g_claim_mu;
std::unordered_set<const clspmv_dmat*> g_claimed;

bool try_claim(const clspmv_dmat* A) {
    std::lock_guard<std::mutex> lk(g_claim_mu);
    return g_claimed.insert(A).second;
}

void release_claim(const clspmv_dmat* A) {
    std::lock_guard<std::mutex> lk(g_claim_mu);
    g_claimed.erase(A);
}

// ---- first-fit allocator over one buffer (offsets, 256-byte aligned) ----
struct Arena {
    size_t size = 0, used = 0;
    std::map<size_t, size_t> free_;  // offset -> length

    void init(size_t n) {
        size = n / kPoolAlign * kPoolAlign;
        free_.clear();
        if (size) free_[0] = size;
    }
    size_t alloc(size_t n) {
        n = (n + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used += n;
            return off;
        }
        return kNone;
    }
    void release(size_t off, size_t n) {
        n = (n + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct SMatrix {
    const clspmv_dmat* A = nullptr;
    int method = 0, rows = 0, cols = 0, kmax = 0, ldx = 0, ldy = 0, stream = 0;
    int path = kSingle;
    size_t x_bytes = 0, y_bytes = 0, s_bytes = 0;  // working buffers
    // Working buffers in the pool (slot == kNone: not resident).
    size_t slot = kNone;
    void* X = nullptr;
    float* Y = nullptr;
    float* S = nullptr;  // host x lands here before conversion (multi / SpMM paths)
    std::vector<hipGraph_t> graph;  // per K, captured against the buffers
    std::vector<hipGraphExec_t> exec;
    std::deque<long long> waiting;  // tickets, arrival order
    bool running = false;           // a batch (or the plain baseline) is using the buffers
    long long last_use = 0;

    size_t slot_bytes() const { return x_bytes + y_bytes + s_bytes; }
};

struct SRequest {
    clspmv_request r{};
    double arrival = 0.0, done = -1.0;
    int state = 0;  // 0 waiting, 1 running, 2 done, < 0 failed (status)
    int col = 0;    // first column in its batch
    size_t pin = kNone;  // pinned staging of host_y
};

struct SBatch {
    int matrix = 0;
    std::vector<long long> tickets;
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
    std::unordered_map<long long, SRequest> reqs;
    std::vector<SBatch> running;
    std::vector<hipEvent_t> free_events;
    void* pool = nullptr;
    Arena pool_arena;
    float* pinned = nullptr;
    Arena pin_arena;
    long long next_ticket = 1, clock_use = 0;
    long long n_requests = 0, n_passes = 0, n_columns = 0, n_evictions = 0;
    int pending = 0;  // requests waiting in some matrix's queue
    bool stop = false;
    std::thread sched;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void drop_graphs(SMatrix& M) {
    for (auto& e : M.exec)
        if (e) (void)hipGraphExecDestroy(e), e = nullptr;
    for (auto& g : M.graph)
        if (g) (void)hipGraphDestroy(g), g = nullptr;
}

void evict(clspmv_server* S, SMatrix& M) {
    drop_graphs(M);
    S->pool_arena.release(M.slot, M.slot_bytes());
    M.slot = kNone;
    M.X = nullptr;
    M.Y = nullptr;
    M.S = nullptr;
    ++S->n_evictions;
}

// Makes M's working buffers resident, evicting idle matrices' buffers (least
// recently used first) when the pool is full.  False: not possible now.
bool ensure_slot(clspmv_server* S, SMatrix& M) {
    if (M.slot != kNone) return true;
    size_t off = S->pool_arena.alloc(M.slot_bytes());
    while (off == kNone) {
        SMatrix* lru = nullptr;
        for (SMatrix& o : S->mats)
            if (&o != &M && o.slot != kNone && !o.running && (!lru || o.last_use < lru->last_use)) lru = &o;
        if (!lru) return false;
        evict(S, *lru);
        off = S->pool_arena.alloc(M.slot_bytes());
    }
    char* base = static_cast<char*>(S->pool) + off;
    M.slot = off;
    M.X = base;
    M.Y = reinterpret_cast<float*>(base + M.x_bytes);
    M.S = M.s_bytes ? reinterpret_cast<float*>(base + M.x_bytes + M.y_bytes) : nullptr;
    return true;
}

// The products of K right-hand sides into y: the SpMM writes Y = A X
// (the bits a zeroed Y plus A X gives); the other paths add into a zeroed
// y.  Runs on `s` (captured or not).
int enqueue_pass(hipStream_t s, const SMatrix& M, int K) {
    if (M.path == kSpmm) {
        const clspmv_dmat* A = M.A;
        CLSPMV_HIP(launch::bspmm4(s, A->dtype == CLSPMV_DT_BF16, A->long_brows, A->brows, M.rows, M.cols, K, A->rp,
                                  A->col, A->lpval, static_cast<const unsigned short*>(M.X), K, M.Y, M.ldy,
                                  /*tiled=*/true, /*accumulate=*/false,
                                  launch::SpmmWindows{A->wtile, A->widx, A->win_max}));
        return CLSPMV_OK;
    }
    const long long ny = (long long)M.ldy * K;
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(ny)), dim3(256), 0, s, ny, M.Y);
    CLSPMV_HIP(hipGetLastError());
    if (M.path == kMulti) {
        const int kb = M.A->info.format == CLSPMV_FMT_CSR ? kMultiCols : 0;
        for (int c = 0; c * kMultiCols < K; ++c) {
            const int kc = std::min(kMultiCols, K - c * kMultiCols);
            int st = multi_launch(s, M.A, kc, static_cast<const float*>(M.X) + (size_t)c * kMultiCols * M.ldx,
                                  M.Y + (size_t)c * kMultiCols * M.ldy, M.ldy, kb);
            if (st) return st;
        }
        return CLSPMV_OK;
    }
    for (int j = 0; j < K; ++j) {
        int st = dmat_launch(s, M.A, M.method, static_cast<const float*>(M.X) + (size_t)j * M.ldx,
                             M.Y + (size_t)j * M.ldy);
        if (st) return st;
    }
    return CLSPMV_OK;
}

int layout_of(const SMatrix& M, bool plain) {
    if (M.path == kSpmm) return plain ? kCol16 : kTile16;
    if (M.path == kMulti && !plain) return kChunkF32;
    return kColF32;
}

// x of one request into columns [g0, g0 + k) of a batch of K.
int stage_x(hipStream_t s, const SMatrix& M, const clspmv_request& r, int K, int g0, bool plain) {
    const int layout = layout_of(M, plain);
    const float* src = nullptr;
    if (r.x) {
        if (r.x_on_device) {
            src = r.x;
        } else if (layout == kColF32) {  // straight into X
            CLSPMV_HIP(hipMemcpy2DAsync(static_cast<float*>(M.X) + (size_t)g0 * M.ldx, sizeof(float) * M.ldx, r.x,
                                        sizeof(float) * M.cols, sizeof(float) * M.cols, r.k,
                                        hipMemcpyHostToDevice, s));
            return CLSPMV_OK;
        } else {  // through the staging columns, then converted
            CLSPMV_HIP(hipMemcpy2DAsync(M.S + (size_t)g0 * M.ldx, sizeof(float) * M.ldx, r.x, sizeof(float) * M.cols,
                                        sizeof(float) * M.cols, r.k, hipMemcpyHostToDevice, s));
            src = M.S + (size_t)g0 * M.ldx;
            hipLaunchKernelGGL(stage_x_kernel, dim3(grid_for((long long)r.k * (M.cols + 3))), dim3(256), 0, s,
                               r.seed, src, M.ldx, r.k, M.cols, layout, M.A->dtype, K, g0, M.ldx, M.X);
            CLSPMV_HIP(hipGetLastError());
            return CLSPMV_OK;
        }
    }
    hipLaunchKernelGGL(stage_x_kernel, dim3(grid_for((long long)r.k * (M.cols + 3))), dim3(256), 0, s, r.seed, src,
                       M.cols, r.k, M.cols, layout, M.A->dtype, K, g0, M.ldx, M.X);
    CLSPMV_HIP(hipGetLastError());
    return CLSPMV_OK;
}

// y of one request (columns [g0, g0 + k) of Y) to the caller.
int copy_y(hipStream_t s, const SMatrix& M, const clspmv_request& r, int g0, float* pinned) {
    const float* y = M.Y + (size_t)g0 * M.ldy;
    const size_t row_bytes = sizeof(float) * M.rows;
    if (r.device_y)
        CLSPMV_HIP(hipMemcpy2DAsync(r.device_y, row_bytes, y, sizeof(float) * M.ldy, row_bytes, r.k,
                                    hipMemcpyDeviceToDevice, s));
    if (r.host_y && pinned)
        CLSPMV_HIP(hipMemcpy2DAsync(pinned, row_bytes, y, sizeof(float) * M.ldy, row_bytes, r.k,
                                    hipMemcpyDeviceToHost, s));
    return CLSPMV_OK;
}

int graph_for(clspmv_server* S, SMatrix& M, int K, hipGraphExec_t* out) {
    if (!M.exec[K]) {
        hipStream_t s = S->streams[M.stream];
        CLSPMV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int st = enqueue_pass(s, M, K);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "capture failed: %s", hipGetErrorString(e));
        M.graph[K] = g;
        CLSPMV_HIP(hipGraphInstantiate(&M.exec[K], g, nullptr, nullptr, 0));
    }
    *out = M.exec[K];
    return CLSPMV_OK;
}

// Launches a batch (scheduler thread, lock held): the requests' columns are
// assigned in order.
int launch_batch(clspmv_server* S, SMatrix& M, SBatch& b) {
    hipStream_t s = S->streams[M.stream];
    int K = 0;
    for (long long t : b.tickets) K += S->reqs[t].r.k;
    int g = 0;
    for (long long t : b.tickets) {
        SRequest& q = S->reqs[t];
        q.col = g;
        int st = stage_x(s, M, q.r, K, g, false);
        if (st) return st;
        g += q.r.k;
    }
    if (S->cfg.use_graphs) {
        hipGraphExec_t ex = nullptr;
        int st = graph_for(S, M, K, &ex);
        if (st) return st;
        CLSPMV_HIP(hipGraphLaunch(ex, s));
    } else {
        int st = enqueue_pass(s, M, K);
        if (st) return st;
    }
    for (long long t : b.tickets) {
        SRequest& q = S->reqs[t];
        int st = copy_y(s, M, q.r, q.col, q.pin != kNone ? S->pinned + q.pin / sizeof(float) : nullptr);
        if (st) return st;
    }
    if (S->free_events.empty()) {
        hipEvent_t e = nullptr;
        CLSPMV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        S->free_events.push_back(e);
    }
    b.ev = S->free_events.back();
    S->free_events.pop_back();
    CLSPMV_HIP(hipEventRecord(b.ev, s));
    ++S->n_passes;
    S->n_columns += K;
    return CLSPMV_OK;
}

// Marks a batch's requests done (lock held); their y is in place.
void finish_batch(clspmv_server* S, SBatch& b, int status, double t) {
    SMatrix& M = S->mats[b.matrix];
    for (long long tk : b.tickets) {
        SRequest& q = S->reqs[tk];
        if (q.pin != kNone) {
            S->pin_arena.release(q.pin, sizeof(float) * (size_t)M.rows * q.r.k);
            q.pin = kNone;
        }
        q.done = t;
        q.state = status ? status : 2;
        ++S->n_requests;
    }
    M.running = false;
    M.last_use = ++S->clock_use;
    release_claim(M.A);
    if (b.ev) S->free_events.push_back(b.ev);
    b.ev = nullptr;
}

struct HostCopy {
    float* dst;
    const float* src;
    size_t bytes;
};

void scheduler(clspmv_server* S) {
    (void)hipSetDevice(S->cfg.device_ordinal);
    std::unique_lock<std::mutex> lk(S->mu);
    std::vector<int> order;
    std::vector<std::pair<SBatch, int>> done;
    std::vector<HostCopy> copies;
    for (;;) {
        if (S->stop && S->pending == 0 && S->running.empty()) break;
        bool progress = false;
        // Retire: finished batches leave the running list; their host_y
        // copies (pinned staging -> caller memory) run without the lock --
        // the staging ranges and host_y stay the requests' until finished.
        done.clear();
        copies.clear();
        for (size_t i = 0; i < S->running.size();) {
            const hipError_t e = hipEventQuery(S->running[i].ev);
            if (e == hipErrorNotReady) {
                ++i;
                continue;
            }
            const int st = e == hipSuccess ? 0 : CLSPMV_ERR_HIP;
            if (!st) {
                const SMatrix& M = S->mats[S->running[i].matrix];
                for (long long tk : S->running[i].tickets) {
                    const SRequest& q = S->reqs[tk];
                    if (q.pin != kNone)
                        copies.push_back({q.r.host_y, S->pinned + q.pin / sizeof(float),
                                          sizeof(float) * (size_t)M.rows * q.r.k});
                }
            }
            done.emplace_back(std::move(S->running[i]), st);
            S->running[i] = std::move(S->running.back());
            S->running.pop_back();
        }
        if (!done.empty()) {
            if (!copies.empty()) {
                lk.unlock();
                for (const HostCopy& c : copies) std::memcpy(c.dst, c.src, c.bytes);
                lk.lock();
            }
            const double tdone = S->now();
            for (auto& d : done) finish_batch(S, d.first, d.second, tdone);
            progress = true;
        }
        // Admit: idle matrices whose oldest request has arrived, oldest first.
        const double t = S->now();
        order.clear();
        for (int m = 0; m < (int)S->mats.size(); ++m) {
            const SMatrix& M = S->mats[m];
            if (!M.running && !M.waiting.empty() && S->reqs[M.waiting.front()].arrival <= t) order.push_back(m);
        }
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return S->reqs[S->mats[a].waiting.front()].arrival < S->reqs[S->mats[b].waiting.front()].arrival;
        });
        for (int m : order) {
            SMatrix& M = S->mats[m];
            if (!ensure_slot(S, M)) continue;  // the pool is held by running passes: wait
            if (!try_claim(M.A)) continue;     // another server's pass is on this matrix
            SBatch b;
            b.matrix = m;
            int K = 0;
            while (!M.waiting.empty() && (int)b.tickets.size() < S->cfg.max_batch) {
                SRequest& q = S->reqs[M.waiting.front()];
                if (q.arrival > t || K + q.r.k > M.kmax) break;
                if (q.r.host_y) {
                    q.pin = S->pin_arena.alloc(sizeof(float) * (size_t)M.rows * q.r.k);
                    if (q.pin == kNone) break;  // staging full: later
                }
                K += q.r.k;
                q.state = 1;
                b.tickets.push_back(M.waiting.front());
                M.waiting.pop_front();
                --S->pending;
            }
            if (b.tickets.empty()) {
                release_claim(M.A);
                continue;
            }
            M.running = true;
            const int st = launch_batch(S, M, b);
            if (st) {
                (void)hipStreamSynchronize(S->streams[M.stream]);
                finish_batch(S, b, st, S->now());
            } else {
                S->running.push_back(std::move(b));
            }
            progress = true;
        }
        if (progress) S->cv.notify_all();
        if (!progress) {
            // Nothing to do right now: wait briefly for a submission (the
            // device keeps running; completions are polled).
            if (S->running.empty() && S->pending == 0)
                S->cv.wait_for(lk, std::chrono::milliseconds(1));
            else {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }
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

// Sets A->serve_multi (once per device matrix, under its claim): one
// multi-vector pass against separate SpMVs at min(kmax, 8) right-hand
// sides, on a stream and buffers of its own.
int calibrate(SMatrix M) {
    const int k = std::min(M.kmax, 8);
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CLSPMV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int st = CLSPMV_OK;
    void* buf = nullptr;
    if (hipMalloc(&buf, sizeof(float) * (size_t)k * (M.ldx + M.ldy)) != hipSuccess)
        st = fail(CLSPMV_ERR_ALLOC, "calibration buffers");
    if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = fail(CLSPMV_ERR_HIP, "hipEventCreate failed");
    float t_single = 0.0f, t_multi = 0.0f;
    if (!st) {
        M.X = buf;
        M.Y = static_cast<float*>(buf) + (size_t)k * M.ldx;
        const long long n = (long long)k * M.ldx;
        hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, static_cast<float*>(M.X));
        st = hipGetLastError() == hipSuccess ? CLSPMV_OK : fail(CLSPMV_ERR_HIP, "zero launch failed");
        M.path = kSingle;
        if (!st) st = time_pass(s, e0, e1, [&] { return enqueue_pass(s, M, k); }, &t_single);
        M.path = kMulti;
        if (!st) st = time_pass(s, e0, e1, [&] { return enqueue_pass(s, M, k); }, &t_multi);
    }
    (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (buf) (void)hipFree(buf);
    (void)hipStreamDestroy(s);
    if (st) return st;
    M.A->serve_multi = t_multi < t_single ? 1 : 0;
    return CLSPMV_OK;
}

// The baseline: one request alone on M's buffers, one SpMV launch per
// right-hand side (or one SpMM), synchronised at the end.
int run_plain(clspmv_server* S, SMatrix& M, const clspmv_request& r, float* pinned) {
    hipStream_t s = S->streams[0];
    SMatrix P = M;  // the same buffers, run per column (or SpMM on column-major X)
    if (P.path == kMulti) P.path = kSingle;
    int st = stage_x(s, P, r, r.k, 0, true);
    if (!st) {
        const long long ny = (long long)P.ldy * r.k;
        hipLaunchKernelGGL(zero_kernel, dim3(grid_for(ny)), dim3(256), 0, s, ny, P.Y);
        st = hipGetLastError() == hipSuccess ? CLSPMV_OK : fail(CLSPMV_ERR_HIP, "zero launch failed");
    }
    if (!st && P.path == kSpmm) {
        const clspmv_dmat* A = P.A;
        hipError_t e = launch::bspmm4(s, A->dtype == CLSPMV_DT_BF16, A->long_brows, A->brows, P.rows, P.cols, r.k,
                                      A->rp, A->col, A->lpval, static_cast<const unsigned short*>(P.X), P.ldx, P.Y,
                                      P.ldy, /*tiled=*/false, /*accumulate=*/true,
                                      launch::SpmmWindows{A->wtile, A->widx, A->win_max});
        if (e != hipSuccess) st = fail(CLSPMV_ERR_HIP, "spmm launch failed: %s", hipGetErrorString(e));
    } else {
        for (int j = 0; j < r.k && !st; ++j)
            st = dmat_launch(s, P.A, P.method, static_cast<const float*>(P.X) + (size_t)j * P.ldx,
                             P.Y + (size_t)j * P.ldy);
    }
    if (!st) st = copy_y(s, P, r, 0, pinned);
    if (!st) CLSPMV_HIP(hipStreamSynchronize(s));
    if (!st && r.host_y) std::memcpy(r.host_y, pinned, sizeof(float) * (size_t)P.rows * r.k);
    if (st) (void)hipStreamSynchronize(s);
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

void free_server(clspmv_server* S) {
    for (SMatrix& M : S->mats) drop_graphs(M);
    for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
    for (SBatch& b : S->running)
        if (b.ev) (void)hipEventDestroy(b.ev);
    if (S->pool) (void)hipFree(S->pool);
    if (S->pinned) (void)hipHostFree(S->pinned);
    for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
    delete S;
}

}  // namespace

extern "C" {

int clspmv_server_create(const clspmv_server_config* cfg, clspmv_server** out) {
    if (!cfg || !out) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (cfg->streams < 1 || cfg->max_batch < 1)
        return fail(CLSPMV_ERR_INVALID, "streams and max_batch must be >= 1");
    if (cfg->hbm_budget < kPoolAlign) return fail(CLSPMV_ERR_INVALID, "hbm_budget below %zu bytes", kPoolAlign);
    int n = 0;
    CLSPMV_HIP(hipGetDeviceCount(&n));
    if (cfg->device_ordinal < 0 || cfg->device_ordinal >= n)
        return fail(CLSPMV_ERR_INVALID, "device %d not present", cfg->device_ordinal);
    CLSPMV_HIP(hipSetDevice(cfg->device_ordinal));
    clspmv_server* S = new clspmv_server();
    S->cfg = *cfg;
    if (!S->cfg.host_staging) S->cfg.host_staging = (size_t)256 << 20;
    S->t0 = std::chrono::steady_clock::now();
    hipError_t e = hipSuccess;
    for (int i = 0; i < cfg->streams && e == hipSuccess; ++i) {
        hipStream_t s = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) S->streams.push_back(s);
    }
    if (e == hipSuccess) e = hipMalloc(&S->pool, cfg->hbm_budget);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&S->pinned), S->cfg.host_staging);
    if (e != hipSuccess) {
        free_server(S);
        return fail(CLSPMV_ERR_ALLOC, "server buffers (pool %zu bytes, staging %zu): %s", cfg->hbm_budget,
                    cfg->host_staging, hipGetErrorString(e));
    }
    S->pool_arena.init(cfg->hbm_budget);
    S->pin_arena.init(S->cfg.host_staging);
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
    for (hipStream_t s : S->streams) (void)hipStreamSynchronize(s);
    free_server(S);
    return CLSPMV_OK;
}

int clspmv_server_add_matrix(clspmv_server* S, const clspmv_dmat* A, int method, int kmax, int* id) {
    if (!S || !A || !id) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (kmax < 1) return fail(CLSPMV_ERR_INVALID, "kmax must be >= 1");
    if (A->device != S->cfg.device_ordinal)
        return fail(CLSPMV_ERR_INVALID, "matrix lives on device %d, server on %d", A->device, S->cfg.device_ordinal);
    SMatrix M;
    M.A = A;
    const bool lowp = A->info.format == CLSPMV_FMT_BSPMM;
    M.method = lowp ? 0 : (method == CLSPMV_METHOD_AUTO ? auto_method(A) : method);
    if (!lowp && (M.method < 0 || M.method >= clspmv_method_count(A->info.format)))
        return fail(CLSPMV_ERR_INVALID, "method %d not available for format %d", method, A->info.format);
    M.rows = A->info.rows;
    M.cols = A->info.cols;
    M.kmax = kmax;
    M.ldx = round_up(std::max(M.cols, 1), 4);
    M.ldy = round_up(std::max(M.rows, 1), 4);
    M.path = lowp ? kSpmm : kSingle;
    auto aligned = [](size_t n) { return (n + kPoolAlign - 1) / kPoolAlign * kPoolAlign; };
    M.x_bytes = aligned((size_t)M.ldx * kmax * (lowp ? 2 : 4));
    M.y_bytes = aligned(sizeof(float) * (size_t)M.ldy * kmax);
    M.graph.assign(kmax + 1, nullptr);
    M.exec.assign(kmax + 1, nullptr);
    CLSPMV_HIP(hipSetDevice(S->cfg.device_ordinal));
    if (!lowp && has_multi(A)) {
        // Fixed once per device matrix, while no pass uses it.
        while (!try_claim(A)) std::this_thread::sleep_for(std::chrono::microseconds(50));
        int st = A->serve_multi < 0 ? calibrate(M) : CLSPMV_OK;
        const int multi = A->serve_multi;
        release_claim(A);
        if (st) return st;
        if (multi == 1) M.path = kMulti;
    }
    if (M.path != kSingle) M.s_bytes = aligned(sizeof(float) * (size_t)M.ldx * kmax);  // host x before conversion
    std::lock_guard<std::mutex> lk(S->mu);
    if (M.slot_bytes() > S->pool_arena.size)
        return fail(CLSPMV_ERR_TOO_LARGE, "working buffers of %d right-hand sides (%zu bytes) exceed the pool (%zu)",
                    kmax, M.slot_bytes(), S->pool_arena.size);
    M.stream = (int)(S->mats.size() % S->streams.size());
    *id = (int)S->mats.size();
    S->mats.push_back(std::move(M));
    return CLSPMV_OK;
}

double clspmv_server_clock(clspmv_server* S) { return S ? S->now() : 0.0; }

int clspmv_server_submit(clspmv_server* S, const clspmv_request* r, long long* ticket) {
    if (!S || !r || !ticket) return fail(CLSPMV_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(CLSPMV_ERR_INVALID, "server is shutting down");
    if (r->matrix < 0 || r->matrix >= (int)S->mats.size())
        return fail(CLSPMV_ERR_INVALID, "matrix %d not registered", r->matrix);
    SMatrix& M = S->mats[r->matrix];
    if (r->k < 1 || r->k > M.kmax) return fail(CLSPMV_ERR_INVALID, "k = %d outside 1..%d", r->k, M.kmax);
    if (r->host_y && sizeof(float) * (size_t)M.rows * r->k > S->pin_arena.size)
        return fail(CLSPMV_ERR_TOO_LARGE, "host_y of %d x %d floats exceeds the pinned staging (%zu bytes)", M.rows,
                    r->k, S->pin_arena.size);
    SRequest q;
    q.r = *r;
    q.arrival = r->arrival > 0 ? r->arrival : S->now();
    const long long t = S->next_ticket++;
    S->reqs.emplace(t, q);
    // Keep the matrix's queue in arrival order (submissions usually arrive
    // in order).
    auto pos = M.waiting.end();
    while (pos != M.waiting.begin() && S->reqs[*(pos - 1)].arrival > q.arrival) --pos;
    M.waiting.insert(pos, t);
    ++S->pending;
    *ticket = t;
    S->cv.notify_all();
    return CLSPMV_OK;
}

int clspmv_server_wait(clspmv_server* S, long long ticket, double* latency_s) {
    if (!S) return fail(CLSPMV_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return fail(CLSPMV_ERR_INVALID, "unknown ticket %lld", ticket);
    S->cv.wait(lk, [&] {
        const SRequest& q = S->reqs[ticket];
        return q.state >= 2 || q.state < 0;
    });
    SRequest q = S->reqs[ticket];
    S->reqs.erase(ticket);
    if (latency_s) *latency_s = q.done - q.arrival;
    if (q.state < 0) return fail(q.state, "request %lld failed", ticket);
    return CLSPMV_OK;
}

int clspmv_server_get_stats(clspmv_server* S, clspmv_server_stats* out) {
    if (!S || !out) return fail(CLSPMV_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out->requests = S->n_requests;
    out->passes = S->n_passes;
    out->columns = S->n_columns;
    out->evictions = S->n_evictions;
    out->pool_bytes = S->pool_arena.size;
    out->pool_used = S->pool_arena.used;
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
            const clspmv_request& r = reqs[i];
            if (r.matrix < 0 || r.matrix >= (int)S->mats.size() || r.k < 1 || r.k > S->mats[r.matrix].kmax)
                return fail(CLSPMV_ERR_INVALID, "request %d: bad matrix or k", i);
            const double arr = start + std::max(0.0, r.arrival);
            sleep_until(S, arr);
            // Take the matrix's buffers (and staging) from the serving loop.
            SMatrix* M = nullptr;
            size_t pin = kNone;
            for (;;) {
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    SMatrix& m = S->mats[r.matrix];
                    const size_t pbytes = sizeof(float) * (size_t)m.rows * r.k;
                    if (!m.running && ensure_slot(S, m) && (!r.host_y || (pin = S->pin_arena.alloc(pbytes)) != kNone)) {
                        if (try_claim(m.A)) {
                            m.running = true;
                            M = &m;
                        } else if (pin != kNone) {
                            S->pin_arena.release(pin, pbytes);
                            pin = kNone;
                        }
                    }
                }
                if (M) break;
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
            int st = run_plain(S, *M, r, pin != kNone ? S->pinned + pin / sizeof(float) : nullptr);
            {
                std::lock_guard<std::mutex> lk(S->mu);
                if (pin != kNone) S->pin_arena.release(pin, sizeof(float) * (size_t)M->rows * r.k);
                M->running = false;
                M->last_use = ++S->clock_use;
                release_claim(M->A);
            }
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
            if (st) {
                for (int j = 0; j < i; ++j) (void)clspmv_server_wait(S, tickets[j], nullptr);
                return st;
            }
        }
        int first_err = CLSPMV_OK;
        for (int i = 0; i < n; ++i) {
            double lat = 0.0;
            int st = clspmv_server_wait(S, tickets[i], &lat);
            if (st && !first_err) first_err = st;
            if (latencies) latencies[i] = lat;
        }
        if (first_err) return first_err;
        last = S->now();
    }
    if (wall_s) *wall_s = last - start;
    return CLSPMV_OK;
}

void clspmv_request_x(unsigned long long seed, int j, int n, float* out) {
    for (int i = 0; i < n; ++i) out[i] = uniform_at(seed, j, i);
}

}  // extern "C"
