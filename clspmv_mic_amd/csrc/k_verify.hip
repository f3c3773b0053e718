// FP64 verification SpMV on the GPU and the per-matrix error report.
//
// The reference checks a device result on the host: spmv_only recomputes
// y = A x serially in FP32 and two_vec_compare reports the differences
// (src/util.cpp:85-131, called from src/eval.cpp:79-97).  Here the check
// runs next to the data: an FP32 CSR matrix (the original values) is
// multiplied in FP64 -- products double(a) * double(x) are exact -- and
// every row's result y is scored by the tolerance metric
//     e_i = |y_i - y64_i| / (|y0_i| + sum_j |a_ij x_j|)
// (0 when both sides are 0 and y_i matches exactly, +inf when only the
// bound is 0; a NaN counts as +inf).  Work-groups reduce their rows in a
// fixed tree order; the host combines the per-group results in group
// order, so the report is the same on every run.

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "kernels.hpp"
#include "rng.hpp"
#include "runtime.hpp"

// Per-work-group results of clspmv_verify_spmv_noalloc live in this
// file-scope array instead of a buffer allocated per call: 32 bytes per CSR
// row block, one copy per GPU, held until the process ends.  In host code
// the name `scratch` is a stand-in, not a device address: each GPU's address
// is looked up once with hipGetSymbolAddress (noalloc_scratch_init, run by
// clspmv_ctx_create with that GPU current) and only that pointer reaches
// kernels.
static __device__ __align__(256) unsigned char scratch[CLSPMV_NOALLOC_MAX_ROW_BLOCKS * 32];

namespace clspmv {
namespace {

using namespace dev;

struct VerifyPart {
    double max_err;  // largest row error of the group
    double sum_sq;   // sum of squared row errors
    double max_abs;  // largest |y - y64|
    int worst;       // first row reaching max_err
    int over;        // rows above tol
};
static_assert(sizeof(VerifyPart) == 32, "scratch holds 32 bytes per row block");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per GPU

__device__ inline double row_error(double y, double y64, double bound) {
    const double d = fabs(y - y64);
    double e = bound > 0.0 ? d / bound : (d > 0.0 ? INFINITY : 0.0);
    if (e != e) e = INFINITY;
    return e;
}

__device__ inline void keep_worst(double e, int r, double& be, int& br) {
    if (e > be || (e == be && r < br)) {
        be = e;
        br = r;
    }
}

__device__ inline double wave_sum(double v, int g) {
    for (int m = g >> 1; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// One work-group per adaptive row block (csr_adaptive's blocks): a run of
// rows whose nonzeros fit kAdaptiveNnz, or one longer row.
__global__ __launch_bounds__(kBlock) void verify_csr_kernel(
    const int* __restrict__ blocks, const int* __restrict__ rp, const int* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ x, const float* __restrict__ y0,
    const float* __restrict__ y, double tol, VerifyPart* __restrict__ part) {
    __shared__ double s_max[kBlock], s_sq[kBlock], s_abs[kBlock];
    __shared__ int s_row[kBlock], s_over[kBlock];
    __shared__ double s_wave[2][kBlock / kWave];
    const int b = blockIdx.x, t = threadIdx.x;
    const int r0 = blocks[b], r1 = blocks[b + 1];
    const int nrows = r1 - r0;
    double me = -1.0, msq = 0.0, mabs = 0.0;
    int mrow = 0x7FFFFFFF, mover = 0;
    auto score = [&](int r, double acc, double bnd) {
        const double yy0 = y0 ? (double)y0[r] : 0.0;
        const double y64 = yy0 + acc;
        const double bound = fabs(yy0) + bnd;
        const double e = row_error((double)y[r], y64, bound);
        keep_worst(e, r, me, mrow);
        msq += e * e;
        mabs = fmax(mabs, fabs((double)y[r] - y64));
        mover += e > tol;
    };
    if (nrows == 1 && rp[r1] - rp[r0] > launch::kAdaptiveNnz) {
        double acc = 0.0, bnd = 0.0;
        for (int j = rp[r0] + t; j < rp[r1]; j += kBlock) {
            const double p = (double)val[j] * (double)x[col[j]];
            acc += p;
            bnd += fabs(p);
        }
        acc = wave_sum(acc, kWave);
        bnd = wave_sum(bnd, kWave);
        if ((t & (kWave - 1)) == 0) {
            s_wave[0][t / kWave] = acc;
            s_wave[1][t / kWave] = bnd;
        }
        __syncthreads();
        if (t == 0) {
            double a = 0.0, c = 0.0;
            for (int w = 0; w < kBlock / kWave; ++w) {
                a += s_wave[0][w];
                c += s_wave[1][w];
            }
            score(r0, a, c);
        }
    } else {
        int g = kBlock / (nrows > 0 ? nrows : 1);
        g = g >= 64 ? 64 : g >= 32 ? 32 : g >= 16 ? 16 : g >= 8 ? 8 : g >= 4 ? 4 : g >= 2 ? 2 : 1;
        const int lane = t & (g - 1);
        const int r = t / g;  // nrows <= kAdaptiveRows = kBlock, so one pass
        double acc = 0.0, bnd = 0.0;
        if (r < nrows) {
            for (int j = rp[r0 + r] + lane; j < rp[r0 + r + 1]; j += g) {
                const double p = (double)val[j] * (double)x[col[j]];
                acc += p;
                bnd += fabs(p);
            }
        }
        acc = wave_sum(acc, g);
        bnd = wave_sum(bnd, g);
        if (r < nrows && lane == 0) score(r0 + r, acc, bnd);
    }
    s_max[t] = me;
    s_row[t] = mrow;
    s_sq[t] = msq;
    s_abs[t] = mabs;
    s_over[t] = mover;
    __syncthreads();
    for (int h = kBlock / 2; h >= 1; h >>= 1) {
        if (t < h) {
            keep_worst(s_max[t + h], s_row[t + h], s_max[t], s_row[t]);
            s_sq[t] += s_sq[t + h];
            s_abs[t] = fmax(s_abs[t], s_abs[t + h]);
            s_over[t] += s_over[t + h];
        }
        __syncthreads();
    }
    if (t == 0) part[b] = VerifyPart{s_max[0], s_sq[0], s_abs[0], s_row[0], s_over[0]};
}

// clspmv_verify_spmv_noalloc: one thread combines the work-groups' results
// in group (row) order, the order and arithmetic verify_into uses on the
// host, and writes the report to device memory.
__global__ void verify_combine_kernel(const VerifyPart* __restrict__ part, int nparts, int rows, double tol,
                                      clspmv_error_report* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    clspmv_error_report r{};
    r.rows = rows;
    r.tol = tol;
    r.worst_row = -1;
    if (rows > 0) {
        double best = -1.0, sq = 0.0, mabs = 0.0;
        int row = -1;
        long long over = 0;
        for (int i = 0; i < nparts; ++i) {
            const VerifyPart p = part[i];
            if (p.max_err > best) {
                best = p.max_err;
                row = p.worst;
            }
            sq += p.sum_sq;
            mabs = fmax(mabs, p.max_abs);
            over += p.over;
        }
        r.max_scaled_error = best;
        r.rms_scaled_error = sqrt(sq / rows);
        r.max_abs_error = mabs;
        r.worst_row = row;
        r.rows_over_tol = over;
    }
    *out = r;
}

__global__ void fill_uniform_kernel(unsigned long long seed, int n, float* x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] = uniform_at(seed, 0, i);
}

__global__ void round16_kernel(int n, bool bf16, const float* __restrict__ x, unsigned short* __restrict__ h) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        h[i] = bf16 ? to_bf16(x[i]) : to_f16(x[i]);
}

int grid_of(long long n) { return (int)std::min<long long>((n + kBlock - 1) / kBlock, 4096); }

// Device temporaries freed on every exit path (after the stream drains).
struct DeviceTemps {
    std::vector<void*> ptrs;
    ~DeviceTemps() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int get(size_t n, T** out) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
        if (e != hipSuccess) return fail(CLSPMV_ERR_ALLOC, "hipMalloc failed: %s", hipGetErrorString(e));
        ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return CLSPMV_OK;
    }
};

int verify_into(clspmv_ctx* ctx, const clspmv_dmat* ref, const float* d_x, const float* d_y0,
                const float* d_y, double tol, clspmv_error_report* rep) {
    const clspmv_dmat_info& I = ref->info;
    *rep = clspmv_error_report{};
    rep->rows = I.rows;
    rep->tol = tol;
    rep->worst_row = -1;
    if (I.rows == 0) return CLSPMV_OK;
    DeviceTemps sc;
    VerifyPart* d_part = nullptr;
    int st = sc.get(ref->nblocks, &d_part);
    if (st) return st;
    hipLaunchKernelGGL(verify_csr_kernel, dim3(ref->nblocks), dim3(kBlock), 0, ctx->stream, ref->blocks,
                       ref->rp, ref->col, ref->val, d_x, d_y0, d_y, tol, d_part);
    CLSPMV_HIP(hipGetLastError());
    std::vector<VerifyPart> h(ref->nblocks);
    CLSPMV_HIP(hipMemcpyAsync(h.data(), d_part, sizeof(VerifyPart) * h.size(), hipMemcpyDeviceToHost,
                              ctx->stream));
    CLSPMV_HIP(hipStreamSynchronize(ctx->stream));
    double best = -1.0, sq = 0.0, mabs = 0.0;
    int row = -1;
    long long over = 0;
    for (const VerifyPart& p : h) {  // blocks are in row order
        if (p.max_err > best) {
            best = p.max_err;
            row = p.worst;
        }
        sq += p.sum_sq;
        mabs = std::fmax(mabs, p.max_abs);
        over += p.over;
    }
    rep->max_scaled_error = best;
    rep->rms_scaled_error = std::sqrt(sq / I.rows);
    rep->max_abs_error = mabs;
    rep->worst_row = row;
    rep->rows_over_tol = over;
    return CLSPMV_OK;
}

int check_ref(const clspmv_ctx* ctx, const clspmv_dmat* ref) {
    if (ref->info.format != CLSPMV_FMT_CSR || ref->dtype != CLSPMV_DT_F32 || !ref->plain)
        return fail(CLSPMV_ERR_INVALID, "the reference must be an FP32 CSR matrix holding its own arrays "
                                        "(clspmv_upload_csr with panel off or both encodings)");
    if (ref->device != ctx->device)
        return fail(CLSPMV_ERR_INVALID, "reference lives on device %d, context on %d", ref->device, ctx->device);
    return CLSPMV_OK;
}

void* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(CLSPMV_ERR_UNSUPPORTED, "device ordinal %d: at most %d devices", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return CLSPMV_OK;
    void* p = nullptr;
    CLSPMV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
    return CLSPMV_OK;
}

}  // namespace clspmv

using namespace clspmv;

extern "C" {

int clspmv_verify_spmv(clspmv_ctx* ctx, const clspmv_dmat* ref, const float* d_x, const float* d_y0,
                       const float* d_y, double tol, clspmv_error_report* rep) {
    if (!ctx || !ref || !rep) return fail(CLSPMV_ERR_INVALID, "null argument");
    int st = check_ref(ctx, ref);
    if (st) return st;
    if (ref->info.rows > 0 && (!d_y || (ref->info.nnz > 0 && !d_x))) return fail(CLSPMV_ERR_INVALID, "null vector");
    if ((st = use_ctx_device(ctx))) return st;
    return verify_into(ctx, ref, d_x, d_y0, d_y, tol, rep);
}

int clspmv_verify_spmv_noalloc(clspmv_ctx* ctx, const clspmv_dmat* ref, const float* d_x, const float* d_y0,
                               const float* d_y, double tol, clspmv_error_report* d_report) {
    if (!ctx || !ref || !d_report) return fail(CLSPMV_ERR_INVALID, "null argument");
    int st = check_ref(ctx, ref);
    if (st) return st;
    const int rows = ref->info.rows;
    if (rows > 0 && (!d_y || (ref->info.nnz > 0 && !d_x))) return fail(CLSPMV_ERR_INVALID, "null vector");
    if (ref->nblocks > CLSPMV_NOALLOC_MAX_ROW_BLOCKS)
        return fail(CLSPMV_ERR_TOO_LARGE, "the reference has %d row blocks; the no-allocation scratch holds %d",
                    ref->nblocks, CLSPMV_NOALLOC_MAX_ROW_BLOCKS);
    if ((st = use_ctx_device(ctx))) return st;
    VerifyPart* part = static_cast<VerifyPart*>(scratch_of(ctx->device));
    if (!part) return fail(CLSPMV_ERR_INVALID, "no scratch set up for device %d", ctx->device);
    if (rows > 0) {
        hipLaunchKernelGGL(verify_csr_kernel, dim3(ref->nblocks), dim3(kBlock), 0, ctx->stream, ref->blocks,
                           ref->rp, ref->col, ref->val, d_x, d_y0, d_y, tol, part);
        CLSPMV_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(verify_combine_kernel, dim3(1), dim3(kWave), 0, ctx->stream, part,
                       rows > 0 ? ref->nblocks : 0, rows, tol, d_report);
    CLSPMV_HIP(hipGetLastError());
    return CLSPMV_OK;
}

int clspmv_dmat_error_report(clspmv_ctx* ctx, const clspmv_dmat* A, const clspmv_dmat* ref,
                             unsigned long long seed, double tol, clspmv_error_report* rep) {
    if (!ctx || !A || !ref || !rep) return fail(CLSPMV_ERR_INVALID, "null argument");
    int st = check_ref(ctx, ref);
    if (st) return st;
    const clspmv_dmat_info& I = A->info;
    if (I.rows != ref->info.rows || I.cols != ref->info.cols)
        return fail(CLSPMV_ERR_INVALID, "matrix is %dx%d, reference %dx%d", I.rows, I.cols, ref->info.rows,
                    ref->info.cols);
    if ((st = use_ctx_device(ctx))) return st;
    const bool block = I.format == CLSPMV_FMT_BSPMM;
    const int ldx = (I.cols + 3) / 4 * 4, ldy = (I.rows + 3) / 4 * 4;
    DeviceTemps sc;
    float *x = nullptr, *y = nullptr;
    unsigned short* x16 = nullptr;
    if ((st = sc.get(ldx, &x)) || (st = sc.get(ldy, &y))) return st;
    if (block && (st = sc.get(ldx, &x16))) return st;
    hipStream_t s = ctx->stream;
    CLSPMV_HIP(hipMemsetAsync(x, 0, sizeof(float) * ldx, s));
    CLSPMV_HIP(hipMemsetAsync(y, 0, sizeof(float) * ldy, s));
    if (I.cols > 0) {
        hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_of(I.cols)), dim3(kBlock), 0, s, seed, I.cols, x);
        CLSPMV_HIP(hipGetLastError());
    }
    if (block) {
        hipLaunchKernelGGL(round16_kernel, dim3(grid_of(ldx)), dim3(kBlock), 0, s, ldx,
                           A->dtype == CLSPMV_DT_BF16, x, x16);
        CLSPMV_HIP(hipGetLastError());
        st = clspmv_spmm(ctx, A, 1, x16, ldx, y, ldy);
    } else {
        st = clspmv_spmv(ctx, A, CLSPMV_METHOD_AUTO, x, y);
    }
    if (!st) st = verify_into(ctx, ref, x, nullptr, y, tol, rep);
    if (st) (void)hipStreamSynchronize(s);  // before DeviceTemps frees what a launch may still use
    return st;
}

}  // extern "C"
