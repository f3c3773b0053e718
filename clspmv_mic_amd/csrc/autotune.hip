// MI355X autotuner: measures every kernel variant on synthetic matrices and
// writes the ten .ben tables the Cocktail partitioner reads.
//
// Replaces the reference's offline microbenchmarks (src/bench_*.cpp, run by
// the `benchmarking` script) with one library call.  The probe matrices are
// generated directly in HBM by fill kernels, so even the largest probes (a
// 2M-row matrix with up to 256M stored entries) cost no host work.  Probe
// patterns:
//   DIA / BDIA : `width` consecutive diagonals around the main diagonal (one
//                band for BDIA), as the reference's bench_dia/bench_bdia;
//   ELL / SELL / CSR / COO : `width` entries per row at pseudo-random
//                columns, i.e. the gather pattern of an unstructured
//                remainder;
//   BCSR / BELL / SBELL : `blocks` dense blocks per block row at
//                pseudo-random block columns.
// Each table row holds GFLOP/s = 2 * nnz / time for every method; a method
// that cannot run on the probe's shape gets 0 and is never chosen.
//
// The knobs a table column stands for are tuned inside it: CSR method 0 and
// BCSR at the fastest lanes-per-row (powers of two around the shape's own
// count, as the runtime tuner tries them); CSR and COO at the faster of the
// format's own arrays and the x-panel encoding (the Cocktail runtime
// uploads both and keeps the faster), the panel probe built on the host
// for probes of up to kPanelProbeMax entries and its rate carried to the
// larger dimensions of the same width; SELL's two columns are its two slice
// heights (32 and 128, the partitioner's sell_meth_num); block shapes are
// the eight rows of each block table.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>

#include "cocktail.hpp"
#include "device_common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace clspmv;

namespace {

using namespace clspmv::dev;

__host__ __device__ __forceinline__ unsigned mix(unsigned long long v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return (unsigned)(v ^ (v >> 31));
}

__global__ void fill_f32(float* p, long long n, float v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// p[i] = (hash(i / per, i % per) % mod) * scale: pseudo-random column ids.
__global__ void fill_cols(int* p, long long n, int per, int mod, int scale, int salt) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = (int)(mix(((unsigned long long)(i / per) << 20) ^ (unsigned long long)(i % per) ^
                         ((unsigned long long)salt << 52)) % (unsigned)mod) * scale;
}

// p[i] = i * step (row pointers of constant-width rows).
__global__ void fill_ramp(int* p, long long n, int step) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = (int)(i * step);
}

// p[i] = i / per (COO row ids of constant-width rows).
__global__ void fill_div(int* p, long long n, int per) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = (int)(i / per);
}

__global__ void empty_kernel() {}

int grid_of(long long n) { return (int)std::min<long long>((n + 255) / 256, 8192); }

struct Bufs {
    clspmv_dmat* m;
    template <class T>
    int alloc(long long n, T** out) {
        return upload_array<T>(m, nullptr, (size_t)std::max<long long>(n, 1), out);
    }
};

int fillf(hipStream_t s, float* p, long long n, float v) {
    hipLaunchKernelGGL(fill_f32, dim3(grid_of(n)), dim3(256), 0, s, p, n, v);
    CLSPMV_HIP(hipGetLastError());
    return CLSPMV_OK;
}
int fillc(hipStream_t s, int* p, long long n, int per, int mod, int salt, int scale = 1) {
    hipLaunchKernelGGL(fill_cols, dim3(grid_of(n)), dim3(256), 0, s, p, n, per, mod, scale, salt);
    CLSPMV_HIP(hipGetLastError());
    return CLSPMV_OK;
}
int fillr(hipStream_t s, int* p, long long n, int step) {
    hipLaunchKernelGGL(fill_ramp, dim3(grid_of(n)), dim3(256), 0, s, p, n, step);
    CLSPMV_HIP(hipGetLastError());
    return CLSPMV_OK;
}
int filld(hipStream_t s, int* p, long long n, int per) {
    hipLaunchKernelGGL(fill_div, dim3(grid_of(n)), dim3(256), 0, s, p, n, per);
    CLSPMV_HIP(hipGetLastError());
    return CLSPMV_OK;
}

clspmv_dmat* new_probe(int format, int rows, int cols, long long nnz) {
    clspmv_dmat* m = new clspmv_dmat();
    m->info.format = format;
    m->info.rows = rows;
    m->info.cols = cols;
    m->info.nnz = (int)nnz;
    m->info.bheight = m->info.bwidth = 1;
    (void)hipGetDevice(&m->device);
    return m;
}

// Probe generators: each builds a device matrix and returns its true nnz.
int make_dia(hipStream_t s, int dim, int width, bool banded, clspmv_dmat** out) {
    clspmv_dmat* m = new_probe(banded ? CLSPMV_FMT_BDIA : CLSPMV_FMT_DIA, dim, dim, 0);
    Bufs b{m};
    const int len = round_up(dim, CLSPMV_GPU_ALIGNMENT);
    const int first = -(width / 2);
    long long nnz = 0;
    for (int j = 0; j < width; ++j) nnz += dim - std::abs(first + j);
    m->info.nnz = (int)nnz;
    m->info.stored = (long long)width * len;
    m->len = len;
    std::vector<int> offs(width);
    for (int j = 0; j < width; ++j) offs[j] = first + j;
    int st = b.alloc((long long)width * len, &m->val);
    if (!st) st = fillf(s, m->val, (long long)width * len, 1.0f);
    if (!st && !banded) {
        m->ndiag = width;
        m->min_off = first;
        m->max_off = first + width - 1;
        st = upload_array(m, offs.data(), offs.size(), &m->offs);
    } else if (!st) {
        m->nband = 1;
        m->ndiag = width;
        m->max_band = width;
        int bptr[2] = {0, width};
        m->min_off = first;
        m->max_off = first + width - 1;
        st = upload_array(m, &first, 1, &m->offs);
        if (!st) st = upload_array(m, bptr, 2, &m->bptr);
        if (!st) st = upload_array(m, offs.data(), offs.size(), &m->offs_exp);
    }
    if (st) { clspmv_dmat_free(m); return st; }
    *out = m;
    return CLSPMV_OK;
}

int make_flat(hipStream_t s, int format, int dim, int width, int slice_h, clspmv_dmat** out) {
    const long long nnz = (long long)dim * width;
    clspmv_dmat* m = new_probe(format, dim, dim, nnz);
    Bufs b{m};
    int st = CLSPMV_OK;
    switch (format) {
        case CLSPMV_FMT_ELL: {
            m->hal = round_up(dim, CLSPMV_GPU_ALIGNMENT);
            m->ellnum = width;
            const long long tot = (long long)m->hal * width;
            m->info.stored = tot;
            st = b.alloc(tot, &m->col);
            if (!st) st = b.alloc(tot, &m->val);
            if (!st) st = fillc(s, m->col, tot, 1, dim, 1);
            if (!st) st = fillf(s, m->val, tot, 1.0f);
            break;
        }
        case CLSPMV_FMT_SELL: {
            const int ns = (dim + slice_h - 1) / slice_h;
            const long long tot = (long long)ns * slice_h * width;
            m->info.slice_height = slice_h;
            m->info.stored = tot;
            st = b.alloc(ns + 1, &m->sptr);
            if (!st) st = fillr(s, m->sptr, ns + 1, slice_h * width);
            if (!st) st = b.alloc(tot, &m->col);
            if (!st) st = b.alloc(tot, &m->val);
            if (!st) st = fillc(s, m->col, tot, 1, dim, 2);
            if (!st) st = fillf(s, m->val, tot, 1.0f);
            break;
        }
        case CLSPMV_FMT_CSR: {
            m->info.stored = nnz;
            m->lanes = 1;
            while (m->lanes < width && m->lanes < 64) m->lanes <<= 1;
            m->max_row = width;
            st = b.alloc(dim + 1LL, &m->rp);
            if (!st) st = fillr(s, m->rp, dim + 1LL, width);
            if (!st) st = b.alloc(nnz, &m->col);
            if (!st) st = b.alloc(nnz, &m->val);
            if (!st) st = fillc(s, m->col, nnz, width, dim, 3);
            if (!st) st = fillf(s, m->val, nnz, 1.0f);
            // Adaptive row blocks for constant row length.
            std::vector<int> blocks{0};
            if (width > launch::kAdaptiveNnz) {
                for (int r = 1; r <= dim; ++r) blocks.push_back(r);
            } else {
                const int per = std::max(1, std::min(launch::kAdaptiveRows, launch::kAdaptiveNnz / std::max(width, 1)));
                for (int r = per; r < dim; r += per) blocks.push_back(r);
                blocks.push_back(dim);
            }
            m->nblocks = (int)blocks.size() - 1;
            if (!st) st = upload_array(m, blocks.data(), blocks.size(), &m->blocks);
            break;
        }
        case CLSPMV_FMT_COO: {
            m->info.stored = nnz;
            const int chunks = (int)((nnz + launch::kCooChunk - 1) / launch::kCooChunk);
            st = b.alloc(nnz, &m->row);
            if (!st) st = b.alloc(nnz, &m->col);
            if (!st) st = b.alloc(nnz, &m->val);
            if (!st) st = filld(s, m->row, nnz, width);
            if (!st) st = fillc(s, m->col, nnz, width, dim, 4);
            if (!st) st = fillf(s, m->val, nnz, 1.0f);
            if (!st) st = b.alloc(2LL * chunks, &m->carry_row);
            if (!st) st = b.alloc(2LL * chunks, &m->carry_val);
            break;
        }
        default:
            st = fail(CLSPMV_ERR_INVALID, "not a flat format");
    }
    if (st) { clspmv_dmat_free(m); return st; }
    *out = m;
    return CLSPMV_OK;
}

// `blocks` dense bh x bw blocks per block row; rows = dim.
int make_block(hipStream_t s, int format, int dim, int bh, int bw, int blocks, clspmv_dmat** out) {
    const int nbr = (dim + bh - 1) / bh;
    const long long nblk = (long long)nbr * blocks;
    const int bw4 = bw / 4;
    clspmv_dmat* m = new_probe(format, dim, dim, nblk * bh * bw);
    m->info.bheight = bh;
    m->info.bwidth = bw;
    m->info.stored = nblk * bh * bw;
    m->brows = nbr;
    Bufs b{m};
    int st = CLSPMV_OK;
    // Block columns: a random block index times bw, kept in float4 units, so
    // a block never crosses the width.
    auto cols = [&](int* p, long long n, int per) {
        return fillc(s, p, n, per, std::max(1, dim / bw), 5, bw4);
    };
    switch (format) {
        case CLSPMV_FMT_BCSR: {
            // Every block row holds `blocks` blocks: the upload would pick the
            // slot-major encoding (no padding), so the probe times that.
            m->hal = round_up(nbr, 64);
            m->bslots = blocks;
            const long long nc = (long long)m->hal * blocks;
            st = b.alloc(nc, &m->col);
            if (!st) st = cols(m->col, nc, 1);
            const long long nd = nc * bh * bw4 * 4;
            if (!st) st = b.alloc(nd, &m->val);
            if (!st) st = fillf(s, m->val, nd, 1.0f);
            break;
        }
        case CLSPMV_FMT_BELL: {
            m->hal = round_up(nbr, CLSPMV_GPU_ALIGNMENT);
            m->f4 = round_up(4 * nbr, CLSPMV_GPU_ALIGNMENT);
            m->ellnum = blocks;
            const long long nc = (long long)m->hal * blocks;
            st = b.alloc(nc, &m->col);
            if (!st) st = cols(m->col, nc, 1);
            const long long nd = (long long)m->f4 * bh * bw4 * blocks;
            if (!st) st = b.alloc(nd, &m->val);
            if (!st) st = fillf(s, m->val, nd, 1.0f);
            break;
        }
        case CLSPMV_FMT_SBELL: {
            const int sh = CLSPMV_WARPSIZE;
            const int ns = (nbr + sh - 1) / sh;
            m->info.slice_height = sh;
            const long long tot = (long long)ns * sh * blocks;
            st = b.alloc(ns + 1LL, &m->sptr);
            if (!st) st = fillr(s, m->sptr, ns + 1LL, sh * blocks);
            if (!st) st = b.alloc(tot, &m->col);
            if (!st) st = cols(m->col, tot, 1);
            if (!st) st = b.alloc(tot * bh * bw, &m->val);
            if (!st) st = fillf(s, m->val, tot * bh * bw, 1.0f);
            break;
        }
        default:
            st = fail(CLSPMV_ERR_INVALID, "not a block format");
    }
    if (st) { clspmv_dmat_free(m); return st; }
    *out = m;
    return CLSPMV_OK;
}

constexpr long long kPanelProbeMax = 8LL << 20;

// The x-panel encoding of a CSR / COO probe (`width` entries per row at the
// columns fill_cols gives with `salt`), built on the host and added to A.
// False when panels do not pay for this shape (auto_panel_width).
int add_panel_probe(clspmv_dmat* A, int dim, int width, int salt, bool* made) {
    *made = false;
    const long long nnz = (long long)dim * width;
    const int pw = auto_panel_width(dim, nnz);
    if (!pw) return CLSPMV_OK;
    std::vector<int> row(nnz), col(nnz);
    std::vector<float> val(nnz, 1.0f);
    for (long long i = 0; i < nnz; ++i) {
        row[i] = (int)(i / width);
        col[i] = (int)(mix(((unsigned long long)(i / width) << 20) ^ (unsigned long long)(i % width) ^
                           ((unsigned long long)salt << 52)) %
                       (unsigned)dim);
    }
    // Entries of a row must be column-sorted for the encoding's row runs.
    for (int r = 0; r < dim; ++r) std::sort(col.begin() + (long long)r * width, col.begin() + (long long)(r + 1) * width);
    int st = build_panel(A, dim, dim, nnz, row.data(), col.data(), val.data(), pw, CLSPMV_DT_F32, 0);
    if (st) return st;
    *made = true;
    return CLSPMV_OK;
}

// Lane counts worth timing around a shape's own (powers of two).
std::vector<int> lane_sweep(int lanes) {
    std::vector<int> c;
    for (int l = std::max(1, lanes / 4); l <= std::min(64, lanes * 4); l <<= 1) c.push_back(l);
    return c;
}

struct Timer {
    clspmv_ctx* ctx;
    float* x;
    float* y;
    int reps;
    // The fastest over `lanes` (A->lanes set to each in turn, then restored).
    double gflops_lanes(clspmv_dmat* A, int method, long long nnz, const std::vector<int>& lanes) {
        const int keep = A->lanes;
        double best = 0.0;
        for (int l : lanes) {
            A->lanes = l;
            best = std::max(best, gflops(A, method, nnz));
        }
        A->lanes = keep;
        return best;
    }
    // GFLOP/s of `method` on A, 0 if the method cannot run on it.
    double gflops(const clspmv_dmat* A, int method, long long nnz) {
        hipStream_t s = ctx->stream;
        if (dmat_launch(s, A, method, x, y)) {
            clear_error();
            return 0.0;
        }
        if (hipEventRecord(ctx->ev_start, s) != hipSuccess) return 0.0;
        for (int i = 0; i < reps; ++i)
            if (dmat_launch(s, A, method, x, y)) return 0.0;
        if (hipEventRecord(ctx->ev_stop, s) != hipSuccess) return 0.0;
        if (hipEventSynchronize(ctx->ev_stop) != hipSuccess) return 0.0;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop) != hipSuccess || ms <= 0) return 0.0;
        return 2.0 * (double)nnz / ((double)ms / reps * 1e-3) / 1e9;
    }
};

}  // namespace

extern "C" int clspmv_autotune(int device, const char* out_dir, int level, int verbose) {
    if (!out_dir) return fail(CLSPMV_ERR_INVALID, "null output directory");
    clspmv_ctx* ctx = nullptr;
    int st = clspmv_ctx_create(device, &ctx);
    if (st) return st;
    hipStream_t s = ctx->stream;
    const int kMaxDim = 2 * 1024 * 1024;
    float *x = nullptr, *y = nullptr;
    if (hipMalloc(&x, sizeof(float) * (kMaxDim + 64)) != hipSuccess ||
        hipMalloc(&y, sizeof(float) * (kMaxDim + 64)) != hipSuccess) {
        clspmv_ctx_destroy(ctx);
        return fail(CLSPMV_ERR_ALLOC, "autotune vectors");
    }
    fillf(s, x, kMaxDim + 64, 1.0f);
    fillf(s, y, kMaxDim + 64, 0.0f);
    // level -1: a reduced grid (three dimensions, 1M entries per probe, 3
    // launches) for tests of the tables' form.
    const bool quick = level < 0;
    Timer T{ctx, x, y, level > 0 ? 20 : quick ? 3 : 8};
    const long long cap = level > 0 ? (256LL << 20) : quick ? (1LL << 20) : (64LL << 20);  // stored entries per probe
    std::vector<int> dims;
    if (quick)
        dims = {1024, 8192, 65536};
    else
        for (int d = 1024; d <= kMaxDim; d *= 2) dims.push_back(d);
    ckt::Tables t;
    auto note = [&](const char* what, int a, int b, const double* gf, int n) {
        if (!verbose) return;
        std::printf("%s %d %d", what, a, b);
        for (int i = 0; i < n; ++i) std::printf(" %.3f", gf[i]);
        std::printf("\n");
        std::fflush(stdout);
    };
    auto done = [&](clspmv_dmat* A) { clspmv_dmat_free(A); };

    // overhead.ben: back-to-back launches of an empty kernel.
    for (int g = 256; g <= 256 * 1024; g *= 2)
        for (int th : {64, 128, 256}) {
            hipLaunchKernelGGL(empty_kernel, dim3(g), dim3(th), 0, s);
            (void)hipEventRecord(ctx->ev_start, s);
            const int n = 200;
            for (int i = 0; i < n; ++i) hipLaunchKernelGGL(empty_kernel, dim3(g), dim3(th), 0, s);
            (void)hipEventRecord(ctx->ev_stop, s);
            (void)hipEventSynchronize(ctx->ev_stop);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
            ckt::OverheadRow r{(unsigned)g, (unsigned)th, (double)ms / n * 1000.0};
            t.overhead.push_back(r);
            note("overhead", g, th, &r.us, 1);
        }

    // bdia.ben and dia.ben on one grid.
    const int bw_list[] = {1, 3, 5, 7, 9, 11, 13, 15, 19, 23, 27, 31, 35, 39, 43, 47, 51, 55, 59, 63};
    for (int d : dims)
        for (int w : bw_list) {
            ckt::WidthRow<CLSPMV_BDIA_IMP_NUM> rb{(unsigned)d, (unsigned)w, {0}};
            ckt::WidthRow<CLSPMV_DIA_IMP_NUM> rd{(unsigned)d, (unsigned)w, {0}};
            if ((long long)d * w <= cap) {
                clspmv_dmat* A = nullptr;
                if ((st = make_dia(s, d, w, true, &A))) goto out;
                for (int m = 0; m < CLSPMV_BDIA_IMP_NUM; ++m) rb.gf[m] = T.gflops(A, m, A->info.nnz);
                done(A);
                if ((st = make_dia(s, d, w, false, &A))) goto out;
                for (int m = 0; m < CLSPMV_DIA_IMP_NUM; ++m) rd.gf[m] = T.gflops(A, m, A->info.nnz);
                done(A);
            }
            t.bdia.push_back(rb);
            t.dia.push_back(rd);
            note("bdia", d, w, rb.gf, CLSPMV_BDIA_IMP_NUM);
            note("dia", d, w, rd.gf, CLSPMV_DIA_IMP_NUM);
        }

    // sell.ben and ell.ben
    for (int d : dims)
        for (int w : {1, 2, 4, 8, 16, 32, 64, 128}) {
            if ((long long)d * w > cap) break;
            ckt::WidthRow<CLSPMV_SELL_IMP_NUM> rs{(unsigned)d, (unsigned)w, {0}};
            ckt::WidthRow<CLSPMV_ELL_IMP_NUM> re{(unsigned)d, (unsigned)w, {0}};
            const long long nnz = (long long)d * w;
            clspmv_dmat* A = nullptr;
            for (int m = 0; m < CLSPMV_SELL_IMP_NUM; ++m) {
                if ((st = make_flat(s, CLSPMV_FMT_SELL, d, w, m == 0 ? CLSPMV_WARPSIZE : CLSPMV_SELL_GROUP_SIZE, &A))) goto out;
                rs.gf[m] = T.gflops(A, m, nnz);
                done(A);
            }
            if ((st = make_flat(s, CLSPMV_FMT_ELL, d, w, 0, &A))) goto out;
            for (int m = 0; m < CLSPMV_ELL_IMP_NUM; ++m) re.gf[m] = T.gflops(A, m, nnz);
            done(A);
            t.sell.push_back(rs);
            t.ell.push_back(re);
            note("sell", d, w, rs.gf, CLSPMV_SELL_IMP_NUM);
            note("ell", d, w, re.gf, CLSPMV_ELL_IMP_NUM);
        }

    // csr.ben and coo.ben (one work-group count per row: the count the
    // partitioner's load-balance model assumes for these kernels).
    {
    // The x-panel rate per width at the largest dimension measured, for the
    // probes too large to encode on the host.
    std::vector<std::pair<int, double>> panel_csr, panel_coo;
    auto carried = [](const std::vector<std::pair<int, double>>& v, int w) {
        double g = 0.0;
        for (const auto& p : v)
            if (p.first == w) g = p.second;
        return g;
    };
    for (int d : dims)
        for (int w : {2, 8, 32, 128, 512, 2048}) {
            if ((long long)d * w > cap) break;
            const long long nnz = (long long)d * w;
            const unsigned groups = (unsigned)std::max(1, std::min(1024, d / 256));
            ckt::GroupRow<CLSPMV_CSR_IMP_NUM> rc{(unsigned)d, (unsigned)w, groups, {0}};
            ckt::GroupRow<CLSPMV_COO_IMP_NUM> ro{(unsigned)d, (unsigned)w, groups, {0}};
            clspmv_dmat* A = nullptr;
            bool panel = false;
            if ((st = make_flat(s, CLSPMV_FMT_CSR, d, w, 0, &A))) goto out;
            rc.gf[0] = T.gflops_lanes(A, 0, nnz, lane_sweep(A->lanes));
            for (int m = 1; m < CLSPMV_CSR_IMP_NUM; ++m) rc.gf[m] = T.gflops(A, m, nnz);
            double gp = carried(panel_csr, w);
            if (nnz <= kPanelProbeMax) {
                if ((st = add_panel_probe(A, d, w, 3, &panel))) {
                    done(A);
                    goto out;
                }
                if (panel) {
                    A->use_panel = true;
                    gp = T.gflops(A, 0, nnz);
                    panel_csr.push_back({w, gp});
                }
            }
            for (int m = 0; m < CLSPMV_CSR_IMP_NUM; ++m) rc.gf[m] = std::max(rc.gf[m], gp);
            done(A);
            if ((st = make_flat(s, CLSPMV_FMT_COO, d, w, 0, &A))) goto out;
            ro.gf[0] = T.gflops(A, 0, nnz);
            gp = carried(panel_coo, w);
            if (nnz <= kPanelProbeMax) {
                if ((st = add_panel_probe(A, d, w, 4, &panel))) {
                    done(A);
                    goto out;
                }
                if (panel) {
                    A->use_panel = true;
                    gp = T.gflops(A, 0, nnz);
                    panel_coo.push_back({w, gp});
                }
            }
            ro.gf[0] = std::max(ro.gf[0], gp);
            done(A);
            t.csr.push_back(rc);
            t.coo.push_back(ro);
            note("csr", d, w, rc.gf, CLSPMV_CSR_IMP_NUM);
            note("coo", d, w, ro.gf, CLSPMV_COO_IMP_NUM);
        }
    }

    // sbell.ben, bell.ben, bcsr.ben: shapes in the partitioner's order.
    {
        const int BH[8] = {1, 2, 4, 8, 1, 2, 4, 8}, BW[8] = {4, 4, 4, 4, 8, 8, 8, 8};
        for (int sh = 0; sh < 8; ++sh)
            for (int d : dims)
                for (int nb : {1, 2, 4, 8, 16, 32}) {
                    const long long stored = (long long)d * nb * BW[sh];
                    if (stored > cap) break;
                    ckt::BlockRow<CLSPMV_SBELL_IMP_NUM> r1{(unsigned)BH[sh], (unsigned)BW[sh], (unsigned)d, (unsigned)nb, {0}};
                    ckt::BlockRow<CLSPMV_BELL_IMP_NUM> r2{(unsigned)BH[sh], (unsigned)BW[sh], (unsigned)d, (unsigned)nb, {0}};
                    ckt::BlockRow<CLSPMV_BCSR_IMP_NUM> r3{(unsigned)BH[sh], (unsigned)BW[sh], (unsigned)d, (unsigned)nb, {0}};
                    clspmv_dmat* A = nullptr;
                    if ((st = make_block(s, CLSPMV_FMT_SBELL, d, BH[sh], BW[sh], nb, &A))) goto out;
                    r1.gf[0] = T.gflops(A, 0, A->info.nnz);
                    done(A);
                    if ((st = make_block(s, CLSPMV_FMT_BELL, d, BH[sh], BW[sh], nb, &A))) goto out;
                    for (int m = 0; m < CLSPMV_BELL_IMP_NUM; ++m) r2.gf[m] = T.gflops(A, m, A->info.nnz);
                    done(A);
                    if ((st = make_block(s, CLSPMV_FMT_BCSR, d, BH[sh], BW[sh], nb, &A))) goto out;
                    r3.gf[0] = T.gflops_lanes(A, 0, A->info.nnz, lane_sweep(A->lanes));
                    done(A);
                    t.sbell.push_back(r1);
                    t.bell.push_back(r2);
                    t.bcsr.push_back(r3);
                    note("sbell", d, nb, r1.gf, 1);
                    note("bell", d, nb, r2.gf, 2);
                    note("bcsr", d, nb, r3.gf, 1);
                }
    }
    st = ckt::write_tables(out_dir, t);
out:
    (void)hipFree(x);
    (void)hipFree(y);
    clspmv_ctx_destroy(ctx);
    return st;
}
