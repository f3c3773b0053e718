// Shared host-side helpers of the clspmv engine: error reporting and a few
// array utilities used by the format converters, the Cocktail partitioner and
// the device runtime.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "clspmv_formats.h"

namespace clspmv {

// Records a printf-style message as the calling thread's last error and
// returns `code`, so call sites read `return fail(CLSPMV_ERR_INVALID, ...)`.
int fail(int code, const char* fmt, ...);
void clear_error();

// malloc() wrapper for arrays handed to the caller (released with free()).
template <class T>
T* host_alloc(size_t n) {
    // malloc(0) may legally return NULL; hand out one element so callers can
    // tell "empty array" from "not allocated".
    return static_cast<T*>(std::malloc(sizeof(T) * (n ? n : 1)));
}

template <class T>
T* host_alloc_zero(size_t n) {
    return static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
}

template <class T>
T* host_copy(const std::vector<T>& v) {
    T* p = host_alloc<T>(v.size());
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

inline int round_up(int length, int alignment) {
    // aligned_length (include/matrix_storage.h:818): smallest multiple of
    // `alignment` that is >= length.
    int r = length % alignment;
    return r == 0 ? length : length + alignment - r;
}

inline long long round_up_ll(long long length, long long alignment) {
    long long r = length % alignment;
    return r == 0 ? length : length + alignment - r;
}

// Row pointer of a COO matrix whose entries are ordered by row: rp[r] is the
// index of the first entry of row r, rp[height] == nnz.  Rows without
// entries repeat the next row's start (the reference's construction at
// include/matrix_storage.h:1290-1306 gives the same array for sorted input).
void row_pointer(const int* rows, int nnz, int height, std::vector<int>& rp);

// True when (row, col) pairs are non-decreasing.
bool coo_sorted(const clspmv_coo_matrix& m);

// Sort entries by (row, col), keeping the input order of equal keys.
void coo_sort(clspmv_coo_matrix& m);

// The reference's correctness report of a result against the serial one
// (two_vec_compare, src/util.cpp:93-131, with distance :74): L2 norm of the
// differences (each taken in float, then widened), the largest difference
// of magnitudes, and the largest such difference relative to the larger
// magnitude.
struct VecCompare {
    double distance = 0.0, max_diff = 0.0, max_ratio = 0.0;
};
VecCompare two_vec_compare(const float* coovec, const float* newvec, int n);

}  // namespace clspmv
