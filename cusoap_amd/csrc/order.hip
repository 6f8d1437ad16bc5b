// Space-filling-curve order of the centers (no reference counterpart): the adjoint edge
// pass takes groups of consecutive centers in this order, so that a group's
// neighbourhoods overlap (its coefficient rows, radial table rows and force sums per
// neighbour atom are shared) whatever order the caller gives the atoms in.
//
// Morton (Z-order) codes of the center coordinates, kMortonBits per axis within their
// bounding box (32 cells per axis), sorted with their indices by a stable radix sort
// (rocPRIM): the order is a function of the coordinates alone, the same on every run,
// and within a cell it keeps the caller's order.
#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kBoxThreads = 256;
constexpr int kMortonBits = 5;              // per axis: the sort runs over 15 bits
constexpr int kBoxPer = 4096;               // centers per workgroup of the first pass

// min and max of each coordinate: partial boxes of kBoxPer centers (pass 1), then one
// workgroup over the partials (pass 2)
__device__ __forceinline__ void box_reduce(double (&v)[6], double* __restrict__ out) {
    __shared__ double red[6][kBoxThreads / 64];
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], __shfl_xor(v[d], o, 64));
            v[3 + d] = fmax(v[3 + d], __shfl_xor(v[3 + d], o, 64));
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 6; ++d) red[d][w] = v[d];
    __syncthreads();
    if (threadIdx.x < 6) {
        double r = red[threadIdx.x][0];
        for (int k = 1; k < kBoxThreads / 64; ++k) r = threadIdx.x < 3 ? fmin(r, red[threadIdx.x][k]) : fmax(r, red[threadIdx.x][k]);
        out[threadIdx.x] = r;
    }
}

__global__ __launch_bounds__(kBoxThreads) void box_partial_kernel(const double* __restrict__ centers, int n,
                                                                 double* __restrict__ part) {
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    const int i0 = blockIdx.x * kBoxPer, i1 = min(n, i0 + kBoxPer);
    for (int i = i0 + threadIdx.x; i < i1; i += kBoxThreads)
        for (int d = 0; d < 3; ++d) {
            const double x = centers[3 * (size_t)i + d];
            v[d] = fmin(v[d], x);
            v[3 + d] = fmax(v[3 + d], x);
        }
    box_reduce(v, part + 6 * (size_t)blockIdx.x);
}

__global__ __launch_bounds__(kBoxThreads) void box_final_kernel(const double* __restrict__ part, int n_part,
                                                               double* __restrict__ box) {
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < n_part; i += kBoxThreads)
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], part[6 * (size_t)i + d]);
            v[3 + d] = fmax(v[3 + d], part[6 * (size_t)i + 3 + d]);
        }
    box_reduce(v, box);
}

// 10 bits of x spread to every third bit
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void morton_kernel(const double* __restrict__ centers, int n, const double* __restrict__ box,
                              uint32_t* __restrict__ code, int32_t* __restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int d = 0; d < 3; ++d) {
        const double lo = box[d], ext = box[3 + d] - lo;
        const double t = ext > 0.0 ? (centers[3 * (size_t)i + d] - lo) / ext : 0.0;   // in [0, 1]
        constexpr int kMax = (1 << kMortonBits) - 1;
        const int k = (int)(t * (kMax + 1));
        q[d] = (uint32_t)(k < 0 ? 0 : (k > kMax ? kMax : k));
    }
    code[i] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    idx[i] = i;
}

int box_parts(int n) { return (std::max(n, 1) + kBoxPer - 1) / kBoxPer; }

struct OrderWs {
    double* part;
    double* box;
    uint32_t* code_in;
    uint32_t* code_out;
    int32_t* idx;
    void* tmp;
    size_t tmp_bytes;
};

size_t sort_tmp_bytes(int n) {
    size_t bytes = 0;
    // a size query (no launch); a failure leaves 0, and the sort then reports its error
    if (rocprim::radix_sort_pairs((void*)nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (const int32_t*)nullptr, (int32_t*)nullptr, (unsigned)std::max(n, 1), 0,
                                  3 * kMortonBits) !=
        hipSuccess)
        bytes = 0;
    return bytes;
}

OrderWs carve_order(void* base, int n) {
    char* p = (char*)base;
    OrderWs w;
    auto take = [&](size_t bytes) {
        void* r = p;
        p += align_up(bytes, 256);
        return r;
    };
    w.part = (double*)take(sizeof(double) * 6 * (size_t)box_parts(n));
    w.box = (double*)take(sizeof(double) * 6);
    w.code_in = (uint32_t*)take(sizeof(uint32_t) * (size_t)n);
    w.code_out = (uint32_t*)take(sizeof(uint32_t) * (size_t)n);
    w.idx = (int32_t*)take(sizeof(int32_t) * (size_t)n);
    w.tmp_bytes = sort_tmp_bytes(n);
    w.tmp = take(w.tmp_bytes);
    return w;
}

}  // namespace

// bytes of center_order's workspace for n centers (the order itself not included)
size_t center_order_ws_bytes(int n) {
    const size_t m = (size_t)std::max(n, 1);
    return align_up(sizeof(double) * 6 * (size_t)box_parts(n), 256) + align_up(sizeof(double) * 6, 256) +
           2 * align_up(sizeof(uint32_t) * m, 256) +
           align_up(sizeof(int32_t) * m, 256) + align_up(sort_tmp_bytes(n), 256);
}

// order (n) int32: the centers' indices along the Morton curve of their coordinates
int center_order(const double* centers, int n, void* ws, int32_t* order, hipStream_t st) {
    if (n <= 0) return CUSOAP_OK;
    OrderWs w = carve_order(ws, n);
    prof_begin("center_order", st);
    box_partial_kernel<<<box_parts(n), kBoxThreads, 0, st>>>(centers, n, w.part);
    CUSOAP_LAUNCH_CHECK();
    box_final_kernel<<<1, kBoxThreads, 0, st>>>(w.part, box_parts(n), w.box);
    CUSOAP_LAUNCH_CHECK();
    morton_kernel<<<(n + 255) / 256, 256, 0, st>>>(centers, n, w.box, w.code_in, w.idx);
    CUSOAP_LAUNCH_CHECK();
    size_t tmp_bytes = w.tmp_bytes;
    const hipError_t e = rocprim::radix_sort_pairs(w.tmp, tmp_bytes, w.code_in, w.code_out, w.idx, order,
                                                   (unsigned)n, 0, 3 * kMortonBits, st);
    if (e != hipSuccess) return fail(CUSOAP_ERR_HIP, std::string("center order: radix sort: ") + hipGetErrorString(e));
    prof_end("center_order", st);
    return CUSOAP_OK;
}

}  // namespace cusoap
