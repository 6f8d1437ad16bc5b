// Cell-list neighbour search with periodic images (replaces the reference's
// chunked GEMM radius search, soap_generator.py:1042-1181).
//
// Pipeline (all on the caller's stream):
//   plan   : bounding box of the centers -> uniform grid of cells >= cutoff
//   count  : bin every image atom (ext = img * N + atom) that falls in the grid
//            (counting sort), then one wavefront per center scans the 27
//            surrounding cells, counts edges with a ballot and, up to 384 edges,
//            stores them in the workspace sorted by ext index (the order of the
//            reference's `nonzero` over (center, ext)) with a bitonic network in
//            the wavefront's registers
//   scan   : exclusive scan of the counts -> CSR offsets
//   fill   : a copy of the sorted rows to the CSR ranges; a center with more
//            than 384 edges is searched again by a workgroup of its own, which
//            places its edges by rank
// The distance test restates the reference formula
//   d^2 = ((-2 c.n) + |c|^2) + |n|^2 <= cutoff^2
// in float64 with every operation rounded separately (no fused multiply-add), so
// the edge set is the one that formula defines.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include <type_traits>

#include "common.hpp"

namespace cusoap {

namespace {

// Each product and sum is rounded on its own: contraction into fused multiply-adds
// is switched off here (the HIP default would fuse them, which changes the edge
// set at the cutoff boundary).
__device__ __forceinline__ double sqnorm_rn(double x, double y, double z) {
#pragma clang fp contract(off)
    return (x * x + y * y) + z * z;
}

__device__ __forceinline__ double dist2_formula(double cx, double cy, double cz, double cn,
                                                double nx, double ny, double nz, double nn) {
#pragma clang fp contract(off)
    const double dot = (cx * nx + cy * ny) + cz * nz;
    return (dot * -2.0 + cn) + nn;
}

struct Grid {
    double ox, oy, oz, hx, hy, hz;
    int nx, ny, nz;
};

__device__ __forceinline__ Grid make_grid(const cusoap_nl_plan_t& p) {
    Grid g;
    g.ox = p.origin[0]; g.oy = p.origin[1]; g.oz = p.origin[2];
    g.hx = p.cell[0];   g.hy = p.cell[1];   g.hz = p.cell[2];
    g.nx = p.dims[0];   g.ny = p.dims[1];   g.nz = p.dims[2];
    return g;
}

__device__ __forceinline__ int cell_coord(double v, double o, double h, int n) {
    const double f = floor((v - o) / h);
    if (!(f >= 0.0) || f >= (double)n) return -1;
    return (int)f;
}

// --- workspace layout -------------------------------------------------------
struct NLWs {
    int32_t* ext_cell;     // (n_ext)
    int32_t* cell_count;   // (n_cells + 1)
    int32_t* cursor;       // (n_cells + 1) fill cursors of the scatter pass
    int64_t* cell_start;   // (n_cells + 1)
    double4* binned;       // (n_ext)  x, y, z, |p|^2
    int32_t* binned_ext;   // (n_ext)
    int32_t* counts;       // (n_centers)
    int32_t* sorted;       // (n_centers, kSortCap) sorted edges of centers with <= kSortCap edges
    int64_t* partials;     // scan partials
    int32_t* max_count;    // (4): [0] the largest edge count; bytes 8..15 the edge total (int64)
    size_t bytes;
};

// the count pass also sorts the edges of every center with at most this many into the
// workspace, so that the fill pass is a copy when no center has more (the 98k-atom
// water box at r_cut 6 + 1.86 padding has at most ~240 edges per center; 1.5 KiB of
// workspace per center)
constexpr int kSortCap = 384;
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

size_t scan_partials(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

NLWs carve(void* base, int64_t n_ext, int32_t n_cells, int32_t n_centers) {
    NLWs w;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return (char*)base + o;
    };
    w.ext_cell = (int32_t*)take(sizeof(int32_t) * n_ext);
    w.cell_count = (int32_t*)take(sizeof(int32_t) * (n_cells + 1));
    w.cursor = (int32_t*)take(sizeof(int32_t) * (n_cells + 1));
    w.cell_start = (int64_t*)take(sizeof(int64_t) * (n_cells + 1));
    w.binned = (double4*)take(sizeof(double4) * n_ext);
    w.binned_ext = (int32_t*)take(sizeof(int32_t) * n_ext);
    w.counts = (int32_t*)take(sizeof(int32_t) * (n_centers + 1));
    w.sorted = (int32_t*)take(sizeof(int32_t) * (size_t)kSortCap * n_centers);
    int64_t mx = std::max<int64_t>(n_cells + 1, n_centers + 1);
    w.partials = (int64_t*)take(sizeof(int64_t) * (scan_partials(mx) + kScanTile));
    w.max_count = (int32_t*)take(sizeof(int32_t) * 4);
    w.bytes = off;
    return w;
}

// --- bounding box -------------------------------------------------------------
__global__ void bbox_partial_kernel(const double* __restrict__ c, int n, double* __restrict__ part) {
    __shared__ double red[6][256];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int d = 0; d < 3; ++d) {
            const double v = c[3 * (size_t)i + d];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    }
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = mn[d];
        red[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; ++d) part[6 * blockIdx.x + d] = red[d][0];
}

// one wavefront: lane-strided partial boxes, then a shuffle reduction (min and max are
// exact, so the result does not depend on the order)
__global__ void bbox_final_kernel(double* part, int nblk) {
    const int lane = threadIdx.x;
    double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = lane; b < nblk; b += 64)
        for (int d = 0; d < 3; ++d) {
            r[d] = fmin(r[d], part[6 * b + d]);
            r[3 + d] = fmax(r[3 + d], part[6 * b + 3 + d]);
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            r[d] = fmin(r[d], __shfl_xor(r[d], o, 64));
            r[3 + d] = fmax(r[3 + d], __shfl_xor(r[3 + d], o, 64));
        }
    __syncthreads();                  // every lane has read its partials before lane 0 overwrites them
    if (lane == 0)
        for (int d = 0; d < 6; ++d) part[d] = r[d];
}

// --- exclusive scan int32 -> int64 ---------------------------------------------
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t t = __shfl_up(v, o, 64);
        if ((int)(threadIdx.x & 63) >= o) v += t;
    }
    return v;
}

// block-wide exclusive scan of one value per thread; returns exclusive prefix, sets total
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* wsum, int64_t& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int64_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        int64_t s = lane < nw ? wsum[lane] : 0;
        s = wave_incl_scan(s);
        if (lane < nw) wsum[lane] = s;
    }
    __syncthreads();
    const int64_t base = wid > 0 ? wsum[wid - 1] : 0;
    total = wsum[nw - 1];
    __syncthreads();
    return base + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_tiles_kernel(const int32_t* __restrict__ in, int64_t n,
                                                                int64_t* __restrict__ out,
                                                                int64_t* __restrict__ partials) {
    __shared__ int64_t wsum[kScanBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? (int64_t)in[base + k] : 0;
        s += v[k];
    }
    int64_t total;
    int64_t pre = block_excl_scan(s, wsum, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_partials_kernel(int64_t* partials, int64_t nb) {
    __shared__ int64_t wsum[kScanBlock / 64];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kScanBlock) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < nb ? partials[i] : 0;
        int64_t total;
        const int64_t pre = block_excl_scan(v, wsum, total);
        if (i < nb) partials[i] = carry + pre;
        carry += total;
    }
    if (threadIdx.x == 0) partials[nb] = carry;
}

__global__ void scan_add_kernel(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ partials,
                                int64_t nb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += partials[i / kScanTile];
    if (i == 0) out[n] = partials[nb];
}

int scan_exclusive(const int32_t* in, int64_t n, int64_t* out, int64_t* partials, hipStream_t st) {
    const int64_t nb = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
    scan_tiles_kernel<<<(unsigned)nb, kScanBlock, 0, st>>>(in, n, out, partials);
    scan_partials_kernel<<<1, kScanBlock, 0, st>>>(partials, nb);
    const int64_t threads = std::max<int64_t>(n, 1);
    scan_add_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, st>>>(out, n, partials, nb);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

// --- binning --------------------------------------------------------------------
__global__ void bin_kernel(cusoap_nl_plan_t plan, const double* __restrict__ pos, int n_atoms,
                           const double* __restrict__ shifts, int64_t n_ext,
                           int32_t* __restrict__ ext_cell, int32_t* __restrict__ cell_count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_ext) return;
    const Grid g = make_grid(plan);
    const int img = (int)(i / n_atoms), a = (int)(i - (int64_t)img * n_atoms);
    const double px = pos[3 * (size_t)a + 0] + shifts[3 * img + 0];
    const double py = pos[3 * (size_t)a + 1] + shifts[3 * img + 1];
    const double pz = pos[3 * (size_t)a + 2] + shifts[3 * img + 2];
    const int ix = cell_coord(px, g.ox, g.hx, g.nx);
    const int iy = cell_coord(py, g.oy, g.hy, g.ny);
    const int iz = cell_coord(pz, g.oz, g.hz, g.nz);
    int cell = -1;
    if (ix >= 0 && iy >= 0 && iz >= 0) {
        cell = (ix * g.ny + iy) * g.nz + iz;
        atomicAdd(&cell_count[cell], 1);
    }
    ext_cell[i] = cell;
}

__global__ void scatter_kernel(const double* __restrict__ pos, int n_atoms, const double* __restrict__ shifts,
                               int64_t n_ext, const int32_t* __restrict__ ext_cell,
                               const int64_t* __restrict__ cell_start, int32_t* __restrict__ cursor,
                               double4* __restrict__ binned, int32_t* __restrict__ binned_ext) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_ext) return;
    const int cell = ext_cell[i];
    if (cell < 0) return;
    const int img = (int)(i / n_atoms), a = (int)(i - (int64_t)img * n_atoms);
    const double px = pos[3 * (size_t)a + 0] + shifts[3 * img + 0];
    const double py = pos[3 * (size_t)a + 1] + shifts[3 * img + 1];
    const double pz = pos[3 * (size_t)a + 2] + shifts[3 * img + 2];
    const int64_t slot = cell_start[cell] + atomicAdd(&cursor[cell], 1);
    binned[slot] = make_double4(px, py, pz, sqnorm_rn(px, py, pz));
    binned_ext[slot] = (int32_t)i;
}

struct CenterCells {
    int x0, x1, y0, y1, z0, z1;
};

__device__ __forceinline__ CenterCells center_cells(const Grid& g, double cx, double cy, double cz) {
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    int ix = clampi((int)floor((cx - g.ox) / g.hx), 0, g.nx - 1);
    int iy = clampi((int)floor((cy - g.oy) / g.hy), 0, g.ny - 1);
    int iz = clampi((int)floor((cz - g.oz) / g.hz), 0, g.nz - 1);
    CenterCells r;
    r.x0 = ix > 0 ? ix - 1 : 0; r.x1 = ix < g.nx - 1 ? ix + 1 : g.nx - 1;
    r.y0 = iy > 0 ? iy - 1 : 0; r.y1 = iy < g.ny - 1 ? iy + 1 : g.ny - 1;
    r.z0 = iz > 0 ? iz - 1 : 0; r.z1 = iz < g.nz - 1 ? iz + 1 : g.nz - 1;
    return r;
}

__device__ __forceinline__ bool edge_test(double cx, double cy, double cz, double cn, const double4& p,
                                          double cutoff2, int drop_self) {
    const double d2 = dist2_formula(cx, cy, cz, cn, p.x, p.y, p.z, p.w);
    if (!(d2 <= cutoff2)) return false;
    if (drop_self) {
        const double rx = p.x - cx, ry = p.y - cy, rz = p.z - cz;
        if (!(sqrt(sqnorm_rn(rx, ry, rz)) > 1e-12)) return false;
    }
    return true;
}

// The candidate cells of a center: up to 3 x 3 rows of cells that are
// consecutive in z, hence contiguous ranges binned[start[r], start[r] + len[r])
// of the binned arrays.  Every field is uniform across the wavefront.
struct CandRows {
    int64_t start[9];
    int len[9];                 // candidates in each cell row (0 for unused rows)
};

__device__ __forceinline__ CandRows cand_rows(const Grid& g, const CenterCells& cc,
                                              const int64_t* __restrict__ cell_start) {
    CandRows cr;
    int r = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        cr.start[i] = 0;
        cr.len[i] = 0;
    }
    for (int ix = cc.x0; ix <= cc.x1; ++ix)
        for (int iy = cc.y0; iy <= cc.y1; ++iy) {
            const int c0 = (ix * g.ny + iy) * g.nz + cc.z0, c1 = (ix * g.ny + iy) * g.nz + cc.z1 + 1;
            const int64_t s0 = cell_start[c0], s1 = cell_start[c1];
#pragma unroll
            for (int i = 0; i < 9; ++i)
                if (i == r) {
                    cr.start[i] = s0;
                    cr.len[i] = (int)(s1 - s0);
                }
            ++r;
        }
    return cr;
}

// zero the cell counts, the scatter cursors and the count summary (one launch)
__global__ void nl_zero_kernel(int32_t* __restrict__ cell_count, int32_t* __restrict__ cursor, int n_cells1,
                               int32_t* __restrict__ summary) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_cells1; i += gridDim.x * blockDim.x) {
        cell_count[i] = 0;
        cursor[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) summary[threadIdx.x] = 0;
}

// the largest per-center edge count, one atomic per workgroup (a per-center atomic on one
// address serialises, ~1 ms per 10^5 centers), and (block 0) the edge total offsets[n]
// beside it, so that one 16-byte copy brings both to the host
__global__ __launch_bounds__(256) void max_count_kernel(const int32_t* __restrict__ counts, int n,
                                                        const int64_t* __restrict__ offsets,
                                                        int32_t* __restrict__ max_count) {
    __shared__ int red[4];
    if (blockIdx.x == 0 && threadIdx.x == 0) ((int64_t*)max_count)[1] = offsets[n];
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, counts[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(max_count, max(max(red[0], red[1]), max(red[2], red[3])));
}

template <int CAP>
__global__ __launch_bounds__(256) void fill_kernel(cusoap_nl_plan_t plan, const double* __restrict__ centers,
                                                   int n_centers, double cutoff2, int drop_self,
                                                   const int64_t* __restrict__ cell_start,
                                                   const double4* __restrict__ binned,
                                                   const int32_t* __restrict__ binned_ext,
                                                   const int64_t* __restrict__ offsets,
                                                   const int32_t* __restrict__ sorted,
                                                   int32_t* __restrict__ nbr) {
    __shared__ __align__(16) int32_t keys[CAP + 4];
    __shared__ int nkeys;
    const int c = blockIdx.x;
    if (c >= n_centers) return;
    {
        // a center with at most kSortCap edges: the row the count pass sorted
        const int64_t b0 = offsets[c];
        const int cnt = (int)(offsets[c + 1] - b0);
        if (cnt <= kSortCap) {
            const int32_t* src = sorted + (size_t)c * kSortCap;
            for (int i = threadIdx.x; i < cnt; i += blockDim.x) nbr[b0 + i] = src[i];
            return;
        }
    }
    if (threadIdx.x == 0) nkeys = 0;
    __syncthreads();
    const Grid g = make_grid(plan);
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const double cn = sqnorm_rn(cx, cy, cz);
    const CandRows cr = cand_rows(g, center_cells(g, cx, cy, cz), cell_start);
    // wavefront q takes cell rows q, q + 4, q + 8 (uniform bounds, no per-candidate row
    // search); the keys are put in order below
    const int wq = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if ((i & 3) != wq) continue;
        const int64_t s0 = cr.start[i];
        for (int k = lane; k < cr.len[i]; k += 64) {
            if (edge_test(cx, cy, cz, cn, binned[s0 + k], cutoff2, drop_self)) {
                const int slot = atomicAdd(&nkeys, 1);
                if (slot < CAP) keys[slot] = binned_ext[s0 + k];
            }
        }
    }
    __syncthreads();
    // ascending order (the reference's nonzero order) by rank: the keys of one center
    // are distinct extended indices, so a key's position is the number of smaller keys;
    // four keys per 16-byte LDS read (the same address in every lane), no barrier
    const int n = nkeys < CAP ? nkeys : CAP;
    const int n4 = (n + 3) & ~3;
    for (int i = n + threadIdx.x; i < n4; i += blockDim.x) keys[i] = INT_MAX;
    __syncthreads();
    const int64_t base = offsets[c];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int k = keys[i];
        int rank = 0;
        for (int j = 0; j < n4; j += 4) {
            const int4 q = *(const int4*)(keys + j);
            rank += (q.x < k) + (q.y < k) + (q.z < k) + (q.w < k);
        }
        nbr[base + rank] = k;
    }
}

// Ascending bitonic sort of the 64 KPL values v[t] (element e = 64 t + lane) of one
// wavefront.  Stage (k, j) pairs e with e ^ j; the lower element keeps the minimum when
// e & k == 0.  Strides below 64 exchange across lanes (__shfl_xor), larger ones between
// a lane's own registers.
template <int KPL>
__device__ __forceinline__ void bitonic_wave(int (&v)[KPL], int lane) {
    constexpr int N = 64 * KPL;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int tj = j >> 6;
#pragma unroll
                for (int t = 0; t < KPL; ++t) {
                    if (t & tj) continue;              // t is the lower element of (t, t | tj)
                    const bool asc = ((64 * t + lane) & k) == 0;
                    const int a = v[t], b = v[t | tj];
                    const int lo = a < b ? a : b, hi = a < b ? b : a;
                    v[t] = asc ? lo : hi;
                    v[t | tj] = asc ? hi : lo;
                }
            } else {
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int t = 0; t < KPL; ++t) {
                    const bool asc = ((64 * t + lane) & k) == 0;
                    const int o = __shfl_xor(v[t], j, 64);
                    const int mn = v[t] < o ? v[t] : o, mx = v[t] < o ? o : v[t];
                    v[t] = (lower == asc) ? mn : mx;
                }
            }
        }
    }
}

// Count pass, one wavefront per center (4 centers per workgroup): counts the
// center's edges and, when there are at most CAP, also puts them in order (the list
// fill_kernel makes) in the center's row of `sorted`.  The passing candidates are
// compacted with ballot + mbcnt (no LDS atomics), and the rank placement runs on the
// wavefront's own keys with no workgroup barrier.
template <int CAP>
__global__ __launch_bounds__(256) void count_sort_kernel(cusoap_nl_plan_t plan, const double* __restrict__ centers,
                                                         int n_centers, double cutoff2, int drop_self,
                                                         const int64_t* __restrict__ cell_start,
                                                         const double4* __restrict__ binned,
                                                         const int32_t* __restrict__ binned_ext,
                                                         int32_t* __restrict__ counts, int32_t* __restrict__ sorted) {
    __shared__ __align__(16) int32_t keys_all[4][CAP + 4];
    const int wq = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + wq;
    if (c >= n_centers) return;     // whole wavefront; no workgroup barrier below
    int32_t* keys = keys_all[wq];
    const Grid g = make_grid(plan);
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const double cn = sqnorm_rn(cx, cy, cz);
    const CandRows cr = cand_rows(g, center_cells(g, cx, cy, cz), cell_start);
    int n = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int64_t s0 = cr.start[i];
        const int len = cr.len[i];
        for (int k0 = 0; k0 < len; k0 += 64) {
            const int k = k0 + lane;
            const bool pass = k < len && edge_test(cx, cy, cz, cn, binned[s0 + k], cutoff2, drop_self);
            const uint64_t m = __ballot(pass);
            const int slot = n + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (pass && slot < CAP) keys[slot] = binned_ext[s0 + k];
            n += __popcll(m);
        }
    }
    if (lane == 0) counts[c] = n;
    if (n > CAP) return;            // the fill pass searches this center again
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int32_t* dst = sorted + (size_t)c * CAP;
    // the keys i = lane + 64 t, padded with INT_MAX to 64 KPL (a power of two), sorted by
    // a bitonic network in registers; the row is then stored in element order
    auto place = [&](auto kpl_c) {
        constexpr int KPL = decltype(kpl_c)::value;
        int kk[KPL];
#pragma unroll
        for (int t = 0; t < KPL; ++t) {
            const int i = lane + 64 * t;
            kk[t] = i < n ? keys[i] : INT_MAX;
        }
        bitonic_wave<KPL>(kk, lane);
#pragma unroll
        for (int t = 0; t < KPL; ++t)
            if (lane + 64 * t < n) dst[lane + 64 * t] = kk[t];
    };
    static_assert(CAP <= 8 * 64, "count_sort_kernel: at most 8 keys per lane");
    const int nt = (n + 63) >> 6;
    if (nt == 1) place(std::integral_constant<int, 1>());
    else if (nt == 2) place(std::integral_constant<int, 2>());
    else if (nt <= 4) place(std::integral_constant<int, 4>());
    else if (nt > 4) place(std::integral_constant<int, 8>());
}

// Fill pass when every center has at most kSortCap edges: one wavefront per center
// copies its sorted row to its CSR range.
__global__ __launch_bounds__(256) void copy_sorted_kernel(int n_centers, const int32_t* __restrict__ sorted,
                                                          const int64_t* __restrict__ offsets,
                                                          int32_t* __restrict__ nbr) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= n_centers) return;
    const int64_t b = offsets[c];
    const int cnt = (int)(offsets[c + 1] - b);
    const int32_t* src = sorted + (size_t)c * kSortCap;
    for (int i = lane; i < cnt; i += 64) nbr[b + i] = src[i];
}

// Pinned host landing space for the plan's bounding box and the count's summary (64 bytes
// per host thread, allocated on first use and kept until the process ends): a device-to-host
// copy into pageable memory is staged through a runtime buffer, into pinned memory it is
// one transfer that the following synchronisation waits for
int host_landing(double** out) {
    static thread_local double* pinned = nullptr;
    if (!pinned) {
        void* p = nullptr;
        const hipError_t e = hipHostMalloc(&p, 64, hipHostMallocPortable);
        if (e != hipSuccess) return fail(CUSOAP_ERR_HIP, std::string("nl: pinned host buffer: ") + hipGetErrorString(e));
        pinned = (double*)p;
    }
    *out = pinned;
    return CUSOAP_OK;
}

}  // namespace

int nl_plan_impl(const double* centers, int32_t n_centers, int32_t n_atoms, int32_t n_img, double cutoff,
                 void* tmp, cusoap_nl_plan_t* plan, hipStream_t st) {
    if (!plan || !tmp || n_centers < 0 || n_atoms < 0 || n_img < 1 || !(cutoff > 0))
        return fail(CUSOAP_ERR_ARG, "nl_plan: invalid argument");
    const int64_t n_ext = (int64_t)n_img * n_atoms;
    if (n_ext >= INT_MAX) return fail(CUSOAP_ERR_UNSUPPORTED, "nl_plan: n_images * n_atoms must be < 2^31");
    double bb[6] = {0, 0, 0, 0, 0, 0};
    if (n_centers > 0) {
        double* land = nullptr;
        if (int rc = host_landing(&land)) return rc;
        const int nblk = std::min(256, std::max(1, (n_centers + 255) / 256));
        bbox_partial_kernel<<<nblk, 256, 0, st>>>(centers, n_centers, (double*)tmp);
        bbox_final_kernel<<<1, 64, 0, st>>>((double*)tmp, nblk);
        CUSOAP_LAUNCH_CHECK();
        CUSOAP_HIP(hipMemcpyAsync(land, tmp, sizeof(bb), hipMemcpyDeviceToHost, st));
        CUSOAP_HIP(hipStreamSynchronize(st));
        std::copy(land, land + 6, bb);
    }
    const double margin = cutoff * (1.0 + 1e-9) + 1e-6;
    int64_t dims[3];
    double lo[3], ext[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = bb[d] - margin;
        ext[d] = (bb[3 + d] - bb[d]) + 2.0 * margin;
        if (!std::isfinite(lo[d]) || !std::isfinite(ext[d]))
            return fail(CUSOAP_ERR_ARG, "nl_plan: non-finite center coordinates");
        dims[d] = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor(ext[d] / margin), 1 << 12));
    }
    const int64_t cap = std::max<int64_t>(64, std::min<int64_t>(1 << 24, 2 * std::max<int64_t>(n_ext, 1)));
    while (dims[0] * dims[1] * dims[2] > cap) {
        int d = 0;
        if (dims[1] > dims[d]) d = 1;
        if (dims[2] > dims[d]) d = 2;
        dims[d] = (dims[d] + 1) / 2;
    }
    for (int d = 0; d < 3; ++d) {
        plan->origin[d] = lo[d];
        plan->dims[d] = (int32_t)dims[d];
        plan->cell[d] = ext[d] / (double)dims[d];
    }
    plan->n_cells = (int32_t)(dims[0] * dims[1] * dims[2]);
    plan->n_ext = n_ext;
    plan->ws_bytes = carve(nullptr, n_ext, plan->n_cells, n_centers).bytes;
    return CUSOAP_OK;
}

int nl_count_impl(const cusoap_nl_plan_t* plan, const double* pos, int32_t n_atoms, const double* shifts,
                  int32_t n_img, const double* centers, int32_t n_centers, double cutoff, int32_t drop_self,
                  void* ws, size_t ws_bytes, int64_t* offsets, int64_t* total_host, int32_t* max_host,
                  hipStream_t st) {
    if (!plan || (int64_t)n_img * n_atoms != plan->n_ext)
        return fail(CUSOAP_ERR_ARG, "nl_count: plan does not match n_atoms * n_img");
    NLWs w = carve(ws, plan->n_ext, plan->n_cells, n_centers);
    if (ws_bytes < w.bytes) return fail(CUSOAP_ERR_WORKSPACE, "nl_count: workspace too small");
    nl_zero_kernel<<<std::max(1, std::min(1024, (plan->n_cells + 256) / 256)), 256, 0, st>>>(
        w.cell_count, w.cursor, plan->n_cells + 1, w.max_count);
    CUSOAP_LAUNCH_CHECK();
    const int64_t n_ext = plan->n_ext;
    if (n_ext > 0) {
        bin_kernel<<<(unsigned)((n_ext + 255) / 256), 256, 0, st>>>(*plan, pos, n_atoms, shifts, n_ext, w.ext_cell,
                                                                    w.cell_count);
        CUSOAP_LAUNCH_CHECK();
    }
    int rc = scan_exclusive(w.cell_count, plan->n_cells, w.cell_start, w.partials, st);
    if (rc) return rc;
    if (n_ext > 0) {
        scatter_kernel<<<(unsigned)((n_ext + 255) / 256), 256, 0, st>>>(pos, n_atoms, shifts, n_ext, w.ext_cell,
                                                                        w.cell_start, w.cursor, w.binned,
                                                                        w.binned_ext);
        CUSOAP_LAUNCH_CHECK();
    }
    const double cutoff2 = cutoff * cutoff;
    if (n_centers > 0) {
        prof_begin("nl_count", st);
        count_sort_kernel<kSortCap><<<(n_centers + 3) / 4, 256, 0, st>>>(*plan, centers, n_centers, cutoff2,
                                                                          drop_self, w.cell_start, w.binned,
                                                                          w.binned_ext, w.counts, w.sorted);
        prof_end("nl_count", st);
        CUSOAP_LAUNCH_CHECK();
    }
    rc = scan_exclusive(w.counts, n_centers, offsets, w.partials, st);
    if (rc) return rc;
    if (n_centers > 0) {
        max_count_kernel<<<std::min(256, (n_centers + 255) / 256), 256, 0, st>>>(w.counts, n_centers, offsets,
                                                                                   w.max_count);
        CUSOAP_LAUNCH_CHECK();
    }
    int64_t summary[2] = {0, 0};                 // [0] low 32 bits: max count, [1] edge total
    double* land = nullptr;
    if (int rc = host_landing(&land)) return rc;
    CUSOAP_HIP(hipMemcpyAsync(land, w.max_count, sizeof(summary), hipMemcpyDeviceToHost, st));
    CUSOAP_HIP(hipStreamSynchronize(st));
    std::memcpy(summary, land, sizeof(summary));
    if (total_host) *total_host = n_centers > 0 ? summary[1] : 0;
    if (max_host) *max_host = (int32_t)(summary[0] & 0xffffffff);
    return CUSOAP_OK;
}

int nl_fill_impl(const cusoap_nl_plan_t* plan, const double* pos, int32_t n_atoms, const double* shifts,
                 int32_t n_img, const double* centers, int32_t n_centers, double cutoff, int32_t drop_self,
                 void* ws, size_t ws_bytes, const int64_t* offsets, int32_t max_count, int32_t* nbr,
                 hipStream_t st) {
    (void)pos; (void)shifts; (void)n_img; (void)n_atoms;
    NLWs w = carve(ws, plan->n_ext, plan->n_cells, n_centers);
    if (ws_bytes < w.bytes) return fail(CUSOAP_ERR_WORKSPACE, "nl_fill: workspace too small");
    if (n_centers == 0 || max_count == 0) return CUSOAP_OK;
    const double cutoff2 = cutoff * cutoff;
    // up to kSortCap edges per center the count pass has sorted them already; above
    // that the search runs again for those centers, with a whole workgroup per center
    // sharing the rank placement (each key compared with every key of its center), which one wavefront
    // does too slowly at ~450 edges ((H2O)_1000 at r_cut 10: 0.073 ms by workgroup,
    // 0.113 ms by wavefront)
    if (max_count <= kSortCap) {
        prof_begin("nl_fill", st);
        copy_sorted_kernel<<<(n_centers + 3) / 4, 256, 0, st>>>(n_centers, w.sorted, offsets, nbr);
        prof_end("nl_fill", st);
    } else if (max_count <= 1024) {
        prof_begin("nl_fill", st);
        fill_kernel<1024><<<n_centers, 256, 0, st>>>(*plan, centers, n_centers, cutoff2, drop_self, w.cell_start,
                                                     w.binned, w.binned_ext, offsets, w.sorted, nbr);
        prof_end("nl_fill", st);
    } else if (max_count <= 4096) {
        prof_begin("nl_fill", st);
        fill_kernel<4096><<<n_centers, 256, 0, st>>>(*plan, centers, n_centers, cutoff2, drop_self, w.cell_start,
                                                     w.binned, w.binned_ext, offsets, w.sorted, nbr);
        prof_end("nl_fill", st);
    } else if (max_count <= 16384) {
        prof_begin("nl_fill", st);
        fill_kernel<16384><<<n_centers, 256, 0, st>>>(*plan, centers, n_centers, cutoff2, drop_self,
                                                      w.cell_start, w.binned, w.binned_ext, offsets, w.sorted, nbr);
        prof_end("nl_fill", st);
    } else {
        return fail(CUSOAP_ERR_CAPACITY, "nl_fill: a center has more than 16384 neighbours within the cutoff");
    }
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace cusoap
