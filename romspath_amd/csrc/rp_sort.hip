// rp_sort.hip -- periodic re-sort of the stored particle columns by grid cell.
//
// Counting sort on the key (grid, tile row, tile column) of each particle's
// rho cell, tiles of tile_x x tile_y cells: count, exclusive scan, scatter
// to a permutation, gather every column through it.
// Particles that no longer move (status 9, or out through an open boundary)
// share the last key.  The order inside a cell is whatever the scatter's
// atomics give; no result depends on it: every particle's step reads only
// its own columns, its deviates are keyed by its release id, and the output
// is returned in release order through id[].
#include "rp_kernels.h"

namespace rp {

__device__ __forceinline__ uint32_t cell_key(const SortArgs& K, int64_t q)
{
    const int8_t st = K.status[q];
    if (st == ST_UNSET || (K.open_boundary && !K.isin[q])) return K.base[K.ngrid];
    const int g = K.ngrid > 1 ? (int)K.gid[q] - 1 : 0;
    const int L = K.L[g], M = K.M[g];
    // clamp into [1, L-1] x [1, M-1] before converting (fmax(NaN, 1) = 1)
    const double X = fmin(fmax(K.x[q], 1.0), (double)(L - 1));
    const double Y = fmin(fmax(K.y[q], 1.0), (double)(M - 1));
    const int I = (int)X, J = (int)Y;
    return K.base[g] + (uint32_t)((J - 1) / K.tile_y * K.ntx[g] + (I - 1) / K.tile_x);
}

__global__ void __launch_bounds__(256) sort_fill(uint32_t* a, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ void __launch_bounds__(256) sort_count(SortArgs K)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < K.n;
         q += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&K.cnt[cell_key(K, q)], 1u);
}

// In-place exclusive scan of cnt[0..nb) by one block of 1024 threads: each
// thread sums a contiguous run, the run sums are scanned across the block,
// and each thread writes its run's prefixes.
__global__ void __launch_bounds__(1024) sort_scan(uint32_t* cnt, int64_t nb)
{
    __shared__ uint32_t wsum[16];
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = min(nb, (int64_t)t * per), b1 = min(nb, b0 + per);
    uint32_t s = 0;
    for (int64_t i = b0; i < b1; i++) s += cnt[i];
    uint32_t v = s;                                   // inclusive scan within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    if (wv == 0) {
        uint32_t w = lane < 16 ? wsum[lane] : 0u;
        for (int o = 1; o < 16; o <<= 1) {
            const uint32_t u = __shfl_up(w, o, 64);
            if (lane >= o) w += u;
        }
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    uint32_t run = (v - s) + (wv > 0 ? wsum[wv - 1] : 0u);
    for (int64_t i = b0; i < b1; i++) {
        const uint32_t c = cnt[i];
        cnt[i] = run;
        run += c;
    }
}

// Exclusive scan of cnt[0..nb) in tiles of kScanTile values (1024 threads x
// 4 consecutive values), three launches: sort_scan_tiles scans each tile in
// place and writes its total to sums[tile]; sort_scan (one block) scans the
// tile totals; sort_scan_add adds each tile's offset.
__device__ __forceinline__ uint32_t block_excl(uint32_t s, uint32_t* wsum, uint32_t& total)
{
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t v = s;                                   // inclusive scan within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    if (wv == 0) {
        uint32_t w = lane < 16 ? wsum[lane] : 0u;
        for (int o = 1; o < 16; o <<= 1) {
            const uint32_t u = __shfl_up(w, o, 64);
            if (lane >= o) w += u;
        }
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    total = wsum[15];
    return (v - s) + (wv > 0 ? wsum[wv - 1] : 0u);
}

__global__ void __launch_bounds__(1024) sort_scan_tiles(uint32_t* cnt, int64_t nb, uint32_t* sums)
{
    __shared__ uint32_t wsum[16];
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * 4;
    uint32_t x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = i0 + k < nb ? cnt[i0 + k] : 0u;
    uint32_t total;
    uint32_t run = block_excl(x[0] + x[1] + x[2] + x[3], wsum, total);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (i0 + k < nb) cnt[i0 + k] = run;
        run += x[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) sort_scan_add(uint32_t* cnt, int64_t nb, const uint32_t* sums)
{
    const uint32_t off = sums[blockIdx.x];
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * 4;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (i0 + k < nb) cnt[i0 + k] += off;
}

__global__ void __launch_bounds__(256) sort_scatter(SortArgs K)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < K.n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t pos = atomicAdd(&K.cnt[cell_key(K, q)], 1u);
        K.perm[pos] = (uint32_t)q;
    }
}

__global__ void __launch_bounds__(256) sort_permute(PermArgs P)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t s = P.perm[i];
#pragma unroll
        for (int k = 0; k < PermArgs::kF64; k++) P.d64[k][i] = P.s64[k][s];
#pragma unroll
        for (int k = 0; k < PermArgs::kU8; k++) P.d8[k][i] = P.s8[k][s];
        P.did[i] = P.sid[s];
    }
}

// write the gathered columns back over the source (sort of a sub-range)
__global__ void __launch_bounds__(256) sort_copy_back(PermArgs P)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n;
         i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < PermArgs::kF64; k++) const_cast<double*>(P.s64[k])[i] = P.d64[k][i];
#pragma unroll
        for (int k = 0; k < PermArgs::kU8; k++) const_cast<uint8_t*>(P.s8[k])[i] = P.d8[k][i];
        const_cast<uint32_t*>(P.sid)[i] = P.did[i];
    }
}

// back to release order: the particle with id r goes to position r
__global__ void __launch_bounds__(256) sort_restore(PermArgs P)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r = P.sid[i];
#pragma unroll
        for (int k = 0; k < PermArgs::kF64; k++) P.d64[k][r] = P.s64[k][i];
#pragma unroll
        for (int k = 0; k < PermArgs::kU8; k++) P.d8[k][r] = P.s8[k][i];
        P.did[r] = r;
    }
}

// counter reset as a kernel, so that a captured launch sequence holds no memset
__global__ void zero_u64(unsigned long long* p)
{
    if (threadIdx.x == 0) p[0] = 0ull;
}

}  // namespace rp
