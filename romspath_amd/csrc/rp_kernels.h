// rp_kernels.h -- kernel argument blocks and kernel declarations shared by
// the step kernels and the C-ABI host layer.
#pragma once

#include "rp_device.h"
#include "rp_modal.h"

namespace rp {

// The reference's secant search for the smoothing parameter has no iteration
// cap (pppack.f90:5270-5292); a GPU lane must terminate, so it stops here and
// the stop is counted (rp_secant_capped reports it).
constexpr int kSecantCap = 64;

// qty rows of the FP32-field vertical walk kept in LDS per lane: the FP32
// step kernels take blockDim x kLdsQ x 4 bytes of dynamic LDS
constexpr int kLdsQ = 25;

struct SlotSel {
    int s[3];   // physical slot (0..3) holding t_b, t_c, t_f
};

struct StepArgs {
    double ex[3], ix[3];
    double pw[3];                // Lagrange weights of polintd(ex, ., ix(2))
    int ets;
    SlotSel slots;
    int idt;
    double deltat, serr, smth, aksback;
    int vturb_on, hturb_on, scheme, nsb;
    double vertdist;
    int open_boundary, ngrid;
    int np, loop;                // np: int(s_w * sub) of the grid with the most levels
    int np_mixed;                // the grids' np differ: each lane takes its grid's np
    int64_t q0, n;               // stored particles q0 .. q0+n-1 take this launch
    int64_t nmid;                // row stride of mid (>= n); row r holds particle q0 + r
    int64_t step_index;
    uint64_t seed;
    double *x, *y, *z, *age;
    const double* dob;
    int8_t* status;
    int8_t* gid;
    uint8_t* isin;
    double *wd, *zeta, *bath;
    int want_diag;
    const uint32_t* id;          // release-order id of each stored particle
    const double* dev;           // injected deviates [id][ndev] or null
    int ndev;
    double* scratch;             // VTurb work arrays, column per lane
    uint32_t scratch_bytes;      // extent of scratch addressed as FP32 rows (< 2^31)
    double* mid;                 // [4][nmid] newX, newY, AdvectZ, clamped z between phases
    unsigned long long* active;  // += particles that took the step
    unsigned long long* work;    // vwalk_kernel's next unclaimed particle (zeroed per launch)
    unsigned* capped;            // += secant searches stopped at the cap
    const double* pert;          // velocity perturbation coefficients [2][ny][nx] or null
    int pert_nx, pert_ny;
    // the launch's active set (compact_active): particle q0 + act[r] for
    // r < *act_n; null: every particle q0 .. q0+n-1, inactive ones skipped
    const uint32_t* act;
    const unsigned long long* act_n;
    // eigenbasis tables of the FP32 form's vertical walk (every grid has the
    // same resampling size), or null
    const ModalTab* modal;
};

// NESTED=false: one grid, all lanes read grid 0 (uniform, scalar loads).
// step_kernel: the whole step (no vertical walk); with the vertical walk,
// advect_kernel (advection + horizontal walk) then vwalk_kernel (the rest).
template <typename F, bool NESTED>
__global__ void step_kernel(const DGrid* __restrict__ grids, StepArgs A);
template <typename F, bool NESTED>
__global__ void advect_kernel(const DGrid* __restrict__ grids, StepArgs A);
template <typename F, bool NESTED>
__global__ void vwalk_kernel(const DGrid* __restrict__ grids, StepArgs A);
// the FP32-field vertical walk with the secant search in the eigenbasis of
// chol1d's pencil (A.modal set): static LDS for the two basis tables, no
// dynamic LDS
template <bool NESTED>
__global__ void vwalk_modal_kernel(const DGrid* __restrict__ grids, StepArgs A);

// The active set of a launch (the particles that pass the skip tests of
// ROMSPath.f90:672-690), in particle order: compact_count writes each
// 256-particle block's count, an exclusive scan of the counts (sort_scan)
// gives the block offsets, compact_scatter writes act[] by wave votes and
// prefixes and the count to act_n.  Blocks of 256 threads.
__global__ void compact_count(StepArgs A, uint32_t* bcnt);
__global__ void compact_scatter(StepArgs A, const uint32_t* boff, uint32_t* act,
                                unsigned long long* act_n);

template <typename F>
__global__ void pack_field(const double* __restrict__ src, const double* __restrict__ add, F* dst,
                           int nx, int ny, int nk, int slot);

__global__ void sample_deviates(uint64_t seed, int64_t step, int64_t n, int ndev, double* out);

// getInterp2D("lon"/"lat") at every particle's position on its grid
// (interpolation_module.f90:421-507, called from dataOutput, ROMSPath.f90:1012-1025)
struct LonLatArgs {
    const double* lon[kMaxGrid];
    const double* lat[kMaxGrid];
    int L[kMaxGrid], M[kMaxGrid];
    const double *x, *y;
    const int8_t* gid;
    int64_t n;
    double *out_lon, *out_lat;
};
__global__ void sample_lonlat(LonLatArgs A);

// placement of released particles (ini_ROMSPath, ROMSPath.f90:322-403)
struct LocateArgs {
    const double* lon[kMaxGrid];
    const double* lat[kMaxGrid];
    const double* angle[kMaxGrid];   // null: 0
    int ngrid, spherical, tb_lane;   // tb_lane: physical slot of t_b
    int64_t n;
    const double *inlon, *inlat;
    double *z, *x, *y, *status;
    int32_t* gid;
    uint8_t* isin;
};
template <typename F>
__global__ void locate_kernel(const DGrid* __restrict__ grids, LocateArgs A);

// salt and temp at the particles (SaltTempOn, ROMSPath.f90:958-981).
// mode 0: every particle, into out_salt/out_temp in stored order; mode 1
// (release, ROMSPath.f90:408-424): the particles in a grid, P = the sample,
// mean = P, and last_age = age for all; mode 2 (after a step): the particles
// whose age changed, P = the sample, mean += P.  Modes 1 and 2 index the
// release-order columns through id; mean_salt/mean_temp may be null.
// Modes 0 and 1 use the depth at the particle; mode 2 the step's own depth
// (bath), which for a rejected move is the attempted position's
// (ROMSPath.f90:849, 964-971).
struct TracerArgs {
    const void* salt[kMaxGrid];   // slot-interleaved, field precision, (L, M, N)
    const void* temp[kMaxGrid];
    SlotSel slots;                 // physical slots of t_b, t_c, t_f
    double ex[3], ix2;
    int64_t n;
    const double *x, *y, *z;
    const int8_t* gid;
    double *out_salt, *out_temp;
    int mode;
    const uint32_t* id;
    const double* age;
    const uint8_t* isin;
    // mode 2 reads bath, the step's pBath (-tdepth of the step); mode 1
    // writes pWD, pZeta, pBath of the particles in a grid (ROMSPath.f90:415-417)
    double *bath, *wd, *pzeta;
    double *last_age, *p_salt, *p_temp, *mean_salt, *mean_temp;
};
template <typename F>
__global__ void tracer_kernel(const DGrid* __restrict__ grids, TracerArgs A);

// WriteBottom (ROMSPath.f90:935-951): the running minimum height above the
// bed and the bottom stresses where it was reached; the per-particle columns
// are in release order (indexed through id)
struct BottomArgs {
    const void* bustr[kMaxGrid];   // slot-interleaved, field precision, (L-1, M)
    const void* bvstr[kMaxGrid];   // (L, M-1)
    SlotSel slots;
    double ex[3], ix2;
    int64_t n;
    const double *x, *y, *z, *age, *bath;
    const int8_t* gid;
    const uint32_t* id;
    double *last_age, *hob, *bustr_out, *bvstr_out;
};
template <typename F>
__global__ void bottom_kernel(const DGrid* __restrict__ grids, BottomArgs A);

// tables of the FP32-field form's p = 0 spline sweep; once per device, 0 on success
int upload_p0_tables();

// current fields of a step: dst[e] = sum_t pw[t] src[4 e' + slot[t]] in FP64,
// stored FP32; one job per (grid, field), blockIdx.y selects the job.  A job
// copies columns of kin levels into columns of kout >= kin levels: output e
// = col kout + k takes source point e' = col kin + k for k < kin and is 0
// in the padding.  n counts output points.  With wet/dry masks (u and v
// under -DWETDRY), slot t's term is taken only where that slot's mask at
// the point's column, wd[slot[t] * wdp + e' / nk], is wet.
constexpr int kMaxCurJobs = 5 * kMaxGrid;
struct CurJobs {
    const float* src[kMaxCurJobs];
    float* dst[kMaxCurJobs];
    int64_t n[kMaxCurJobs];
    int kin[kMaxCurJobs], kout[kMaxCurJobs];
    int slot[kMaxCurJobs][3];
    const uint8_t* wd[kMaxCurJobs];   // null: no masks
    int64_t wdp[kMaxCurJobs];
    int nk[kMaxCurJobs];
    double pw[3];
    int njobs;
};
__global__ void current_fields(CurJobs J);

// re-sort of the particle columns by cell (rp_sort.hip)
struct SortArgs {
    const double *x, *y;
    const int8_t *status, *gid;
    const uint8_t* isin;
    int64_t n;
    int open_boundary, ngrid;
    int L[kMaxGrid], M[kMaxGrid];
    int tile_x, tile_y;            // cells per sort tile
    int ntx[kMaxGrid];             // tiles per row of each grid
    uint32_t base[kMaxGrid + 1];   // first key of each grid; base[ngrid] = parked particles
    uint32_t* cnt;                 // base[ngrid] + 1 bins
    uint32_t* perm;                // n: stored index of the particle at each sorted position
};

struct PermArgs {
    static constexpr int kF64 = 8;   // x, y, z, age, dob, wd, zeta, bath
    static constexpr int kU8 = 3;    // status, gid, isin
    const uint32_t* perm;
    int64_t n;
    const double* s64[kF64];
    double* d64[kF64];
    const uint8_t* s8[kU8];
    uint8_t* d8[kU8];
    const uint32_t* sid;
    uint32_t* did;
};

__global__ void sort_fill(uint32_t* a, int64_t n, uint32_t v);
__global__ void sort_count(SortArgs K);
__global__ void sort_scan(uint32_t* cnt, int64_t nb);
constexpr int kScanTile = 4096;
__global__ void sort_scan_tiles(uint32_t* cnt, int64_t nb, uint32_t* sums);
__global__ void sort_scan_add(uint32_t* cnt, int64_t nb, const uint32_t* sums);
__global__ void sort_scatter(SortArgs K);
__global__ void sort_permute(PermArgs P);
__global__ void sort_copy_back(PermArgs P);
__global__ void sort_restore(PermArgs P);
__global__ void zero_u64(unsigned long long* p);

// ---------------------------------------------------------------------------
// adjoint (rp_adjoint.hip)
// ---------------------------------------------------------------------------
struct AdjState {                // particle state at the start of one step
    double *x, *y, *z;
    int8_t *status, *gid;
    uint8_t* isin;
};

struct AdjArgs {
    double *lx, *ly, *lz;        // lambda = dL/d(x, y, z), stored order, updated in place
    double* gpert;               // dL/dc [2][ny][nx] accumulated, or null
};

template <typename F, bool NESTED>
__global__ void adj_forward_kernel(const DGrid* __restrict__ grids, StepArgs A);
template <typename F, bool NESTED>
__global__ void adj_step_kernel(const DGrid* __restrict__ grids, StepArgs A, AdjArgs J);
__global__ void adj_save_kernel(AdjState dst, const double* x, const double* y, const double* z,
                                const int8_t* status, const int8_t* gid, const uint8_t* isin,
                                int64_t n);
__global__ void adj_fill(double* a, int64_t n, double v);

// rows of VTurb scratch per lane for np resampled points and nw w-levels
inline int64_t vturb_rows(int np, int nw) { return 9LL * np + 2LL * nw; }

}  // namespace rp
