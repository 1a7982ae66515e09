// rp_abi.hip -- host side of librp.so: context, device memory, uploads and
// the extern "C" entry points declared in include/rp_abi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rp_abi.h"
#include "rp_kernels.h"

using namespace rp;

namespace {

// ---------------------------------------------------------------------------
// Scratch of rp_step_noalloc, which allocates nothing: one file-scope device
// array, held by every GPU the library runs on until the process ends.
//   [0, kNaHead)                  the vertical walk's work-queue counter
//   [kNaHead, +kNaMidBytes)       hand-over rows of one launch pair:
//                                 4 FP64 rows of kNaChunk particles
//   [kNaHead + kNaMidBytes, end)  the vertical walk's per-lane work rows
// Larger particle sets are stepped in launch pairs of kNaChunk particles;
// the row region bounds the lanes of the persistent vertical walk.  Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU race.
// Host code uses the address hipGetSymbolAddress returned (na_ptr), never
// the array's name.
// ---------------------------------------------------------------------------
constexpr size_t kNaBytes = 64ull << 20;
constexpr size_t kNaHead = 256;
constexpr int64_t kNaChunk = 131072;
constexpr size_t kNaMidBytes = 4ull * sizeof(double) * kNaChunk;
constexpr size_t kNaRowBytes = kNaBytes - kNaHead - kNaMidBytes;
static_assert(kNaRowBytes < 0x7FFFFFFFull, "row region addressed through one buffer descriptor");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDev = 64;
std::mutex na_mu;
char* na_ptr[kNaMaxDev] = {};

// the array's device address on the current device, looked up once per device
int na_setup(int dev, char** out)
{
    if (dev < 0 || dev >= kNaMaxDev) return -1;
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        na_ptr[dev] = (char*)p;
    }
    *out = na_ptr[dev];
    return 0;
}

struct GridDev {
    int L = 0, M = 0, N = 0, vtransform = 0;
    double hc = 0;
    double *s_rho = nullptr, *cs_r = nullptr, *s_w = nullptr, *cs_w = nullptr;
    double *h = nullptr, *pm = nullptr, *pn = nullptr;
    int8_t *mrho = nullptr, *mu = nullptr, *mv = nullptr;
    void *zeta = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *aks = nullptr;
    float* cur[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // FP32 form: fields at ix(2)
    ModalTab* modal = nullptr;                // FP32 form: eigenbasis tables of this grid's n
    double *lon = nullptr, *lat = nullptr;   // lon_rho / lat_rho (x_rho / y_rho)
    double* angle = nullptr;                  // grid angle (spherical grids), may stay null
    void *salt = nullptr, *temp = nullptr;    // SaltTempOn fields, slot-interleaved (L, M, N)
    bool tracer_set[3] = {false, false, false};
    void *bustr = nullptr, *bvstr = nullptr;  // WriteBottom fields, slot-interleaved (L-1, M), (L, M-1)
    bool bstress_set[3] = {false, false, false};
    // wet/dry masks (rp_set_hydro_wetdry): per slot, rho then u then v
    // points, one byte each; null until the first call
    uint8_t* wd = nullptr;
    double scl[kMaxGrid][2], off[kMaxGrid][2];
    bool set = false;
    bool slot_set[3] = {false, false, false};
};

struct Ctx {
    rp_config cfg;
    // allocation-failure test knob: RP_FAIL_ALLOC_AFTER=k in the environment
    // at rp_create lets the context's first k device or pinned allocations
    // succeed and refuses the rest with RP_E_ALLOC; -1 (unset): no limit
    int64_t alloc_budget = -1;
    char* na = nullptr;          // this device's rp_step_noalloc scratch (na_setup)
    int fbytes = 4;
    int spherical = 0;           // GRIDS(1)%spherical, for LL2ij
    int device = 0;
    int ncu = 256;
    hipStream_t stream = nullptr;
    GridDev grids[kMaxGrid];
    // particles (stored order; id[] gives release order)
    int64_t n = 0;
    double *x = nullptr, *y = nullptr, *z = nullptr, *age = nullptr, *dob = nullptr;
    double *wd = nullptr, *zeta = nullptr, *bath = nullptr;
    int8_t *status = nullptr, *gid = nullptr;
    uint8_t* isin = nullptr;
    uint32_t* id = nullptr;
    double* mid = nullptr;       // [4][n] hand-over between the step's two kernels
    uint32_t* act = nullptr;     // [n] the active set of a launch (compact_scatter)
    uint32_t* act_blk = nullptr; // [(n + 255) / 256] block counts, then offsets
    unsigned long long* act_n = nullptr;   // its count
    // injected deviates
    double* dev = nullptr;
    int ndev = 0;
    bool use_dev = false;
    // hydro uploads (upload_field): the caller's FP64 arrays are copied into
    // pinned host buffers before the call returns, the host-to-device copies
    // run on their own stream while queued steps execute, and each field's
    // pack kernel runs in order on the step stream.  Device staging is kept
    // per (grid, field), so one record's copies never wait for the steps
    // still reading the previous one.
    hipStream_t up_stream = nullptr;
    struct Pinned {
        double* p = nullptr;
        size_t elems = 0;
        hipEvent_t done = nullptr;      // its last host-to-device copy finished
        bool used = false;
    };
    Pinned pinned[2];
    int pinned_next = 0;
    struct FieldStage {
        double* dev = nullptr;
        size_t elems = 0;
        hipEvent_t copied = nullptr;    // on up_stream: the record is on the device
        hipEvent_t packed = nullptr;    // on stream: the pack kernel has read it
        bool used = false;
    };
    FieldStage fstage[kMaxGrid][9];   // zeta, u, v, w, aks, salt, temp, bustr, bvstr
    // SaltTempOn per step (rp_tracers_begin): release order, trc_n particles
    double *trc_last_age = nullptr, *trc_salt = nullptr, *trc_temp = nullptr;
    double *trc_mean_salt = nullptr, *trc_mean_temp = nullptr;   // SaltTempMean, else null
    int64_t trc_n = -1;
    bool trc_mean = false;
    int64_t trc_mI = 0;
    // WriteBottom (rp_bottom_begin): release order, bot_n particles
    double *bot_last_age = nullptr, *bot_hob = nullptr, *bot_bustr = nullptr, *bot_bvstr = nullptr;
    int64_t bot_n = -1;
    // VTurb scratch
    double* scratch = nullptr;
    size_t scratch_bytes = 0;
    // device copy of the grid table read by the step kernel
    DGrid* dgrids = nullptr;
    DGrid tab_host[kMaxGrid];
    // counters
    unsigned long long* active = nullptr;
    unsigned long long* work = nullptr;
    unsigned* capped = nullptr;
    // [form: FP32 band sweeps, FP64, FP32 eigenbasis][nested], 0 = not queried
    int vwalk_blocks_per_cu[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    // RP_VWALK_FORM=band at rp_create: the FP32 form keeps its band-sweep
    // vertical walk even where the eigenbasis one applies
    bool vwalk_band = false;
    // per step: events before the step, between its two phases, after it;
    // a captured batch of steps has one entry covering all of its steps
    struct StepEvents {
        hipEvent_t e[3];
        int steps = 1;
    };
    std::vector<StepEvents> ev_used, ev_free;
    // re-sort by cell (cfg.sort_interval > 0)
    int64_t steps_done = 0;
    uint32_t* sort_cnt = nullptr;
    int64_t sort_nb = 0;
    uint32_t* perm = nullptr;
    double* alt64[PermArgs::kF64] = {};
    uint8_t* alt8[PermArgs::kU8] = {};
    uint32_t* alt_id = nullptr;
    int64_t alt_n = -1;
    int sort_tile[2] = {1, 1};           // cells per sort tile (x, y)
    bool ranges_intact = true;           // each scenario's particles within its range
    // ensembles: scenario s holds stored positions first[s] .. first[s]+count[s]-1
    std::vector<int64_t> scen_first, scen_count, scen_steps;
    hipGraphExec_t gexec = nullptr;      // last captured batch of steps
    // current fields (FP32 form): what they were last built for
    uint64_t field_gen = 0;              // bumped by every hydro upload
    uint64_t cur_gen = ~0ull;
    double cur_pw[3] = {0.0, 0.0, 0.0};
    int cur_slots[3] = {-1, -1, -1};
    int cur_first = -1;
    // velocity perturbation (rp_set_perturbation), device copy or null
    double* pert = nullptr;
    int pert_nx = 0, pert_ny = 0;
    // adjoint recording (rp_adjoint_record)
    struct AdjSnap {
        void* f[kMaxGrid][5];         // copies of zeta, u, v, w, aks of every grid
        uint8_t* wd[kMaxGrid];        // and of its wet/dry masks, when it has them
        DGrid* dtab;                  // device grid table pointing at them
    };
    struct AdjRec {
        int K = 0;                    // checkpoint interval, 0 = not recording
        bool dirty = true;            // hydro slots changed since the last snapshot
        std::vector<StepArgs> args;   // per recorded step (particle pointers set at launch)
        std::vector<int> snap;        // snapshot of each step
        std::vector<AdjSnap> snaps;
        std::vector<char*> ckpt;      // state at steps 0, K, 2K, ...
        std::vector<double*> devs;    // injected deviates of each step, or null
        unsigned long long* active = nullptr;   // counter for recomputed steps
    } adj;
    double phase_ms[2] = {0.0, 0.0};     // averages of the last rp_kernel_timing call
    std::string err;
};

int fail(Ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((c), RP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

template <typename T>
int dalloc(Ctx* c, T** p, size_t count)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    if (c->alloc_budget == 0)
        return fail(c, RP_E_ALLOC, "allocation of " + std::to_string(count * sizeof(T)) +
                                       " bytes refused (RP_FAIL_ALLOC_AFTER)");
    if (c->alloc_budget > 0) c->alloc_budget--;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, RP_E_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                       " bytes failed: " + hipGetErrorString(e));
    }
    return RP_OK;
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int put(Ctx* c, void* dst, const void* src, size_t bytes)
{
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return RP_OK;
}

// The reference's status literals are default REAL (single precision):
// par(n,pStatus)=1.1 and =1.2 (ROMSPath.f90:880-881) store 1.1f and 1.2f
// widened to double.  Those are the values returned; the double 1.1 and 1.2
// are accepted on input too.
constexpr double kStTop = (double)1.1f, kStBot = (double)1.2f;

int status_code(double s, int8_t* out)
{
    if (s == 9.0) *out = ST_UNSET;
    else if (s == 0.0) *out = ST_NEW;
    else if (s == 1.0) *out = ST_MOVED;
    else if (s == kStTop || s == 1.1) *out = ST_TOP;
    else if (s == kStBot || s == 1.2) *out = ST_BOT;
    else if (s == -8.0) *out = ST_EXIT;
    else if (s == -9.0) *out = ST_LAND;
    else return -1;
    return 0;
}

double status_value(int8_t c)
{
    switch (c) {
    case ST_UNSET: return 9.0;
    case ST_NEW: return 0.0;
    case ST_MOVED: return 1.0;
    case ST_TOP: return kStTop;
    case ST_BOT: return kStBot;
    case ST_EXIT: return -8.0;
    case ST_LAND: return -9.0;
    default: return NAN;
    }
}

int make_event(Ctx* c, hipEvent_t& e)
{
    if (!e) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return RP_OK;
}

// Upload one FP64 Fortran (nx, ny, nk) field into physical slot `slot` of
// the interleaved device array, adding `add` (may be null) in FP64 first.
// Field f of grid gi has its own device staging buffer.  The call returns
// once the caller's arrays have been read; the conversion is ordered on the
// step stream, after the steps already queued there and before later ones.
const char* const kFieldName[9] = {"zeta", "u", "v", "w", "AKs", "salt", "temp", "bustr", "bvstr"};

// the first non-finite value of a caller's array, or -1
int64_t first_nonfinite(const double* a, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(a[i])) return (int64_t)i;
    return -1;
}

int upload_field(Ctx* c, int gi, int f, void* dst, const double* src, const double* add, int nx,
                 int ny, int nk, int slot)
{
    const size_t elems = (size_t)nx * ny * nk;
    const size_t need = add ? 2 * elems : elems;
    int rc;
    // a NaN or Inf in a hydro record is rejected before anything is queued
    for (int k = 0; k < (add ? 2 : 1); k++) {
        const int64_t bad = first_nonfinite(k ? add : src, elems);
        if (bad >= 0)
            return fail(c, RP_E_ARG, std::string("non-finite value in the ") + (k ? "Stokes drift of " : "") +
                                         kFieldName[f] + " field of grid " + std::to_string(gi + 1) +
                                         " at element " + std::to_string(bad));
    }
    if (!c->up_stream) HIPCHK(c, hipStreamCreateWithFlags(&c->up_stream, hipStreamNonBlocking));
    Ctx::FieldStage& fs = c->fstage[gi][f];
    if ((rc = make_event(c, fs.copied)) || (rc = make_event(c, fs.packed))) return rc;
    if (fs.elems < need) {
        if (fs.used) HIPCHK(c, hipEventSynchronize(fs.packed));
        fs.elems = 0;
        if ((rc = dalloc(c, &fs.dev, need))) return rc;
        fs.elems = need;
    } else if (fs.used) {
        // the copy must not overwrite what the previous pack still reads
        HIPCHK(c, hipStreamWaitEvent(c->up_stream, fs.packed, 0));
    }
    const double* part[2] = {src, add};
    for (int k = 0; k < (add ? 2 : 1); k++) {
        Ctx::Pinned& h = c->pinned[c->pinned_next];
        c->pinned_next ^= 1;
        if ((rc = make_event(c, h.done))) return rc;
        if (h.used) HIPCHK(c, hipEventSynchronize(h.done));
        if (h.elems < elems) {
            if (h.p) (void)hipHostFree(h.p);
            h.p = nullptr;
            h.elems = 0;
            if (c->alloc_budget == 0)
                return fail(c, RP_E_ALLOC, "pinned staging allocation refused (RP_FAIL_ALLOC_AFTER)");
            if (c->alloc_budget > 0) c->alloc_budget--;
            HIPCHK(c, hipHostMalloc((void**)&h.p, elems * sizeof(double), hipHostMallocDefault));
            h.elems = elems;
        }
        std::memcpy(h.p, part[k], elems * sizeof(double));
        HIPCHK(c, hipMemcpyAsync(fs.dev + k * elems, h.p, elems * sizeof(double),
                                 hipMemcpyHostToDevice, c->up_stream));
        HIPCHK(c, hipEventRecord(h.done, c->up_stream));
        h.used = true;
    }
    HIPCHK(c, hipEventRecord(fs.copied, c->up_stream));
    HIPCHK(c, hipStreamWaitEvent(c->stream, fs.copied, 0));
    const double* dadd = add ? fs.dev + elems : nullptr;
    const int block = 256;
    const int64_t want = ((int64_t)elems + block - 1) / block;
    const int grid = (int)std::min<int64_t>(want, (int64_t)c->ncu * 16);
    if (c->fbytes == 4)
        hipLaunchKernelGGL(pack_field<float>, dim3(grid), dim3(block), 0, c->stream, fs.dev, dadd,
                           (float*)dst, nx, ny, nk, slot);
    else
        hipLaunchKernelGGL(pack_field<double>, dim3(grid), dim3(block), 0, c->stream, fs.dev, dadd,
                           (double*)dst, nx, ny, nk, slot);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(fs.packed, c->stream));
    fs.used = true;
    return RP_OK;
}

int np_of(const Ctx* c, int N) { return (int)((double)(N + 1) * c->cfg.sub); }

// The eigenbasis tables (rp_modal.h, rp_modal_basis) of n interior
// unknowns; the form applies to even n from kMh to 2 kMh (the kernel reads
// rows j and n-1-j, j < kMh, without a branch).
bool modal_applies(int n) { return n >= kMh && n % 2 == 0 && n / 2 <= kMh; }
int build_modal_tab(int n, ModalTab& T)
{
    std::memset(&T, 0, sizeof T);
    T.n = n;
    T.half = n / 2;
    return rp_modal_basis(n, &T.basis[0][0][0], &T.lam[0][0]);
}

// bytes of one slot's wet/dry masks: rho, u and v points
int64_t wd_plane(const GridDev& g)
{
    return (int64_t)g.L * g.M + (int64_t)(g.L - 1) * g.M + (int64_t)g.L * (g.M - 1);
}

void free_grid(GridDev& g)
{
    dfree(g.s_rho); dfree(g.cs_r); dfree(g.s_w); dfree(g.cs_w);
    dfree(g.h); dfree(g.pm); dfree(g.pn);
    dfree(g.mrho); dfree(g.mu); dfree(g.mv);
    dfree(g.zeta); dfree(g.u); dfree(g.v); dfree(g.w); dfree(g.aks);
    for (auto& p : g.cur) dfree(p);
    dfree(g.modal);
    dfree(g.lon); dfree(g.lat); dfree(g.angle);
    dfree(g.salt); dfree(g.temp);
    for (auto& t : g.tracer_set) t = false;
    dfree(g.bustr); dfree(g.bvstr);
    for (auto& t : g.bstress_set) t = false;
    dfree(g.wd);
    g.set = false;
}


// ---------------------------------------------------------------------------
// re-sort of the stored particle columns
// ---------------------------------------------------------------------------

// key layout of the sort: tiles per row and first key of each grid; returns
// the number of bins (the last bin holds the particles that no longer move)
int64_t sort_keys(const Ctx* c, SortArgs& K)
{
    uint32_t nb = 0;
    K.tile_x = c->sort_tile[0];
    K.tile_y = c->sort_tile[1];
    for (int g = 0; g < c->cfg.ngrid; g++) {
        const int L = c->grids[g].L, M = c->grids[g].M;
        K.base[g] = nb;
        K.L[g] = L;
        K.M[g] = M;
        K.ntx[g] = (L - 2) / K.tile_x + 1;
        nb += (uint32_t)K.ntx[g] * (uint32_t)((M - 2) / K.tile_y + 1);
    }
    K.base[c->cfg.ngrid] = nb;
    return (int64_t)nb + 1;
}

int sort_buffers(Ctx* c, int64_t bins)
{
    if (c->sort_nb < bins) {
        // the counts, then (from a multiple of 4) one total per scan tile
        const int64_t tiles = (bins + kScanTile - 1) / kScanTile;
        if (dalloc(c, &c->sort_cnt, (size_t)(((bins + 3) & ~3LL) + tiles))) return RP_E_ALLOC;
        c->sort_nb = bins;
    }
    if (c->alt_n != c->n) {
        int rc = dalloc(c, &c->perm, (size_t)c->n);
        for (auto& a : c->alt64) rc |= dalloc(c, &a, (size_t)c->n);
        for (auto& a : c->alt8) rc |= dalloc(c, &a, (size_t)c->n);
        rc |= dalloc(c, &c->alt_id, (size_t)c->n);
        if (rc) return RP_E_ALLOC;
        c->alt_n = c->n;
    }
    return RP_OK;
}

PermArgs perm_args(Ctx* c, int64_t q0, int64_t cnt, double** cur64[], uint8_t** cur8[])
{
    PermArgs P;
    P.perm = c->perm;
    P.n = cnt;
    cur64[0] = &c->x; cur64[1] = &c->y; cur64[2] = &c->z; cur64[3] = &c->age;
    cur64[4] = &c->dob; cur64[5] = &c->wd; cur64[6] = &c->zeta; cur64[7] = &c->bath;
    cur8[0] = (uint8_t**)&c->status; cur8[1] = (uint8_t**)&c->gid; cur8[2] = &c->isin;
    for (int k = 0; k < PermArgs::kF64; k++) {
        P.s64[k] = *cur64[k] + q0;
        P.d64[k] = c->alt64[k] + q0;
    }
    for (int k = 0; k < PermArgs::kU8; k++) {
        P.s8[k] = *cur8[k] + q0;
        P.d8[k] = c->alt8[k] + q0;
    }
    P.sid = c->id + q0;
    P.did = c->alt_id + q0;
    return P;
}

void swap_columns(Ctx* c, double** cur64[], uint8_t** cur8[])
{
    for (int k = 0; k < PermArgs::kF64; k++) std::swap(*cur64[k], c->alt64[k]);
    for (int k = 0; k < PermArgs::kU8; k++) std::swap(*cur8[k], c->alt8[k]);
    std::swap(c->id, c->alt_id);
}

// Counting sort of stored particles q0 .. q0+cnt-1 by (grid, tile) on the
// context stream, launches only (buffers from sort_buffers).  The whole set
// swaps with the spare set; a sub-range is copied back into place.
int sort_range(Ctx* c, int64_t q0, int64_t cnt)
{
    SortArgs K;
    const int64_t bins = sort_keys(c, K);
    if (c->sort_nb < bins || c->alt_n != c->n) return fail(c, RP_E_STATE, "sort buffers missing");
    K.x = c->x + q0;
    K.y = c->y + q0;
    K.status = c->status + q0;
    K.gid = c->gid + q0;
    K.isin = c->isin + q0;
    K.n = cnt;
    K.open_boundary = c->cfg.open_boundary != 0;
    K.ngrid = c->cfg.ngrid;
    K.cnt = c->sort_cnt;
    K.perm = c->perm;
    const unsigned gb = (unsigned)std::min<int64_t>((bins + 255) / 256, 4096);
    const unsigned gp = (unsigned)std::min<int64_t>((cnt + 255) / 256, 8192);
    hipLaunchKernelGGL(sort_fill, dim3(gb), dim3(256), 0, c->stream, c->sort_cnt, bins, 0u);
    hipLaunchKernelGGL(sort_count, dim3(gp), dim3(256), 0, c->stream, K);
    const int64_t tiles = (bins + kScanTile - 1) / kScanTile;
    uint32_t* sums = c->sort_cnt + ((bins + 3) & ~3LL);
    hipLaunchKernelGGL(sort_scan_tiles, dim3((unsigned)tiles), dim3(1024), 0, c->stream, c->sort_cnt, bins, sums);
    hipLaunchKernelGGL(sort_scan, dim3(1), dim3(1024), 0, c->stream, sums, tiles);
    hipLaunchKernelGGL(sort_scan_add, dim3((unsigned)tiles), dim3(1024), 0, c->stream, c->sort_cnt, bins, sums);
    hipLaunchKernelGGL(sort_scatter, dim3(gp), dim3(256), 0, c->stream, K);
    double** cur64[PermArgs::kF64];
    uint8_t** cur8[PermArgs::kU8];
    const PermArgs P = perm_args(c, q0, cnt, cur64, cur8);
    hipLaunchKernelGGL(sort_permute, dim3(gp), dim3(256), 0, c->stream, P);
    if (q0 == 0 && cnt == c->n) {
        HIPCHK(c, hipGetLastError());
        swap_columns(c, cur64, cur8);
    } else {
        hipLaunchKernelGGL(sort_copy_back, dim3(gp), dim3(256), 0, c->stream, P);
        HIPCHK(c, hipGetLastError());
    }
    return RP_OK;
}

// stored order back to release order (so scenario ranges hold again)
int restore_order(Ctx* c)
{
    double** cur64[PermArgs::kF64];
    uint8_t** cur8[PermArgs::kU8];
    const PermArgs P = perm_args(c, 0, c->n, cur64, cur8);
    const unsigned gp = (unsigned)std::min<int64_t>((c->n + 255) / 256, 8192);
    hipLaunchKernelGGL(sort_restore, dim3(gp), dim3(256), 0, c->stream, P);
    HIPCHK(c, hipGetLastError());
    swap_columns(c, cur64, cur8);
    c->ranges_intact = true;
    return RP_OK;
}

// ---------------------------------------------------------------------------
// one step: host preparation (checks, buffers, grid table) apart from the
// launches, so that a sequence of steps can be captured into a HIP graph
// ---------------------------------------------------------------------------
using KFn = void (*)(const DGrid*, StepArgs);

struct StepPlan {
    StepArgs A;
    int block = 256;
    int64_t grid_all = 0, grid_vt = 0;
    int64_t vt_cap = 0;                        // blocks of the persistent vertical walk at most
    bool vt = false;
    KFn k[3] = {nullptr, nullptr, nullptr};   // step / advect / vertical walk
    bool cur = false;                          // FP32 form: build the current fields first
    bool modal = false;                        // vertical walk: vwalk_modal_kernel
    CurJobs jobs;
};

// dynamic LDS of the vertical walk: the band-sweep FP32 form keeps qty rows
// there; the eigenbasis form's tables are static LDS
size_t vwalk_lds(const StepPlan& P)
{
    return P.cur && !P.modal ? (size_t)P.block * kLdsQ * 4 : 0;
}

// points of field f (zeta, u, v, w, aks) of a grid
int64_t field_points(const GridDev& g, int f)
{
    switch (f) {
    case 0: return (int64_t)g.L * g.M;
    case 1: return (int64_t)(g.L - 1) * g.M * g.N;
    case 2: return (int64_t)g.L * (g.M - 1) * g.N;
    default: return (int64_t)g.L * g.M * (g.N + 1);
    }
}

// levels per column of the current AKs field
int cur_column(const GridDev& g) { return (g.N + 1 + 3) & ~3; }

int check_step(Ctx* c, int t_b, int t_c, int t_f, int64_t step_index)
{
    if (t_b < 1 || t_b > 3 || t_c < 1 || t_c > 3 || t_f < 1 || t_f > 3)
        return fail(c, RP_E_ARG, "slot numbers must be 1..3");
    if (step_index < 0) return fail(c, RP_E_ARG, "negative step index");
    for (int i = 0; i < c->cfg.ngrid; i++) {
        const GridDev& g = c->grids[i];
        if (!g.set) return fail(c, RP_E_STATE, "rp_step: grid " + std::to_string(i + 1) + " not set");
        if (!g.slot_set[t_b - 1] || !g.slot_set[t_c - 1] || !g.slot_set[t_f - 1])
            return fail(c, RP_E_STATE, "rp_step: hydro slot not loaded");
    }
    return RP_OK;
}

// the device view of every grid, as the step kernels read it
void fill_tab(const Ctx* c, DGrid tab[kMaxGrid])
{
    std::memset(tab, 0, sizeof(DGrid) * kMaxGrid);
    for (int i = 0; i < c->cfg.ngrid; i++) {
        const GridDev& g = c->grids[i];
        DGrid& d = tab[i];
        d.L = g.L; d.M = g.M; d.N = g.N; d.vtransform = g.vtransform; d.hc = g.hc;
        d.s_rho = g.s_rho; d.cs_r = g.cs_r; d.s_w = g.s_w; d.cs_w = g.cs_w;
        d.h = g.h; d.pm = g.pm; d.pn = g.pn;
        d.mrho = g.mrho; d.mu = g.mu; d.mv = g.mv;
        d.zeta = g.zeta; d.u = g.u; d.v = g.v; d.w = g.w; d.aks = g.aks;
        d.czeta = g.cur[0]; d.cu = g.cur[1]; d.cv = g.cur[2]; d.cw = g.cur[3]; d.caks = g.cur[4];
        d.nwc = cur_column(g);
        d.np = np_of(c, g.N);
        d.kh = c->cfg.kh[i];
        if (g.wd) {
            const int64_t nr = (int64_t)g.L * g.M, nu = (int64_t)(g.L - 1) * g.M;
            d.wdr = g.wd;
            d.wdu = g.wd + nr;
            d.wdv = g.wd + nr + nu;
            d.wdp = wd_plane(g);
        }
        for (int a = 0; a < kMaxGrid; a++)
            for (int k = 0; k < 2; k++) {
                d.scl[a][k] = g.scl[a][k];
                d.off[a][k] = g.off[a][k];
            }
    }
}

// the device grid table (c->dgrids) brought up to date with the grids
int refresh_tab(Ctx* c)
{
    DGrid tab[kMaxGrid];
    fill_tab(c, tab);
    if (std::memcmp(tab, c->tab_host, sizeof(tab)) != 0) {
        std::memcpy(c->tab_host, tab, sizeof(tab));
        HIPCHK(c, hipMemcpyAsync(c->dgrids, c->tab_host, sizeof(tab), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return RP_OK;
}

// na: the no-allocation scratch (rp_step_noalloc), or null.  With it the
// hand-over rows, the work-queue counter and the vertical walk's work rows
// come from that array, nothing is allocated and nothing is re-sorted.
int plan_step(Ctx* c, const double ex[3], const double ix[3], int ets, int t_b, int t_c, int t_f,
              int64_t step_index, int want_diag, int64_t q0, int64_t cnt, StepPlan& P,
              char* na = nullptr)
{
    int rc = check_step(c, t_b, t_c, t_f, step_index);
    if (rc) return rc;
    int maxN = 0;
    for (int i = 0; i < c->cfg.ngrid; i++) maxN = std::max(maxN, c->grids[i].N);
    const bool vt = c->cfg.vturb_on != 0;
    StepArgs& A = P.A;
    for (int k = 0; k < 3; k++) {
        A.ex[k] = ex[k];
        A.ix[k] = ix[k];
    }
    A.ets = ets;
    {
        // Lagrange weights of the quadratic through (ex, .) at ix(2)
        const double x = ix[1], x0 = ex[0], x1 = ex[1], x2 = ex[2];
        A.pw[0] = (x - x1) * (x - x2) / ((x0 - x1) * (x0 - x2));
        A.pw[1] = (x - x0) * (x - x2) / ((x1 - x0) * (x1 - x2));
        A.pw[2] = (x - x0) * (x - x1) / ((x2 - x0) * (x2 - x1));
    }
    A.slots.s[0] = t_b - 1;
    A.slots.s[1] = t_c - 1;
    A.slots.s[2] = t_f - 1;
    A.idt = c->cfg.idt;
    A.deltat = c->cfg.deltat;
    A.serr = c->cfg.serr;
    A.smth = c->cfg.smth;
    A.aksback = c->cfg.aksback;
    A.vturb_on = vt;
    A.hturb_on = c->cfg.hturb_on != 0;
    A.scheme = c->cfg.scheme;
    A.nsb = c->cfg.nsb;
    A.vertdist = c->cfg.vertdist;
    A.open_boundary = c->cfg.open_boundary != 0;
    A.ngrid = c->cfg.ngrid;
    A.np = np_of(c, maxN);
    A.np_mixed = 0;
    for (int i = 0; i < c->cfg.ngrid; i++) A.np_mixed |= np_of(c, c->grids[i].N) != A.np;
    A.loop = vt ? c->cfg.idt / (int)c->cfg.deltat : 0;
    A.q0 = q0;
    A.n = cnt;
    A.nmid = na ? kNaChunk : c->n;
    A.step_index = step_index;
    A.seed = c->cfg.seed;
    // the per-step samplers read the step's own depth (pBath) every step
    A.want_diag = want_diag != 0 ? 1 : ((c->bot_n == c->n || c->trc_n == c->n) ? 2 : 0);
    A.dev = c->use_dev ? c->dev : nullptr;
    A.ndev = c->ndev;
    A.active = c->active;
    A.work = na ? (unsigned long long*)na : c->work;
    A.capped = c->capped;
    A.mid = na ? (double*)(na + kNaHead) : c->mid;
    A.pert = c->pert;
    A.pert_nx = c->pert_nx;
    A.pert_ny = c->pert_ny;
    // the step kernels run over the launch's active set, built by
    // compact_count / sort_scan / compact_scatter (the no-allocation form
    // runs over all particles)
    A.act = na ? nullptr : c->act;
    A.act_n = na ? nullptr : c->act_n;
    if (vt && A.np < 3) return fail(c, RP_E_ARG, "smoothing spline needs sub*(N+1) >= 3");

    const int block = c->cfg.block > 0 ? c->cfg.block : 256;
    if (block % 64 || block > 256) return fail(c, RP_E_ARG, "block must be a multiple of 64, <= 256");
    P.block = block;
    P.vt = vt;
    P.grid_all = (cnt + block - 1) / block;
    const bool nested = c->cfg.ngrid > 1;
    const bool f32 = c->fbytes == 4;
    // the FP32 form's kernels with the wet/dry mask code where a grid has masks
    bool wet = false;
    for (int i = 0; i < c->cfg.ngrid; i++) wet = wet || c->grids[i].wd;
    if (f32 && wet && !nested) {
        P.k[0] = step_kernel<CurWet, false>; P.k[1] = advect_kernel<CurWet, false>;
        P.k[2] = vwalk_kernel<CurWet, false>;
    } else if (f32 && wet) {
        P.k[0] = step_kernel<CurWet, true>; P.k[1] = advect_kernel<CurWet, true>;
        P.k[2] = vwalk_kernel<CurWet, true>;
    } else if (f32 && !nested) {
        P.k[0] = step_kernel<Cur, false>; P.k[1] = advect_kernel<Cur, false>; P.k[2] = vwalk_kernel<Cur, false>;
    } else if (f32) {
        P.k[0] = step_kernel<Cur, true>; P.k[1] = advect_kernel<Cur, true>; P.k[2] = vwalk_kernel<Cur, true>;
    } else if (!nested) {
        P.k[0] = step_kernel<double, false>; P.k[1] = advect_kernel<double, false>; P.k[2] = vwalk_kernel<double, false>;
    } else {
        P.k[0] = step_kernel<double, true>; P.k[1] = advect_kernel<double, true>; P.k[2] = vwalk_kernel<double, true>;
    }
    // the FP32 form's vertical walk in the eigenbasis when every grid has
    // the same resampling size (a wave's transforms use one table)
    P.cur = f32;
    A.modal = nullptr;
    P.modal = false;
    if (f32 && vt && !c->vwalk_band && c->grids[0].modal) {
        bool same = true;
        for (int i = 1; i < c->cfg.ngrid; i++)
            same = same && c->grids[i].modal && c->grids[i].N == c->grids[0].N;
        if (same) {
            A.modal = c->grids[0].modal;
            P.modal = true;
            P.k[2] = nested ? vwalk_modal_kernel<true> : vwalk_modal_kernel<false>;
        }
    }
    P.grid_vt = P.grid_all;
    A.scratch = nullptr;
    if (vt) {
        // persistent grid for the vertical walk: exactly the resident blocks
        int& bpc = c->vwalk_blocks_per_cu[P.modal ? 2 : (f32 ? 0 : 1)][nested ? 1 : 0];
        if (bpc == 0) {
            int nb = 0;
            HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, P.k[2], block, vwalk_lds(P)));
            bpc = std::max(nb, 1);
        }
        P.vt_cap = (int64_t)c->ncu * bpc;
        if (f32) {
            // the FP32 rows (6 np per lane) are addressed through one buffer
            // descriptor, whose extent is below 2^31 bytes
            const int64_t per_block = 6LL * A.np * block * 4;
            P.vt_cap = std::max<int64_t>(1, std::min<int64_t>(P.vt_cap, 0x7FFFFFFFLL / per_block));
        }
        if (na) {
            // as many blocks as the array's row region holds
            const int64_t lane_bytes = f32 ? 6LL * A.np * 4 : (int64_t)vturb_rows(A.np, maxN + 1) * 8;
            const int64_t fit = (int64_t)kNaRowBytes / (lane_bytes * block);
            if (fit < 1)
                return fail(c, RP_E_ARG, "rp_step_noalloc: the vertical walk's work rows of one block (" +
                                             std::to_string(lane_bytes * block) + " bytes) exceed the " +
                                             std::to_string(kNaRowBytes) + "-byte row region");
            P.vt_cap = std::min(P.vt_cap, fit);
            A.scratch = (double*)(na + kNaHead + kNaMidBytes);
            A.scratch_bytes = (uint32_t)kNaRowBytes;
        }
        P.grid_vt = std::min<int64_t>(P.grid_all, P.vt_cap);
        if (!na) {
            const int64_t T = P.grid_vt * block;
            const size_t need = (size_t)vturb_rows(A.np, maxN + 1) * T * sizeof(double);
            if (c->scratch_bytes < need) {
                if ((rc = dalloc(c, (char**)&c->scratch, need))) return rc;
                c->scratch_bytes = need;
            }
            A.scratch = c->scratch;
            A.scratch_bytes = (uint32_t)std::min<size_t>(c->scratch_bytes, 0x7FFFFFFF);
        }
    }
    if (c->cfg.sort_interval > 0 && !na) {
        SortArgs K;
        if ((rc = sort_buffers(c, sort_keys(c, K)))) return rc;
    }
    P.cur = f32;
    if (f32) {
        // jobs of the current-field kernel: zeta and AKs at (b, c, f), u, v, w
        // at (b, b, c) on the first external step, (b, c, f) after it
        CurJobs& J = P.jobs;
        J.njobs = 0;
        for (int k = 0; k < 3; k++) J.pw[k] = A.pw[k];
        for (int gi = 0; gi < c->cfg.ngrid; gi++) {
            GridDev& g = c->grids[gi];
            void* src[5] = {g.zeta, g.u, g.v, g.w, g.aks};
            for (int f = 0; f < 5; f++) {
                // AKs columns padded to a multiple of four levels
                const int kin = f == 4 ? g.N + 1 : 1;
                const int kout = f == 4 ? cur_column(g) : 1;
                const int64_t npt = field_points(g, f) / kin * kout;
                if (!g.cur[f] && na) return fail(c, RP_E_STATE, "rp_step_noalloc: current fields not allocated");
                if (!g.cur[f] && (rc = dalloc(c, &g.cur[f], (size_t)npt))) return rc;
                const int j = J.njobs++;
                J.src[j] = (const float*)src[f];
                J.dst[j] = g.cur[f];
                J.n[j] = npt;
                J.kin[j] = kin;
                J.kout[j] = kout;
                const bool first = ets == 1 && f >= 1 && f <= 3;
                // u and v under wet/dry masks: each slot's dry points
                // contribute 0 (the masks' u then v sections follow the
                // rho section of each slot's plane)
                const int64_t nr = (int64_t)g.L * g.M, nu = (int64_t)(g.L - 1) * g.M;
                J.wd[j] = (g.wd && (f == 1 || f == 2)) ? g.wd + nr + (f == 2 ? nu : 0) : nullptr;
                J.wdp[j] = wd_plane(g);
                J.nk[j] = g.N;
                J.slot[j][0] = A.slots.s[0];
                J.slot[j][1] = first ? A.slots.s[0] : A.slots.s[1];
                J.slot[j][2] = first ? A.slots.s[1] : A.slots.s[2];
            }
        }
    }
    return refresh_tab(c);
}

// events for one timed entry, taken from the pool; the entry goes back to
// the pool (ev_free) until the caller moves it to ev_used
int take_events(Ctx* c, Ctx::StepEvents& ev)
{
    if (!c->ev_free.empty()) {
        ev = c->ev_free.back();
        c->ev_free.pop_back();
    } else {
        for (auto& e : ev.e) HIPCHK(c, hipEventCreate(&e));
    }
    ev.steps = 1;
    c->ev_free.push_back(ev);
    return RP_OK;
}

// the step's launches on the context stream (particle columns as they stand
// now, after any re-sort); ev records the phase boundaries when given
int enqueue_step(Ctx* c, StepPlan& P, const Ctx::StepEvents* ev)
{
    StepArgs& A = P.A;
    A.x = c->x; A.y = c->y; A.z = c->z; A.age = c->age; A.dob = c->dob;
    A.status = c->status; A.gid = c->gid; A.isin = c->isin;
    A.wd = c->wd; A.zeta = c->zeta; A.bath = c->bath;
    A.id = c->id;
    const dim3 bd(P.block);
    if (P.cur && P.jobs.njobs > 0) {
        // built once per step time and slot set (the scenarios of a step share them)
        const int first = A.ets == 1 ? 1 : 0;
        if (c->cur_gen != c->field_gen || c->cur_pw[0] != A.pw[0] || c->cur_pw[1] != A.pw[1] ||
            c->cur_pw[2] != A.pw[2] || c->cur_first != first ||
            c->cur_slots[0] != A.slots.s[0] || c->cur_slots[1] != A.slots.s[1] ||
            c->cur_slots[2] != A.slots.s[2]) {
            hipLaunchKernelGGL(current_fields, dim3(512, (unsigned)P.jobs.njobs), dim3(256), 0,
                               c->stream, P.jobs);
            HIPCHK(c, hipGetLastError());
            c->cur_gen = c->field_gen;
            for (int k = 0; k < 3; k++) c->cur_pw[k] = A.pw[k];
            c->cur_first = first;
            for (int k = 0; k < 3; k++) c->cur_slots[k] = A.slots.s[k];
        }
    }
    const size_t lds = c->fbytes == 4 ? (size_t)P.block * kLdsQ * 4 : 0;   // qty rows (FP32 form)
    if (A.act) {
        const int64_t nb = (A.n + 255) / 256;
        hipLaunchKernelGGL(compact_count, dim3((unsigned)nb), dim3(256), 0, c->stream, A, c->act_blk);
        hipLaunchKernelGGL(sort_scan, dim3(1), dim3(1024), 0, c->stream, c->act_blk, nb);
        hipLaunchKernelGGL(compact_scatter, dim3((unsigned)nb), dim3(256), 0, c->stream, A,
                           (const uint32_t*)c->act_blk, c->act, c->act_n);
        HIPCHK(c, hipGetLastError());
    }
    if (!P.vt) {
        hipLaunchKernelGGL(P.k[0], dim3((unsigned)P.grid_all), bd, lds, c->stream, c->dgrids, A);
        HIPCHK(c, hipGetLastError());
        if (ev) HIPCHK(c, hipEventRecord(ev->e[1], c->stream));
    } else {
        hipLaunchKernelGGL(P.k[1], dim3((unsigned)P.grid_all), bd, 0, c->stream, c->dgrids, A);
        HIPCHK(c, hipGetLastError());
        if (ev) HIPCHK(c, hipEventRecord(ev->e[1], c->stream));
        hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, c->stream, A.work);
        hipLaunchKernelGGL(P.k[2], dim3((unsigned)P.grid_vt), bd, vwalk_lds(P), c->stream, c->dgrids, A);
    }
    HIPCHK(c, hipGetLastError());
    if (ev) HIPCHK(c, hipEventRecord(ev->e[2], c->stream));
    return RP_OK;
}


// ---------------------------------------------------------------------------
// adjoint recording and backward sweep
// ---------------------------------------------------------------------------
size_t field_bytes(const Ctx* c, const GridDev& g, int f)
{
    const size_t fb = (size_t)c->fbytes * 4;   // four interleaved slots
    switch (f) {
    case 0: return (size_t)g.L * g.M * fb;
    case 1: return (size_t)(g.L - 1) * g.M * g.N * fb;
    case 2: return (size_t)g.L * (g.M - 1) * g.N * fb;
    default: return (size_t)g.L * g.M * (g.N + 1) * fb;
    }
}

// bytes of one saved particle state: x y z age (FP64), status gid isin (bytes)
size_t state_bytes(int64_t n) { return (size_t)n * (4 * sizeof(double) + 3); }

void copy_state(Ctx* c, char* buf, bool save)
{
    const int64_t n = c->n;
    double* d[4] = {c->x, c->y, c->z, c->age};
    void* b[3] = {c->status, c->gid, c->isin};
    for (int k = 0; k < 4; k++) {
        char* p = buf + (size_t)k * n * sizeof(double);
        if (save) (void)hipMemcpyAsync(p, d[k], n * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
        else (void)hipMemcpyAsync(d[k], p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    }
    for (int k = 0; k < 3; k++) {
        char* p = buf + (size_t)4 * n * sizeof(double) + (size_t)k * n;
        if (save) (void)hipMemcpyAsync(p, b[k], n, hipMemcpyDeviceToDevice, c->stream);
        else (void)hipMemcpyAsync(b[k], p, n, hipMemcpyDeviceToDevice, c->stream);
    }
}

void free_recording(Ctx* c)
{
    Ctx::AdjRec& R = c->adj;
    for (auto& sn : R.snaps) {
        for (int g = 0; g < kMaxGrid; g++) {
            for (int f = 0; f < 5; f++) dfree(sn.f[g][f]);
            dfree(sn.wd[g]);
        }
        dfree(sn.dtab);
    }
    for (auto& p : R.ckpt) dfree(p);
    for (auto& p : R.devs) dfree(p);
    R.snaps.clear();
    R.ckpt.clear();
    R.devs.clear();
    R.args.clear();
    R.snap.clear();
    R.dirty = true;
}

// copy of the current hydro slots of every grid, with its device grid table
int take_snapshot(Ctx* c)
{
    Ctx::AdjSnap sn;
    std::memset(&sn, 0, sizeof(sn));
    DGrid tab[kMaxGrid];
    std::memcpy(tab, c->tab_host, sizeof(tab));
    for (int g = 0; g < c->cfg.ngrid; g++) {
        GridDev& G = c->grids[g];
        void* src[5] = {G.zeta, G.u, G.v, G.w, G.aks};
        for (int f = 0; f < 5; f++) {
            const size_t bytes = field_bytes(c, G, f);
            if (dalloc(c, (char**)&sn.f[g][f], bytes)) return RP_E_ALLOC;
            HIPCHK(c, hipMemcpyAsync(sn.f[g][f], src[f], bytes, hipMemcpyDeviceToDevice, c->stream));
        }
        tab[g].zeta = sn.f[g][0]; tab[g].u = sn.f[g][1]; tab[g].v = sn.f[g][2];
        tab[g].w = sn.f[g][3]; tab[g].aks = sn.f[g][4];
        if (G.wd) {
            const int64_t plane = wd_plane(G), nr = (int64_t)G.L * G.M, nu = (int64_t)(G.L - 1) * G.M;
            if (dalloc(c, (char**)&sn.wd[g], (size_t)(3 * plane))) return RP_E_ALLOC;
            HIPCHK(c, hipMemcpyAsync(sn.wd[g], G.wd, (size_t)(3 * plane), hipMemcpyDeviceToDevice, c->stream));
            tab[g].wdr = sn.wd[g];
            tab[g].wdu = sn.wd[g] + nr;
            tab[g].wdv = sn.wd[g] + nr + nu;
        }
    }
    if (dalloc(c, &sn.dtab, kMaxGrid)) return RP_E_ALLOC;
    HIPCHK(c, hipMemcpyAsync(sn.dtab, tab, sizeof(tab), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->adj.snaps.push_back(sn);
    c->adj.dirty = false;
    return RP_OK;
}

using AdjFwdFn = void (*)(const DGrid*, StepArgs);
using AdjStepFn = void (*)(const DGrid*, StepArgs, AdjArgs);

AdjFwdFn adj_forward_fn(const Ctx* c)
{
    const bool f32 = c->fbytes == 4, nested = c->cfg.ngrid > 1;
    if (f32) return nested ? adj_forward_kernel<float, true> : adj_forward_kernel<float, false>;
    return nested ? adj_forward_kernel<double, true> : adj_forward_kernel<double, false>;
}

AdjStepFn adj_step_fn(const Ctx* c)
{
    const bool f32 = c->fbytes == 4, nested = c->cfg.ngrid > 1;
    if (f32) return nested ? adj_step_kernel<float, true> : adj_step_kernel<float, false>;
    return nested ? adj_step_kernel<double, true> : adj_step_kernel<double, false>;
}

// persistent grid of a queue kernel and the scratch it needs (rows x elem bytes per lane)
int queue_grid(Ctx* c, const void* fn, int block, size_t dyn_lds, int64_t n, int64_t& grid)
{
    int nb = 0;
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, dyn_lds));
    grid = std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, (int64_t)c->ncu * std::max(nb, 1)));
    return RP_OK;
}

int ensure_scratch(Ctx* c, size_t need)
{
    if (c->scratch_bytes < need) {
        const int rc = dalloc(c, (char**)&c->scratch, need);
        if (rc) return rc;
        c->scratch_bytes = need;
    }
    return RP_OK;
}

int max_levels(const Ctx* c)
{
    int maxN = 0;
    for (int i = 0; i < c->cfg.ngrid; i++) maxN = std::max(maxN, c->grids[i].N);
    return maxN;
}

// one forward step of a recorded run (spline on the reference-order path)
int adj_forward_launch(Ctx* c, StepArgs A, const DGrid* tab)
{
    const int block = 256;
    const AdjFwdFn fn = adj_forward_fn(c);
    int64_t grid = 0;
    int rc = queue_grid(c, (const void*)fn, block, 0, A.n, grid);
    if (rc) return rc;
    if (A.vturb_on) {
        const int64_t T = grid * block;
        if ((rc = ensure_scratch(c, (size_t)vturb_rows(A.np, max_levels(c) + 1) * T * sizeof(double))))
            return rc;
        A.scratch = c->scratch;
    }
    A.x = c->x; A.y = c->y; A.z = c->z; A.age = c->age; A.dob = c->dob;
    A.status = c->status; A.gid = c->gid; A.isin = c->isin;
    A.wd = c->wd; A.zeta = c->zeta; A.bath = c->bath;
    A.id = c->id;
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, c->stream, c->work);
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(block), 0, c->stream, tab, A);
    HIPCHK(c, hipGetLastError());
    return RP_OK;
}

// recorded rp_step: checkpoint, field snapshot and deviates as needed, then the step
int record_step(Ctx* c, StepArgs A)
{
    Ctx::AdjRec& R = c->adj;
    const int64_t k = (int64_t)R.args.size();
    if (k % R.K == 0) {
        char* buf = nullptr;
        if (dalloc(c, &buf, state_bytes(c->n))) return RP_E_ALLOC;
        copy_state(c, buf, true);
        R.ckpt.push_back(buf);
    }
    if (R.dirty) {
        const int rc = take_snapshot(c);
        if (rc) return rc;
    }
    double* dv = nullptr;
    if (A.dev) {
        if (dalloc(c, &dv, (size_t)c->n * c->ndev)) return RP_E_ALLOC;
        HIPCHK(c, hipMemcpyAsync(dv, A.dev, (size_t)c->n * c->ndev * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
    }
    R.devs.push_back(dv);
    R.snap.push_back((int)R.snaps.size() - 1);
    R.args.push_back(A);
    return adj_forward_launch(c, A, c->dgrids);
}

}  // namespace

extern "C" {

int rp_create(void** out, const rp_config* cfg)
{
    if (!out || !cfg) return RP_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RP_E_NODEV;
    Ctx* c = new Ctx();
    c->cfg = *cfg;
    if (const char* k = std::getenv("RP_FAIL_ALLOC_AFTER")) c->alloc_budget = std::max<int64_t>(0, std::atoll(k));
    if (const char* f = std::getenv("RP_VWALK_FORM")) c->vwalk_band = std::strcmp(f, "band") == 0;
    std::memset(c->tab_host, 0, sizeof(c->tab_host));
    if (cfg->precision != RP_PREC_FP32_FIELDS && cfg->precision != RP_PREC_FP64) {
        delete c;
        return RP_E_ARG;
    }
    if (cfg->ngrid < 1 || cfg->ngrid > RP_MAXGRID || cfg->idt <= 0 || cfg->dt <= 0 ||
        (cfg->vturb_on && (cfg->deltat < 1.0 || cfg->sub <= 0.0))) {
        delete c;
        return RP_E_ARG;
    }
    c->fbytes = cfg->precision == RP_PREC_FP64 ? 8 : 4;
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
            hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RP_E_HIP;
    }
    if (upload_p0_tables() != 0 || na_setup(c->device, &c->na) != 0) {
        rp_destroy(c);
        return RP_E_HIP;
    }
    if (dalloc(c, &c->active, 1) || dalloc(c, &c->capped, 1) || dalloc(c, &c->dgrids, kMaxGrid) ||
        dalloc(c, &c->work, 1) || dalloc(c, &c->act_n, 1)) {
        rp_destroy(c);
        return RP_E_ALLOC;
    }
    if (hipMemsetAsync(c->active, 0, sizeof(unsigned long long), c->stream) != hipSuccess ||
        hipMemsetAsync(c->capped, 0, sizeof(unsigned), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        rp_destroy(c);
        return RP_E_HIP;
    }
    *out = c;
    return RP_OK;
}

int rp_set_grid(void* ctx, int32_t ng, const rp_grid* gh)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !gh) return RP_E_ARG;
    if (ng < 1 || ng > c->cfg.ngrid) return fail(c, RP_E_ARG, "grid number out of range");
    if (gh->L < 4 || gh->M < 4 || gh->N < 1 || gh->vtransform < 1 || gh->vtransform > 3 ||
        !gh->s_rho || !gh->cs_r || !gh->s_w || !gh->cs_w || !gh->h || !gh->pm || !gh->pn ||
        !gh->mask_rho || !gh->mask_u || !gh->mask_v)
        return fail(c, RP_E_ARG, "invalid grid description");
    GridDev& g = c->grids[ng - 1];
    free_grid(g);
    g.L = gh->L; g.M = gh->M; g.N = gh->N;
    g.vtransform = gh->vtransform;
    g.hc = gh->hc;
    const size_t nr = (size_t)g.L * g.M, nu = (size_t)(g.L - 1) * g.M, nv = (size_t)g.L * (g.M - 1);
    int rc = 0;
    rc |= dalloc(c, &g.s_rho, g.N); rc |= dalloc(c, &g.cs_r, g.N);
    rc |= dalloc(c, &g.s_w, g.N + 1); rc |= dalloc(c, &g.cs_w, g.N + 1);
    rc |= dalloc(c, &g.h, nr); rc |= dalloc(c, &g.pm, nr); rc |= dalloc(c, &g.pn, nr);
    rc |= dalloc(c, &g.mrho, nr); rc |= dalloc(c, &g.mu, nu); rc |= dalloc(c, &g.mv, nv);
    const size_t fb = c->fbytes;
    rc |= dalloc(c, (char**)&g.zeta, nr * 4 * fb);
    rc |= dalloc(c, (char**)&g.u, nu * g.N * 4 * fb);
    rc |= dalloc(c, (char**)&g.v, nv * g.N * 4 * fb);
    rc |= dalloc(c, (char**)&g.w, nr * (g.N + 1) * 4 * fb);
    rc |= dalloc(c, (char**)&g.aks, nr * (g.N + 1) * 4 * fb);
    if (c->fbytes == 4) {
        // the FP32 form's current fields (one value per point at ix(2)),
        // allocated with the grid so that no step needs to
        for (int f = 0; f < 5; f++) {
            const int kin = f == 4 ? g.N + 1 : 1;
            const int kout = f == 4 ? cur_column(g) : 1;
            rc |= dalloc(c, &g.cur[f], (size_t)(field_points(g, f) / kin * kout));
        }
    }
    if (rc) {
        free_grid(g);
        return RP_E_ALLOC;
    }
    // zero the fields so unused slot lanes hold finite values
    HIPCHK(c, hipMemsetAsync(g.zeta, 0, nr * 4 * fb, c->stream));
    HIPCHK(c, hipMemsetAsync(g.u, 0, nu * g.N * 4 * fb, c->stream));
    HIPCHK(c, hipMemsetAsync(g.v, 0, nv * g.N * 4 * fb, c->stream));
    HIPCHK(c, hipMemsetAsync(g.w, 0, nr * (g.N + 1) * 4 * fb, c->stream));
    HIPCHK(c, hipMemsetAsync(g.aks, 0, nr * (g.N + 1) * 4 * fb, c->stream));
    if ((rc = put(c, g.s_rho, gh->s_rho, g.N * 8)) || (rc = put(c, g.cs_r, gh->cs_r, g.N * 8)) ||
        (rc = put(c, g.s_w, gh->s_w, (g.N + 1) * 8)) || (rc = put(c, g.cs_w, gh->cs_w, (g.N + 1) * 8)) ||
        (rc = put(c, g.h, gh->h, nr * 8)) || (rc = put(c, g.pm, gh->pm, nr * 8)) ||
        (rc = put(c, g.pn, gh->pn, nr * 8)))
        return rc;
    std::vector<int8_t> m8(nr);
    auto conv = [&](const int32_t* src, size_t cnt, int8_t* dst) -> int {
        for (size_t i = 0; i < cnt; i++) {
            if (src[i] != 0 && src[i] != 1) return -1;
            m8[i] = (int8_t)src[i];
        }
        hipError_t e = hipMemcpy(dst, m8.data(), cnt, hipMemcpyHostToDevice);
        return e == hipSuccess ? 0 : -2;
    };
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (conv(gh->mask_rho, nr, g.mrho) || conv(gh->mask_u, nu, g.mu) || conv(gh->mask_v, nv, g.mv))
        return fail(c, RP_E_ARG, "masks must be 0/1");
    for (int i = 0; i < kMaxGrid; i++)
        for (int k = 0; k < 2; k++) {
            g.scl[i][k] = gh->scl[i][k];
            g.off[i][k] = gh->off[i][k];
        }
    for (int s = 0; s < 3; s++) g.slot_set[s] = false;
    dfree(g.modal);
    if (c->fbytes == 4 && c->cfg.vturb_on && modal_applies(np_of(c, g.N) - 2)) {
        ModalTab T;
        if (build_modal_tab(np_of(c, g.N) - 2, T) != RP_OK) return fail(c, RP_E_ARG, "eigenbasis tables");
        if ((rc = dalloc(c, &g.modal, 1)) || (rc = put(c, g.modal, &T, sizeof T))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    g.set = true;
    return RP_OK;
}

int rp_set_hydro_slot(void* ctx, int32_t ng, int32_t slot, const double* zeta, const double* u,
                      const double* v, const double* w, const double* aks, const double* su,
                      const double* sv)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (ng < 1 || ng > c->cfg.ngrid || !c->grids[ng - 1].set)
        return fail(c, RP_E_STATE, "rp_set_hydro_slot: grid not set");
    if (slot < 1 || slot > 3) return fail(c, RP_E_ARG, "slot must be 1..3");
    if (!zeta || !u || !v || !w || !aks) return fail(c, RP_E_ARG, "null field");
    GridDev& g = c->grids[ng - 1];
    const int s = slot - 1;
    int rc;
    c->adj.dirty = true;
    c->field_gen++;
    const int gi = ng - 1;
    if ((rc = upload_field(c, gi, 0, g.zeta, zeta, nullptr, g.L, g.M, 1, s)) ||
        (rc = upload_field(c, gi, 1, g.u, u, su, g.L - 1, g.M, g.N, s)) ||
        (rc = upload_field(c, gi, 2, g.v, v, sv, g.L, g.M - 1, g.N, s)) ||
        (rc = upload_field(c, gi, 3, g.w, w, nullptr, g.L, g.M, g.N + 1, s)) ||
        (rc = upload_field(c, gi, 4, g.aks, aks, nullptr, g.L, g.M, g.N + 1, s)))
        return rc;
    g.slot_set[s] = true;
    return RP_OK;
}

int rp_set_particles(void* ctx, int64_t n, const double* x, const double* y, const double* z,
                     const double* age, const double* dob, const double* status, const int32_t* gid,
                     const uint8_t* isin)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (n < 0 || n > (int64_t)0xFFFFFFFFLL) return fail(c, RP_E_RANGE, "particle count out of range");
    if (n > 0 && (!x || !y || !z || !age || !dob || !status || !gid || !isin))
        return fail(c, RP_E_ARG, "null particle column");
    std::vector<int8_t> st(n), gd(n);
    std::vector<uint32_t> ids(n);
    for (int64_t i = 0; i < n; i++) {
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]) ||
            !std::isfinite(age[i]) || !std::isfinite(dob[i]))
            return fail(c, RP_E_ARG, "non-finite particle column value at particle " + std::to_string(i));
        if (status_code(status[i], &st[i])) return fail(c, RP_E_ARG, "unknown status value");
        if (gid[i] < 1 || gid[i] > c->cfg.ngrid) return fail(c, RP_E_ARG, "grid id out of range");
        gd[i] = (int8_t)gid[i];
        ids[i] = (uint32_t)i;
    }
    int rc = 0;
    rc |= dalloc(c, &c->x, n); rc |= dalloc(c, &c->y, n); rc |= dalloc(c, &c->z, n);
    rc |= dalloc(c, &c->age, n); rc |= dalloc(c, &c->dob, n);
    rc |= dalloc(c, &c->wd, n); rc |= dalloc(c, &c->zeta, n); rc |= dalloc(c, &c->bath, n);
    rc |= dalloc(c, &c->status, n); rc |= dalloc(c, &c->gid, n); rc |= dalloc(c, &c->isin, n);
    rc |= dalloc(c, &c->id, n);
    rc |= dalloc(c, &c->mid, 4 * (size_t)n);
    rc |= dalloc(c, &c->act, (size_t)n);
    rc |= dalloc(c, &c->act_blk, (size_t)((n + 255) / 256 + 1));
    if (rc) return RP_E_ALLOC;
    c->n = n;
    c->bot_n = -1;
    c->trc_n = -1;
    c->steps_done = 0;
    c->ranges_intact = true;
    free_recording(c);
    c->adj.K = 0;
    c->scen_first.clear();
    c->scen_count.clear();
    c->scen_steps.clear();
    if (n == 0) return RP_OK;
    if ((rc = put(c, c->x, x, n * 8)) || (rc = put(c, c->y, y, n * 8)) || (rc = put(c, c->z, z, n * 8)) ||
        (rc = put(c, c->age, age, n * 8)) || (rc = put(c, c->dob, dob, n * 8)) ||
        (rc = put(c, c->status, st.data(), n)) || (rc = put(c, c->gid, gd.data(), n)) ||
        (rc = put(c, c->isin, isin, n)) || (rc = put(c, c->id, ids.data(), n * 4)))
        return rc;
    std::vector<double> fill(n, 9999.0);
    if ((rc = put(c, c->wd, fill.data(), n * 8)) || (rc = put(c, c->zeta, fill.data(), n * 8)) ||
        (rc = put(c, c->bath, fill.data(), n * 8)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RP_OK;
}

int rp_set_deviates(void* ctx, const double* normals, int64_t n, int32_t ndev)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (!normals) {
        c->use_dev = false;
        return RP_OK;
    }
    const int need = 2 + (c->cfg.vturb_on ? c->cfg.idt / (int)c->cfg.deltat : 0);
    if (n != c->n || ndev < need) return fail(c, RP_E_ARG, "deviate block has the wrong shape");
    if (first_nonfinite(normals, (size_t)n * ndev) >= 0) return fail(c, RP_E_ARG, "non-finite deviate");
    int rc = dalloc(c, &c->dev, (size_t)n * ndev);
    if (rc) return rc;
    if ((rc = put(c, c->dev, normals, (size_t)n * ndev * 8))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ndev = ndev;
    c->use_dev = true;
    return RP_OK;
}

int rp_step(void* ctx, const double ex[3], const double ix[3], int32_t ets, int32_t t_b,
            int32_t t_c, int32_t t_f, int64_t step_index, int32_t want_diag)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !ex || !ix) return RP_E_ARG;
    if (c->n == 0) return check_step(c, t_b, t_c, t_f, step_index);
    StepPlan P;
    int rc = plan_step(c, ex, ix, ets, t_b, t_c, t_f, step_index, want_diag, 0, c->n, P);
    if (rc) return rc;
    Ctx::StepEvents ev;
    if ((rc = take_events(c, ev))) return rc;
    if (c->adj.K > 0) {
        // recorded run: no re-sort (the adjoint indexes the stored order)
        HIPCHK(c, hipEventRecord(ev.e[0], c->stream));
        if ((rc = record_step(c, P.A))) return rc;
        HIPCHK(c, hipEventRecord(ev.e[1], c->stream));
        HIPCHK(c, hipEventRecord(ev.e[2], c->stream));
        c->steps_done++;
        c->ev_free.pop_back();
        c->ev_used.push_back(ev);
        return RP_OK;
    }
    HIPCHK(c, hipEventRecord(ev.e[0], c->stream));
    if (c->cfg.sort_interval > 0 && c->steps_done % c->cfg.sort_interval == 0) {
        if ((rc = sort_range(c, 0, c->n))) return rc;
        c->ranges_intact = c->scen_first.size() <= 1;
    }
    c->steps_done++;
    if ((rc = enqueue_step(c, P, &ev))) return rc;
    c->ev_free.pop_back();
    c->ev_used.push_back(ev);
    return RP_OK;
}

int rp_step_noalloc(void* ctx, const double ex[3], const double ix[3], int32_t ets, int32_t t_b,
                    int32_t t_c, int32_t t_f, int64_t step_index, int32_t want_diag)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !ex || !ix) return RP_E_ARG;
    if (!c->na) return fail(c, RP_E_STATE, "rp_step_noalloc: no scratch address for this device");
    if (c->adj.K > 0)
        return fail(c, RP_E_STATE, "rp_step_noalloc: a recorded run (rp_adjoint_record) stores "
                                   "checkpoints; use rp_step");
    if (c->n == 0) return check_step(c, t_b, t_c, t_f, step_index);
    StepPlan P;
    const int64_t first = std::min<int64_t>(c->n, kNaChunk);
    int rc = plan_step(c, ex, ix, ets, t_b, t_c, t_f, step_index, want_diag, 0, first, P, c->na);
    if (rc) return rc;
    // launch pairs of at most kNaChunk particles; each zeroes the shared
    // work counter before its vertical walk and writes the hand-over rows
    // before reading them
    for (int64_t q0 = 0; q0 < c->n; q0 += kNaChunk) {
        P.A.q0 = q0;
        P.A.n = std::min<int64_t>(kNaChunk, c->n - q0);
        P.grid_all = (P.A.n + P.block - 1) / P.block;
        P.grid_vt = std::min<int64_t>(P.grid_all, P.vt_cap);
        if ((rc = enqueue_step(c, P, nullptr))) return rc;
    }
    c->steps_done++;
    return RP_OK;
}

int rp_set_tuning(void* ctx, int32_t block, int32_t sort_interval, int32_t tile_x, int32_t tile_y)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (block != 0 && (block < 64 || block > 256 || block % 64))
        return fail(c, RP_E_ARG, "block must be 0 or a multiple of 64 in 64..256");
    if (sort_interval < 0) return fail(c, RP_E_ARG, "negative sort interval");
    if (tile_x < 1 || tile_y < 1 || tile_x > 4096 || tile_y > 4096)
        return fail(c, RP_E_ARG, "sort tile must be 1..4096 cells");
    c->cfg.block = block;
    c->cfg.sort_interval = sort_interval;
    c->sort_tile[0] = tile_x;
    c->sort_tile[1] = tile_y;
    for (auto& f : c->vwalk_blocks_per_cu) f[0] = f[1] = 0;
    // the next step sorts
    c->steps_done = 0;
    for (auto& k : c->scen_steps) k = 0;
    return RP_OK;
}

int rp_set_scenarios(void* ctx, int32_t nscen, const int64_t* counts)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || nscen < 1 || !counts) return RP_E_ARG;
    int64_t tot = 0;
    for (int32_t s = 0; s < nscen; s++) {
        if (counts[s] < 0) return fail(c, RP_E_ARG, "negative scenario count");
        tot += counts[s];
    }
    if (tot != c->n) return fail(c, RP_E_ARG, "scenario counts must add up to the particle count");
    c->scen_first.assign(nscen, 0);
    c->scen_count.assign(counts, counts + nscen);
    c->scen_steps.assign(nscen, 0);
    for (int32_t s = 1; s < nscen; s++) c->scen_first[s] = c->scen_first[s - 1] + counts[s - 1];
    return RP_OK;
}

int rp_step_scenario(void* ctx, int32_t scen, const double ex[3], const double ix[3], int32_t ets,
                     int32_t t_b, int32_t t_c, int32_t t_f, int64_t step_index, int32_t want_diag)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !ex || !ix) return RP_E_ARG;
    if (scen < 0 || scen >= (int32_t)c->scen_first.size())
        return fail(c, RP_E_ARG, "scenario index out of range (rp_set_scenarios)");
    if (c->adj.K > 0) return fail(c, RP_E_STATE, "rp_step_scenario: a recorded run steps with rp_step");
    const int64_t q0 = c->scen_first[scen], cnt = c->scen_count[scen];
    if (cnt == 0) return check_step(c, t_b, t_c, t_f, step_index);
    StepPlan P;
    int rc = plan_step(c, ex, ix, ets, t_b, t_c, t_f, step_index, want_diag, q0, cnt, P);
    if (rc) return rc;
    Ctx::StepEvents ev;
    if ((rc = take_events(c, ev))) return rc;
    HIPCHK(c, hipEventRecord(ev.e[0], c->stream));
    if (!c->ranges_intact) {
        if ((rc = restore_order(c))) return rc;
    }
    if (c->cfg.sort_interval > 0 && c->scen_steps[scen] % c->cfg.sort_interval == 0) {
        if ((rc = sort_range(c, q0, cnt))) return rc;
    }
    c->scen_steps[scen]++;
    if ((rc = enqueue_step(c, P, &ev))) return rc;
    c->ev_free.pop_back();
    c->ev_used.push_back(ev);
    return RP_OK;
}

int rp_step_batch(void* ctx, int32_t nsteps, const double* ex, const double* ix, const int32_t* ets,
                  const int32_t* slots, const int64_t* step_index, const int32_t* want_diag,
                  int32_t use_graph)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || nsteps < 0 || (nsteps > 0 && (!ex || !ix || !ets || !slots || !step_index)))
        return RP_E_ARG;
    if (nsteps == 0) return RP_OK;
    if (c->adj.K > 0) return fail(c, RP_E_STATE, "rp_step_batch: a recorded run steps with rp_step");
    if (c->n == 0) {
        for (int32_t i = 0; i < nsteps; i++) {
            const int rc = check_step(c, slots[3 * i], slots[3 * i + 1], slots[3 * i + 2], step_index[i]);
            if (rc) return rc;
        }
        return RP_OK;
    }
    std::vector<StepPlan> plans(nsteps);
    for (int32_t i = 0; i < nsteps; i++) {
        const int rc = plan_step(c, ex + 3 * i, ix + 3 * i, ets[i], slots[3 * i], slots[3 * i + 1],
                                 slots[3 * i + 2], step_index[i], want_diag ? want_diag[i] : 0, 0,
                                 c->n, plans[i]);
        if (rc) return rc;
    }
    if (c->n > 0 && c->cfg.sort_interval > 0) {
        SortArgs K;
        const int rc = sort_buffers(c, sort_keys(c, K));
        if (rc) return rc;
    }
    Ctx::StepEvents ev;
    int rc = take_events(c, ev);
    if (rc) return rc;
    ev.steps = nsteps;
    HIPCHK(c, hipEventRecord(ev.e[0], c->stream));
    auto enqueue_all = [&]() -> int {
        for (int32_t i = 0; i < nsteps; i++) {
            if (c->cfg.sort_interval > 0 && c->steps_done % c->cfg.sort_interval == 0) {
                const int r = sort_range(c, 0, c->n);
                if (r) return r;
                c->ranges_intact = c->scen_first.size() <= 1;
            }
            c->steps_done++;
            const int r = enqueue_step(c, plans[i], nullptr);
            if (r) return r;
        }
        return RP_OK;
    };
    if (use_graph) {
        // the previous batch may still run from the executable graph
        if (c->gexec) HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_all();
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(c, RP_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        // batches of the same shape (steps, re-sorts) differ only in kernel
        // arguments: the executable graph is updated in place; otherwise it
        // is instantiated anew
        bool updated = false;
        if (c->gexec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult ur = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(c->gexec, graph, &bad, &ur) == hipSuccess &&
                      ur == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(c->gexec);
                c->gexec = nullptr;
            }
        }
        hipError_t ei = hipSuccess;
        if (!updated) ei = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            c->gexec = nullptr;
            return fail(c, RP_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    } else {
        if ((rc = enqueue_all())) return rc;
    }
    HIPCHK(c, hipEventRecord(ev.e[1], c->stream));
    HIPCHK(c, hipEventRecord(ev.e[2], c->stream));
    c->ev_free.pop_back();
    c->ev_used.push_back(ev);
    return RP_OK;
}

int rp_set_perturbation(void* ctx, int32_t nx, int32_t ny, const double* coef)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (!coef) {
        dfree(c->pert);
        c->pert_nx = c->pert_ny = 0;
        return RP_OK;
    }
    if (nx < 4 || ny < 4 || nx > 64 || ny > 64)
        return fail(c, RP_E_ARG, "perturbation control net must be 4..64 points each way");
    if (c->adj.K > 0 && (nx != c->pert_nx || ny != c->pert_ny || !c->pert))
        return fail(c, RP_E_STATE, "perturbation shape fixed while recording");
    const size_t cnt = 2 * (size_t)nx * ny;
    if (!c->pert || nx != c->pert_nx || ny != c->pert_ny) {
        if (dalloc(c, &c->pert, cnt)) return RP_E_ALLOC;
    }
    c->pert_nx = nx;
    c->pert_ny = ny;
    HIPCHK(c, hipMemcpyAsync(c->pert, coef, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RP_OK;
}

int rp_adjoint_record(void* ctx, int32_t ckpt_every)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || ckpt_every < 0) return RP_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_recording(c);
    c->adj.K = ckpt_every;
    if (ckpt_every > 0) {
        if (c->n == 0) return fail(c, RP_E_STATE, "rp_adjoint_record: no particles");
        if (!c->adj.active && dalloc(c, &c->adj.active, 1)) return RP_E_ALLOC;
    }
    return RP_OK;
}

int rp_adjoint(void* ctx, int64_t n, const double* lam_x, const double* lam_y, const double* lam_z,
               double* g_x, double* g_y, double* g_z, double* g_pert)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !lam_x || !lam_y || !lam_z || !g_x || !g_y || !g_z) return RP_E_ARG;
    Ctx::AdjRec& R = c->adj;
    if (R.K <= 0 || R.args.empty()) return fail(c, RP_E_STATE, "rp_adjoint: no recorded steps");
    if (n != c->n) return fail(c, RP_E_ARG, "particle count differs from rp_set_particles");
    if (g_pert && !c->pert) return fail(c, RP_E_STATE, "rp_adjoint: no perturbation set");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t nsteps = (int64_t)R.args.size();
    const int K = R.K;
    // lambda in stored order
    std::vector<uint32_t> ids(n);
    HIPCHK(c, hipMemcpy(ids.data(), c->id, n * 4, hipMemcpyDeviceToHost));
    std::vector<double> hl(3 * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        hl[i] = lam_x[ids[i]];
        hl[n + i] = lam_y[ids[i]];
        hl[2 * n + i] = lam_z[ids[i]];
    }
    double *lam = nullptr, *gp = nullptr;
    char *fin = nullptr, *seg = nullptr;
    const size_t npert = c->pert ? 2 * (size_t)c->pert_nx * c->pert_ny : 0;
    // per saved step: x y z (FP64) + status gid isin
    const size_t seg_step = (size_t)n * (3 * sizeof(double) + 3);
    auto cleanup = [&]() { dfree(lam); dfree(gp); dfree(fin); dfree(seg); };
    if (dalloc(c, &lam, 3 * (size_t)n) || (g_pert && dalloc(c, &gp, npert)) ||
        dalloc(c, &fin, state_bytes(n)) || dalloc(c, &seg, seg_step * K)) {
        cleanup();
        return RP_E_ALLOC;
    }
    int rc = RP_OK;
    auto run = [&]() -> int {
        HIPCHK(c, hipMemcpyAsync(lam, hl.data(), hl.size() * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        if (gp) hipLaunchKernelGGL(adj_fill, dim3(8), dim3(256), 0, c->stream, gp, (int64_t)npert, 0.0);
        copy_state(c, fin, true);
        const int block = 256;
        const AdjStepFn sfn = adj_step_fn(c);
        const size_t lds = gp ? npert * sizeof(double) : 0;
        int64_t grid_b = 0;
        int r = queue_grid(c, (const void*)sfn, block, lds, n, grid_b);
        if (r) return r;
        const int64_t Tb = grid_b * block;
        const int np = R.args[0].np;
        auto state_at = [&](int64_t j) {
            AdjState s;
            char* b = seg + seg_step * j;
            s.x = (double*)b;
            s.y = s.x + n;
            s.z = s.y + n;
            s.status = (int8_t*)(s.z + n);
            s.gid = s.status + n;
            s.isin = (uint8_t*)(s.gid + n);
            return s;
        };
        const int64_t nseg = (nsteps + K - 1) / K;
        const unsigned gsave = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
        for (int64_t sgi = nseg - 1; sgi >= 0; sgi--) {
            const int64_t k0 = sgi * K, k1 = std::min<int64_t>(k0 + K, nsteps);
            copy_state(c, R.ckpt[sgi], false);
            for (int64_t k = k0; k < k1; k++) {
                const AdjState st = state_at(k - k0);
                hipLaunchKernelGGL(adj_save_kernel, dim3(gsave), dim3(256), 0, c->stream, st, c->x,
                                   c->y, c->z, c->status, c->gid, c->isin, n);
                if (k + 1 < k1) {
                    StepArgs A = R.args[k];
                    A.dev = R.devs[k];
                    A.active = R.active;
                    if ((r = adj_forward_launch(c, A, R.snaps[R.snap[k]].dtab))) return r;
                }
            }
            for (int64_t k = k1 - 1; k >= k0; k--) {
                StepArgs A = R.args[k];
                const AdjState st = state_at(k - k0);
                A.x = st.x; A.y = st.y; A.z = st.z;
                A.status = st.status; A.gid = st.gid; A.isin = st.isin;
                A.age = c->age; A.dob = c->dob; A.id = c->id;
                A.wd = c->wd; A.zeta = c->zeta; A.bath = c->bath;
                A.want_diag = 0;
                A.dev = R.devs[k];
                A.active = R.active;
                if (A.vturb_on) {
                    const size_t need = (size_t)vturb_rows(np, max_levels(c) + 1) * Tb * sizeof(Dual<3>);
                    if ((r = ensure_scratch(c, need))) return r;
                }
                A.scratch = c->scratch;
                AdjArgs J;
                J.lx = lam;
                J.ly = lam + n;
                J.lz = lam + 2 * n;
                J.gpert = gp;
                hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, c->stream, c->work);
                hipLaunchKernelGGL(sfn, dim3((unsigned)grid_b), dim3(block), lds, c->stream,
                                   R.snaps[R.snap[k]].dtab, A, J);
                HIPCHK(c, hipGetLastError());
            }
        }
        copy_state(c, fin, false);
        HIPCHK(c, hipMemcpyAsync(hl.data(), lam, hl.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        if (gp) HIPCHK(c, hipMemcpyAsync(g_pert, gp, npert * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return RP_OK;
    };
    rc = run();
    cleanup();
    if (rc) return rc;
    for (int64_t i = 0; i < n; i++) {
        g_x[ids[i]] = hl[i];
        g_y[ids[i]] = hl[n + i];
        g_z[ids[i]] = hl[2 * n + i];
    }
    return RP_OK;
}

int rp_get_particles(void* ctx, int64_t n, double* x, double* y, double* z, double* age,
                     double* status, int32_t* gid, uint8_t* isin, double* wd, double* zeta,
                     double* bath)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (n != c->n) return fail(c, RP_E_ARG, "particle count differs from rp_set_particles");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n == 0) return RP_OK;
    std::vector<uint32_t> ids(n);
    HIPCHK(c, hipMemcpy(ids.data(), c->id, n * 4, hipMemcpyDeviceToHost));
    std::vector<double> tmp(n);
    auto getd = [&](double* dst, const double* src) -> int {
        if (!dst) return 0;
        if (hipMemcpy(tmp.data(), src, n * 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        for (int64_t i = 0; i < n; i++) dst[ids[i]] = tmp[i];
        return 0;
    };
    if (getd(x, c->x) || getd(y, c->y) || getd(z, c->z) || getd(age, c->age) || getd(wd, c->wd) ||
        getd(zeta, c->zeta) || getd(bath, c->bath))
        return fail(c, RP_E_HIP, "device-to-host copy failed");
    std::vector<int8_t> b8(n);
    if (status) {
        HIPCHK(c, hipMemcpy(b8.data(), c->status, n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) status[ids[i]] = status_value(b8[i]);
    }
    if (gid) {
        HIPCHK(c, hipMemcpy(b8.data(), c->gid, n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) gid[ids[i]] = b8[i];
    }
    if (isin) {
        std::vector<uint8_t> u8(n);
        HIPCHK(c, hipMemcpy(u8.data(), c->isin, n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) isin[ids[i]] = u8[i];
    }
    return RP_OK;
}

int rp_set_lonlat(void* ctx, int32_t ng, const double* lon, const double* lat)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (ng < 1 || ng > c->cfg.ngrid || !c->grids[ng - 1].set)
        return fail(c, RP_E_STATE, "rp_set_lonlat: grid not set");
    if (!lon || !lat) return fail(c, RP_E_ARG, "null lon/lat");
    GridDev& g = c->grids[ng - 1];
    const size_t nr = (size_t)g.L * g.M;
    int rc;
    if ((rc = dalloc(c, &g.lon, nr)) || (rc = dalloc(c, &g.lat, nr))) return rc;
    if ((rc = put(c, g.lon, lon, nr * 8)) || (rc = put(c, g.lat, lat, nr * 8))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RP_OK;
}

int rp_set_hydro_tracers(void* ctx, int32_t ng, int32_t slot, const double* salt, const double* temp)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (ng < 1 || ng > c->cfg.ngrid || !c->grids[ng - 1].set)
        return fail(c, RP_E_STATE, "rp_set_hydro_tracers: grid not set");
    if (slot < 1 || slot > 3) return fail(c, RP_E_ARG, "slot must be 1..3");
    if (!salt || !temp) return fail(c, RP_E_ARG, "null field");
    GridDev& g = c->grids[ng - 1];
    const size_t elems = (size_t)g.L * g.M * g.N * 4 * c->fbytes;
    int rc;
    if (!g.salt && ((rc = dalloc(c, (char**)&g.salt, elems)) || (rc = dalloc(c, (char**)&g.temp, elems))))
        return rc;
    const int gi = ng - 1, s = slot - 1;
    if ((rc = upload_field(c, gi, 5, g.salt, salt, nullptr, g.L, g.M, g.N, s)) ||
        (rc = upload_field(c, gi, 6, g.temp, temp, nullptr, g.L, g.M, g.N, s)))
        return rc;
    g.tracer_set[s] = true;
    return RP_OK;
}

int rp_set_hydro_wetdry(void* ctx, int32_t ng, int32_t slot, const double* mask_rho, const double* mask_u,
                        const double* mask_v)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (ng < 1 || ng > c->cfg.ngrid || !c->grids[ng - 1].set)
        return fail(c, RP_E_STATE, "rp_set_hydro_wetdry: grid not set");
    if (slot < 1 || slot > 3) return fail(c, RP_E_ARG, "slot must be 1..3");
    if (!mask_rho || !mask_u || !mask_v) return fail(c, RP_E_ARG, "null mask");
    GridDev& g = c->grids[ng - 1];
    const int64_t cnt[3] = {(int64_t)g.L * g.M, (int64_t)(g.L - 1) * g.M, (int64_t)g.L * (g.M - 1)};
    const double* src[3] = {mask_rho, mask_u, mask_v};
    const int64_t plane = wd_plane(g);
    std::vector<uint8_t> b((size_t)plane);
    int64_t o = 0;
    for (int f = 0; f < 3; f++)
        for (int64_t i = 0; i < cnt[f]; i++, o++) {
            const double m = src[f][i];
            if (!(m == 0.0 || m == 1.0)) return fail(c, RP_E_ARG, "rp_set_hydro_wetdry: masks must be 0 or 1");
            b[(size_t)o] = m == 1.0 ? 1 : 0;
        }
    int rc;
    if (!g.wd) {
        if ((rc = dalloc(c, (char**)&g.wd, (size_t)(3 * plane)))) return rc;
        // slots not given masks yet are dry, as the reference allocates them
        HIPCHK(c, hipMemsetAsync(g.wd, 0, (size_t)(3 * plane), c->stream));
    }
    // steps already queued may still read this slot's previous masks
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(g.wd + (slot - 1) * plane, b.data(), (size_t)plane, hipMemcpyHostToDevice));
    c->adj.dirty = true;   // a recorded run snapshots the masks with the slots
    c->field_gen++;        // the FP32 form's current u and v fields carry the masks
    return RP_OK;
}

namespace {
// checks and fills what every tracer launch shares: the slots' salt/temp
// and hydro fields on every grid, the stored particle columns
int tracer_args(Ctx* c, int32_t t_b, int32_t t_c, int32_t t_f, const char* who, TracerArgs& A)
{
    if (t_b < 1 || t_b > 3 || t_c < 1 || t_c > 3 || t_f < 1 || t_f > 3)
        return fail(c, RP_E_ARG, "slot numbers must be 1..3");
    std::memset(&A, 0, sizeof(A));
    for (int i = 0; i < c->cfg.ngrid; i++) {
        const GridDev& g = c->grids[i];
        if (!g.set) return fail(c, RP_E_STATE, std::string(who) + ": grid not set");
        if (!g.tracer_set[t_b - 1] || !g.tracer_set[t_c - 1] || !g.tracer_set[t_f - 1] ||
            !g.slot_set[t_b - 1] || !g.slot_set[t_c - 1] || !g.slot_set[t_f - 1])
            return fail(c, RP_E_STATE, std::string(who) + ": salt/temp or hydro slot not loaded");
        A.salt[i] = g.salt;
        A.temp[i] = g.temp;
    }
    A.slots.s[0] = t_b - 1; A.slots.s[1] = t_c - 1; A.slots.s[2] = t_f - 1;
    A.n = c->n;
    A.x = c->x; A.y = c->y; A.z = c->z; A.gid = c->gid;
    A.id = c->id; A.age = c->age; A.isin = c->isin;
    A.bath = c->bath;
    A.wd = c->wd;
    A.pzeta = c->zeta;
    return RP_OK;
}

int launch_tracers(Ctx* c, const TracerArgs& A)
{
    int rc;
    if ((rc = refresh_tab(c))) return rc;
    if (A.n == 0) return RP_OK;
    const int grid = (int)std::min<int64_t>((A.n + 255) / 256, (int64_t)c->ncu * 16);
    if (c->fbytes == 4) hipLaunchKernelGGL(tracer_kernel<float>, dim3(grid), dim3(256), 0, c->stream, c->dgrids, A);
    else hipLaunchKernelGGL(tracer_kernel<double>, dim3(grid), dim3(256), 0, c->stream, c->dgrids, A);
    HIPCHK(c, hipGetLastError());
    return RP_OK;
}
}  // namespace

int rp_sample_tracers(void* ctx, int64_t n, const double ex[3], const double ix[3], int32_t t_b,
                      int32_t t_c, int32_t t_f, double* salt, double* temp)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !ex || !ix) return RP_E_ARG;
    if (n != c->n) return fail(c, RP_E_ARG, "particle count differs from rp_set_particles");
    if (n > 0 && (!salt || !temp)) return fail(c, RP_E_ARG, "null output");
    TracerArgs A;
    int rc;
    if ((rc = tracer_args(c, t_b, t_c, t_f, "rp_sample_tracers", A))) return rc;
    if (n == 0) return RP_OK;
    for (int k = 0; k < 3; k++) A.ex[k] = ex[k];
    A.ix2 = ix[1];
    double* d = nullptr;
    if ((rc = dalloc(c, &d, 2 * (size_t)n))) return rc;
    A.out_salt = d; A.out_temp = d + n;
    A.mode = 0;
    if ((rc = launch_tracers(c, A))) {
        dfree(d);
        return rc;
    }
    std::vector<double> h(2 * (size_t)n);
    std::vector<uint32_t> ids(n);
    hipError_t e = hipMemcpyAsync(h.data(), d, 2 * (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ids.data(), c->id, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d);
    if (e != hipSuccess) return fail(c, RP_E_HIP, std::string("rp_sample_tracers: ") + hipGetErrorString(e));
    for (int64_t i = 0; i < n; i++) {
        salt[ids[i]] = h[i];
        temp[ids[i]] = h[n + i];
    }
    return RP_OK;
}

int rp_tracers_begin(void* ctx, int32_t mean, int32_t t_c)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    TracerArgs A;
    int rc;
    if ((rc = tracer_args(c, t_c, t_c, t_c, "rp_tracers_begin", A))) return rc;
    const size_t n = (size_t)c->n;
    c->trc_n = -1;
    if ((rc = dalloc(c, &c->trc_last_age, n)) || (rc = dalloc(c, &c->trc_salt, n)) ||
        (rc = dalloc(c, &c->trc_temp, n)))
        return rc;
    dfree(c->trc_mean_salt); dfree(c->trc_mean_temp);
    if (mean && ((rc = dalloc(c, &c->trc_mean_salt, n)) || (rc = dalloc(c, &c->trc_mean_temp, n))))
        return rc;
    // P_Salt = P_Temp = 0 and the means 0 (ROMSPath.f90:247-258), then the
    // particles in a grid sampled at t_c
    HIPCHK(c, hipMemsetAsync(c->trc_salt, 0, n * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(c->trc_temp, 0, n * 8, c->stream));
    if (mean) {
        HIPCHK(c, hipMemsetAsync(c->trc_mean_salt, 0, n * 8, c->stream));
        HIPCHK(c, hipMemsetAsync(c->trc_mean_temp, 0, n * 8, c->stream));
    }
    A.ex[0] = 0.0; A.ex[1] = 1.0; A.ex[2] = 2.0;   // three equal samples: polintd returns the sample
    A.ix2 = 1.0;
    A.mode = 1;
    A.last_age = c->trc_last_age;
    A.p_salt = c->trc_salt; A.p_temp = c->trc_temp;
    A.mean_salt = c->trc_mean_salt; A.mean_temp = c->trc_mean_temp;
    if ((rc = launch_tracers(c, A))) return rc;
    // mI = 1 once a particle in a grid has been sampled (ROMSPath.f90:422)
    c->trc_mI = 0;
    if (mean && n > 0) {
        std::vector<uint8_t> isin(n);
        HIPCHK(c, hipMemcpyAsync(isin.data(), c->isin, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (uint8_t v : isin)
            if (v) { c->trc_mI = 1; break; }
    }
    c->trc_n = c->n;
    c->trc_mean = mean != 0;
    return RP_OK;
}

int rp_tracers_update(void* ctx, const double ex[3], const double ix[3], int32_t t_b, int32_t t_c,
                      int32_t t_f)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !ex || !ix) return RP_E_ARG;
    if (c->trc_n != c->n) return fail(c, RP_E_STATE, "rp_tracers_update: rp_tracers_begin not called for these particles");
    TracerArgs A;
    int rc;
    if ((rc = tracer_args(c, t_b, t_c, t_f, "rp_tracers_update", A))) return rc;
    for (int k = 0; k < 3; k++) A.ex[k] = ex[k];
    A.ix2 = ix[1];
    A.mode = 2;
    A.last_age = c->trc_last_age;
    A.p_salt = c->trc_salt; A.p_temp = c->trc_temp;
    A.mean_salt = c->trc_mean_salt; A.mean_temp = c->trc_mean_temp;
    if ((rc = launch_tracers(c, A))) return rc;
    c->trc_mI++;                      // mI = mI + 1 after every update_particles (ROMSPath.f90:533)
    return RP_OK;
}

int rp_tracers_get(void* ctx, int64_t n, double* salt, double* temp, int32_t reset_mean)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (c->trc_n != c->n) return fail(c, RP_E_STATE, "rp_tracers_get: rp_tracers_begin not called for these particles");
    if (n != c->n) return fail(c, RP_E_ARG, "particle count differs from rp_set_particles");
    if (n > 0 && (!salt || !temp)) return fail(c, RP_E_ARG, "null output");
    if (n == 0) return RP_OK;
    const double* ds = c->trc_mean ? c->trc_mean_salt : c->trc_salt;
    const double* dt = c->trc_mean ? c->trc_mean_temp : c->trc_temp;
    hipError_t e = hipMemcpyAsync(salt, ds, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(temp, dt, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && c->trc_mean && reset_mean) {
        e = hipMemsetAsync(c->trc_mean_salt, 0, (size_t)n * 8, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->trc_mean_temp, 0, (size_t)n * 8, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, RP_E_HIP, std::string("rp_tracers_get: ") + hipGetErrorString(e));
    if (c->trc_mean) {
        // mean_salt(n) / mI (ROMSPath.f90:2010-2014)
        const double mI = (double)c->trc_mI;
        for (int64_t i = 0; i < n; i++) {
            salt[i] = salt[i] / mI;
            temp[i] = temp[i] / mI;
        }
        if (reset_mean) c->trc_mI = 0;
    }
    return RP_OK;
}

int rp_set_hydro_bstress(void* ctx, int32_t ng, int32_t slot, const double* bustr, const double* bvstr)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (ng < 1 || ng > c->cfg.ngrid || !c->grids[ng - 1].set)
        return fail(c, RP_E_STATE, "rp_set_hydro_bstress: grid not set");
    if (slot < 1 || slot > 3) return fail(c, RP_E_ARG, "slot must be 1..3");
    if (!bustr || !bvstr) return fail(c, RP_E_ARG, "null field");
    GridDev& g = c->grids[ng - 1];
    const size_t nu = (size_t)(g.L - 1) * g.M, nv = (size_t)g.L * (g.M - 1);
    int rc;
    if (!g.bustr && ((rc = dalloc(c, (char**)&g.bustr, nu * 4 * c->fbytes)) ||
                     (rc = dalloc(c, (char**)&g.bvstr, nv * 4 * c->fbytes))))
        return rc;
    const int gi = ng - 1, s = slot - 1;
    if ((rc = upload_field(c, gi, 7, g.bustr, bustr, nullptr, g.L - 1, g.M, 1, s)) ||
        (rc = upload_field(c, gi, 8, g.bvstr, bvstr, nullptr, g.L, g.M - 1, 1, s)))
        return rc;
    g.bstress_set[s] = true;
    return RP_OK;
}

namespace {
__global__ void bottom_init(int64_t n, const double* __restrict__ age, const uint32_t* __restrict__ id,
                            double* last_age, double* hob, double* bu, double* bv)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r = id[q];
        last_age[r] = age[q];
        hob[r] = 9999.0;
        bu[r] = 0.0;
        bv[r] = 0.0;
    }
}
__global__ void bottom_reset_hob(int64_t n, double* hob)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        hob[q] = 9999.0;
}
}  // namespace

int rp_bottom_begin(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    const int64_t n = c->n;
    int rc;
    if ((rc = dalloc(c, &c->bot_last_age, (size_t)n)) || (rc = dalloc(c, &c->bot_hob, (size_t)n)) ||
        (rc = dalloc(c, &c->bot_bustr, (size_t)n)) || (rc = dalloc(c, &c->bot_bvstr, (size_t)n)))
        return rc;
    if (n > 0) {
        const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->ncu * 16);
        hipLaunchKernelGGL(bottom_init, dim3(grid), dim3(256), 0, c->stream, n, c->age, c->id,
                           c->bot_last_age, c->bot_hob, c->bot_bustr, c->bot_bvstr);
        HIPCHK(c, hipGetLastError());
    }
    c->bot_n = n;
    return RP_OK;
}

int rp_bottom_update(void* ctx, const double ex[3], const double ix[3], int32_t t_b, int32_t t_c,
                     int32_t t_f)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !ex || !ix) return RP_E_ARG;
    if (c->bot_n != c->n) return fail(c, RP_E_STATE, "rp_bottom_update: rp_bottom_begin not called for these particles");
    if (t_b < 1 || t_b > 3 || t_c < 1 || t_c > 3 || t_f < 1 || t_f > 3)
        return fail(c, RP_E_ARG, "slot numbers must be 1..3");
    BottomArgs A;
    std::memset(&A, 0, sizeof(A));
    for (int i = 0; i < c->cfg.ngrid; i++) {
        const GridDev& g = c->grids[i];
        if (!g.set) return fail(c, RP_E_STATE, "rp_bottom_update: grid not set");
        if (!g.bstress_set[t_b - 1] || !g.bstress_set[t_c - 1] || !g.bstress_set[t_f - 1])
            return fail(c, RP_E_STATE, "rp_bottom_update: bustr/bvstr slot not loaded");
        A.bustr[i] = g.bustr;
        A.bvstr[i] = g.bvstr;
    }
    if (c->n == 0) return RP_OK;
    A.slots.s[0] = t_b - 1; A.slots.s[1] = t_c - 1; A.slots.s[2] = t_f - 1;
    for (int k = 0; k < 3; k++) A.ex[k] = ex[k];
    A.ix2 = ix[1];
    A.n = c->n;
    A.x = c->x; A.y = c->y; A.z = c->z; A.age = c->age; A.bath = c->bath;
    A.gid = c->gid; A.id = c->id;
    A.last_age = c->bot_last_age; A.hob = c->bot_hob;
    A.bustr_out = c->bot_bustr; A.bvstr_out = c->bot_bvstr;
    int rc;
    if ((rc = refresh_tab(c))) return rc;
    const DGrid* dtab = c->dgrids;
    const int grid = (int)std::min<int64_t>((c->n + 255) / 256, (int64_t)c->ncu * 16);
    if (c->fbytes == 4) hipLaunchKernelGGL(bottom_kernel<float>, dim3(grid), dim3(256), 0, c->stream, dtab, A);
    else hipLaunchKernelGGL(bottom_kernel<double>, dim3(grid), dim3(256), 0, c->stream, dtab, A);
    HIPCHK(c, hipGetLastError());
    return RP_OK;
}

int rp_bottom_get(void* ctx, int64_t n, double* hob, double* bustr, double* bvstr, int32_t reset_hob)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (c->bot_n != c->n) return fail(c, RP_E_STATE, "rp_bottom_get: rp_bottom_begin not called for these particles");
    if (n != c->n) return fail(c, RP_E_ARG, "particle count differs from rp_set_particles");
    if (n > 0 && (!hob || !bustr || !bvstr)) return fail(c, RP_E_ARG, "null output");
    if (n == 0) return RP_OK;
    hipError_t e = hipMemcpyAsync(hob, c->bot_hob, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(bustr, c->bot_bustr, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(bvstr, c->bot_bvstr, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && reset_hob) {
        const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->ncu * 16);
        hipLaunchKernelGGL(bottom_reset_hob, dim3(grid), dim3(256), 0, c->stream, n, c->bot_hob);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, RP_E_HIP, std::string("rp_bottom_get: ") + hipGetErrorString(e));
    return RP_OK;
}

int rp_set_grid_angle(void* ctx, int32_t ng, int32_t spherical, const double* angle)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (ng < 1 || ng > c->cfg.ngrid || !c->grids[ng - 1].set)
        return fail(c, RP_E_STATE, "rp_set_grid_angle: grid not set");
    if (spherical != 0 && spherical != 1) return fail(c, RP_E_ARG, "spherical must be 0 or 1");
    GridDev& g = c->grids[ng - 1];
    if (ng == 1) c->spherical = spherical;
    int rc;
    if (angle) {
        const size_t nr = (size_t)g.L * g.M;
        if ((rc = dalloc(c, &g.angle, nr)) || (rc = put(c, g.angle, angle, nr * 8))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        dfree(g.angle);
    }
    return RP_OK;
}

int rp_locate_particles(void* ctx, int64_t n, const double* lon, const double* lat, double* z,
                        int32_t t_b, double* x, double* y, double* status, int32_t* gid,
                        uint8_t* isin)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (n < 0 || (n > 0 && (!lon || !lat || !z || !x || !y || !status || !gid || !isin)))
        return fail(c, RP_E_ARG, "rp_locate_particles: null column");
    if (t_b < 1 || t_b > 3) return fail(c, RP_E_ARG, "slot numbers must be 1..3");
    LocateArgs A;
    std::memset(&A, 0, sizeof(A));
    for (int i = 0; i < c->cfg.ngrid; i++) {
        const GridDev& g = c->grids[i];
        if (!g.set) return fail(c, RP_E_STATE, "rp_locate_particles: grid not set");
        if (!g.lon || !g.lat) return fail(c, RP_E_STATE, "rp_locate_particles: rp_set_lonlat missing");
        if (!g.slot_set[t_b - 1]) return fail(c, RP_E_STATE, "rp_locate_particles: hydro slot not loaded");
        A.lon[i] = g.lon;
        A.lat[i] = g.lat;
        A.angle[i] = g.angle;
    }
    if (n == 0) return RP_OK;
    A.ngrid = c->cfg.ngrid;
    A.spherical = c->spherical;
    A.tb_lane = t_b - 1;
    A.n = n;
    DGrid tab[kMaxGrid];
    fill_tab(c, tab);
    // inputs lon, lat, z; outputs z, x, y, status (FP64), gid, isin
    double* d = nullptr;
    DGrid* dtab = nullptr;
    int32_t* dg = nullptr;
    uint8_t* di = nullptr;
    int rc = 0;
    if ((rc = dalloc(c, &d, 7 * (size_t)n)) || (rc = dalloc(c, &dtab, kMaxGrid)) ||
        (rc = dalloc(c, &dg, (size_t)n)) || (rc = dalloc(c, &di, (size_t)n))) {
        dfree(d); dfree(dtab); dfree(dg); dfree(di);
        return rc;
    }
    A.inlon = d; A.inlat = d + n; A.z = d + 2 * n;
    A.x = d + 3 * n; A.y = d + 4 * n; A.status = d + 5 * n;
    A.gid = dg; A.isin = di;
    hipError_t e = hipMemcpyAsync(d, lon, n * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + n, lat, n * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + 2 * n, z, n * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dtab, tab, sizeof(tab), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->ncu * 16);
        if (c->fbytes == 4) hipLaunchKernelGGL(locate_kernel<float>, dim3(grid), dim3(256), 0, c->stream, dtab, A);
        else hipLaunchKernelGGL(locate_kernel<double>, dim3(grid), dim3(256), 0, c->stream, dtab, A);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(z, d + 2 * n, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(x, d + 3 * n, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(y, d + 4 * n, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(status, d + 5 * n, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(gid, dg, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(isin, di, n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d); dfree(dtab); dfree(dg); dfree(di);
    if (e != hipSuccess) return fail(c, RP_E_HIP, std::string("rp_locate_particles: ") + hipGetErrorString(e));
    return RP_OK;
}

int rp_sample_lonlat(void* ctx, int64_t n, double* lon, double* lat)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !lon || !lat) return RP_E_ARG;
    if (n != c->n) return fail(c, RP_E_ARG, "particle count differs from rp_set_particles");
    LonLatArgs A;
    std::memset(&A, 0, sizeof(A));
    for (int i = 0; i < c->cfg.ngrid; i++) {
        const GridDev& g = c->grids[i];
        if (!g.lon || !g.lat) return fail(c, RP_E_STATE, "rp_sample_lonlat: rp_set_lonlat missing");
        A.lon[i] = g.lon;
        A.lat[i] = g.lat;
        A.L[i] = g.L;
        A.M[i] = g.M;
    }
    if (n == 0) return RP_OK;
    double* d = nullptr;
    int rc = dalloc(c, &d, 2 * (size_t)n);
    if (rc) return rc;
    A.x = c->x; A.y = c->y; A.gid = c->gid; A.n = n;
    A.out_lon = d; A.out_lat = d + n;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)c->ncu * 8);
    hipLaunchKernelGGL(sample_lonlat, dim3(grid), dim3(256), 0, c->stream, A);
    hipError_t e = hipGetLastError();
    std::vector<double> h(2 * (size_t)n);
    std::vector<uint32_t> ids(n);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, 2 * (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ids.data(), c->id, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d);
    if (e != hipSuccess) return fail(c, RP_E_HIP, std::string("rp_sample_lonlat: ") + hipGetErrorString(e));
    for (int64_t i = 0; i < n; i++) {
        lon[ids[i]] = h[i];
        lat[ids[i]] = h[n + i];
    }
    return RP_OK;
}

int rp_sample_deviates(void* ctx, int64_t step_index, int64_t n, int32_t ndev, double* out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !out || n < 0 || ndev < 1) return RP_E_ARG;
    if (n == 0) return RP_OK;
    double* d = nullptr;
    int rc = dalloc(c, &d, (size_t)n * ndev);
    if (rc) return rc;
    hipLaunchKernelGGL(sample_deviates, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       (uint64_t)c->cfg.seed, step_index, n, ndev, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d, (size_t)n * ndev * 8, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(c, RP_E_HIP, hipGetErrorString(e));
    return RP_OK;
}

int rp_active_steps(void* ctx, int64_t* count)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !count) return RP_E_ARG;
    unsigned long long v = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&v, c->active, sizeof(v), hipMemcpyDeviceToHost));
    *count = (int64_t)v;
    return RP_OK;
}

int rp_secant_capped(void* ctx, int64_t* count)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !count) return RP_E_ARG;
    unsigned v = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&v, c->capped, sizeof(v), hipMemcpyDeviceToHost));
    *count = (int64_t)v;
    return RP_OK;
}

int rp_synchronize(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RP_OK;
}

int rp_kernel_timing(void* ctx, double* avg_ms, int64_t* launches)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !avg_ms || !launches) return RP_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0, ph[2] = {0.0, 0.0};
    int64_t steps = 0;
    for (auto& e : c->ev_used) {
        float a = 0.f, b = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, e.e[0], e.e[1]));
        HIPCHK(c, hipEventElapsedTime(&b, e.e[1], e.e[2]));
        ph[0] += a;
        ph[1] += b;
        tot += (double)a + (double)b;
        steps += e.steps;
        c->ev_free.push_back(e);
    }
    const double k = steps == 0 ? 0.0 : 1.0 / (double)steps;
    *launches = steps;
    *avg_ms = tot * k;
    c->phase_ms[0] = ph[0] * k;
    c->phase_ms[1] = ph[1] * k;
    c->ev_used.clear();
    return RP_OK;
}

int rp_phase_timing(void* ctx, double* advect_ms, double* walk_ms)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !advect_ms || !walk_ms) return RP_E_ARG;
    *advect_ms = c->phase_ms[0];
    *walk_ms = c->phase_ms[1];
    return RP_OK;
}

const char* rp_last_error(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return "null context";
    return c->err.c_str();
}

int rp_destroy(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return RP_E_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->grids) free_grid(g);
    dfree(c->x); dfree(c->y); dfree(c->z); dfree(c->age); dfree(c->dob);
    dfree(c->wd); dfree(c->zeta); dfree(c->bath);
    dfree(c->status); dfree(c->gid); dfree(c->isin); dfree(c->id); dfree(c->mid); dfree(c->act); dfree(c->act_blk); dfree(c->act_n);
    dfree(c->bot_last_age); dfree(c->bot_hob); dfree(c->bot_bustr); dfree(c->bot_bvstr);
    dfree(c->trc_last_age); dfree(c->trc_salt); dfree(c->trc_temp);
    dfree(c->trc_mean_salt); dfree(c->trc_mean_temp);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    free_recording(c);
    dfree(c->adj.active);
    dfree(c->pert);
    dfree(c->sort_cnt); dfree(c->perm); dfree(c->alt_id);
    for (auto& a : c->alt64) dfree(a);
    for (auto& a : c->alt8) dfree(a);
    if (c->up_stream) (void)hipStreamSynchronize(c->up_stream);
    for (auto& h : c->pinned) {
        if (h.p) (void)hipHostFree(h.p);
        if (h.done) (void)hipEventDestroy(h.done);
    }
    for (auto& row : c->fstage)
        for (auto& fs : row) {
            dfree(fs.dev);
            if (fs.copied) (void)hipEventDestroy(fs.copied);
            if (fs.packed) (void)hipEventDestroy(fs.packed);
        }
    if (c->up_stream) (void)hipStreamDestroy(c->up_stream);
    dfree(c->dev); dfree(c->scratch);
    dfree(c->active); dfree(c->capped); dfree(c->dgrids); dfree(c->work);
    for (auto& e : c->ev_used) for (auto& x : e.e) (void)hipEventDestroy(x);
    for (auto& e : c->ev_free) for (auto& x : e.e) (void)hipEventDestroy(x);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return RP_OK;
}

}  // extern "C"
