// scn_engine.hip -- event-based multigroup transport engine for gfx950 behind the C ABI of
// include/scone_gpu.h.
//
// Cycle structure (eigenPhysicsPackage_class.f90:208-310 restated as device passes):
//   1. k_transport(first)   bank (AoS, HBM) -> live SoA: per-history RNG stride, placeCoord, then up
//                           to K (flight, collision) pairs per particle in registers; survivors are
//                           compacted into the next live SoA with a wave ballot + prefix popcount.
//   2. k_transport(...)     repeated over the compacted queue (K grows geometrically) until empty.
//   3. sort sites           fission sites were appended with key broodID<<24 | emission index;
//                           a radix sort restores the sequential detain order (sortByBroodID).
//   4. canonical sums       population weights and keffImplicitClerk scores (fixed summation tree).
//   5. k_close_bins         scoreMemory%closeCycle on the collision-clerk BIN array.
//   6. normSize_Repr        per-site skip-ahead random numbers, radix-sorted threshold, stable
//                           keep / copy / duplicate compaction.
#include <cstring>
#include <cstdio>
#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <chrono>
#include <string>
#include <vector>
#include <algorithm>
#include <functional>
#include <mutex>

#include "scone_gpu.h"
#include "scn_fpmath.h"
#include "scn_device.h"
#include "scn_host.h"
#include "scn_transport.h"

using namespace scn;

// pairwise levels a canonical sum may take above its chunk partials (2^40 partials)
#define SCN_DSUM_MAX_LEVELS 40

// ---- allocation-free population control ----------------------------------------------------
// Scratch of scn_norm_size_repr_noalloc: per-site draws, their sorted bits, flags, prefix sums,
// brood bounds and rocPRIM's temporary storage, carved from one static device array of fixed size
// (each GPU holds its own copy for the life of the process).  Host code never passes the array's
// name to a kernel: it takes the array's device address with hipGetSymbolAddress once per device.
constexpr int NOALLOC_SORT_TMP = 24 << 20;
constexpr int NOALLOC_SCAN_TMP = 1 << 20;
constexpr size_t noalloc_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NOALLOC_BYTES =
    3 * noalloc_round(sizeof(double) * SCN_NOALLOC_MAX_SITES) + 2 * noalloc_round(sizeof(int) * SCN_NOALLOC_MAX_SITES) +
    2 * noalloc_round(sizeof(int) * (SCN_NOALLOC_MAX_SITES + 2)) + noalloc_round(sizeof(int) * 64) + NOALLOC_SORT_TMP +
    NOALLOC_SCAN_TMP;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

using namespace scn_eng;

// gather the key-sorted fission sites into the next bank (particleState fields of pTemp,
// neutronMGstd_class.f90:177-198)
__global__ void k_gather_sites(Sites S, const uint32_t* __restrict__ order, int n, scn_particle_state* __restrict__ out,
                               const uint64_t* __restrict__ keys_sorted)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int s = (int)order[j];
    scn_particle_state q;
    q.wgt = S.w[s];
    for (int k = 0; k < 3; ++k) {
        q.r[k] = S.r[(size_t)k * S.cap + s];
        q.dir[k] = S.u[(size_t)k * S.cap + s];
    }
    q.E = 0.0;
    q.G = S.G[s];
    q.isMG = 1;
    q.type = 1;
    q.pad0 = 0;
    q.time = S.t[s];
    q.lambda = SCN_INF;
    q.matIdx = S.mat[s];
    q.cellIdx = 0;
    q.uniqueID = S.uid[s];
    q.collisionN = S.coll[s];
    q.broodID = (int)(keys_sorted[j] >> 24);
    q.pad1 = 0;
    out[j] = q;
}

// ---- canonical sums ------------------------------------------------------------------------
struct SumJob {
    const double* x;
    int64_t n;        // values (with n_dev: the capacity, n_dev's value is clamped to it)
    int64_t stride;   // in doubles
    const int* n_dev; // null, or the number of values in device memory
};
struct SumJobs {
    SumJob j[8];
    int njobs;
};

__global__ void k_dsum_chunks(SumJobs J, double* y, int64_t ystride, int64_t* m_tab)
{
    int job = blockIdx.y;
    const SumJob& q = J.j[job];
    int64_t n = q.n;
    if (q.n_dev) n = min((int64_t)max(*q.n_dev, 0), q.n);
    int64_t m = (n + SCN_DSUM_CHUNK - 1) / SCN_DSUM_CHUNK;
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m_tab && c == 0) m_tab[job] = m;   // device plan: level 0 sizes
    if (c >= m) return;
    double s = 0.0;
    int64_t e = (c + 1) * SCN_DSUM_CHUNK < n ? (c + 1) * SCN_DSUM_CHUNK : n;
    for (int64_t i = c * SCN_DSUM_CHUNK; i < e; ++i) s = s + q.x[i * q.stride];
    y[job * ystride + c] = s;
}

__global__ void k_dsum_level(const double* y, double* y2, const int64_t* m_in, int64_t ystride)
{
    int job = blockIdx.y;
    int64_t m = m_in[job];
    int64_t h = (m + 1) / 2;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m <= 1 || i >= h) return;
    const double* a = y + job * ystride;
    y2[job * ystride + i] = (2 * i + 1 < m) ? a[2 * i] + a[2 * i + 1] : a[2 * i];
}

// one pairwise level with sizes in device memory: row `level` of m_tab in, row level + 1 out; a
// value standing alone (m == 1) is carried forward, so any number of levels gives the same tree
__global__ void k_dsum_level_dev(const double* y, double* y2, int64_t* m_tab, int level, int64_t ystride)
{
    int job = blockIdx.y;
    int64_t m = m_tab[level * 8 + job];
    int64_t h = (m + 1) / 2;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) m_tab[(level + 1) * 8 + job] = h;
    if (i >= h) return;
    const double* a = y + job * ystride;
    y2[job * ystride + i] = (2 * i + 1 < m) ? a[2 * i] + a[2 * i + 1] : a[2 * i];
}

__global__ void __launch_bounds__(1024) k_dsum_final(const double* y, const int64_t* m_in, int64_t ystride, double* out)
{
    __shared__ double a[2048];
    __shared__ double b[2048];
    int job = blockIdx.x;
    int64_t m = m_in[job];
    for (int i = threadIdx.x; i < m; i += blockDim.x) a[i] = y[job * ystride + i];
    __syncthreads();
    double* src = a;
    double* dst = b;
    while (m > 1) {
        int64_t h = (m + 1) / 2;
        for (int i = threadIdx.x; i < h; i += blockDim.x) dst[i] = (2 * i + 1 < m) ? src[2 * i] + src[2 * i + 1] : src[2 * i];
        __syncthreads();
        double* t = src;
        src = dst;
        dst = t;
        m = h;
    }
    if (threadIdx.x == 0) out[job] = (m_in[job] > 0) ? src[0] : 0.0;
}

// ---- scoreMemory%closeCycle ------------------------------------------------------------------
__global__ void k_close_bins(double* bins, double* csum, double* csum2, int64_t n, const double* norm_score,
                             double norm_value)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double nf = norm_score ? norm_value / *norm_score : 1.0;
    double res = bins[i] * nf;
    bins[i] = 0.0;
    csum[i] = csum[i] + res;
    csum2[i] = csum2[i] + res * res;
}

__global__ void k_copy_scalar(const double* src, double* dst) { *dst = *src; }

// the same with a fixed-point score memory: four limb words per bin (scn_fix_value), cleared
// after the read
__global__ void k_close_bins_fix(unsigned long long* words, double* csum, double* csum2, int64_t n,
                                 const double* norm_score, double norm_value)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double nf = norm_score ? norm_value / *norm_score : 1.0;
    unsigned long long* w = words + 4 * i;
    const uint64_t wv[4] = {w[0], w[1], w[2], w[3]};
    double res = scn_fix_value(wv) * nf;
    w[0] = w[1] = w[2] = w[3] = 0ull;
    csum[i] = csum[i] + res;
    csum2[i] = csum2[i] + res * res;
}

__global__ void k_fix_scalar(const unsigned long long* w, double* dst)
{
    const uint64_t wv[4] = {w[0], w[1], w[2], w[3]};
    *dst = scn_fix_value(wv);
}

__global__ void k_fix_values(const unsigned long long* words, int64_t n, double* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long* w = words + 4 * i;
    const uint64_t wv[4] = {w[0], w[1], w[2], w[3]};
    out[i] = scn_fix_value(wv);
}

// ---- normSize_Repr ---------------------------------------------------------------------------
__global__ void k_site_rn(uint64_t master, int n, double* rn, uint64_t* rn_bits)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint64_t s = rng_skip(master, (int64_t)j);
    double u = scn_rng_get(&s);
    rn[j] = u;
    rn_bits[j] = scn_dbits(u);
}

__global__ void k_norm_flags(const double* rn, int n, const uint64_t* thr_bits, int keep_mode, int* flag)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double thr = scn_bitsd(*thr_bits);
    flag[j] = keep_mode ? (rn[j] > thr) : (rn[j] <= thr);
}

__global__ void k_norm_keep(const scn_particle_state* in, const int* flag, const int* pos, int n, scn_particle_state* out,
                            int out_cap)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (flag[j] && pos[j] < out_cap) out[pos[j]] = in[j];
}

__global__ void k_fill_int(int* p, int n, int v)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = v;
}

// broodIDs in [0, max_brood] and non-decreasing (sortByBroodID order); bad[0] = 1 otherwise
__global__ void k_check_broods(const scn_particle_state* in, int n, int max_brood, int* bad)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int b = in[j].broodID;
    if (b < 0 || b > max_brood || (j > 0 && in[j - 1].broodID > b)) bad[0] = 1;
}

__global__ void k_brood_bounds(const scn_particle_state* in, int n, int max_brood, int* start_of, int* end_of)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int b = in[j].broodID;
    if (b < 0 || b > max_brood) return;   // excluded by the host check before launch
    if (j == 0 || in[j - 1].broodID != b) start_of[b] = j;
    if (j == n - 1 || in[j + 1].broodID != b) end_of[b] = j + 1;
}

__global__ void k_norm_dup(const scn_particle_state* in, const int* dflag, const int* D, const int* start_of,
                           const int* end_of, int n, int n_copies, int max_brood, scn_particle_state* out, int out_cap)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int b = in[j].broodID;
    if (b < 0 || b > max_brood) return;   // excluded by k_check_broods / the host check before launch
    int s = start_of[b];
    int kb = end_of[b] - s;
    int64_t base = (int64_t)(n_copies + 1) * s + D[s];
    int rho = j - s;
    for (int c = 0; c <= n_copies; ++c) {
        const int64_t o = base + (int64_t)c * kb + rho;
        if (o < out_cap) out[o] = in[j];
    }
    if (dflag[j]) {
        const int64_t o = base + (int64_t)(n_copies + 1) * kb + (D[j] - D[s]);
        if (o < out_cap) out[o] = in[j];
    }
}

// ---- device-resident power iteration ----------------------------------------------------------
// Every scalar a cycle needs (master RNG state, k, the analog/implicit accumulators, the site
// count, the normalisation plan) lives in one device record, so a cycle is a fixed sequence of
// launches with no host round trip and can be replayed from a captured hipGraph.  The arithmetic
// is the host loop's, operation for operation (scn_eigen_step), so the k history is the same bits.
struct EigCtl {
    uint64_t master;
    double k_new;          // k loaded into the histories of the next cycle
    double keff_0;
    double dk[2];
    double kA[2], kI[2], imp[4][2];
    int nA, nI, cur, cycle;
    int pop;
    int n_sites;           // sites banked by the last transport (clamped to the capacity)
    int excess, heap, n_copies, n_dup;
    int n_cand;            // draws inside the selection window (gathered)
    int pad_sel;
    double thr;            // the heap-th smallest draw (normSize_Repr's threshold)
    double sel_sigmas;     // half-width of the selection window in standard deviations
    double sel_lo, sel_hi; // the window of the last cycle
    unsigned long long sites_total;
};

// zero the transport's queue counters and the site count (before the first cycle; every cycle
// then clears them for the next one in k_dcycle_end)
__global__ void k_dcycle_begin(int* counters_i, int* site_count)
{
    if (threadIdx.x < 4) counters_i[threadIdx.x] = 0;
    if (threadIdx.x == 0) *site_count = 0;
}

__global__ void k_bank_weights(const scn_particle_state* __restrict__ bank, int n, double* __restrict__ w)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) w[j] = bank[j].wgt;
}

// window [lo, hi) that holds the k-th smallest of n uniform draws with overwhelming probability:
// its expected value k/n plus or minus s standard deviations of the order statistic and 2s
// spacings (s = 8 by default); the selection is exact either way (a miss falls back to all draws)
__device__ __forceinline__ void sel_window(int k, int n, double s, double* lo, double* hi)
{
    const double p = (double)k / (double)n;
    const double h = s * (sqrt(p * (1.0 - p) / (double)n) + 2.0 / (double)n);
    *lo = p - h;
    *hi = p + h;
}

// exclusive prefix sum of n ints (n in device memory), 1024 values per block: block scans, a scan
// of the block totals in one block, then the block offsets added
constexpr int SCAN_PER_BLOCK = 1024;
__global__ void __launch_bounds__(256) k_scan_blocks(const int* __restrict__ in, int* __restrict__ out, int* bsum,
                                                     const int* n_ptr)
{
    __shared__ int w[256];
    const int n = *n_ptr;
    const int base = blockIdx.x * SCAN_PER_BLOCK;
    if (base >= n) return;
    int v[4], t = 0;
    for (int k = 0; k < 4; ++k) {
        const int j = base + threadIdx.x * 4 + k;
        v[k] = j < n ? in[j] : 0;
        t += v[k];
    }
    w[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int x = threadIdx.x >= off ? w[threadIdx.x - off] : 0;
        __syncthreads();
        w[threadIdx.x] += x;
        __syncthreads();
    }
    int run = w[threadIdx.x] - t;
    for (int k = 0; k < 4; ++k) {
        const int j = base + threadIdx.x * 4 + k;
        if (j < n) out[j] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = w[255];
}

__global__ void __launch_bounds__(1024) k_scan_top(int* bsum, const int* n_ptr)
{
    __shared__ int w[1024];
    const int nb = (*n_ptr + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK;
    const int per = (nb + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    int t = 0;
    for (int b = b0; b < min(b0 + per, nb); ++b) t += bsum[b];
    w[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int x = threadIdx.x >= off ? w[threadIdx.x - off] : 0;
        __syncthreads();
        w[threadIdx.x] += x;
        __syncthreads();
    }
    int run = w[threadIdx.x] - t;
    for (int b = b0; b < min(b0 + per, nb); ++b) {
        const int x = bsum[b];
        bsum[b] = run;
        run += x;
    }
}

// history b's sites start at the exclusive prefix sum of the site counts of histories 1..b-1
// (sortByBroodID of the detain order, without a sort); block sums of that scan are added here.
// Each site's weight goes to its position in the next bank (a contiguous array for the END sum).
__global__ void __launch_bounds__(256) k_site_index(Sites S, const int* __restrict__ site_off,
                                                    const int* __restrict__ bsum, int pop, int out_cap, EigCtl* ctl,
                                                    double* __restrict__ w_out, DevStatus* st)
{
    const int count = *S.count;
    const int n = min(count, min(S.cap, out_cap));
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {
        ctl->n_sites = n;
        ctl->sites_total += (unsigned long long)n;
        if (count > n) raise_error(st, SCN_ERR_OVERFLOW, count);
    }
    if (j >= n) return;
    const uint64_t key = S.key[j];
    const int b = (int)(key >> 24);
    if (b < 1 || b > pop) return;
    const int64_t dst = (int64_t)site_off[b - 1] + bsum[(b - 1) / SCAN_PER_BLOCK] + (int64_t)(key & 0xFFFFFF);
    if (dst < n) w_out[dst] = S.w[j];
}

// when normSize_Repr must draw (heap > 0): the draw of every position of the next bank
// (particleDungeon_class.f90:471-487: one number per site from the master stream skipped to it;
// k_dcycle_close has applied the package's stride of pop + 1), in bank order; draws below the selection window are
// counted per block (into below[block], summed by k_sel_pick), draws inside it gathered (one atomic
// per wave that has any)
__global__ void __launch_bounds__(256) k_site_draws(EigCtl* ctl, double* __restrict__ rn, double* __restrict__ cand,
                                                    int cand_cap, int* __restrict__ below_part)
{
    __shared__ int below;
    const int n = ctl->n_sites;
    const int heap = ctl->heap;
    if (heap == 0 || (int64_t)blockIdx.x * blockDim.x >= n) return;   // uniform over the block
    const double lo = ctl->sel_lo, hi = ctl->sel_hi;
    const uint64_t master = ctl->master;
    if (threadIdx.x == 0) below = 0;
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    int is_below = 0, is_cand = 0;
    double u = 0.0;
    if (j < n) {
        uint64_t s = rng_skip(master, j);
        u = scn_rng_get(&s);
        rn[j] = u;
        is_below = u < lo;
        is_cand = !is_below && u < hi;
    }
    const unsigned long long mb = __ballot(is_below), mc = __ballot(is_cand);
    const int lane = threadIdx.x & 63;
    if (mb && lane == 0) atomicAdd(&below, __popcll(mb));
    if (mc) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&ctl->n_cand, __popcll(mc));
        base = __shfl(base, 0);
        if (is_cand) {
            const int i = base + __popcll(mc & ((1ull << lane) - 1ull));
            if (i < cand_cap) cand[i] = u;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) below_part[blockIdx.x] = below;
}

// canonical sum, first stage with coalesced loads: block b of a job sums chunks 256b .. 256b + 255
// (SCN_DSUM_CHUNK consecutive values each, in order) and pairs them 8 levels deep, which is
// element b of level 8 of the tree (pairs never straddle two blocks: a block's first element is
// even at every level below 9, and the odd element a level carries alone is its last, which lies
// in the last block).  m_tab[job] receives level 8's size.
__global__ void __launch_bounds__(256) k_dsum_tree(SumJobs J, double* y, int64_t ystride, int64_t* m_tab)
{
    constexpr int HALF = SCN_DSUM_CHUNK / 2;
    __shared__ double x[256 * (HALF + 1)];
    __shared__ double a[256];
    const int job = blockIdx.y;
    const SumJob& q = J.j[job];
    int64_t n = q.n;
    if (q.n_dev) n = min((int64_t)max(*q.n_dev, 0), q.n);
    const int64_t m = (n + SCN_DSUM_CHUNK - 1) / SCN_DSUM_CHUNK;
    if (blockIdx.x == 0 && threadIdx.x == 0) m_tab[job] = (m + 255) / 256;
    const int64_t c0 = (int64_t)blockIdx.x * 256;
    if (c0 >= m) return;   // uniform over the block
    const int64_t v0 = c0 * SCN_DSUM_CHUNK;
    const int64_t c = c0 + threadIdx.x;
    const int64_t e = c < m ? min((c + 1) * SCN_DSUM_CHUNK, n) - c * SCN_DSUM_CHUNK : 0;   // values of chunk c
    double s = 0.0;
    for (int half = 0; half < 2; ++half) {   // the chunk's first 16 values, then its last 16
        for (int k = 0; k < HALF; ++k) {
            const int l = k * 256 + threadIdx.x;           // 16 consecutive values of one chunk per 16 threads
            const int ch = l / HALF, off = l % HALF;
            const int64_t i = v0 + (int64_t)ch * SCN_DSUM_CHUNK + half * HALF + off;
            x[ch * (HALF + 1) + off] = i < n ? q.x[i * q.stride] : 0.0;
        }
        __syncthreads();
        const double* xc = x + threadIdx.x * (HALF + 1);
        for (int i = 0; i < HALF; ++i)
            if (half * HALF + i < e) s = s + xc[i];
        __syncthreads();
    }
    a[threadIdx.x] = s;
    __syncthreads();
    int64_t ml = m;   // size of the level being paired
    for (int lev = 0; lev < 8; ++lev) {
        const int64_t base = c0 >> lev;               // this block's first element at level lev
        const int64_t h = 128 >> lev;                 // outputs of this block at level lev + 1
        double v = 0.0;
        const int i = threadIdx.x;
        if (i < h) {
            const int64_t g = base + 2 * i;           // global index at level lev
            if (g < ml) v = (g + 1 < ml) ? a[2 * i] + a[2 * i + 1] : a[2 * i];
        }
        __syncthreads();
        if (i < h) a[i] = v;
        __syncthreads();
        ml = (ml + 1) / 2;
    }
    if (threadIdx.x == 0) y[job * ystride + blockIdx.x] = a[0];
}

// close of a cycle's scalars in one block: the canonical sums' last levels for every job
// (START weight, END weight, the 4 keffImplicit scores), then keffAnalogClerk
// (keffAnalogClerk_class.f90:162-181) or keffImplicitClerk (:306-326) after the package's stride
// of the master state by pop + 1, and the normSize_Repr plan from the site count
// (particleDungeon_class.f90:431-470)
__global__ void __launch_bounds__(1024) k_dcycle_close(const double* y, const int64_t* m_in, int64_t ystride,
                                                       int njobs, EigCtl* ctl, int phase, DevStatus* st)
{
    __shared__ double a[2048];
    __shared__ double b[2048];
    __shared__ double sums[8];
    for (int job = 0; job < njobs; ++job) {
        int64_t m = m_in[job];
        for (int i = threadIdx.x; i < m; i += blockDim.x) a[i] = y[job * ystride + i];
        __syncthreads();
        double* src = a;
        double* dst = b;
        while (m > 1) {
            const int64_t h = (m + 1) / 2;
            for (int i = threadIdx.x; i < h; i += blockDim.x)
                dst[i] = (2 * i + 1 < m) ? src[2 * i] + src[2 * i + 1] : src[2 * i];
            __syncthreads();
            double* t = src;
            src = dst;
            dst = t;
            m = h;
        }
        if (threadIdx.x == 0) sums[job] = (m_in[job] > 0) ? src[0] : 0.0;
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    ctl->master = rng_stride(ctl->master, (int64_t)ctl->pop + 1);
    const double startW = sums[0], endW = sums[1];
    const double k_norm = ctl->dk[1 - ctl->cur];
    if (phase == 0) {
        const double k_eff = endW / startW * k_norm;
        ctl->kA[0] = ctl->kA[0] + k_eff;
        ctl->kA[1] = ctl->kA[1] + k_eff * k_eff;
        ctl->nA++;
    } else {
        const double k_est = sums[2] / (sums[4] + sums[5] - sums[3]);
        ctl->kI[0] = ctl->kI[0] + k_est;
        ctl->kI[1] = ctl->kI[1] + k_est * k_est;
        ctl->nI++;
        for (int k = 0; k < 4; ++k) {
            ctl->imp[k][0] = ctl->imp[k][0] + sums[2 + k];
            ctl->imp[k][1] = ctl->imp[k][1] + sums[2 + k] * sums[2 + k];
        }
    }
    const int n = ctl->n_sites;
    if (n == 0) raise_error(st, SCN_ERR_NORM, 0);
    const int excess = n - ctl->pop;
    ctl->excess = excess;
    ctl->heap = n == 0 ? 0 : (excess < 0 ? ((-excess) % n) : excess);
    ctl->n_copies = (n == 0 || excess >= 0) ? 0 : (-excess) / n;
    ctl->n_dup = (n == 0 || excess >= 0) ? 0 : (-excess) % n;
    if (ctl->heap) sel_window(ctl->heap, n, ctl->sel_sigmas, &ctl->sel_lo, &ctl->sel_hi);   // k_site_index's window
}

// digit (0..255) of a 256-bin histogram h where the cumulative count reaches k, and the rank
// left within it; the first 256 threads of the block scan, all threads get the result
__device__ void pick_digit(const int* h, int k, int* digit, int* rank)
{
    __shared__ int c[256];
    __shared__ int res[2];
    const int t = threadIdx.x;
    const int v = t < 256 ? h[t] : 0;
    if (t < 256) c[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int x = (t < 256 && t >= off) ? c[t - off] : 0;
        __syncthreads();
        if (t < 256) c[t] += x;
        __syncthreads();
    }
    if (t < 256) {
        const int incl = c[t], excl = incl - v;
        if (excl < k && k <= incl) {
            res[0] = t;
            res[1] = k - excl;
        }
    }
    __syncthreads();
    *digit = res[0];
    *rank = res[1];
    __syncthreads();
}

// the heap-th smallest draw, exactly, in one block of 1024 threads.  When the window held it (the
// draws below it fewer than heap, with the window's draws at least heap) and every draw inside was
// gathered, it is the (heap - below)-th smallest of the gathered draws.  Up to PICK_LDS of them:
// a histogram of PICK_BUCKETS equal sub-intervals of the window (the bucket index is monotone in
// the draw) gives the bucket and the rank within it, then ranks are counted among that bucket's
// draws.  More: a radix select on their bit patterns (non-negative doubles order as their bits).
// A miss of the window: the radix select over all draws.
constexpr int PICK_LDS = 4096;
constexpr int PICK_BUCKETS = 4096;
constexpr int PICK_MEMBERS = 512;
__global__ void __launch_bounds__(1024) k_sel_pick(const double* __restrict__ rn, EigCtl* ctl,
                                                   const double* __restrict__ cand, int cand_cap,
                                                   const int* __restrict__ below_part)
{
    __shared__ double v[PICK_LDS];
    __shared__ int hb[PICK_BUCKETS];
    __shared__ int part[1024];
    __shared__ double mem[PICK_MEMBERS];
    __shared__ int h[256];
    __shared__ int s_bucket, s_rank, s_m, s_below;
    if (ctl->heap == 0) return;
    const int t = threadIdx.x;
    // draws below the window: the per-block counts of k_site_draws
    const int nblk = (ctl->n_sites + 255) / 256;
    int bl = 0;
    for (int i = t; i < nblk; i += blockDim.x) bl += below_part[i];
    if (t == 0) s_below = 0;
    __syncthreads();
    if (bl) atomicAdd(&s_below, bl);
    __syncthreads();
    const int heap = ctl->heap, nb = s_below, nc = ctl->n_cand;
    const bool in_window = nb < heap && heap <= nb + nc && nc <= cand_cap;
    if (in_window && nc <= PICK_LDS) {
        const int k = heap - nb;
        const double lo = ctl->sel_lo, scale = (double)PICK_BUCKETS / (ctl->sel_hi - ctl->sel_lo);
        auto bucket = [&](double u) {
            const double x = (u - lo) * scale;   // >= 0, monotone in u
            return x >= (double)(PICK_BUCKETS - 1) ? PICK_BUCKETS - 1 : (int)x;
        };
        for (int i = t; i < PICK_BUCKETS; i += blockDim.x) hb[i] = 0;
        if (t == 0) s_m = 0;
        __syncthreads();
        for (int i = t; i < nc; i += blockDim.x) {
            const double u = cand[i];
            v[i] = u;
            atomicAdd(&hb[bucket(u)], 1);
        }
        __syncthreads();
        constexpr int PER = PICK_BUCKETS / 1024;
        int c[PER], tot = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            c[q] = hb[t * PER + q];
            tot += c[q];
        }
        part[t] = tot;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int x = t >= off ? part[t - off] : 0;
            __syncthreads();
            part[t] += x;
            __syncthreads();
        }
        const int incl = part[t], excl = incl - tot;
        if (excl < k && k <= incl) {
            int run = excl;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (run + c[q] >= k) {
                    s_bucket = t * PER + q;
                    s_rank = k - run;
                    break;
                }
                run += c[q];
            }
        }
        __syncthreads();
        const int B = s_bucket, r = s_rank;
        for (int i = t; i < nc; i += blockDim.x)
            if (bucket(v[i]) == B) {
                const int m = atomicAdd(&s_m, 1);
                if (m < PICK_MEMBERS) mem[m] = v[i];
            }
        __syncthreads();
        const int m = s_m;
        if (m <= PICK_MEMBERS) {
            for (int i = t; i < m; i += blockDim.x) {
                const double x = mem[i];
                int less = 0, leq = 0;
                for (int q = 0; q < m; ++q) {
                    less += mem[q] < x;
                    leq += mem[q] <= x;
                }
                if (less < r && r <= leq) ctl->thr = x;   // equal draws write the same value
            }
        } else {   // a crowded bucket: count over the whole candidate set restricted to it
            for (int i = t; i < nc; i += blockDim.x) {
                const double x = v[i];
                if (bucket(x) != B) continue;
                int less = 0, leq = 0;
                for (int q = 0; q < nc; ++q) {
                    if (bucket(v[q]) != B) continue;
                    less += v[q] < x;
                    leq += v[q] <= x;
                }
                if (less < r && r <= leq) ctl->thr = x;
            }
        }
        return;
    }
    const double* src = in_window ? cand : rn;
    const int n = in_window ? nc : ctl->n_sites;
    int k = in_window ? heap - nb : heap;
    uint64_t prefix = 0, mask = 0;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        if (t < 256) h[t] = 0;
        __syncthreads();
        for (int j = t; j < n; j += blockDim.x) {
            const uint64_t x = scn_dbits(src[j]);
            if ((x & mask) == prefix) atomicAdd(&h[(int)((x >> shift) & 255)], 1);
        }
        __syncthreads();
        int d, rr;
        pick_digit(h, k, &d, &rr);
        prefix |= (uint64_t)d << shift;
        mask |= (uint64_t)255 << shift;
        k = rr;
    }
    if (t == 0) ctl->thr = scn_bitsd(prefix);
}

// normSize_Repr's selection flags (keep the draws above the threshold when there are too many
// sites, duplicate those at or below it when there are too few; particleDungeon_class.f90:488-
// 560) and their exclusive prefix sum within blocks of SCAN_PER_BLOCK, block totals to bsum
__global__ void __launch_bounds__(256) k_norm_flag_scan(const double* __restrict__ rn, const EigCtl* ctl,
                                                        int* __restrict__ flag, int* __restrict__ pos, int* bsum)
{
    __shared__ int w[256];
    const int n = ctl->n_sites;
    const int base = blockIdx.x * SCAN_PER_BLOCK;
    if (base >= n) return;
    const double thr = ctl->thr;
    const int ex = ctl->excess;
    const bool dup = ex < 0 && ctl->n_dup != 0;
    int v[4], t = 0;
    for (int k = 0; k < 4; ++k) {
        const int j = base + threadIdx.x * 4 + k;
        v[k] = 0;
        if (j < n) {
            v[k] = ex > 0 ? (rn[j] > thr) : dup ? (rn[j] <= thr) : 0;
            flag[j] = v[k];
        }
        t += v[k];
    }
    w[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int x = threadIdx.x >= off ? w[threadIdx.x - off] : 0;
        __syncthreads();
        w[threadIdx.x] += x;
        __syncthreads();
    }
    int run = w[threadIdx.x] - t;
    for (int k = 0; k < 4; ++k) {
        const int j = base + threadIdx.x * 4 + k;
        if (j < n) pos[j] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = w[255];
}

// the next cycle's bank, straight from the site buffer: keep (excess > 0), copies and duplicates
// (excess < 0, each history's sites as a group: (n_copies + 1) copies of the group, then its
// duplicated sites) or all (excess = 0), with the weights to a contiguous array for the next
// cycle's START sum
__global__ void __launch_bounds__(256) k_norm_out(Sites S, const int* __restrict__ site_off,
                                                  const int* __restrict__ bsum_site, const int* __restrict__ hist_nsite,
                                                  const int* __restrict__ flag, const int* __restrict__ pos,
                                                  const int* __restrict__ bsum_pos, const EigCtl* ctl, int pop,
                                                  scn_particle_state* __restrict__ out, double* __restrict__ w_out,
                                                  int out_cap)
{
    const int n = ctl->n_sites;
    const int count = min(*S.count, S.cap);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const uint64_t key = S.key[j];
    const int b = (int)(key >> 24);
    if (b < 1 || b > pop) return;
    const int64_t s = (int64_t)site_off[b - 1] + bsum_site[(b - 1) / SCAN_PER_BLOCK];
    const int64_t e = (int64_t)(key & 0xFFFFFF);
    const int64_t dst = s + e;
    if (dst >= n) return;
    scn_particle_state q;
    q.wgt = S.w[j];
    for (int k = 0; k < 3; ++k) {
        q.r[k] = S.r[(size_t)k * S.cap + j];
        q.dir[k] = S.u[(size_t)k * S.cap + j];
    }
    q.E = 0.0;
    q.G = S.G[j];
    q.isMG = 1;
    q.type = 1;
    q.pad0 = 0;
    q.time = S.t[j];
    q.lambda = SCN_INF;
    q.matIdx = S.mat[j];
    q.cellIdx = 0;
    q.uniqueID = S.uid[j];
    q.collisionN = S.coll[j];
    q.broodID = b;
    q.pad1 = 0;
    const int ex = ctl->excess;
    auto put = [&](int64_t o) {
        if (o >= 0 && o < out_cap) {
            out[o] = q;
            w_out[o] = q.wgt;
        }
    };
    if (ex == 0) {
        put(dst);
    } else if (ex > 0) {
        if (flag[dst]) put((int64_t)pos[dst] + bsum_pos[dst / SCAN_PER_BLOCK]);
    } else {
        const int64_t kb = min((int64_t)hist_nsite[b - 1], (int64_t)n - s);   // the group's sites in the bank
        const int nc = ctl->n_copies;
        const int64_t ps = (int64_t)pos[s] + bsum_pos[s / SCAN_PER_BLOCK];
        const int64_t base = (int64_t)(nc + 1) * s + ps;
        for (int c = 0; c <= nc; ++c) put(base + (int64_t)c * kb + e);
        if (flag[dst]) put(base + (int64_t)(nc + 1) * kb + ((int64_t)pos[dst] + bsum_pos[dst / SCAN_PER_BLOCK] - ps));
    }
}

// end of the cycle: population check, the package's stride by 1, k for the next cycle
// (eigenPhysicsPackage_class.f90:262-310); the transport's counters, the site count and the draw
// histogram are cleared for the next cycle
__global__ void k_dcycle_end(EigCtl* ctl, const int* flag, const int* pos, const int* bsum_pos, double* k_hist,
                             int k_hist_cap, int phase, int* counters_i, int* site_count, DevStatus* st)
{
    if (threadIdx.x < 4) counters_i[threadIdx.x] = 0;
    if (threadIdx.x != 0) return;
    *site_count = 0;
    const int n = ctl->n_sites;
    if (n > 0) {
        const int ex = ctl->excess;
        const int last = flag[n - 1] + pos[n - 1] + bsum_pos[(n - 1) / SCAN_PER_BLOCK];
        const int n_new = ex == 0 ? n : ex > 0 ? last : n * (ctl->n_copies + 1) + last;
        if (n_new != ctl->pop) raise_error(st, SCN_ERR_NORM, n_new);
    }
    ctl->n_cand = 0;
    ctl->master = rng_stride(ctl->master, 1);
    ctl->cur = 1 - ctl->cur;
    const double k_new = phase == 0 ? ctl->kA[0] / ctl->nA : ctl->kI[0] / ctl->nI;
    ctl->k_new = k_new;
    ctl->dk[1 - ctl->cur] = k_new;
    ctl->keff_0 = k_new;
    if (ctl->cycle < k_hist_cap) k_hist[ctl->cycle] = k_new;
    ctl->cycle++;
}

// ---- initial fission source (fissionSource_class.f90:149-271) ---------------------------------
template <int NL, int GEO>
__global__ void k_fission_source(DevProblem P, int n, uint64_t master, double b0, double b1, double b2, double t0,
                                 double t1, double t2, int attempts, scn_particle_state* bank, DevStatus* st)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t rng = rng_stride(master, (int64_t)i + 1);
    const double bottom[3] = {b0, b1, b2}, top[3] = {t0, t1, t2};
    scn_particle_state s;
    memset(&s, 0, sizeof s);
    int it = 0;
    for (;;) {
        if (++it > attempts) {
            raise_error(st, SCN_ERR_SAMPLING, 3);
            break;
        }
        Pt<NL, GEO> p;
        double r3[3];
        for (int k = 0; k < 3; ++k) r3[k] = scn_rng_get(&rng);
        for (int k = 0; k < 3; ++k) {
            p.r[0][k] = (top[k] - bottom[k]) * r3[k] + bottom[k];
            s.r[k] = p.r[0][k];
        }
        p.u[0] = 1.0;
        p.u[1] = 0.0;
        p.u[2] = 0.0;
        place_coord(P, p, st);
        int mat = p.mat;
        if (mat == SCN_VOID_MAT || mat == SCN_OUTSIDE_MAT) continue;
        if (mat == SCN_UNDEF_MAT || mat == SCN_OVERLAP_MAT) {
            raise_error(st, SCN_ERR_MATERIAL, mat);
            break;
        }
        if (!P.fissile[mat - 1]) continue;
        s.matIdx = mat;
        s.uniqueID = p.uid;
        s.wgt = 1.0;
        s.time = 0.0;
        s.type = 1;
        double mu = 2.0 * scn_rng_get(&rng) - 1.0;
        double phi = SCN_TWO_PI * scn_rng_get(&rng);
        double rem = scn_rng_get(&rng);
        const double* chi = P.chi + (size_t)(mat - 1) * P.nG;
        int G;
        for (G = 1; G <= P.nG; ++G) {
            rem = rem - chi[G - 1];
            if (rem < 0.0) break;
        }
        if (G > P.nG) {
            raise_error(st, SCN_ERR_SAMPLING, 2);
            break;
        }
        s.G = G;
        s.isMG = 1;
        const double x1[3] = {1.0, 0.0, 0.0};
        scn_rotate(x1, mu, phi, s.dir);
        s.lambda = SCN_INF;
        s.cellIdx = -1;
        break;
    }
    bank[i] = s;
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct EigenState {
    int open = 0;
    scn_eigen_settings S{};
    uint64_t master = 0;
    double kA[2] = {0, 0}, kI[2] = {0, 0}, imp[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    int nA = 0, nI = 0;
    double dk[2] = {1.0, 1.0};
    int cur = 0;
    double keff_0 = 1.0, k_new = 1.0;
    int phase = 0, c_in_phase = 0;
    int n_this = 0;
    int64_t n_hist = 0;
    double transport_s = 0.0;
    std::vector<double> k_history, cycle_time;
    Counters c0{};
    unsigned long long sites_dev = 0;   // cycle_graph runs: sites counted on the device so far
};

struct Engine {
    std::string err;
    EigenState eig;
    hipStream_t stream = nullptr;
    int nl = 1;             // compiled nesting depth
    int geo = 0;            // general-geometry kernels (cells, rotations, non-box border)
    int transport = SCN_TRANSPORT_ST;
    DevProblem P{};
    int64_t n_bins = 0;
    int64_t norm_bin = -1;
    double norm_value = 1.0;
    int nG = 0;
    int cap = 0;            // live capacity (histories per cycle)
    int site_cap = 0;
    int bank_cap = 0;
    std::vector<DevBuf> owned;
    Live live[2]{};
    Sites sites{};
    uint64_t* keys_alt = nullptr;
    uint32_t* order_alt = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    scn_particle_state* bankA = nullptr;   // thisCycle
    scn_particle_state* bankB = nullptr;   // nextCycle
    double* hist_k = nullptr;
    double* bins = nullptr;                // BIN: n_bins doubles, or 4 limb words per bin (P.fix)
    double* bins_out = nullptr;            // fixed point: BIN values read out for scn_eigen_cycle
    int bin_words = 1;
    double* csum[2] = {nullptr, nullptr};
    double* csum2[2] = {nullptr, nullptr};
    double* norm_score = nullptr;
    int* counters_i = nullptr;             // [0]: queue A count, [1]: queue B count
    Counters* counters = nullptr;
    DevStatus* status = nullptr;
    double* sum_y = nullptr;
    double* sum_y2 = nullptr;
    int64_t* sum_m = nullptr;
    double* sum_out = nullptr;
    int64_t sum_cap = 0;
    // normSize scratch
    double* rn = nullptr;
    uint64_t* rn_bits = nullptr;
    uint64_t* rn_sorted = nullptr;
    int* flag = nullptr;
    int* pos = nullptr;
    int* start_of = nullptr;
    int* end_of = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    void* rsort_tmp = nullptr;
    size_t rsort_tmp_bytes = 0;
    int n_cu = 256;
    int use_lds = 0;
    int lds_bytes = 0;
    scn_stats stats{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device-resident cycles (cycle_graph): control record, per-history site counts and offsets,
    // scan block sums, contiguous bank and site weights, the draw histogram and candidates of the
    // threshold selection, one captured graph per phase (0 inactive, 1 active)
    EigCtl* ctl = nullptr;
    int* hist_nsite = nullptr;
    int* site_off = nullptr;
    int* scan_bsum = nullptr;
    int* scan_bsum2 = nullptr;
    double* wgt_bank = nullptr;
    double* wgt_sites = nullptr;
    double* sel_cand = nullptr;
    int* sel_below = nullptr;
    int sel_cap = 0;
    int64_t* sum_mtab = nullptr;
    double* k_hist_dev = nullptr;
    int k_hist_cap = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> cyc_ev;
    int dev_cycles = 0;     // 1 while a cycle_graph run is open
    int dc_occ[2] = {0, 0}; // resident transport blocks per CU in each phase's graph
    // tuning
    int k0 = 0;             // 0: one launch per cycle runs every history to completion
    int tail_frac = 16;
    int chunk = 256;        // histories a wave takes from the input at once
    int refill = DEFAULT_EIGEN_REFILL;   // idle lanes a wave gathers before it starts new histories
};

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            E->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
            return SCN_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)


// every copy goes through the engine stream and waits for it (kernels run on that stream)
#define COPY_SYNC(dst, src, bytes, kind)                                                  \
    do {                                                                                 \
        HIPCHK(hipMemcpyAsync((dst), (src), (bytes), (kind), E->stream));                \
        HIPCHK(hipStreamSynchronize(E->stream));                                         \
    } while (0)

template <typename T>
int dalloc(Engine* E, T** p, size_t count)
{
    size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        E->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
        *p = nullptr;
        return SCN_ERR_HIP;
    }
    E->owned.push_back({q, bytes});
    *p = (T*)q;
    return 0;
}

template <typename T>
int upload(Engine* E, T** d, const T* h, size_t count)
{
    int rc = dalloc(E, d, count);
    if (rc) return rc;
    if (count && h) COPY_SYNC(*d, h, count * sizeof(T), hipMemcpyHostToDevice);
    return 0;
}

int alloc_live(Engine* E, Live& L, int cap, int nl)
{
    int rc = 0;
    rc |= dalloc(E, &L.r, (size_t)nl * 3 * cap);
    rc |= dalloc(E, &L.u, (size_t)3 * cap);
    rc |= dalloc(E, &L.w, cap);
    rc |= dalloc(E, &L.w0, cap);
    rc |= dalloc(E, &L.t, cap);
    rc |= dalloc(E, &L.k, (size_t)4 * cap);
    rc |= dalloc(E, &L.rng, cap);
    rc |= dalloc(E, &L.loc, (size_t)nl * cap);
    rc |= dalloc(E, &L.nest, cap);
    rc |= dalloc(E, &L.mat, cap);
    rc |= dalloc(E, &L.uid, cap);
    rc |= dalloc(E, &L.G, cap);
    rc |= dalloc(E, &L.brood, cap);
    rc |= dalloc(E, &L.nsite, cap);
    rc |= dalloc(E, &L.coll, cap);
    return rc;
}

// the engine error a device status read back to the host names (0: none); clears the device copy
int status_error(Engine* E, const DevStatus& h)
{
    if (h.code) {
        char buf[256];
        const char* what = "engine error";
        switch (h.code) {
        case SCN_ERR_OVERFLOW: what = "Run out of space for particles (site or queue buffer)"; break;
        case SCN_ERR_MATERIAL: what = "Particle is in an undefined or overlapping material"; break;
        case SCN_ERR_NAN: what = "Distance is NaN"; break;
        case SCN_ERR_MAJORANT: what = "Majorant is 0"; break;
        case SCN_ERR_SAMPLING: what = "Sampling failed: unnormalised CDF (scatter or chi)"; break;
        case SCN_ERR_UNSUPPORTED: what = "Unsupported reaction (MT) in collision"; break;
        case SCN_ERR_ARG: what = "Failed to find material cell (diveToMat)"; break;
        }
        snprintf(buf, sizeof buf, "%s [detail %d]", what, h.detail);
        E->err = buf;
        DevStatus z{};
        COPY_SYNC(E->status, &z, sizeof z, hipMemcpyHostToDevice);
        return h.code;
    }
    return 0;
}

int check_status(Engine* E)
{
    DevStatus h{};
    HIPCHK(hipMemcpyAsync(&h, E->status, sizeof h, hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return status_error(E, h);
}

int ensure_sum(Engine* E, int64_t n)
{
    int64_t m = (n + SCN_DSUM_CHUNK - 1) / SCN_DSUM_CHUNK + 1;
    if (m <= E->sum_cap) return 0;
    int rc = 0;
    rc |= dalloc(E, &E->sum_y, (size_t)8 * m);
    rc |= dalloc(E, &E->sum_y2, (size_t)8 * m);
    E->sum_cap = m;
    return rc;
}

// canonical sums of up to 8 strided arrays; results land in E->sum_out[0..njobs)
int dsum_batch(Engine* E, SumJobs J)
{
    int64_t maxn = 0;
    for (int q = 0; q < J.njobs; ++q) maxn = std::max(maxn, J.j[q].n);
    int rc = ensure_sum(E, maxn);
    if (rc) return rc;
    const int64_t ys = E->sum_cap;
    int64_t m[8];
    int64_t mmax = 0;
    for (int q = 0; q < J.njobs; ++q) {
        m[q] = (J.j[q].n + SCN_DSUM_CHUNK - 1) / SCN_DSUM_CHUNK;
        mmax = std::max(mmax, m[q]);
    }
    if (mmax > 0) {
        dim3 g((unsigned)((mmax + 255) / 256), J.njobs);
        hipLaunchKernelGGL(k_dsum_chunks, g, dim3(256), 0, E->stream, J, E->sum_y, ys, (int64_t*)nullptr);
    }
    // the element counts of every pairwise level, worked out on the host and uploaded at once, so
    // the levels run back to back with one synchronisation at the end
    int64_t mt[SCN_DSUM_MAX_LEVELS][8];
    int levels = 0;
    for (int64_t mm = mmax; mm > 2048; mm = (mm + 1) / 2) {
        if (levels + 1 >= SCN_DSUM_MAX_LEVELS) {
            E->err = "sum too long";
            return SCN_ERR_ARG;
        }
        for (int q = 0; q < J.njobs; ++q) {
            mt[levels][q] = m[q];
            m[q] = (m[q] + 1) / 2;
        }
        ++levels;
    }
    for (int q = 0; q < J.njobs; ++q) mt[levels][q] = m[q];
    HIPCHK(hipMemcpyAsync(E->sum_m, mt, sizeof(int64_t) * 8 * (levels + 1), hipMemcpyHostToDevice, E->stream));
    double* a = E->sum_y;
    double* b = E->sum_y2;
    for (int l = 0; l < levels; ++l) {
        int64_t h = 0;
        for (int q = 0; q < J.njobs; ++q) h = std::max(h, (mt[l][q] + 1) / 2);
        dim3 g((unsigned)((h + 255) / 256), J.njobs);
        hipLaunchKernelGGL(k_dsum_level, g, dim3(256), 0, E->stream, a, b, E->sum_m + 8 * l, ys);
        // jobs already at m <= 1 keep their value in place: copy forward through the level kernel
        for (int q = 0; q < J.njobs; ++q)
            if (mt[l][q] == 1)
                HIPCHK(hipMemcpyAsync(b + q * ys, a + q * ys, sizeof(double), hipMemcpyDeviceToDevice, E->stream));
        std::swap(a, b);
    }
    hipLaunchKernelGGL(k_dsum_final, dim3(J.njobs), dim3(1024), 0, E->stream, a, E->sum_m + 8 * levels, ys,
                       E->sum_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(E->stream));
    return 0;
}

// kernel variant: register depth NL and, for cell / rotated geometries, the general kernels
template <int MODE>
TransportFn pick_nl(int nl, int geo, const void** kern)
{
    if (geo) return nl <= 3 ? variant_fn<3, MODE, 1>(kern) : variant_fn<8, MODE, 1>(kern);
    switch (nl) {
    case 1: return variant_fn<1, MODE, 0>(kern);
    case 3: return variant_fn<3, MODE, 0>(kern);
    case 4: return variant_fn<4, MODE, 0>(kern);
    default: return variant_fn<8, MODE, 0>(kern);
    }
}

// transport one cycle's bank (E->bankA, n_in histories) into E->bankB (sites sorted by brood);
// returns the number of sites in *n_sites and the per-history keff scores in E->hist_k.
int transport_cycle(Engine* E, int n_in, uint64_t master, double k_eff, int* n_sites, int score_bins)
{
    if (n_in > E->cap) {
        E->err = "bank larger than the engine capacity";
        return SCN_ERR_OVERFLOW;
    }
    HIPCHK(hipMemsetAsync(E->counters_i, 0, sizeof(int) * 4, E->stream));
    HIPCHK(hipMemsetAsync(E->sites.count, 0, sizeof(int), E->stream));
    const void* kern = nullptr;
    TransportFn fn = (E->transport == SCN_TRANSPORT_DT)   ? pick_nl<SCN_TRANSPORT_DT>(E->nl, E->geo, &kern)
                     : (E->transport == SCN_TRANSPORT_HT) ? pick_nl<SCN_TRANSPORT_HT>(E->nl, E->geo, &kern)
                                                          : pick_nl<SCN_TRANSPORT_ST>(E->nl, E->geo, &kern);
    double* bins = score_bins ? E->bins : nullptr;
    // a null BIN pointer means "do not score": point at a scratch copy instead
    double* bins_arg = bins ? bins : E->bins;
    int use_lds = score_bins ? E->use_lds : 0;
    size_t sh = use_lds ? (size_t)E->lds_bytes : EQUEUE_BYTES;
    int occ = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BLOCK, sh));
    const int persistent = E->n_cu * std::max(1, occ);
    DevProblem P = E->P;
    if (!score_bins) P.n_clerks = 0;
    int n_live = n_in;
    int n_sites_host = 0;
    int cur = 0;
    int it = 0;
    int K = E->k0;
    int first = 1;
    while (n_live > 0) {
        int grid = std::max(1, std::min((n_live + BLOCK - 1) / BLOCK, persistent));
        int tail = K <= 0 || n_live <= std::max(65536, n_in / E->tail_frac);
        int steps = tail ? (1 << 30) : K;
        int* nin_ptr = E->counters_i + cur;
        int* nout_ptr = E->counters_i + (1 - cur);
        HIPCHK(hipMemsetAsync(nout_ptr, 0, sizeof(int), E->stream));
        HIPCHK(hipMemsetAsync(E->counters_i + 2, 0, sizeof(int), E->stream));
        HIPCHK(hipEventRecord(E->ev0, E->stream));
        fn(dim3(grid), dim3(BLOCK), sh, E->stream, P, E->bankA, first, n_in, E->live[cur], nin_ptr, E->live[1 - cur],
           nout_ptr, E->cap, steps, master, k_eff, E->sites, bins_arg, use_lds, E->hist_k, E->cap, E->counters,
           E->status, E->counters_i + 2, E->chunk, CycleIn{nullptr, nullptr, nullptr, E->refill});
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(E->ev1, E->stream));
        // survivors, status and site count come back together, behind one synchronisation
        int host_n = 0;
        DevStatus hs{};
        HIPCHK(hipMemcpyAsync(&host_n, nout_ptr, sizeof(int), hipMemcpyDeviceToHost, E->stream));
        HIPCHK(hipMemcpyAsync(&hs, E->status, sizeof hs, hipMemcpyDeviceToHost, E->stream));
        HIPCHK(hipMemcpyAsync(&n_sites_host, E->sites.count, sizeof(int), hipMemcpyDeviceToHost, E->stream));
        HIPCHK(hipStreamSynchronize(E->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
        E->stats.transport_kernel_ms += ms;
        E->stats.transport_launches += 1;
        int rc = status_error(E, hs);
        if (rc) return rc;
        n_live = host_n;
        cur = 1 - cur;
        first = 0;
        K *= 2;
        ++it;
    }
    const int ns = n_sites_host;   // read with the last launch's survivors
    if (ns > E->site_cap || ns > E->bank_cap) {
        E->err = "Run out of space for fission sites";
        return SCN_ERR_OVERFLOW;
    }
    if (ns > 0) {
        int end_bit = 24;
        while (((int64_t)1 << (end_bit - 24)) <= (int64_t)n_in + 1) ++end_bit;
        size_t need = 0;
        HIPCHK(rocprim::radix_sort_pairs(nullptr, need, E->sites.key, E->keys_alt, E->sites.order, E->order_alt,
                                         (size_t)ns, 0, end_bit, E->stream));
        if (need > E->sort_tmp_bytes) {
            int rc = dalloc(E, (char**)&E->sort_tmp, need);
            if (rc) return rc;
            E->sort_tmp_bytes = need;
        }
        HIPCHK(rocprim::radix_sort_pairs(E->sort_tmp, need, E->sites.key, E->keys_alt, E->sites.order, E->order_alt,
                                         (size_t)ns, 0, end_bit, E->stream));
        hipLaunchKernelGGL(k_gather_sites, dim3((ns + 255) / 256), dim3(256), 0, E->stream, E->sites, E->order_alt, ns,
                           E->bankB, E->keys_alt);
        HIPCHK(hipGetLastError());
    }
    *n_sites = ns;
    E->stats.n_histories += n_in;
    E->stats.n_sites += ns;
    return 0;
}

// sums after a cycle: START (bankA weights), END (bankB weights), 4 keff scores
int cycle_sums(Engine* E, int n_in, int n_out, double out[6])
{
    SumJobs J{};
    J.njobs = 6;
    J.j[0] = {&E->bankA[0].wgt, n_in, (int64_t)(sizeof(scn_particle_state) / sizeof(double))};
    J.j[1] = {&E->bankB[0].wgt, n_out, (int64_t)(sizeof(scn_particle_state) / sizeof(double))};
    for (int k = 0; k < 4; ++k) J.j[2 + k] = {E->hist_k + (size_t)k * E->cap, n_in, 1};
    int rc = dsum_batch(E, J);
    if (rc) return rc;
    COPY_SYNC(out, E->sum_out, sizeof(double) * 6, hipMemcpyDeviceToHost);
    return 0;
}

int close_bins(Engine* E, int which)
{
    if (E->n_bins == 0) return 0;
    const double* ns = nullptr;
    const dim3 g((unsigned)((E->n_bins + 255) / 256));
    if (E->P.fix) {
        unsigned long long* w = (unsigned long long*)E->bins;
        if (E->norm_bin >= 0) {
            hipLaunchKernelGGL(k_fix_scalar, dim3(1), dim3(1), 0, E->stream, (const unsigned long long*)w + 4 * E->norm_bin,
                               E->norm_score);
            ns = E->norm_score;
        }
        hipLaunchKernelGGL(k_close_bins_fix, g, dim3(256), 0, E->stream, w, E->csum[which], E->csum2[which], E->n_bins,
                           ns, E->norm_value);
        HIPCHK(hipGetLastError());
        return 0;
    }
    if (E->norm_bin >= 0) {
        hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, E->stream, E->bins + E->norm_bin, E->norm_score);
        ns = E->norm_score;
    }
    hipLaunchKernelGGL(k_close_bins, g, dim3(256), 0, E->stream, E->bins, E->csum[which], E->csum2[which], E->n_bins,
                       ns, E->norm_value);
    HIPCHK(hipGetLastError());
    return 0;
}

// device buffers of one normSize_Repr pass over at most `n_max` sites
struct NormBufs {
    double* rn;
    uint64_t* rn_bits;
    uint64_t* rn_sorted;
    int* flag;
    int* pos;
    int* start_of;       // [max_brood + 1]
    int* end_of;
    int max_brood;
    void* sort_tmp;      // rocPRIM temporary storage
    size_t sort_tmp_bytes;
    void* scan_tmp;
    size_t scan_tmp_bytes;
};

// temporary storage rocPRIM needs for n sites (a size query: nothing is allocated or launched)
hipError_t norm_tmp_needs(const NormBufs& B, int n, hipStream_t stream, size_t* sort_need, size_t* scan_need)
{
    hipError_t e = rocprim::radix_sort_keys(nullptr, *sort_need, B.rn_bits, B.rn_sorted, (size_t)n, 0, 64, stream);
    if (e != hipSuccess) return e;
    return rocprim::exclusive_scan(nullptr, *scan_need, B.flag, B.pos, 0, (size_t)n, rocprim::plus<int>(), stream);
}

#define NCHK(expr)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return SCN_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

// particleDungeon%normSize_Repr (particleDungeon_class.f90:431-602) of in[0..n), sorted by
// broodID with every broodID in [0, B.max_brood], into out (capacity out_cap).  The temporary
// storage in B must hold what norm_tmp_needs reports for n.  *n_new = population after.
int norm_size_run(const NormBufs& B, hipStream_t stream, const scn_particle_state* in, int n, int tot_pop,
                  uint64_t master, scn_particle_state* out, int out_cap, int* n_new, std::string& err)
{
    if (n == 0) {
        err = "Normalisation failed: no fission sites";
        return SCN_ERR_NORM;
    }
    if (tot_pop > out_cap) {
        err = "normSize_Repr: bank capacity too small";
        return SCN_ERR_OVERFLOW;
    }
    int excess = n - tot_pop;
    int heap = excess < 0 ? ((-excess) % n) : excess;
    if (excess == 0) {
        NCHK(hipMemcpyAsync(out, in, sizeof(scn_particle_state) * n, hipMemcpyDeviceToDevice, stream));
        *n_new = n;
        return 0;
    }
    const unsigned g = (unsigned)((n + 255) / 256);
    uint64_t* thr = nullptr;
    size_t sort_need = B.sort_tmp_bytes, scan_need = B.scan_tmp_bytes;
    if (heap != 0) {
        hipLaunchKernelGGL(k_site_rn, dim3(g), dim3(256), 0, stream, master, n, B.rn, B.rn_bits);
        NCHK(rocprim::radix_sort_keys(B.sort_tmp, sort_need, B.rn_bits, B.rn_sorted, (size_t)n, 0, 64, stream));
        thr = B.rn_sorted + (heap - 1);
    }
    int last[2];
    if (excess > 0) {
        hipLaunchKernelGGL(k_norm_flags, dim3(g), dim3(256), 0, stream, B.rn, n, thr, 1, B.flag);
        NCHK(rocprim::exclusive_scan(B.scan_tmp, scan_need, B.flag, B.pos, 0, (size_t)n, rocprim::plus<int>(), stream));
        hipLaunchKernelGGL(k_norm_keep, dim3(g), dim3(256), 0, stream, in, B.flag, B.pos, n, out, out_cap);
        NCHK(hipMemcpyAsync(&last[0], B.flag + (n - 1), sizeof(int), hipMemcpyDeviceToHost, stream));
        NCHK(hipMemcpyAsync(&last[1], B.pos + (n - 1), sizeof(int), hipMemcpyDeviceToHost, stream));
        NCHK(hipStreamSynchronize(stream));
        *n_new = last[0] + last[1];
    } else {
        int n_copies = (-excess) / n;
        int n_dup = (-excess) % n;
        if (n_dup != 0) {
            hipLaunchKernelGGL(k_norm_flags, dim3(g), dim3(256), 0, stream, B.rn, n, thr, 0, B.flag);
        } else {
            NCHK(hipMemsetAsync(B.flag, 0, sizeof(int) * n, stream));
        }
        NCHK(rocprim::exclusive_scan(B.scan_tmp, scan_need, B.flag, B.pos, 0, (size_t)n, rocprim::plus<int>(), stream));
        hipLaunchKernelGGL(k_brood_bounds, dim3(g), dim3(256), 0, stream, in, n, B.max_brood, B.start_of, B.end_of);
        hipLaunchKernelGGL(k_norm_dup, dim3(g), dim3(256), 0, stream, in, B.flag, B.pos, B.start_of, B.end_of, n,
                           n_copies, B.max_brood, out, out_cap);
        NCHK(hipMemcpyAsync(&last[0], B.flag + (n - 1), sizeof(int), hipMemcpyDeviceToHost, stream));
        NCHK(hipMemcpyAsync(&last[1], B.pos + (n - 1), sizeof(int), hipMemcpyDeviceToHost, stream));
        NCHK(hipStreamSynchronize(stream));
        *n_new = n * (n_copies + 1) + last[0] + last[1];
    }
    NCHK(hipGetLastError());
    if (*n_new != tot_pop) {
        err = "Normalisation failed!";
        return SCN_ERR_NORM;
    }
    return 0;
}
#undef NCHK

// particleDungeon%normSize_Repr on E->bankB[0..n) (sorted by brood); result in E->bankA
int norm_size_device(Engine* E, int n, int tot_pop, uint64_t master, int* n_new)
{
    if ((int64_t)tot_pop > E->bank_cap) {
        E->err = "normSize_Repr: bank capacity too small";
        return SCN_ERR_OVERFLOW;
    }
    NormBufs B{E->rn, E->rn_bits, E->rn_sorted, E->flag, E->pos, E->start_of, E->end_of, E->cap + 1,
               E->rsort_tmp, E->rsort_tmp_bytes, E->scan_tmp, E->scan_tmp_bytes};
    if (n > 0) {
        size_t sort_need = 0, scan_need = 0;
        HIPCHK(norm_tmp_needs(B, n, E->stream, &sort_need, &scan_need));
        if (sort_need > E->rsort_tmp_bytes) {
            int rc = dalloc(E, (char**)&E->rsort_tmp, sort_need);
            if (rc) return rc;
            E->rsort_tmp_bytes = sort_need;
        }
        if (scan_need > E->scan_tmp_bytes) {
            int rc = dalloc(E, (char**)&E->scan_tmp, scan_need);
            if (rc) return rc;
            E->scan_tmp_bytes = scan_need;
        }
        B.sort_tmp = E->rsort_tmp;
        B.sort_tmp_bytes = E->rsort_tmp_bytes;
        B.scan_tmp = E->scan_tmp;
        B.scan_tmp_bytes = E->scan_tmp_bytes;
    }
    return norm_size_run(B, E->stream, E->bankB, n, tot_pop, master, E->bankA, E->bank_cap, n_new, E->err);
}


// ---- device-resident power iteration: host side --------------------------------------------------
// canonical sums (the tree of dsum_batch) with counts in device memory; J.j[q].n are capacities.
// Level 8 of the tree comes out of k_dsum_tree, the rest is paired here while more than 2048
// values remain and finished by k_dcycle_close.  Returns the row of sum_mtab with the last sizes.
int dsum_dev_levels(Engine* E, const SumJobs& J, double** last)
{
    int64_t maxn = 0;
    for (int q = 0; q < J.njobs; ++q) maxn = std::max(maxn, J.j[q].n);
    const int64_t ys = E->sum_cap;
    const int64_t m0 = (maxn + SCN_DSUM_CHUNK - 1) / SCN_DSUM_CHUNK;
    hipLaunchKernelGGL(k_dsum_tree, dim3((unsigned)std::max<int64_t>(1, (m0 + 255) / 256), J.njobs), dim3(256), 0,
                       E->stream, J, E->sum_y, ys, E->sum_mtab);
    int64_t mm = (m0 + 255) / 256;
    double* a = E->sum_y;
    double* b = E->sum_y2;
    int l = 0;
    for (; mm > 2048; mm = (mm + 1) / 2, ++l) {
        const int64_t h = (mm + 1) / 2;
        hipLaunchKernelGGL(k_dsum_level_dev, dim3((unsigned)((h + 255) / 256), J.njobs), dim3(256), 0, E->stream, a, b,
                           E->sum_mtab, l, ys);
        std::swap(a, b);
    }
    *last = a;
    return l;
}

// one cycle as a fixed launch sequence on E->stream (captured into a graph by dcycle_setup)
int enqueue_dcycle(Engine* E, int phase, int score_bins)
{
    const int pop = E->eig.S.pop;
    const void* kern = nullptr;
    TransportFn fn = (E->transport == SCN_TRANSPORT_DT)   ? pick_nl<SCN_TRANSPORT_DT>(E->nl, E->geo, &kern)
                     : (E->transport == SCN_TRANSPORT_HT) ? pick_nl<SCN_TRANSPORT_HT>(E->nl, E->geo, &kern)
                                                          : pick_nl<SCN_TRANSPORT_ST>(E->nl, E->geo, &kern);
    const int use_lds = score_bins ? E->use_lds : 0;
    const size_t sh = use_lds ? (size_t)E->lds_bytes : EQUEUE_BYTES;
    const int grid = std::max(1, std::min((pop + BLOCK - 1) / BLOCK, E->n_cu * std::max(1, E->dc_occ[phase])));
    DevProblem P = E->P;
    if (!score_bins) P.n_clerks = 0;
    EigCtl* c = E->ctl;
    // the transport's queue counters and the site count were cleared by the previous cycle's end
    fn(dim3(grid), dim3(BLOCK), sh, E->stream, P, E->bankA, 1, pop, E->live[0], E->counters_i, E->live[1],
       E->counters_i + 1, E->cap, 1 << 30, 0, 0.0, E->sites, E->bins, use_lds, E->hist_k, E->cap, E->counters,
       E->status, E->counters_i + 2, E->chunk, CycleIn{&c->master, &c->k_new, E->hist_nsite, E->refill});
    // offsets of each history's sites in the next bank (broodID, emission order), their weights
    const int hb = (pop + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK;
    hipLaunchKernelGGL(k_scan_blocks, dim3(std::max(1, hb)), dim3(256), 0, E->stream, (const int*)E->hist_nsite,
                       E->site_off, E->scan_bsum, (const int*)&c->pop);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, E->stream, E->scan_bsum, (const int*)&c->pop);
    const int sg = (E->site_cap + 255) / 256;
    hipLaunchKernelGGL(k_site_index, dim3(sg), dim3(256), 0, E->stream, E->sites, (const int*)E->site_off,
                       (const int*)E->scan_bsum, pop, E->bank_cap, c, E->wgt_sites, E->status);
    // START / END population weights and the keffImplicit scores, the k estimates, the plan
    SumJobs J{};
    J.njobs = 6;
    J.j[0] = {E->wgt_bank, pop, 1, nullptr};
    J.j[1] = {E->wgt_sites, E->bank_cap, 1, &c->n_sites};
    for (int k = 0; k < 4; ++k) J.j[2 + k] = {E->hist_k + (size_t)k * E->cap, pop, 1, nullptr};
    double* last = nullptr;
    const int l = dsum_dev_levels(E, J, &last);
    hipLaunchKernelGGL(k_dcycle_close, dim3(1), dim3(1024), 0, E->stream, (const double*)last,
                       (const int64_t*)(E->sum_mtab + 8 * l), E->sum_cap, J.njobs, c, phase, E->status);
    if (score_bins) {
        int rc = close_bins(E, phase == 1 ? 0 : 1);
        if (rc) return rc;
    }
    // normSize_Repr: the draws, the heap-th smallest of them, the flags, the next bank
    const int bg = (E->bank_cap + 255) / 256;
    hipLaunchKernelGGL(k_site_draws, dim3(bg), dim3(256), 0, E->stream, c, E->rn, E->sel_cand, E->sel_cap,
                       E->sel_below);
    hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(1024), 0, E->stream, (const double*)E->rn, c,
                       (const double*)E->sel_cand, E->sel_cap, (const int*)E->sel_below);
    const int fb = (E->bank_cap + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK;
    hipLaunchKernelGGL(k_norm_flag_scan, dim3(std::max(1, fb)), dim3(256), 0, E->stream, (const double*)E->rn,
                       (const EigCtl*)c, E->flag, E->pos, E->scan_bsum2);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, E->stream, E->scan_bsum2, (const int*)&c->n_sites);
    hipLaunchKernelGGL(k_norm_out, dim3(sg), dim3(256), 0, E->stream, E->sites, (const int*)E->site_off,
                       (const int*)E->scan_bsum, (const int*)E->hist_nsite, (const int*)E->flag, (const int*)E->pos,
                       (const int*)E->scan_bsum2, (const EigCtl*)c, pop, E->bankA, E->wgt_bank, E->bank_cap);
    hipLaunchKernelGGL(k_dcycle_end, dim3(1), dim3(256), 0, E->stream, c, (const int*)E->flag, (const int*)E->pos,
                       (const int*)E->scan_bsum2, E->k_hist_dev, E->k_hist_cap, phase, E->counters_i,
                       E->sites.count, E->status);
    HIPCHK(hipGetLastError());
    return 0;
}

void dcycle_release_graphs(Engine* E)
{
    for (int ph = 0; ph < 2; ++ph) {
        if (E->gexec[ph]) (void)hipGraphExecDestroy(E->gexec[ph]);
        if (E->graph[ph]) (void)hipGraphDestroy(E->graph[ph]);
        E->gexec[ph] = nullptr;
        E->graph[ph] = nullptr;
    }
}

// buffers, initial control record and the two phase graphs of a cycle_graph run
int dcycle_setup(Engine* E, const scn_eigen_settings& S)
{
    int rc = 0;
    if (!E->ctl) {
        rc |= dalloc(E, &E->ctl, 1);
        rc |= dalloc(E, &E->hist_nsite, E->cap);
        rc |= dalloc(E, &E->site_off, E->cap);
        rc |= dalloc(E, &E->scan_bsum, (size_t)std::max(E->site_cap, E->cap) / SCAN_PER_BLOCK + 2);
        rc |= dalloc(E, &E->scan_bsum2, (size_t)E->bank_cap / SCAN_PER_BLOCK + 2);
        rc |= dalloc(E, &E->wgt_bank, (size_t)E->bank_cap);
        rc |= dalloc(E, &E->wgt_sites, (size_t)E->bank_cap);
        // the selection window holds at most 8 sqrt(n) + 32 draws on average: room for twice that
        // (more falls back to a pass over all draws, with the same result)
        E->sel_cap = std::max(1 << 16, (int)(16.0 * std::sqrt((double)E->bank_cap)) + 4096);
        rc |= dalloc(E, &E->sel_cand, (size_t)E->sel_cap);
        rc |= dalloc(E, &E->sel_below, (size_t)E->bank_cap / 256 + 1);
        rc |= dalloc(E, &E->sum_mtab, (size_t)8 * (SCN_DSUM_MAX_LEVELS + 2));
        if (rc) return rc;
    }
    const int ncyc = S.n_inactive + S.n_active;
    if (ncyc > E->k_hist_cap) {
        rc = dalloc(E, &E->k_hist_dev, (size_t)ncyc);
        if (rc) return rc;
        E->k_hist_cap = ncyc;
    }
    while ((int)E->cyc_ev.size() < ncyc + 1) {
        hipEvent_t ev;
        HIPCHK(hipEventCreate(&ev));
        E->cyc_ev.push_back(ev);
    }
    rc = ensure_sum(E, (int64_t)E->bank_cap);
    if (rc) return rc;
    const EigenState& X = E->eig;
    EigCtl h{};
    h.master = X.master;
    h.k_new = X.k_new;
    h.keff_0 = X.keff_0;
    h.dk[0] = X.dk[0];
    h.dk[1] = X.dk[1];
    h.cur = X.cur;
    h.pop = S.pop;
    h.sel_sigmas = S.select_window > 0.0 ? S.select_window : 8.0;
    COPY_SYNC(E->ctl, &h, sizeof h, hipMemcpyHostToDevice);
    // state the first captured cycle expects: cleared counters (the selection counts start at zero
    // in the control record), the weights of the starting bank
    hipLaunchKernelGGL(k_dcycle_begin, dim3(1), dim3(64), 0, E->stream, E->counters_i, E->sites.count);
    hipLaunchKernelGGL(k_bank_weights, dim3((S.pop + 255) / 256), dim3(256), 0, E->stream,
                       (const scn_particle_state*)E->bankA, S.pop, E->wgt_bank);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(E->stream));
    dcycle_release_graphs(E);
    for (int ph = 0; ph < 2; ++ph) {
        const int score_bins = (ph == 1 || S.inactive_tally) && E->n_bins > 0;
        // resident blocks of the transport kernel, worked out before the capture starts
        const void* kern = nullptr;
        if (E->transport == SCN_TRANSPORT_DT) pick_nl<SCN_TRANSPORT_DT>(E->nl, E->geo, &kern);
        else if (E->transport == SCN_TRANSPORT_HT) pick_nl<SCN_TRANSPORT_HT>(E->nl, E->geo, &kern);
        else pick_nl<SCN_TRANSPORT_ST>(E->nl, E->geo, &kern);
        const size_t sh = (score_bins && E->use_lds) ? (size_t)E->lds_bytes : EQUEUE_BYTES;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&E->dc_occ[ph], kern, BLOCK, sh));
        HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_dcycle(E, ph, score_bins);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(E->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) {
            E->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
            return SCN_ERR_HIP;
        }
        E->graph[ph] = g;
        HIPCHK(hipGraphInstantiate(&E->gexec[ph], g, nullptr, nullptr, 0));
    }
    return 0;
}

// run up to n cycles from the graphs; the control record comes back to the host once at the end
int dcycle_step(Engine* E, int n_cycles, int* done)
{
    EigenState& X = E->eig;
    const scn_eigen_settings& S = X.S;
    int ran = 0;
    HIPCHK(hipEventRecord(E->cyc_ev[0], E->stream));
    while (ran < n_cycles && X.phase < 2) {
        const int ncyc = X.phase == 0 ? S.n_inactive : S.n_active;
        if (X.c_in_phase >= ncyc) {
            X.phase += 1;
            X.c_in_phase = 0;
            continue;
        }
        HIPCHK(hipGraphLaunch(E->gexec[X.phase], E->stream));
        HIPCHK(hipEventRecord(E->cyc_ev[ran + 1], E->stream));
        X.c_in_phase += 1;
        ++ran;
    }
    HIPCHK(hipStreamSynchronize(E->stream));
    int rc = check_status(E);
    if (rc) return rc;
    EigCtl h{};
    COPY_SYNC(&h, E->ctl, sizeof h, hipMemcpyDeviceToHost);
    const int first = (int)X.k_history.size();
    std::vector<double> kh(ran > 0 ? ran : 1);
    if (ran > 0) COPY_SYNC(kh.data(), E->k_hist_dev + first, sizeof(double) * ran, hipMemcpyDeviceToHost);
    for (int c = 0; c < ran; ++c) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, E->cyc_ev[c], E->cyc_ev[c + 1]));
        X.k_history.push_back(kh[c]);
        X.cycle_time.push_back(ms * 1e-3);
        X.transport_s += ms * 1e-3;
        E->stats.cycle_ms += ms;
    }
    X.master = h.master;
    X.k_new = h.k_new;
    X.keff_0 = h.keff_0;
    X.dk[0] = h.dk[0];
    X.dk[1] = h.dk[1];
    X.cur = h.cur;
    X.kA[0] = h.kA[0];
    X.kA[1] = h.kA[1];
    X.kI[0] = h.kI[0];
    X.kI[1] = h.kI[1];
    for (int k = 0; k < 4; ++k) {
        X.imp[k][0] = h.imp[k][0];
        X.imp[k][1] = h.imp[k][1];
    }
    X.nA = h.nA;
    X.nI = h.nI;
    X.n_this = S.pop;
    X.n_hist += (int64_t)S.pop * ran;
    E->stats.n_histories += (int64_t)S.pop * ran;
    E->stats.n_sites += (int64_t)(h.sites_total - X.sites_dev);
    X.sites_dev = h.sites_total;
    E->stats.n_cycles += ran;
    if (done) *done = ran;
    return 0;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

static thread_local std::string g_create_err;

const char* scn_last_error(const void* engine)
{
    if (!engine) return g_create_err.c_str();
    return ((const Engine*)engine)->err.c_str();
}

int scn_device_info(char* name, int32_t name_len, int32_t* n_cu)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SCN_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return SCN_ERR_HIP;
    if (name && name_len > 0) {
        if (prop.name[0]) snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
        else snprintf(name, name_len, "%s", prop.gcnArchName);   // the runtime may leave the marketing name empty
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    return 0;
}

int scn_create(const scn_problem* pr, int32_t capacity, void** out)
{
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_err = "no GPU visible to HIP";
        return SCN_ERR_NO_DEVICE;
    }
    if (!pr || capacity <= 0 || pr->mg.n_mat <= 0 || pr->mg.n_groups <= 0 || pr->n_universes <= 0) {
        g_create_err = "invalid problem tables";
        return SCN_ERR_ARG;
    }
    Engine* E = new Engine();
    const int nG = pr->mg.n_groups, nM = pr->mg.n_mat;
    E->nG = nG;
    E->transport = pr->transport;
    E->n_bins = pr->n_bins;
    E->norm_bin = pr->norm_bin;
    E->norm_value = pr->norm_value;
    if (const char* m = check_cell_tables(pr)) {
        g_create_err = m;
        delete E;
        return SCN_ERR_ARG;
    }
    if (pr->ww_n > 0) {   // the eigenvalue engine has no per-history buffer for split particles
        g_create_err = "weight windows are supported in fixed-source (scn_batch_*) runs only";
        delete E;
        return SCN_ERR_UNSUPPORTED;
    }
    {
        int d = problem_depth(pr);
        if (d < 1 || d > SCN_MAX_NEST) {
            g_create_err = "geometry nesting deeper than SCN_MAX_NEST or malformed graph";
            delete E;
            return SCN_ERR_UNSUPPORTED;
        }
        E->geo = needs_full_geometry(pr);
        E->nl = compiled_nl(d, E->geo);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) E->n_cu = prop.multiProcessorCount;
    int rc = 0;
    if (hipStreamCreate(&E->stream) != hipSuccess) rc = SCN_ERR_HIP;
    if (!rc && (hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess)) rc = SCN_ERR_HIP;
    scn_universe* d_unis = nullptr;
    scn_box* d_boxes = nullptr;
    int *d_gf = nullptr, *d_gi = nullptr, *d_ints = nullptr, *d_fiss = nullptr, *d_goff = nullptr, *d_gidx = nullptr;
    double *d_real = nullptr, *d_xs = nullptr, *d_P0 = nullptr, *d_prod = nullptr, *d_chi = nullptr, *d_maj = nullptr;
    scn_clerk* d_clerks = nullptr;
    // clerks admitted per group (energyFilter evaluated once, host side)
    std::vector<int> goff, gidx, tgoff, tgidx;
    group_clerks(pr, goff, gidx);
    group_clerks(pr, tgoff, tgidx, SCN_CLERK_TRACK);
    int *d_tgoff = nullptr, *d_tgidx = nullptr;
    scn_surface* d_surfs = nullptr;
    if (!rc) rc |= upload(E, &d_unis, pr->universes, pr->n_universes);
    if (!rc) rc |= upload(E, &d_surfs, pr->surfaces, (size_t)std::max(1, pr->n_surfaces));
    if (!rc) rc |= upload(E, &d_boxes, pr->boxes, pr->n_boxes);
    if (!rc) rc |= upload(E, &d_gf, pr->graph_fill, pr->graph_size);
    if (!rc) rc |= upload(E, &d_gi, pr->graph_id, pr->graph_size);
    if (!rc) rc |= upload(E, &d_real, pr->real_data, std::max(1, pr->n_real));
    if (!rc) rc |= upload(E, &d_ints, pr->int_data, std::max(1, pr->n_int));
    if (!rc) rc |= upload(E, &d_xs, pr->mg.xs, (size_t)nM * nG * 6);
    if (!rc) rc |= upload(E, &d_P0, pr->mg.P0, (size_t)nM * nG * nG);
    if (!rc) rc |= upload(E, &d_prod, pr->mg.prod, (size_t)nM * nG * nG);
    if (!rc) rc |= upload(E, &d_chi, pr->mg.chi, (size_t)nM * nG);
    if (!rc) rc |= upload(E, &d_fiss, pr->mg.fissile, (size_t)nM);
    const std::vector<double> maj_tab = majorant_table(pr);
    if (!rc) rc |= upload(E, &d_maj, maj_tab.data(), maj_tab.size());
    const std::vector<double> st_tab = st_track_table(pr);
    double* d_stxs = nullptr;
    if (!rc) rc |= upload(E, &d_stxs, st_tab.data(), st_tab.size());
    if (!rc) rc |= upload(E, &d_clerks, pr->clerks, std::max(1, pr->n_clerks));
    if (!rc) rc |= upload(E, &d_goff, goff.data(), goff.size());
    if (!rc) rc |= upload(E, &d_gidx, gidx.data(), std::max<size_t>(1, gidx.size()));
    if (!rc) rc |= upload(E, &d_tgoff, tgoff.data(), tgoff.size());
    if (!rc) rc |= upload(E, &d_tgidx, tgidx.data(), std::max<size_t>(1, tgidx.size()));
    std::vector<int64_t> g_addr;
    const FastColl fc = fast_collision_layout(pr, goff, gidx, g_addr);
    int64_t* d_gaddr = nullptr;
    if (!rc) rc |= upload(E, &d_gaddr, g_addr.data(), g_addr.size());
    const std::vector<double> uinv = universe_inverse_table(pr);
    double* d_uinv = nullptr;
    if (!rc) rc |= upload(E, &d_uinv, uinv.data(), uinv.size());
    std::vector<int> flat;
    const int flat_lat = flat_lattice(pr, flat);
    int* d_flat = nullptr;
    if (!rc) rc |= upload(E, &d_flat, flat.data(), flat.size());
    DevProblem& P = E->P;
    P.fc = fc;
    P.fc.g_addr = d_gaddr;
    P.uni_inv = d_uinv;
    P.flat_lat = flat_lat;
    P.flat = d_flat;
    P.transport = pr->transport;
    P.st_cache = pr->st_cache;
    P.root_uni = pr->root_uni;
    P.border_box = pr->border_box;
    P.n_mat = nM;
    P.nG = nG;
    P.n_clerks = pr->n_clerks;
    P.graph_size = pr->graph_size;
    P.unis = d_unis;
    P.boxes = d_boxes;
    P.gfill = d_gf;
    P.gid = d_gi;
    P.real = d_real;
    P.ints = d_ints;
    P.xs = d_xs;
    P.P0 = d_P0;
    P.prod = d_prod;
    P.chi = d_chi;
    P.fissile = d_fiss;
    P.maj = d_maj;
    P.st_xs = d_stxs;
    P.collision_xs = pr->mg.collision_xs;
    P.clerks = d_clerks;
    P.gclerk_off = d_goff;
    P.gclerk_idx = d_gidx;
    P.n_bins = pr->n_bins;
    P.ht_cutoff = pr->ht_cutoff;
    P.surfs = d_surfs;
    P.collision = pr->collision;
    P.max_split = pr->max_split;
    P.ww_n = 0;   // weight windows are a fixed-source feature (rejected for eigenvalue runs)
    P.any_rot = any_rotated(pr);
    P.n_track = (int)tgidx.size();
    P.gtrack_off = d_tgoff;
    P.gtrack_idx = d_tgidx;
    if (pr->mg.P1) {
        double* d_P1 = nullptr;
        if (!rc) rc |= upload(E, &d_P1, pr->mg.P1, (size_t)nM * nG * nG);
        P.P1 = d_P1;
    }
    // buffers
    const int cap = capacity;
    E->cap = cap;
    E->site_cap = 4 * cap + 1024;
    E->bank_cap = 4 * cap + 1024;
    if (!rc) rc |= alloc_live(E, E->live[0], cap, E->nl);
    if (!rc) rc |= alloc_live(E, E->live[1], cap, E->nl);
    Sites& S = E->sites;
    S.cap = E->site_cap;
    if (!rc) rc |= dalloc(E, &S.r, (size_t)3 * S.cap);
    if (!rc) rc |= dalloc(E, &S.u, (size_t)3 * S.cap);
    if (!rc) rc |= dalloc(E, &S.w, S.cap);
    if (!rc) rc |= dalloc(E, &S.t, S.cap);
    if (!rc) rc |= dalloc(E, &S.G, S.cap);
    if (!rc) rc |= dalloc(E, &S.mat, S.cap);
    if (!rc) rc |= dalloc(E, &S.uid, S.cap);
    if (!rc) rc |= dalloc(E, &S.coll, S.cap);
    if (!rc) rc |= dalloc(E, &S.key, S.cap);
    if (!rc) rc |= dalloc(E, &S.order, S.cap);
    if (!rc) rc |= dalloc(E, &S.count, 1);
    if (!rc) rc |= dalloc(E, &E->keys_alt, S.cap);
    if (!rc) rc |= dalloc(E, &E->order_alt, S.cap);
    if (!rc) rc |= dalloc(E, &E->bankA, E->bank_cap);
    if (!rc) rc |= dalloc(E, &E->bankB, E->bank_cap);
    if (!rc) rc |= dalloc(E, &E->hist_k, (size_t)4 * cap);
    P.fix = pr->fixed_point_tally ? 1 : 0;
    E->bin_words = P.fix ? 4 : 1;   // doubles, or four limb words per bin
    if (!rc) rc |= dalloc(E, &E->bins, (size_t)E->bin_words * std::max<int64_t>(1, E->n_bins));
    if (!rc && P.fix) rc |= dalloc(E, &E->bins_out, std::max<int64_t>(1, E->n_bins));
    for (int w = 0; w < 2 && !rc; ++w) {
        rc |= dalloc(E, &E->csum[w], std::max<int64_t>(1, E->n_bins));
        rc |= dalloc(E, &E->csum2[w], std::max<int64_t>(1, E->n_bins));
    }
    if (!rc) rc |= dalloc(E, &E->norm_score, 1);
    if (!rc) rc |= dalloc(E, &E->counters_i, 4);
    if (!rc) rc |= dalloc(E, &E->counters, 1);
    if (!rc) rc |= dalloc(E, &E->status, 1);
    if (!rc) rc |= dalloc(E, &E->sum_m, 8 * SCN_DSUM_MAX_LEVELS);
    if (!rc) rc |= dalloc(E, &E->sum_out, 8);
    if (!rc) rc |= dalloc(E, &E->rn, E->bank_cap);
    if (!rc) rc |= dalloc(E, &E->rn_bits, E->bank_cap);
    if (!rc) rc |= dalloc(E, &E->rn_sorted, E->bank_cap);
    if (!rc) rc |= dalloc(E, &E->flag, E->bank_cap);
    if (!rc) rc |= dalloc(E, &E->pos, E->bank_cap);
    if (!rc) rc |= dalloc(E, &E->start_of, (size_t)cap + 2);
    if (!rc) rc |= dalloc(E, &E->end_of, (size_t)cap + 2);
    if (!rc) rc |= ensure_sum(E, (int64_t)E->bank_cap);
    if (!rc) {
        if (hipMemset(E->bins, 0, sizeof(double) * E->bin_words * std::max<int64_t>(1, E->n_bins)) != hipSuccess ||
            hipMemset(E->status, 0, sizeof(DevStatus)) != hipSuccess ||
            hipMemset(E->counters, 0, sizeof(Counters)) != hipSuccess)
            rc = SCN_ERR_HIP;
    }
    E->use_lds = (E->n_bins > 0 && E->n_bins * E->bin_words <= LDS_BINS_MAX) ? 1 : 0;
    E->lds_bytes = E->use_lds ? (int)(E->n_bins * E->bin_words * sizeof(double)) : 0;
    if (rc) {
        g_create_err = E->err.empty() ? "engine creation failed" : E->err;
        scn_destroy(E);
        return rc;
    }
    *out = E;
    return 0;
}

void scn_destroy(void* engine)
{
    Engine* E = (Engine*)engine;
    if (!E) return;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    dcycle_release_graphs(E);
    for (hipEvent_t ev : E->cyc_ev) (void)hipEventDestroy(ev);
    for (auto& b : E->owned) (void)hipFree(b.p);
    if (E->ev0) (void)hipEventDestroy(E->ev0);
    if (E->ev1) (void)hipEventDestroy(E->ev1);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
}

int scn_get_stats(void* engine, scn_stats* out)
{
    Engine* E = (Engine*)engine;
    if (!E || !out) return SCN_ERR_ARG;
    Counters c{};
    COPY_SYNC(&c, E->counters, sizeof c, hipMemcpyDeviceToHost);
    *out = E->stats;
    out->n_segments = (int64_t)c.segments;
    out->n_collisions = (int64_t)c.collisions;
    out->n_scores = (int64_t)c.scores;
    return 0;
}

int scn_reset_stats(void* engine)
{
    Engine* E = (Engine*)engine;
    if (!E) return SCN_ERR_ARG;
    E->stats = scn_stats{};
    HIPCHK(hipMemsetAsync(E->counters, 0, sizeof(Counters), E->stream));
    return 0;
}

int scn_eigen_cycle(void* engine, const scn_particle_state* bank_in, int32_t n_in, int64_t rng_state, double k_eff,
                    scn_particle_state* bank_out, int32_t bank_out_cap, int32_t* n_out, double* bin_scores,
                    double* keff_scores)
{
    Engine* E = (Engine*)engine;
    if (!E || !bank_in || n_in < 0 || !n_out) return SCN_ERR_ARG;
    if (n_in > E->cap) {
        E->err = "bank larger than the engine capacity";
        return SCN_ERR_OVERFLOW;
    }
    COPY_SYNC(E->bankA, bank_in, sizeof(scn_particle_state) * n_in, hipMemcpyHostToDevice);
    if (E->n_bins) HIPCHK(hipMemsetAsync(E->bins, 0, sizeof(double) * E->bin_words * E->n_bins, E->stream));
    int ns = 0;
    int rc = transport_cycle(E, n_in, (uint64_t)rng_state, k_eff, &ns, bin_scores != nullptr);
    if (rc) return rc;
    if (ns > bank_out_cap) {
        E->err = "Run out of space for particles in the output bank";
        return SCN_ERR_OVERFLOW;
    }
    COPY_SYNC(bank_out, E->bankB, sizeof(scn_particle_state) * ns, hipMemcpyDeviceToHost);
    *n_out = ns;
    if (keff_scores) {
        double sums[6];
        rc = cycle_sums(E, n_in, ns, sums);
        if (rc) return rc;
        for (int k = 0; k < 4; ++k) keff_scores[k] = sums[2 + k];
    }
    if (bin_scores && E->n_bins && E->P.fix) {
        hipLaunchKernelGGL(k_fix_values, dim3((unsigned)((E->n_bins + 255) / 256)), dim3(256), 0, E->stream,
                           (const unsigned long long*)E->bins, E->n_bins, E->bins_out);
        HIPCHK(hipGetLastError());
        COPY_SYNC(bin_scores, E->bins_out, sizeof(double) * E->n_bins, hipMemcpyDeviceToHost);
    } else if (bin_scores && E->n_bins) {
        COPY_SYNC(bin_scores, E->bins, sizeof(double) * E->n_bins, hipMemcpyDeviceToHost);
    }
    return 0;
}

int scn_norm_size_repr(void* engine, scn_particle_state* bank, int32_t* n, int32_t cap, int32_t tot_pop,
                       int64_t rng_state)
{
    Engine* E = (Engine*)engine;
    if (!E || !bank || !n) return SCN_ERR_ARG;
    if (*n > E->bank_cap || tot_pop > cap) {
        E->err = "normSize_Repr: bank capacity too small";
        return SCN_ERR_OVERFLOW;
    }
    // sortByBroodID (particleDungeon_class.f90:923-981): stable, host side for caller-supplied banks
    for (int i = 0; i < *n; ++i) {
        if (bank[i].broodID < 0 || bank[i].broodID > E->cap + 1) {
            E->err = "Brood ID out of range";
            return SCN_ERR_ARG;
        }
    }
    std::stable_sort(bank, bank + *n, [](const scn_particle_state& a, const scn_particle_state& b) {
        return a.broodID < b.broodID;
    });
    COPY_SYNC(E->bankB, bank, sizeof(scn_particle_state) * (*n), hipMemcpyHostToDevice);
    int nn = 0;
    int rc = norm_size_device(E, *n, tot_pop, (uint64_t)rng_state, &nn);
    if (rc) return rc;
    COPY_SYNC(bank, E->bankA, sizeof(scn_particle_state) * nn, hipMemcpyDeviceToHost);
    *n = nn;
    return 0;
}

int scn_fission_source(void* engine, int32_t n, int64_t rng_state, const double bottom[3], const double top[3],
                       int32_t g_in, int32_t attempts, scn_particle_state* bank)
{
    Engine* E = (Engine*)engine;
    if (!E || n < 0 || n > E->bank_cap || !bank) return SCN_ERR_ARG;
    (void)g_in;
    const unsigned g = (unsigned)((n + 255) / 256);
#define SCN_FS(NLV, GEOV)                                                                                    \
    hipLaunchKernelGGL((k_fission_source<NLV, GEOV>), dim3(g), dim3(256), 0, E->stream, E->P, n, (uint64_t)rng_state, \
                       bottom[0], bottom[1], bottom[2], top[0], top[1], top[2], attempts, E->bankA, E->status)
    if (E->geo) {
        if (E->nl <= 3) SCN_FS(3, 1);
        else SCN_FS(8, 1);
    } else {
        switch (E->nl) {
        case 1: SCN_FS(1, 0); break;
        case 3: SCN_FS(3, 0); break;
        case 4: SCN_FS(4, 0); break;
        default: SCN_FS(8, 0);
        }
    }
#undef SCN_FS
    HIPCHK(hipGetLastError());
    int rc = check_status(E);
    if (rc) return rc;
    COPY_SYNC(bank, E->bankA, sizeof(scn_particle_state) * n, hipMemcpyDeviceToHost);
    return 0;
}

int scn_eigen_start(void* engine, const scn_eigen_settings* S, const scn_particle_state* initial_bank)
{
    Engine* E = (Engine*)engine;
    if (!E || !S || !initial_bank) return SCN_ERR_ARG;
    const int pop = S->pop;
    if (pop <= 0 || pop > E->cap) {
        E->err = "pop larger than the engine capacity";
        return SCN_ERR_OVERFLOW;
    }
    EigenState& X = E->eig;
    X = EigenState{};
    X.S = *S;
    if (S->events_per_launch > 0) E->k0 = S->events_per_launch;
    if (S->refill < 0 || S->refill > WAVE) {
        E->err = "refill must be 0 (default) or 1..64";
        return SCN_ERR_ARG;
    }
    E->refill = S->refill > 0 ? S->refill : DEFAULT_EIGEN_REFILL;
    COPY_SYNC(E->bankA, initial_bank, sizeof(scn_particle_state) * pop, hipMemcpyHostToDevice);
    for (int w = 0; w < 2 && E->n_bins; ++w) {
        HIPCHK(hipMemsetAsync(E->csum[w], 0, sizeof(double) * E->n_bins, E->stream));
        HIPCHK(hipMemsetAsync(E->csum2[w], 0, sizeof(double) * E->n_bins, E->stream));
    }
    if (E->n_bins) HIPCHK(hipMemsetAsync(E->bins, 0, sizeof(double) * E->bin_words * E->n_bins, E->stream));
    // eigenPhysicsPackage%run: stride(getOffset) = 0 on one rank; generateInitialState: stride(totalPop)
    X.master = scn_rng_stride((uint64_t)S->seed, 0);
    X.master = scn_rng_stride(X.master, pop);
    X.dk[0] = X.dk[1] = 1.0;
    X.keff_0 = S->keff0;
    X.k_new = X.keff_0;
    X.n_this = pop;
    X.open = 1;
    COPY_SYNC(&X.c0, E->counters, sizeof(Counters), hipMemcpyDeviceToHost);
    E->dev_cycles = 0;
    if (S->cycle_graph) {
        if (E->k0 != 0) {
            E->err = "cycle_graph needs one transport launch per cycle (events_per_launch 0)";
            X.open = 0;
            return SCN_ERR_ARG;
        }
        int rc = dcycle_setup(E, *S);
        if (rc) {
            X.open = 0;
            return rc;
        }
        E->dev_cycles = 1;
    }
    return 0;
}

int scn_eigen_step(void* engine, int32_t n_cycles, int32_t* done)
{
    Engine* E = (Engine*)engine;
    if (!E || !E->eig.open) return SCN_ERR_ARG;
    if (E->dev_cycles) return dcycle_step(E, n_cycles, done);
    EigenState& X = E->eig;
    const scn_eigen_settings& S = X.S;
    const int pop = S.pop;
    int ran = 0;
    while (ran < n_cycles && X.phase < 2) {
        int ncyc = X.phase == 0 ? S.n_inactive : S.n_active;
        if (X.c_in_phase >= ncyc) {
            X.phase += 1;
            X.c_in_phase = 0;
            X.k_new = X.keff_0;   // cycles(): k_new = keff_0 at entry
            continue;
        }
        int score_bins = (X.phase == 1 || S.inactive_tally) && E->n_bins > 0;
        auto t0 = std::chrono::steady_clock::now();
        int ns = 0;
        int rc = transport_cycle(E, X.n_this, X.master, X.k_new, &ns, score_bins);
        if (rc) return rc;
        X.n_hist += X.n_this;
        X.master = scn_rng_stride(X.master, (int64_t)pop + 1);
        double sums[6];
        rc = cycle_sums(E, X.n_this, ns, sums);
        if (rc) return rc;
        double startW = sums[0], endW = sums[1];
        double k_norm = X.dk[1 - X.cur];
        if (X.phase == 0) {   // keffAnalogClerk_class.f90:162-181
            double k_eff = endW / startW * k_norm;
            X.kA[0] = X.kA[0] + k_eff;
            X.kA[1] = X.kA[1] + k_eff * k_eff;
            X.nA++;
        } else {              // keffImplicitClerk_class.f90:306-326
            double k_est = sums[2] / (sums[4] + sums[5] - sums[3]);
            X.kI[0] = X.kI[0] + k_est;
            X.kI[1] = X.kI[1] + k_est * k_est;
            X.nI++;
            for (int k = 0; k < 4; ++k) {
                X.imp[k][0] = X.imp[k][0] + sums[2 + k];
                X.imp[k][1] = X.imp[k][1] + sums[2 + k] * sums[2 + k];
            }
        }
        if (score_bins) {
            rc = close_bins(E, X.phase == 1 ? 0 : 1);
            if (rc) return rc;
        }
        int nn = 0;
        rc = norm_size_device(E, ns, pop, X.master, &nn);
        if (rc) return rc;
        X.master = scn_rng_stride(X.master, 1);
        X.n_this = nn;
        X.cur = 1 - X.cur;
        double kr[2];
        if (X.phase == 0) score_result(X.kA[0], X.kA[1], X.nA, kr);
        else score_result(X.kI[0], X.kI[1], X.nI, kr);
        X.k_new = kr[0];
        X.dk[1 - X.cur] = X.k_new;
        X.keff_0 = X.k_new;
        X.k_history.push_back(X.k_new);
        HIPCHK(hipStreamSynchronize(E->stream));
        auto t1 = std::chrono::steady_clock::now();
        double dt = std::chrono::duration<double>(t1 - t0).count();
        X.cycle_time.push_back(dt);
        X.transport_s += dt;
        E->stats.cycle_ms += dt * 1e3;
        E->stats.n_cycles += 1;
        X.c_in_phase += 1;
        ++ran;
    }
    if (done) *done = ran;
    return 0;
}

int scn_eigen_finish(void* engine, scn_eigen_result* res)
{
    Engine* E = (Engine*)engine;
    if (!E || !res || !E->eig.open) return SCN_ERR_ARG;
    EigenState& X = E->eig;
    res->n_histories = X.n_hist;
    res->transport_time = X.transport_s;
    if (X.nA) score_result(X.kA[0], X.kA[1], X.nA, res->k_analog);
    if (X.nI) {
        score_result(X.kI[0], X.kI[1], X.nI, res->k_implicit);
        for (int k = 0; k < 4; ++k) score_result(X.imp[k][0], X.imp[k][1], X.nI, res->imp[k]);
    }
    for (size_t i = 0; i < X.k_history.size(); ++i) {
        if (res->k_history) res->k_history[i] = X.k_history[i];
        if (res->cycle_time) res->cycle_time[i] = X.cycle_time[i];
    }
    Counters c1{};
    COPY_SYNC(&c1, E->counters, sizeof c1, hipMemcpyDeviceToHost);
    res->n_collisions = (int64_t)(c1.collisions - X.c0.collisions);
    res->n_events = (int64_t)((c1.collisions - X.c0.collisions) + (c1.segments - X.c0.segments));
    for (int w = 0; w < 2 && E->n_bins; ++w) {
        double* dst = w == 0 ? res->bins_active : res->bins_inactive;
        int N = w == 0 ? X.nI : X.nA;
        if (!dst || N == 0 || (w == 1 && !X.S.inactive_tally)) continue;
        std::vector<double> cs(E->n_bins), cs2(E->n_bins);
        COPY_SYNC(cs.data(), E->csum[w], sizeof(double) * E->n_bins, hipMemcpyDeviceToHost);
        COPY_SYNC(cs2.data(), E->csum2[w], sizeof(double) * E->n_bins, hipMemcpyDeviceToHost);
        for (int64_t i = 0; i < E->n_bins; ++i) score_result(cs[i], cs2[i], N, dst + 2 * i);
    }
    X.open = 0;
    return 0;
}

int scn_eigen_run(void* engine, const scn_eigen_settings* S, const scn_particle_state* initial_bank,
                  scn_eigen_result* res)
{
    int rc = scn_eigen_start(engine, S, initial_bank);
    if (rc) return rc;
    int32_t done = 0;
    rc = scn_eigen_step(engine, S->n_inactive + S->n_active, &done);
    if (rc) return rc;
    return scn_eigen_finish(engine, res);
}

// Allocation-free population control (scratch in the static `scratch` array, see above).
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[64];

static int noalloc_scratch(char** base)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        g_create_err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return SCN_ERR_NO_DEVICE;
    }
    if (dev < 0 || dev >= 64) {
        g_create_err = "device index above 63";
        return SCN_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return SCN_ERR_HIP;
        }
        g_noalloc_ptr[dev] = p;
    }
    *base = (char*)g_noalloc_ptr[dev];
    return 0;
}

int scn_norm_size_repr_noalloc(const scn_particle_state* bank_in, int32_t n, scn_particle_state* bank_out,
                               int32_t out_cap, int32_t tot_pop, int64_t rng_state, void* stream, int32_t* n_out)
{
    if (!bank_in || !bank_out || !n_out || n < 0 || tot_pop < 0 || out_cap < 0) {
        g_create_err = "scn_norm_size_repr_noalloc: invalid argument";
        return SCN_ERR_ARG;
    }
    if (n > SCN_NOALLOC_MAX_SITES) {
        g_create_err = "scn_norm_size_repr_noalloc: " + std::to_string(n) + " sites exceed the scratch maximum of " +
                       std::to_string(SCN_NOALLOC_MAX_SITES) + " (SCN_NOALLOC_MAX_SITES)";
        return SCN_ERR_ARG;
    }
    char* q = nullptr;
    int rc = noalloc_scratch(&q);
    if (rc) return rc;
    auto take = [&q](size_t bytes) {
        char* r = q;
        q += noalloc_round(bytes);
        return r;
    };
    const size_t M = SCN_NOALLOC_MAX_SITES;
    NormBufs B{};
    B.rn = (double*)take(sizeof(double) * M);
    B.rn_bits = (uint64_t*)take(sizeof(uint64_t) * M);
    B.rn_sorted = (uint64_t*)take(sizeof(uint64_t) * M);
    B.flag = (int*)take(sizeof(int) * M);
    B.pos = (int*)take(sizeof(int) * M);
    B.start_of = (int*)take(sizeof(int) * (M + 2));
    B.end_of = (int*)take(sizeof(int) * (M + 2));
    B.max_brood = (int)M + 1;
    int* bad = (int*)take(sizeof(int) * 64);
    B.sort_tmp = take(NOALLOC_SORT_TMP);
    B.sort_tmp_bytes = NOALLOC_SORT_TMP;
    B.scan_tmp = take(NOALLOC_SCAN_TMP);
    B.scan_tmp_bytes = NOALLOC_SCAN_TMP;
    hipStream_t st = (hipStream_t)stream;
    std::string& err = g_create_err;
    if (n > 0) {
        // the brood bounds index start_of/end_of by broodID: check range and order first
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, bad, 1, 0);
        hipLaunchKernelGGL(k_check_broods, dim3((n + 255) / 256), dim3(256), 0, st, bank_in, n, B.max_brood, bad);
        int hbad = 0;
        hipError_t e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            err = std::string("scn_norm_size_repr_noalloc: ") + hipGetErrorString(e);
            return SCN_ERR_HIP;
        }
        if (hbad) {
            err = "scn_norm_size_repr_noalloc: broodIDs must be sorted (sortByBroodID) and within [0, " +
                  std::to_string(B.max_brood) + "]";
            return SCN_ERR_ARG;
        }
        size_t sort_need = 0, scan_need = 0;
        e = norm_tmp_needs(B, n, st, &sort_need, &scan_need);
        if (e != hipSuccess) {
            err = std::string("rocprim size query: ") + hipGetErrorString(e);
            return SCN_ERR_HIP;
        }
        if (sort_need > B.sort_tmp_bytes || scan_need > B.scan_tmp_bytes) {
            err = "scn_norm_size_repr_noalloc: rocPRIM temporary storage exceeds the scratch reserve";
            return SCN_ERR_ARG;
        }
    }
    int nn = 0;
    rc = norm_size_run(B, st, bank_in, n, tot_pop, (uint64_t)rng_state, bank_out, out_cap, &nn, err);
    if (rc) return rc;
    *n_out = nn;
    return 0;
}

}  // extern "C"
