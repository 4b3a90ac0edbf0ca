// scn_btransport.h -- the batch engine's transport kernel (k_btransport) and the slot types its
// host code shares with it.
//
// The kernel templates are compiled only where SCN_BTRANSPORT_KERNELS is defined: one translation
// unit per (tracking mode, geometry variant) instantiates pick_bt_variant<MODE, GEO>, so the
// variants compile in parallel and the host code of scn_batch.hip never parses the kernel.
#ifndef SCN_BTRANSPORT_H
#define SCN_BTRANSPORT_H

#include <hip/hip_runtime.h>

#include "scone_gpu.h"
#include "scn_fpmath.h"
#include "scn_device.h"

namespace scn_bat {

using namespace scn;

constexpr int BBLOCK = 256;
constexpr int BWAVE = 64;
constexpr int BQUEUE = 4;   // evicted global scores a lane queues in LDS before its wave adds them
constexpr int MAX_SLOTS = 4096;
constexpr int MATERIAL_SOURCE_ATTEMPTS = 200;   // materialSource_class.f90:168
constexpr int MAX_PROBG = 4096;
constexpr int DEFAULT_BUFFER = 50;   // fixedSourcePhysicsPackage_class.f90:372 'buffer' default
constexpr int MAX_BUFFER = 1024;
constexpr int MAX_WINDOW = 16;        // batches of one solve in flight in one cycle (power of two)
constexpr int DEFAULT_REFILL = 32;    // idle lanes a wave gathers before it starts new histories
constexpr int WINDOW_SHIFT = 4;
// a history still alive after this many (flight, collision) pairs in one launch is reported as lost
// (an error, not a hang)
constexpr int MAX_PAIRS_PER_LAUNCH = 1 << 22;

// live queue of one slot, structure of arrays; position arrays level-major
struct BLive {
    double* r;     // [depth*3][cap]
    double* u;     // [3][cap]
    double* w;
    double* w0;
    double* t;
    uint64_t* rng;
    int* loc;      // [depth][cap]
    int* nest;
    int* mat;
    int* uid;
    int* G;
    int* brood;
    int* coll;
    int* split;    // neutronMGimp splitCount of a parked particle
};

// read-only description of a resident solve
struct SlotDev {
    DevProblem P;
    BLive q[2];
    double* bin;
    double* csum;
    double* csum2;
    uint64_t* hist_rng;
    int* hist_coll;
    const double* probG;
    double src_r[3], src_dir[3], bottom[3], top[3];
    double t_low, t_high;
    double stop_rel, norm_value;
    int64_t stop_bin, norm_bin;
    uint64_t seed;
    int src_type, src_iso, src_G, src_mat, n_probG;
    int pop, n_batches;
    int buffer;         // 'buffer' of the package: depth of a history's LIFO buffer of secondaries
    int window;         // batches transported per cycle; BIN holds one copy of the score memory each
    int adapt;          // > 0: each cycle's window follows the stop bin's trend (projection factor in %)
    uint8_t* src_att;   // material source: attempts rejected before each history's source point,
                        // [window][pop], written by k_bsource (MATERIAL_SOURCE_ATTEMPTS: sampling failed)
};

// mutable state of a slot, written by the kernels
struct SlotState {
    int running;        // 1 while batches remain
    int batch;          // batches closed (scoreMemory batchN)
    int count[2];       // particles written into queue side k by the last launch
    int n_in;           // particles the current launch reads (per batch)
    int n_run;          // batches transported this cycle (batch .. batch + n_run - 1)
    int n_close;        // of those, batches closed (fewer after an early stop)
    int stopped;        // 1 if the early-stop target ended the solve
    int win;            // batches the next cycle transports (adaptive windows); 0 = the slot's window
    double norm[MAX_WINDOW];   // normalisation factor of each batch being closed
    unsigned long long segments, collisions, scores, histories;
};

// control words: [0] work items of the current launch, [2] running slots after the cycle,
// [5] items whose particles are all handed out, [6] chunks of the source pre-pass handed out
constexpr int CTL_WORDS = 8;

// work item of one launch: the histories of one batch of one slot's window (first launch of a
// cycle) or one slot's parked particles (later launches), key = slot * MAX_WINDOW + w
struct Item {
    int key;
    int n;                 // particles
    long long prefix;      // particles of the items before this one
};

typedef void (*BTransportFn)(dim3, dim3, size_t, hipStream_t, const SlotDev*, SlotState*, const Item*, int*, int*, int, int,
                             int, int, int, DevStatus*, int, Sites, int);
typedef void (*BSourceFn)(dim3, dim3, hipStream_t, const SlotDev*, const SlotState*, const Item*, int*, DevStatus*);

// launchers and kernel address of one variant: register depth NL, tracking mode MODE, geometry
// variant GEO (0: lattice/pin/box, 1: cells, rotations, CSG borders); src launches the
// material-source pre-pass of the same geometry variant
template <int NL, int MODE, int GEO>
void bt_variant(BTransportFn* fn, const void** kern, BSourceFn* src);

#ifdef SCN_BTRANSPORT_KERNELS
template <int NL, int GEO>
__device__ __forceinline__ void bload(const BLive& L, int cap, int i, Pt<NL, GEO>& p)
{
    p.nest = L.nest[i];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        if (l < p.nest) {
#pragma unroll
            for (int k = 0; k < 3; ++k) p.r[l][k] = L.r[(size_t)(l * 3 + k) * cap + i];
            p.loc[l] = L.loc[(size_t)l * cap + i];
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) p.r[l][k] = 0.0;
            p.loc[l] = 0;
        }
        p.uni[l] = 0;
        p.root[l] = 0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) p.u[k] = L.u[(size_t)k * cap + i];
    p.w = L.w[i];
    p.w0 = L.w0[i];
    p.t = L.t[i];
    p.rng = L.rng[i];
    p.mat = L.mat[i];
    p.uid = L.uid[i];
    p.G = L.G[i];
    p.brood = L.brood[i];
    p.coll = L.coll[i];
    p.split = L.split[i];
    p.nsite = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) p.k[k] = 0.0;
    p.dead = 0;
    p.fate = FATE_NO;
    p.track = 0.0;
}

template <int NL, int GEO>
__device__ __forceinline__ void bstore(const BLive& L, int cap, int i, const Pt<NL, GEO>& p)
{
    L.nest[i] = p.nest;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        if (l < p.nest) {
#pragma unroll
            for (int k = 0; k < 3; ++k) L.r[(size_t)(l * 3 + k) * cap + i] = p.r[l][k];
            L.loc[(size_t)l * cap + i] = p.loc[l];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) L.u[(size_t)k * cap + i] = p.u[k];
    L.w[i] = p.w;
    L.w0[i] = p.w0;
    L.t[i] = p.t;
    L.rng[i] = p.rng;
    L.mat[i] = p.mat;
    L.uid[i] = p.uid;
    L.G[i] = p.G;
    L.brood[i] = p.brood;
    L.coll[i] = p.coll;
    L.split[i] = p.split;
}

// the point and time of one attempt of materialSource%sampleParticle
// (materialSource_class.f90:160-190): uniform in the bounding box, 4 draws of the source stream
__device__ __forceinline__ void src_point(const SlotDev& D, uint64_t& srng, double r[3], double& time)
{
    double r3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) r3[k] = scn_rng_get(&srng);
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = (D.top[k] - D.bottom[k]) * r3[k] + D.bottom[k];
    time = D.t_low + scn_rng_get(&srng) * (D.t_high - D.t_low);
}

// one attempt: the point placed with u = [1,0,0]; returns the material there
template <int NL, int GEO>
__device__ __forceinline__ int src_attempt(const SlotDev& D, const DevProblem& P, uint64_t& srng, Pt<NL, GEO>& p,
                                           double r[3], double& time, DevStatus* st)
{
    src_point(D, srng, r, time);
#pragma unroll
    for (int k = 0; k < 3; ++k) p.r[0][k] = r[k];
    p.u[0] = 1.0;
    p.u[1] = 0.0;
    p.u[2] = 0.0;
    place_coord(P, p, st);
    return p.mat;
}

// Source particle j = i + 1 of batch `batch` (batch w of the slot's window), then the start of
// its history (fixedSourcePhysicsPackage_class.f90:187-213): the source is sampled on the seed
// strided by 2*pop*batch + j (generate, source_inter.f90:110-116), `src` here, and the history
// runs on the seed strided by 2*pop*batch + pop + j, which is src strided by pop: the map
// (run_a, run_c).  Returns 0 if sampling failed (error raised).
template <int NL, int GEO>
__device__ __forceinline__ int born(const SlotDev& D, const DevProblem& P, int w, int i, uint64_t src, uint64_t run_a,
                                    uint64_t run_c, Pt<NL, GEO>& p, DevStatus* st)
{
    const int64_t j = (int64_t)i + 1;
    uint64_t srng = src;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        p.loc[l] = 0;
        p.uni[l] = 0;
        p.root[l] = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) p.r[l][k] = 0.0;
    }
    double r[3], dir[3], time = 0.0;
    int G;
    if (D.src_type == SCN_SRC_POINT) {
        // configSource%sampleParticle: position, energy-angle, energy (pointSource_class.f90:183-246)
#pragma unroll
        for (int k = 0; k < 3; ++k) r[k] = D.src_r[k];
        if (D.src_iso) {
            double mu = 2.0 * scn_rng_get(&srng) - 1.0;
            double phi = SCN_TWO_PI * scn_rng_get(&srng);
            const double x1[3] = {1.0, 0.0, 0.0};
            scn_rotate(x1, mu, phi, dir);
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) dir[k] = D.src_dir[k];
        }
        if (D.n_probG > 0) {
            double u = scn_rng_get(&srng);
            int g;
            for (g = 1; g <= D.n_probG; ++g) {
                u = u - D.probG[g - 1];
                if (u < 0.0) break;
            }
            if (g > D.n_probG) {
                raise_error(st, SCN_ERR_SAMPLING, 5);
                return 0;
            }
            G = g;
        } else {
            G = D.src_G;
        }
    } else {
        // materialSource%sampleParticle: rejection on whatIsAt(r).  k_bsource ran the rejection on
        // this stream before the launch and kept how many attempts failed; skip their draws and
        // take the accepted attempt's point and time (its material was checked there), so no wave
        // waits on its lanes' longest rejection run
        const int k = D.src_att[(size_t)w * D.pop + i];
        if (k >= MATERIAL_SOURCE_ATTEMPTS) return 0;   // the pre-pass raised the error
        srng = (RNG_SKIP4.a[k] * srng + RNG_SKIP4.c[k]) & SCN_LCG_MASK;   // 4k draws on (k < 256)
        src_point(D, srng, r, time);
        double mu = 2.0 * scn_rng_get(&srng) - 1.0;
        double phi = SCN_TWO_PI * scn_rng_get(&srng);
        const double x1[3] = {1.0, 0.0, 0.0};
        scn_rotate(x1, mu, phi, dir);
        G = D.src_G;
    }
    p.rng = (run_a * src + run_c) & SCN_LCG_MASK;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p.r[0][k] = r[k];
        p.u[k] = dir[k];
    }
    p.w = 1.0;
    p.w0 = 1.0;
    p.G = G;
    p.t = time;
    p.coll = 0;
    p.split = 0;
    p.brood = (int)j;
    p.nsite = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) p.k[k] = 0.0;
    p.dead = 0;
    p.fate = FATE_NO;
    p.track = 0.0;
    place_coord(P, p, st);
    if (p.mat == SCN_OUTSIDE_MAT) {   // pointSource%init: 'Source has been placed outside geometry'
        raise_error(st, SCN_ERR_ARG, 6);
        return 0;
    }
    return 1;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// counters of the wave's current slot, summed over the wave by lane 0
__device__ __forceinline__ void flush_counts(SlotState* SS, Cnt& cn, int lane)
{
    const unsigned long long seg = wave_sum((unsigned long long)cn.seg);
    const unsigned long long col = wave_sum((unsigned long long)cn.col);
    const unsigned long long sc = wave_sum((unsigned long long)cn.sc);
    if (lane == 0) {
        if (seg) atomicAdd(&SS->segments, seg);
        if (col) atomicAdd(&SS->collisions, col);
        if (sc) atomicAdd(&SS->scores, sc);
    }
    cn.seg = cn.col = cn.sc = 0;
}

// add the wave's LDS copy of the low bins to BIN and clear it (all lanes of the wave call this
// at the same point; LDS operations of one wave complete in order)
__device__ __forceinline__ void flush_lds_bins(Bins& B, int lane)
{
    __builtin_amdgcn_wave_barrier();
    if (B.fix) {
        unsigned long long* lw = (unsigned long long*)B.lds;
        for (int64_t i = lane; i < 4 * B.lds_n; i += BWAVE) {
            const unsigned long long v = lw[i];
            if (v) {
                atomicAdd((unsigned long long*)B.glob + i, v);
                lw[i] = 0ull;
            }
        }
    } else {
        for (int64_t i = lane; i < B.lds_n; i += BWAVE) {
            const double v = B.lds[i];
            if (v != 0.0) {
                unsafeAtomicAdd(&B.glob[i], v);
                B.lds[i] = 0.0;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// One launch of the event loop over the tile table.  A wave works on one slot at a time (its
// tables are then wave-uniform).  Every lane holds one particle and advances it by one tracking
// event per iteration (plus the collision that event may lead to), so lanes whose flights are
// short do not wait for long ones.  A lane whose particle dies (or, in a K-limited launch, has run
// K (flight, collision) pairs and is parked in the slot's other queue) takes the next particle of
// the tile at once, so lanes do not idle while others finish.
// When the tile is used up the wave takes the next tile: one of the same slot continues the
// refill directly, one of another slot starts once the live lanes have finished.
template <int NL, int MODE, int GEO>
__global__ void __launch_bounds__(BBLOCK)
k_btransport(const SlotDev* __restrict__ slots, SlotState* ss, const Item* __restrict__ items, int* item_cnt, int* ctl,
             int first, int in_side, int out_side, int chunk, int max_steps, DevStatus* st, int lds_per_wave, Sites SB,
             int refill)
{
    extern __shared__ double lds_all[];
    const int lane = threadIdx.x & (BWAVE - 1);
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    // this wave's private copy of the first lds_per_wave bins of its current slot
    double* wl = lds_all + (size_t)(threadIdx.x / BWAVE) * lds_per_wave;
    for (int i = lane; i < lds_per_wave; i += BWAVE) wl[i] = 0.0;
    // after the block's bin copies, each wave's queues of evicted scores: 2 BQUEUE words per lane
    double* wq = lds_all + (size_t)(BBLOCK / BWAVE) * lds_per_wave + (size_t)(threadIdx.x / BWAVE) * 2 * BQUEUE * BWAVE;
    const int n_items = ctl[0];
    const int lane_gid = blockIdx.x * BBLOCK + threadIdx.x;   // owner of one LIFO buffer in SB
    int key = -1;                    // current slot * MAX_WINDOW + batch of its window (wave-uniform)
    int s = -1;                      // current slot
    int next = 0, end = 0;           // particles [next, end) of the current chunk not yet started
    int pend_s = -1, pend_b = 0, pend_e = 0;   // a chunk of another item, started once the lanes are idle
    int exhausted = n_items == 0;    // every item has been found used up
    // Waves start spread over the items in proportion to their particles (the item holding
    // particle g * total / n_waves) and pull chunks from an item's counter until it is used up,
    // then move to the next item; so a wave switches items (and drains) rarely, and items are
    // shared by as many waves as their size calls for.  A wave leaves after finding every item
    // used up once (or all of them closed), so the walk is bounded.
    int it = 0, visited = 0;
    if (n_items > 0) {
        const int n_waves = gridDim.x * (BBLOCK / BWAVE);
        const int g = blockIdx.x * (BBLOCK / BWAVE) + (int)(threadIdx.x / BWAVE);
        const long long total = items[n_items - 1].prefix + items[n_items - 1].n;
        const long long target = total * (long long)g / (long long)n_waves;
        int lo = 0, hi = n_items;   // items[lo].prefix <= target < items[hi].prefix
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (items[mid].prefix <= target) lo = mid;
            else hi = mid;
        }
        it = __builtin_amdgcn_readfirstlane(lo);
    }
    const SlotDev* D = slots;
    SlotState* SS = ss;
    int n_in = 0, batch = 0, cap = 0;
    Bins B;
    B.lds = wl;
    B.glob = nullptr;
    B.lds_n = 0;
    B.c_addr = -1;
    B.c_acc = 0.0;
    B.q = wq + lane;
    B.qn = 0;
    B.qcap = BQUEUE;
    B.fix = 0;
    Cnt cn{0, 0, 0};
    Pt<NL, GEO> p;
    DistCache<NL> cache;   // surface tracking: the current flight's distance cache
    cache.lvl = 0;
    int alive = 0, steps = 0, nseg = 0;
    uint64_t run_a = 1, run_c = 0;   // stride by pop histories, the slot's (wave-uniform)
    int use_dt = MODE == SCN_TRANSPORT_DT;   // this flight's tracking method (HT: chosen per flight)
    int fresh = 1;                           // the next event starts a flight
    for (int iter = 0;; ++iter) {
        while (next >= end && !exhausted && pend_s < 0) {
            int b = 0;
            if (lane == 0) b = atomicAdd(&item_cnt[it], chunk);
            b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));   // wave-uniform: the slot record loads as scalars
            const int ik = __builtin_amdgcn_readfirstlane(items[it].key);
            const int in_ = __builtin_amdgcn_readfirstlane(items[it].n);
            if (b < in_) {
                if (ik == key) {
                    next = b;
                    end = min(b + chunk, in_);
                } else {
                    pend_s = ik;
                    pend_b = b;
                    pend_e = min(b + chunk, in_);
                }
            } else {
                // the first empty pull of an item closes it
                if (b < in_ + chunk && lane == 0) atomicAdd(&ctl[5], 1);
                if (++visited >= n_items) {
                    exhausted = 1;
                    break;
                }
                it = (it + 1 == n_items) ? 0 : it + 1;
                int closed = 0;
                if (lane == 0) closed = __atomic_load_n(&ctl[5], __ATOMIC_RELAXED);
                if (__builtin_amdgcn_readfirstlane(__shfl(closed, 0)) >= n_items) {
                    exhausted = 1;
                    break;
                }
            }
        }
        const unsigned long long live = __ballot(alive);
        if (pend_s >= 0 && live == 0) {
            if (key >= 0) {
                flush_counts(SS, cn, lane);
                bins_flush(B);
                flush_lds_bins(B, lane);
            }
            key = pend_s;
            pend_s = -1;
            s = key >> WINDOW_SHIFT;
            const int w = key & (MAX_WINDOW - 1);
            D = slots + s;
            SS = ss + s;
            n_in = SS->n_in;
            batch = SS->batch + w;
            cap = D->pop;
            if (first) {   // (a, c) of the stride by pop: s -> a s + c
                run_c = rng_stride(0, D->pop);
                run_a = (rng_stride(1, D->pop) - run_c) & SCN_LCG_MASK;
            }
            // fixed point: four limb words per bin, in BIN and in the wave's LDS copy
            B.fix = D->P.fix;
            B.glob = D->bin + (size_t)w * (size_t)D->P.n_bins * (B.fix ? 4 : 1);
            B.lds_n = min((int64_t)(B.fix ? lds_per_wave / 4 : lds_per_wave), D->P.n_bins);
            SB.depth = D->buffer;
            next = pend_b;
            end = pend_e;
        }
        // new histories start once `refill` lanes are idle (or none is alive): a start runs for
        // many lanes at a time rather than once for every lane that frees up
        if (next < end && (__popcll(~live) >= refill || live == 0)) {
            const unsigned long long idle = ~live;
            const int rank = __popcll(idle & lt_mask);
            const int i = next + rank;
            // the source stream of history `next` (wave-uniform), then each lane's own: its
            // rank more strides on, one map of the first byte table
            uint64_t src = 0;
            if (first) src = rng_stride(D->seed, 2 * (int64_t)D->pop * batch + next + 1);
            if (!alive && i < end) {
                if (first) {
                    src = (RNG_STRIDES.a[0][rank] * src + RNG_STRIDES.c[0][rank]) & SCN_LCG_MASK;
                    alive = born(*D, D->P, key & (MAX_WINDOW - 1), i, src, run_a, run_c, p, st);
                } else {
                    bload(D->q[in_side], cap, i, p);
                    relink(D->P, p);
                    alive = 1;
                }
                p.sp = 0;
                p.lane = lane_gid;
                steps = 0;
                nseg = 0;
                cache.lvl = 0;
                fresh = 1;
            }
            next = min(next + __popcll(idle), end);
        }
        if (__ballot(alive) == 0) {
            if (exhausted && pend_s < 0 && next >= end) break;
            continue;
        }
        int park = 0;
        if (alive) {
            // one tracking event; a real collision is processed in the same iteration
            const DevProblem& P = D->P;
            if (MODE == SCN_TRANSPORT_HT && fresh) use_dt = ht_use_dt(P, p, st);
            fresh = 0;
            int r = use_dt ? dt_step(P, p, B, st, cn) : st_step(P, p, cache, B, st, cn);
            if (r == STEP_MORE && ++nseg >= MAX_SEGMENTS_PER_FLIGHT) {
                raise_error(st, SCN_ERR_ARG, SCN_ERR_LOST_DETAIL);
                r = STEP_DEAD;
            }
            if (r == STEP_COLLIDE) {
                if (collide(P, p, 1.0, B, SB, st, cn)) {
                    cache.lvl = 0;
                    nseg = 0;
                    fresh = 1;
                    if (++steps >= max_steps && p.sp == 0) {
                        park = 1;
                        alive = 0;
                    } else if (steps >= MAX_PAIRS_PER_LAUNCH) {
                        raise_error(st, SCN_ERR_ARG, SCN_ERR_LOST_DETAIL);
                        alive = 0;
                    }
                } else {
                    r = STEP_DEAD;
                }
            }
            if (r == STEP_DEAD && p.sp > 0 && !st->code) {
                // buffer%release: the last detained secondary continues the history on the same
                // RNG stream (fixedSourcePhysicsPackage_class.f90:239-243, particle_class.f90:290-310)
                const size_t e = (size_t)(--p.sp) * SB.stride + p.lane;
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    p.loc[l] = 0;
                    p.uni[l] = 0;
                    p.root[l] = 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) p.r[l][k] = 0.0;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    p.r[0][k] = SB.r[(size_t)k * SB.cap + e];
                    p.u[k] = SB.u[(size_t)k * SB.cap + e];
                }
                p.w = SB.w[e];
                p.w0 = p.w;
                p.t = SB.t[e];
                p.G = SB.G[e];
                p.dead = 0;
                p.fate = FATE_NO;
                p.track = 0.0;
                p.split = 0;   // particle_fromParticleState resets splitCount
                place_coord(P, p, st);
                cache.lvl = 0;
                nseg = 0;
                fresh = 1;
                r = STEP_MORE;
            }
            if (r == STEP_DEAD) {
                alive = 0;
                if (D->hist_rng) {
                    const size_t h = (size_t)batch * D->pop + (size_t)(p.brood - 1);
                    D->hist_rng[h] = p.rng;
                    D->hist_coll[h] = p.coll;
                }
            }
        }
        if (__ballot(B.qn >= B.qcap)) bins_drain(B);   // a full queue: the wave adds all of them
        const unsigned long long pm = __ballot(park);
        if (pm) {
            int obase = 0;
            if (lane == 0) obase = atomicAdd(&SS->count[out_side], __popcll(pm));
            obase = __shfl(obase, 0);
            if (park) {
                const int o = obase + __popcll(pm & lt_mask);
                if (o < cap) bstore(D->q[out_side], cap, o, p);
                else raise_error(st, SCN_ERR_OVERFLOW, s);
            }
        }
        if ((iter & 63) == 63) {
            int e = 0;
            if (lane == 0) e = __atomic_load_n(&st->code, __ATOMIC_RELAXED);
            if (__shfl(e, 0)) break;
        }
    }
    if (key >= 0) {
        flush_counts(SS, cn, lane);
        bins_flush(B);
        flush_lds_bins(B, lane);
    }
}

// Material-source pre-pass of a cycle's first launch: for every history of every item of a
// material-source slot, the rejection loop of materialSource%sampleParticle on the history's
// source stream (as born() would run it), keeping only how many attempts failed.  Waves claim
// chunks of SRC_CHUNK consecutive histories from a counter (chunks of point-source items take no
// time, so a fixed deal would leave waves idle) and work through a chunk one item at a time, so
// the slot (and its tables) is wave-uniform; each lane takes a contiguous run of the
// item's histories and makes one attempt per iteration, moving to its next history (its stream
// one stride further: one LCG step) as soon as one is accepted.
constexpr int SRC_CHUNK = 2048;
template <int NL, int GEO>
__global__ void __launch_bounds__(256)
k_bsource(const SlotDev* __restrict__ slots, const SlotState* __restrict__ ss, const Item* __restrict__ items, int* ctl,
          DevStatus* st)
{
    const int n_items = ctl[0];
    if (n_items <= 0) return;
    const int lane = threadIdx.x & (BWAVE - 1);
    const long long total = items[n_items - 1].prefix + items[n_items - 1].n;
    const long long n_chunks = (total + SRC_CHUNK - 1) / SRC_CHUNK;
    const uint64_t A1 = RNG_STRIDES.a[0][1], C1 = RNG_STRIDES.c[0][1];   // one stride of the LCG
    Pt<NL, GEO> p;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        p.loc[l] = 0;
        p.uni[l] = 0;
        p.root[l] = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) p.r[l][q] = 0.0;
    }
    for (;;) {
        int claim = 0;
        if (lane == 0) claim = atomicAdd(&ctl[6], 1);
        const long long c = __builtin_amdgcn_readfirstlane(__shfl(claim, 0));
        if (c >= n_chunks) break;
        long long a = c * SRC_CHUNK;
        const long long e = min(a + SRC_CHUNK, total);
        while (a < e) {   // one item of the chunk at a time (wave-uniform)
            int lo = 0, hi = n_items;   // items[lo].prefix <= a < items[lo].prefix + items[lo].n
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (items[mid].prefix <= a) lo = mid;
                else hi = mid;
            }
            lo = __builtin_amdgcn_readfirstlane(lo);
            const long long prefix = items[lo].prefix;
            const long long seg_end = min(e, prefix + items[lo].n);
            const int key = __builtin_amdgcn_readfirstlane(items[lo].key);
            const int s = key >> WINDOW_SHIFT, w = key & (MAX_WINDOW - 1);
            const SlotDev& D = slots[s];
            if (D.src_type != SCN_SRC_POINT) {
                const long long len = seg_end - a, per = (len + BWAVE - 1) / BWAVE;
                long long g = a + lane * per;
                const long long g_stop = min(g + per, seg_end);
                int i = (int)(g - prefix), k = 0, active = g < g_stop;
                uint64_t start = 0, srng = 0;
                if (active) {
                    const int64_t b2 = 2 * (int64_t)D.pop * (ss[s].batch + w);
                    start = rng_stride(D.seed, b2 + (int64_t)i + 1);
                    srng = start;
                }
                while (active) {
                    double r[3], time;
                    const int mat = src_attempt(D, D.P, srng, p, r, time, st);
                    int done = 0;
                    if (mat == SCN_OUTSIDE_MAT) {
                    } else if (mat < 1 || mat > D.P.n_mat) {
                        raise_error(st, SCN_ERR_MATERIAL, mat);
                        k = MATERIAL_SOURCE_ATTEMPTS;
                        done = 1;
                    } else if (mat == D.src_mat) {
                        done = 1;
                    }
                    if (!done && ++k >= MATERIAL_SOURCE_ATTEMPTS) {
                        raise_error(st, SCN_ERR_SAMPLING, 4);
                        done = 1;
                    }
                    if (done) {
                        D.src_att[(size_t)w * D.pop + i] = (uint8_t)k;
                        if (++g < g_stop) {   // the next history: its stream is one stride on
                            ++i;
                            start = (A1 * start + C1) & SCN_LCG_MASK;
                            srng = start;
                            k = 0;
                        } else {
                            active = 0;
                        }
                    }
                }
            }
            a = seg_end;
        }
    }
}

// the pre-pass over a grid of every resident block of its kernel on g.x CUs
template <int NL, int GEO>
void launch_bsource(dim3 g, dim3 b, hipStream_t s, const SlotDev* slots, const SlotState* ss, const Item* items,
                    int* ctl, DevStatus* st)
{
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_bsource<NL, GEO>, (int)b.x, 0) != hipSuccess)
        occ = 2;
    hipLaunchKernelGGL((k_bsource<NL, GEO>), dim3(g.x * (unsigned)std::max(1, occ)), b, 0, s, slots, ss, items, ctl,
                       st);
}

template <int NL, int MODE, int GEO>
void launch_bt(dim3 g, dim3 b, size_t sh, hipStream_t s, const SlotDev* slots, SlotState* ss, const Item* items,
               int* item_cnt, int* ctl, int first, int in_side, int out_side, int chunk, int max_steps, DevStatus* st,
               int lds_per_wave, Sites SB, int refill)
{
    hipLaunchKernelGGL((k_btransport<NL, MODE, GEO>), g, b, sh, s, slots, ss, items, item_cnt, ctl, first, in_side,
                       out_side, chunk, max_steps, st, lds_per_wave, SB, refill);
}

template <int NL, int MODE, int GEO>
const void* bt_kernel()
{
    return (const void*)k_btransport<NL, MODE, GEO>;
}

template <int NL, int MODE, int GEO>
void bt_variant(BTransportFn* fn, const void** kern, BSourceFn* src)
{
    *fn = launch_bt<NL, MODE, GEO>;
    *kern = bt_kernel<NL, MODE, GEO>();
    if (src) *src = launch_bsource<NL, GEO>;
}
#endif  // SCN_BTRANSPORT_KERNELS

}  // namespace scn_bat

#endif  // SCN_BTRANSPORT_H
