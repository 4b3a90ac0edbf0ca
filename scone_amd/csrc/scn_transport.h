// scn_transport.h -- the eigenvalue engine's transport kernel (k_transport) and the types its
// host code shares with it.
//
// The kernel templates are compiled only where SCN_TRANSPORT_KERNELS is defined: one translation
// unit per (tracking mode, geometry variant) instantiates pick_variant<MODE, GEO>, so the variants
// compile in parallel and the host code of scn_engine.hip never parses the kernel.
#ifndef SCN_TRANSPORT_H
#define SCN_TRANSPORT_H

#include <hip/hip_runtime.h>

#include "scone_gpu.h"
#include "scn_fpmath.h"
#include "scn_device.h"

namespace scn_eng {

using namespace scn;

constexpr int BLOCK = 256;
// without an LDS copy of the score memory, the dynamic LDS holds each lane's queue of EQUEUE
// evicted global scores (Bins::q)
constexpr int EQUEUE = 4;
constexpr size_t EQUEUE_BYTES = sizeof(double) * 2 * EQUEUE * BLOCK;
constexpr int WAVE = 64;
constexpr int LDS_BINS_MAX = 8192;   // collision-clerk memories up to 64 KiB are privatised in LDS
constexpr int DEFAULT_EIGEN_REFILL = 1;    // idle lanes a wave gathers before it starts new histories

// live particle queue, structure of arrays; level-major position arrays
struct Live {
    double* r;     // [NL*3][cap]
    double* u;     // [3][cap]
    double* w;
    double* w0;
    double* t;
    double* k;     // [4][cap]
    uint64_t* rng;
    int* loc;      // [NL][cap]
    int* nest;
    int* mat;
    int* uid;
    int* G;
    int* brood;
    int* nsite;
    int* coll;
};

struct Counters {
    unsigned long long segments;
    unsigned long long collisions;
    unsigned long long scores;
    unsigned long long pad;
};

// master RNG state and k of a cycle read from device memory (cycles replayed from a graph)
struct CycleIn {
    const uint64_t* master;   // null: use the by-value master and k_eff
    const double* k_eff;
    int* hist_nsite;          // [hist_cap]: fission sites each history emitted (null: not recorded)
    int refill;               // idle lanes a wave gathers before it starts new histories (1: at once)
};

typedef void (*TransportFn)(dim3, dim3, size_t, hipStream_t, const DevProblem&, const scn_particle_state*, int, int,
                            Live, const int*, Live, int*, int, int, uint64_t, double, Sites, double*, int, double*,
                            int, Counters*, DevStatus*, int*, int, CycleIn);

// launcher and kernel address of one variant: register depth NL, tracking mode MODE, geometry
// variant GEO (0: lattice/pin/box, 1: cells, rotations, CSG borders)
template <int NL, int MODE, int GEO>
TransportFn variant_fn(const void** kern);

#ifdef SCN_TRANSPORT_KERNELS
template <int NL, int GEO>
__device__ inline void load_live(const Live& L, int cap, int i, Pt<NL, GEO>& p)
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
#pragma unroll
    for (int k = 0; k < 4; ++k) p.k[k] = L.k[(size_t)k * cap + i];
    p.rng = L.rng[i];
    p.mat = L.mat[i];
    p.uid = L.uid[i];
    p.G = L.G[i];
    p.brood = L.brood[i];
    p.nsite = L.nsite[i];
    p.coll = L.coll[i];
    p.split = 0;
    p.dead = 0;
    p.fate = FATE_NO;
    p.track = 0.0;
}

template <int NL, int GEO>
__device__ inline void store_live(const Live& L, int cap, int i, const Pt<NL, GEO>& p)
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
#pragma unroll
    for (int k = 0; k < 4; ++k) L.k[(size_t)k * cap + i] = p.k[k];
    L.rng[i] = p.rng;
    L.mat[i] = p.mat;
    L.uid[i] = p.uid;
    L.G[i] = p.G;
    L.brood[i] = p.brood;
    L.nsite[i] = p.nsite;
    L.coll[i] = p.coll;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One launch of the event loop (persistent grid).  first != 0: the input is the cycle's bank
// (AoS) and history n = i + 1 starts on the master strided by n.  Waves take chunks of the input
// from an atomic counter; every lane holds one particle and advances it by one tracking event per
// iteration (plus the collision it may lead to), and a lane whose particle dies -- or, with a
// step limit, has run max_steps (flight, collision) pairs and is parked in the output queue --
// takes the next particle of the chunk at once.  Lanes therefore stay busy until the input is
// used up, whatever the spread of history lengths.
template <int NL, int MODE, int GEO>
__global__ void __launch_bounds__(BLOCK)
k_transport(DevProblem P, const scn_particle_state* __restrict__ bank, int first, int n_first, Live in,
            const int* __restrict__ n_in_ptr, Live out, int* __restrict__ n_out, int cap, int max_steps,
            uint64_t master, double k_eff, Sites S, double* __restrict__ bins, int use_lds,
            double* __restrict__ hist_k, int hist_cap, Counters* __restrict__ cnt, DevStatus* st, int* work,
            int chunk, CycleIn cin)
{
    extern __shared__ double lds_bins[];
    const int n_in = first ? n_first : *n_in_ptr;
    if (cin.master) {
        master = *cin.master;
        k_eff = *cin.k_eff;
    }
    // LDS copy of the whole score memory: doubles, or four limb words per bin (fixed point)
    const int lds_words = use_lds ? (int)P.n_bins * (P.fix ? 4 : 1) : 0;
    for (int i = threadIdx.x; i < lds_words; i += BLOCK) lds_bins[i] = 0.0;
    __syncthreads();
    Bins B;
    B.lds = use_lds ? lds_bins : nullptr;
    B.glob = bins;
    B.lds_n = use_lds ? P.n_bins : 0;
    B.c_addr = -1;
    B.c_acc = 0.0;
    B.q = use_lds ? nullptr : lds_bins + (size_t)(threadIdx.x / WAVE) * 2 * EQUEUE * WAVE + (threadIdx.x & (WAVE - 1));
    B.qn = 0;
    B.qcap = use_lds ? 0 : EQUEUE;
    B.fix = P.fix;
    Cnt cn{0, 0, 0};
    const int lane = threadIdx.x & (WAVE - 1);
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    int next = 0, end = 0, exhausted = 0;
    Pt<NL, GEO> p;
    DistCache<NL> cache;
    cache.lvl = 0;
    int alive = 0, steps = 0, nseg = 0;
    int use_dt = MODE == SCN_TRANSPORT_DT;   // this flight's tracking method (HT: chosen per flight)
    int fresh = 1;                           // the next event starts a flight
    for (int iter = 0;; ++iter) {
        if (next >= end && !exhausted) {
            int c = 0;
            if (lane == 0) c = atomicAdd(work, 1);
            c = __shfl(c, 0);
            const long long base = (long long)c * chunk;
            if (base >= n_in) {
                exhausted = 1;
            } else {
                next = (int)base;
                end = (int)min((long long)n_in, base + chunk);
            }
        }
        const unsigned long long live = __ballot(alive);
        // histories start once cin.refill lanes are idle (or none is alive), many lanes at a time
        if (next < end && (__popcll(~live) >= cin.refill || live == 0)) {
            const unsigned long long idle = ~live;
            const int i = next + __popcll(idle & lt_mask);
            if (!alive && i < end) {
                if (first) {
                    const scn_particle_state& s = bank[i];
                    p.rng = rng_stride(master, (int64_t)i + 1);
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
                        p.r[0][k] = s.r[k];
                        p.u[k] = s.dir[k];
                    }
                    p.w = s.wgt;
                    p.w0 = s.wgt;
                    p.G = s.G;
                    p.t = s.time;
                    p.coll = s.collisionN;
                    p.brood = i + 1;
                    p.nsite = 0;
                    p.k[0] = p.k[1] = p.k[2] = p.k[3] = 0.0;
                    p.dead = 0;
                    p.fate = FATE_NO;
                    p.split = 0;
                    p.track = 0.0;
                    place_coord(P, p, st);
                } else {
                    load_live(in, cap, i, p);
                    relink(P, p);
                }
                alive = 1;
                steps = 0;
                nseg = 0;
                cache.lvl = 0;
                fresh = 1;
            }
            next = min(next + __popcll(idle), end);
        }
        if (__ballot(alive) == 0) {
            if (exhausted) break;
            continue;
        }
        int park = 0;
        if (alive) {
            if (MODE == SCN_TRANSPORT_HT && fresh) use_dt = ht_use_dt(P, p, st);
            fresh = 0;
            int r = use_dt ? dt_step(P, p, B, st, cn) : st_step(P, p, cache, B, st, cn);
            if (r == STEP_MORE && ++nseg >= MAX_SEGMENTS_PER_FLIGHT) {
                raise_error(st, SCN_ERR_ARG, SCN_ERR_LOST_DETAIL);
                r = STEP_DEAD;
            }
            if (r == STEP_COLLIDE) {
                if (collide(P, p, k_eff, B, S, st, cn)) {
                    cache.lvl = 0;
                    nseg = 0;
                    fresh = 1;
                    if (++steps >= max_steps) {
                        park = 1;
                        alive = 0;
                    }
                } else {
                    r = STEP_DEAD;
                }
            }
            if (r == STEP_DEAD) {
                alive = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) hist_k[(size_t)k * hist_cap + (p.brood - 1)] = p.k[k];
                if (cin.hist_nsite) cin.hist_nsite[p.brood - 1] = p.nsite;
            }
        }
        if (__ballot(B.qn >= B.qcap && B.qcap > 0)) bins_drain(B);   // a full queue: the wave adds all of them
        // park step-limited survivors: wave ballot, one atomic per wave
        const unsigned long long pm = __ballot(park);
        if (pm) {
            int obase = 0;
            if (lane == 0) obase = atomicAdd(n_out, __popcll(pm));
            obase = __shfl(obase, 0);
            if (park) {
                const int o = obase + __popcll(pm & lt_mask);
                if (o < cap) store_live(out, cap, o, p);
                else raise_error(st, SCN_ERR_OVERFLOW, o);
            }
        }
        if ((iter & 63) == 63) {
            int e = 0;
            if (lane == 0) e = __atomic_load_n(&st->code, __ATOMIC_RELAXED);
            if (__shfl(e, 0)) break;
        }
    }
    bins_flush(B);
    const unsigned long long seg = wave_sum64((unsigned long long)cn.seg);
    const unsigned long long col = wave_sum64((unsigned long long)cn.col);
    const unsigned long long sc = wave_sum64((unsigned long long)cn.sc);
    if (lane == 0) {
        if (seg) atomicAdd(&cnt->segments, seg);
        if (col) atomicAdd(&cnt->collisions, col);
        if (sc) atomicAdd(&cnt->scores, sc);
    }
    __syncthreads();
    if (use_lds && P.fix) {
        const unsigned long long* lw = (const unsigned long long*)lds_bins;
        for (int i = threadIdx.x; i < lds_words; i += BLOCK) {
            const unsigned long long v = lw[i];
            if (v) atomicAdd((unsigned long long*)bins + i, v);
        }
    } else if (use_lds) {
        for (int i = threadIdx.x; i < (int)P.n_bins; i += BLOCK) {
            double v = lds_bins[i];
            if (v != 0.0) unsafeAtomicAdd(&bins[i], v);
        }
    }
}

template <int NL, int MODE, int GEO>
void launch_transport(dim3 g, dim3 b, size_t sh, hipStream_t s, const DevProblem& P, const scn_particle_state* bank,
                      int first, int n_first, Live in, const int* n_in_ptr, Live out, int* n_out, int cap,
                      int max_steps, uint64_t master, double k_eff, Sites S, double* bins, int use_lds,
                      double* hist_k, int hist_cap, Counters* cnt, DevStatus* st, int* work, int chunk, CycleIn cin)
{
    hipLaunchKernelGGL((k_transport<NL, MODE, GEO>), g, b, sh, s, P, bank, first, n_first, in, n_in_ptr, out, n_out, cap,
                       max_steps, master, k_eff, S, bins, use_lds, hist_k, hist_cap, cnt, st, work, chunk, cin);
}

template <int NL, int MODE, int GEO>
const void* transport_kernel()
{
    return (const void*)k_transport<NL, MODE, GEO>;
}

template <int NL, int MODE, int GEO>
TransportFn variant_fn(const void** kern)
{
    *kern = transport_kernel<NL, MODE, GEO>();
    return launch_transport<NL, MODE, GEO>;
}
#endif  // SCN_TRANSPORT_KERNELS

}  // namespace scn_eng

#endif  // SCN_TRANSPORT_H
