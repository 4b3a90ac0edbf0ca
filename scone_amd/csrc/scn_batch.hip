// scn_batch.hip -- fixed-source engine with resident problem slots (batched and served solves)
// behind the scn_batch_* entry points of include/scone_gpu.h.
//
// Every admitted solve owns a slot: its tables, two live-particle queues (SoA), its score memory
// (BIN, CSUM, CSUM2) and optional per-history records, all in one block of a device arena that
// is sized by the HBM budget.  A cycle advances every running slot by `window` batches of
// fixedSourcePhysicsPackage%cycles (PhysicsPackages/fixedSourcePhysicsPackage_class.f90:187-275):
//   k_items          one work item per batch of every running slot's window (slot, batch, count)
//   k_btransport     persistent kernel: waves start spread over the items in proportion to their
//                    particles and pull chunks from an item's atomic counter until it is used up;
//                    each lane runs one particle in registers and takes the next particle of the
//                    chunk as soon as its own dies, so lanes stay busy; with a step limit K,
//                    survivors are parked in the slot's other queue (wave ballot, one atomic per
//                    wave).  The first launch samples the batch's source particles in place.
//   (k_items, k_btransport) x (n_launch - 1), K doubling, the last launch runs to completion
//   k_bnorm/k_bclose tallyAdmin%reportCycleEnd normalisation and scoreMemory%closeCycle
//   k_bend           batchN + closed batches, early stop on a bin's relative sd, slot completion
// Nothing in a cycle depends on host values, so the sequence can be replayed from a hipGraph.
#include <cstring>
#include <cstdio>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "scone_gpu.h"
#include "scn_fpmath.h"
#include "scn_device.h"
#include "scn_host.h"
#include "scn_btransport.h"

using namespace scn;

namespace {

using namespace scn_bat;

// item table of one launch.  One block of 1024 threads scans the slots: the first launch of a
// cycle covers batches batch .. batch + n_run - 1 of every running slot (pop histories each), later
// launches the particles parked in each slot's queue.  Each item's pull counter starts at zero.
__global__ void __launch_bounds__(1024)
k_items(const SlotDev* __restrict__ slots, SlotState* ss, int n_slots, int first, int in_side, int out_side,
        Item* __restrict__ items, int* __restrict__ item_cnt, int* ctl)
{
    __shared__ int s_scan[1024];
    __shared__ long long s_pscan[1024];
    int carry = 0;
    long long pcarry = 0;
    for (int base = 0; base < n_slots; base += 1024) {
        const int s = base + (int)threadIdx.x;
        int n = 0, nb = 0;
        if (s < n_slots) {
            SlotState& S = ss[s];
            const SlotDev& D = slots[s];
            const int pop = D.pop;
            if (first) {
                const int w = S.win > 0 ? min(S.win, D.window) : D.window;
                nb = S.running ? min(w, D.n_batches - S.batch) : 0;
                n = nb > 0 ? pop : 0;
                S.n_run = nb;
            } else {
                n = S.count[in_side] < pop ? S.count[in_side] : pop;   // a queue never holds more than pop
                nb = n > 0 ? 1 : 0;
            }
            S.n_in = n;
            S.count[out_side] = 0;
        }
        s_scan[threadIdx.x] = nb;
        s_pscan[threadIdx.x] = (long long)nb * n;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int v = (int)threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
            long long pv = (int)threadIdx.x >= off ? s_pscan[threadIdx.x - off] : 0;
            __syncthreads();
            s_scan[threadIdx.x] += v;
            s_pscan[threadIdx.x] += pv;
            __syncthreads();
        }
        if (s < n_slots && nb > 0) {
            const int i0 = carry + s_scan[threadIdx.x] - nb;                  // this slot's first item
            const long long p0 = pcarry + s_pscan[threadIdx.x] - (long long)nb * n;
            for (int w = 0; w < nb; ++w) {
                items[i0 + w] = Item{s * MAX_WINDOW + w, n, p0 + (long long)w * n};
                item_cnt[i0 + w] = 0;
            }
        }
        carry += s_scan[1023];
        pcarry += s_pscan[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctl[0] = carry;      // items (at most n_slots * MAX_WINDOW: the table's size)
        ctl[5] = 0;          // items whose histories are all handed out
        ctl[6] = 0;          // chunks of the material-source pre-pass handed out
        if (first) ctl[2] = 0;
    }
}

// End of a cycle for each running slot (one thread per slot), over the n_run batches its window
// transported, in batch order: tallyAdmin_class.f90:768-784 normFactor = normValue / BIN(normBin)
// (or 1) of each batch, then the early-stop test on the stop bin's relative sd after each batch
// (scoreMemory%getResult, scoreMemory_class.f90:537-573) evaluated on the CSUM / CSUM2 values that
// closing the batches one by one gives, so the solve stops after the same batch as with one
// batch per cycle; batches after the stop are not closed.
// value of bin i of BIN copy w: a double, or the four limb words of a fixed-point score memory
__device__ __forceinline__ double bin_value(const SlotDev& D, int w, int64_t i)
{
    const int64_t nb = D.P.n_bins;
    if (!D.P.fix) return D.bin[(size_t)w * (size_t)nb + i];
    const unsigned long long* b = (const unsigned long long*)D.bin + 4 * ((size_t)w * (size_t)nb + i);
    const uint64_t wv[4] = {b[0], b[1], b[2], b[3]};
    return scn_fix_value(wv);
}

__global__ void k_bnorm(const SlotDev* __restrict__ slots, SlotState* ss, int n_slots, DevStatus* st)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    SlotState& S = ss[s];
    if (!S.running) return;
    const SlotDev& D = slots[s];
    const int n_run = S.n_run;
    int n_close = n_run;
    double cs = 0.0, cs2 = 0.0;
    if (D.stop_bin >= 0) {
        cs = D.csum[D.stop_bin];
        cs2 = D.csum2[D.stop_bin];
    }
    for (int w = 0; w < n_run; ++w) {
        double f = 1.0;
        if (D.norm_bin >= 0) {
            double v = bin_value(D, w, D.norm_bin);
            if (v == 0.0) {
                raise_error(st, SCN_ERR_NORM, s);
                f = 0.0;
            } else {
                f = D.norm_value / v;
            }
        }
        S.norm[w] = f;
        const int N = S.batch + w + 1;
        if (D.stop_bin >= 0 && N < D.n_batches && N >= 2) {
            const double r = bin_value(D, w, D.stop_bin) * f;
            cs = cs + r;
            cs2 = cs2 + r * r;
            double mean = cs / N;
            double inv_N = 1.0 / N;
            double inv_Nm1 = 1.0 / (N - 1);
            double var = cs2 * inv_N * inv_Nm1 - mean * mean * inv_Nm1;
            double sd = sqrt(var);
            if (mean != 0.0 && sd / fabs(mean) <= D.stop_rel) {
                n_close = w + 1;
                S.stopped = 1;
                break;
            }
        } else if (D.stop_bin >= 0) {
            const double r = bin_value(D, w, D.stop_bin) * f;
            cs = cs + r;
            cs2 = cs2 + r * r;
        }
    }
    S.n_close = n_close;
    // adaptive window: the relative sd of the mean falls like 1/sqrt(N), so the target is projected
    // at N* = N (rel / stop_rel)^2 batches; the next cycle transports adapt% of the batches still
    // missing to N* (at least one): the solve approaches the projection from below, in a few
    // cycles, and few batches are transported past the stop.  Which batches are transported in
    // which cycle changes nothing in the results.
    if (D.adapt > 0 && D.stop_bin >= 0 && !S.stopped) {
        const int N = S.batch + n_close;
        int w = D.window;
        if (N < 2) {
            w = 2 - N;
        } else if (D.stop_rel > 0.0) {
            const double mean = cs / N;
            const double var = cs2 * (1.0 / N) * (1.0 / (N - 1)) - mean * mean * (1.0 / (N - 1));
            const double sd = sqrt(var);
            if (mean != 0.0 && sd == sd) {
                const double q = sd / fabs(mean) / D.stop_rel;
                const double need = ceil(0.01 * D.adapt * ((double)N * q * q - N));
                w = need < 1.0 ? 1 : (need > (double)D.window ? D.window : (int)need);
            }
        }
        S.win = w;
    }
}

// scoreMemory_class.f90:309-342 closeCycle of each closed batch of every running slot, in batch
// order (grid.y = slot); the BIN copies of every transported batch are cleared
__global__ void k_bclose(const SlotDev* __restrict__ slots, const SlotState* __restrict__ ss)
{
    const int s = blockIdx.y;
    const SlotState& S = ss[s];
    if (!S.running) return;
    const SlotDev& D = slots[s];
    const int64_t n = D.P.n_bins;
    const int n_run = S.n_run, n_close = S.n_close;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double c = D.csum[i], c2 = D.csum2[i];
        for (int w = 0; w < n_run; ++w) {
            if (w < n_close) {
                double r = bin_value(D, w, i) * S.norm[w];
                c = c + r;
                c2 = c2 + r * r;
            }
            if (D.P.fix) {
                unsigned long long* b = (unsigned long long*)D.bin + 4 * ((size_t)w * (size_t)n + i);
                b[0] = b[1] = b[2] = b[3] = 0ull;
            } else {
                D.bin[(size_t)w * (size_t)n + i] = 0.0;
            }
        }
        D.csum[i] = c;
        D.csum2[i] = c2;
    }
}

// end of the cycle: batchN + n_close, completion (all batches, or the early stop found above)
__global__ void k_bend(const SlotDev* __restrict__ slots, SlotState* ss, int n_slots, int* ctl)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    SlotState& S = ss[s];
    if (!S.running) return;
    const SlotDev& D = slots[s];
    const int N = S.batch + S.n_close;
    S.batch = N;
    S.histories += (unsigned long long)D.pop * (unsigned long long)S.n_close;
    if (N >= D.n_batches || S.stopped) S.running = 0;
    else atomicAdd(&ctl[2], 1);
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
// kernel variant key: register depth NL, plus GEO_KEY for the general-geometry kernels
constexpr int GEO_KEY = 16;

template <int MODE>
void pick_bt(int key, BTransportFn* fn, const void** kern, BSourceFn* src = nullptr)
{
    switch (key) {
    case 1: bt_variant<1, MODE, 0>(fn, kern, src); break;
    case 3: bt_variant<3, MODE, 0>(fn, kern, src); break;
    case 4: bt_variant<4, MODE, 0>(fn, kern, src); break;
    case 3 + GEO_KEY: bt_variant<3, MODE, 1>(fn, kern, src); break;
    case 8 + GEO_KEY: bt_variant<8, MODE, 1>(fn, kern, src); break;
    default: bt_variant<8, MODE, 0>(fn, kern, src); break;
    }
}

struct SlotHost {
    int used = 0;
    size_t off = 0, bytes = 0;
    int depth = 0;
    int64_t n_bins = 0;
    int pop = 0, n_batches = 0, record = 0;
    int fissile = 0, buffer = 0;
    int geo = 0;        // needs the general-geometry kernels
    int ww = 0;         // weight windows: split copies use the history's buffer
    int mat_src = 0;    // material source: the cycle's first launch runs the sampling pre-pass
    int window_max = 1; // window the solve was admitted with (BIN copies allocated)
    int done_known = 0; // a query saw the solve done: its counters (done_st) no longer change
    SlotState done_st{};
};

struct Batch {
    std::string err;
    scn_batch_settings S{};
    hipStream_t stream = nullptr;
    char* arena = nullptr;
    size_t arena_size = 0;
    std::map<size_t, size_t> freel;      // offset -> size of free blocks
    int n_slots = 0;
    SlotDev* d_slots = nullptr;
    SlotState* d_state = nullptr;
    Item* d_items = nullptr;         // work items of a launch (n_slots * MAX_WINDOW at most)
    int* d_item_cnt = nullptr;
    int* d_ctl = nullptr;
    DevStatus* d_status = nullptr;
    std::vector<SlotDev> h_slots;
    std::vector<SlotHost> hs;
    int n_cu = 256;
    // kernel variant of the recorded cycle
    int nl_cur = 0;
    int ms_cur = 0;     // the captured cycle runs the material-source pre-pass
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<hipEvent_t> ev;
    hipEvent_t c0 = nullptr, c1 = nullptr;
    scn_stats stats{};
    int64_t rel_seg = 0, rel_col = 0, rel_sc = 0, rel_hist = 0;   // counters of released slots
    // pinned host staging of the small readbacks (slot states, score sums): a pageable copy
    // goes through the runtime's own staging and costs several times more per call
    char* pin = nullptr;
    size_t pin_bytes = 0;
    int S_lds_per_wave = 2048;   // low bins of a slot privatised per wave in LDS (16 KiB per wave)
    // per-lane LIFO buffers of fixed-source secondaries: a block of the arena (so inside the HBM
    // budget), reserved when a solve with secondaries is admitted, for max_lanes lanes of the
    // deepest such solve's buffer, and returned when none is resident
    Sites stack{};
    size_t stack_off = SIZE_MAX, stack_bytes = 0;
    int stack_depth = 0;
    int max_lanes = 0;   // lanes of the largest persistent grid of any transport kernel variant
};

#define BHIPCHK(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            B->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
            return SCN_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// first-fit allocation in the arena; returns SIZE_MAX if nothing fits
size_t arena_alloc(Batch* B, size_t bytes)
{
    bytes = align256(std::max<size_t>(bytes, 256));
    for (auto it = B->freel.begin(); it != B->freel.end(); ++it) {
        if (it->second >= bytes) {
            size_t off = it->first, sz = it->second;
            B->freel.erase(it);
            if (sz > bytes) B->freel[off + bytes] = sz - bytes;
            return off;
        }
    }
    return SIZE_MAX;
}

void arena_free(Batch* B, size_t off, size_t bytes)
{
    bytes = align256(std::max<size_t>(bytes, 256));
    auto it = B->freel.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != B->freel.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        B->freel.erase(nx);
    }
    if (it != B->freel.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            B->freel.erase(it);
        }
    }
}

// byte layout of one slot block
struct SlotLayout {
    size_t tables = 0;          // bytes of the table region (copied from the host image)
    size_t queue = 0;           // offset of queue 0; queue 1 follows
    size_t queue_bytes = 0;
    size_t bins = 0;            // BIN, CSUM, CSUM2
    size_t hist = 0;
    size_t src = 0;             // material-source attempt counts
    size_t total = 0;
};

size_t live_bytes(int depth, int cap)
{
    size_t c = (size_t)cap;
    size_t b = 0;
    b += align256(sizeof(double) * depth * 3 * c);
    b += align256(sizeof(double) * 3 * c);
    b += 3 * align256(sizeof(double) * c);
    b += align256(sizeof(uint64_t) * c);
    b += align256(sizeof(int) * depth * c);
    b += 7 * align256(sizeof(int) * c);
    return b;
}

void carve_live(char* base, int depth, int cap, BLive& L)
{
    size_t c = (size_t)cap;
    char* p = base;
    auto take = [&](size_t bytes) {
        char* q = p;
        p += align256(bytes);
        return q;
    };
    L.r = (double*)take(sizeof(double) * depth * 3 * c);
    L.u = (double*)take(sizeof(double) * 3 * c);
    L.w = (double*)take(sizeof(double) * c);
    L.w0 = (double*)take(sizeof(double) * c);
    L.t = (double*)take(sizeof(double) * c);
    L.rng = (uint64_t*)take(sizeof(uint64_t) * c);
    L.loc = (int*)take(sizeof(int) * depth * c);
    L.nest = (int*)take(sizeof(int) * c);
    L.mat = (int*)take(sizeof(int) * c);
    L.uid = (int*)take(sizeof(int) * c);
    L.G = (int*)take(sizeof(int) * c);
    L.brood = (int*)take(sizeof(int) * c);
    L.coll = (int*)take(sizeof(int) * c);
    L.split = (int*)take(sizeof(int) * c);
}

// host image of a slot's tables: appends arrays, returns their offsets
struct Image {
    std::vector<char> bytes;
    template <typename T>
    size_t put(const T* src, size_t count)
    {
        size_t off = align256(bytes.size());
        size_t n = sizeof(T) * std::max<size_t>(count, 1);
        bytes.resize(off + n, 0);
        if (src && count) memcpy(bytes.data() + off, src, sizeof(T) * count);
        return off;
    }
};

int validate(Batch* B, const scn_fixed_request* R, int* depth)
{
    const scn_problem* pr = R->problem;
    const scn_source* sr = R->source;
    if (!pr || !sr) {
        B->err = "request without problem or source";
        return SCN_ERR_ARG;
    }
    if (pr->mg.n_mat <= 0 || pr->mg.n_groups <= 0 || pr->n_universes <= 0) {
        B->err = "invalid problem tables";
        return SCN_ERR_ARG;
    }
    if (pr->transport != B->S.transport) {
        B->err = "problem transport operator differs from the batch engine's";
        return SCN_ERR_ARG;
    }
    if (R->buffer < 0 || R->buffer > MAX_BUFFER) {
        B->err = "buffer size outside 0..MAX_BUFFER";
        return SCN_ERR_ARG;
    }
    if (R->window < 0 || R->window > MAX_WINDOW) {
        B->err = "window outside 0..16";
        return SCN_ERR_ARG;
    }
    if (R->adapt < 0 || R->adapt > 100) {
        B->err = "adapt outside 0..100";
        return SCN_ERR_ARG;
    }
    if (R->window > 1 && B->S.n_launch > 1) {
        B->err = "a window of several batches needs an engine with n_launch 1";
        return SCN_ERR_ARG;
    }
    if (const char* m = check_cell_tables(pr)) {
        B->err = m;
        return SCN_ERR_ARG;
    }
    int d = problem_depth(pr);
    if (d < 1 || d > SCN_MAX_NEST) {
        B->err = "geometry nesting deeper than SCN_MAX_NEST or malformed graph";
        return SCN_ERR_UNSUPPORTED;
    }
    *depth = d;
    if (R->pop < 1 || R->n_batches < 1) {
        B->err = "pop and n_batches must be >= 1";
        return SCN_ERR_ARG;
    }
    if ((int64_t)R->pop * 2 * (int64_t)R->n_batches > ((int64_t)1 << 40)) {
        B->err = "pop * n_batches too large";
        return SCN_ERR_ARG;
    }
    if (R->stop_bin >= pr->n_bins) {
        B->err = "stop_bin outside the score memory";
        return SCN_ERR_ARG;
    }
    const int nG = pr->mg.n_groups;
    if (sr->type == SCN_SRC_POINT) {
        if (sr->n_probG > 0) {
            if (sr->n_probG != nG || !sr->probG || sr->n_probG > MAX_PROBG) {
                B->err = "Source energy group distribution must have nG groups";
                return SCN_ERR_ARG;
            }
        } else if (sr->G < 1 || sr->G > nG) {
            B->err = "source group outside 1..nG";
            return SCN_ERR_ARG;
        }
    } else if (sr->type == SCN_SRC_MATERIAL) {
        if (sr->mat < 1 || sr->mat > pr->mg.n_mat) {
            B->err = "Source material was not found in the material definitions";
            return SCN_ERR_ARG;
        }
        if (sr->G < 1 || sr->G > nG) {
            B->err = "source group outside 1..nG";
            return SCN_ERR_ARG;
        }
        if (sr->t_high < sr->t_low) {
            B->err = "tHigh is less than tLow";
            return SCN_ERR_ARG;
        }
    } else {
        B->err = "unknown source type";
        return SCN_ERR_ARG;
    }
    return 0;
}

int check_bstatus(Batch* B)
{
    DevStatus h{};
    BHIPCHK(hipMemcpyAsync(&h, B->d_status, sizeof h, hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    if (h.code) {
        char buf[256];
        const char* what = "engine error";
        switch (h.code) {
        case SCN_ERR_OVERFLOW: what = "Run out of space for particles (buffer of secondaries or slot queue)"; break;
        case SCN_ERR_MATERIAL: what = "Particle is in an undefined or overlapping material"; break;
        case SCN_ERR_NAN: what = "Distance is NaN"; break;
        case SCN_ERR_MAJORANT: what = "Majorant is 0"; break;
        case SCN_ERR_SAMPLING: what = "Sampling failed (source rejection, probG, scatter CDF)"; break;
        case SCN_ERR_UNSUPPORTED: what = "Unsupported reaction in collision"; break;
        case SCN_ERR_ARG: what = "Source outside the geometry or particle lost"; break;
        case SCN_ERR_NORM: what = "Normalisation score is 0"; break;
        }
        snprintf(buf, sizeof buf, "%s [detail %d]", what, h.detail);
        B->err = buf;
        DevStatus z{};
        BHIPCHK(hipMemcpyAsync(B->d_status, &z, sizeof z, hipMemcpyHostToDevice, B->stream));
        BHIPCHK(hipStreamSynchronize(B->stream));
        return h.code;
    }
    return 0;
}

// kernel variant key for the resident slots: the deepest geometry, general kernels if any slot
// needs them
int max_depth_nl(Batch* B)
{
    int d = 1, geo = 0;
    for (const SlotHost& h : B->hs)
        if (h.used) {
            d = std::max(d, h.depth);
            geo |= h.geo;
        }
    return compiled_nl(d, geo) + (geo ? GEO_KEY : 0);
}

// a wave's copy of a slot's low bins, then its lanes' queues of evicted scores
size_t lds_bytes_for(int lds_per_wave) { return sizeof(double) * ((size_t)lds_per_wave + 2 * BQUEUE * BWAVE) * (BBLOCK / BWAVE); }
size_t lds_bytes(const Batch* B) { return lds_bytes_for(B->S_lds_per_wave); }

// persistent grid of the transport kernel variant: resident blocks per CU x CUs
// a resident solve samples a material source (the cycle then runs the sampling pre-pass)
int any_material_source(const Batch* B)
{
    for (const SlotHost& h : B->hs)
        if (h.used && h.mat_src) return 1;
    return 0;
}

int transport_variant(Batch* B, int nl, BTransportFn* fn, int* grid, BSourceFn* src)
{
    const void* kern;
    if (B->S.transport == SCN_TRANSPORT_DT) pick_bt<SCN_TRANSPORT_DT>(nl, fn, &kern, src);
    else if (B->S.transport == SCN_TRANSPORT_HT) pick_bt<SCN_TRANSPORT_HT>(nl, fn, &kern, src);
    else pick_bt<SCN_TRANSPORT_ST>(nl, fn, &kern, src);
    int occ = 0;
    BHIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BBLOCK, lds_bytes(B)));
    *grid = B->n_cu * std::max(1, occ);
    return 0;
}

// enqueue one cycle on B->stream (also the body of the captured graph); timed = per-launch events
int enqueue_cycle(Batch* B, BTransportFn fn, BSourceFn src, int mat_src, int grid, int timed)
{
    const int ns = B->n_slots;
    const int tile = B->S.tile;
    int side = 0;
    for (int L = 0; L < B->S.n_launch; ++L) {
        const int first = L == 0;
        const int in_side = first ? 0 : side;
        const int out_side = first ? 0 : 1 - side;
        hipLaunchKernelGGL(k_items, dim3(1), dim3(1024), 0, B->stream, B->d_slots, B->d_state, ns, first, in_side,
                           out_side, B->d_items, B->d_item_cnt, B->d_ctl);
        if (first && mat_src)
            src(dim3(B->n_cu), dim3(256), B->stream, B->d_slots, B->d_state, B->d_items, B->d_ctl, B->d_status);
        const int last = L == B->S.n_launch - 1;
        const int steps = last ? (1 << 30) : (B->S.k0 << L);
        if (timed) BHIPCHK(hipEventRecord(B->ev[2 * L], B->stream));
        fn(dim3(grid), dim3(BBLOCK), lds_bytes(B), B->stream, B->d_slots, B->d_state, B->d_items, B->d_item_cnt,
           B->d_ctl, first, in_side, out_side, tile, steps, B->d_status, B->S_lds_per_wave, B->stack, B->S.refill);
        if (timed) BHIPCHK(hipEventRecord(B->ev[2 * L + 1], B->stream));
        side = out_side;
    }
    const int sb = (ns + 255) / 256;
    hipLaunchKernelGGL(k_bnorm, dim3(sb), dim3(256), 0, B->stream, B->d_slots, B->d_state, ns, B->d_status);
    hipLaunchKernelGGL(k_bclose, dim3(16, ns), dim3(256), 0, B->stream, B->d_slots, B->d_state);
    hipLaunchKernelGGL(k_bend, dim3(sb), dim3(256), 0, B->stream, B->d_slots, B->d_state, ns, B->d_ctl);
    BHIPCHK(hipGetLastError());
    return 0;
}

void drop_graph(Batch* B)
{
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->gexec = nullptr;
    B->graph = nullptr;
}

constexpr size_t STACK_ENTRY_BYTES = 8 * sizeof(double) + sizeof(int);   // r, u, w, t, G

// arena bytes of per-lane buffers of `depth` entries for `lanes` lanes
size_t stack_bytes_for(int64_t lanes, int depth)
{
    return align256((size_t)lanes * (size_t)depth * STACK_ENTRY_BYTES + 1024);
}

// the largest persistent grid (in lanes) any transport kernel variant of `transport` takes on a
// device of n_cu CUs with `lds` bytes of LDS per block
int64_t max_grid_lanes(int transport, int n_cu, size_t lds)
{
    int64_t lanes = 0;
    for (int key : {1, 3, 4, 8, 3 + GEO_KEY, 8 + GEO_KEY}) {
        BTransportFn f;
        const void* kern;
        if (transport == SCN_TRANSPORT_DT) pick_bt<SCN_TRANSPORT_DT>(key, &f, &kern);
        else if (transport == SCN_TRANSPORT_HT) pick_bt<SCN_TRANSPORT_HT>(key, &f, &kern);
        else pick_bt<SCN_TRANSPORT_ST>(key, &f, &kern);
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BBLOCK, lds) != hipSuccess) occ = 8;
        lanes = std::max(lanes, (int64_t)n_cu * std::max(1, occ) * BBLOCK);
    }
    return lanes;
}

// the per-lane LIFO buffers of this cycle's grid inside the reserved arena block
int ensure_stack(Batch* B, int grid)
{
    if (B->stack_depth == 0 || B->stack_off == SIZE_MAX) {
        B->stack = Sites{};
        return 0;
    }
    const int lanes = grid * BBLOCK;
    if (lanes > B->max_lanes) {
        B->err = "persistent grid larger than the lanes the secondary buffers were sized for";
        return SCN_ERR_OVERFLOW;
    }
    const int depth = B->stack_depth;
    if (B->stack.r && B->stack.stride == lanes && B->stack.depth == depth) return 0;
    drop_graph(B);
    const size_t n = (size_t)lanes * depth;
    char* q = B->arena + B->stack_off;
    Sites S{};
    S.r = (double*)q;
    S.u = S.r + 3 * n;
    S.w = S.u + 3 * n;
    S.t = S.w + n;
    S.G = (int*)(S.t + n);
    S.cap = (int)n;
    S.stride = lanes;
    S.depth = depth;
    B->stack = S;
    return 0;
}

// reserve (grow) or return the secondary buffers for the resident solves plus `extra_depth`
// (a solve being admitted, 0 for none); SCN_ERR_BUSY if the budget cannot hold them
int reserve_stack(Batch* B, int extra_depth)
{
    int depth = extra_depth;
    for (const SlotHost& h : B->hs)
        if (h.used && (h.fissile || h.ww)) depth = std::max(depth, h.buffer);
    if (depth == B->stack_depth) return 0;
    if (depth < B->stack_depth && depth > 0) return 0;   // keep the deeper block while it is in use
    size_t off = SIZE_MAX, bytes = 0;
    if (depth > 0) {   // the new block first: the resident solves keep theirs if it does not fit
        bytes = stack_bytes_for(B->max_lanes, depth);
        off = arena_alloc(B, bytes);
        if (off == SIZE_MAX) return SCN_ERR_BUSY;
    }
    drop_graph(B);
    B->stack = Sites{};
    if (B->stack_off != SIZE_MAX) arena_free(B, B->stack_off, B->stack_bytes);
    B->stack_off = off;
    B->stack_bytes = bytes;
    B->stack_depth = depth;
    return 0;
}

int run_cycle(Batch* B, int* n_running_out)
{
    const int nl = max_depth_nl(B);
    const int msrc = any_material_source(B);
    BTransportFn fn;
    BSourceFn src = nullptr;
    int grid = 0;
    int rc0 = transport_variant(B, nl, &fn, &grid, &src);
    if (rc0) return rc0;
    rc0 = ensure_stack(B, grid);
    if (rc0) return rc0;
    BHIPCHK(hipEventRecord(B->c0, B->stream));
    if (B->S.use_graph) {
        if (!B->gexec || B->nl_cur != nl || B->ms_cur != msrc) {
            drop_graph(B);
            BHIPCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_cycle(B, fn, src, msrc, grid, 0);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(B->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) {
                B->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
                return SCN_ERR_HIP;
            }
            B->graph = g;
            BHIPCHK(hipGraphInstantiate(&B->gexec, B->graph, nullptr, nullptr, 0));
            B->nl_cur = nl;
            B->ms_cur = msrc;
        }
        BHIPCHK(hipGraphLaunch(B->gexec, B->stream));
    } else {
        int rc = enqueue_cycle(B, fn, src, msrc, grid, 1);
        if (rc) return rc;
    }
    BHIPCHK(hipEventRecord(B->c1, B->stream));
    int n_running = 0;
    BHIPCHK(hipMemcpyAsync(&n_running, B->d_ctl + 2, sizeof(int), hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    float ms = 0.f;
    BHIPCHK(hipEventElapsedTime(&ms, B->c0, B->c1));
    B->stats.cycle_ms += ms;
    B->stats.n_cycles += 1;
    if (!B->S.use_graph) {
        for (int L = 0; L < B->S.n_launch; ++L) {
            float m = 0.f;
            BHIPCHK(hipEventElapsedTime(&m, B->ev[2 * L], B->ev[2 * L + 1]));
            B->stats.transport_kernel_ms += m;
        }
    }
    B->stats.transport_launches += B->S.n_launch;
    int rc = check_bstatus(B);
    if (rc) return rc;
    *n_running_out = n_running;
    return 0;
}

}  // namespace

// the pinned staging buffer, grown to at least `bytes` (null if the allocation fails)
static char* pinned(Batch* B, size_t bytes)
{
    if (bytes <= B->pin_bytes) return B->pin;
    size_t want = std::max<size_t>(size_t(1) << 16, B->pin_bytes);
    while (want < bytes) want *= 2;
    if (B->pin) (void)hipHostFree(B->pin);
    B->pin = nullptr;
    B->pin_bytes = 0;
    if (hipHostMalloc((void**)&B->pin, want, hipHostMallocDefault) != hipSuccess) {
        B->pin = nullptr;
        return nullptr;
    }
    B->pin_bytes = want;
    return B->pin;
}

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

static thread_local std::string g_batch_create_err;

const char* scn_batch_last_error(const void* batch)
{
    if (!batch) return g_batch_create_err.c_str();
    return ((const Batch*)batch)->err.c_str();
}

int scn_batch_create(const scn_batch_settings* settings, void** out)
{
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_batch_create_err = "no GPU visible to HIP";
        return SCN_ERR_NO_DEVICE;
    }
    if (!settings || settings->max_slots < 1 || settings->max_slots > MAX_SLOTS || settings->hbm_budget < (1 << 20) ||
        (settings->transport < SCN_TRANSPORT_ST || settings->transport > SCN_TRANSPORT_HT) ||
        (settings->tile != 0 && (settings->tile % BWAVE != 0 || settings->tile > 4096)) || settings->k0 < 0 ||
        settings->n_launch < 0 || settings->n_launch > 16 || settings->refill < 0 || settings->refill > BWAVE) {
        g_batch_create_err = "invalid batch settings";
        return SCN_ERR_ARG;
    }
    Batch* B = new Batch();
    B->S = *settings;
    if (B->S.tile == 0) B->S.tile = 256;
    if (B->S.k0 == 0) B->S.k0 = 8;
    if (B->S.n_launch == 0) B->S.n_launch = 1;
    if (B->S.refill == 0) B->S.refill = DEFAULT_REFILL;
    B->n_slots = settings->max_slots;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) B->n_cu = prop.multiProcessorCount;
    // the largest persistent grid any transport variant of this engine can take (sizes the
    // per-lane buffers of secondaries)
    B->max_lanes = (int)max_grid_lanes(B->S.transport, B->n_cu, lds_bytes(B));
    auto fail = [&](const std::string& m, int code) {
        g_batch_create_err = m;
        scn_batch_destroy(B);
        return code;
    };
    if (hipStreamCreate(&B->stream) != hipSuccess) return fail("hipStreamCreate failed", SCN_ERR_HIP);
    if (hipEventCreate(&B->c0) != hipSuccess || hipEventCreate(&B->c1) != hipSuccess)
        return fail("hipEventCreate failed", SCN_ERR_HIP);
    B->ev.assign(2 * B->S.n_launch, nullptr);
    for (auto& e : B->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail("hipEventCreate failed", SCN_ERR_HIP);
    B->arena_size = (size_t)settings->hbm_budget;
    if (hipMalloc((void**)&B->arena, B->arena_size) != hipSuccess)
        return fail("hipMalloc of the HBM budget failed", SCN_ERR_HIP);
    B->freel[0] = B->arena_size & ~(size_t)255;
    // work items: at most one per batch of every slot's window
    const size_t n_items = (size_t)B->n_slots * MAX_WINDOW;
    if (hipMalloc((void**)&B->d_items, sizeof(Item) * n_items) != hipSuccess ||
        hipMalloc((void**)&B->d_item_cnt, sizeof(int) * n_items) != hipSuccess ||
        hipMalloc((void**)&B->d_slots, sizeof(SlotDev) * B->n_slots) != hipSuccess ||
        hipMalloc((void**)&B->d_state, sizeof(SlotState) * B->n_slots) != hipSuccess ||
        hipMalloc((void**)&B->d_ctl, sizeof(int) * CTL_WORDS) != hipSuccess ||
        hipMalloc((void**)&B->d_status, sizeof(DevStatus)) != hipSuccess)
        return fail("hipMalloc failed", SCN_ERR_HIP);
    B->h_slots.assign(B->n_slots, SlotDev{});
    B->hs.assign(B->n_slots, SlotHost{});
    if (hipMemsetAsync(B->d_state, 0, sizeof(SlotState) * B->n_slots, B->stream) != hipSuccess ||
        hipMemsetAsync(B->d_slots, 0, sizeof(SlotDev) * B->n_slots, B->stream) != hipSuccess ||
        hipMemsetAsync(B->d_ctl, 0, sizeof(int) * CTL_WORDS, B->stream) != hipSuccess ||
        hipMemsetAsync(B->d_status, 0, sizeof(DevStatus), B->stream) != hipSuccess ||
        hipStreamSynchronize(B->stream) != hipSuccess)
        return fail("device initialisation failed", SCN_ERR_HIP);
    *out = B;
    return 0;
}

void scn_batch_destroy(void* batch)
{
    Batch* B = (Batch*)batch;
    if (!B) return;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    drop_graph(B);
    if (B->pin) (void)hipHostFree(B->pin);
    if (B->arena) (void)hipFree(B->arena);
    if (B->d_items) (void)hipFree(B->d_items);
    if (B->d_item_cnt) (void)hipFree(B->d_item_cnt);
    if (B->d_slots) (void)hipFree(B->d_slots);
    if (B->d_state) (void)hipFree(B->d_state);
    if (B->d_ctl) (void)hipFree(B->d_ctl);
    if (B->d_status) (void)hipFree(B->d_status);
    for (auto e : B->ev)
        if (e) (void)hipEventDestroy(e);
    if (B->c0) (void)hipEventDestroy(B->c0);
    if (B->c1) (void)hipEventDestroy(B->c1);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
}

int scn_batch_add(void* batch, const scn_fixed_request* R, int32_t* slot)
{
    Batch* B = (Batch*)batch;
    if (!B || !R || !slot) return SCN_ERR_ARG;
    *slot = -1;
    int depth = 0;
    int rc = validate(B, R, &depth);
    if (rc) return rc;
    int s = -1;
    for (int i = 0; i < B->n_slots; ++i)
        if (!B->hs[i].used) {
            s = i;
            break;
        }
    if (s < 0) {
        B->err = "no free slot";
        return SCN_ERR_BUSY;
    }
    const scn_problem* pr = R->problem;
    const scn_source* sr = R->source;
    const int nG = pr->mg.n_groups, nM = pr->mg.n_mat;
    // table image
    std::vector<int> goff, gidx, tgoff, tgidx;
    group_clerks(pr, goff, gidx);
    group_clerks(pr, tgoff, tgidx, SCN_CLERK_TRACK);
    Image I;
    size_t o_unis = I.put(pr->universes, pr->n_universes);
    size_t o_boxes = I.put(pr->boxes, pr->n_boxes);
    size_t o_surfs = I.put(pr->surfaces, (size_t)std::max(0, pr->n_surfaces));
    size_t o_wwmap = I.put(pr->ww_map, pr->ww_n > 0 ? (size_t)1 : (size_t)0);
    size_t o_wwlo = I.put(pr->ww_lower, (size_t)std::max(0, pr->ww_n));
    size_t o_wwup = I.put(pr->ww_upper, (size_t)std::max(0, pr->ww_n));
    size_t o_gf = I.put(pr->graph_fill, pr->graph_size);
    size_t o_gi = I.put(pr->graph_id, pr->graph_size);
    size_t o_real = I.put(pr->real_data, pr->n_real);
    size_t o_ints = I.put(pr->int_data, pr->n_int);
    size_t o_xs = I.put(pr->mg.xs, (size_t)nM * nG * 6);
    size_t o_P0 = I.put(pr->mg.P0, (size_t)nM * nG * nG);
    size_t o_prod = I.put(pr->mg.prod, (size_t)nM * nG * nG);
    size_t o_chi = I.put(pr->mg.chi, (size_t)nM * nG);
    size_t o_fiss = I.put(pr->mg.fissile, (size_t)nM);
    const std::vector<double> maj_tab = majorant_table(pr);
    size_t o_maj = I.put(maj_tab.data(), maj_tab.size());
    const std::vector<double> st_tab = st_track_table(pr);
    size_t o_stxs = I.put(st_tab.data(), st_tab.size());
    size_t o_clerks = I.put(pr->clerks, (size_t)pr->n_clerks);
    size_t o_goff = I.put(goff.data(), goff.size());
    size_t o_gidx = I.put(gidx.data(), gidx.size());
    size_t o_tgoff = I.put(tgoff.data(), tgoff.size());
    std::vector<int64_t> g_addr;
    const FastColl fc = fast_collision_layout(pr, goff, gidx, g_addr);
    size_t o_gaddr = I.put(g_addr.data(), g_addr.size());
    const std::vector<double> uinv = universe_inverse_table(pr);
    size_t o_uinv = I.put(uinv.data(), uinv.size());
    std::vector<int> flat;
    const int flat_lat = flat_lattice(pr, flat);
    size_t o_flat = I.put(flat.data(), flat.size());
    size_t o_tgidx = I.put(tgidx.data(), tgidx.size());
    // pointSource%init normalisation (pointSource_class.f90:120-128,150-152)
    std::vector<double> probG;
    double dir[3] = {sr->dir[0], sr->dir[1], sr->dir[2]};
    if (sr->type == SCN_SRC_POINT && !sr->isotropic) {
        double n = scn_norm2(sr->dir);
        for (int k = 0; k < 3; ++k) dir[k] = sr->dir[k] / n;
    }
    if (sr->type == SCN_SRC_POINT && sr->n_probG > 0) {
        double t = 0.0;
        for (int g = 0; g < sr->n_probG; ++g) t = t + sr->probG[g];
        probG.resize(sr->n_probG);
        for (int g = 0; g < sr->n_probG; ++g) probG[g] = sr->probG[g] / t;
    }
    size_t o_probG = I.put(probG.data(), probG.size());
    size_t o_P1 = pr->mg.P1 ? I.put(pr->mg.P1, (size_t)nM * nG * nG) : 0;
    SlotLayout Ly;
    Ly.tables = align256(I.bytes.size());
    Ly.queue = Ly.tables;
    Ly.queue_bytes = live_bytes(depth, R->pop);
    Ly.bins = Ly.queue + 2 * Ly.queue_bytes;
    const size_t nb = (size_t)std::max<int64_t>(1, pr->n_bins);
    const int window = R->window > 0 ? R->window : 1;
    // BIN (one copy per batch of the window), CSUM, CSUM2
    const size_t words = pr->fixed_point_tally ? 4 : 1;   // per bin of a BIN copy (fixed point: limb words)
    size_t bins_bytes = align256(sizeof(double) * words * nb * window) + 2 * align256(sizeof(double) * nb);
    Ly.hist = Ly.bins + bins_bytes;
    size_t nh = R->record ? (size_t)R->n_batches * (size_t)R->pop : 0;
    size_t hist_bytes = R->record ? align256(sizeof(uint64_t) * nh) + align256(sizeof(int) * nh) : 0;
    // material source: rejected attempts of every history of the window, from the pre-pass
    Ly.src = Ly.hist + hist_bytes;
    const size_t src_bytes = sr->type == SCN_SRC_POINT ? 0 : align256((size_t)window * (size_t)R->pop);
    Ly.total = Ly.src + src_bytes;
    size_t off = arena_alloc(B, Ly.total);
    if (off == SIZE_MAX) {
        B->err = "not enough of the HBM budget free for this request";
        return SCN_ERR_BUSY;
    }
    // secondaries (fission sites, weight-window splits) need the per-lane buffers, which come out
    // of the same budget
    int needs_stack = pr->ww_n > 0;
    for (int m = 0; m < nM; ++m) needs_stack |= pr->mg.fissile[m] ? 1 : 0;
    if (needs_stack) {
        const int rs = reserve_stack(B, R->buffer > 0 ? R->buffer : DEFAULT_BUFFER);
        if (rs) {
            arena_free(B, off, Ly.total);
            B->err = "not enough of the HBM budget free for this request's buffers of secondaries";
            return rs;
        }
    }
    char* base = B->arena + off;
    SlotDev D{};
    DevProblem& P = D.P;
    P.transport = pr->transport;
    P.st_cache = pr->st_cache;
    P.root_uni = pr->root_uni;
    P.border_box = pr->border_box;
    P.n_mat = nM;
    P.nG = nG;
    P.n_clerks = pr->n_clerks;
    P.graph_size = pr->graph_size;
    P.unis = (const scn_universe*)(base + o_unis);
    P.boxes = (const scn_box*)(base + o_boxes);
    P.surfs = (const scn_surface*)(base + o_surfs);
    P.collision = pr->collision;
    P.max_split = pr->max_split;
    P.ww_n = pr->ww_n;
    P.ww_map = (const scn_clerk*)(base + o_wwmap);
    P.ww_lower = (const double*)(base + o_wwlo);
    P.ww_upper = (const double*)(base + o_wwup);
    P.ww_survival = pr->ww_survival;
    P.any_rot = any_rotated(pr);
    P.n_track = (int)tgidx.size();
    P.gtrack_off = (const int*)(base + o_tgoff);
    P.gtrack_idx = (const int*)(base + o_tgidx);
    P.gfill = (const int*)(base + o_gf);
    P.gid = (const int*)(base + o_gi);
    P.real = (const double*)(base + o_real);
    P.ints = (const int*)(base + o_ints);
    P.xs = (const double*)(base + o_xs);
    P.P0 = (const double*)(base + o_P0);
    P.prod = (const double*)(base + o_prod);
    P.chi = (const double*)(base + o_chi);
    P.fissile = (const int*)(base + o_fiss);
    P.maj = (const double*)(base + o_maj);
    P.st_xs = (const double*)(base + o_stxs);
    P.collision_xs = pr->mg.collision_xs;
    P.clerks = (const scn_clerk*)(base + o_clerks);
    P.gclerk_off = (const int*)(base + o_goff);
    P.gclerk_idx = (const int*)(base + o_gidx);
    P.fc = fc;
    P.fc.g_addr = (const int64_t*)(base + o_gaddr);
    P.uni_inv = (const double*)(base + o_uinv);
    P.flat_lat = flat_lat;
    P.flat = (const int*)(base + o_flat);
    P.fix = pr->fixed_point_tally ? 1 : 0;
    P.n_bins = pr->n_bins;
    P.ht_cutoff = pr->ht_cutoff;
    P.P1 = pr->mg.P1 ? (const double*)(base + o_P1) : nullptr;
    carve_live(base + Ly.queue, depth, R->pop, D.q[0]);
    carve_live(base + Ly.queue + Ly.queue_bytes, depth, R->pop, D.q[1]);
    D.bin = (double*)(base + Ly.bins);
    D.csum = (double*)(base + Ly.bins + align256(sizeof(double) * words * nb * window));
    D.csum2 = (double*)(base + Ly.bins + align256(sizeof(double) * words * nb * window) + align256(sizeof(double) * nb));
    D.hist_rng = R->record ? (uint64_t*)(base + Ly.hist) : nullptr;
    D.hist_coll = R->record ? (int*)(base + Ly.hist + align256(sizeof(uint64_t) * nh)) : nullptr;
    D.probG = probG.empty() ? nullptr : (const double*)(base + o_probG);
    D.src_att = sr->type == SCN_SRC_POINT ? nullptr : (uint8_t*)(base + Ly.src);
    for (int k = 0; k < 3; ++k) {
        D.src_r[k] = sr->r[k];
        D.src_dir[k] = dir[k];
        D.bottom[k] = sr->bottom[k];
        D.top[k] = sr->top[k];
    }
    D.t_low = sr->t_low;
    D.t_high = sr->t_high;
    D.stop_rel = R->stop_rel;
    D.stop_bin = R->stop_bin;
    D.norm_bin = pr->norm_bin;
    D.norm_value = pr->norm_value;
    D.seed = (uint64_t)R->seed;
    D.src_type = sr->type;
    D.src_iso = sr->isotropic;
    D.src_G = sr->G;
    D.src_mat = sr->mat;
    D.n_probG = (int)probG.size();
    D.pop = R->pop;
    D.n_batches = R->n_batches;
    D.buffer = R->buffer > 0 ? R->buffer : DEFAULT_BUFFER;
    D.window = window;
    D.adapt = R->stop_bin >= 0 ? R->adapt : 0;
    SlotState st{};
    st.running = 1;
    st.win = D.adapt > 0 ? std::min(2, window) : 0;
    auto undo = [&](int code) {
        arena_free(B, off, Ly.total);
        return code;
    };
    if (hipMemcpyAsync(base, I.bytes.data(), I.bytes.size(), hipMemcpyHostToDevice, B->stream) != hipSuccess ||
        hipMemsetAsync(base + Ly.bins, 0, bins_bytes, B->stream) != hipSuccess ||
        hipMemcpyAsync(B->d_slots + s, &D, sizeof D, hipMemcpyHostToDevice, B->stream) != hipSuccess ||
        hipMemcpyAsync(B->d_state + s, &st, sizeof st, hipMemcpyHostToDevice, B->stream) != hipSuccess ||
        hipStreamSynchronize(B->stream) != hipSuccess) {
        B->err = "upload of the request failed";
        return undo(SCN_ERR_HIP);
    }
    B->h_slots[s] = D;
    SlotHost& h = B->hs[s];
    h.used = 1;
    h.done_known = 0;
    h.off = off;
    h.bytes = Ly.total;
    h.depth = depth;
    h.geo = needs_full_geometry(pr);
    h.ww = pr->ww_n > 0;
    h.mat_src = sr->type != SCN_SRC_POINT;
    h.n_bins = pr->n_bins;
    h.pop = R->pop;
    h.n_batches = R->n_batches;
    h.record = R->record;
    h.fissile = 0;
    for (int m = 0; m < nM; ++m) h.fissile |= pr->mg.fissile[m] ? 1 : 0;
    h.buffer = D.buffer;
    h.window_max = window;
    *slot = s;
    return 0;
}

int scn_batch_cycle(void* batch, int32_t n_cycles, int32_t* n_running)
{
    Batch* B = (Batch*)batch;
    if (!B || n_cycles < 0) return SCN_ERR_ARG;
    int running = 0;
    for (int c = 0; c < n_cycles; ++c) {
        int rc = run_cycle(B, &running);
        if (rc) return rc;
        if (running == 0) break;
    }
    if (n_cycles == 0) {
        std::vector<SlotState> ss(B->n_slots);
        BHIPCHK(hipMemcpyAsync(ss.data(), B->d_state, sizeof(SlotState) * B->n_slots, hipMemcpyDeviceToHost, B->stream));
        BHIPCHK(hipStreamSynchronize(B->stream));
        for (int i = 0; i < B->n_slots; ++i) running += (B->hs[i].used && ss[i].running) ? 1 : 0;
    }
    if (n_running) *n_running = running;
    return 0;
}

int scn_batch_query(void* batch, int32_t slot, scn_fixed_result* res, double* bins, uint64_t* hist_rng,
                    int32_t* hist_coll)
{
    Batch* B = (Batch*)batch;
    if (!B || slot < 0 || slot >= B->n_slots) return SCN_ERR_ARG;
    SlotHost& h = B->hs[slot];
    scn_fixed_result r{};
    if (!h.used) {
        r.state = SCN_SLOT_FREE;
        if (res) *res = r;
        return 0;
    }
    // the slot's counters and, when asked for, its score sums come back behind one synchronisation
    // (the sums of a slot still running are read and dropped)
    const SlotDev& D = B->h_slots[slot];
    const bool want_bins = bins && h.n_bins > 0;
    const size_t nb = want_bins ? (size_t)h.n_bins : 0;
    char* stage = pinned(B, sizeof(SlotState) + 2 * sizeof(double) * nb);
    if (!stage) {
        B->err = "hipHostMalloc of the readback staging failed";
        return SCN_ERR_HIP;
    }
    double* cs = (double*)(stage + sizeof(SlotState));
    double* cs2 = cs + nb;
    BHIPCHK(hipMemcpyAsync(stage, B->d_state + slot, sizeof(SlotState), hipMemcpyDeviceToHost, B->stream));
    if (want_bins) {
        BHIPCHK(hipMemcpyAsync(cs, D.csum, sizeof(double) * nb, hipMemcpyDeviceToHost, B->stream));
        BHIPCHK(hipMemcpyAsync(cs2, D.csum2, sizeof(double) * nb, hipMemcpyDeviceToHost, B->stream));
    }
    BHIPCHK(hipStreamSynchronize(B->stream));
    SlotState st;
    std::memcpy(&st, stage, sizeof st);
    if (!st.running) {
        h.done_st = st;
        h.done_known = 1;
    }
    r.state = st.running ? SCN_SLOT_RUNNING : SCN_SLOT_DONE;
    r.n_batches = st.batch;
    r.n_bins = h.n_bins;
    r.n_histories = (int64_t)st.histories;
    r.n_collisions = (int64_t)st.collisions;
    r.n_segments = (int64_t)st.segments;
    if (res) *res = r;
    if (r.state != SCN_SLOT_DONE) return 0;
    if (want_bins) {
        const int N = std::max(1, st.batch);
        for (int64_t i = 0; i < h.n_bins; ++i) score_result(cs[i], cs2[i], N, bins + 2 * i);
    }
    if (h.record && (hist_rng || hist_coll)) {
        size_t nh = (size_t)st.batch * (size_t)h.pop;
        if (hist_rng) BHIPCHK(hipMemcpyAsync(hist_rng, D.hist_rng, sizeof(uint64_t) * nh, hipMemcpyDeviceToHost, B->stream));
        if (hist_coll) BHIPCHK(hipMemcpyAsync(hist_coll, D.hist_coll, sizeof(int) * nh, hipMemcpyDeviceToHost, B->stream));
        BHIPCHK(hipStreamSynchronize(B->stream));
    }
    return 0;
}

int scn_batch_set_window(void* batch, int32_t slot, int32_t window)
{
    Batch* B = (Batch*)batch;
    if (!B || slot < 0 || slot >= B->n_slots || !B->hs[slot].used) return SCN_ERR_ARG;
    if (window < 1 || window > B->hs[slot].window_max) {
        B->err = "window outside 1 .. the window the solve was admitted with";
        return SCN_ERR_ARG;
    }
    SlotDev& D = B->h_slots[slot];
    D.window = window;
    BHIPCHK(hipMemcpyAsync(&B->d_slots[slot].window, &D.window, sizeof(int), hipMemcpyHostToDevice, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    return 0;
}

int scn_batch_bin(void* batch, int32_t slot, int64_t bin, double result[2], int32_t* n_batches)
{
    Batch* B = (Batch*)batch;
    if (!B || !result || slot < 0 || slot >= B->n_slots || !B->hs[slot].used) return SCN_ERR_ARG;
    const SlotHost& h = B->hs[slot];
    if (bin < 0 || bin >= h.n_bins) {
        B->err = "bin outside the solve's score memory";
        return SCN_ERR_ARG;
    }
    const SlotDev& D = B->h_slots[slot];
    SlotState st{};
    double c[2] = {0.0, 0.0};
    BHIPCHK(hipMemcpyAsync(&st, B->d_state + slot, sizeof st, hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipMemcpyAsync(&c[0], D.csum + bin, sizeof(double), hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipMemcpyAsync(&c[1], D.csum2 + bin, sizeof(double), hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    if (n_batches) *n_batches = st.batch;
    if (st.batch < 1) {
        result[0] = result[1] = 0.0;
        return 0;
    }
    score_result(c[0], c[1], st.batch, result);
    return 0;
}

int scn_batch_states(void* batch, int32_t* states, int32_t n)
{
    Batch* B = (Batch*)batch;
    if (!B || !states || n < 0 || n > B->n_slots) return SCN_ERR_ARG;
    if (n == 0) return 0;
    SlotState* ss = (SlotState*)pinned(B, sizeof(SlotState) * n);
    if (!ss) {
        B->err = "hipHostMalloc of the readback staging failed";
        return SCN_ERR_HIP;
    }
    BHIPCHK(hipMemcpyAsync(ss, B->d_state, sizeof(SlotState) * n, hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    for (int i = 0; i < n; ++i)
        states[i] = !B->hs[i].used ? SCN_SLOT_FREE : (ss[i].running ? SCN_SLOT_RUNNING : SCN_SLOT_DONE);
    return 0;
}

int scn_batch_release(void* batch, int32_t slot)
{
    Batch* B = (Batch*)batch;
    if (!B || slot < 0 || slot >= B->n_slots) return SCN_ERR_ARG;
    SlotHost& h = B->hs[slot];
    if (!h.used) return 0;
    SlotState st;
    if (h.done_known) {
        st = h.done_st;   // read by the query that saw it done
    } else {
        char* stage = pinned(B, sizeof(SlotState));
        if (!stage) {
            B->err = "hipHostMalloc of the readback staging failed";
            return SCN_ERR_HIP;
        }
        BHIPCHK(hipMemcpyAsync(stage, B->d_state + slot, sizeof(SlotState), hipMemcpyDeviceToHost, B->stream));
        BHIPCHK(hipStreamSynchronize(B->stream));
        std::memcpy(&st, stage, sizeof st);
    }
    if (st.running) {
        B->err = "slot still running";
        return SCN_ERR_ARG;
    }
    B->rel_seg += (int64_t)st.segments;
    B->rel_col += (int64_t)st.collisions;
    B->rel_sc += (int64_t)st.scores;
    B->rel_hist += (int64_t)st.histories;
    // cleared in stream order (before any later upload or cycle); the source outlives the copy
    static const SlotState zero_state{};
    BHIPCHK(hipMemcpyAsync(B->d_state + slot, &zero_state, sizeof zero_state, hipMemcpyHostToDevice, B->stream));
    arena_free(B, h.off, h.bytes);
    h = SlotHost{};
    B->h_slots[slot] = SlotDev{};
    return reserve_stack(B, 0);   // shrinks nothing in use; frees the buffers when no solve needs them
}

int64_t scn_batch_secondary_bytes(int32_t transport, int32_t buffer)
{
    if (transport < SCN_TRANSPORT_ST || transport > SCN_TRANSPORT_HT || buffer < 0) return -1;
    int ndev = 0, n_cu = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return -1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return -1;
    n_cu = prop.multiProcessorCount;
    const size_t lds = lds_bytes_for(Batch{}.S_lds_per_wave);
    return (int64_t)stack_bytes_for(max_grid_lanes(transport, n_cu, lds), buffer > 0 ? buffer : DEFAULT_BUFFER);
}

int scn_batch_stats(void* batch, scn_stats* out)
{
    Batch* B = (Batch*)batch;
    if (!B || !out) return SCN_ERR_ARG;
    std::vector<SlotState> ss(B->n_slots);
    BHIPCHK(hipMemcpyAsync(ss.data(), B->d_state, sizeof(SlotState) * B->n_slots, hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    *out = B->stats;
    int64_t seg = B->rel_seg, col = B->rel_col, sc = B->rel_sc, hist = B->rel_hist;
    for (int i = 0; i < B->n_slots; ++i) {
        if (!B->hs[i].used) continue;
        seg += (int64_t)ss[i].segments;
        col += (int64_t)ss[i].collisions;
        sc += (int64_t)ss[i].scores;
        hist += (int64_t)ss[i].histories;
    }
    out->n_segments = seg;
    out->n_collisions = col;
    out->n_scores = sc;
    out->n_histories = hist;
    return 0;
}

}  // extern "C"
