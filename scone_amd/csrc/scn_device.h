// scn_device.h -- device-side physics of the multigroup tracking loop (gfx950).
//
// One particle lives in registers as a Pt<NL, GEO> (NL = compile-time nesting depth of the geometry,
// so every level array is statically indexed and stays in VGPRs).  The arithmetic of every
// routine reproduces the reference's operation order -- the same order the CPU oracle uses --
// so a history computed here is bit-identical to the oracle's:
//   surfaces      Geometry/Surfaces/CompositeSurfaces/{box,squareCylinder}_class.f90,
//                 QuadSurfaces/cylinder_class.f90
//   universes     Geometry/Universes/{root,lat,pin}Universe_class.f90, universe_inter.f90 enter
//   geometry      Geometry/geometryStd_class.f90 placeCoord/diveToMat/closestDist_cache/
//                 move_withCache/teleport, coord_class.f90 moveLocal/moveGlobal
//   transport     TransportOperator/transportOperator{ST,DT}_class.f90
//   collision     CollisionOperator/CollisionProcessors/{collisionProcessor_inter,neutronMGstd_class}.f90
//   tallies       Tallies/TallyClerks/{collisionClerk,keffImplicitClerk}_class.f90 with the maps,
//                 energyFilter and flux/macro responses
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scone_gpu.h"
#include "scn_fpmath.h"

// place_coord takes the flat-lattice branch (place_flat) when the problem has one; a translation
// unit may compile it out (the result is the same, by the general walk)
#ifndef SCN_FLAT_PLACEMENT
#define SCN_FLAT_PLACEMENT 1
#endif

namespace scn {

// ---- RNG jumps --------------------------------------------------------------------------------
// A_i, C_i with s -> A_i*s + C_i (mod 2^63) the LCG advanced by 2^i steps: A_0 = g, C_0 = 1,
// A_{i+1} = A_i^2, C_{i+1} = C_i*(A_i + 1).  Skipping k steps applies the maps of k's set bits
// (they commute), one 64-bit multiply per set bit, where Brown's algorithm (scn_rng_skip,
// RNG_class.f90:251-304) composes the maps with up to five per bit.  Same states: both are the
// exact k-fold composition of the generator mod 2^63.
struct RngJumps {
    uint64_t a[63];
    uint64_t c[63];
};

constexpr RngJumps make_rng_jumps()
{
    RngJumps t{};
    uint64_t a = SCN_LCG_G, c = 1ULL;
    for (int i = 0; i < 63; ++i) {
        t.a[i] = a;
        t.c[i] = c;
        c = (c * (a + 1ULL)) & SCN_LCG_MASK;
        a = (a * a) & SCN_LCG_MASK;
    }
    return t;
}

static __constant__ RngJumps RNG_JUMPS = make_rng_jumps();

// scn_rng_skip on the device: k < 0 wraps through the period 2^63 the same way
__device__ __forceinline__ uint64_t rng_skip(uint64_t s, int64_t k_in)
{
    uint64_t k = (k_in >= 0) ? (uint64_t)k_in : (uint64_t)(INT64_MAX - (-k_in) + 1);
    k &= SCN_LCG_MASK;
    for (int i = 0; k != 0; ++i, k >>= 1)
        if (k & 1ULL) s = (RNG_JUMPS.a[i] * s + RNG_JUMPS.c[i]) & SCN_LCG_MASK;
    return s;
}

static __constant__ scn_rng_stride_table RNG_STRIDES = scn_make_rng_stride_table();
static __constant__ scn_rng_skip4_table RNG_SKIP4 = scn_make_rng_skip4_table();

// stride by n histories: the five byte maps of RNG_STRIDES (scn_rng_stride_bytes), the same state
// as rng_skip(s, 152917 n)
__device__ __forceinline__ uint64_t rng_stride(uint64_t s, int64_t n)
{
    if (n < 0 || n >= ((int64_t)1 << 40)) return rng_skip(s, SCN_STRIDE * n);
    const uint64_t u = (uint64_t)n;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int b = (int)((u >> (8 * k)) & 255ULL);
        s = (RNG_STRIDES.a[k][b] * s + RNG_STRIDES.c[k][b]) & SCN_LCG_MASK;
    }
    return s;
}

// ---- collision clerks without a table walk ----------------------------------------------------
// When the collision clerks fall into at most FAST_CLERKS layouts (maps, responses, virtual-
// collision flag), with at most 3 maps each, all lin-grid spaceMaps (a total-flux clerk has none),
// and each group is admitted by at most one clerk of a layout (a mesh tally split into one clerk
// per group is one layout), a score needs no walk of the clerk table: each layout's map
// parameters, responses and virtual-collision flag live in the problem record (wave-uniform
// loads), and the admitting clerk's first bin comes from a per-(layout, group) table (-1: no clerk
// of the layout admits the group).  The bins and the arithmetic are the general path's (map1d,
// collisionClerk reportInColl).
constexpr int FAST_CLERKS = 4;
struct FastClerk {
    int n_maps;
    int width;
    int handle_virtual;
    int pad;
    int resp[8];         // SCN_MT_* of each response
    int axis[3];
    int nbins[3];
    int multi[3];
    double g0[3];        // grid minimum (real_data[data])
    double step[3];      // grid step (real_data[data + 1])
    double step_inv[3];  // 1 / step (only steers scn_floor_div, which checks its answer)
};
struct FastColl {
    int on;
    int n_clerks;
    FastClerk c[FAST_CLERKS];
    const int64_t* g_addr;   // [n_clerks][nG]: mem_addr of the layout's clerk admitting group g, or -1
};

// ---- problem tables as one by-value kernel argument ---------------------------------------
struct DevProblem {
    int transport, st_cache, root_uni, border_box, n_mat, nG, n_clerks, graph_size;
    const scn_universe* unis;
    const scn_box* boxes;
    const int* gfill;
    const int* gid;
    const double* real;
    const int* ints;
    const double* xs;
    const double* P0;
    const double* prod;
    const double* chi;
    const int* fissile;
    const double* maj;   // [nG][2]: max(majorant(G) + 0, collisionXS) and its reciprocal
    double collision_xs;
    const scn_clerk* clerks;
    const int* gclerk_off;   // [nG+1] clerks whose energy filter admits group g
    const int* gclerk_idx;
    int64_t n_bins;
    double ht_cutoff;
    const double* P1;    // [mat][G_in][G_out] 3 P1/P0, or null (P0 scattering)
    const scn_surface* surfs;   // surfaces of cells and of a non-box root border
    int any_rot;         // some universe is rotated: lower-level directions differ from the global one
    int n_track;         // track-length clerks present (scored by the general kernels only)
    int collision;       // SCN_COLL_STD / SCN_COLL_IMP
    int max_split;       // neutronMGimp maxSplit
    int ww_n;            // weight-window bins (0: none; general kernels only)
    const scn_clerk* ww_map;
    const double* ww_lower;
    const double* ww_upper;
    double ww_survival;
    const int* gtrack_off;   // [nG+1] track clerks whose energy filter admits group g
    const int* gtrack_idx;
    FastColl fc;             // shared-layout collision clerks (fc.on = 0: walk the clerk table)
    const double* uni_inv;   // [n_universes][6]: 1/pitch(3) and 1/size(0..1) of lattices (scn_floor_div, scn_idiv)
    int fix;                 // 1: BIN holds four 64-bit limb words per bin (scn_fix), order-independent sums
    int flat_lat;            // flat lattice (flat_lattice in scn_host.h): its universe index, 0 otherwise
    const int* flat;         // [2 + 5 n_local]: root-outside (mat, uid), then per lattice cell (mat, uid,
                             // nest, pin universe, pin graph root)
    const double* st_xs;     // [n_mat*nG + 1][2]: surface tracking's sigmaTrack and its reciprocal
};

enum { EV_COLL = 1, EV_BOUNDARY = 2, EV_CROSS = 3 };
enum { SURF_BORDER = 1, PIN_IN = -1, PIN_OUT = -2, LAT_OUTLINE = -7 };
enum { XS_TOTAL = 0, XS_SCATTER = 1, XS_CAPTURE = 2, XS_FISSION = 3, XS_NUFISSION = 4, XS_KAPPA = 5 };
enum { FATE_NO = 5000, FATE_ABS = 5001, FATE_LEAK = 5002 };
enum { K_IMP_PROD = 0, K_SCATTER_PROD = 1, K_IMP_ABS = 2, K_ANA_LEAK = 3 };
// a flight that needs more segments than this is treated as a lost particle (error, not a hang)
constexpr int MAX_SEGMENTS_PER_FLIGHT = 1 << 22;
constexpr int SCN_ERR_LOST_DETAIL = -77;

// error flags raised by kernels (first one wins, host maps to SCN_ERR_*)
struct DevStatus {
    int code;
    int detail;
};

__device__ __forceinline__ void raise_error(DevStatus* st, int code, int detail)
{
    atomicCAS(&st->code, 0, code);
    st->detail = detail;
}

// per-thread event counters (segments, real collisions, clerk score additions)
struct Cnt {
    int64_t seg, col, sc;
};

template <int NL, int GEO = 0>
struct Pt {
    double r[NL][3];   // position at each nesting level (coordList%lvl(l)%r)
    double u[3];       // direction (identical at every level: no rotated universes)
    int nest;
    int uni[NL], root[NL], loc[NL];
    int mat, uid;
    int G;
    double w, w0, t;
    uint64_t rng;
    int brood, nsite, coll;
    int sp, lane;      // fixed source: depth of the lane's LIFO buffer of secondaries, lane index
    int split;         // neutronMGimp splitCount (0 for a particle taken from a bank or buffer)
    double k[4];       // keffImplicitClerk scores of this history
    double track;      // trackingCache(1)%xs
    int fate;
    int dead;
};

// register arrays are only ever indexed with compile-time constants (a runtime index would move
// the whole array to scratch); these select the component of a runtime axis instead
// (blended through the bit patterns: a plain select of the three loads gets folded into a load from
// a selected address, which forces the array out of registers)
__device__ __forceinline__ double get3(const double v[3], int i)
{
    uint64_t a = scn_dbits(v[0]), b = scn_dbits(v[1]), c = scn_dbits(v[2]);
    uint64_t m0 = 0ull - (uint64_t)(i == 0), m1 = 0ull - (uint64_t)(i == 1), m2 = 0ull - (uint64_t)(i == 2);
    return scn_bitsd((a & m0) | (b & m1) | (c & m2));
}
__device__ __forceinline__ void set3(double v[3], int i, double x)
{
    if (i == 0) v[0] = x;
    else if (i == 1) v[1] = x;
    else v[2] = x;
}

// ---------------------------------------------------------------------------------------------
// surfaces
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double box_evaluate(const scn_box& b, const double r[3])
{
    double c = -SCN_HUGE;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i >= b.n_axes) break;
        double t = fabs(get3(r, b.axes[i]) - b.origin[i]) - b.halfwidth[i];
        if (t > c) c = t;
    }
    return c;
}

__device__ __forceinline__ double box_distance(const scn_box& b, const double r[3], const double u[3])
{
    double far = SCN_HUGE, near = -SCN_HUGE;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i >= b.n_axes) break;
        int ax = b.axes[i];
        double rax = get3(r, ax), uax = get3(u, ax);
        double rb = rax - b.origin[i];
        double a_far = scn_sign(b.halfwidth[i], uax);
        double a_near = -a_far;
        double tn, tf;
        if (uax != 0.0) {
            tn = (a_near - rb) / uax;
            tf = (a_far - rb) / uax;
        } else {
            tn = scn_sign(SCN_INF, a_near - rb);
            tf = scn_sign(SCN_INF, a_far - rb);
            if (tn > tf) {
                double t = tn;
                tn = tf;
                tf = t;
            }
        }
        far = scn_fmin(far, tf);
        near = scn_fmax(near, tn);
    }
    double d;
    if (far <= near * SCN_MISS_TOL) {
        d = SCN_INF;
    } else if (fabs(box_evaluate(b, r)) < SCN_SURF_TOL) {
        d = (fabs(far) >= fabs(near)) ? far : near;
    } else {
        d = (near <= 0.0) ? far : near;
    }
    if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
    return d;
}

__device__ __forceinline__ int box_halfspace(const scn_box& b, const double r[3], const double u[3])
{
    double c = box_evaluate(b, r);
    int hs = c > 0.0;
    if (fabs(c) < SCN_SURF_TOL) {
        int maxCom = 0;
        double best = 0.0, rl = 0.0, ul = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (i >= b.n_axes) break;
            double ri = get3(r, b.axes[i]) - b.origin[i];
            double t = fabs(ri) - b.halfwidth[i];
            if (i == 0 || t > best) {
                best = t;
                maxCom = i;
                rl = ri;
                ul = get3(u, b.axes[i]);
            }
        }
        (void)maxCom;
        double proj = ul * scn_sign(1.0, rl);
        hs = proj > 0.0;
        if (proj == 0.0) hs = c >= 0.0;
    }
    return hs;
}

__device__ __forceinline__ void box_explicit_bc(const scn_box& b, double r[3], double u[3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i >= b.n_axes) break;
        int ax = b.axes[i];
        double rax = get3(r, ax);
        double r0 = rax - b.origin[i];
        if (fabs(r0) <= b.halfwidth[i] * (1.0 - SCN_SURF_TOL)) continue;
        int bc = (r0 < 0.0) ? b.bc[2 * ax] : b.bc[2 * ax + 1];
        if (bc == SCN_BC_REFLECTIVE) set3(u, ax, -get3(u, ax));
        else if (bc == SCN_BC_PERIODIC) set3(r, ax, rax - 2.0 * scn_sign(b.halfwidth[i], r0));
    }
}

__device__ __forceinline__ void box_transform_bc(const scn_box& b, double r[3], double u[3])
{
    int Ri[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i >= b.n_axes) break;
        double a_bar = b.halfwidth[i] * (1.0 - SCN_SURF_TOL);
        double q = ceil(fabs(get3(r, b.axes[i]) - b.origin[i]) / a_bar);
        Ri[i] = (q < 2.0e6) ? (int)q / 2 : 1000000;   // bounded image count (finite flights only)
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i >= b.n_axes) break;
        int ax = b.axes[i];
        double rax = get3(r, ax), uax = get3(u, ax);
        for (int t = 1; t <= Ri[i]; ++t) {
            double r0 = rax - b.origin[i];
            int bc = (r0 < 0.0) ? b.bc[2 * ax] : b.bc[2 * ax + 1];
            if (bc == SCN_BC_REFLECTIVE) {
                double a0 = scn_sign(b.halfwidth[i], r0) + b.origin[i];
                double d = rax - a0;
                rax = rax - 2.0 * d;
                uax = -uax;
            } else if (bc == SCN_BC_PERIODIC) {
                double d = scn_sign(b.halfwidth[i], r0);
                rax = rax - 2.0 * d;
            }
        }
        set3(r, ax, rax);
        set3(u, ax, uax);
    }
}

__device__ __forceinline__ double cyl_distance(double radius, double r_sq, const double r[3], const double u[3])
{
    double tol = 2.0 * radius * SCN_SURF_TOL;
    double c = (0.0 + r[0] * r[0] + r[1] * r[1]) - r_sq;
    double k = 0.0 + r[0] * u[0] + r[1] * u[1];
    double a = 1.0 - u[2] * u[2];
    double delta = k * k - a * c;
    double d;
    if (delta < 0.0 || a == 0.0) {
        d = SCN_INF;
    } else if (fabs(c) < tol) {
        if (k >= 0.0) {
            d = SCN_INF;
        } else {
            d = -k + sqrt(delta);
            d = d / a;
        }
    } else if (c < 0.0) {
        d = -k + sqrt(delta);
        d = d / a;
    } else {
        d = -k - sqrt(delta);
        d = d / a;
        if (d <= 0.0) d = SCN_INF;
    }
    return scn_fmin(d, SCN_INF);
}


// ---- surfaces of cells and of a non-box root border (Geometry/Surfaces) --------------------
// dot_product / sum as a running sum from zero in index order; x**2 as x*x.
__device__ __forceinline__ double dot3(const double a[3], const double b[3])
{
    return 0.0 + a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// coordinates across an axis (cylinder_class.f90 build: plane)
__device__ __forceinline__ void plane_of(int axis, int& p0, int& p1)
{
    p0 = (axis == 0) ? 1 : 0;
    p1 = (axis == 2) ? 1 : 2;
}

// shared quadratic root choice of cylinder_class.f90 / quadric_class.f90 distance
__device__ __forceinline__ double quad_root(double a, double k, double c, double tol)
{
    const double delta = k * k - a * c;
    double d;
    if (delta < 0.0 || a == 0.0) {
        d = SCN_INF;
    } else if (fabs(c) < tol) {
        if (k >= 0.0) {
            d = SCN_INF;
        } else {
            d = -k + sqrt(delta);
            d = d / a;
        }
    } else if (c < 0.0) {
        d = -k + sqrt(delta);
        d = d / a;
    } else {
        d = -k - sqrt(delta);
        d = d / a;
        if (d <= 0.0) d = SCN_INF;
    }
    return scn_fmin(d, SCN_INF);
}

// ---- truncCone (truncCone_class.f90) and wedge (wedge_class.f90); same operation order as the
// oracle's restatement of each routine
#define SCN_EPS 2.220446049250313e-16   // epsilon(ONE)

__device__ __forceinline__ double cone_evaluate(const scn_surface& S, const double r[3])
{
    const double* c = S.c;
    int p0, p1;
    plane_of(S.axis, p0, p1);
    const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1], da = get3(r, S.axis) - c[S.axis];
    const double v = (0.0 + d0 * d0 + d1 * d1) - c[3] * (da * da);
    return scn_fmax(v, scn_fmax(-da + c[4], da - c[5]));
}

__device__ __forceinline__ double cone_distance(const scn_surface& S, const double r[3], const double u[3])
{
    const double* c = S.c;
    int p0, p1;
    plane_of(S.axis, p0, p1);
    const int ax = S.axis;
    double diff[3], orient[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) diff[i] = r[i] - c[i];
    set3(orient, ax, scn_sign(1.0, c[4]));
    const double tsq = c[3];
    const double up0 = get3(u, p0), up1 = get3(u, p1), ua = get3(u, ax);
    const double dp0 = get3(diff, p0), dp1 = get3(diff, p1), da = get3(diff, ax);
    const double a = (0.0 + up0 * up0 + up1 * up1) - tsq * (ua * ua);
    const double b = (0.0 + dp0 * up0 + dp1 * up1) - tsq * da * ua;
    const double cc = (0.0 + dp0 * dp0 + dp1 * dp1) - tsq * (da * da);
    const double delta = b * b - a * cc;
    double d1, d2, t[3];
    if (fabs(a) < SCN_EPS) {
        d1 = -(0.5 * cc / b);
        d2 = scn_sign(SCN_INF, dot3(u, orient));
#pragma unroll
        for (int i = 0; i < 3; ++i) t[i] = diff[i] + d1 * u[i];
        if (!(dot3(t, orient) > 0.0)) d1 = d2;
    } else if (delta < SCN_EPS) {
        d1 = -SCN_INF;
        d2 = SCN_INF;
    } else {
        d1 = -(b + sqrt(delta)) / a;
        d2 = -(b - sqrt(delta)) / a;
#pragma unroll
        for (int i = 0; i < 3; ++i) t[i] = diff[i] + d1 * u[i];
        if (!(dot3(t, orient) > 0.0)) d1 = scn_sign(SCN_INF, dot3(u, orient));
#pragma unroll
        for (int i = 0; i < 3; ++i) t[i] = diff[i] + d2 * u[i];
        if (!(dot3(t, orient) > 0.0)) d2 = scn_sign(SCN_INF, dot3(u, orient));
    }
    double near = scn_fmin(d1, d2), far = scn_fmax(d1, d2);
    if (fabs(ua) > SCN_EPS) {
        d1 = (c[5] - da) / ua;
        d2 = (c[4] - da) / ua;
    } else if (da > c[4] && da < c[5]) {
        d1 = -SCN_INF;
        d2 = SCN_INF;
    } else {
        d1 = -SCN_INF;
        d2 = -SCN_INF;
    }
    far = scn_fmin(far, scn_fmax(d1, d2));
    near = scn_fmax(near, scn_fmin(d1, d2));
    double dist;
    if (far <= near * SCN_MISS_TOL) dist = SCN_INF;
    else if (fabs(cone_evaluate(S, r)) < S.tol) dist = (fabs(far) >= fabs(near)) ? far : near;
    else dist = (near <= 0.0) ? far : near;
    if (dist <= 0.0 || dist > SCN_INF) dist = SCN_INF;
    return dist;
}

__device__ __forceinline__ int cone_going(const scn_surface& S, const double r[3], const double u[3])
{
    const double* c = S.c;
    int p0, p1;
    plane_of(S.axis, p0, p1);
    const int ax = S.axis;
    const double da = get3(r, ax) - c[ax];
    double n[3] = {0.0, 0.0, 0.0};
    if (fabs(da - c[4]) < S.tol) {
        set3(n, ax, -1.0);
    } else if (fabs(da - c[5]) < S.tol) {
        set3(n, ax, 1.0);
    } else {
        set3(n, p0, get3(r, p0) - c[p0]);
        set3(n, p1, get3(r, p1) - c[p1]);
        set3(n, ax, -c[3] * da);
    }
    const double nn = scn_norm2(n);
#pragma unroll
    for (int i = 0; i < 3; ++i) n[i] = n[i] / nn;
    const double proj = dot3(n, u);
    if (proj == 0.0) return cone_evaluate(S, r) >= 0.0;
    return proj > 0.0;
}

__device__ __forceinline__ double wedge_evaluate(const scn_surface& S, const double r[3])
{
    const double* c = S.c;
    int p0, p1;
    plane_of(S.axis, p0, p1);
    const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1];
    const double cAxis = fabs(get3(r, S.axis) - c[S.axis]) - c[3];
    const double cSide1 = 0.0 + c[5] * d0 + c[6] * d1;
    const double cSide2 = 0.0 + c[7] * d0 + c[8] * d1;
    const double cBase = (0.0 + c[9] * d0 + c[10] * d1) - c[4];
    return scn_fmax(scn_fmax(scn_fmax(cAxis, cBase), cSide1), cSide2);
}

__device__ __forceinline__ void wedge_side(double k, double cc, double& near, double& far)
{
    if (fabs(k) < SCN_EPS) {
        if (cc > 0.0) {
            near = SCN_INF;
            far = -SCN_INF;
        }
    } else {
        const double d = -cc / k;
        if (k > 0.0) far = scn_fmin(far, d);
        else near = scn_fmax(near, d);
    }
}

__device__ __forceinline__ double wedge_distance(const scn_surface& S, const double r[3], const double u[3])
{
    const double* c = S.c;
    int p0, p1;
    plane_of(S.axis, p0, p1);
    const int ax = S.axis;
    const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1], da = get3(r, ax) - c[ax];
    const double up0 = get3(u, p0), up1 = get3(u, p1), ua = get3(u, ax);
    double near = -SCN_INF, far = SCN_INF;
    wedge_side(0.0 + up0 * c[5] + up1 * c[6], 0.0 + d0 * c[5] + d1 * c[6], near, far);
    wedge_side(0.0 + up0 * c[7] + up1 * c[8], 0.0 + d0 * c[7] + d1 * c[8], near, far);
    {
        const double k = 0.0 + up0 * c[9] + up1 * c[10];
        const double cc = 0.0 + d0 * c[9] + d1 * c[10];
        if (fabs(k) < SCN_EPS) {
            if (cc > c[4]) {
                near = SCN_INF;
                far = -SCN_INF;
            }
        } else {
            const double d = -(cc - c[4]) / k;
            if (k > 0.0) far = scn_fmin(far, d);
            else near = scn_fmax(near, d);
        }
    }
    double tn, tf;
    if (fabs(ua) > SCN_EPS) {
        const double e1 = (c[3] - da) / ua;
        const double e2 = (-c[3] - da) / ua;
        tn = scn_fmin(e1, e2);
        tf = scn_fmax(e1, e2);
    } else if (fabs(da) < c[3]) {
        tn = -SCN_INF;
        tf = SCN_INF;
    } else {
        tn = -SCN_INF;
        tf = -SCN_INF;
    }
    far = scn_fmin(far, tf);
    near = scn_fmax(near, tn);
    const double ev = wedge_evaluate(S, r);
    double dist;
    if (far <= near * SCN_MISS_TOL) dist = SCN_INF;
    else if (fabs(ev) < S.tol) dist = (fabs(far) >= fabs(near)) ? far : near;
    else if (ev < 0.0) dist = far;
    else dist = near;
    if (dist <= 0.0 || dist > SCN_INF) dist = SCN_INF;
    return dist;
}

__device__ __forceinline__ int wedge_going(const scn_surface& S, const double r[3], const double u[3])
{
    const double* c = S.c;
    int p0, p1;
    plane_of(S.axis, p0, p1);
    const int ax = S.axis;
    const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1], da = get3(r, ax) - c[ax];
    const double cSide1 = 0.0 + c[5] * d0 + c[6] * d1;
    const double cSide2 = 0.0 + c[7] * d0 + c[8] * d1;
    double n[3] = {0.0, 0.0, 0.0};
    if (fabs(da - c[3]) < S.tol) {
        set3(n, ax, 1.0);
    } else if (fabs(da + c[3]) < S.tol) {
        set3(n, ax, -1.0);
    } else {
        int k = 9;
        if (fabs(cSide1) < S.tol) k = 5;
        else if (fabs(cSide2) < S.tol) k = 7;
        const double a0 = (k == 5) ? c[5] : (k == 7) ? c[7] : c[9];
        const double a1 = (k == 5) ? c[6] : (k == 7) ? c[8] : c[10];
        set3(n, p0, a0);
        set3(n, p1, a1);
    }
    const double nn = scn_norm2(n);
#pragma unroll
    for (int i = 0; i < 3; ++i) n[i] = n[i] / nn;
    const double proj = dot3(n, u);
    if (fabs(proj) < SCN_EPS) return wedge_evaluate(S, r) >= 0.0;
    return proj > 0.0;
}

__device__ __forceinline__ double surf_evaluate(const DevProblem& P, const scn_surface& S, const double r[3])
{
    const double* c = S.c;
    switch (S.type) {
    case SCN_SURF_APLANE: return get3(r, S.axis) - c[0];
    case SCN_SURF_PLANE: {
        const double n[3] = {c[0], c[1], c[2]};
        return dot3(r, n) - c[3];
    }
    case SCN_SURF_SPHERE: {
        const double d[3] = {r[0] - c[0], r[1] - c[1], r[2] - c[2]};
        return dot3(d, d) - c[4];
    }
    case SCN_SURF_CYLINDER: {
        int p0, p1;
        plane_of(S.axis, p0, p1);
        const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1];
        return (0.0 + d0 * d0 + d1 * d1) - c[4];
    }
    case SCN_SURF_QUADRIC: {
        const double x = r[0], y = r[1], z = r[2];
        return x * (c[0] * x + c[3] * y + c[5] * z + c[6]) + y * (c[1] * y + c[4] * z + c[7]) +
               z * (c[2] * z + c[8]) + c[9];
    }
    case SCN_SURF_BOX: return box_evaluate(P.boxes[S.axis], r);
    case SCN_SURF_TRUNC_CONE: return cone_evaluate(S, r);
    case SCN_SURF_WEDGE: return wedge_evaluate(S, r);
    default: {   // SCN_SURF_TRUNC_CYLINDER
        int p0, p1;
        plane_of(S.axis, p0, p1);
        const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1];
        const double v = ((0.0 + d0 * d0 + d1 * d1) - c[3] * c[3]) / c[3] * 0.5;
        return scn_fmax(v, fabs(get3(r, S.axis) - c[S.axis]) - c[4]);
    }
    }
}

__device__ __forceinline__ double surf_distance(const DevProblem& P, const scn_surface& S, const double r[3],
                                                const double u[3])
{
    const double* c = S.c;
    switch (S.type) {
    case SCN_SURF_APLANE: {
        const double ra = c[0] - get3(r, S.axis);
        const double ua = get3(u, S.axis);
        double d = SCN_INF;
        if (!(fabs(ra) < S.tol) && ua != 0.0) d = ra / ua;
        if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
        return d;
    }
    case SCN_SURF_PLANE: {
        const double n[3] = {c[0], c[1], c[2]};
        const double k = dot3(u, n);
        const double cc = dot3(r, n) - c[3];
        if (k == 0.0 || fabs(cc) < S.tol) return SCN_INF;
        double d = -cc / k;
        if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
        return d;
    }
    case SCN_SURF_SPHERE: {
        const double dr[3] = {r[0] - c[0], r[1] - c[1], r[2] - c[2]};
        const double cc = dot3(dr, dr) - c[4];
        const double k = dot3(dr, u);
        const double delta = k * k - cc;
        double d;
        if (delta < 0.0) {
            d = SCN_INF;
        } else if (fabs(cc) < S.tol) {
            d = (k >= 0.0) ? SCN_INF : -k + sqrt(delta);
        } else if (cc < 0.0) {
            d = -k + sqrt(delta);
        } else {
            d = -k - sqrt(delta);
            if (d <= 0.0) d = SCN_INF;
        }
        return d;
    }
    case SCN_SURF_CYLINDER: {
        int p0, p1;
        plane_of(S.axis, p0, p1);
        const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1];
        const double cc = (0.0 + d0 * d0 + d1 * d1) - c[4];
        const double k = 0.0 + d0 * get3(u, p0) + d1 * get3(u, p1);
        const double ua = get3(u, S.axis);
        return quad_root(1.0 - ua * ua, k, cc, S.tol);
    }
    case SCN_SURF_QUADRIC: {
        const double x = r[0], y = r[1], z = r[2], u1 = u[0], u2 = u[1], u3 = u[2];
        const double cc = x * (c[0] * x + c[3] * y + c[5] * z + c[6]) + y * (c[1] * y + c[4] * z + c[7]) +
                          z * (c[2] * z + c[8]) + c[9];
        const double k = c[0] * u1 * x + c[1] * u2 * y + c[2] * u3 * z +
                         0.5 * (c[3] * (u1 * y + u2 * x) + c[4] * (u2 * z + u3 * y) + c[5] * (u3 * x + u1 * z) +
                                c[6] * u1 + c[7] * u2 + c[8] * u3);
        const double a = u1 * (c[0] * u1 + c[3] * u2 + c[5] * u3) + u2 * (c[1] * u2 + c[4] * u3) + c[2] * u3 * u3;
        return quad_root(a, k, cc, S.tol);
    }
    case SCN_SURF_BOX: return box_distance(P.boxes[S.axis], r, u);
    case SCN_SURF_TRUNC_CONE: return cone_distance(S, r, u);
    case SCN_SURF_WEDGE: return wedge_distance(S, r, u);
    default: {   // SCN_SURF_TRUNC_CYLINDER: cylinder slab intersected with the axial slab
        int p0, p1;
        plane_of(S.axis, p0, p1);
        const double rad = c[3], hw = c[4];
        const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1];
        const double c1 = (0.0 + d0 * d0 + d1 * d1) - rad * rad;
        const double k = 0.0 + d0 * get3(u, p0) + d1 * get3(u, p1);
        const double uax = get3(u, S.axis);
        const double a = 1.0 - uax * uax;
        const double delta = k * k - a * c1;
        double far, near;
        if (delta <= 0.0 || a == 0.0) {
            far = SCN_INF;
            near = scn_sign(SCN_INF, c1);
        } else {
            far = (-k + sqrt(delta)) / a;
            near = (-k - sqrt(delta)) / a;
            if (far < near) {
                const double t = far;
                far = near;
                near = t;
            }
        }
        const double rb = get3(r, S.axis) - c[S.axis];
        double tn, tf;
        if (uax != 0.0) {
            tn = (-hw - rb) / uax;
            tf = (hw - rb) / uax;
        } else {
            tn = scn_sign(SCN_INF, -hw - rb);
            tf = scn_sign(SCN_INF, hw - rb);
        }
        if (tf < tn) {
            const double t = tf;
            tf = tn;
            tn = t;
        }
        far = scn_fmin(far, tf);
        near = scn_fmax(near, tn);
        const double cc = scn_fmax(c1 / rad * 0.5, fabs(rb) - hw);
        double d;
        if (far <= near * SCN_MISS_TOL) d = SCN_INF;
        else if (fabs(cc) < S.tol) d = (fabs(far) >= fabs(near)) ? far : near;
        else d = (near <= 0.0) ? far : near;
        if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
        return d;
    }
    }
}

// surface_inter.f90 halfspace with each type's going (box types: box_halfspace)
__device__ __forceinline__ int surf_halfspace(const DevProblem& P, const scn_surface& S, const double r[3],
                                              const double u[3])
{
    if (S.type == SCN_SURF_BOX) return box_halfspace(P.boxes[S.axis], r, u);
    const double cv = surf_evaluate(P, S, r);
    if (!(fabs(cv) < S.tol)) return cv > 0.0;
    const double* c = S.c;
    switch (S.type) {
    case SCN_SURF_APLANE: {
        const double ua = get3(u, S.axis);
        return (ua == 0.0) ? (get3(r, S.axis) - c[0]) >= 0.0 : ua > 0.0;
    }
    case SCN_SURF_PLANE: {
        const double n[3] = {c[0], c[1], c[2]};
        const double proj = dot3(u, n);
        return (proj == 0.0) ? (dot3(r, n) - c[3]) >= 0.0 : proj > 0.0;
    }
    case SCN_SURF_SPHERE: {
        const double dr[3] = {r[0] - c[0], r[1] - c[1], r[2] - c[2]};
        return dot3(dr, u) >= 0.0;
    }
    case SCN_SURF_CYLINDER: {
        int p0, p1;
        plane_of(S.axis, p0, p1);
        const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1];
        return (0.0 + d0 * get3(u, p0) + d1 * get3(u, p1)) >= 0.0;
    }
    case SCN_SURF_QUADRIC: {   // quadric_class.f90 going: normalised gradient . u
        double n[3];
        n[0] = 2.0 * c[0] * r[0] + c[3] * r[1] + c[5] * r[2] + c[6];
        n[1] = 2.0 * c[1] * r[1] + c[3] * r[0] + c[4] * r[2] + c[7];
        n[2] = 2.0 * c[2] * r[2] + c[4] * r[1] + c[5] * r[0] + c[8];
        const double nn = scn_norm2(n);
#pragma unroll
        for (int i = 0; i < 3; ++i) n[i] = n[i] / nn;
        return dot3(n, u) >= 0.0;
    }
    case SCN_SURF_TRUNC_CONE: return cone_going(S, r, u);
    case SCN_SURF_WEDGE: return wedge_going(S, r, u);
    default: {   // SCN_SURF_TRUNC_CYLINDER
        int p0, p1;
        plane_of(S.axis, p0, p1);
        const double rad = c[3];
        const double d0 = get3(r, p0) - c[p0], d1 = get3(r, p1) - c[p1];
        const double c1 = ((0.0 + d0 * d0 + d1 * d1) - rad * rad) / rad * 0.5;
        const double rv = get3(r, S.axis) - c[S.axis];
        const double c2 = fabs(rv) - c[4];
        double proj, cc;
        if (c1 >= 2.0 * rad * c2) {
            proj = 0.0 + get3(u, p0) * d0 + get3(u, p1) * d1;
            cc = c1;
        } else {
            proj = get3(u, S.axis) * rv;
            cc = c2;
        }
        return (proj == 0.0) ? cc >= 0.0 : proj > 0.0;
    }
    }
}

// simpleCell_class.f90 inside / distance over the record [n, +-surfIdx ...] at ints[rec]
// unionCell_class.f90 insideComplex over the record's tokens (shortCircuit skips to the group's
// closing bracket; the host checked that brackets balance)
__device__ __noinline__ int union_inside_complex(const DevProblem& P, const int* tok, int m, const double r[3],
                                                 const double u[3])
{
    int isIt = 1, depth = 0;
    for (int i = 0; i < m; ++i) {
        const int el = tok[i];
        if (el < SCN_UCELL_UNION) {
            const int hs = surf_halfspace(P, P.surfs[abs(el) - 1], r, u);
            isIt = hs == (el > 0);
        } else if ((el == SCN_UCELL_UNION && isIt) || (el == SCN_UCELL_INTERSECT && !isIt)) {
            if (depth == 0) return isIt;
            depth -= 1;
            int d = 1;
            while (d > 0 && i + 1 < m) {
                ++i;
                if (tok[i] == SCN_UCELL_OPEN) d++;
                else if (tok[i] == SCN_UCELL_CLOSE) d--;
            }
        } else if (el == SCN_UCELL_OPEN) {
            depth += 1;
        } else if (el == SCN_UCELL_CLOSE) {
            depth -= 1;
        }
    }
    return isIt;
}

__device__ __forceinline__ int cell_inside(const DevProblem& P, int rec, const double r[3], const double u[3])
{
    const int* c = P.ints + rec;
    const int n = c[0];
    if (n == SCN_UCELL) {
        if (!c[1]) return union_inside_complex(P, c + 3, c[2], r, u);
        int isIt = 0;
        for (int i = 0; i < c[2]; ++i) {
            const int s = c[3 + i];
            if (s >= SCN_UCELL_UNION) continue;
            const int hs = surf_halfspace(P, P.surfs[abs(s) - 1], r, u);
            isIt = hs == (s > 0);
            if (!isIt) return 0;
        }
        return isIt;
    }
    if (n <= 0) return 0;
    for (int i = 1; i <= n; ++i) {
        const int s = c[i];
        const int hs = surf_halfspace(P, P.surfs[abs(s) - 1], r, u);
        if (hs != (s > 0)) return 0;
    }
    return 1;
}

__device__ __forceinline__ double cell_distance(const DevProblem& P, int rec, const double r[3], const double u[3],
                                                int* surfIdx)
{
    const int* c = P.ints + rec;
    double d = SCN_INF;
    int si = 0;
    if (c[0] == SCN_UCELL) {   // unionCell: every surface of the definition, in order
        for (int i = 0; i < c[2]; ++i) {
            const int s = c[3 + i];
            if (s >= SCN_UCELL_UNION) continue;
            const double t = surf_distance(P, P.surfs[abs(s) - 1], r, u);
            if (t < d) {
                d = t;
                si = abs(s);
            }
        }
        *surfIdx = si;
        return d;
    }
    for (int i = 1; i <= c[0]; ++i) {
        const double t = surf_distance(P, P.surfs[abs(c[i]) - 1], r, u);
        if (t < d) {
            d = t;
            si = abs(c[i]);
        }
    }
    *surfIdx = si;
    return d;
}

// a non-box root border
__device__ __forceinline__ int border_halfspace(const DevProblem& P, const scn_surface& S, const double r[3],
                                             const double u[3])
{
    return surf_halfspace(P, S, r, u);
}
__device__ __forceinline__ double border_distance(const DevProblem& P, const scn_surface& S, const double r[3],
                                               const double u[3])
{
    return surf_distance(P, S, r, u);
}

// matmul(rotMat, v) with rotMat row-major (universe_inter.f90 enter, coord_class.f90 rotate)
__device__ __forceinline__ void rot_apply(const double* R, double v[3])
{
    double t[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] = 0.0 + R[3 * i] * v[0] + R[3 * i + 1] * v[1] + R[3 * i + 2] * v[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = t[i];
}

// ---------------------------------------------------------------------------------------------
// universes
// ---------------------------------------------------------------------------------------------
// latUniverse get_ijk (latUniverse_class.f90:489-506), its integer divisions by scn_idiv
__device__ __forceinline__ void get_ijk(int localID, const int32_t sizeN[3], const double* inv, int ijk[3])
{
    int temp = localID - 1;
    int base = scn_idiv(temp, inv[3]);
    ijk[0] = temp - sizeN[0] * base + 1;
    temp = base;
    base = scn_idiv(temp, inv[4]);
    ijk[1] = temp - sizeN[1] * base + 1;
    ijk[2] = base + 1;
}

// latUniverse findCell (latUniverse_class.f90:280-312)
__device__ __forceinline__ int lat_findcell(const DevProblem& P, const scn_universe& U, int uniIdx, const double r[3],
                                            const double u[3])
{
    const double* inv = P.uni_inv + 6 * (size_t)(uniIdx - 1);
    int ijk[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ijk[i] = (int)scn_floor_div(r[i] - U.corner[i], U.pitch[i], inv[i]) + 1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double r_bar = r[i] - U.corner[i] - (double)ijk[i] * U.pitch[i] + 0.5 * U.pitch[i];
        if (fabs(r_bar) > U.a_bar[i] && r_bar * u[i] > 0.0) ijk[i] += (u[i] < 0.0) ? -1 : 1;
    }
    int out = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) out |= (ijk[i] <= 0 || ijk[i] > U.size[i]);
    if (out) return U.out_local;
    return ijk[0] + U.size[0] * (ijk[1] - 1 + U.size[1] * (ijk[2] - 1));
}

template <int GEO>
__device__ __forceinline__ int uni_findcell(const DevProblem& P, int uniIdx, const double r[3], const double u[3])
{
    const scn_universe& U = P.unis[uniIdx - 1];
    if (U.type == SCN_UNI_LAT) {
        return lat_findcell(P, U, uniIdx, r, u);
    } else if (U.type == SCN_UNI_PIN) {
        double r_sq = r[0] * r[0] + r[1] * r[1];
        double mul = (r[0] * u[0] + r[1] * u[1] >= 0.0) ? -1.0 : 1.0;
        const double* rad = P.real + U.annuli;
        const double* rsq = rad + U.n_annuli;
        int l;
        for (l = 1; l <= U.n_annuli; ++l)
            if (r_sq < rsq[l - 1] + mul * (2.0 * rad[l - 1] * SCN_SURF_TOL)) return l;
        return l;
    }
    if (GEO && U.type == SCN_UNI_CELL) {
        // cellUniverse_class.f90:203-250 findCell in definition order; 262-297 findCellOverlap
        const int* cl = P.ints + U.cells;
        if (U.check_overlap) {
            int found = 0, id = 0;
            for (int i = 1; i <= U.n_cells; ++i) {
                if (cell_inside(P, cl[i - 1], r, u)) {
                    ++found;
                    id = i;
                }
            }
            return (found == 0) ? U.n_cells + 1 : (found > 1) ? U.n_cells + 2 : id;
        }
        for (int i = 1; i <= U.n_cells; ++i)
            if (cell_inside(P, cl[i - 1], r, u)) return i;
        return U.n_cells + 1;
    }
    if (GEO && U.border_surf >= 0) return border_halfspace(P, P.surfs[U.border_surf], r, u) ? 2 : 1;
    return box_halfspace(P.boxes[U.border], r, u) ? 2 : 1;
}

template <int GEO>
__device__ __forceinline__ double uni_distance(const DevProblem& P, int uniIdx, int localID, const double r[3],
                                      const double u[3], int* surfIdx, DevStatus* st)
{
    const scn_universe& U = P.unis[uniIdx - 1];
    if (U.type == SCN_UNI_LAT) {
        if (localID == U.out_local) {
            scn_box outline;
            outline.n_axes = 3;
            outline.axes[0] = 0;
            outline.axes[1] = 1;
            outline.axes[2] = 2;
            for (int i = 0; i < 3; ++i) {
                outline.origin[i] = 0.0;
                outline.halfwidth[i] = U.outline_hw[i];
            }
            *surfIdx = LAT_OUTLINE;
            return box_distance(outline, r, u);
        }
        int ijk[3];
        get_ijk(localID, U.size, P.uni_inv + 6 * (size_t)(uniIdx - 1), ijk);
        double d = SCN_INF;
        int ax = 0;
        double uax = u[0];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double r_bar = r[i] - U.corner[i];
            r_bar = r_bar - ((double)ijk[i] - 0.5) * U.pitch[i];
            double bound = scn_sign(U.pitch[i] * 0.5, u[i]);
            double test_d = (bound - r_bar) / u[i];
            if (test_d < d) {
                d = test_d;
                ax = i;
                uax = u[i];
            }
        }
        d = scn_fmax(0.0, d);
        d = scn_fmin(SCN_INF, d);
        int s = (ax + 1) * 2;
        if (uax < 0.0) s = s - 1;
        *surfIdx = -s;
        return d;
    } else if (U.type == SCN_UNI_PIN) {
        const double* rad = P.real + U.annuli;
        const double* rsq = rad + U.n_annuli;
        int id = localID;
        double d_out = (id > U.n_annuli) ? SCN_INF : cyl_distance(rad[id - 1], rsq[id - 1], r, u);
        double d_in = (id == 1) ? SCN_INF : cyl_distance(rad[id - 2], rsq[id - 2], r, u);
        if (d_in < d_out) {
            *surfIdx = PIN_IN;
            return d_in;
        }
        *surfIdx = PIN_OUT;
        return d_out;
    }
    if (GEO && U.type == SCN_UNI_CELL) {   // cellUniverse_class.f90:359-378
        if (localID > U.n_cells) {   // UNDEF or OVERLAP local cell
            raise_error(st, SCN_ERR_MATERIAL, localID);
            *surfIdx = 0;
            return SCN_INF;
        }
        return cell_distance(P, P.ints[U.cells + localID - 1], r, u, surfIdx);
    }
    *surfIdx = SURF_BORDER;
    if (GEO && U.border_surf >= 0) return border_distance(P, P.surfs[U.border_surf], r, u);
    return box_distance(P.boxes[U.border], r, u);
}

__device__ __forceinline__ void uni_offset(const DevProblem& P, int uniIdx, int localID, double off[3])
{
    const scn_universe& U = P.unis[uniIdx - 1];
    off[0] = off[1] = off[2] = 0.0;
    if (U.type != SCN_UNI_LAT) return;
    int doOffset = (U.offset_mode == 2) ? (P.ints[U.offset_map + localID - 1] == 1) : (U.offset_mode == 1);
    if (doOffset && localID != U.out_local) {
        int ijk[3];
        get_ijk(localID, U.size, P.uni_inv + 6 * (size_t)(uniIdx - 1), ijk);
        for (int i = 0; i < 3; ++i) off[i] = ((double)ijk[i] - 0.5) * U.pitch[i] + U.corner[i];
    }
}

// ---------------------------------------------------------------------------------------------
// geometryStd
// ---------------------------------------------------------------------------------------------

// direction of level l from the one above it (coord_class.f90 rotate/assignDirection: a rotated
// universe applies its rotation matrix, every other level shares the direction above)
__device__ __forceinline__ void level_dir(const DevProblem& P, int uniIdx, double u[3])
{
    const scn_universe& U = P.unis[uniIdx - 1];
    if (U.rotated) rot_apply(U.rot, u);
}

// diveToMat from level `start` (0-based); levels are statically unrolled.
template <int NL, int GEO>
__device__ __forceinline__ void dive(const DevProblem& P, Pt<NL, GEO>& p, int start, DevStatus* st)
{
    double u[3] = {p.u[0], p.u[1], p.u[2]};   // direction at level l
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        if (GEO && l > 0 && l <= start && P.any_rot) level_dir(P, p.uni[l], u);
        if (l < start) continue;
        int pos = p.root[l] + p.loc[l] - 1;
        int fill = P.gfill[pos - 1];
        int id = P.gid[pos - 1];
        if (fill >= 0) {
            p.mat = fill;
            p.uid = id;
            return;
        }
        if (l + 1 < NL) {
            int child = -fill;
            double off[3];
            uni_offset(P, p.uni[l], p.loc[l], off);
            const scn_universe& Uc = P.unis[child - 1];
            p.nest = l + 2;
            p.uni[l + 1] = child;
            p.root[l + 1] = id;
            // universe%enter (universe_inter.f90:400-424): rotate, then translate; a 'global'
            // universe is entered from the global position
            double rr[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) rr[k] = (GEO && Uc.global) ? p.r[0][k] : p.r[l][k] - off[k];
            if (GEO && Uc.rotated) {
                rot_apply(Uc.rot, rr);
                rot_apply(Uc.rot, u);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) p.r[l + 1][k] = rr[k] - Uc.origin[k];
            p.loc[l + 1] = uni_findcell<GEO>(P, child, p.r[l + 1], u);
        }
    }
    raise_error(st, SCN_ERR_ARG, 1);
    p.mat = SCN_UNDEF_MAT;
}

// placeCoord in a flat lattice (flat_lattice in scn_host.h): the root's halfspace test, the
// lattice's findCell and, for a pin cell, the pin level's position, computed as place_coord and
// diveToMat compute them (the root has no offset; a one-region pin's findCell is always 1); the
// material, uniqueID and depth of each lattice cell come from the table built on the host.
template <int NL, int GEO>
__device__ __forceinline__ void place_flat(const DevProblem& P, Pt<NL, GEO>& p)
{
    const scn_universe& R = P.unis[P.root_uni - 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) p.r[0][k] = p.r[0][k] - R.origin[k];
    const int l0 = box_halfspace(P.boxes[R.border], p.r[0], p.u) ? 2 : 1;
    const int L = P.flat_lat;
    const scn_universe& U = P.unis[L - 1];
    double r1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) r1[k] = (p.r[0][k] - 0.0) - U.origin[k];
    // every level is written unconditionally (levels past the particle's depth are never read):
    // outside the root the lattice cell found is simply not used
    const int l1 = lat_findcell(P, U, L, r1, p.u);
    const int* c = P.flat + 2 + 5 * (l1 - 1);
    const int inside = l0 == 1;
    p.mat = inside ? c[0] : P.flat[0];
    p.uid = inside ? c[1] : P.flat[1];
    p.nest = inside ? c[2] : 1;
    constexpr int L1 = NL > 1 ? 1 : 0, L2 = NL > 2 ? 2 : 0;
    p.uni[0] = P.root_uni;
    p.root[0] = 1;
    p.loc[0] = l0;
    double off[3];
    uni_offset(P, L, l1, off);
    const int pin = c[3] > 0 ? c[3] : L;
    const scn_universe& Uc = P.unis[pin - 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p.r[L1][k] = r1[k];
        if (NL > 2) p.r[L2][k] = (r1[k] - off[k]) - Uc.origin[k];
    }
    p.uni[L1] = L;
    p.root[L1] = P.gid[0];
    p.loc[L1] = l1;
    if (NL > 2) {
        p.uni[L2] = c[3];
        p.root[L2] = c[4];
        p.loc[L2] = 1;
    }
}

template <int NL, int GEO>
__device__ __forceinline__ void place_coord(const DevProblem& P, Pt<NL, GEO>& p, DevStatus* st)
{
    if (SCN_FLAT_PLACEMENT && !GEO && NL >= 3 && P.flat_lat) {
        place_flat(P, p);
        return;
    }
    const scn_universe& R = P.unis[P.root_uni - 1];
    p.nest = 1;
    p.mat = SCN_UNDEF_MAT;
    p.uid = -3;
#pragma unroll
    for (int k = 0; k < 3; ++k) p.r[0][k] = p.r[0][k] - R.origin[k];
    p.uni[0] = P.root_uni;
    p.root[0] = 1;
    p.loc[0] = uni_findcell<GEO>(P, P.root_uni, p.r[0], p.u);
    dive(P, p, 0, st);
}

// rebuild uni/root of every level from the stored local IDs (geomGraph walk)
template <int NL, int GEO>
__device__ __forceinline__ void relink(const DevProblem& P, Pt<NL, GEO>& p)
{
    p.uni[0] = P.root_uni;
    p.root[0] = 1;
#pragma unroll
    for (int l = 0; l + 1 < NL; ++l) {
        if (l + 1 < p.nest) {
            int pos = p.root[l] + p.loc[l] - 1;
            p.uni[l + 1] = -P.gfill[pos - 1];
            p.root[l + 1] = P.gid[pos - 1];
        }
    }
}

template <int NL, int GEO>
__device__ __forceinline__ void move_local(const DevProblem& P, Pt<NL, GEO>& p, double d, int n)
{
    p.nest = n;
    double u[3] = {p.u[0], p.u[1], p.u[2]};
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        if (l < n) {
            if (GEO && l > 0 && P.any_rot) level_dir(P, p.uni[l], u);
#pragma unroll
            for (int k = 0; k < 3; ++k) p.r[l][k] = p.r[l][k] + d * u[k];
        }
    }
}

template <int NL, int GEO>
__device__ __forceinline__ void move_global(Pt<NL, GEO>& p, double d)
{
    p.nest = 1;
    p.mat = SCN_UNDEF_MAT;
    p.uid = -3;
#pragma unroll
    for (int k = 0; k < 3; ++k) p.r[0][k] = p.r[0][k] + d * p.u[k];
}

template <int NL>
struct DistCache {
    int lvl;
    double dist[NL];
    int surf[NL];
};

// geometryStd%move_withCache (cache on) / move_noCache (cache off); returns the event
template <int NL, int GEO>
__device__ __forceinline__ int move(const DevProblem& P, Pt<NL, GEO>& p, double& maxDist, DistCache<NL>& cache, DevStatus* st)
{
    double dist = SCN_INF;
    int surfIdx = 0, level = 0;
    double ul[3] = {p.u[0], p.u[1], p.u[2]};   // direction at level l
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        if (l < p.nest) {
            if (GEO && l > 0 && P.any_rot) level_dir(P, p.uni[l], ul);
            double td;
            int ts;
            if (P.st_cache) {
                if (cache.lvl < l + 1) {
                    cache.dist[l] = uni_distance<GEO>(P, p.uni[l], p.loc[l], p.r[l], ul, &cache.surf[l], st);
                    cache.lvl += 1;
                }
                td = cache.dist[l];
                ts = cache.surf[l];
            } else {
                td = uni_distance<GEO>(P, p.uni[l], p.loc[l], p.r[l], ul, &ts, st);
            }
            if ((dist - td) >= dist * SCN_FP_REL_TOL) {
                dist = td;
                surfIdx = ts;
                level = l + 1;
            }
        }
    }
    if (maxDist < dist && maxDist < SCN_INF) {
        move_local(P, p, maxDist, p.nest);
        cache.lvl = 0;
        return EV_COLL;
    } else if (surfIdx == SURF_BORDER && level == 1) {
        move_global(p, dist);
        maxDist = dist;
        cache.lvl = 0;
        if (P.border_box >= 0) box_explicit_bc(P.boxes[P.border_box], p.r[0], p.u);
        place_coord(P, p, st);
        return EV_BOUNDARY;
    }
    move_local(P, p, dist, level);
    maxDist = dist;
#pragma unroll
    for (int l = 0; l < NL; ++l)
        if (l < level - 1) cache.dist[l] = cache.dist[l] - dist;
    cache.lvl = level - 1;
    // universe%cross at `level`, then diveToMat(level).  The crossed level is a runtime value, so
    // its data are gathered with masks (an `if (l == level-1)` lets the compiler substitute the
    // runtime index into the register arrays, which would spill them to scratch).
    {
        const int lv = level - 1;
        double rl[3] = {0.0, 0.0, 0.0};
        double uc[3] = {p.u[0], p.u[1], p.u[2]};   // direction at the crossed level
        int ul = 0, ll = 0;
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            const uint64_t m = 0ull - (uint64_t)(l == lv);
            const uint32_t m32 = (uint32_t)m;
            if (GEO && l > 0 && l <= lv && P.any_rot) level_dir(P, p.uni[l], uc);
#pragma unroll
            for (int k = 0; k < 3; ++k) rl[k] = scn_bitsd((scn_dbits(rl[k]) & ~m) | (scn_dbits(p.r[l][k]) & m));
            ul = (int)(((uint32_t)ul & ~m32) | ((uint32_t)p.uni[l] & m32));
            ll = (int)(((uint32_t)ll & ~m32) | ((uint32_t)p.loc[l] & m32));
        }
        const scn_universe& U = P.unis[ul - 1];
        int nl;
        if (U.type == SCN_UNI_PIN) {
            nl = ll + ((surfIdx == PIN_IN) ? -1 : 1);
        } else if (GEO && U.type == SCN_UNI_CELL) {
            // cellUniverse_class.f90:388-436: NUDGE the local position across, then search
#pragma unroll
            for (int k = 0; k < 3; ++k) rl[k] = rl[k] + uc[k] * SCN_NUDGE;
            nl = uni_findcell<GEO>(P, ul, rl, uc);
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                const uint64_t m = 0ull - (uint64_t)(l == lv);
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    p.r[l][k] = scn_bitsd((scn_dbits(p.r[l][k]) & ~m) | (scn_dbits(rl[k]) & m));
            }
        } else {
            nl = uni_findcell<GEO>(P, ul, rl, uc);
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) p.loc[l] = (l == lv) ? nl : p.loc[l];
    }
    dive(P, p, level - 1, st);
    return EV_CROSS;
}

template <int NL, int GEO>
__device__ __forceinline__ void teleport(const DevProblem& P, Pt<NL, GEO>& p, double dist, DevStatus* st)
{
    move_global(p, dist);
    place_coord(P, p, st);
    if (p.mat == SCN_OUTSIDE_MAT) {
        if (P.border_box >= 0) box_transform_bc(P.boxes[P.border_box], p.r[0], p.u);
        place_coord(P, p, st);
    }
}

// ---------------------------------------------------------------------------------------------
// nuclear data and tallies
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ const double* xs_of(const DevProblem& P, int mat, int G)
{
    return P.xs + ((size_t)(mat - 1) * P.nG + (G - 1)) * 6;
}

__device__ __forceinline__ double track_mat_xs(const DevProblem& P, int mat, int G, DevStatus* st)
{
    if (mat == SCN_VOID_MAT) return 0.0;
    if (mat < 1 || mat > P.n_mat) {
        raise_error(st, SCN_ERR_MATERIAL, mat);
        return 1.0;
    }
    return xs_of(P, mat, G)[XS_TOTAL] + 0.0;
}

__device__ __forceinline__ double macro_get(const double* x, int MT)
{
    switch (MT) {
    case SCN_MT_TOTAL: return x[XS_TOTAL];
    case SCN_MT_DISAPPEARANCE: return x[XS_CAPTURE];
    case SCN_MT_ESCATTER: return 0.0;
    case SCN_MT_NONELASTIC: return x[XS_SCATTER] + x[XS_FISSION] + x[XS_CAPTURE];
    case SCN_MT_IESCATTER: return x[XS_SCATTER];
    case SCN_MT_ALLSCATTER: return 0.0 + x[XS_SCATTER];
    case SCN_MT_FISSION: return x[XS_FISSION];
    case SCN_MT_NUFISSION: return x[XS_NUFISSION];
    case SCN_MT_KAPPA: return x[XS_KAPPA];
    case SCN_MT_ABSORPTION: return x[XS_FISSION] + x[XS_CAPTURE];
    default: return 0.0;
    }
}

// coordList%cell: cellShelf index of the lowest-level cell (0 unless that level is a cell universe
// in one of its defined cells; always 0 in the lattice/pin kernels, which have no cell universes)
template <int NL, int GEO>
__device__ __forceinline__ int lowest_cell(const DevProblem& P, const Pt<NL, GEO>& p)
{
    if (!GEO) return 0;
    int u = p.uni[0], lc = p.loc[0];
#pragma unroll
    for (int l = 1; l < NL; ++l) {
        if (l == p.nest - 1) {
            u = p.uni[l];
            lc = p.loc[l];
        }
    }
    const scn_universe& U = P.unis[u - 1];
    if (U.type != SCN_UNI_CELL || lc < 1 || lc > U.n_cells) return 0;
    return P.ints[U.cells + U.n_cells + lc - 1];
}

// tally maps: spaceMap, materialMap, energyMap (multigroup: bin 0), cellMap
__device__ __forceinline__ int map1d(const DevProblem& P, const scn_map1d& m, const double* rg, int mat, int cell)
{
    if (m.type == SCN_MAP_ENERGY) return 0;
    if (m.type == SCN_MAP_CELL) return P.ints[m.data + ((cell < 0 || cell > m.axis) ? 0 : cell)];
    if (m.type == SCN_MAP_SPACE) {
        double v = get3(rg, m.axis);
        const double* g = P.real + m.data;
        int idx;
        if (m.grid_type == SCN_GRID_LIN) {
            idx = (int)floor((v - g[0]) / g[1]) + 1;
        } else {
            if (v < g[0] || v > g[m.nbins]) {
                idx = -1;
            } else {
                int lo = 0, hi = m.nbins;
                while (hi - lo > 1) {
                    int mid = (lo + hi) / 2;
                    if (v >= g[mid]) lo = mid;
                    else hi = mid;
                }
                idx = lo + 1;
            }
        }
        if (idx < 1 || idx >= m.nbins + 1) idx = 0;
        return idx;
    }
    int nm = P.n_mat;
    return (mat >= 0 && mat <= nm) ? P.ints[m.data + mat] : P.ints[m.data + nm + 1];
}

// score memory: bins [0, lds_n) go to a private LDS copy (flushed to BIN by the kernel); the
// rest go to the global BIN array through a one-entry per-lane cache: consecutive scores of a
// lane to the same bin are summed in a register and added with one atomic when the lane scores
// another bin (bins_flush at the end).  Only the summation order of BIN changes.
// A score evicted from the cache waits in the lane's LDS queue (qcap entries, when the kernel
// gives one) and the wave adds every lane's queue at once (bins_drain): a global atomic is
// performed beyond L2, and the in-order count of vector memory operations makes the next load
// wait for it, so one wait per drain replaces one in every iteration that evicts.
// With fix = 1 every score goes straight to its bin's four limb words (scn_fix_limbs, integer
// atomics), in LDS or global memory; the cache is not used, and the sums are order-independent.
constexpr int BINQ_STRIDE = 64;   // lanes of a wave: entry k of a lane is q[2k * 64] (bin), q[(2k + 1) * 64]
struct Bins {
    double* lds;
    double* glob;
    int64_t lds_n;
    int64_t c_addr;    // bin held in the lane's cache, -1 none
    double c_acc;
    double* q;         // the lane's queue of evicted scores in LDS (null: none)
    int qn, qcap;      // entries queued, capacity (wave-uniform)
    int fix;
};

__device__ __forceinline__ void fix_add(unsigned long long* w, double v)
{
    uint64_t l[4];
    scn_fix_limbs(scn_fix_of(v), l);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (l[k]) atomicAdd(w + k, (unsigned long long)l[k]);
}

// add the lane's queued scores to BIN
__device__ __forceinline__ void bins_drain(Bins& B)
{
    for (int k = 0; k < B.qn; ++k)
        unsafeAtomicAdd(&B.glob[__double_as_longlong(B.q[2 * k * BINQ_STRIDE])], B.q[(2 * k + 1) * BINQ_STRIDE]);
    B.qn = 0;
}

__device__ __forceinline__ void bins_flush(Bins& B)
{
    if (B.c_addr >= 0) unsafeAtomicAdd(&B.glob[B.c_addr], B.c_acc);
    B.c_addr = -1;
    B.c_acc = 0.0;
    bins_drain(B);
}

__device__ __forceinline__ void bin_evict(Bins& B, int64_t addr, double v)
{
    if (B.qn < B.qcap) {
        B.q[2 * B.qn * BINQ_STRIDE] = __longlong_as_double(addr);
        B.q[(2 * B.qn + 1) * BINQ_STRIDE] = v;
        ++B.qn;
    } else {
        unsafeAtomicAdd(&B.glob[addr], v);
    }
}

__device__ __forceinline__ void bin_add(Bins& B, int64_t addr, double v)
{
    if (B.fix) {
        fix_add((unsigned long long*)(addr < B.lds_n ? B.lds : B.glob) + 4 * addr, v);
        return;
    }
    if (addr < B.lds_n) {
        atomicAdd(&B.lds[addr], v);
    } else if (addr == B.c_addr) {
        B.c_acc = B.c_acc + v;
    } else {
        if (B.c_addr >= 0) bin_evict(B, B.c_addr, B.c_acc);
        B.c_addr = addr;
        B.c_acc = v;
    }
}

// tallyAdmin%reportInColl: keffImplicitClerk first, then the collision clerks
template <int NL, int GEO>
__device__ __forceinline__ int report_in_coll(const DevProblem& P, Pt<NL, GEO>& p, int isVirtual, Bins& B, DevStatus* st)
{
    int nsc = 0;
    int mat = p.mat;
    if (mat != SCN_VOID_MAT) {
        double flux = p.w / p.track;
        const double* x = xs_of(P, mat, p.G);
        double nuFiss = x[XS_NUFISSION];
        double absXS = x[XS_CAPTURE] + x[XS_FISSION];
        p.k[K_IMP_PROD] = p.k[K_IMP_PROD] + nuFiss * flux;
        p.k[K_IMP_ABS] = p.k[K_IMP_ABS] + absXS * flux;
    }
    if (P.n_clerks == 0) return 0;
    if (P.fc.on) {
        const double flux_v = p.w / p.track;   // flux of clerks that score virtual collisions
        int nsc = 0;
        // unrolled over the FAST_CLERKS slots: the rolled loop, inside the flights' divergent
        // branches, lost scores in the delta-tracking kernels (tests/test_batch_gpu.py, slab-DT)
#pragma unroll
        for (int c = 0; c < FAST_CLERKS; ++c) {
            if (c >= P.fc.n_clerks) break;
            const FastClerk& F = P.fc.c[c];
            if (!F.handle_virtual && isVirtual) continue;   // wave-uniform: no address load
            const int64_t a0 = P.fc.g_addr[(size_t)c * P.nG + (p.G - 1)];
            if (a0 < 0) continue;
            int bin = 1, inside = 1;
            for (int k = 0; k < F.n_maps; ++k) {
                const double v = get3(p.r[0], F.axis[k]);
                const int idx = (int)scn_floor_div(v - F.g0[k], F.step[k], F.step_inv[k]) + 1;
                if (idx < 1 || idx >= F.nbins[k] + 1) {
                    inside = 0;
                    break;
                }
                bin = bin + (idx - 1) * F.multi[k];
            }
            if (!inside) continue;
            const double flux = F.handle_virtual ? flux_v : p.w / track_mat_xs(P, mat, p.G, st);
            const int64_t addr = a0 + (int64_t)F.width * (bin - 1);
            for (int i = 0; i < F.width; ++i) {
                double resp;
                if (F.resp[i] == SCN_MT_FLUX) resp = 1.0;
                else resp = (mat == SCN_VOID_MAT) ? 0.0 : macro_get(xs_of(P, mat, p.G), F.resp[i]);
                bin_add(B, addr + i, resp * flux);
            }
            nsc += F.width;
        }
        return nsc;
    }
    const int cell = lowest_cell(P, p);
    int c0 = P.gclerk_off[p.G - 1], c1 = P.gclerk_off[p.G];   // collision clerks only
    for (int ci = c0; ci < c1; ++ci) {
        const scn_clerk& C = P.clerks[P.gclerk_idx[ci]];
        if (!C.handle_virtual && isVirtual) continue;
        int bin = 1;
        for (int k = 0; k < C.n_maps; ++k) {
            int b = map1d(P, C.maps[k], p.r[0], mat, cell);
            if (b == 0) {
                bin = 0;
                break;
            }
            bin = bin + (b - 1) * C.multi[k];
        }
        if (bin == 0) continue;
        double flux = C.handle_virtual ? p.w / p.track : p.w / track_mat_xs(P, mat, p.G, st);
        int64_t addr = C.mem_addr + (int64_t)C.width * (bin - 1);
        for (int i = 0; i < C.width; ++i) {
            double resp;
            if (C.resp[i] == SCN_MT_FLUX) resp = 1.0;
            else resp = (mat == SCN_VOID_MAT) ? 0.0 : macro_get(xs_of(P, mat, p.G), C.resp[i]);
            bin_add(B, addr + i, resp * flux);
        }
        nsc += C.width;
    }
    return nsc;
}

// tallyAdmin%reportPath to the trackClerks (trackClerk_class.f90 reportPath): a surface-tracking
// path of length L scores response(pre-path material, G) * w * L in the bin of the pre-path
// state (global position r0 and material mat0 before the move)
template <int NL, int GEO>
__device__ __forceinline__ int report_path(const DevProblem& P, const Pt<NL, GEO>& p, double L, const double r0[3],
                                           int mat0, int cell0, Bins& B)
{
    int nsc = 0;
    const int c0 = P.gtrack_off[p.G - 1], c1 = P.gtrack_off[p.G];
    for (int ci = c0; ci < c1; ++ci) {
        const scn_clerk& C = P.clerks[P.gtrack_idx[ci]];
        int bin = 1;
        for (int k = 0; k < C.n_maps; ++k) {
            const int b = map1d(P, C.maps[k], r0, mat0, cell0);
            if (b == 0) {
                bin = 0;
                break;
            }
            bin = bin + (b - 1) * C.multi[k];
        }
        if (bin == 0) continue;
        const int64_t addr = C.mem_addr + (int64_t)C.width * (bin - 1);
        for (int i = 0; i < C.width; ++i) {
            double resp;
            if (C.resp[i] == SCN_MT_FLUX) resp = 1.0;
            else resp = (mat0 == SCN_VOID_MAT) ? 0.0 : macro_get(xs_of(P, mat0, p.G), C.resp[i]);
            bin_add(B, addr + i, resp * p.w * L);
        }
        nsc += C.width;
    }
    return nsc;
}

// ---------------------------------------------------------------------------------------------
// flights
// ---------------------------------------------------------------------------------------------

// outcome of one tracking event (an ST segment or a DT tentative flight)
enum { STEP_DEAD = 0, STEP_COLLIDE = 1, STEP_MORE = 2 };

// one segment of surfaceTracking (transportOperatorST_class.f90:59-160): to the nearest of the
// sampled collision point and the closest surface; `cache` is the flight's distCache (fresh at the
// start of every flight)
template <int NL, int GEO>
__device__ __forceinline__ int st_step(const DevProblem& P, Pt<NL, GEO>& p, DistCache<NL>& cache, Bins& B,
                                       DevStatus* st, Cnt& cnt)
{
    double sigmaTrack, sigmaT, dist, invSigmaTrack;
    int mat = p.mat;
    // (sigmaTrack, 1 / sigmaTrack) of the material and group, tabled on the host with the same
    // IEEE operations (the void pair last); an index outside the table takes the checked path
    const bool tabled = mat == SCN_VOID_MAT || (mat >= 1 && mat <= P.n_mat);
    const size_t ti = !tabled ? 0 : (mat == SCN_VOID_MAT) ? (size_t)P.n_mat * P.nG
                                                          : (size_t)(mat - 1) * P.nG + (p.G - 1);
    const double* tt = P.st_xs + 2 * ti;
    if (tabled) sigmaTrack = tt[0];
    else sigmaTrack = scn_fmax(track_mat_xs(P, mat, p.G, st), P.collision_xs);
    p.track = sigmaTrack;
    if (sigmaTrack < SCN_ST_TOL) {
        dist = SCN_INF;
        invSigmaTrack = SCN_INF;
        sigmaT = 0.0;
    } else {
        invSigmaTrack = tabled ? tt[1] : 1.0 / sigmaTrack;
        dist = -scn_log(scn_rng_get(&p.rng)) * invSigmaTrack;
        sigmaT = track_mat_xs(P, mat, p.G, st);
        if (dist != dist) {
            raise_error(st, SCN_ERR_NAN, p.brood);
            p.dead = 1;
            return STEP_DEAD;
        }
    }
    double r0[3];   // savePrePath: the pre-path state of track clerks
    int mat0 = 0, cell0 = 0;
    if (GEO && P.n_track) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r0[k] = p.r[0][k];
        mat0 = p.mat;
        cell0 = lowest_cell(P, p);
    }
    int event = move(P, p, dist, cache, st);
    cnt.seg++;
    p.t = p.t + dist / 1.0;
    if (event == EV_COLL) p.fate = FATE_NO;
    if (GEO && P.n_track) cnt.sc += report_path(P, p, dist, r0, mat0, cell0, B);
    mat = p.mat;
    if (mat == SCN_OUTSIDE_MAT) {
        p.dead = 1;
        p.fate = FATE_LEAK;
        p.k[K_ANA_LEAK] = p.k[K_ANA_LEAK] + p.w;
        return STEP_DEAD;
    } else if (mat == SCN_UNDEF_MAT || mat == SCN_OVERLAP_MAT || st->code) {
        raise_error(st, SCN_ERR_MATERIAL, mat);
        p.dead = 1;
        return STEP_DEAD;
    }
    if (event == EV_COLL) {
        if (scn_rng_get(&p.rng) < sigmaT * invSigmaTrack) return STEP_COLLIDE;
        cnt.sc += report_in_coll(P, p, 1, B, st);
    }
    return STEP_MORE;
}

// one tentative flight of deltaTracking (transportOperatorDT_class.f90:47-130).  The majorant
// and its inverse depend on the group only, so evaluating them per tentative flight gives the
// values the flight computes once.
template <int NL, int GEO>
__device__ __forceinline__ int dt_step(const DevProblem& P, Pt<NL, GEO>& p, Bins& B, DevStatus* st, Cnt& cnt)
{
    // (majorant, 1 / majorant) of the group, evaluated on the host with the same IEEE operations
    const double maj = P.maj[2 * (p.G - 1)];
    p.track = maj;
    const double majorant_inv = P.maj[2 * (p.G - 1) + 1];
    if (fabs(majorant_inv) > SCN_HUGE) {
        raise_error(st, SCN_ERR_MAJORANT, p.G);
        p.dead = 1;
        return STEP_DEAD;
    }
    double distance = -scn_log(scn_rng_get(&p.rng)) * majorant_inv;
    teleport(P, p, distance, st);
    cnt.seg++;
    p.t = p.t + distance / 1.0;
    int mat = p.mat;
    if (mat == SCN_OUTSIDE_MAT) {
        p.fate = FATE_LEAK;
        p.dead = 1;
        p.k[K_ANA_LEAK] = p.k[K_ANA_LEAK] + p.w;
        return STEP_DEAD;
    } else if (mat == SCN_VOID_MAT) {
        cnt.sc += report_in_coll(P, p, 1, B, st);
        return STEP_MORE;
    } else if (mat == SCN_UNDEF_MAT || mat == SCN_OVERLAP_MAT || st->code) {
        raise_error(st, SCN_ERR_MATERIAL, mat);
        p.dead = 1;
        return STEP_DEAD;
    }
    double sigmaT = track_mat_xs(P, mat, p.G, st);
    if (scn_rng_get(&p.rng) < sigmaT * majorant_inv) return STEP_COLLIDE;
    cnt.sc += report_in_coll(P, p, 1, B, st);
    return STEP_MORE;
}

// transportOperatorHT tracking_selection (transportOperatorHT_class.f90:41-64): delta tracking for
// this flight when sigmaT * (1 / majorant) > 1 - cutoff (sigmaT = 0 in void)
template <int NL, int GEO>
__device__ __forceinline__ int ht_use_dt(const DevProblem& P, const Pt<NL, GEO>& p, DevStatus* st)
{
    const double majorant_inv = P.maj[2 * (p.G - 1) + 1];
    double sigmaT = (p.mat == SCN_VOID_MAT) ? 0.0 : track_mat_xs(P, p.mat, p.G, st);
    double ratio = sigmaT * majorant_inv;
    return ratio > (1.0 - P.ht_cutoff);
}

// surfaceTracking: segments until a real collision or a leak; returns 1 if alive
template <int NL, int GEO>
__device__ __forceinline__ int flight_st(const DevProblem& P, Pt<NL, GEO>& p, Bins& B, DevStatus* st, Cnt& cnt)
{
    DistCache<NL> cache;
    cache.lvl = 0;
    for (int seg = 0;; ++seg) {
        if (seg >= MAX_SEGMENTS_PER_FLIGHT) {
            raise_error(st, SCN_ERR_ARG, SCN_ERR_LOST_DETAIL);
            p.dead = 1;
            return 0;
        }
        int r = st_step(P, p, cache, B, st, cnt);
        if (r != STEP_MORE) return r == STEP_COLLIDE;
    }
}

// deltaTracking: tentative flights until a real collision or a leak; returns 1 if alive
template <int NL, int GEO>
__device__ __forceinline__ int flight_dt(const DevProblem& P, Pt<NL, GEO>& p, Bins& B, DevStatus* st, Cnt& cnt)
{
    for (int seg = 0;; ++seg) {
        if (seg >= MAX_SEGMENTS_PER_FLIGHT) {
            raise_error(st, SCN_ERR_ARG, SCN_ERR_LOST_DETAIL);
            p.dead = 1;
            return 0;
        }
        int r = dt_step(P, p, B, st, cnt);
        if (r != STEP_MORE) return r == STEP_COLLIDE;
    }
}

// fission-site sink.  Eigenvalue (stride == 0): structure-of-arrays append buffer, sorted by key
// afterwards.  Fixed source (stride > 0): every lane owns a LIFO buffer of `depth` entries (the
// package's per-thread `buffer`); entry j of lane l is element j*stride + l, and only r, u, w, t, G
// are kept (the rest of the released state is fixed by particle_fromParticleState).
struct Sites {
    double* r;        // [3][cap]
    double* u;        // [3][cap]
    double* w;
    double* t;
    int* G;
    int* mat;
    int* uid;
    int* coll;        // collisionN of the site (neutronMGimp keeps the parent's; 0 otherwise)
    uint64_t* key;    // broodID << 24 | emission index
    uint32_t* order;  // identity, sorted along with the keys
    int* count;
    int cap;
    int stride;       // lanes of the per-lane buffers (0: append buffer)
    int depth;        // capacity of one lane's buffer
};

// CDF scan of multiScatterMG%sampleGout / fissionMG%sampleOut: subtract x[0], x[1], ... from
// rem until it drops below zero; returns the 1-based index, or n + 1 if it never does.  The
// subtractions are the reference's sequence; the loads are issued four at a time (all in bounds)
// so the scan is not one memory latency per group.
__device__ __forceinline__ int cdf_scan(const double* __restrict__ x, int n, double rem)
{
    int g = 0;
    for (; g + 4 <= n; g += 4) {
        const double a0 = x[g], a1 = x[g + 1], a2 = x[g + 2], a3 = x[g + 3];
        rem = rem - a0;
        if (rem < 0.0) return g + 1;
        rem = rem - a1;
        if (rem < 0.0) return g + 2;
        rem = rem - a2;
        if (rem < 0.0) return g + 3;
        rem = rem - a3;
        if (rem < 0.0) return g + 4;
    }
    for (; g < n; ++g) {
        rem = rem - x[g];
        if (rem < 0.0) return g + 1;
    }
    return n + 1;
}

// legendrePoly_func.f90:35-99 sampleLegendre_P1: mu with density 1/2 (1 + c mu), c = 3 P1/P0, as a
// mixture of uniform / linear (|c| < 1) or linear / delta at 1 (1 <= |c| <= 3), mirrored for c < 0.
// Returns 0 if |c| > 3.
__device__ __forceinline__ int legendre_p1(double c, uint64_t* rng, double* mu)
{
    const double a = fabs(c);
    double threshold;
    int top, low;   // 1 uniform, 2 linear, 3 delta
    if (a < 1.0) {
        threshold = a;
        top = 2;
        low = 1;
    } else if (a <= 3.0) {
        threshold = 0.5 * (a - 1.0);
        top = 3;
        low = 2;
    } else {
        return 0;
    }
    const int ex = (scn_rng_get(rng) < threshold) ? top : low;
    double x;
    if (ex == 1) x = 2.0 * scn_rng_get(rng) - 1.0;
    else if (ex == 2) x = 2.0 * sqrt(scn_rng_get(rng)) - 1.0;
    else x = 1.0;
    if (c < 0.0) x = -x;
    *mu = x;
    return 1;
}

// collisionProcessor%collide with neutronMGstd; returns 1 if the particle survives
template <int NL, int GEO>
__device__ __forceinline__ int collide(const DevProblem& P, Pt<NL, GEO>& p, double k_eff, Bins& B, const Sites& S,
                              DevStatus* st, Cnt& cnt)
{
    const int nG = P.nG;
    int mat = p.mat;
    cnt.col++;
    (void)track_mat_xs(P, mat, p.G, st);   // its material check
    // the alpha-reaction draw (collisionProcessor_inter.f90 collide): probAlpha = 0 / sigmaT is
    // +-0 or NaN in multigroup data, which no draw in [0, 1) falls below; the draw is still made
    (void)scn_rng_get(&p.rng);
    const double* x = xs_of(P, mat, p.G);
    double rr = scn_rng_get(&p.rng);
    int C = 1;
    double xs = x[XS_TOTAL] * rr - 0.0;
    if (xs > 0.0) C++;
    xs = xs - x[XS_SCATTER];
    if (xs > 0.0) C++;
    xs = xs - x[XS_CAPTURE];
    if (xs > 0.0) C++;
    cnt.sc += report_in_coll(P, p, 0, B, st);
    double preColl_w = p.w;
    if (P.fissile[mat - 1]) {
        double wgt = p.w, w0 = p.w0;
        double rand1 = scn_rng_get(&p.rng);
        double sig_tot = x[XS_TOTAL], sig_nuFiss = x[XS_NUFISSION];
        double nd_ = fabs((wgt * sig_nuFiss) / (w0 * sig_tot * k_eff)) + rand1;
        if (!(nd_ < (double)S.cap)) {
            raise_error(st, SCN_ERR_OVERFLOW, -1);
            p.dead = 1;
            return 0;
        }
        int n = (int)nd_;
        if (n >= 1) {
            wgt = scn_sign(w0, wgt);
            const double* chi = P.chi + (size_t)(mat - 1) * nG;
            int base;
            if (S.stride > 0) {
                // detain into the history's buffer (particleDungeon_class.f90:159-176)
                if (p.sp + n > S.depth) {
                    raise_error(st, SCN_ERR_OVERFLOW, -2);
                    p.dead = 1;
                    return 0;
                }
                base = p.sp;
                p.sp += n;
            } else {
                base = atomicAdd(S.count, n);
            }
            for (int i = 0; i < n; ++i) {
                double mu = 2.0 * scn_rng_get(&p.rng) - 1.0;
                double phi = SCN_TWO_PI * scn_rng_get(&p.rng);
                double rem = scn_rng_get(&p.rng);
                int G_out = cdf_scan(chi, nG, rem);
                if (G_out > nG) {
                    raise_error(st, SCN_ERR_SAMPLING, 2);
                    p.dead = 1;
                    return 0;
                }
                double d[3];
                scn_rotate(p.u, mu, phi, d);
                if (S.stride > 0) {
                    const size_t e = (size_t)(base + i) * S.stride + p.lane;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        S.r[(size_t)k * S.cap + e] = p.r[0][k];
                        S.u[(size_t)k * S.cap + e] = d[k];
                    }
                    S.w[e] = wgt;
                    S.t[e] = p.t;
                    S.G[e] = G_out;
                    continue;
                }
                int slot = base + i;
                if (slot < S.cap) {
                    for (int k = 0; k < 3; ++k) {
                        S.r[(size_t)k * S.cap + slot] = p.r[0][k];
                        S.u[(size_t)k * S.cap + slot] = d[k];
                    }
                    S.w[slot] = wgt;
                    S.t[slot] = p.t;
                    S.G[slot] = G_out;
                    S.mat[slot] = mat;
                    S.uid[slot] = p.uid;
                    S.coll[slot] = (P.collision == SCN_COLL_IMP) ? p.coll : 0;
                    S.key[slot] = ((uint64_t)p.brood << 24) | (uint64_t)p.nsite;
                    S.order[slot] = (uint32_t)slot;
                } else {
                    raise_error(st, SCN_ERR_OVERFLOW, slot);
                }
                p.nsite++;
            }
        }
    }
    if (C == 2) {
        double rem = scn_rng_get(&p.rng) * x[XS_SCATTER];
        const double* P0 = P.P0 + ((size_t)(mat - 1) * nG + (p.G - 1)) * nG;
        const double* prod = P.prod + ((size_t)(mat - 1) * nG + (p.G - 1)) * nG;
        int G_out = cdf_scan(P0, nG, rem);
        if (G_out > nG) {
            raise_error(st, SCN_ERR_SAMPLING, 1);
            p.dead = 1;
            return 0;
        }
        double mu;
        if (P.P1) {
            // multiScatterP1MG_class.f90:69-85 sampleOut -> legendrePoly_func.f90:35-99
            const double c = P.P1[((size_t)(mat - 1) * nG + (p.G - 1)) * nG + (G_out - 1)];
            if (!legendre_p1(c, &p.rng, &mu)) {
                raise_error(st, SCN_ERR_SAMPLING, 3);
                p.dead = 1;
                return 0;
            }
        } else {
            mu = 2.0 * scn_rng_get(&p.rng) - 1.0;
        }
        double phi = SCN_TWO_PI * scn_rng_get(&p.rng);
        double w_mul = prod[G_out - 1];
        p.G = G_out;
        p.w = p.w * w_mul;
        double nd[3];
        scn_rotate(p.u, mu, phi, nd);
        p.u[0] = nd[0];
        p.u[1] = nd[1];
        p.u[2] = nd[2];
    } else if (C == 3 || C == 4) {
        p.dead = 1;
    } else {
        raise_error(st, SCN_ERR_UNSUPPORTED, -3);
        p.dead = 1;
        return 0;
    }
    if (GEO && P.ww_n > 0 && !p.dead) {
        // neutronMGimp%cutoffs with weight windows (neutronMGimp_class.f90 split / russianRoulette,
        // weightWindowsField_class.f90 atP): the window of the particle's state bin
        const scn_clerk& M = *P.ww_map;
        int bin = 1;
        for (int k = 0; k < M.n_maps; ++k) {
            const int b = map1d(P, M.maps[k], p.r[0], p.mat, lowest_cell(P, p));
            if (b == 0) {
                bin = 0;
                break;
            }
            bin = bin + (b - 1) * M.multi[k];
        }
        double minW = 0.0, maxW = 0.0, avW = 0.0;
        if (bin > 0) {
            minW = P.ww_lower[bin - 1];
            maxW = P.ww_upper[bin - 1];
            avW = minW * P.ww_survival;
        }
        if (p.w > maxW && maxW != 0.0 && p.split < P.max_split) {
            int mult = (int)ceil(p.w / maxW);
            if (mult + p.split > P.max_split) mult = P.max_split - p.split + 1;
            const double wc = p.w / mult;
            if (S.stride == 0 || p.sp + (mult - 1) > S.depth) {   // thisCycle%detain beyond the buffer
                raise_error(st, S.stride == 0 ? SCN_ERR_UNSUPPORTED : SCN_ERR_OVERFLOW, -4);
                p.dead = 1;
                return 0;
            }
            for (int i = 0; i < mult - 1; ++i) {
                const size_t e = (size_t)(p.sp++) * S.stride + p.lane;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    S.r[(size_t)k * S.cap + e] = p.r[0][k];
                    S.u[(size_t)k * S.cap + e] = p.u[k];
                }
                S.w[e] = wc;
                S.t[e] = p.t;
                S.G[e] = p.G;
            }
            p.split = p.split + mult;
            p.w = p.w / mult;
        } else if (p.w < minW) {
            if (scn_rng_get(&p.rng) < (1.0 - p.w / avW)) p.dead = 1;
            else p.w = avW;
        }
    }
    p.coll += 1;
    if (C == 2) {
        double score = scn_fmax(p.w - preColl_w, 0.0);
        if (score > 0.0) p.k[K_SCATTER_PROD] = p.k[K_SCATTER_PROD] + score;
    }
    if (p.dead) {
        p.fate = FATE_ABS;
        return 0;
    }
    return 1;
}

}  // namespace scn
