/*
 * scn_oracle.c -- CPU restatement of SCONE's multigroup tracking loop (TEST INFRASTRUCTURE).
 *
 * History-based and scalar like the reference: one particle at a time walks
 * transport -> collide -> transport ... with the tallyAdmin report order of the Fortran.
 * Only the tests, the smoke check and bench.py's CPU-baseline leg use this file, as the checker.
 * Every routine names the reference routine it restates (paths relative to the SCONE tree).
 */
#include "scn_oracle.h"
#include "scn_fpmath.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- error handling (errors_mod.f90 fatalError -> error code + message) -------------------- */
static int g_err = 0;
static char g_msg[512];

static void o_fail(int code, const char* fmt, ...)
{
#pragma omp critical(orc_err)
    {
        if (!g_err) {
            va_list ap;
            va_start(ap, fmt);
            vsnprintf(g_msg, sizeof g_msg, fmt, ap);
            va_end(ap);
            g_err = code;
        }
    }
}
const char* orc_last_error(void) { return g_msg; }
int orc_error_code(void) { return g_err; }
static void o_clear(void) { g_err = 0; g_msg[0] = 0; }

/* ---- shared math exported for tests --------------------------------------------------------- */
double orc_rng_get(uint64_t* s) { return scn_rng_get(s); }
uint64_t orc_rng_skip(uint64_t s, int64_t k) { return scn_rng_skip(s, k); }
double orc_log(double x) { return scn_log(x); }
void orc_sincos(double x, double* s, double* c) { scn_sincos(x, s, c); }
double orc_norm2(const double v[3]) { return scn_norm2(v); }
void orc_rotate(const double dir[3], double mu, double phi, double out[3]) { scn_rotate(dir, mu, phi, out); }

/* canonical cross-history sum: chunks of SCN_DSUM_CHUNK summed left to right, then pairwise */
double orc_dsum(const double* x, int64_t n)
{
    if (n <= 0) return 0.0;
    int64_t m = (n + SCN_DSUM_CHUNK - 1) / SCN_DSUM_CHUNK;
    double* y = (double*)malloc(sizeof(double) * (size_t)m);
    for (int64_t c = 0; c < m; ++c) {
        double s = 0.0;
        int64_t e = (c + 1) * SCN_DSUM_CHUNK < n ? (c + 1) * SCN_DSUM_CHUNK : n;
        for (int64_t i = c * SCN_DSUM_CHUNK; i < e; ++i) s = s + x[i];
        y[c] = s;
    }
    while (m > 1) {
        int64_t h = (m + 1) / 2;
        for (int64_t i = 0; i < h; ++i) y[i] = (2 * i + 1 < m) ? y[2 * i] + y[2 * i + 1] : y[2 * i];
        m = h;
    }
    double r = y[0];
    free(y);
    return r;
}

/* ======================================================================================
 * Surfaces
 * ====================================================================================== */

/* box_class.f90:134-144 evaluate (squareCylinder_class.f90 same over its two plane axes) */
double orc_box_evaluate(const scn_box* b, const double r[3])
{
    double c = -SCN_HUGE;
    for (int i = 0; i < b->n_axes; ++i) {
        double t = fabs(r[b->axes[i]] - b->origin[i]) - b->halfwidth[i];
        if (t > c) c = t;
    }
    return c;
}

/* box_class.f90:168-237 distance */
double orc_box_distance(const scn_box* b, const double r[3], const double u[3])
{
    double far = SCN_HUGE, near = -SCN_HUGE;
    for (int i = 0; i < b->n_axes; ++i) {
        int ax = b->axes[i];
        double rb = r[ax] - b->origin[i];
        double a_far = scn_sign(b->halfwidth[i], u[ax]);
        double a_near = -a_far;
        double tn, tf;
        if (u[ax] != 0.0) {
            tn = (a_near - rb) / u[ax];
            tf = (a_far - rb) / u[ax];
        } else {
            tn = scn_sign(SCN_INF, a_near - rb);
            tf = scn_sign(SCN_INF, a_far - rb);
            if (tn > tf) { double t = tn; tn = tf; tf = t; }
        }
        far = scn_fmin(far, tf);
        near = scn_fmax(near, tn);
    }
    double d;
    if (far <= near * SCN_MISS_TOL) {
        d = SCN_INF;
    } else if (fabs(orc_box_evaluate(b, r)) < SCN_SURF_TOL) {
        d = (fabs(far) >= fabs(near)) ? far : near;
    } else {
        d = (near <= 0.0) ? far : near;
    }
    if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
    return d;
}

/* box_class.f90:252-278 going */
static int o_box_going(const scn_box* b, const double r[3], const double u[3])
{
    int maxCom = 0;
    double best = 0.0;
    for (int i = 0; i < b->n_axes; ++i) {
        double t = fabs(r[b->axes[i]] - b->origin[i]) - b->halfwidth[i];
        if (i == 0 || t > best) { best = t; maxCom = i; }
    }
    double rl = r[b->axes[maxCom]] - b->origin[maxCom];
    double proj = u[b->axes[maxCom]] * scn_sign(1.0, rl);
    int hs = proj > 0.0;
    if (proj == 0.0) hs = orc_box_evaluate(b, r) >= 0.0;
    return hs;
}

/* surface_inter.f90:363-378 halfspace */
int orc_box_halfspace(const scn_box* b, const double r[3], const double u[3])
{
    double c = orc_box_evaluate(b, r);
    int hs = c > 0.0;
    if (fabs(c) < SCN_SURF_TOL) hs = o_box_going(b, r, u);
    return hs;
}

/* box_class.f90:380-420 explicitBC */
void orc_box_explicit_bc(const scn_box* b, double r[3], double u[3])
{
    for (int i = 0; i < b->n_axes; ++i) {
        int ax = b->axes[i];
        double r0 = r[ax] - b->origin[i];
        if (fabs(r0) <= b->halfwidth[i] * (1.0 - SCN_SURF_TOL)) continue;
        int bc = (r0 < 0.0) ? b->bc[2 * ax] : b->bc[2 * ax + 1];
        if (bc == SCN_BC_REFLECTIVE) u[ax] = -u[ax];
        else if (bc == SCN_BC_PERIODIC) r[ax] = r[ax] - 2.0 * scn_sign(b->halfwidth[i], r0);
    }
}

/* box_class.f90:431-487 transformBC */
void orc_box_transform_bc(const scn_box* b, double r[3], double u[3])
{
    int Ri[3];
    for (int i = 0; i < b->n_axes; ++i) {
        double a_bar = b->halfwidth[i] * (1.0 - SCN_SURF_TOL);
        Ri[i] = (int)ceil(fabs(r[b->axes[i]] - b->origin[i]) / a_bar) / 2;
    }
    for (int i = 0; i < b->n_axes; ++i) {
        int ax = b->axes[i];
        for (int t = 1; t <= Ri[i]; ++t) {
            double r0 = r[ax] - b->origin[i];
            int bc = (r0 < 0.0) ? b->bc[2 * ax] : b->bc[2 * ax + 1];
            if (bc == SCN_BC_REFLECTIVE) {
                double a0 = scn_sign(b->halfwidth[i], r0) + b->origin[i];
                double d = r[ax] - a0;
                r[ax] = r[ax] - 2.0 * d;
                u[ax] = -u[ax];
            } else if (bc == SCN_BC_PERIODIC) {
                double d = scn_sign(b->halfwidth[i], r0);
                r[ax] = r[ax] - 2.0 * d;
            }
        }
    }
}

/* cylinder_class.f90:209-270 distance, z-axis cylinder centred at the origin */
static double o_cyl_distance(double radius, double r_sq, const double r[3], const double u[3])
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
        if (k >= 0.0) d = SCN_INF;
        else { d = -k + sqrt(delta); d = d / a; }
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
double orc_cyl_distance(double radius, const double r[3], const double u[3])
{
    return o_cyl_distance(radius, radius * radius, r, u);
}

/* ---- surfaces of cells and of a non-box root border ---------------------------------------
 * Each type restates its class in Geometry/Surfaces; dot_product and sum are evaluated as a
 * running sum from zero in index order, x**2 as x*x. */

/* the two coordinates across an axis (cylinder_class.f90 build: plane) */
static void o_plane_of(int axis, int p[2])
{
    p[0] = (axis == 0) ? 1 : 0;
    p[1] = (axis == 2) ? 1 : 2;
}

static double o_dot3(const double a[3], const double b[3])
{
    double s = 0.0;
    for (int i = 0; i < 3; ++i) s = s + a[i] * b[i];
    return s;
}

/* quadric_class.f90 normal: gradient / norm2(gradient) */
static void o_quadric_normal(const double* a, const double r[3], double n[3])
{
    n[0] = 2.0 * a[0] * r[0] + a[3] * r[1] + a[5] * r[2] + a[6];
    n[1] = 2.0 * a[1] * r[1] + a[3] * r[0] + a[4] * r[2] + a[7];
    n[2] = 2.0 * a[2] * r[2] + a[4] * r[1] + a[5] * r[0] + a[8];
    double nn = scn_norm2(n);
    for (int i = 0; i < 3; ++i) n[i] = n[i] / nn;
}

#define O_EPS 2.220446049250313e-16   /* epsilon(ONE) */

/* truncCone_class.f90:159-170 evaluate; c = vertex[3], tan^2, hMin, hMax */
static double o_cone_evaluate(const scn_surface* S, const double r[3])
{
    const double* c = S->c;
    int p[2];
    o_plane_of(S->axis, p);
    int a = S->axis;
    double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]], da = r[a] - c[a];
    double v = (0.0 + d0 * d0 + d1 * d1) - c[3] * (da * da);
    double cMin = -da + c[4];
    double cMax = da - c[5];
    return scn_fmax(v, scn_fmax(cMin, cMax));
}

/* truncCone_class.f90:183-262 distance */
static double o_cone_distance(const scn_surface* S, const double r[3], const double u[3])
{
    const double* c = S->c;
    int p[2];
    o_plane_of(S->axis, p);
    int ax = S->axis;
    double diff[3], orient[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i) diff[i] = r[i] - c[i];
    orient[ax] = scn_sign(1.0, c[4]);
    double tsq = c[3];
    double a = (0.0 + u[p[0]] * u[p[0]] + u[p[1]] * u[p[1]]) - tsq * (u[ax] * u[ax]);
    double b = (0.0 + diff[p[0]] * u[p[0]] + diff[p[1]] * u[p[1]]) - tsq * diff[ax] * u[ax];
    double cc = (0.0 + diff[p[0]] * diff[p[0]] + diff[p[1]] * diff[p[1]]) - tsq * (diff[ax] * diff[ax]);
    double delta = b * b - a * cc;
    double d1, d2, t[3];
    if (fabs(a) < O_EPS) {
        d1 = -(0.5 * cc / b);
        d2 = scn_sign(SCN_INF, o_dot3(u, orient));
        for (int i = 0; i < 3; ++i) t[i] = diff[i] + d1 * u[i];
        if (!(o_dot3(t, orient) > 0.0)) d1 = d2;
    } else if (delta < O_EPS) {
        d1 = -SCN_INF;
        d2 = SCN_INF;
    } else {
        d1 = -(b + sqrt(delta)) / a;
        d2 = -(b - sqrt(delta)) / a;
        for (int i = 0; i < 3; ++i) t[i] = diff[i] + d1 * u[i];
        if (!(o_dot3(t, orient) > 0.0)) d1 = scn_sign(SCN_INF, o_dot3(u, orient));
        for (int i = 0; i < 3; ++i) t[i] = diff[i] + d2 * u[i];
        if (!(o_dot3(t, orient) > 0.0)) d2 = scn_sign(SCN_INF, o_dot3(u, orient));
    }
    double near = scn_fmin(d1, d2), far = scn_fmax(d1, d2);
    if (fabs(u[ax]) > O_EPS) {
        d1 = (c[5] - diff[ax]) / u[ax];
        d2 = (c[4] - diff[ax]) / u[ax];
    } else if (diff[ax] > c[4] && diff[ax] < c[5]) {
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
    else if (fabs(o_cone_evaluate(S, r)) < S->tol) dist = (fabs(far) >= fabs(near)) ? far : near;
    else dist = (near <= 0.0) ? far : near;
    if (dist <= 0.0 || dist > SCN_INF) dist = SCN_INF;
    return dist;
}

/* truncCone_class.f90:272-316 going through normal (normalised by norm2) */
static int o_cone_going(const scn_surface* S, const double r[3], const double u[3])
{
    const double* c = S->c;
    int p[2];
    o_plane_of(S->axis, p);
    int ax = S->axis;
    double n[3] = {0.0, 0.0, 0.0};
    double da = r[ax] - c[ax];
    if (fabs(da - c[4]) < S->tol) {
        n[ax] = -1.0;
    } else if (fabs(da - c[5]) < S->tol) {
        n[ax] = 1.0;
    } else {
        n[p[0]] = r[p[0]] - c[p[0]];
        n[p[1]] = r[p[1]] - c[p[1]];
        n[ax] = -c[3] * da;
    }
    double nn = scn_norm2(n);
    for (int i = 0; i < 3; ++i) n[i] = n[i] / nn;
    double proj = o_dot3(n, u);
    if (proj == 0.0) return o_cone_evaluate(S, r) >= 0.0;
    return proj > 0.0;
}

/* wedge_class.f90:195-210 evaluate; c = origin[3], halfwidth, altitude, n1[2], n2[2], n3[2] */
static double o_wedge_evaluate(const scn_surface* S, const double r[3])
{
    const double* c = S->c;
    int p[2];
    o_plane_of(S->axis, p);
    int ax = S->axis;
    double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]];
    double cAxis = fabs(r[ax] - c[ax]) - c[3];
    double cSide1 = 0.0 + c[5] * d0 + c[6] * d1;
    double cSide2 = 0.0 + c[7] * d0 + c[8] * d1;
    double cBase = (0.0 + c[9] * d0 + c[10] * d1) - c[4];
    return scn_fmax(scn_fmax(scn_fmax(cAxis, cBase), cSide1), cSide2);
}

/* one side plane of the wedge distance: k = u . n, cc = diff . n (minus the altitude for the base) */
static void o_wedge_side(double k, double cc, double* near, double* far)
{
    if (fabs(k) < O_EPS) {
        if (cc > 0.0) {
            *near = SCN_INF;
            *far = -SCN_INF;
        }
    } else {
        double d = -cc / k;
        if (k > 0.0) *far = scn_fmin(*far, d);
        else *near = scn_fmax(*near, d);
    }
}

/* wedge_class.f90:222-320 distance */
static double o_wedge_distance(const scn_surface* S, const double r[3], const double u[3])
{
    const double* c = S->c;
    int p[2];
    o_plane_of(S->axis, p);
    int ax = S->axis;
    double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]], da = r[ax] - c[ax];
    double near = -SCN_INF, far = SCN_INF;
    o_wedge_side(0.0 + u[p[0]] * c[5] + u[p[1]] * c[6], 0.0 + d0 * c[5] + d1 * c[6], &near, &far);
    o_wedge_side(0.0 + u[p[0]] * c[7] + u[p[1]] * c[8], 0.0 + d0 * c[7] + d1 * c[8], &near, &far);
    {   /* base: c > altitude is outside; d = -(c - altitude) / k */
        double k = 0.0 + u[p[0]] * c[9] + u[p[1]] * c[10];
        double cc = 0.0 + d0 * c[9] + d1 * c[10];
        if (fabs(k) < O_EPS) {
            if (cc > c[4]) {
                near = SCN_INF;
                far = -SCN_INF;
            }
        } else {
            double d = -(cc - c[4]) / k;
            if (k > 0.0) far = scn_fmin(far, d);
            else near = scn_fmax(near, d);
        }
    }
    double tn, tf;
    if (fabs(u[ax]) > O_EPS) {
        double e1 = (c[3] - da) / u[ax];
        double e2 = (-c[3] - da) / u[ax];
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
    double dist;
    if (far <= near * SCN_MISS_TOL) dist = SCN_INF;
    else if (fabs(o_wedge_evaluate(S, r)) < S->tol) dist = (fabs(far) >= fabs(near)) ? far : near;
    else if (o_wedge_evaluate(S, r) < 0.0) dist = far;
    else dist = near;
    if (dist <= 0.0 || dist > SCN_INF) dist = SCN_INF;
    return dist;
}

/* wedge_class.f90:330-380 going through normal (normalised by norm2) */
static int o_wedge_going(const scn_surface* S, const double r[3], const double u[3])
{
    const double* c = S->c;
    int p[2];
    o_plane_of(S->axis, p);
    int ax = S->axis;
    double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]], da = r[ax] - c[ax];
    double cSide1 = 0.0 + c[5] * d0 + c[6] * d1;
    double cSide2 = 0.0 + c[7] * d0 + c[8] * d1;
    double n[3] = {0.0, 0.0, 0.0};
    if (fabs(da - c[3]) < S->tol) {
        n[ax] = 1.0;
    } else if (fabs(da + c[3]) < S->tol) {
        n[ax] = -1.0;
    } else if (fabs(cSide1) < S->tol) {
        n[p[0]] = c[5];
        n[p[1]] = c[6];
    } else if (fabs(cSide2) < S->tol) {
        n[p[0]] = c[7];
        n[p[1]] = c[8];
    } else {
        n[p[0]] = c[9];
        n[p[1]] = c[10];
    }
    double nn = scn_norm2(n);
    for (int i = 0; i < 3; ++i) n[i] = n[i] / nn;
    double proj = o_dot3(n, u);
    if (fabs(proj) < O_EPS) return o_wedge_evaluate(S, r) >= 0.0;
    return proj > 0.0;
}

/* surface%evaluate of every type */
double orc_surf_evaluate(const scn_problem* P, const scn_surface* S, const double r[3])
{
    const double* c = S->c;
    switch (S->type) {
    case SCN_SURF_APLANE: /* aPlane_class.f90 evaluate */
        return r[S->axis] - c[0];
    case SCN_SURF_PLANE: /* plane_class.f90 evaluate */
        return o_dot3(r, c) - c[3];
    case SCN_SURF_SPHERE: { /* sphere_class.f90 evaluate */
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = r[i] - c[i];
        return o_dot3(d, d) - c[4];
    }
    case SCN_SURF_CYLINDER: { /* cylinder_class.f90 evaluate */
        int p[2];
        o_plane_of(S->axis, p);
        double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]];
        return (0.0 + d0 * d0 + d1 * d1) - c[4];
    }
    case SCN_SURF_QUADRIC: { /* quadric_class.f90 evaluate */
        double x = r[0], y = r[1], z = r[2];
        return x * (c[0] * x + c[3] * y + c[5] * z + c[6]) + y * (c[1] * y + c[4] * z + c[7]) +
               z * (c[2] * z + c[8]) + c[9];
    }
    case SCN_SURF_BOX:
        return orc_box_evaluate(&P->boxes[S->axis], r);
    case SCN_SURF_TRUNC_CYLINDER: { /* truncCylinder_class.f90 evaluate */
        int p[2];
        o_plane_of(S->axis, p);
        int a = S->axis;
        double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]];
        double v = ((0.0 + d0 * d0 + d1 * d1) - c[3] * c[3]) / c[3] * 0.5;
        return scn_fmax(v, fabs(r[a] - c[a]) - c[4]);
    }
    case SCN_SURF_TRUNC_CONE:
        return o_cone_evaluate(S, r);
    case SCN_SURF_WEDGE:
        return o_wedge_evaluate(S, r);
    }
    o_fail(SCN_ERR_ARG, "unknown surface type %d", S->type);
    return 0.0;
}

/* the quadratic branch shared by cylinder and quadric distance (a: coefficient of d^2) */
static double o_quad_distance(double a, double k, double c, double tol)
{
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

/* surface%distance of every type */
double orc_surf_distance(const scn_problem* P, const scn_surface* S, const double r[3], const double u[3])
{
    const double* c = S->c;
    switch (S->type) {
    case SCN_SURF_APLANE: { /* aPlane_class.f90 distance */
        double ra = c[0] - r[S->axis];
        double ua = u[S->axis];
        double d;
        if (fabs(ra) < S->tol) d = SCN_INF;
        else if (ua != 0.0) d = ra / ua;
        else d = SCN_INF;
        if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
        return d;
    }
    case SCN_SURF_PLANE: { /* plane_class.f90 distance */
        double k = o_dot3(u, c);
        double cc = o_dot3(r, c) - c[3];
        double d;
        if (k == 0.0 || fabs(cc) < S->tol) {
            d = SCN_INF;
        } else {
            d = -cc / k;
            if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
        }
        return d;
    }
    case SCN_SURF_SPHERE: { /* sphere_class.f90 distance (no cap at INF) */
        double dr[3];
        for (int i = 0; i < 3; ++i) dr[i] = r[i] - c[i];
        double cc = o_dot3(dr, dr) - c[4];
        double k = o_dot3(dr, u);
        double delta = k * k - cc;
        double d;
        if (delta < 0.0) {
            d = SCN_INF;
        } else if (fabs(cc) < S->tol) {
            d = (k >= 0.0) ? SCN_INF : -k + sqrt(delta);
        } else if (cc < 0.0) {
            d = -k + sqrt(delta);
        } else {
            d = -k - sqrt(delta);
            if (d <= 0.0) d = SCN_INF;
        }
        return d;
    }
    case SCN_SURF_CYLINDER: { /* cylinder_class.f90 distance */
        int p[2];
        o_plane_of(S->axis, p);
        double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]];
        double cc = (0.0 + d0 * d0 + d1 * d1) - c[4];
        double k = 0.0 + d0 * u[p[0]] + d1 * u[p[1]];
        double ua = u[S->axis];
        double a = 1.0 - ua * ua;
        return o_quad_distance(a, k, cc, S->tol);
    }
    case SCN_SURF_QUADRIC: { /* quadric_class.f90 distance */
        double x = r[0], y = r[1], z = r[2], u1 = u[0], u2 = u[1], u3 = u[2];
        double cc = orc_surf_evaluate(P, S, r);
        double k = c[0] * u1 * x + c[1] * u2 * y + c[2] * u3 * z +
                   0.5 * (c[3] * (u1 * y + u2 * x) + c[4] * (u2 * z + u3 * y) + c[5] * (u3 * x + u1 * z) +
                          c[6] * u1 + c[7] * u2 + c[8] * u3);
        double a = u1 * (c[0] * u1 + c[3] * u2 + c[5] * u3) + u2 * (c[1] * u2 + c[4] * u3) + c[2] * u3 * u3;
        return o_quad_distance(a, k, cc, S->tol);
    }
    case SCN_SURF_BOX:
        return orc_box_distance(&P->boxes[S->axis], r, u);
    case SCN_SURF_TRUNC_CYLINDER: { /* truncCylinder_class.f90 distance */
        int p[2];
        o_plane_of(S->axis, p);
        int ax = S->axis;
        double rad = c[3], hw = c[4];
        double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]];
        double c1 = (0.0 + d0 * d0 + d1 * d1) - rad * rad;
        double k = 0.0 + d0 * u[p[0]] + d1 * u[p[1]];
        double a = 1.0 - u[ax] * u[ax];
        double delta = k * k - a * c1;
        double far, near;
        if (delta <= 0.0 || a == 0.0) {
            far = SCN_INF;
            near = scn_sign(SCN_INF, c1);
        } else {
            far = (-k + sqrt(delta)) / a;
            near = (-k - sqrt(delta)) / a;
            if (far < near) { double t = far; far = near; near = t; }
        }
        double rb = r[ax] - c[ax];
        double tn, tf;
        if (u[ax] != 0.0) {
            tn = (-hw - rb) / u[ax];
            tf = (hw - rb) / u[ax];
        } else {
            tn = scn_sign(SCN_INF, -hw - rb);
            tf = scn_sign(SCN_INF, hw - rb);
        }
        if (tf < tn) { double t = tf; tf = tn; tn = t; }
        far = scn_fmin(far, tf);
        near = scn_fmax(near, tn);
        double cc = scn_fmax(c1 / rad * 0.5, fabs(rb) - hw);
        double d;
        if (far <= near * SCN_MISS_TOL) d = SCN_INF;
        else if (fabs(cc) < S->tol) d = (fabs(far) >= fabs(near)) ? far : near;
        else d = (near <= 0.0) ? far : near;
        if (d <= 0.0 || d > SCN_INF) d = SCN_INF;
        return d;
    }
    case SCN_SURF_TRUNC_CONE:
        return o_cone_distance(S, r, u);
    case SCN_SURF_WEDGE:
        return o_wedge_distance(S, r, u);
    }
    o_fail(SCN_ERR_ARG, "unknown surface type %d", S->type);
    return SCN_INF;
}

/* surface%going: the halfspace a point on the surface moves into */
static int o_surf_going(const scn_problem* P, const scn_surface* S, const double r[3], const double u[3])
{
    const double* c = S->c;
    switch (S->type) {
    case SCN_SURF_APLANE: { /* aPlane_class.f90 going */
        double ua = u[S->axis];
        if (ua == 0.0) return (r[S->axis] - c[0]) >= 0.0;
        return ua > 0.0;
    }
    case SCN_SURF_PLANE: { /* plane_class.f90 going */
        double proj = o_dot3(u, c);
        if (proj == 0.0) return (o_dot3(r, c) - c[3]) >= 0.0;
        return proj > 0.0;
    }
    case SCN_SURF_SPHERE: { /* sphere_class.f90 going */
        double dr[3];
        for (int i = 0; i < 3; ++i) dr[i] = r[i] - c[i];
        return o_dot3(dr, u) >= 0.0;
    }
    case SCN_SURF_CYLINDER: { /* cylinder_class.f90 going */
        int p[2];
        o_plane_of(S->axis, p);
        double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]];
        return (0.0 + d0 * u[p[0]] + d1 * u[p[1]]) >= 0.0;
    }
    case SCN_SURF_QUADRIC: { /* quadric_class.f90 going */
        double n[3];
        o_quadric_normal(c, r, n);
        return o_dot3(n, u) >= 0.0;
    }
    case SCN_SURF_TRUNC_CYLINDER: { /* truncCylinder_class.f90 going */
        int p[2];
        o_plane_of(S->axis, p);
        int a = S->axis;
        double rad = c[3];
        double d0 = r[p[0]] - c[p[0]], d1 = r[p[1]] - c[p[1]];
        double c1 = ((0.0 + d0 * d0 + d1 * d1) - rad * rad) / rad * 0.5;
        double rv = r[a] - c[a];
        double c2 = fabs(rv) - c[4];
        double proj, cc;
        if (c1 >= 2.0 * rad * c2) {
            proj = 0.0 + u[p[0]] * d0 + u[p[1]] * d1;
            cc = c1;
        } else {
            proj = u[a] * rv;
            cc = c2;
        }
        if (proj == 0.0) return cc >= 0.0;
        return proj > 0.0;
    }
    case SCN_SURF_TRUNC_CONE:
        return o_cone_going(S, r, u);
    case SCN_SURF_WEDGE:
        return o_wedge_going(S, r, u);
    }
    o_fail(SCN_ERR_ARG, "unknown surface type %d", S->type);
    return 0;
}

/* surface_inter.f90:363-378 halfspace (box types keep their own, box_class.f90) */
int orc_surf_halfspace(const scn_problem* P, const scn_surface* S, const double r[3], const double u[3])
{
    if (S->type == SCN_SURF_BOX) return orc_box_halfspace(&P->boxes[S->axis], r, u);
    double c = orc_surf_evaluate(P, S, r);
    int hs = c > 0.0;
    if (fabs(c) < S->tol) hs = o_surf_going(P, S, r, u);
    return hs;
}

/* ---- cells: simpleCell_class.f90 (record [n, +-surfIdx ...] in int_data) ------------------ */

/* unionCell_class.f90 insideComplex: walk the tokens; an operator that decides its group (a union
 * after a true operand, an intersection after a false one) returns at the top level and skips to
 * the group's closing bracket inside one (shortCircuit) */
static int o_union_inside_complex(const scn_problem* P, const int32_t* tok, int m, const double r[3],
                                  const double u[3])
{
    int isIt = 1, depth = 0;
    for (int i = 0; i < m; ++i) {
        int el = tok[i];
        if (el < SCN_UCELL_UNION) {
            int hs = orc_surf_halfspace(P, &P->surfaces[abs(el) - 1], r, u);
            isIt = (hs != 0) == (el > 0);
        } else if ((el == SCN_UCELL_UNION && isIt) || (el == SCN_UCELL_INTERSECT && !isIt)) {
            if (depth == 0) return isIt;
            depth -= 1;
            int d = 1;
            while (d > 0) {
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

/* inside: every surface halfspace agrees with the cell's sense (simpleCell, and a unionCell
 * without unions); a unionCell with unions evaluates its expression */
static int o_cell_inside(const scn_problem* P, int rec, const double r[3], const double u[3])
{
    const int32_t* c = P->int_data + rec;
    if (c[0] == SCN_UCELL) {
        const int32_t* tok = c + 3;
        if (!c[1]) return o_union_inside_complex(P, tok, c[2], r, u);
        int isIt = 0;
        for (int i = 0; i < c[2]; ++i) {
            int s = tok[i];
            if (s >= SCN_UCELL_UNION) continue;
            int hs = orc_surf_halfspace(P, &P->surfaces[abs(s) - 1], r, u);
            isIt = (hs != 0) == (s > 0);
            if (!isIt) return 0;
        }
        return isIt;
    }
    int isIt = 0;
    for (int i = 1; i <= c[0]; ++i) {
        int s = c[i];
        int sense = s > 0;
        int hs = orc_surf_halfspace(P, &P->surfaces[abs(s) - 1], r, u);
        isIt = (hs != 0) == sense;
        if (!isIt) return 0;
    }
    return isIt;
}

/* distance: the nearest surface of the cell (the first of equal distances) */
static double o_cell_distance(const scn_problem* P, int rec, const double r[3], const double u[3], int32_t* surfIdx)
{
    const int32_t* c = P->int_data + rec;
    double d = SCN_INF;
    *surfIdx = 0;
    if (c[0] == SCN_UCELL) {   /* unionCell: every surface of the definition, in order */
        for (int i = 0; i < c[2]; ++i) {
            int s = c[3 + i];
            if (s >= SCN_UCELL_UNION) continue;
            double t = orc_surf_distance(P, &P->surfaces[abs(s) - 1], r, u);
            if (t < d) {
                d = t;
                *surfIdx = abs(s);
            }
        }
        return d;
    }
    for (int i = 1; i <= c[0]; ++i) {
        double t = orc_surf_distance(P, &P->surfaces[abs(c[i]) - 1], r, u);
        if (t < d) {
            d = t;
            *surfIdx = abs(c[i]);
        }
    }
    return d;
}

/* ======================================================================================
 * Universes
 * ====================================================================================== */
enum { SURF_BORDER = 1 };                                   /* rootUniverse: border surface    */
enum { PIN_MOVING_IN = -1, PIN_MOVING_OUT = -2 };           /* pinUniverse_class.f90           */
enum { LAT_OUTLINE = -7 };                                  /* latUniverse_class.f90 OUTLINE_SURF */

typedef struct o_coord {
    double r[3];
    double dir[3];
    int uniIdx, uniRootID, localID, cellIdx;
} o_coord;

typedef struct o_coords {
    int nesting;
    o_coord lvl[SCN_MAX_NEST];
    int matIdx;
    int uniqueID;
} o_coords;

static const scn_universe* o_uni(const scn_problem* P, int idx) { return &P->universes[idx - 1]; }

/* latUniverse_class.f90:489-506 get_ijk */
static void o_get_ijk(int localID, const int32_t sizeN[3], int ijk[3])
{
    int temp = localID - 1;
    int base = temp / sizeN[0];
    ijk[0] = temp - sizeN[0] * base + 1;
    temp = base;
    base = temp / sizeN[1];
    ijk[1] = temp - sizeN[1] * base + 1;
    ijk[2] = base + 1;
}

/* findCell of every universe type; cellIdx is 0 for root/lattice/pin */
int orc_uni_findcell(const scn_problem* P, int32_t uniIdx, const double r[3], const double u[3])
{
    const scn_universe* U = o_uni(P, uniIdx);
    switch (U->type) {
    case SCN_UNI_ROOT: /* rootUniverse_class.f90:127-145 */
        if (U->border_surf >= 0) return orc_surf_halfspace(P, &P->surfaces[U->border_surf], r, u) ? 2 : 1;
        return orc_box_halfspace(&P->boxes[U->border], r, u) ? 2 : 1;
    case SCN_UNI_LAT: { /* latUniverse_class.f90:270-312 */
        int ijk[3];
        for (int i = 0; i < 3; ++i) ijk[i] = (int)floor((r[i] - U->corner[i]) / U->pitch[i]) + 1;
        for (int i = 0; i < 3; ++i) {
            double r_bar = r[i] - U->corner[i] - (double)ijk[i] * U->pitch[i] + 0.5 * U->pitch[i];
            if (fabs(r_bar) > U->a_bar[i] && r_bar * u[i] > 0.0) ijk[i] += (u[i] < 0.0) ? -1 : 1;
        }
        for (int i = 0; i < 3; ++i)
            if (ijk[i] <= 0 || ijk[i] > U->size[i]) return U->out_local;
        return ijk[0] + U->size[0] * (ijk[1] - 1 + U->size[1] * (ijk[2] - 1));
    }
    case SCN_UNI_PIN: { /* pinUniverse_class.f90:150-175 */
        double r_sq = r[0] * r[0] + r[1] * r[1];
        double mul = (r[0] * u[0] + r[1] * u[1] >= 0.0) ? -1.0 : 1.0;
        const double* rad = P->real_data + U->annuli;
        const double* rsq = rad + U->n_annuli;
        int l;
        for (l = 1; l <= U->n_annuli; ++l)
            if (r_sq < rsq[l - 1] + mul * (2.0 * rad[l - 1] * SCN_SURF_TOL)) return l;
        return l;
    }
    case SCN_UNI_CELL: { /* cellUniverse_class.f90:203-250 findCell, 262-297 findCellOverlap */
        const int32_t* cl = P->int_data + U->cells;
        if (U->check_overlap) {
            int found = 0, id = 0;
            for (int i = 1; i <= U->n_cells; ++i)
                if (o_cell_inside(P, cl[i - 1], r, u)) { ++found; id = i; }
            if (found == 0) return U->n_cells + 1;
            if (found > 1) return U->n_cells + 2;
            return id;
        }
        for (int i = 1; i <= U->n_cells; ++i)
            if (o_cell_inside(P, cl[i - 1], r, u)) return i;
        return U->n_cells + 1;
    }
    }
    o_fail(SCN_ERR_ARG, "findCell: unknown universe type %d", U->type);
    return 1;
}

double orc_uni_distance(const scn_problem* P, int32_t uniIdx, int32_t localID, const double r[3],
                        const double u[3], int32_t* surfIdx)
{
    const scn_universe* U = o_uni(P, uniIdx);
    switch (U->type) {
    case SCN_UNI_ROOT: /* rootUniverse_class.f90:152-162 */
        *surfIdx = SURF_BORDER;
        if (U->border_surf >= 0) return orc_surf_distance(P, &P->surfaces[U->border_surf], r, u);
        return orc_box_distance(&P->boxes[U->border], r, u);
    case SCN_UNI_LAT: { /* latUniverse_class.f90:319-373 */
        if (localID == U->out_local) {
            scn_box outline = {3, {0, 1, 2}, {0, 0, 0, 0, 0, 0}, {0.0, 0.0, 0.0},
                               {U->outline_hw[0], U->outline_hw[1], U->outline_hw[2]}};
            *surfIdx = LAT_OUTLINE;
            return orc_box_distance(&outline, r, u);
        }
        int ijk[3];
        o_get_ijk(localID, U->size, ijk);
        double d = SCN_INF;
        int ax = 0;
        for (int i = 0; i < 3; ++i) {
            double r_bar = r[i] - U->corner[i];
            r_bar = r_bar - ((double)ijk[i] - 0.5) * U->pitch[i];
            double bound = scn_sign(U->pitch[i] * 0.5, u[i]);
            double test_d = (bound - r_bar) / u[i];
            if (test_d < d) { d = test_d; ax = i; }
        }
        d = scn_fmax(0.0, d);
        d = scn_fmin(SCN_INF, d);
        int s = (ax + 1) * 2;
        if (u[ax] < 0.0) s = s - 1;
        *surfIdx = -s;
        return d;
    }
    case SCN_UNI_PIN: { /* pinUniverse_class.f90:185-226 */
        const double* rad = P->real_data + U->annuli;
        const double* rsq = rad + U->n_annuli;
        int id = localID;
        double d_out, d_in;
        if (id < 1 || id > U->n_annuli + 1) { o_fail(SCN_ERR_ARG, "pin distance: invalid local ID %d", id); return SCN_INF; }
        d_out = (id > U->n_annuli) ? SCN_INF : o_cyl_distance(rad[id - 1], rsq[id - 1], r, u);
        d_in = (id == 1) ? SCN_INF : o_cyl_distance(rad[id - 2], rsq[id - 2], r, u);
        if (d_in < d_out) { *surfIdx = PIN_MOVING_IN; return d_in; }
        *surfIdx = PIN_MOVING_OUT;
        return d_out;
    }
    case SCN_UNI_CELL: { /* cellUniverse_class.f90:359-378 */
        if (localID == U->n_cells + 1) {
            o_fail(SCN_ERR_MATERIAL, "Particle is in undefined local cell. Local ID: %d", localID);
            *surfIdx = 0;
            return SCN_INF;
        }
        if (localID == U->n_cells + 2) {
            o_fail(SCN_ERR_MATERIAL, "Particle is in an overlapping local cell. Local ID: %d", localID);
            *surfIdx = 0;
            return SCN_INF;
        }
        return o_cell_distance(P, P->int_data[U->cells + localID - 1], r, u, surfIdx);
    }
    }
    o_fail(SCN_ERR_ARG, "distance: unknown universe type %d", U->type);
    return SCN_INF;
}

/* universe%cross: root and lattice re-run findCell, pin steps the annulus index */
static void o_uni_cross(const scn_problem* P, o_coord* c, int surfIdx)
{
    const scn_universe* U = o_uni(P, c->uniIdx);
    if (U->type == SCN_UNI_PIN) { /* pinUniverse_class.f90:236-253 */
        if (surfIdx == PIN_MOVING_IN) c->localID -= 1;
        else if (surfIdx == PIN_MOVING_OUT) c->localID += 1;
        else o_fail(SCN_ERR_ARG, "pin cross: unknown surface memento %d", surfIdx);
    } else if (U->type == SCN_UNI_CELL) {
        /* cellUniverse_class.f90:388-436: NUDGE the local position over the surface, then search
         * (the neighbour lists only speed the search up: cells do not overlap) */
        for (int i = 0; i < 3; ++i) c->r[i] = c->r[i] + c->dir[i] * SCN_NUDGE;
        c->localID = orc_uni_findcell(P, c->uniIdx, c->r, c->dir);
        c->cellIdx = 0;
    } else {
        c->localID = orc_uni_findcell(P, c->uniIdx, c->r, c->dir);
        c->cellIdx = 0;
    }
}

/* universe%cellOffset: lattice cell centre (latUniverse_class.f90:389-414), zero otherwise */
static void o_uni_offset(const scn_problem* P, const o_coord* c, double off[3])
{
    const scn_universe* U = o_uni(P, c->uniIdx);
    off[0] = off[1] = off[2] = 0.0;
    if (U->type != SCN_UNI_LAT) return;
    int doOffset = (U->offset_mode == 2) ? (P->int_data[U->offset_map + c->localID - 1] == 1)
                                         : (U->offset_mode == 1);
    if (doOffset && c->localID != U->out_local) {
        int ijk[3];
        o_get_ijk(c->localID, U->size, ijk);
        for (int i = 0; i < 3; ++i) off[i] = ((double)ijk[i] - 0.5) * U->pitch[i] + U->corner[i];
    }
}

/* matmul(rotMat, v) (rotMat row-major): each component a running sum from zero */
static void o_matvec(const double R[9], const double v[3], double out[3])
{
    double t[3];
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s = s + R[3 * i + j] * v[j];
        t[i] = s;
    }
    for (int i = 0; i < 3; ++i) out[i] = t[i];
}

/* universe_inter.f90:400-424 enter: rotate position and direction, then translate */
static void o_uni_enter(const scn_problem* P, int uniIdx, o_coord* nw, const double r[3], const double u[3])
{
    const scn_universe* U = o_uni(P, uniIdx);
    for (int i = 0; i < 3; ++i) {
        nw->r[i] = r[i];
        nw->dir[i] = u[i];
    }
    nw->uniIdx = uniIdx;
    if (U->rotated) {
        o_matvec(U->rot, nw->r, nw->r);
        o_matvec(U->rot, nw->dir, nw->dir);
    }
    for (int i = 0; i < 3; ++i) nw->r[i] = nw->r[i] - U->origin[i];
    nw->localID = orc_uni_findcell(P, uniIdx, nw->r, nw->dir);
    nw->cellIdx = 0;
}

/* ======================================================================================
 * geometryStd (Geometry/geometryStd_class.f90)
 * ====================================================================================== */

/* coord_class.f90 takeAboveGeom / moveGlobal / moveLocal */
static void o_take_above(o_coords* c) { c->nesting = 1; c->matIdx = SCN_UNDEF_MAT; c->uniqueID = -3; }
static void o_move_global(o_coords* c, double d)
{
    o_take_above(c);
    for (int i = 0; i < 3; ++i) c->lvl[0].r[i] = c->lvl[0].r[i] + d * c->lvl[0].dir[i];
}
static void o_move_local(o_coords* c, double d, int n)
{
    c->nesting = n;
    for (int l = 0; l < n; ++l)
        for (int i = 0; i < 3; ++i) c->lvl[l].r[i] = c->lvl[l].r[i] + d * c->lvl[l].dir[i];
}

/* coord_class.f90 rotate / assignDirection: new global direction, lower levels through their
 * universes' rotation matrices */
static void o_assign_direction(const scn_problem* P, o_coords* c, const double u[3])
{
    for (int k = 0; k < 3; ++k) c->lvl[0].dir[k] = u[k];
    for (int l = 1; l < c->nesting; ++l) {
        const scn_universe* U = o_uni(P, c->lvl[l].uniIdx);
        if (U->rotated) o_matvec(U->rot, c->lvl[l - 1].dir, c->lvl[l].dir);
        else for (int k = 0; k < 3; ++k) c->lvl[l].dir[k] = c->lvl[l - 1].dir[k];
    }
}

/* geometryStd_class.f90:565-619 diveToMat (level is 1-based) */
static void o_dive(const scn_problem* P, o_coords* c, int start)
{
    for (int i = start; i <= SCN_MAX_NEST; ++i) {
        o_coord* L = &c->lvl[i - 1];
        int pos = L->uniRootID + L->localID - 1;
        if (pos < 1 || pos > P->graph_size) { o_fail(SCN_ERR_ARG, "diveToMat: graph index %d", pos); return; }
        int fill = P->graph_fill[pos - 1];
        int id = P->graph_id[pos - 1];
        if (fill >= 0) {
            c->matIdx = fill;
            c->uniqueID = id;
            return;
        }
        if (i == SCN_MAX_NEST) break;
        double off[3], r[3];
        o_uni_offset(P, L, off);
        if (o_uni(P, -fill)->global) /* the nested universe reverts to the global frame */
            for (int k = 0; k < 3; ++k) r[k] = c->lvl[0].r[k];
        else
            for (int k = 0; k < 3; ++k) r[k] = L->r[k] - off[k];
        c->nesting += 1;
        o_uni_enter(P, -fill, &c->lvl[i], r, L->dir);
        c->lvl[i].uniRootID = id;
    }
    o_fail(SCN_ERR_ARG, "diveToMat: failed to find material cell");
}

/* geometryStd_class.f90:119-147 placeCoord */
static void o_place_coord(const scn_problem* P, o_coords* c)
{
    double r[3], u[3];
    o_take_above(c);
    for (int i = 0; i < 3; ++i) { r[i] = c->lvl[0].r[i]; u[i] = c->lvl[0].dir[i]; }
    o_uni_enter(P, P->root_uni, &c->lvl[0], r, u);
    c->lvl[0].uniRootID = 1;
    o_dive(P, c, 1);
}

typedef struct o_cache { int lvl; double dist[SCN_MAX_NEST]; int surf[SCN_MAX_NEST]; } o_cache;

/* geometryStd_class.f90:633-717 closestDist / closestDist_cache (tie favours the outer level) */
static void o_closest(const scn_problem* P, const o_coords* c, o_cache* cache, int use_cache,
                      double* dist, int* surfIdx, int* lvl)
{
    *dist = SCN_INF;
    *surfIdx = 0;
    *lvl = 0;
    for (int l = 1; l <= c->nesting; ++l) {
        double td;
        int ts;
        if (use_cache) {
            if (cache->lvl < l) {
                cache->dist[l - 1] = orc_uni_distance(P, c->lvl[l - 1].uniIdx, c->lvl[l - 1].localID,
                                                      c->lvl[l - 1].r, c->lvl[l - 1].dir, &cache->surf[l - 1]);
                cache->lvl += 1;
            }
            td = cache->dist[l - 1];
            ts = cache->surf[l - 1];
        } else {
            td = orc_uni_distance(P, c->lvl[l - 1].uniIdx, c->lvl[l - 1].localID, c->lvl[l - 1].r,
                                  c->lvl[l - 1].dir, &ts);
        }
        if ((*dist - td) >= *dist * SCN_FP_REL_TOL) {
            *dist = td;
            *surfIdx = ts;
            *lvl = l;
        }
    }
}

enum { EV_COLL = 1, EV_BOUNDARY = 2, EV_CROSS = 3 };

/* geometryStd_class.f90:214-352 move_withCache / move_noCache (no fields: fieldDist = INF) */
static int o_move(const scn_problem* P, o_coords* c, double* maxDist, o_cache* cache, int use_cache)
{
    double dist;
    int surfIdx, level;
    o_closest(P, c, cache, use_cache, &dist, &surfIdx, &level);
    if (*maxDist < dist && *maxDist < SCN_INF) { /* fieldDist = INF without fields */
        o_move_local(c, *maxDist, c->nesting);
        cache->lvl = 0;
        return EV_COLL;
    } else if (surfIdx == SURF_BORDER && level == 1) {
        o_move_global(c, dist);
        *maxDist = dist;
        cache->lvl = 0;
        if (P->border_box >= 0) orc_box_explicit_bc(&P->boxes[P->border_box], c->lvl[0].r, c->lvl[0].dir);
        o_place_coord(P, c);
        return EV_BOUNDARY;
    } else {
        o_move_local(c, dist, level);
        *maxDist = dist;
        for (int l = 0; l < level - 1; ++l) cache->dist[l] = cache->dist[l] - dist;
        cache->lvl = level - 1;
        o_uni_cross(P, &c->lvl[level - 1], surfIdx);
        o_dive(P, c, level);
        return EV_CROSS;
    }
}

/* geometryStd_class.f90:492-514 teleport */
static void o_teleport(const scn_problem* P, o_coords* c, double dist)
{
    o_move_global(c, dist);
    o_place_coord(P, c);
    if (c->matIdx == SCN_OUTSIDE_MAT) {
        if (P->border_box >= 0) orc_box_transform_bc(&P->boxes[P->border_box], c->lvl[0].r, c->lvl[0].dir);
        o_place_coord(P, c);
    }
}

static void o_export_place(const o_coords* c, orc_place_out* out)
{
    out->matIdx = c->matIdx;
    out->uniqueID = c->uniqueID;
    out->nesting = c->nesting;
    for (int l = 0; l < SCN_MAX_NEST; ++l) {
        out->localID[l] = (l < c->nesting) ? c->lvl[l].localID : 0;
        out->uniIdx[l] = (l < c->nesting) ? c->lvl[l].uniIdx : 0;
        for (int i = 0; i < 3; ++i) out->r[l][i] = (l < c->nesting) ? c->lvl[l].r[i] : 0.0;
    }
}

static void o_init_coords(o_coords* c, const double r[3], const double u[3])
{
    memset(c, 0, sizeof *c);
    o_take_above(c);
    for (int i = 0; i < 3; ++i) { c->lvl[0].r[i] = r[i]; c->lvl[0].dir[i] = u[i]; }
}

int orc_place(const scn_problem* P, const double r[3], const double u[3], orc_place_out* out)
{
    o_coords c;
    o_clear();
    o_init_coords(&c, r, u);
    o_place_coord(P, &c);
    o_export_place(&c, out);
    return g_err;
}

int orc_move(const scn_problem* P, const double r[3], const double u[3], double max_dist,
             orc_place_out* out, double u_out[3], double* dist_out, int32_t* event)
{
    o_coords c;
    o_cache cache;
    o_clear();
    cache.lvl = 0;
    o_init_coords(&c, r, u);
    o_place_coord(P, &c);
    double d = max_dist;
    *event = o_move(P, &c, &d, &cache, 1);
    *dist_out = d;
    o_export_place(&c, out);
    for (int i = 0; i < 3; ++i) u_out[i] = c.lvl[0].dir[i];
    return g_err;
}

int orc_teleport(const scn_problem* P, const double r[3], const double u[3], double dist,
                 orc_place_out* out, double u_out[3])
{
    o_coords c;
    o_clear();
    o_init_coords(&c, r, u);
    o_place_coord(P, &c);
    o_teleport(P, &c, dist);
    o_export_place(&c, out);
    for (int i = 0; i < 3; ++i) u_out[i] = c.lvl[0].dir[i];
    return g_err;
}

/* ======================================================================================
 * Nuclear data (baseMgNeutronDatabase_class.f90, baseMgNeutronMaterial_class.f90)
 * ====================================================================================== */
enum { XS_TOTAL = 0, XS_SCATTER = 1, XS_CAPTURE = 2, XS_FISSION = 3, XS_NUFISSION = 4, XS_KAPPA = 5 };

static const double* o_xs(const scn_problem* P, int mat, int G)
{
    return P->mg.xs + ((size_t)(mat - 1) * P->mg.n_groups + (G - 1)) * 6;
}

/* getTrackMatXS (baseMgNeutronDatabase_class.f90:143-165) */
static double o_track_mat_xs(const scn_problem* P, int mat, int G)
{
    if (mat == SCN_VOID_MAT) return 0.0;
    if (mat < 1 || mat > P->mg.n_mat) {
        o_fail(SCN_ERR_MATERIAL, "Particle is in an undefined material with index: %d", mat);
        return 0.0;
    }
    return o_xs(P, mat, G)[XS_TOTAL] + 0.0; /* + getAlphaAbsorption() = 0 */
}

/* neutronMacroXSs%get (neutronXsPackages_class.f90) restricted to MG macro codes */
static double o_macro_get(const double* x, int MT)
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

/* ======================================================================================
 * Particle, tallies and the per-history loop
 * ====================================================================================== */
typedef struct o_site_buf { scn_particle_state* v; int64_t n, cap; } o_site_buf;

typedef struct o_part {
    o_coords coords;
    int G;
    double w, w0, time, k_eff, track_xs, preColl_w;
    double E, lambda;
    int type, isMG;
    uint64_t rng;
    int isDead, fate, broodID, collisionN;
    int splitCount;        /* neutronMGimp splits of this particle (0 when it leaves the buffer) */
} o_part;

typedef struct o_hctx {
    const scn_problem* P;
    double* bins;          /* collision-clerk BIN memory, or NULL */
    double keff[4];        /* per-history IMP_PROD, SCATTER_PROD, IMP_ABS, ANA_LEAK */
    o_site_buf* sites;
    int64_t stack_cap;     /* > 0: sites form the history's LIFO buffer of this size (fixed source) */
    orc_hist* h;
    int64_t n_events;
} o_hctx;

enum { K_IMP_PROD = 0, K_SCATTER_PROD = 1, K_IMP_ABS = 2, K_ANA_LEAK = 3 };
enum { FATE_NO = 5000, FATE_ABS = 5001, FATE_LEAK = 5002 };

/* coordList%cell: the cellShelf index of the cell at the lowest level (0 unless that level is a
 * cell universe in one of its defined cells) */
static int o_cell_of(const scn_problem* P, const o_coords* c)
{
    const o_coord* l = &c->lvl[(c->nesting > 1 ? c->nesting : 1) - 1];
    const scn_universe* U = o_uni(P, l->uniIdx);
    if (U->type != SCN_UNI_CELL || l->localID < 1 || l->localID > U->n_cells) return 0;
    return P->int_data[U->cells + U->n_cells + l->localID - 1];
}

/* spaceMap / materialMap / energyMap / cellMap / multiMap (Tallies/TallyMaps) on the particle's
 * state: global position, material, lowest-level cell */
static int o_map1d_at(const scn_problem* P, const scn_map1d* m, const double rg[3], int mat, int cell)
{
    if (m->type == SCN_MAP_ENERGY) return 0;   /* energyMap_class.f90 map: multigroup -> 0 */
    if (m->type == SCN_MAP_CELL) {             /* cellMap_class.f90 map: binMap%getOrDefault */
        if (cell < 0 || cell > m->axis) cell = 0;
        return P->int_data[m->data + cell];
    }
    if (m->type == SCN_MAP_SPACE) {
        double v = rg[m->axis];
        int idx;
        if (m->grid_type == SCN_GRID_LIN) { /* grid_class.f90:154-175 */
            const double* g = P->real_data + m->data;
            idx = (int)floor((v - g[0]) / g[1]) + 1;
        } else { /* binarySearch: last bin edge <= value */
            const double* g = P->real_data + m->data;
            int lo = 0, hi = m->nbins; /* edges g[0..nbins] */
            if (v < g[0] || v > g[m->nbins]) idx = -1;
            else {
                while (hi - lo > 1) {
                    int mid = (lo + hi) / 2;
                    if (v >= g[mid]) lo = mid; else hi = mid;
                }
                idx = lo + 1;
            }
        }
        if (idx < 1 || idx >= m->nbins + 1) idx = 0;
        return idx;
    } else {
        int nm = P->mg.n_mat;
        return (mat >= 0 && mat <= nm) ? P->int_data[m->data + mat] : P->int_data[m->data + nm + 1];
    }
}

/* bin of a clerk's (or weight-window field's) map for a state: multiMap_class.f90:153-173 -- the
 * product of the 1D maps' bins with strides multi[k] (first map fastest), 0 when any map misses.
 * A map-less clerk has the single bin 1. */
static int o_clerk_bin(const scn_problem* P, const scn_clerk* C, const double rg[3], int mat, int cell)
{
    int bin = 1;
    for (int k = 0; k < C->n_maps; ++k) {
        int b = o_map1d_at(P, &C->maps[k], rg, mat, cell);
        if (b == 0) return 0;
        bin = bin + (b - 1) * C->multi[k];
    }
    return bin;
}

/* tallyAdmin%reportInColl (tallyAdmin_class.f90:466-486) for a collision at global position rg in
 * material mat (lowest-level cell `cell`) of a particle of group G and weight w whose tracking
 * cross section was track_xs: the attachment (keffImplicitClerk, keffImplicitClerk_class.f90:
 * 180-223) first, then every collision clerk (collisionClerk_class.f90:192-242): energyFilter
 * (energyFilter_class.f90:109-129), map bin, flux = w / sigma (tracking xs when virtual collisions
 * are handled, the material total otherwise), address mem + width*(bin-1) + i. */
static void o_score_coll(const scn_problem* P, double* keff, double* bins, const double rg[3], int mat, int cell,
                         int G, double w, double track_xs, int isVirtual)
{
    if (mat != SCN_VOID_MAT) {
        double flux = w / track_xs;
        const double* x = o_xs(P, mat, G);
        double nuFiss = x[XS_NUFISSION];
        double absXS = x[XS_CAPTURE] + x[XS_FISSION];
        keff[K_IMP_PROD] = keff[K_IMP_PROD] + nuFiss * flux;
        keff[K_IMP_ABS] = keff[K_IMP_ABS] + absXS * flux;
    }
    if (!bins) return;
    for (int ci = 0; ci < P->n_clerks; ++ci) {
        const scn_clerk* C = &P->clerks[ci];
        if (C->kind != SCN_CLERK_COLLISION) continue;
        if (!C->handle_virtual && isVirtual) continue;
        if (C->gtop > 0 && !(C->gtop <= G && G <= C->glow)) continue;
        int bin = o_clerk_bin(P, C, rg, mat, cell);
        if (bin == 0) continue;
        double flux = C->handle_virtual ? w / track_xs : w / o_track_mat_xs(P, mat, G);
        int64_t addr = C->mem_addr + (int64_t)C->width * (bin - 1);
        for (int i = 0; i < C->width; ++i) {
            double resp;
            if (C->resp[i] == SCN_MT_FLUX) resp = 1.0;
            else resp = (mat == SCN_VOID_MAT) ? 0.0 : o_macro_get(o_xs(P, mat, G), C->resp[i]);
            bins[addr + i] = bins[addr + i] + resp * flux;
        }
    }
}

static void o_report_in_coll(o_hctx* X, const o_part* p, int isVirtual)
{
    o_score_coll(X->P, X->keff, X->bins, p->coords.lvl[0].r, p->coords.matIdx, o_cell_of(X->P, &p->coords), p->G,
                 p->w, p->track_xs, isVirtual);
}

/* tallyAdmin%reportHist -> keffImplicitClerk%reportHist (keffImplicitClerk_class.f90:289-299) */
static void o_report_hist(o_hctx* X, const o_part* p)
{
    if (p->fate == FATE_LEAK) X->keff[K_ANA_LEAK] = X->keff[K_ANA_LEAK] + p->w;
}

/* transportOperatorST_class.f90:48-166 surfaceTracking */
/* tallyAdmin_class.f90:545-568 reportPath to trackClerks (trackClerk_class.f90 reportPath): the
 * path of length L scores response(pre-path material, group) * w * L in the bin of the pre-path
 * state (position and material before the move) */
static void o_report_path(o_hctx* X, const o_part* p, double L, const double r0[3], int mat0, int cell0)
{
    const scn_problem* P = X->P;
    if (!X->bins) return;
    for (int ci = 0; ci < P->n_clerks; ++ci) {
        const scn_clerk* C = &P->clerks[ci];
        if (C->kind != SCN_CLERK_TRACK) continue;
        if (C->gtop > 0 && !(C->gtop <= p->G && p->G <= C->glow)) continue;
        int bin = o_clerk_bin(P, C, r0, mat0, cell0);
        if (bin == 0) continue;
        int64_t addr = C->mem_addr + (int64_t)C->width * (bin - 1);
        for (int i = 0; i < C->width; ++i) {
            double resp;
            if (C->resp[i] == SCN_MT_FLUX) resp = 1.0;
            else resp = (mat0 == SCN_VOID_MAT) ? 0.0 : o_macro_get(o_xs(P, mat0, p->G), C->resp[i]);
            X->bins[addr + i] = X->bins[addr + i] + resp * p->w * L;
        }
    }
}

static void o_surface_tracking(o_hctx* X, o_part* p)
{
    const scn_problem* P = X->P;
    o_cache cache;
    cache.lvl = 0;
    for (;;) {
        double sigmaTrack, sigmaT, dist, invSigmaTrack;
        int mat = p->coords.matIdx;
        /* getTrackingXS(MATERIAL_XS), baseMgNeutronDatabase_class.f90:95-137 */
        if (mat == SCN_VOID_MAT) sigmaTrack = P->mg.collision_xs;
        else sigmaTrack = scn_fmax(o_track_mat_xs(P, mat, p->G), P->mg.collision_xs);
        p->track_xs = sigmaTrack;
        if (sigmaTrack < SCN_ST_TOL) {
            dist = SCN_INF;
            invSigmaTrack = SCN_INF;
            sigmaT = 0.0;
        } else {
            invSigmaTrack = 1.0 / sigmaTrack;
            dist = -scn_log(scn_rng_get(&p->rng)) * invSigmaTrack;
            sigmaT = o_track_mat_xs(P, mat, p->G);
            if (dist != dist) { o_fail(SCN_ERR_NAN, "Distance is NaN"); p->isDead = 1; return; }
        }
        double r0[3] = {p->coords.lvl[0].r[0], p->coords.lvl[0].r[1], p->coords.lvl[0].r[2]};  /* savePrePath */
        int mat0 = p->coords.matIdx;
        int cell0 = o_cell_of(P, &p->coords);
        int event = o_move(P, &p->coords, &dist, &cache, P->st_cache);
        X->n_events++;
        p->time = p->time + dist / 1.0;
        if (event == EV_COLL) p->fate = FATE_NO;
        o_report_path(X, p, dist, r0, mat0, cell0);
        mat = p->coords.matIdx;
        if (mat == SCN_OUTSIDE_MAT) {
            p->isDead = 1;
            p->fate = FATE_LEAK;
        } else if (mat == SCN_UNDEF_MAT) {
            o_fail(SCN_ERR_MATERIAL, "Particle is in undefined material");
            p->isDead = 1;
            return;
        } else if (mat == SCN_OVERLAP_MAT) {
            o_fail(SCN_ERR_MATERIAL, "Particle is in overlapping cells");
            p->isDead = 1;
            return;
        }
        if (p->isDead) break;
        if (event == EV_COLL) {
            if (scn_rng_get(&p->rng) < sigmaT * invSigmaTrack) break;
            o_report_in_coll(X, p, 1);
        }
        if (g_err) { p->isDead = 1; return; }
    }
}

/* transportOperatorDT_class.f90:47-130 deltaTracking */
static void o_delta_tracking(o_hctx* X, o_part* p)
{
    const scn_problem* P = X->P;
    /* getTrackingXS(MAJORANT_XS) */
    double maj = scn_fmax(P->mg.majorant[p->G - 1] + 0.0, P->mg.collision_xs);
    p->track_xs = maj;
    double majorant_inv = 1.0 / maj;
    if (fabs(majorant_inv) > SCN_HUGE) { o_fail(SCN_ERR_MAJORANT, "Majorant is 0"); p->isDead = 1; return; }
    for (;;) {
        double distance = -scn_log(scn_rng_get(&p->rng)) * majorant_inv;
        o_teleport(P, &p->coords, distance);
        X->n_events++;
        p->time = p->time + distance / 1.0;
        int mat = p->coords.matIdx;
        if (mat == SCN_OUTSIDE_MAT) {
            p->fate = FATE_LEAK;
            p->isDead = 1;
            break;
        } else if (mat == SCN_VOID_MAT) {
            o_report_in_coll(X, p, 1);
            continue;
        } else if (mat == SCN_UNDEF_MAT || mat == SCN_OVERLAP_MAT) {
            o_fail(SCN_ERR_MATERIAL, "Particle is in undefined or overlapping material");
            p->isDead = 1;
            return;
        }
        double sigmaT = o_track_mat_xs(P, mat, p->G);
        if (scn_rng_get(&p->rng) < sigmaT * majorant_inv) break;
        o_report_in_coll(X, p, 1);
        if (g_err) { p->isDead = 1; return; }
    }
}

/* transportOperatorHT_class.f90:41-64 tracking_selection: delta tracking when
 * sigmaT * (1 / majorant) > 1 - cutoff, surface tracking otherwise (sigmaT = 0 in void) */
static int o_ht_use_dt(const scn_problem* P, const o_part* p)
{
    double majorant_inv = 1.0 / scn_fmax(P->mg.majorant[p->G - 1] + 0.0, P->mg.collision_xs);
    int mat = p->coords.matIdx;
    double sigmaT = (mat == SCN_VOID_MAT) ? 0.0 : o_track_mat_xs(P, mat, p->G);
    double ratio = sigmaT * majorant_inv;
    return ratio > (1.0 - P->ht_cutoff);
}

/* transportOperator_inter.f90:105-130 transport */
static void o_transport(o_hctx* X, o_part* p)
{
    int dt = X->P->transport == SCN_TRANSPORT_DT ||
             (X->P->transport == SCN_TRANSPORT_HT && o_ht_use_dt(X->P, p));
    if (dt) o_delta_tracking(X, p);
    else o_surface_tracking(X, p);
    if (p->isDead) o_report_hist(X, p);
}

static void o_site_push(o_site_buf* b, const scn_particle_state* s)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->v = (scn_particle_state*)realloc(b->v, sizeof(scn_particle_state) * (size_t)b->cap);
    }
    b->v[b->n++] = *s;
}

/* legendrePoly_func.f90:35-99 sampleLegendre_P1: mu from 1/2 (1 + c mu) on [-1, 1] as a mixture
 * of uniform / linear (|c| < 1) or linear / delta at 1 (1 <= |c| <= 3), mirrored for c < 0 */
static int o_legendre_p1(double c, uint64_t* rng, double* mu)
{
    enum { UNIFORM = 1, LIN = 2, DELTA = 3 };
    double a = fabs(c), threshold;
    int top, low;
    if (a < 1.0) {
        threshold = a;
        top = LIN;
        low = UNIFORM;
    } else if (a <= 3.0) {
        threshold = 0.5 * (a - 1.0);
        top = DELTA;
        low = LIN;
    } else {
        o_fail(SCN_ERR_SAMPLING, "P1 must have absolute value < 3.0");
        return g_err;
    }
    int ex = (scn_rng_get(rng) < threshold) ? top : low;
    double x;
    if (ex == UNIFORM) x = 2.0 * scn_rng_get(rng) - 1.0;
    else if (ex == LIN) x = 2.0 * sqrt(scn_rng_get(rng)) - 1.0;
    else x = 1.0;
    if (c < 0.0) x = -x;
    *mu = x;
    return 0;
}

/* fissionMG%sampleOut group (fissionMG_class.f90:183-206): subtract chi(G_out) from the draw u
 * until it falls below 0; nG + 1 when the scan falls through (unnormalised chi) */
static int o_chi_gout(const double* chi, int nG, double u)
{
    double rem = u;
    int G_out;
    for (G_out = 1; G_out <= nG; ++G_out) {
        rem = rem - chi[G_out - 1];
        if (rem < 0.0) break;
    }
    return G_out;
}

/* multiScatterMG%sampleGout (multiScatterMG_class.f90:265-290) for incident group G_in with draw u:
 * u * scatterXS(G_in) minus P0(G_out, G_in) over G_out until below 0; *w_mul = production(G_in,
 * G_out) = scatteringMultiplicity(G_out, G_in) (multiScatterMG_class.f90:199-222, sampleOut's
 * weight change).  nG + 1 when the scan falls through. */
static int o_scatter_gout(const scn_problem* P, int mat, int G_in, double u, double* w_mul)
{
    const int nG = P->mg.n_groups;
    const double* x = o_xs(P, mat, G_in);
    const double* P0 = P->mg.P0 + ((size_t)(mat - 1) * nG + (G_in - 1)) * nG;
    const double* prod = P->mg.prod + ((size_t)(mat - 1) * nG + (G_in - 1)) * nG;
    double rem = u * x[XS_SCATTER];
    int G_out;
    for (G_out = 1; G_out <= nG; ++G_out) {
        rem = rem - P0[G_out - 1];
        if (rem < 0.0) break;
    }
    if (G_out <= nG) *w_mul = prod[G_out - 1];
    return G_out;
}

/* collisionProcessor_inter.f90:114-194 collide with neutronMGstd_class.f90:85-282 */
/* neutronMGimp_class.f90 cutoffs with weight windows (weightWindowsField_class.f90 atP): the window
 * of the particle's state bin; above the upper weight the particle is split (split), below the
 * lower weight it plays Russian roulette (russianRoulette) for weight lower*constSurvival */
static void o_weight_window(o_hctx* X, o_part* p, int mat)
{
    const scn_problem* P = X->P;
    const scn_clerk* M = P->ww_map;
    int bin = o_clerk_bin(P, M, p->coords.lvl[0].r, p->coords.matIdx, o_cell_of(P, &p->coords));
    double minW = 0.0, maxW = 0.0, avW = 0.0;
    if (bin > 0) {
        minW = P->ww_lower[bin - 1];
        maxW = P->ww_upper[bin - 1];
        avW = minW * P->ww_survival;
    }
    if (p->w > maxW && maxW != 0.0 && p->splitCount < P->max_split) {
        int mult = (int)ceil(p->w / maxW);
        if (mult + p->splitCount > P->max_split) mult = P->max_split - p->splitCount + 1;
        scn_particle_state s;   /* pTemp = p */
        memset(&s, 0, sizeof s);
        for (int k = 0; k < 3; ++k) { s.r[k] = p->coords.lvl[0].r[k]; s.dir[k] = p->coords.lvl[0].dir[k]; }
        s.wgt = p->w / mult;
        s.G = p->G;
        s.E = p->E;
        s.isMG = 1;
        s.type = p->type;
        s.time = p->time;
        s.lambda = p->lambda;
        s.matIdx = mat;
        s.uniqueID = p->coords.uniqueID;
        s.collisionN = p->collisionN;
        s.broodID = p->broodID;
        for (int i = 1; i <= mult - 1; ++i) {
            if (X->stack_cap <= 0) { o_fail(SCN_ERR_UNSUPPORTED, "weight-window splitting needs a buffer"); return; }
            if (X->sites->n >= X->stack_cap) {
                o_fail(SCN_ERR_OVERFLOW, "Run out of space for particles. Max size:%lld", (long long)X->stack_cap);
                p->isDead = 1;
                return;
            }
            o_site_push(X->sites, &s);
        }
        p->splitCount = p->splitCount + mult;
        p->w = p->w / mult;
    } else if (p->w < minW) {
        if (scn_rng_get(&p->rng) < (1.0 - p->w / avW)) p->isDead = 1;
        else p->w = avW;
    }
}

static void o_collide(o_hctx* X, o_part* p)
{
    const scn_problem* P = X->P;
    const int nG = P->mg.n_groups;
    int mat = p->coords.matIdx;
    X->n_events++;
    /* sampleCollision */
    double denom = o_track_mat_xs(P, mat, p->G);
    double probAlpha = 0.0 / denom;
    if (scn_rng_get(&p->rng) < probAlpha) { o_fail(SCN_ERR_UNSUPPORTED, "alpha absorption"); p->isDead = 1; return; }
    const double* x = o_xs(P, mat, p->G);
    double r = scn_rng_get(&p->rng);
    /* invert_macroXSs (neutronXsPackages_class.f90:211-…) */
    int C = 1;
    double xs = x[XS_TOTAL] * r - 0.0;
    if (xs > 0.0) C++;
    xs = xs - x[XS_SCATTER];
    if (xs > 0.0) C++;
    xs = xs - x[XS_CAPTURE];
    if (xs > 0.0) C++;
    o_report_in_coll(X, p, 0);
    p->preColl_w = p->w;
    /* implicit: fission site production */
    if (P->mg.fissile[mat - 1]) {
        double wgt = p->w, w0 = p->w0, k_eff = p->k_eff;
        double rand1 = scn_rng_get(&p->rng);
        double sig_tot = x[XS_TOTAL], sig_nuFiss = x[XS_NUFISSION];
        int n = (int)(fabs((wgt * sig_nuFiss) / (w0 * sig_tot * k_eff)) + rand1);
        if (n >= 1) {
            wgt = scn_sign(w0, wgt);
            const double* chi = P->mg.chi + (size_t)(mat - 1) * nG;
            for (int i = 0; i < n; ++i) {
                double mu = 2.0 * scn_rng_get(&p->rng) - 1.0;
                double phi = SCN_TWO_PI * scn_rng_get(&p->rng);
                int G_out = o_chi_gout(chi, nG, scn_rng_get(&p->rng));
                if (G_out > nG) { o_fail(SCN_ERR_SAMPLING, "fissionMG sampleOut: unnormalised chi"); p->isDead = 1; return; }
                scn_particle_state s;
                memset(&s, 0, sizeof s);
                scn_rotate(p->coords.lvl[0].dir, mu, phi, s.dir);
                for (int k = 0; k < 3; ++k) s.r[k] = p->coords.lvl[0].r[k];
                s.G = G_out;
                s.wgt = wgt;
                /* neutronMGstd restarts the site's count; neutronMGimp keeps the parent's */
                s.collisionN = (P->collision == SCN_COLL_IMP) ? p->collisionN : 0;
                s.E = p->E;
                s.isMG = 1;
                s.type = p->type;
                s.time = p->time;
                s.lambda = p->lambda;
                s.matIdx = mat;
                s.uniqueID = p->coords.uniqueID;
                s.cellIdx = 0;
                s.broodID = p->broodID;
                if (X->stack_cap > 0 && X->sites->n >= X->stack_cap) {
                    /* particleDungeon_class.f90:167-172 detain beyond the buffer size */
                    o_fail(SCN_ERR_OVERFLOW, "Run out of space for particles. Max size:%lld", (long long)X->stack_cap);
                    p->isDead = 1;
                    return;
                }
                o_site_push(X->sites, &s);
                if (X->h) X->h->n_sites++;
            }
        }
    }
    if (C == 2) { /* inelastic: multiScatterMG sampleOut */
        double w_mul = 0.0;
        int G_out = o_scatter_gout(P, mat, p->G, scn_rng_get(&p->rng), &w_mul);
        if (G_out > nG) { o_fail(SCN_ERR_SAMPLING, "multiScatterMG sampleGout: sampling failed"); p->isDead = 1; return; }
        double mu;
        if (P->mg.P1) {
            /* multiScatterP1MG_class.f90:69-85 sampleOut -> legendrePoly_func.f90:35-99 */
            double c = P->mg.P1[((size_t)(mat - 1) * nG + (p->G - 1)) * nG + (G_out - 1)];
            if (o_legendre_p1(c, &p->rng, &mu)) { p->isDead = 1; return; }
        } else {
            mu = 2.0 * scn_rng_get(&p->rng) - 1.0;
        }
        double phi = SCN_TWO_PI * scn_rng_get(&p->rng);
        p->G = G_out;
        p->w = p->w * w_mul;
        double nd[3];
        scn_rotate(p->coords.lvl[0].dir, mu, phi, nd);
        o_assign_direction(P, &p->coords, nd);
    } else if (C == 3 || C == 4) {
        p->isDead = 1;
    } else {
        o_fail(SCN_ERR_UNSUPPORTED, "Unsupported MT number: %d", -3);
        p->isDead = 1;
        return;
    }
    if (P->collision == SCN_COLL_IMP && P->ww_n > 0 && !p->isDead) {
        o_weight_window(X, p, mat);
        if (g_err) return;
    }
    p->collisionN += 1;
    /* reportOutColl -> keffImplicitClerk (scattering weight gain) */
    if (C == 2) {
        double score = scn_fmax(p->w - p->preColl_w, 0.0);
        if (score > 0.0) X->keff[K_SCATTER_PROD] = X->keff[K_SCATTER_PROD] + score;
    }
    if (X->h) X->h->n_coll++;
    if (p->isDead) {
        p->fate = FATE_ABS;
        o_report_hist(X, p);
    }
}

/* eigenPhysicsPackage_class.f90:211-252, one history n (1-based) */
static void o_history(o_hctx* X, const scn_particle_state* s, int n, uint64_t master, double k_eff)
{
    o_part p;
    memset(&p, 0, sizeof p);
    p.rng = scn_rng_stride(master, n);
    o_init_coords(&p.coords, s->r, s->dir);
    p.w = s->wgt;
    p.w0 = s->wgt;
    p.G = s->G;
    p.E = s->E;
    p.isMG = 1;
    p.type = s->type;
    p.time = s->time;
    p.lambda = s->lambda;
    p.collisionN = s->collisionN;
    p.broodID = n;
    p.fate = FATE_NO;
    o_place_coord(X->P, &p.coords);
    p.k_eff = k_eff;
    p.preColl_w = p.w;
    X->keff[0] = X->keff[1] = X->keff[2] = X->keff[3] = 0.0;
    for (;;) {
        o_transport(X, &p);
        if (p.isDead || g_err) break;
        o_collide(X, &p);
        if (p.isDead || g_err) break;
    }
    if (X->h) {
        X->h->n_events = X->n_events;
        X->h->rng_end = p.rng;
        X->h->w_end = p.w;
        X->h->fate = p.fate;
        X->h->G_end = p.G;
        for (int k = 0; k < 3; ++k) X->h->r_end[k] = p.coords.lvl[0].r[k];
    }
}

/* fixedSourcePhysicsPackage_class.f90:201-258, one history n (1-based) with its secondaries: fission
 * sites go to the per-thread buffer (collide(p, tally, buffer, buffer)); when a particle dies the
 * last detained site is released (particleDungeon_class.f90:266-281, LIFO) and transported on the
 * same RNG stream, until the buffer is empty.  k_eff = 1. */
static void o_history_fixed(o_hctx* X, const scn_particle_state* s, int n, uint64_t master)
{
    o_part p;
    memset(&p, 0, sizeof p);
    p.rng = scn_rng_stride(master, n);
    scn_particle_state cur = *s;
    cur.broodID = n;                 /* thisCycle%copy(p, n) */
    X->sites->n = 0;
    X->keff[0] = X->keff[1] = X->keff[2] = X->keff[3] = 0.0;
    int32_t n_coll = 0;
    for (;;) {
        /* particle_fromParticleState (particle_class.f90:290-310) */
        o_init_coords(&p.coords, cur.r, cur.dir);
        p.w = cur.wgt;
        p.w0 = cur.wgt;
        p.G = cur.G;
        p.E = cur.E;
        p.isMG = 1;
        p.type = cur.type;
        p.time = cur.time;
        p.lambda = cur.lambda;
        p.fate = FATE_NO;
        p.collisionN = cur.collisionN;
        p.broodID = cur.broodID;
        p.splitCount = 0;
        p.isDead = 0;
        p.k_eff = 1.0;
        o_place_coord(X->P, &p.coords);
        p.preColl_w = p.w;
        for (;;) {
            o_transport(X, &p);
            if (p.isDead || g_err) break;
            o_collide(X, &p);
            n_coll++;
            if (p.isDead || g_err) break;
        }
        if (g_err || X->sites->n == 0) break;
        cur = X->sites->v[--X->sites->n];
    }
    if (X->h) {
        X->h->n_events = X->n_events;
        X->h->n_coll = n_coll;
        X->h->rng_end = p.rng;
        X->h->w_end = p.w;
        X->h->fate = p.fate;
        X->h->G_end = p.G;
        for (int k = 0; k < 3; ++k) X->h->r_end[k] = p.coords.lvl[0].r[k];
    }
}

/* the histories of one fixed-source batch (fixedSourcePhysicsPackage_class.f90:199-261) into BIN */
static int o_fixed_batch(const scn_problem* P, const scn_particle_state* bank, int32_t n_in, uint64_t master,
                         double* bins, int32_t buffer_size, orc_hist* hist, int32_t n_threads)
{
    int nt = 1;
#ifdef _OPENMP
    if (n_threads > 0) nt = n_threads;
#else
    (void)n_threads;
#endif
    o_site_buf* bufs = (o_site_buf*)calloc((size_t)nt, sizeof(o_site_buf));
    double** tbins = (double**)calloc((size_t)nt, sizeof(double*));
    tbins[0] = bins;
    for (int t = 1; t < nt; ++t) tbins[t] = (double*)calloc((size_t)P->n_bins + 1, sizeof(double));
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int n = 1; n <= n_in; ++n) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        if (g_err) continue;
        o_hctx X;
        memset(&X, 0, sizeof X);
        X.P = P;
        X.bins = P->n_bins ? tbins[t] : NULL;
        X.sites = &bufs[t];
        X.stack_cap = buffer_size;
        X.h = hist ? &hist[n - 1] : NULL;
        if (X.h) memset(X.h, 0, sizeof *X.h);
        o_history_fixed(&X, &bank[n - 1], n, master);
    }
    for (int t = 1; t < nt; ++t)
        for (int64_t i = 0; i < P->n_bins; ++i) bins[i] = bins[i] + tbins[t][i];
    for (int t = 0; t < nt; ++t) free(bufs[t].v);
    for (int t = 1; t < nt; ++t) free(tbins[t]);
    free(tbins);
    free(bufs);
    return g_err;
}

int orc_eigen_cycle(const scn_problem* P, const scn_particle_state* bank_in, int32_t n_in,
                    int64_t rng_state, double k_eff, scn_particle_state* bank_out, int32_t cap,
                    int32_t* n_out, double* bin_scores, double* keff_scores, orc_hist* hist,
                    int32_t n_threads)
{
    o_clear();
    if (n_in < 0 || !bank_in) { o_fail(SCN_ERR_ARG, "bad bank"); return g_err; }
    int nt = 1;
#ifdef _OPENMP
    if (n_threads > 0) nt = n_threads;
#else
    (void)n_threads;
#endif
    double* hk = (double*)calloc((size_t)n_in * 4 + 1, sizeof(double));
    int64_t* hs_start = (int64_t*)calloc((size_t)n_in + 1, sizeof(int64_t));
    int32_t* hs_count = (int32_t*)calloc((size_t)n_in + 1, sizeof(int32_t));
    int32_t* hs_thr = (int32_t*)calloc((size_t)n_in + 1, sizeof(int32_t));
    o_site_buf* bufs = (o_site_buf*)calloc((size_t)nt, sizeof(o_site_buf));
    double** tbins = (double**)calloc((size_t)nt, sizeof(double*));
    if (bin_scores) {
        memset(bin_scores, 0, sizeof(double) * (size_t)P->n_bins);
        tbins[0] = bin_scores;
        for (int t = 1; t < nt; ++t) tbins[t] = (double*)calloc((size_t)P->n_bins + 1, sizeof(double));
    }
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int n = 1; n <= n_in; ++n) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        if (g_err) continue;
        o_hctx X;
        memset(&X, 0, sizeof X);
        X.P = P;
        X.bins = bin_scores ? tbins[t] : NULL;
        X.sites = &bufs[t];
        X.h = hist ? &hist[n - 1] : NULL;
        if (X.h) memset(X.h, 0, sizeof *X.h);
        int64_t start = bufs[t].n;
        o_history(&X, &bank_in[n - 1], n, (uint64_t)rng_state, k_eff);
        hs_start[n - 1] = start;
        hs_count[n - 1] = (int32_t)(bufs[t].n - start);
        hs_thr[n - 1] = t;
        for (int k = 0; k < 4; ++k) hk[(size_t)k * n_in + (n - 1)] = X.keff[k];
    }
    int64_t total = 0;
    for (int n = 0; n < n_in; ++n) total += hs_count[n];
    if (!g_err && total > cap) o_fail(SCN_ERR_OVERFLOW, "Run out of space for particles. Max size: %d", cap);
    if (!g_err) {
        int64_t pos = 0;
        for (int n = 0; n < n_in; ++n)
            for (int j = 0; j < hs_count[n]; ++j) bank_out[pos++] = bufs[hs_thr[n]].v[hs_start[n] + j];
        *n_out = (int32_t)total;
        if (keff_scores)
            for (int k = 0; k < 4; ++k) keff_scores[k] = orc_dsum(hk + (size_t)k * n_in, n_in);
        if (bin_scores)
            for (int t = 1; t < nt; ++t)
                for (int64_t i = 0; i < P->n_bins; ++i) bin_scores[i] = bin_scores[i] + tbins[t][i];
    }
    for (int t = 0; t < nt; ++t) free(bufs[t].v);
    for (int t = 1; t < nt; ++t) free(tbins[t]);
    free(tbins);
    free(bufs);
    free(hk);
    free(hs_start);
    free(hs_count);
    free(hs_thr);
    return g_err;
}

/* ======================================================================================
 * Population control: particleDungeon_class.f90:431-602 normSize_Repr (one rank) with
 * heapQueue_class.f90:55-178 and sortByBroodID :923-981
 * ====================================================================================== */
typedef struct o_heap { double* v; int n, cap; } o_heap;

static void o_heap_sift_up(o_heap* h, int i)
{
    while (i > 0) {
        int par = (i - 1) / 2;
        if (h->v[par] >= h->v[i]) break;
        double t = h->v[par]; h->v[par] = h->v[i]; h->v[i] = t;
        i = par;
    }
}
static void o_heap_sift_down(o_heap* h, int i)
{
    for (;;) {
        int l = 2 * i + 1, r = l + 1, m = i;
        if (l < h->n && h->v[l] > h->v[m]) m = l;
        if (r < h->n && h->v[r] > h->v[m]) m = r;
        if (m == i) break;
        double t = h->v[m]; h->v[m] = h->v[i]; h->v[i] = t;
        i = m;
    }
}
/* pushReplace: push while not full, otherwise replace the maximum */
static void o_heap_push_replace(o_heap* h, double x)
{
    if (h->n < h->cap) {
        h->v[h->n++] = x;
        o_heap_sift_up(h, h->n - 1);
    } else {
        h->v[0] = x;
        o_heap_sift_down(h, 0);
    }
}

static void o_sort_by_brood(scn_particle_state* bank, int n)
{
    int maxB = 0;
    for (int i = 0; i < n; ++i) if (bank[i].broodID > maxB) maxB = bank[i].broodID;
    if (maxB == 0) return;
    int* cnt = (int*)calloc((size_t)maxB + 2, sizeof(int));
    for (int i = 0; i < n; ++i) cnt[bank[i].broodID]++;
    int loc = 0;
    for (int b = 0; b <= maxB; ++b) { int c = cnt[b]; cnt[b] = loc; loc += c; }
    scn_particle_state* tmp = (scn_particle_state*)malloc(sizeof(scn_particle_state) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) tmp[cnt[bank[i].broodID]++] = bank[i];
    memcpy(bank, tmp, sizeof(scn_particle_state) * (size_t)n);
    free(tmp);
    free(cnt);
}

int orc_norm_size_repr(scn_particle_state* bank, int32_t* np, int32_t cap, int32_t tot_pop, int64_t rng_state)
{
    o_clear();
    int pop = *np;
    o_sort_by_brood(bank, pop);
    int totSites = pop;
    if (totSites == 0) { o_fail(SCN_ERR_NORM, "Normalisation failed: no fission sites"); return g_err; }
    int excess = totSites - tot_pop;
    int heapSize = (excess < 0) ? ((-excess) % totSites) : excess;
    double threshold = 1.0;
    uint64_t seed = 0;
    if (heapSize != 0) {
        uint64_t mr = (uint64_t)rng_state;
        o_heap h;
        h.cap = heapSize;
        h.n = 0;
        h.v = (double*)malloc(sizeof(double) * (size_t)heapSize);
        o_heap_push_replace(&h, 2.0);
        seed = mr;
        for (int j = 0; j < pop; ++j) {
            double rn = scn_rng_get(&mr);
            if (rn < h.v[0]) o_heap_push_replace(&h, rn);
        }
        threshold = h.v[0];
        free(h.v);
    }
    uint64_t rr = seed;
    if (excess > 0) {
        int count = 0;
        for (int i = 0; i < pop; ++i)
            if (scn_rng_get(&rr) > threshold) bank[count++] = bank[i];
        pop = count;
    } else if (excess < 0) {
        int n_copies = (-excess) / totSites;
        int n_dup = (-excess) % totSites;
        int64_t need = (int64_t)pop * (n_copies + 1) + n_dup;
        if (need > cap) { o_fail(SCN_ERR_OVERFLOW, "normSize_Repr: bank capacity %d too small", cap); return g_err; }
        for (int c = 1; c <= n_copies; ++c) memcpy(bank + (size_t)pop * c, bank, sizeof(scn_particle_state) * (size_t)pop);
        int count = pop * (n_copies + 1);
        if (n_dup != 0)
            for (int i = 0; i < pop; ++i)
                if (scn_rng_get(&rr) <= threshold) bank[count++] = bank[i];
        pop = count;
        o_sort_by_brood(bank, pop);
    }
    if (pop != tot_pop) { o_fail(SCN_ERR_NORM, "Normalisation failed!"); return g_err; }
    *np = pop;
    return 0;
}

/* ======================================================================================
 * Sources
 * ====================================================================================== */

static void o_result(double csum, double csum2, int N, double out[2]);

/* fissionSource_class.f90:149-271 via source_inter.f90:98-119 (particle i uses stride(i)) */
int orc_fission_source(const scn_problem* P, int32_t n, int64_t rng_state, const double bottom[3],
                       const double top[3], int32_t g_in, int32_t attempts, scn_particle_state* bank)
{
    o_clear();
    (void)g_in;
    const int nG = P->mg.n_groups;
    for (int i = 1; i <= n && !g_err; ++i) {
        uint64_t rng = scn_rng_stride((uint64_t)rng_state, i);
        scn_particle_state* s = &bank[i - 1];
        memset(s, 0, sizeof *s);
        int it = 0;
        for (;;) {
            if (++it > attempts) { o_fail(SCN_ERR_SAMPLING, "Failed to find a fissile material"); break; }
            double r3[3], r[3];
            for (int k = 0; k < 3; ++k) r3[k] = scn_rng_get(&rng);
            for (int k = 0; k < 3; ++k) r[k] = (top[k] - bottom[k]) * r3[k] + bottom[k];
            o_coords c;
            double u0[3] = {1.0, 0.0, 0.0};
            o_init_coords(&c, r, u0);
            o_place_coord(P, &c);
            int mat = c.matIdx;
            if (mat == SCN_VOID_MAT || mat == SCN_OUTSIDE_MAT) continue;
            if (mat == SCN_UNDEF_MAT || mat == SCN_OVERLAP_MAT) { o_fail(SCN_ERR_MATERIAL, "source in undefined material"); break; }
            if (!P->mg.fissile[mat - 1]) continue;
            s->matIdx = mat;
            s->uniqueID = c.uniqueID;
            s->wgt = 1.0;
            s->time = 0.0;
            s->type = 1;
            for (int k = 0; k < 3; ++k) s->r[k] = r[k];
            double mu = 2.0 * scn_rng_get(&rng) - 1.0;
            double phi = SCN_TWO_PI * scn_rng_get(&rng);
            int G = o_chi_gout(P->mg.chi + (size_t)(mat - 1) * nG, nG, scn_rng_get(&rng));
            if (G > nG) { o_fail(SCN_ERR_SAMPLING, "fissionMG sampleOut: unnormalised chi"); break; }
            s->G = G;
            s->isMG = 1;
            double x1[3] = {1.0, 0.0, 0.0};
            scn_rotate(x1, mu, phi, s->dir);
            s->lambda = SCN_INF;
            s->cellIdx = -1;
            break;
        }
    }
    return g_err;
}

/* pointSource%init (pointSource_class.f90:120-128,150-152): dir / norm2(dir) and
 * probG / sum(probG), the sum taken left to right */
static void o_source_prepare(const scn_source* in, scn_source* out, double* probG)
{
    *out = *in;
    if (in->type == SCN_SRC_POINT && !in->isotropic) {
        double n = scn_norm2(in->dir);
        for (int k = 0; k < 3; ++k) out->dir[k] = in->dir[k] / n;
    }
    if (in->type == SCN_SRC_POINT && in->n_probG > 0) {
        double t = 0.0;
        for (int g = 0; g < in->n_probG; ++g) t = t + in->probG[g];
        for (int g = 0; g < in->n_probG; ++g) probG[g] = in->probG[g] / t;
        out->probG = probG;
    }
}

/* configSource_inter.f90:129-141 sampleParticle with pointSource's sampleType, samplePosition,
 * sampleEnergyAngle (:197-212) and sampleEnergy (:221-246), in that order */
static int o_point_sample(const scn_source* S, uint64_t* rng, scn_particle_state* s)
{
    memset(s, 0, sizeof *s);
    s->matIdx = -1;
    s->cellIdx = -1;
    s->uniqueID = -1;
    s->lambda = SCN_INF;
    s->type = 1;
    for (int k = 0; k < 3; ++k) s->r[k] = S->r[k];
    if (S->isotropic) {
        double mu = 2.0 * scn_rng_get(rng) - 1.0;
        double phi = SCN_TWO_PI * scn_rng_get(rng);
        double x1[3] = {1.0, 0.0, 0.0};
        scn_rotate(x1, mu, phi, s->dir);
    } else {
        for (int k = 0; k < 3; ++k) s->dir[k] = S->dir[k];
    }
    if (S->n_probG > 0) {
        double r = scn_rng_get(rng);
        int g;
        for (g = 1; g <= S->n_probG; ++g) {
            r = r - S->probG[g - 1];
            if (r < 0.0) break;
        }
        if (g > S->n_probG) { o_fail(SCN_ERR_SAMPLING, "pointSource sampleEnergy: probG scan fell through"); return g_err; }
        s->G = g;
    } else {
        s->G = S->G;
    }
    s->isMG = 1;
    s->time = 0.0;
    s->wgt = 1.0;
    return 0;
}

/* materialSource_class.f90:150-231 sampleParticle (MG data) */
static int o_material_sample(const scn_problem* P, const scn_source* S, uint64_t* rng, scn_particle_state* s)
{
    memset(s, 0, sizeof *s);
    s->matIdx = -1;
    s->cellIdx = -1;
    s->uniqueID = -1;
    s->lambda = SCN_INF;
    for (int i = 1;; ++i) {
        if (i > 200) {
            o_fail(SCN_ERR_SAMPLING, "Infinite loop in sampling source. Please check that defined volume contains source material.");
            return g_err;
        }
        double rand3[3], r[3];
        for (int k = 0; k < 3; ++k) rand3[k] = scn_rng_get(rng);
        for (int k = 0; k < 3; ++k) r[k] = (S->top[k] - S->bottom[k]) * rand3[k] + S->bottom[k];
        double time = S->t_low + scn_rng_get(rng) * (S->t_high - S->t_low);
        o_coords c;
        double u0[3] = {1.0, 0.0, 0.0};   /* geometryStd%whatIsAt without u */
        o_init_coords(&c, r, u0);
        o_place_coord(P, &c);
        int mat = c.matIdx;
        if (mat == SCN_OUTSIDE_MAT) continue;
        if (mat < 1 || mat > P->mg.n_mat) {
            o_fail(SCN_ERR_MATERIAL, "Nuclear data did not return neutron material.");
            return g_err;
        }
        if (mat != S->mat) continue;
        s->matIdx = mat;
        s->uniqueID = c.uniqueID;
        s->wgt = 1.0;
        s->time = time;
        s->type = 1;
        for (int k = 0; k < 3; ++k) s->r[k] = r[k];
        double mu = 2.0 * scn_rng_get(rng) - 1.0;
        double phi = SCN_TWO_PI * scn_rng_get(rng);
        double x1[3] = {1.0, 0.0, 0.0};
        scn_rotate(x1, mu, phi, s->dir);
        s->G = S->G;
        s->isMG = 1;
        return 0;
    }
}

/* source_inter.f90:98-119 generate: particle i (1-based) is sampled with the master strided by i */
int orc_source_sample(const scn_problem* P, const scn_source* src, int32_t n, int64_t rng_state,
                      scn_particle_state* bank)
{
    o_clear();
    double probG[512];
    scn_source S;
    if (src->n_probG > 512) { o_fail(SCN_ERR_ARG, "probG longer than 512 groups"); return g_err; }
    o_source_prepare(src, &S, probG);
    for (int i = 1; i <= n && !g_err; ++i) {
        uint64_t rng = scn_rng_stride((uint64_t)rng_state, i);
        if (S.type == SCN_SRC_POINT) o_point_sample(&S, &rng, &bank[i - 1]);
        else if (S.type == SCN_SRC_MATERIAL) o_material_sample(P, &S, &rng, &bank[i - 1]);
        else o_fail(SCN_ERR_ARG, "unknown source type %d", S.type);
    }
    return g_err;
}

/* fixedSourcePhysicsPackage_class.f90:135-147 run and :149-289 cycles (one rank, no common
 * buffer), buffer_size = 'buffer' (default 50): per batch,
 * generate(pop, pRNG); pRNG%stride(totalPop); history n on pRNG strided by n; pRNG%stride(totalPop);
 * reportCycleEnd -> scoreMemory%closeCycle(normFactor) (tallyAdmin_class.f90:736-786).
 * stop_bin >= 0 adds an early stop: after a batch (N >= 2 batches closed), the solve ends when the
 * bin's sd / |mean| <= stop_rel.  bins_result gets (mean, sd) per bin from N batches. */
int orc_fixed_run(const scn_problem* P, const scn_source* src, int32_t pop, int32_t n_batches, int64_t seed,
                  int64_t stop_bin, double stop_rel, int32_t buffer_size, double* bins_result, orc_hist* hist,
                  int32_t* n_done, int32_t n_threads)
{
    o_clear();
    if (buffer_size < 1) { o_fail(SCN_ERR_ARG, "buffer size must be >= 1"); return g_err; }
    const int64_t nb = P->n_bins;
    scn_particle_state* bank = (scn_particle_state*)calloc((size_t)pop + 1, sizeof(scn_particle_state));
    double* bins = (double*)calloc((size_t)nb + 1, sizeof(double));
    double* csum = (double*)calloc((size_t)nb + 1, sizeof(double));
    double* csum2 = (double*)calloc((size_t)nb + 1, sizeof(double));
    uint64_t master = scn_rng_stride((uint64_t)seed, 0);   /* stride(getOffset(totalPop)), one rank */
    int done = 0;
    for (int b = 0; b < n_batches && !g_err; ++b) {
        if (orc_source_sample(P, src, pop, (int64_t)master, bank)) break;
        master = scn_rng_stride(master, pop);
        if (o_fixed_batch(P, bank, pop, master, bins, buffer_size, hist ? hist + (size_t)b * pop : NULL, n_threads))
            break;
        master = scn_rng_stride(master, pop);
        double norm = 1.0;
        if (P->norm_bin >= 0) {
            double ns = bins[P->norm_bin];
            if (ns == 0.0) { o_fail(SCN_ERR_NORM, "Normalisation score from clerk is 0"); break; }
            norm = P->norm_value / ns;
        }
        orc_close_cycle(bins, csum, csum2, nb, norm);
        done = b + 1;
        if (stop_bin >= 0 && stop_bin < nb && done >= 2) {
            double o[2];
            o_result(csum[stop_bin], csum2[stop_bin], done, o);
            if (o[0] != 0.0 && o[1] / fabs(o[0]) <= stop_rel) break;
        }
    }
    if (!g_err && bins_result)
        for (int64_t i = 0; i < nb; ++i) o_result(csum[i], csum2[i], done, bins_result + 2 * i);
    if (n_done) *n_done = done;
    free(bank);
    free(bins);
    free(csum);
    free(csum2);
    return g_err;
}

/* ======================================================================================
 * Power iteration: eigenPhysicsPackage_class.f90:135-343 with scoreMemory_class.f90 closes
 * ====================================================================================== */
void orc_score_result(double csum, double csum2, int N, double out[2]);
static void o_result(double csum, double csum2, int N, double out[2])
{
    /* scoreMemory_class.f90:537-573 getResult_withSTD */
    double mean = csum / N;
    double inv_N = 1.0 / N;
    double inv_Nm1 = (N != 1) ? 1.0 / (N - 1) : 1.0;
    double std = csum2 * inv_N * inv_Nm1 - mean * mean * inv_Nm1;
    out[0] = mean;
    out[1] = sqrt(std);
}
void orc_score_result(double csum, double csum2, int N, double out[2]) { o_result(csum, csum2, N, out); }

/* scoreMemory_class.f90:309-342 closeCycle (batchSize 1): res = BIN*norm; CSUM += res; CSUM2 += res^2 */
void orc_close_cycle(double* bins, double* csum, double* csum2, int64_t n, double norm)
{
    for (int64_t i = 0; i < n; ++i) {
        double r = bins[i] * norm;
        bins[i] = 0.0;
        csum[i] = csum[i] + r;
        csum2[i] = csum2[i] + r * r;
    }
}

/* keffImplicitClerk_class.f90:306-326 closeCycle estimate; scores in IMP_PROD, SCATTER_PROD,
 * IMP_ABS, ANA_LEAK order */
double orc_keff_implicit(const double s[4]) { return s[0] / (s[2] + s[3] - s[1]); }

int orc_legendre_p1(double c, uint64_t* rng, double* mu)
{
    o_clear();
    return o_legendre_p1(c, rng, mu);
}

int orc_eigen_run(const scn_problem* P, const scn_eigen_settings* S,
                  const scn_particle_state* initial_bank, scn_eigen_result* res, int32_t n_threads)
{
    o_clear();
    const int pop = S->pop;
    const int cap = 4 * pop + 16;
    scn_particle_state* thisC = (scn_particle_state*)malloc(sizeof(scn_particle_state) * (size_t)cap);
    scn_particle_state* nextC = (scn_particle_state*)malloc(sizeof(scn_particle_state) * (size_t)cap);
    memcpy(thisC, initial_bank, sizeof(scn_particle_state) * (size_t)pop);
    int nThis = pop;
    uint64_t master = (uint64_t)S->seed;
    master = scn_rng_stride(master, 0);      /* run: stride(getOffset(totalPop)) = 0 */
    master = scn_rng_stride(master, pop);    /* generateInitialState: stride(totalPop) */
    double* bins = P->n_bins ? (double*)calloc((size_t)P->n_bins, sizeof(double)) : NULL;
    double* csumA = P->n_bins ? (double*)calloc((size_t)P->n_bins * 2, sizeof(double)) : NULL;
    double* csumI = P->n_bins ? (double*)calloc((size_t)P->n_bins * 2, sizeof(double)) : NULL;
    double kA[2] = {0, 0}, kI[2] = {0, 0}, imp[4][2];
    memset(imp, 0, sizeof imp);
    int nA = 0, nI = 0;
    double dk[2] = {1.0, 1.0};   /* particleDungeon%k_eff of the two dungeons */
    int cur = 0;                 /* index of thisCycle dungeon */
    double keff_0 = S->keff0;
    int cyc = 0;
    res->n_histories = 0;
    res->n_collisions = 0;
    for (int phase = 0; phase < 2 && !g_err; ++phase) {
        int ncyc = phase == 0 ? S->n_inactive : S->n_active;
        int score_bins = phase == 1 || S->inactive_tally;
        double k_new = keff_0;
        for (int c = 0; c < ncyc && !g_err; ++c) {
            double startW;
            {
                double* w = (double*)malloc(sizeof(double) * (size_t)(nThis > 0 ? nThis : 1));
                for (int i = 0; i < nThis; ++i) w[i] = thisC[i].wgt;
                startW = orc_dsum(w, nThis);
                free(w);
            }
            int nNext = 0;
            double ks[4];
            orc_hist* hist = NULL;
            if (orc_eigen_cycle(P, thisC, nThis, (int64_t)master, k_new, nextC, cap, &nNext,
                                score_bins ? bins : NULL, ks, hist, n_threads))
                break;
            res->n_histories += nThis;
            master = scn_rng_stride(master, (int64_t)pop + 1);
            double endW;
            {
                double* w = (double*)malloc(sizeof(double) * (size_t)(nNext > 0 ? nNext : 1));
                for (int i = 0; i < nNext; ++i) w[i] = nextC[i].wgt;
                endW = orc_dsum(w, nNext);
                free(w);
            }
            double k_norm = dk[1 - cur];
            if (phase == 0) {
                orc_keff_analog(startW, endW, k_norm, kA);
                nA++;
            } else { /* keffImplicitClerk_class.f90:306-326 */
                double k_est = ks[0] / (ks[2] + ks[3] - ks[1]);
                kI[0] = kI[0] + k_est;
                kI[1] = kI[1] + k_est * k_est;
                nI++;
                for (int k = 0; k < 4; ++k) {
                    imp[k][0] = imp[k][0] + ks[k];
                    imp[k][1] = imp[k][1] + ks[k] * ks[k];
                }
            }
            if (score_bins) { /* scoreMemory%closeCycle(normFactor = 1) */
                double* cs = phase == 1 ? csumA : csumI;
                for (int64_t i = 0; i < P->n_bins; ++i) {
                    double r = bins[i] * 1.0;
                    bins[i] = 0.0;
                    cs[2 * i] = cs[2 * i] + r;
                    cs[2 * i + 1] = cs[2 * i + 1] + r * r;
                }
            }
            if (orc_norm_size_repr(nextC, &nNext, cap, pop, (int64_t)master)) break;
            master = scn_rng_stride(master, 1);
            scn_particle_state* t = thisC; thisC = nextC; nextC = t;
            nThis = nNext;
            cur = 1 - cur;
            double kr[2];
            if (phase == 0) o_result(kA[0], kA[1], nA, kr);
            else o_result(kI[0], kI[1], nI, kr);
            k_new = kr[0];
            dk[1 - cur] = k_new;
            keff_0 = k_new;
            if (res->k_history) res->k_history[cyc] = k_new;
            cyc++;
        }
    }
    if (!g_err) {
        if (nA) o_result(kA[0], kA[1], nA, res->k_analog);
        if (nI) {
            o_result(kI[0], kI[1], nI, res->k_implicit);
            for (int k = 0; k < 4; ++k) o_result(imp[k][0], imp[k][1], nI, res->imp[k]);
        }
        for (int64_t i = 0; i < P->n_bins; ++i) {
            if (res->bins_active && nI) {
                double o[2];
                o_result(csumA[2 * i], csumA[2 * i + 1], nI, o);
                res->bins_active[2 * i] = o[0];
                res->bins_active[2 * i + 1] = o[1];
            }
            if (res->bins_inactive && nA && S->inactive_tally) {
                double o[2];
                o_result(csumI[2 * i], csumI[2 * i + 1], nA, o);
                res->bins_inactive[2 * i] = o[0];
                res->bins_inactive[2 * i + 1] = o[1];
            }
        }
    }
    free(thisC);
    free(nextC);
    free(bins);
    free(csumA);
    free(csumI);
    return g_err;
}

/* ---- single-operation entry points (tests pin these against the reference's unit tests) ---- */

int orc_clerk_bin(const scn_problem* P, int32_t clerk, const double r[3], int32_t mat, int32_t cell)
{
    if (clerk < 0 || clerk >= P->n_clerks) return -1;
    return o_clerk_bin(P, &P->clerks[clerk], r, mat, cell);
}

void orc_score_collision(const scn_problem* P, const double r[3], int32_t mat, int32_t cell, int32_t G, double w,
                         double track_xs, int32_t is_virtual, double* bins, double keff[4])
{
    o_score_coll(P, keff, bins, r, mat, cell, G, w, track_xs, is_virtual);
}

int orc_scatter_gout(const scn_problem* P, int32_t mat, int32_t G_in, double u, double* w_mul)
{
    *w_mul = 0.0;
    return o_scatter_gout(P, mat, G_in, u, w_mul);
}

int orc_chi_gout(const scn_problem* P, int32_t mat, double u)
{
    const int nG = P->mg.n_groups;
    return o_chi_gout(P->mg.chi + (size_t)(mat - 1) * nG, nG, u);
}

/* keffAnalogClerk%closeCycle (keffAnalogClerk_class.f90:162-181): the cycle's estimate
 * k = endWgt / startWgt * k_eff (k_eff of the dungeon, the k the cycle ran with), accumulated into
 * acc[0] (sum) and acc[1] (sum of squares) as scoreMemory%accumulate does */
void orc_keff_analog(double start_w, double end_w, double k_norm, double acc[2])
{
    double k = end_w / start_w * k_norm;
    acc[0] = acc[0] + k;
    acc[1] = acc[1] + k * k;
}
