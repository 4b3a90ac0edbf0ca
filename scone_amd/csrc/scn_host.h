// scn_host.h -- host-side helpers shared by the eigenvalue engine and the batch engine.
#pragma once

#include <cmath>
#include <functional>
#include <vector>
#include <algorithm>

#include "scone_gpu.h"
#include "scn_device.h"

namespace scn {

// nesting depth of the geometry graph (root = 1): the deepest chain of universe fills
// (geomGraph_class.f90 getFill walked from the root); 0 for a malformed graph
inline int problem_depth(const scn_problem* pr)
{
    if (pr->root_uni < 1 || pr->root_uni > pr->n_universes) return 0;
    std::vector<int> rootpos(pr->n_universes + 1, 0);
    rootpos[pr->root_uni] = 1;
    for (int i = 0; i < pr->graph_size; ++i) {
        int f = pr->graph_fill[i];
        if (f < 0 && -f <= pr->n_universes) rootpos[-f] = pr->graph_id[i];
    }
    std::function<int(int, int)> dep = [&](int u, int guard) -> int {
        if (guard > SCN_MAX_NEST + 1) return SCN_MAX_NEST + 2;
        int best = 0;
        const scn_universe& U = pr->universes[u - 1];
        for (int l = 1; l <= U.n_local; ++l) {
            int pos = rootpos[u] + l - 1;
            if (pos < 1 || pos > pr->graph_size) continue;
            int f = pr->graph_fill[pos - 1];
            if (f < 0 && -f <= pr->n_universes) best = std::max(best, dep(-f, guard + 1));
        }
        return best + 1;
    };
    return dep(pr->root_uni, 1);
}

// some universe applies a rotation matrix on enter
inline int any_rotated(const scn_problem* pr)
{
    for (int i = 0; i < pr->n_universes; ++i)
        if (pr->universes[i].rotated) return 1;
    return 0;
}

// index checks of the surface and cell tables, so that kernels read only inside them;
// returns an error message or nullptr
inline const char* check_cell_tables(const scn_problem* pr)
{
    for (int i = 0; i < pr->n_surfaces; ++i) {
        const scn_surface& S = pr->surfaces[i];
        if (S.type < SCN_SURF_APLANE || S.type > SCN_SURF_WEDGE) return "unknown surface type";
        if (S.type == SCN_SURF_BOX) {
            if (S.axis < 0 || S.axis >= pr->n_boxes) return "box surface refers to a missing box";
        } else if (S.type != SCN_SURF_PLANE && S.type != SCN_SURF_QUADRIC && S.type != SCN_SURF_SPHERE) {
            if (S.axis < 0 || S.axis > 2) return "surface axis outside 0..2";
        }
    }
    if (pr->n_surfaces > 0 && !pr->surfaces) return "surface table missing";
    for (int c = 0; c < pr->n_clerks; ++c)
        if (pr->clerks[c].kind != SCN_CLERK_COLLISION && pr->clerks[c].kind != SCN_CLERK_TRACK)
            return "unknown clerk kind";
    if (pr->collision != SCN_COLL_STD && pr->collision != SCN_COLL_IMP) return "unknown collision processor";
    if (pr->ww_n < 0 || (pr->ww_n > 0 && (!pr->ww_map || !pr->ww_lower || !pr->ww_upper)))
        return "weight-window tables missing";
    if (pr->ww_n > 0) {
        if (pr->ww_map->n_maps < 0 || pr->ww_map->n_maps > 3) return "weight-window map";
        int64_t n = 1;
        for (int k = 0; k < pr->ww_map->n_maps; ++k) n *= pr->ww_map->maps[k].nbins;
        if (n != pr->ww_n) return "weight-window bins do not match the map";
        if (pr->collision != SCN_COLL_IMP) return "weight windows need neutronMGimp";
        if (pr->max_split < 1) return "maxSplit must be at least 1";
    }
    for (int u = 0; u < pr->n_universes; ++u) {
        const scn_universe& U = pr->universes[u];
        if (U.type < SCN_UNI_ROOT || U.type > SCN_UNI_CELL) return "unknown universe type";
        if (U.type == SCN_UNI_ROOT) {
            if (U.border_surf >= pr->n_surfaces) return "root border surface outside the surface table";
            if (U.border_surf < 0 && (U.border < 0 || U.border >= pr->n_boxes)) return "root border box missing";
        }
        if (U.type != SCN_UNI_CELL) continue;
        if (U.n_cells < 0 || U.n_local != U.n_cells + 2) return "cell universe local cell count";
        if (U.cells < 0 || (int64_t)U.cells + 2 * (int64_t)U.n_cells > pr->n_int)
            return "cell universe list outside int_data";
        for (int c = 0; c < U.n_cells; ++c) {
            const int rec = pr->int_data[U.cells + c];
            if (rec < 0 || rec >= pr->n_int) return "cell record outside int_data";
            if (pr->int_data[rec] == SCN_UCELL) {
                // unionCell: [SCN_UCELL, simple, m, tokens]; surfaces in range, operators known,
                // brackets balanced (the short-circuit skip relies on it)
                if ((int64_t)rec + 3 > pr->n_int) return "cell record outside int_data";
                const int m = pr->int_data[rec + 2];
                if (m < 1 || (int64_t)rec + 3 + m > pr->n_int) return "cell record outside int_data";
                int depth = 0;
                for (int k = 0; k < m; ++k) {
                    const int t = pr->int_data[rec + 3 + k];
                    if (t >= SCN_UCELL_UNION) {
                        if (t > SCN_UCELL_COMPLEM) return "unknown unionCell operator";
                        if (t == SCN_UCELL_OPEN) ++depth;
                        if (t == SCN_UCELL_CLOSE && --depth < 0) return "unbalanced unionCell brackets";
                    } else if (t == 0 || t > pr->n_surfaces || -t > pr->n_surfaces) {
                        return "cell surface outside the surface table";
                    }
                }
                if (depth != 0) return "unbalanced unionCell brackets";
                continue;
            }
            const int n = pr->int_data[rec];
            if (n < 0 || (int64_t)rec + 1 + n > pr->n_int) return "cell record outside int_data";
            for (int k = 1; k <= n; ++k) {
                const int s = pr->int_data[rec + k];
                if (s == 0 || s > pr->n_surfaces || -s > pr->n_surfaces) return "cell surface outside the surface table";
            }
        }
    }
    return nullptr;
}

// the problem needs the general kernels: cell universes, rotated or 'global' universes, a root
// border that is not a box, or track-length clerks
inline int needs_full_geometry(const scn_problem* pr)
{
    for (int c = 0; c < pr->n_clerks; ++c)
        if (pr->clerks[c].kind == SCN_CLERK_TRACK) return 1;   // path scoring lives in the general kernels
    if (pr->ww_n > 0) return 1;                                // so do weight windows
    for (int i = 0; i < pr->n_universes; ++i) {
        const scn_universe& U = pr->universes[i];
        if (U.type == SCN_UNI_CELL || U.rotated || U.global) return 1;
        if (U.type == SCN_UNI_ROOT && U.border_surf >= 0) return 1;
    }
    return 0;
}

// compiled register-array depth for a geometry depth
inline int compiled_nl(int depth, int geo = 0)
{
    if (geo) return depth <= 3 ? 3 : 8;   // general-geometry kernels exist for 3 and 8 levels
    if (depth <= 1) return 1;
    if (depth <= 3) return 3;
    if (depth <= 4) return 4;
    return 8;
}

// clerks of one kind whose energyFilter (energyFilter_class.f90:109-129) admits group g, evaluated once:
// goff[g-1] .. goff[g] index into gidx
inline void group_clerks(const scn_problem* pr, std::vector<int>& goff, std::vector<int>& gidx,
                         int kind = SCN_CLERK_COLLISION)
{
    const int nG = pr->mg.n_groups;
    goff.assign(nG + 1, 0);
    gidx.clear();
    for (int g = 1; g <= nG; ++g) {
        for (int c = 0; c < pr->n_clerks; ++c) {
            const scn_clerk& K = pr->clerks[c];
            if (K.kind != kind) continue;
            if (K.gtop > 0 && !(K.gtop <= g && g <= K.glow)) continue;
            gidx.push_back(c);
        }
        goff[g] = (int)gidx.size();
    }
}

// FastColl of the problem's collision clerks, or fc.on = 0 when they do not fit it.  Clerks with
// the same layout (maps, responses, virtual-collision flag) share one of the FAST_CLERKS slots
// when no group is admitted by two clerks of the slot (a mesh tally split into one clerk per
// group takes one slot; a total-flux clerk next to per-group flux clerks takes a second); maps
// must be lin spaceMaps.  g_addr receives the
// per-(slot, group) first-bin table (pointer set by the caller after upload).
inline FastColl fast_collision_layout(const scn_problem* pr, const std::vector<int>& goff,
                                      const std::vector<int>& gidx, std::vector<int64_t>& g_addr)
{
    FastColl F{};
    const int nG = std::max(pr->mg.n_groups, 1);
    g_addr.assign(nG, -1);
    auto same_layout = [&](const scn_clerk& K, const scn_clerk& R) {
        if (K.n_maps != R.n_maps || K.width != R.width || K.handle_virtual != R.handle_virtual) return false;
        for (int i = 0; i < K.width; ++i)
            if (K.resp[i] != R.resp[i]) return false;
        for (int k = 0; k < K.n_maps; ++k) {
            const scn_map1d &m = K.maps[k], &f = R.maps[k];
            if (m.axis != f.axis || m.nbins != f.nbins || K.multi[k] != R.multi[k]) return false;
            if (pr->real_data[m.data] != pr->real_data[f.data] || pr->real_data[m.data + 1] != pr->real_data[f.data + 1])
                return false;
        }
        return true;
    };
    std::vector<int> slot_of(pr->n_clerks, -1), rep;   // slot of each collision clerk, a clerk of each slot
    std::vector<std::vector<char>> taken;               // groups admitted by a clerk of each slot
    std::vector<std::vector<char>> admits(pr->n_clerks, std::vector<char>(nG, 0));
    for (int g = 1; g <= pr->mg.n_groups; ++g)
        for (int e = goff[g - 1]; e < goff[g]; ++e) admits[gidx[e]][g - 1] = 1;
    for (int c = 0; c < pr->n_clerks; ++c) {
        const scn_clerk& K = pr->clerks[c];
        if (K.kind != SCN_CLERK_COLLISION) continue;
        if (K.n_maps < 0 || K.n_maps > 3 || K.width < 1 || K.width > 8) return F;
        for (int k = 0; k < K.n_maps; ++k) {
            const scn_map1d& m = K.maps[k];
            if (m.type != SCN_MAP_SPACE || m.grid_type != SCN_GRID_LIN) return F;
            if (m.data < 0 || m.data + 1 >= pr->n_real) return F;
        }
        auto fits = [&](int q) {
            if (!same_layout(K, pr->clerks[rep[q]])) return false;
            for (int g = 0; g < nG; ++g)
                if (admits[c][g] && taken[q][g]) return false;
            return true;
        };
        int q = 0;
        while (q < (int)rep.size() && !fits(q)) ++q;
        if (q == (int)rep.size()) {
            if (q == FAST_CLERKS) return F;
            rep.push_back(c);
            taken.emplace_back(nG, 0);
        }
        for (int g = 0; g < nG; ++g) taken[q][g] |= admits[c][g];
        slot_of[c] = q;
    }
    if (rep.empty()) return F;
    const int nc = (int)rep.size();
    std::vector<int64_t> addr((size_t)nc * nG, -1);
    for (int g = 1; g <= pr->mg.n_groups; ++g)
        for (int e = goff[g - 1]; e < goff[g]; ++e) {
            const int q = slot_of[gidx[e]];
            if (q < 0 || addr[(size_t)q * nG + g - 1] >= 0) return F;   // two clerks of one slot admit the group
            addr[(size_t)q * nG + g - 1] = pr->clerks[gidx[e]].mem_addr;
        }
    for (int q = 0; q < nc; ++q) {
        const scn_clerk& K = pr->clerks[rep[q]];
        FastClerk& C = F.c[q];
        C.n_maps = K.n_maps;
        C.width = K.width;
        C.handle_virtual = K.handle_virtual;
        for (int i = 0; i < K.width; ++i) C.resp[i] = K.resp[i];
        for (int k = 0; k < K.n_maps; ++k) {
            C.axis[k] = K.maps[k].axis;
            C.nbins[k] = K.maps[k].nbins;
            C.multi[k] = K.multi[k];
            C.g0[k] = pr->real_data[K.maps[k].data];
            C.step[k] = pr->real_data[K.maps[k].data + 1];
            C.step_inv[k] = 1.0 / C.step[k];
        }
    }
    g_addr = addr;
    F.on = 1;
    F.n_clerks = nc;
    return F;
}

// A flat lattice: a box-bordered root filled by one lattice each of whose cells holds a material or
// a one-region pinUniverse (slab, voxel and mesh-like geometries), nothing rotated or 'global'.
// Delta tracking then places a particle with the root's halfspace test and one lattice findCell
// (place_flat on the device), the cell's material, uniqueID and depth coming from `flat`:
// [0..1] the root's outside cell (mat, uid), then (mat, uid, nest, pin universe, pin graph root)
// for lattice local IDs 1..n_local, as diveToMat walks the geomGraph to them (pin 0 for a
// material cell).  Returns the lattice's universe index, or 0.
inline int flat_lattice(const scn_problem* pr, std::vector<int>& flat)
{
    flat.assign(2, 0);
    if (pr->root_uni < 1 || pr->root_uni > pr->n_universes || pr->graph_size < 2) return 0;
    const scn_universe& R = pr->universes[pr->root_uni - 1];
    if (R.type != SCN_UNI_ROOT || R.border_surf >= 0 || R.rotated || R.global) return 0;
    const int f_in = pr->graph_fill[0], f_out = pr->graph_fill[1];   // root local cells 1 and 2
    if (f_out < 0 || f_in >= 0) return 0;
    flat[0] = f_out;
    flat[1] = pr->graph_id[1];
    const int L = -f_in;
    if (L < 1 || L > pr->n_universes) return 0;
    const scn_universe& U = pr->universes[L - 1];
    if (U.type != SCN_UNI_LAT || U.rotated || U.global) return 0;
    const int R1 = pr->graph_id[0];
    for (int l = 1; l <= U.n_local; ++l) {
        const int pos = R1 + l - 1;
        if (pos < 1 || pos > pr->graph_size) return 0;
        int f = pr->graph_fill[pos - 1], id = pr->graph_id[pos - 1], nest = 2, pin = 0, pin_root = 0;
        if (f < 0) {
            const int C = -f;
            if (C < 1 || C > pr->n_universes) return 0;
            const scn_universe& Q = pr->universes[C - 1];
            if (Q.type != SCN_UNI_PIN || Q.n_annuli != 1 || Q.rotated || Q.global) return 0;
            if (id < 1 || id > pr->graph_size) return 0;   // the pin's only cell: graph position id
            pin = C;
            pin_root = id;
            f = pr->graph_fill[id - 1];
            id = pr->graph_id[id - 1];
            if (f < 0) return 0;
            nest = 3;
        }
        flat.push_back(f);
        flat.push_back(id);
        flat.push_back(nest);
        flat.push_back(pin);
        flat.push_back(pin_root);
    }
    return L;
}

// reciprocals of lattice pitches and sizes, 6 per universe (scn_floor_div and scn_idiv on the device)
inline std::vector<double> universe_inverse_table(const scn_problem* pr)
{
    std::vector<double> t(6 * (size_t)std::max(1, pr->n_universes), 0.0);
    for (int u = 0; u < pr->n_universes; ++u) {
        const scn_universe& U = pr->universes[u];
        if (U.type != SCN_UNI_LAT) continue;
        for (int i = 0; i < 3; ++i) t[6 * (size_t)u + i] = 1.0 / U.pitch[i];
        for (int i = 0; i < 2; ++i) t[6 * (size_t)u + 3 + i] = 1.0 / (double)std::max(1, U.size[i]);
    }
    return t;
}

// delta tracking's per-group majorant (baseMgNeutronDatabase getMajorantXS, floored by the
// collision XS) and its reciprocal, as transportOperatorDT_class.f90 computes them each flight
inline std::vector<double> majorant_table(const scn_problem* pr)
{
    const int nG = pr->mg.n_groups;
    std::vector<double> t(2 * (size_t)std::max(nG, 1), 0.0);
    for (int g = 0; g < nG; ++g) {
        const double m = scn_fmax(pr->mg.majorant[g] + 0.0, pr->mg.collision_xs);
        t[2 * g] = m;
        t[2 * g + 1] = 1.0 / m;
    }
    return t;
}

// surface tracking's sigmaTrack = max(total XS, collisionXS) per material and group and its
// reciprocal (transportOperatorST_class.f90 surfaceTracking), then the void pair (collisionXS)
inline std::vector<double> st_track_table(const scn_problem* pr)
{
    const int nG = pr->mg.n_groups, nM = pr->mg.n_mat;
    std::vector<double> t(2 * ((size_t)nM * nG + 1), 0.0);
    for (int m = 0; m < nM; ++m)
        for (int g = 0; g < nG; ++g) {
            const size_t k = (size_t)m * nG + g;
            const double s = scn_fmax(pr->mg.xs[k * 6 + XS_TOTAL] + 0.0, pr->mg.collision_xs);
            t[2 * k] = s;
            t[2 * k + 1] = 1.0 / s;
        }
    t[2 * (size_t)nM * nG] = pr->mg.collision_xs;
    t[2 * (size_t)nM * nG + 1] = 1.0 / pr->mg.collision_xs;
    return t;
}

// scoreMemory_class.f90:537-573 getResult: mean and standard deviation of the mean over N batches
inline void score_result(double csum, double csum2, int N, double out[2])
{
    double mean = csum / N;
    double inv_N = 1.0 / N;
    double inv_Nm1 = (N != 1) ? 1.0 / (N - 1) : 1.0;
    double var = csum2 * inv_N * inv_Nm1 - mean * mean * inv_Nm1;
    out[0] = mean;
    out[1] = sqrt(var);
}

}  // namespace scn
