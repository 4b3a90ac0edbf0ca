// scn_host.h -- host-side helpers shared by the eigenvalue engine and the batch engine.
#pragma once

#include <cmath>
#include <functional>
#include <vector>
#include <algorithm>

#include "scone_gpu.h"

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

// compiled register-array depth for a geometry depth
inline int compiled_nl(int depth)
{
    if (depth <= 1) return 1;
    if (depth <= 3) return 3;
    if (depth <= 4) return 4;
    return 8;
}

// clerks whose energyFilter (energyFilter_class.f90:109-129) admits group g, evaluated once:
// goff[g-1] .. goff[g] index into gidx
inline void group_clerks(const scn_problem* pr, std::vector<int>& goff, std::vector<int>& gidx)
{
    const int nG = pr->mg.n_groups;
    goff.assign(nG + 1, 0);
    gidx.clear();
    for (int g = 1; g <= nG; ++g) {
        for (int c = 0; c < pr->n_clerks; ++c) {
            const scn_clerk& K = pr->clerks[c];
            if (K.gtop > 0 && !(K.gtop <= g && g <= K.glow)) continue;
            gidx.push_back(c);
        }
        goff[g] = (int)gidx.size();
    }
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
