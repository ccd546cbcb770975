// Reverse Cuthill-McKee ordering of the cell adjacency graph (host side).
//
// The reference has no RCM: its only reordering is a Morton z-order in a
// research script (zoomy_jax/gnn_blueprint/mesh_spectral_geometry.py:34-46)
// and an unused mesh field (zoomy_jax/mesh/mesh.py:169).  The permutation
// follows the algorithm of scipy.sparse.csgraph.reverse_cuthill_mckee:
// degrees (a diagonal entry counts twice), nodes visited from the lowest
// degree seed of each component, each BFS level's newly found nodes sorted
// by degree with a stable insertion sort, and the final order reversed.  The
// seed order uses a stable sort of the degrees (lowest index first among
// equal degrees).
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/zoomy_hip.h"

namespace {

// RCM with the component of `start` (>= 0) visited first from `start`; the
// other components from their lowest-degree node.
int rcm_impl(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t start,
             int32_t* perm_out) {
    if (n < 0 || (n > 0 && (!indptr || !perm_out)) || start >= n) return ZJ_EINVAL;
    if (n == 0) return ZJ_OK;
    std::vector<int32_t> degree(n);
    for (int32_t i = 0; i < n; ++i) {
        int32_t lo = indptr[i], hi = indptr[i + 1];
        if (hi < lo) return ZJ_EINVAL;
        degree[i] = hi - lo;
        for (int32_t jj = lo; jj < hi; ++jj) {
            int32_t j = indices[jj];
            if (j < 0 || j >= n) return ZJ_EINVAL;
            if (j == i) {
                degree[i] += 1;
                break;
            }
        }
    }
    std::vector<int32_t> inds(n);
    std::iota(inds.begin(), inds.end(), 0);
    std::stable_sort(inds.begin(), inds.end(),
                     [&](int32_t a, int32_t b) { return degree[a] < degree[b]; });
    std::vector<int32_t> rev(n);
    for (int32_t i = 0; i < n; ++i) rev[inds[i]] = i;
    std::vector<int32_t> order(n);
    int32_t N = 0;
    for (int32_t zz = start >= 0 ? -1 : 0; zz < n; ++zz) {
        if (zz < 0 || inds[zz] != -1) {
            int32_t seed = zz < 0 ? start : inds[zz];
            order[N++] = seed;
            inds[rev[seed]] = -1;
            int32_t level_start = N - 1, level_end = N;
            while (level_start < level_end) {
                for (int32_t ii = level_start; ii < level_end; ++ii) {
                    int32_t i = order[ii];
                    int32_t N_old = N;
                    for (int32_t jj = indptr[i]; jj < indptr[i + 1]; ++jj) {
                        int32_t j = indices[jj];
                        if (inds[rev[j]] != -1) {
                            inds[rev[j]] = -1;
                            order[N++] = j;
                        }
                    }
                    // stable insertion sort of the new nodes by degree
                    for (int32_t kk = N_old + 1; kk < N; ++kk) {
                        int32_t node = order[kk];
                        int32_t d = degree[node];
                        int32_t ll = kk;
                        while (ll > N_old && d < degree[order[ll - 1]]) {
                            order[ll] = order[ll - 1];
                            --ll;
                        }
                        order[ll] = node;
                    }
                }
                level_start = level_end;
                level_end = N;
            }
        }
        if (N == n) break;
    }
    for (int32_t i = 0; i < n; ++i) perm_out[i] = order[n - 1 - i];
    return ZJ_OK;
}

}  // namespace

extern "C" int zj_rcm(int32_t n, const int32_t* indptr, const int32_t* indices,
                      int32_t* perm_out) {
    return rcm_impl(n, indptr, indices, -1, perm_out);
}

extern "C" int zj_rcm_from(int32_t n, const int32_t* indptr, const int32_t* indices,
                           int32_t start, int32_t* perm_out) {
    if (start < 0) return ZJ_EINVAL;
    return rcm_impl(n, indptr, indices, start, perm_out);
}
