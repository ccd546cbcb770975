// Host-side patch decomposition for the LDS-staged stage kernel.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace zj {

struct PatchHost {
    int P = 0, npatch = 0, mf = 0, mx = 0, ntpl = 0;
    int expl = 0;  // own cells listed explicitly (cut from a cell order)
    int max_load = 0, max_slope = 0, max_vert = 0;
    std::vector<int32_t> tpl_id, n_load, n_vert;  // per patch
    std::vector<int32_t> load_ids;                // (npatch, max_load)
    std::vector<int32_t> vert_ids;                // (npatch, max_vert)
    std::vector<uint16_t> tpl_nbr;                // (ntpl, mf, max_slope)
    std::vector<uint16_t> tpl_extra;              // (ntpl, mx, max_slope)
    std::vector<uint16_t> tpl_vert;               // (ntpl, mf, max_slope)
    std::vector<int32_t> tpl_nslope, tpl_nown;    // per template
};

// Returns 0 on success.  cv / nbr are (mf, N), extra is (mx, N).  order
// (N ids, a permutation) or nullptr: patch p owns order[p P .. p P + P) or,
// without an order, the consecutive ids p P .. p P + P - 1.
int build_patches(int N, int mf, int mx, int nverts, const int32_t* cv, const int32_t* nbr,
                  const int32_t* extra, int P, const int32_t* order, PatchHost& out);

// Checks every local id of every patch against the global tables (bounds and
// identity of the referenced cell / vertex) and that every cell is owned by
// exactly one patch.  Returns 0 when consistent.
int validate_patches(int N, int mf, int mx, int nverts, const int32_t* nbr, const int32_t* extra,
                     const int32_t* cv, const PatchHost& ph);

// Cells sorted by the Morton code of their centre (cen: (2, N), SoA), ties
// by id: consecutive runs of this order are compact 2-D clusters whatever
// the mesh's own cell order.
std::vector<int32_t> morton_cluster_order(int N, const double* cen);

}  // namespace zj
