// Patch decomposition of the mesh for the LDS-staged stage kernel (host).
//
// Cells are cut into patches of P consecutive ids, or of P consecutive
// entries of a given cell order (spatial clusters, for cell orders whose id
// runs are thin strips); the own cells are then listed explicitly.  For
// patch p the kernel needs, in LDS:
//   own cells        [p P, p P + n_own) (or their listed ids) local 0..n_own-1
//   face ring        face neighbours of own cells             slopes computed
//   outer ring       face + vertex neighbours of the above    state only
//   vertices         of every own / face-ring cell            geometry
// Local ids are assigned in first-discovery order (own cells in order, then
// each cell's slots and stencil entries in order), so the local tables of
// every interior patch of a regular mesh are identical and are stored once
// as a shared template; a patch carries only its template id and its global
// load / vertex id lists.
#include "zj_patches.hpp"

#include "../../include/zoomy_hip.h"

#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>

namespace zj {

int build_patches(int N, int mf, int mx, int nverts, const int32_t* cv, const int32_t* nbr,
                  const int32_t* extra, int P, const int32_t* order, PatchHost& out) {
    if (P <= 0 || P > 1024) return 1;
    const int npatch = (N + P - 1) / P;
    std::vector<int32_t> gen(N, -1), loc(N, 0), vgen(nverts, -1), vloc(nverts, 0);
    std::vector<std::vector<int32_t>> loads(npatch), verts(npatch);
    std::vector<int32_t> tpl_of(npatch);
    std::unordered_map<std::string, int> tpl_index;
    out.tpl_nbr.clear();
    out.tpl_extra.clear();
    out.tpl_vert.clear();
    out.tpl_nslope.clear();
    out.tpl_nown.clear();
    struct Local {
        std::vector<uint16_t> nbr, extra, vert;
        int n_slope, n_own;
    };
    std::vector<Local> tpls;
    int max_load = 0, max_slope = 0, max_vert = 0;
    std::vector<int32_t> own;
    for (int p = 0; p < npatch; ++p) {
        const int c0 = p * P, n_own = std::min(P, N - c0);
        std::vector<int32_t>& L = loads[p];
        L.clear();
        auto add = [&](int32_t c) {
            if (gen[c] != p) {
                gen[c] = p;
                loc[c] = (int32_t)L.size();
                L.push_back(c);
            }
        };
        // own cells in ascending id order: runs of consecutive ids fall on
        // consecutive threads (coalesced where the cell order allows)
        own.resize(n_own);
        for (int i = 0; i < n_own; ++i) own[i] = order ? order[c0 + i] : c0 + i;
        if (order) std::sort(own.begin(), own.end());
        for (int i = 0; i < n_own; ++i) add(own[i]);
        for (int i = 0; i < n_own; ++i)
            for (int k = 0; k < mf; ++k) {
                int32_t nb = nbr[(size_t)k * N + own[i]];
                if (nb >= 0) add(nb);
            }
        const int n_slope = (int)L.size();
        for (int ls = 0; ls < n_slope; ++ls) {
            const int32_t c = L[ls];
            for (int k = 0; k < mf; ++k) {
                int32_t nb = nbr[(size_t)k * N + c];
                if (nb >= 0) add(nb);
            }
            for (int j = 0; j < mx; ++j) {
                int32_t e = extra[(size_t)j * N + c];
                if (e < 0) break;
                add(e);
            }
        }
        Local T;
        T.n_slope = n_slope;
        T.n_own = n_own;
        T.nbr.assign((size_t)mf * n_slope, 0xFFFF);
        T.extra.assign((size_t)(mx > 0 ? mx : 1) * n_slope, 0xFFFF);
        T.vert.assign((size_t)mf * n_slope, 0);
        std::vector<int32_t>& V = verts[p];
        V.clear();
        for (int ls = 0; ls < n_slope; ++ls) {
            const int32_t c = L[ls];
            for (int k = 0; k < mf; ++k) {
                int32_t nb = nbr[(size_t)k * N + c];
                if (nb >= 0) T.nbr[(size_t)k * n_slope + ls] = (uint16_t)loc[nb];
                int32_t v = cv[(size_t)k * N + c];
                if (vgen[v] != p) {
                    vgen[v] = p;
                    vloc[v] = (int32_t)V.size();
                    V.push_back(v);
                }
                T.vert[(size_t)k * n_slope + ls] = (uint16_t)vloc[v];
            }
            for (int j = 0; j < mx; ++j) {
                int32_t e = extra[(size_t)j * N + c];
                if (e < 0) break;
                T.extra[(size_t)j * n_slope + ls] = (uint16_t)loc[e];
            }
        }
        if (L.size() >= 0xFFFF || V.size() >= 0xFFFF) return 2;
        max_load = std::max(max_load, (int)L.size());
        max_slope = std::max(max_slope, n_slope);
        max_vert = std::max(max_vert, (int)V.size());
        std::string key;
        key.reserve(16 + 2 * (T.nbr.size() + T.extra.size() + T.vert.size()));
        key.append((const char*)&n_slope, sizeof(int));
        key.append((const char*)&n_own, sizeof(int));
        key.append((const char*)T.nbr.data(), 2 * T.nbr.size());
        key.append((const char*)T.extra.data(), 2 * T.extra.size());
        key.append((const char*)T.vert.data(), 2 * T.vert.size());
        auto it = tpl_index.find(key);
        if (it == tpl_index.end()) {
            int id = (int)tpls.size();
            tpl_index.emplace(std::move(key), id);
            tpls.push_back(std::move(T));
            tpl_of[p] = id;
        } else {
            tpl_of[p] = it->second;
        }
    }
    // pad to multiples of 64 so LDS arrays of the kernel stay aligned
    auto up = [](int x) { return (x + 63) / 64 * 64; };
    out.P = P;
    out.expl = order ? 1 : 0;
    out.npatch = npatch;
    out.mf = mf;
    out.mx = mx > 0 ? mx : 1;
    out.max_load = up(max_load);
    out.max_slope = up(max_slope);
    out.max_vert = up(max_vert);
    const int ntpl = (int)tpls.size();
    out.ntpl = ntpl;
    out.tpl_id = tpl_of;
    out.n_load.resize(npatch);
    out.n_vert.resize(npatch);
    out.load_ids.assign((size_t)npatch * out.max_load, 0);
    out.vert_ids.assign((size_t)npatch * out.max_vert, 0);
    for (int p = 0; p < npatch; ++p) {
        out.n_load[p] = (int32_t)loads[p].size();
        out.n_vert[p] = (int32_t)verts[p].size();
        std::memcpy(&out.load_ids[(size_t)p * out.max_load], loads[p].data(),
                    sizeof(int32_t) * loads[p].size());
        std::memcpy(&out.vert_ids[(size_t)p * out.max_vert], verts[p].data(),
                    sizeof(int32_t) * verts[p].size());
    }
    const int S = out.max_slope;
    out.tpl_nbr.assign((size_t)ntpl * mf * S, 0xFFFF);
    out.tpl_extra.assign((size_t)ntpl * out.mx * S, 0xFFFF);
    out.tpl_vert.assign((size_t)ntpl * mf * S, 0);
    out.tpl_nslope.resize(ntpl);
    out.tpl_nown.resize(ntpl);
    for (int t = 0; t < ntpl; ++t) {
        const Local& T = tpls[t];
        out.tpl_nslope[t] = T.n_slope;
        out.tpl_nown[t] = T.n_own;
        for (int k = 0; k < mf; ++k)
            for (int ls = 0; ls < T.n_slope; ++ls) {
                out.tpl_nbr[((size_t)t * mf + k) * S + ls] = T.nbr[(size_t)k * T.n_slope + ls];
                out.tpl_vert[((size_t)t * mf + k) * S + ls] = T.vert[(size_t)k * T.n_slope + ls];
            }
        for (int j = 0; j < mx; ++j)
            for (int ls = 0; ls < T.n_slope; ++ls)
                out.tpl_extra[((size_t)t * out.mx + j) * S + ls] = T.extra[(size_t)j * T.n_slope + ls];
    }
    return 0;
}

int validate_patches(int N, int mf, int mx, int nverts, const int32_t* nbr, const int32_t* extra,
                     const int32_t* cv, const PatchHost& ph) {
    const int S = ph.max_slope;
    std::vector<uint8_t> owned(N, 0);
    for (int p = 0; p < ph.npatch; ++p) {
        const int t = ph.tpl_id[p];
        if (t < 0 || t >= ph.ntpl) return 10;
        const int ns = ph.tpl_nslope[t], no = ph.tpl_nown[t];
        const int nl = ph.n_load[p], nv = ph.n_vert[p];
        if (no > ph.P || ns > S || nl > ph.max_load || nv > ph.max_vert || ns > nl) return 11;
        const int32_t* L = &ph.load_ids[(size_t)p * ph.max_load];
        const int32_t* V = &ph.vert_ids[(size_t)p * ph.max_vert];
        for (int i = 0; i < nl; ++i)
            if (L[i] < 0 || L[i] >= N) return 12;
        for (int i = 0; i < nv; ++i)
            if (V[i] < 0 || V[i] >= nverts) return 13;
        if (!ph.expl && no != std::min(ph.P, N - p * ph.P)) return 11;
        for (int i = 0; i < no; ++i) {
            if (!ph.expl && L[i] != p * ph.P + i) return 14;
            if (owned[L[i]]++) return 14;  // owned twice
        }
        for (int ls = 0; ls < ns; ++ls) {
            const int32_t c = L[ls];
            for (int k = 0; k < mf; ++k) {
                const uint16_t l = ph.tpl_nbr[((size_t)t * mf + k) * S + ls];
                const int32_t g = nbr[(size_t)k * N + c];
                if (l == 0xFFFF) {
                    if (g >= 0) return 15;
                } else {
                    if (l >= nl || L[l] != g) return 16;
                    if (ls < no && l >= ns) return 17;  // flux reads slopes
                }
                const uint16_t lv = ph.tpl_vert[((size_t)t * mf + k) * S + ls];
                if (lv >= nv || V[lv] != cv[(size_t)k * N + c]) return 18;
            }
            for (int j = 0; j < mx; ++j) {
                const uint16_t l = ph.tpl_extra[((size_t)t * ph.mx + j) * S + ls];
                const int32_t g = extra[(size_t)j * N + c];
                if (g < 0) {
                    if (l != 0xFFFF) return 19;
                    break;
                }
                if (l >= nl || L[l] != g) return 20;
            }
        }
    }
    for (int c = 0; c < N; ++c)
        if (!owned[c]) return 21;  // owned by no patch
    return 0;
}

std::vector<int32_t> morton_cluster_order(int N, const double* cen) {
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < N; ++c) {
            lo[a] = std::min(lo[a], cen[(size_t)a * N + c]);
            hi[a] = std::max(hi[a], cen[(size_t)a * N + c]);
        }
    // one scale for both axes, so that clusters are square in space
    const double span = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), 1e-300);
    std::vector<uint64_t> key(N);
    for (int c = 0; c < N; ++c) {
        uint32_t q[2];
        for (int a = 0; a < 2; ++a) {
            const double u = (cen[(size_t)a * N + c] - lo[a]) / span;
            q[a] = (uint32_t)std::min(65535.0, std::max(0.0, u * 65536.0));
        }
        uint64_t code = 0;
        for (int b = 15; b >= 0; --b)
            code = (code << 2) | (((q[1] >> b) & 1u) << 1) | ((q[0] >> b) & 1u);
        key[c] = (code << 32) | (uint32_t)c;
    }
    std::sort(key.begin(), key.end());
    std::vector<int32_t> order(N);
    for (int c = 0; c < N; ++c) order[c] = (int32_t)(key[c] & 0xFFFFFFFFu);
    return order;
}

}  // namespace zj

// Host-only: build and validate the patch decomposition of a mesh without
// touching a device.  stats_out[6] = {P, npatch, ntemplates, max_load,
// max_slope, max_vert}.  Used by the CPU tests.
extern "C" int zj_patch_check(const zj_mesh_desc* d, int32_t P, int32_t* stats_out) {
    if (!d || !stats_out || d->n_cells <= 0) return ZJ_EINVAL;
    const int N = d->n_cells;
    std::vector<int32_t> ex_dummy(N, -1);
    const int mx = d->max_extra > 0 ? d->max_extra : 1;
    const int32_t* exh = d->max_extra > 0 ? d->cell_extra : ex_dummy.data();
    zj::PatchHost ph;
    if (zj::build_patches(N, d->max_faces, mx, d->n_verts, d->cell_vertices, d->cell_nbr, exh, P,
                          nullptr, ph))
        return ZJ_EINVAL;
    int rc = zj::validate_patches(N, d->max_faces, mx, d->n_verts, d->cell_nbr, exh,
                                  d->cell_vertices, ph);
    stats_out[0] = ph.P;
    stats_out[1] = ph.npatch;
    stats_out[2] = ph.ntpl;
    stats_out[3] = ph.max_load;
    stats_out[4] = ph.max_slope;
    stats_out[5] = ph.max_vert;
    return rc ? 100 + rc : ZJ_OK;
}
