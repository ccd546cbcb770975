// Internal interface between the Chorin projection (zj_chorin.hip) and the
// learned preconditioner (zj_mpnn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/zoomy_hip.h"

namespace zj {

// CG scalar slots and the scalar steps of k_cg_fin
enum { CG_B2 = 0, CG_RZ, CG_ALPHA, CG_BETA, CG_RR, CG_NSLOT = 8 };
enum { FIN_B2 = 0, FIN_RR0, FIN_RZ0, FIN_PQ, FIN_RR, FIN_RZ, FIN_RZ_FLEX };

// One level of the multigrid hierarchy (arrays of element size esz; h and b
// are FP64 averages on levels >= 1).
struct HLevel {
    int nx = 0, ny = 0;
    double dx = 0, dy = 0;
    void *diag = nullptr, *cE = nullptr, *cN = nullptr;
    void *x = nullptr, *f = nullptr, *r = nullptr;
    double *h = nullptr, *b = nullptr;
};

struct Hier {
    std::vector<HLevel> levels;
    void* mem = nullptr;
    size_t esz = 8;
    int alloc(int nx, int ny, double dx, double dy, int min_coarse, size_t esz);
    void release();
};

Hier& chorin_h32(zj_chorin* c);
int chorin_nx(zj_chorin* c);
int chorin_ny(zj_chorin* c);

// z = M r with the learned preconditioner (FP32 vectors on level 0 of the
// FP32 hierarchy); kernels return at once while *done != 0.
int mpnn_apply_f32(zj_mpnn* net, zj_chorin* c, const float* r, float* z, const int* done,
                   hipStream_t st);

}  // namespace zj
