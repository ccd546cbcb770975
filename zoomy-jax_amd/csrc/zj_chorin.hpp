// Internal interface between the Chorin projection (zj_chorin.hip) and the
// learned preconditioner (zj_mpnn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/zoomy_hip.h"

namespace zj {

// CG scalar slots (CG_RZ and CG_RZ2 hold r.z of even and odd iterations)
// and the scalar steps of k_fin (FIN_NONE: a reduction folded by the kernel
// that uses it)
enum { CG_B2 = 0, CG_RZ, CG_RR, CG_RZ2, CG_NSLOT = 8 };
enum { FIN_B2 = 0, FIN_RR0, FIN_RR, FIN_NONE };

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
    bool fused = true;  // fused multigrid kernels when nu == 1
    int alloc(int nx, int ny, double dx, double dy, int min_coarse, size_t esz);
    void release();
};

// precision (ZJ_F64 / ZJ_F32), cell count, dimension and numerics of an
// FV solver context
int solver_dtype(const zj_solver* s);
int solver_cells(const zj_solver* s);
int solver_dim(const zj_solver* s);
zj_params solver_params(const zj_solver* s);

// Network of the learned preconditioner on an nx x ny grid with fine-level
// FP32 stencil (diag, cE, cN): z (+)= N(r).  slot 0, 1 keep the activations
// for mpnn_backward_f32; slot 2 is inference.  Kernels return at once while
// *done != 0.
int mpnn_forward_f32(zj_mpnn* net, int nx, int ny, const float* r, const float* diag,
                     const float* cE, const float* cN, float* z, int add, int slot,
                     const int* done, hipStream_t st);
// accumulates the parameter gradients of sum(dz * N(r)); dr (may be NULL)
// receives the input gradient
int mpnn_backward_f32(zj_mpnn* net, int nx, int ny, const float* dz, const float* diag,
                      float* dr, int slot, hipStream_t st);
// the operator features must be recomputed (stencil rebuilt in place)
void mpnn_invalidate_phi(zj_mpnn* net);

}  // namespace zj
