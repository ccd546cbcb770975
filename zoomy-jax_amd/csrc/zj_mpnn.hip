// Learned message-passing preconditioner: forward, backward and Adam kernels
// on BF16 MFMA (gfx950 v_mfma_f32_16x16x32_bf16), behind zj_mpnn_* and the
// learned V-cycle of the Chorin pressure solve.
//
// Graph: 2x2 blocks of fine cells (node (a, b) = cells (2a + di, 2b + dj),
// slot s = 2 di + dj), 4-neighbour adjacency, zero outside the grid; a batch
// of B grids is stacked node-major.  Network (linear in the residual r):
//   h0 = bf16((r / diag) W_in)                        W_in  4 x 64   (FP32)
//   h_{l+1} = bf16(sigmoid(phi U_l) * ([h_l | agg(h_l)] W_l))       l = 0..2
//                                                     W_l 128 x 64, U_l 16 x 64
//   z = h3 W_out                                      W_out 64 x 4   (FP32)
// The layer products run on MFMA with BF16 operands and FP32 accumulation;
// weight gradients are MFMA products over the node dimension, reduced per
// block and then over blocks in a fixed order (deterministic).
//
// MFMA 16x16x32 operand layout (wave64): A row = lane & 15, k = 8 (lane >> 4)
// + e; B column = lane & 15, k = 8 (lane >> 4) + e; D column = lane & 15,
// row = 4 (lane >> 4) + i.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/zoomy_hip.h"
#include "zj_chorin.hpp"

namespace {

constexpr int HID = 64;
constexpr int NL = 3;
constexpr int NF = 16;
constexpr int OFF_WIN = 0;
constexpr int LAYER_STRIDE = 2 * HID * HID + NF * HID;  // W_l then U_l
constexpr int OFF_L0 = 4 * HID;
constexpr int OFF_WOUT = OFF_L0 + NL * LAYER_STRIDE;
constexpr int N_PARAMS = OFF_WOUT + HID * 4;
// bf16 weight copies per layer: forward-swizzled W (16 fragments x 64 lanes
// x 8), forward-swizzled U (4 x 64 x 8), row-major W (128 x 64)
constexpr int WF_SIZE = 16 * 64 * 8;
constexpr int UF_SIZE = 4 * 64 * 8;
constexpr int WR_SIZE = 2 * HID * HID;
constexpr int BF_LAYER = WF_SIZE + UF_SIZE + WR_SIZE;
constexpr int WG_BLOCKS = 256;  // blocks of the weight-gradient kernels
constexpr int PART_STRIDE = 2 * HID * HID + NF * HID;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// round to nearest even
__device__ __host__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ bf16x8 ld8(const bf16_t* p) {
    return __builtin_bit_cast(bf16x8, *(const uint4*)p);
}
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
// hardware conversion (v_cvt_pk_bf16_f32, round to nearest even)
__device__ __forceinline__ bf16x8 pack8(const float* v) {
    f32x8 x = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
    return __builtin_convertvector(x, bf16x8);
}
__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
    f32x4 x = {a, b, c, d};
    return __builtin_bit_cast(uint2, __builtin_convertvector(x, bf16x4));
}
// gate: logistic sigmoid with the hardware reciprocal
__device__ __forceinline__ float gate_of(float x) {
    return __builtin_amdgcn_rcpf(1.f + __expf(-x));
}
__device__ __forceinline__ void unpack8(const bf16_t* p, float* v) {
    uint4 u = *(const uint4*)p;
    v[0] = __uint_as_float(u.x << 16); v[1] = __uint_as_float(u.x & 0xFFFF0000u);
    v[2] = __uint_as_float(u.y << 16); v[3] = __uint_as_float(u.y & 0xFFFF0000u);
    v[4] = __uint_as_float(u.z << 16); v[5] = __uint_as_float(u.z & 0xFFFF0000u);
    v[6] = __uint_as_float(u.w << 16); v[7] = __uint_as_float(u.w & 0xFFFF0000u);
}

struct Grid {
    int B, nx, ny;     // cells per grid
    int nxn, nyn, Nn;  // nodes per grid, total nodes
};

__device__ __forceinline__ int node_nbrs(const Grid& g, int n, int* nb) {
    const int per = g.nxn * g.nyn;
    const int s = n / per, r = n - s * per;
    const int a = r / g.nyn, b = r - a * g.nyn;
    nb[0] = a > 0 ? n - g.nyn : -1;
    nb[1] = a + 1 < g.nxn ? n + g.nyn : -1;
    nb[2] = b > 0 ? n - 1 : -1;
    nb[3] = b + 1 < g.nyn ? n + 1 : -1;
    return 0;
}
// cell of slot q of node n
__device__ __forceinline__ int node_cell(const Grid& g, int n, int q) {
    const int per = g.nxn * g.nyn;
    const int s = n / per, r = n - s * per;
    const int a = r / g.nyn, b = r - a * g.nyn;
    return s * g.nx * g.ny + (2 * a + (q >> 1)) * g.ny + 2 * b + (q & 1);
}

// ───────────────────────── weights ─────────────────────────────────────────
__global__ void k_cast_weights(const float* __restrict__ P, bf16_t* __restrict__ wb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const float* W = P + OFF_L0 + l * LAYER_STRIDE;
    const float* U = W + 2 * HID * HID;
    bf16_t* WF = wb + l * BF_LAYER;
    bf16_t* UF = WF + WF_SIZE;
    bf16_t* WR = UF + UF_SIZE;
    if (t < WF_SIZE) {  // [nt][ks][lane][e]
        int e = t & 7, lane = (t >> 3) & 63, ks = (t >> 9) & 3, nt = t >> 11;
        int k = ks * 32 + (lane >> 4) * 8 + e, col = nt * 16 + (lane & 15);
        WF[t] = f2bf(W[k * HID + col]);
    }
    if (t < UF_SIZE) {  // [nt][lane][e], K padded to 32
        int e = t & 7, lane = (t >> 3) & 63, nt = t >> 9;
        int k = (lane >> 4) * 8 + e, col = nt * 16 + (lane & 15);
        UF[t] = k < NF ? f2bf(U[k * HID + col]) : (bf16_t)0;
    }
    if (t < WR_SIZE) WR[t] = f2bf(W[t]);
}

// ───────────────────────── forward ─────────────────────────────────────────
// phi (16 bf16 per node) from the fine-level stencil (diag, cE, cN)
__global__ void k_phi(Grid g, const float* __restrict__ diag, const float* __restrict__ cE,
                      const float* __restrict__ cN, bf16_t* __restrict__ phi) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.Nn) return;
    float v[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = node_cell(g, n, q);
        const int cl = c % (g.nx * g.ny);
        const int i = cl / g.ny, j = cl - i * g.ny;
        const float d = diag[c];
        float tot = fabsf(cE[c]) + fabsf(cN[c]);
        tot += i > 0 ? fabsf(cE[c - g.ny]) : 0.f;
        tot += j > 0 ? fabsf(cN[c - 1]) : 0.f;
        v[q] = cE[c] / d;
        v[4 + q] = cN[c] / d;
        v[8 + q] = tot / d;
    }
    v[12] = 1.f;
    v[13] = v[14] = v[15] = 0.f;
    uint4* o = (uint4*)(phi + (size_t)n * NF);
    bf16x8 a = pack8(v), b = pack8(v + 8);
    o[0] = __builtin_bit_cast(uint4, a);
    o[1] = __builtin_bit_cast(uint4, b);
}

// h0 = bf16((r / diag) W_in); rh kept for the W_in gradient
__global__ void k_embed(Grid g, const float* __restrict__ P, const float* __restrict__ r,
                        const float* __restrict__ diag, float* __restrict__ rh,
                        bf16_t* __restrict__ h0, const int* __restrict__ done) {
    if (done && *done) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.Nn) return;
    float x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = node_cell(g, n, q);
        x[q] = r[c] / diag[c];
    }
    if (rh) *(float4*)(rh + (size_t)n * 4) = make_float4(x[0], x[1], x[2], x[3]);
    const float* W = P + OFF_WIN;
    for (int k0 = 0; k0 < HID; k0 += 8) {
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + e;
            v[e] = x[0] * W[k] + x[1] * W[HID + k] + x[2] * W[2 * HID + k] + x[3] * W[3 * HID + k];
        }
        *(uint4*)(h0 + (size_t)n * HID + k0) = __builtin_bit_cast(uint4, pack8(v));
    }
}

// One gated MP layer; one wave per 16-node tile (grid-stride), weights held
// in registers.  y, gt (FP32) are stored for the backward when given.
__global__ __launch_bounds__(256) void k_layer_fwd(Grid g, const bf16_t* __restrict__ hin,
                                                   const bf16_t* __restrict__ phi,
                                                   const bf16_t* __restrict__ wl,
                                                   bf16_t* __restrict__ hout,
                                                   float* __restrict__ y_out,
                                                   float* __restrict__ g_out,
                                                   const int* __restrict__ done) {
    if (done && *done) return;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const bf16_t* WF = wl;
    const bf16_t* UF = wl + WF_SIZE;
    bf16x8 wf[4][4], uf[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) wf[nt][ks] = ld8(WF + ((nt * 4 + ks) * 64 + lane) * 8);
        uf[nt] = ld8(UF + (nt * 64 + lane) * 8);
    }
    const int ntiles = (g.Nn + 15) >> 4;
    for (int tile = wave; tile < ntiles; tile += nwaves) {
        const int n = tile * 16 + m;
        const bool valid = n < g.Nn;
        bf16x8 a[4], ap;
        if (valid) {
            const bf16_t* hs = hin + (size_t)n * HID;
            a[0] = ld8(hs + kq * 8);
            a[1] = ld8(hs + 32 + kq * 8);
            int nb[4];
            node_nbrs(g, n, nb);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (nb[q] < 0) continue;
                    float v[8];
                    unpack8(hin + (size_t)nb[q] * HID + half * 32 + kq * 8, v);
#pragma unroll
                    for (int e = 0; e < 8; ++e) s[e] += v[e];
                }
                a[2 + half] = pack8(s);
            }
            if (kq < 2) {
                ap = ld8(phi + (size_t)n * NF + kq * 8);
            } else {
                float z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                ap = pack8(z);
            }
        } else {
            float z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            a[0] = a[1] = a[2] = a[3] = ap = pack8(z);
        }
        // transposed product: D^T = W^T A^T, so lane (m, kq) holds output
        // channels nt*16 + 4 kq + i of node tile*16 + m
        f32x4 acc[4], gac[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
            gac[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) acc[nt] = mfma(wf[nt][ks], a[ks], acc[nt]);
            gac[nt] = mfma(uf[nt], ap, gac[nt]);
        }
        if (valid) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const size_t o = (size_t)n * HID + nt * 16 + kq * 4;
                f32x4 gt, hv;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    gt[i] = gate_of(gac[nt][i]);
                    hv[i] = gt[i] * acc[nt][i];
                }
                *(uint2*)(hout + o) = pack4(hv[0], hv[1], hv[2], hv[3]);
                if (y_out) {
                    *(float4*)(y_out + o) = make_float4(acc[nt][0], acc[nt][1], acc[nt][2],
                                                        acc[nt][3]);
                    *(float4*)(g_out + o) = make_float4(gt[0], gt[1], gt[2], gt[3]);
                }
            }
        }
    }
}

// z = h3 W_out (add != 0: z += ...)
__global__ void k_out(Grid g, const float* __restrict__ P, const bf16_t* __restrict__ h3,
                      float* __restrict__ z, int add, const int* __restrict__ done) {
    if (done && *done) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.Nn) return;
    const float* W = P + OFF_WOUT;
    float o[4] = {0, 0, 0, 0};
    for (int k0 = 0; k0 < HID; k0 += 8) {
        float v[8];
        unpack8(h3 + (size_t)n * HID + k0, v);
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] += v[e] * W[(k0 + e) * 4 + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = node_cell(g, n, q);
        z[c] = add ? z[c] + o[q] : o[q];
    }
}


// ───────────────────────── fused inference ─────────────────────────────────
// z (+)= N(r) for one tile of FT x FT nodes in one launch: embedding, the
// three layers and the output on LDS-resident activations with a 3-node
// halo (regions of side FT+6, FT+4, FT+2, FT), each layer's BF16 weights
// staged in LDS.  Same fragments, same MFMA sequence and same neighbour
// summation order as the layer-by-layer kernels, hence the same result.
constexpr int FT = 16;
constexpr int FR0 = FT + 6;
constexpr int FTH = 1024;       // 16 waves (4 per SIMD)
constexpr int HS = HID + 8;     // padded row (144 B): 16-byte fragment reads of
                                // 16 consecutive rows hit distinct banks
constexpr size_t FUSED_LDS = (size_t)FR0 * FR0 * HS * 2      // region 0 / 2
                             + (size_t)(FT + 4) * (FT + 4) * HS * 2  // region 1 / 3
                             + (size_t)FR0 * FR0 * NF * 2             // phi
                             + (size_t)(WF_SIZE + UF_SIZE) * 2;       // weights

__global__ __launch_bounds__(FTH) void k_mpnn_fused(Grid g, const float* __restrict__ P,
                                                    const bf16_t* __restrict__ wb,
                                                    const bf16_t* __restrict__ phi,
                                                    const float* __restrict__ r,
                                                    const float* __restrict__ diag,
                                                    float* __restrict__ z, int add,
                                                    const int* __restrict__ done) {
    if (done && *done) return;
    extern __shared__ __align__(16) unsigned char fl[];
    bf16_t* hA = (bf16_t*)fl;
    bf16_t* hB = hA + FR0 * FR0 * HS;
    bf16_t* sph = hB + (FT + 4) * (FT + 4) * HS;
    bf16_t* sw = sph + FR0 * FR0 * NF;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kq = lane >> 4;
    // this tile: nodes (a0 .. a0+FT-1, b0 ..) of grid `gs`
    const int tiles_x = (g.nxn + FT - 1) / FT, tiles_y = (g.nyn + FT - 1) / FT;
    const int per_grid = tiles_x * tiles_y;
    const int gs = blockIdx.x / per_grid, tr = blockIdx.x - gs * per_grid;
    const int a0 = (tr / tiles_y) * FT, b0 = (tr % tiles_y) * FT;
    const int node_base = gs * g.nxn * g.nyn;
    const int cell_base = gs * g.nx * g.ny;
    auto in_grid = [&](int a, int b) { return a >= 0 && a < g.nxn && b >= 0 && b < g.nyn; };
    // phi and the embedding of region 0 (origin a0-3, b0-3)
    for (int u = tid; u < FR0 * FR0; u += FTH) {
        const int ia = u / FR0, ib = u - ia * FR0;
        const int a = a0 - 3 + ia, b = b0 - 3 + ib;
        uint4* dp = (uint4*)(sph + u * NF);
        uint4* dh = (uint4*)(hA + u * HS);
        if (!in_grid(a, b)) {
            const uint4 zz = {0u, 0u, 0u, 0u};
            dp[0] = dp[1] = zz;
            for (int k = 0; k < HID / 8; ++k) dh[k] = zz;
            continue;
        }
        const int n = node_base + a * g.nyn + b;
        const uint4* sp = (const uint4*)(phi + (size_t)n * NF);
        dp[0] = sp[0];
        dp[1] = sp[1];
        float x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cell_base + (2 * a + (q >> 1)) * g.ny + 2 * b + (q & 1);
            x[q] = r[c] / diag[c];
        }
        const float* W = P + OFF_WIN;
        for (int k0 = 0; k0 < HID; k0 += 8) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + e;
                v[e] = x[0] * W[k] + x[1] * W[HID + k] + x[2] * W[2 * HID + k] +
                       x[3] * W[3 * HID + k];
            }
            dh[k0 / 8] = __builtin_bit_cast(uint4, pack8(v));
        }
    }
    bf16_t* hin = hA;
    bf16_t* hout = hB;
    for (int l = 0; l < NL; ++l) {
        // stage the layer's forward-swizzled W and U
        {
            const uint4* src = (const uint4*)(wb + l * BF_LAYER);
            uint4* dst = (uint4*)sw;
            for (int u = tid; u < (WF_SIZE + UF_SIZE) / 8; u += FTH) dst[u] = src[u];
        }
        __syncthreads();
        const int sin = FR0 - 2 * l, sout = sin - 2;        // region sides
        const int off_out = 3 - (l + 1);                    // origin of the output region
        const int nrow = sout * sout, ntile = (nrow + 15) / 16;
        // weight fragments are read from LDS per tile (holding all 20 in
        // registers would cost 80 VGPRs and halve the waves per SIMD)
        for (int t = wave; t < ntile; t += FTH / 64) {
            const int row = t * 16 + m;
            bf16x8 a[4], ap;
            if (row < nrow) {
                const int ia = row / sout, ib = row - ia * sout;
                const int ci = (ia + 1) * sin + (ib + 1);  // same node in the input region
                const bf16_t* hs = hin + ci * HS;
                a[0] = ld8(hs + kq * 8);
                a[1] = ld8(hs + 32 + kq * 8);
                const int nb[4] = {ci - sin, ci + sin, ci - 1, ci + 1};
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    float sacc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float v[8];
                        unpack8(hin + nb[q] * HS + half * 32 + kq * 8, v);
#pragma unroll
                        for (int e = 0; e < 8; ++e) sacc[e] += v[e];
                    }
                    a[2 + half] = pack8(sacc);
                }
                const int pi = (ia + l + 1) * FR0 + (ib + l + 1);
                if (kq < 2) {
                    ap = ld8(sph + pi * NF + kq * 8);
                } else {
                    float zz[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    ap = pack8(zz);
                }
            } else {
                float zz[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                a[0] = a[1] = a[2] = a[3] = ap = pack8(zz);
            }
            f32x4 acc[4], gac[4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
                gac[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks)
                    acc[nt] = mfma(ld8(sw + ((nt * 4 + ks) * 64 + lane) * 8), a[ks], acc[nt]);
                gac[nt] = mfma(ld8(sw + WF_SIZE + (nt * 64 + lane) * 8), ap, gac[nt]);
            }
            if (row < nrow) {
                const int ia = row / sout, ib = row - ia * sout;
                const bool ing = in_grid(a0 - off_out + ia, b0 - off_out + ib);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    uint2 pk = {0u, 0u};
                    if (ing) {
                        f32x4 hv;
#pragma unroll
                        for (int i = 0; i < 4; ++i) hv[i] = gate_of(gac[nt][i]) * acc[nt][i];
                        pk = pack4(hv[0], hv[1], hv[2], hv[3]);
                    }
                    *(uint2*)(hout + row * HS + nt * 16 + kq * 4) = pk;
                }
            }
        }
        __syncthreads();
        bf16_t* tmp = hin;
        hin = hout;
        hout = tmp;
    }
    // output region (side FT) is in hin
    for (int u = tid; u < FT * FT; u += FTH) {
        const int ia = u / FT, ib = u - ia * FT;
        const int a = a0 + ia, b = b0 + ib;
        if (!in_grid(a, b)) continue;
        const float* W = P + OFF_WOUT;
        float o[4] = {0, 0, 0, 0};
        for (int k0 = 0; k0 < HID; k0 += 8) {
            float v[8];
            unpack8(hin + u * HS + k0, v);
#pragma unroll
            for (int e = 0; e < 8; ++e)
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] += v[e] * W[(k0 + e) * 4 + q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cell_base + (2 * a + (q >> 1)) * g.ny + 2 * b + (q & 1);
            z[c] = add ? z[c] + o[q] : o[q];
        }
    }
}

// ───────────────────────── backward ────────────────────────────────────────
// dh3 = dz W_out^T (FP32 nodes x 64); do kept (nodes x 4) for the W_out grad
__global__ void k_out_bwd(Grid g, const float* __restrict__ P, const float* __restrict__ dz,
                          float* __restrict__ dov, float* __restrict__ dh) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.Nn) return;
    const float* W = P + OFF_WOUT;
    float d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = dz[node_cell(g, n, q)];
    *(float4*)(dov + (size_t)n * 4) = make_float4(d[0], d[1], d[2], d[3]);
    for (int k = 0; k < HID; k += 4) {
        float4 v;
        v.x = d[0] * W[k * 4] + d[1] * W[k * 4 + 1] + d[2] * W[k * 4 + 2] + d[3] * W[k * 4 + 3];
        v.y = d[0] * W[k * 4 + 4] + d[1] * W[k * 4 + 5] + d[2] * W[k * 4 + 6] + d[3] * W[k * 4 + 7];
        v.z = d[0] * W[k * 4 + 8] + d[1] * W[k * 4 + 9] + d[2] * W[k * 4 + 10] + d[3] * W[k * 4 + 11];
        v.w = d[0] * W[k * 4 + 12] + d[1] * W[k * 4 + 13] + d[2] * W[k * 4 + 14] + d[3] * W[k * 4 + 15];
        *(float4*)(dh + (size_t)n * HID + k) = v;
    }
}

// dy = bf16(dh * g), dp = bf16(dh * y * g (1 - g)); dA = dy W^T (FP32 nodes x
// 128) on MFMA.  One wave per 16-node tile.
__global__ __launch_bounds__(256) void k_layer_bwd(Grid g, const float* __restrict__ dh,
                                                   const float* __restrict__ y,
                                                   const float* __restrict__ gt,
                                                   const bf16_t* __restrict__ wl,
                                                   float* __restrict__ dA,
                                                   bf16_t* __restrict__ dyb,
                                                   bf16_t* __restrict__ dpb) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const bf16_t* WR = wl + WF_SIZE + UF_SIZE;
    const int ntiles = (g.Nn + 15) >> 4;
    for (int tile = wave; tile < ntiles; tile += nwaves) {
        const int n = tile * 16 + m;
        const bool valid = n < g.Nn;
        bf16x8 a[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            float dyv[8], dpv[8];
            const int k0 = ks * 32 + kq * 8;
            if (valid) {
                const size_t o = (size_t)n * HID + k0;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float d = dh[o + e], gg = gt[o + e];
                    dyv[e] = d * gg;
                    dpv[e] = d * y[o + e] * gg * (1.f - gg);
                }
                bf16x8 py = pack8(dyv), pp = pack8(dpv);
                *(uint4*)(dyb + o) = __builtin_bit_cast(uint4, py);
                *(uint4*)(dpb + o) = __builtin_bit_cast(uint4, pp);
                a[ks] = py;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) dyv[e] = 0.f;
                a[ks] = pack8(dyv);
            }
        }
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            const int col = nt * 16 + m;  // input channel of the layer
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = mfma(a[ks], ld8(WR + (size_t)col * HID + ks * 32 + kq * 8), acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = tile * 16 + kq * 4 + i;
                if (row < g.Nn) dA[(size_t)row * 2 * HID + col] = acc[i];
            }
        }
    }
}

// dh_in = dA[:, :64] + sum over neighbours of dA[nb, 64:]
__global__ void k_dh_gather(Grid g, const float* __restrict__ dA, float* __restrict__ dh) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = t >> 4, k = (t & 15) * 4;
    if (n >= g.Nn) return;
    int nb[4];
    node_nbrs(g, n, nb);
    float4 s = *(const float4*)(dA + (size_t)n * 2 * HID + k);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (nb[q] < 0) continue;
        float4 v = *(const float4*)(dA + (size_t)nb[q] * 2 * HID + HID + k);
        s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
    *(float4*)(dh + (size_t)n * HID + k) = s;
}

// Per-block partial sums of dW = A^T dy (128 x 64) and dU = phi^T dp
// (16 x 64) over 32-node chunks staged in LDS (A = [h | agg(h)]).
__global__ __launch_bounds__(256) void k_wgrad(Grid g, const bf16_t* __restrict__ hin,
                                               const bf16_t* __restrict__ phi,
                                               const bf16_t* __restrict__ dyb,
                                               const bf16_t* __restrict__ dpb,
                                               float* __restrict__ part) {
    __shared__ bf16_t sA[32][2 * HID + 8];
    __shared__ bf16_t sY[32][HID + 8];
    __shared__ bf16_t sP[32][HID + 8];
    __shared__ bf16_t sF[32][NF + 8];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int m = lane & 15, kq = lane >> 4;
    f32x4 accW[8], accU;
#pragma unroll
    for (int t = 0; t < 8; ++t) accW[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    accU = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nchunks = (g.Nn + 31) >> 5;
    for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        // stage: 32 nodes x (8 groups of 8 self, 8 of agg, 8 of dy, 8 of dp, 2 of phi)
        for (int u = tid; u < 32 * 34; u += 256) {
            const int nl = u / 34, part_id = u - nl * 34;
            const int n = ch * 32 + nl;
            float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            bf16x8 val = pack8(v);
            if (n < g.Nn) {
                if (part_id < 8) {
                    val = ld8(hin + (size_t)n * HID + part_id * 8);
                } else if (part_id < 16) {
                    const int k0 = (part_id - 8) * 8;
                    int nb[4];
                    node_nbrs(g, n, nb);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (nb[q] < 0) continue;
                        float t8[8];
                        unpack8(hin + (size_t)nb[q] * HID + k0, t8);
#pragma unroll
                        for (int e = 0; e < 8; ++e) v[e] += t8[e];
                    }
                    val = pack8(v);
                } else if (part_id < 24) {
                    val = ld8(dyb + (size_t)n * HID + (part_id - 16) * 8);
                } else if (part_id < 32) {
                    val = ld8(dpb + (size_t)n * HID + (part_id - 24) * 8);
                } else {
                    val = ld8(phi + (size_t)n * NF + (part_id - 32) * 8);
                }
            }
            bf16_t* dst;
            if (part_id < 16) dst = &sA[nl][part_id * 8];
            else if (part_id < 24) dst = &sY[nl][(part_id - 16) * 8];
            else if (part_id < 32) dst = &sP[nl][(part_id - 24) * 8];
            else dst = &sF[nl][(part_id - 32) * 8];
            const uint4 uv = __builtin_bit_cast(uint4, val);
            dst[0] = (bf16_t)(uv.x & 0xFFFF); dst[1] = (bf16_t)(uv.x >> 16);
            dst[2] = (bf16_t)(uv.y & 0xFFFF); dst[3] = (bf16_t)(uv.y >> 16);
            dst[4] = (bf16_t)(uv.z & 0xFFFF); dst[5] = (bf16_t)(uv.z >> 16);
            dst[6] = (bf16_t)(uv.w & 0xFFFF); dst[7] = (bf16_t)(uv.w >> 16);
        }
        __syncthreads();
        // B fragments (k = node 8 kq + e, col = m): dy and dp
        bf16x8 bY[4], bP;
        {
            uint4 u;
            for (int nt = 0; nt < 4; ++nt) {
                bf16_t t8[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) t8[e] = sY[kq * 8 + e][nt * 16 + m];
                memcpy(&u, t8, 16);
                bY[nt] = __builtin_bit_cast(bf16x8, u);
            }
            bf16_t t8[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) t8[e] = sP[kq * 8 + e][w * 16 + m];
            memcpy(&u, t8, 16);
            bP = __builtin_bit_cast(bf16x8, u);
        }
        // this wave: W rows (input channels) of tiles mt = 2w, 2w + 1
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const int mt = 2 * w + tt;
            bf16_t t8[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) t8[e] = sA[kq * 8 + e][mt * 16 + m];
            uint4 u;
            memcpy(&u, t8, 16);
            const bf16x8 aA = __builtin_bit_cast(bf16x8, u);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) accW[tt * 4 + nt] = mfma(aA, bY[nt], accW[tt * 4 + nt]);
        }
        {
            bf16_t t8[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) t8[e] = sF[kq * 8 + e][m];
            uint4 u;
            memcpy(&u, t8, 16);
            accU = mfma(__builtin_bit_cast(bf16x8, u), bP, accU);
        }
        __syncthreads();
    }
    float* pb = part + (size_t)blockIdx.x * PART_STRIDE;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        const int mt = 2 * w + tt;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                pb[(mt * 16 + kq * 4 + i) * HID + nt * 16 + m] = accW[tt * 4 + nt][i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) pb[2 * HID * HID + (kq * 4 + i) * HID + w * 16 + m] = accU[i];
}

// Per-block partials of X^T Y for X (nodes x 4) and Y (nodes x 64) given as
// FP32 (W_in: X = rh, Y = dh0; W_out: X^T Y with X = h3 (bf16), Y = do).
__global__ __launch_bounds__(256) void k_small_grad(Grid g, const float* __restrict__ x4,
                                                    const float* __restrict__ y64,
                                                    const bf16_t* __restrict__ h64,
                                                    float* __restrict__ part) {
    // thread t owns output (q = t >> 6, k = t & 63): sum_n x4[n][q] * y64[n][k]
    // (h64 != NULL: sum_n h64[n][k] * x4[n][q], the W_out layout k * 4 + q)
    const int t = threadIdx.x, q = t >> 6, k = t & 63;
    float s = 0.f;
    const int per = (g.Nn + gridDim.x - 1) / gridDim.x;
    const int n0 = blockIdx.x * per, n1 = min(g.Nn, n0 + per);
    for (int n = n0; n < n1; ++n) {
        const float xv = x4[(size_t)n * 4 + q];
        const float yv = h64 ? bf2f(h64[(size_t)n * HID + k]) : y64[(size_t)n * HID + k];
        s += xv * yv;
    }
    float* pb = part + (size_t)blockIdx.x * PART_STRIDE;
    if (h64) pb[k * 4 + q] = s;
    else pb[q * HID + k] = s;
}

// grad[off + i] += sum over blocks of part[b][i] (fixed order)
__global__ void k_reduce_grad(const float* __restrict__ part, int nblocks, int count,
                              float* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * PART_STRIDE + i];
    grad[i] += s;
}

// dr = (dh0 W_in^T) / diag per cell
__global__ void k_embed_bwd(Grid g, const float* __restrict__ P, const float* __restrict__ dh0,
                            const float* __restrict__ diag, float* __restrict__ dr) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.Nn) return;
    const float* W = P + OFF_WIN;
    float s[4] = {0, 0, 0, 0};
    for (int k = 0; k < HID; ++k) {
        const float d = dh0[(size_t)n * HID + k];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += d * W[q * HID + k];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = node_cell(g, n, q);
        dr[c] = s[q] / diag[c];
    }
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ gr,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_fillf(int n, float v, float* __restrict__ o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = v;
}

inline int nb256(long n) { return (int)((n + 255) / 256); }

}  // namespace

// ═════════════════════════════ host side ═══════════════════════════════════
#define HIPCHK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "zoomy_hip: %s failed: %s\n", #x,         \
                    hipGetErrorString(e_));                           \
            return ZJ_EHIP;                                           \
        }                                                             \
    } while (0)

struct MpnnCache {
    size_t cap = 0;  // nodes
    void* mem = nullptr;
    float* rh = nullptr;
    bf16_t* h[NL + 1] = {};
    float* y[NL] = {};
    float* g[NL] = {};
};

struct zj_mpnn {
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
    bf16_t* wb = nullptr;
    int t = 0;
    // work arrays (grown on demand)
    size_t cap = 0;
    void* mem = nullptr;
    bf16_t* phi = nullptr;
    float *dh = nullptr, *dA = nullptr, *dov = nullptr;
    bf16_t *dyb = nullptr, *dpb = nullptr;
    float* part = nullptr;
    MpnnCache cache[3];  // 0, 1: training slots; 2: inference
    Grid phi_grid{};
    const void* phi_key = nullptr;
    bool fused = true;      // fused single-launch inference
    bool lds_attr = false;  // dynamic LDS limit raised for the fused kernel
};

namespace {

Grid make_grid(int B, int nx, int ny) {
    Grid g;
    g.B = B; g.nx = nx; g.ny = ny;
    g.nxn = nx / 2; g.nyn = ny / 2;
    g.Nn = B * g.nxn * g.nyn;
    return g;
}

int grow_cache(MpnnCache& c, size_t nodes, bool train) {
    if (c.cap >= nodes && (!train || c.y[0])) return ZJ_OK;
    if (c.mem) (void)hipFree(c.mem);
    c = MpnnCache();
    size_t bytes = nodes * 4 * sizeof(float) + (NL + 1) * nodes * HID * sizeof(bf16_t) +
                   (train ? 2 * NL * nodes * HID * sizeof(float) : 0) + 64 * 256;
    if (hipMalloc(&c.mem, bytes) != hipSuccess) return ZJ_ENOMEM;
    char* p = (char*)c.mem;
    auto take = [&](size_t b) { void* r = p; p += (b + 255) & ~(size_t)255; return r; };
    c.rh = (float*)take(nodes * 4 * sizeof(float));
    for (int l = 0; l <= NL; ++l) c.h[l] = (bf16_t*)take(nodes * HID * sizeof(bf16_t));
    if (train)
        for (int l = 0; l < NL; ++l) {
            c.y[l] = (float*)take(nodes * HID * sizeof(float));
            c.g[l] = (float*)take(nodes * HID * sizeof(float));
        }
    c.cap = nodes;
    return ZJ_OK;
}

int grow_work(zj_mpnn* net, size_t nodes) {
    if (net->cap >= nodes) return ZJ_OK;
    if (net->mem) (void)hipFree(net->mem);
    net->mem = nullptr;
    net->cap = 0;
    size_t bytes = nodes * NF * 2 + nodes * HID * 4 + nodes * 2 * HID * 4 + nodes * 4 * 4 +
                   2 * nodes * HID * 2 + 64 * 256;
    if (hipMalloc(&net->mem, bytes) != hipSuccess) return ZJ_ENOMEM;
    char* p = (char*)net->mem;
    auto take = [&](size_t b) { void* r = p; p += (b + 255) & ~(size_t)255; return r; };
    net->phi = (bf16_t*)take(nodes * NF * 2);
    net->dh = (float*)take(nodes * HID * 4);
    net->dA = (float*)take(nodes * 2 * HID * 4);
    net->dov = (float*)take(nodes * 4 * 4);
    net->dyb = (bf16_t*)take(nodes * HID * 2);
    net->dpb = (bf16_t*)take(nodes * HID * 2);
    net->cap = nodes;
    net->phi_key = nullptr;
    return ZJ_OK;
}

int layer_grid(const Grid& g) {
    const int ntiles = (g.Nn + 15) / 16;
    return std::max(1, std::min((ntiles + 3) / 4, 2048));
}

// z (=|+=) N(r) on grid g with fine-level stencil (diag, cE, cN); slot: cache
// index (0, 1 keep the training activations; 2: inference)
int forward(zj_mpnn* net, const Grid& g, const float* r, const float* diag, const float* cE,
            const float* cN, float* z, int add, int slot, const int* done, hipStream_t st) {
    const bool train = slot < 2;
    if (int rc = grow_work(net, g.Nn)) return rc;
    MpnnCache& C = net->cache[slot];
    if (int rc = grow_cache(C, g.Nn, train)) return rc;
    if (net->phi_key != (const void*)diag || net->phi_grid.Nn != g.Nn) {
        k_phi<<<nb256(g.Nn), 256, 0, st>>>(g, diag, cE, cN, net->phi);
        net->phi_key = diag;
        net->phi_grid = g;
    }
    if (!train && net->fused) {
        const int tiles = g.B * ((g.nxn + FT - 1) / FT) * ((g.nyn + FT - 1) / FT);
        k_mpnn_fused<<<tiles, FTH, FUSED_LDS, st>>>(g, net->P, net->wb, net->phi, r, diag, z, add,
                                                     done);
        HIPCHK(hipGetLastError());
        return ZJ_OK;
    }
    k_embed<<<nb256(g.Nn), 256, 0, st>>>(g, net->P, r, diag, train ? C.rh : nullptr, C.h[0],
                                          done);
    const int lg = layer_grid(g);
    for (int l = 0; l < NL; ++l)
        k_layer_fwd<<<lg, 256, 0, st>>>(g, C.h[l], net->phi, net->wb + l * BF_LAYER, C.h[l + 1],
                                         train ? C.y[l] : nullptr, train ? C.g[l] : nullptr,
                                         done);
    k_out<<<nb256(g.Nn), 256, 0, st>>>(g, net->P, C.h[NL], z, add, done);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

// accumulate parameter gradients of sum(dz * N(r)) (activations of `slot`);
// dr (may be NULL) receives the input gradient
int backward(zj_mpnn* net, const Grid& g, const float* dz, const float* diag, float* dr, int slot,
             hipStream_t st) {
    MpnnCache& C = net->cache[slot];
    if (slot > 1 || C.cap < (size_t)g.Nn || !C.y[0]) return ZJ_EINVAL;
    const int nblk = std::min(WG_BLOCKS, std::max(1, (g.Nn + 31) / 32));
    const int sblk = std::min(WG_BLOCKS, std::max(1, g.Nn / 64));
    k_out_bwd<<<nb256(g.Nn), 256, 0, st>>>(g, net->P, dz, net->dov, net->dh);
    k_small_grad<<<sblk, 256, 0, st>>>(g, net->dov, nullptr, C.h[NL], net->part);
    k_reduce_grad<<<1, 256, 0, st>>>(net->part, sblk, HID * 4, net->G + OFF_WOUT);
    const int lg = layer_grid(g);
    for (int l = NL - 1; l >= 0; --l) {
        const bf16_t* wl = net->wb + l * BF_LAYER;
        k_layer_bwd<<<lg, 256, 0, st>>>(g, net->dh, C.y[l], C.g[l], wl, net->dA, net->dyb,
                                         net->dpb);
        k_wgrad<<<nblk, 256, 0, st>>>(g, C.h[l], net->phi, net->dyb, net->dpb, net->part);
        k_reduce_grad<<<nb256(PART_STRIDE), 256, 0, st>>>(net->part, nblk, PART_STRIDE,
                                                          net->G + OFF_L0 + l * LAYER_STRIDE);
        k_dh_gather<<<nb256((long)g.Nn * 16), 256, 0, st>>>(g, net->dA, net->dh);
    }
    k_small_grad<<<sblk, 256, 0, st>>>(g, C.rh, net->dh, nullptr, net->part);
    k_reduce_grad<<<1, 256, 0, st>>>(net->part, sblk, 4 * HID, net->G + OFF_WIN);
    if (dr) k_embed_bwd<<<nb256(g.Nn), 256, 0, st>>>(g, net->P, net->dh, diag, dr);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int cast_weights(zj_mpnn* net, hipStream_t st) {
    k_cast_weights<<<dim3(nb256(WF_SIZE), NL), 256, 0, st>>>(net->P, net->wb);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

}  // namespace

namespace zj {

int mpnn_forward_f32(zj_mpnn* net, int nx, int ny, const float* r, const float* diag,
                     const float* cE, const float* cN, float* z, int add, int slot,
                     const int* done, hipStream_t st) {
    return forward(net, make_grid(1, nx, ny), r, diag, cE, cN, z, add, slot, done, st);
}

int mpnn_backward_f32(zj_mpnn* net, int nx, int ny, const float* dz, const float* diag,
                      float* dr, int slot, hipStream_t st) {
    return backward(net, make_grid(1, nx, ny), dz, diag, dr, slot, st);
}

void mpnn_invalidate_phi(zj_mpnn* net) {
    if (net) net->phi_key = nullptr;
}

}  // namespace zj

extern "C" {

int zj_mpnn_create(zj_mpnn** out) {
    if (!out) return ZJ_EINVAL;
    zj_mpnn* n = new (std::nothrow) zj_mpnn();
    if (!n) return ZJ_ENOMEM;
    const size_t pb = (size_t)N_PARAMS * sizeof(float);
    if (hipMalloc((void**)&n->P, pb) != hipSuccess || hipMalloc((void**)&n->G, pb) != hipSuccess ||
        hipMalloc((void**)&n->M, pb) != hipSuccess || hipMalloc((void**)&n->V, pb) != hipSuccess ||
        hipMalloc((void**)&n->wb, (size_t)NL * BF_LAYER * sizeof(bf16_t)) != hipSuccess ||
        hipMalloc((void**)&n->part, (size_t)WG_BLOCKS * PART_STRIDE * sizeof(float)) !=
            hipSuccess) {
        zj_mpnn_destroy(n);
        return ZJ_ENOMEM;
    }
    if (hipFuncSetAttribute((const void*)k_mpnn_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)FUSED_LDS) != hipSuccess) {
        zj_mpnn_destroy(n);
        return ZJ_EHIP;
    }
    (void)hipMemset(n->P, 0, pb);
    (void)hipMemset(n->G, 0, pb);
    (void)hipMemset(n->M, 0, pb);
    (void)hipMemset(n->V, 0, pb);
    *out = n;
    return ZJ_OK;
}

int zj_mpnn_invalidate_features(zj_mpnn* n) {
    if (!n) return ZJ_EINVAL;
    n->phi_key = nullptr;
    return ZJ_OK;
}

int zj_mpnn_set_fused(zj_mpnn* n, int32_t on) {
    if (!n) return ZJ_EINVAL;
    n->fused = on != 0;
    return ZJ_OK;
}

int zj_mpnn_destroy(zj_mpnn* n) {
    if (!n) return ZJ_OK;
    (void)hipDeviceSynchronize();
    for (void* p : {(void*)n->P, (void*)n->G, (void*)n->M, (void*)n->V, (void*)n->wb,
                    (void*)n->part, n->mem, n->cache[0].mem, n->cache[1].mem, n->cache[2].mem})
        if (p) (void)hipFree(p);
    delete n;
    return ZJ_OK;
}

int64_t zj_mpnn_param_count(void) { return N_PARAMS; }

int zj_mpnn_set_params(zj_mpnn* n, const float* host, int64_t count) {
    if (!n || !host || count != N_PARAMS) return ZJ_EINVAL;
    HIPCHK(hipMemcpy(n->P, host, (size_t)count * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(n->M, 0, (size_t)count * sizeof(float)));
    HIPCHK(hipMemset(n->V, 0, (size_t)count * sizeof(float)));
    n->t = 0;
    int rc = cast_weights(n, nullptr);
    HIPCHK(hipDeviceSynchronize());
    return rc;
}

int zj_mpnn_get_params(zj_mpnn* n, float* host, int64_t count) {
    if (!n || !host || count != N_PARAMS) return ZJ_EINVAL;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(host, n->P, (size_t)count * sizeof(float), hipMemcpyDeviceToHost));
    return ZJ_OK;
}

int zj_mpnn_get_grad(zj_mpnn* n, float* host, int64_t count) {
    if (!n || !host || count != N_PARAMS) return ZJ_EINVAL;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(host, n->G, (size_t)count * sizeof(float), hipMemcpyDeviceToHost));
    return ZJ_OK;
}

int zj_mpnn_zero_grad(zj_mpnn* n, void* stream) {
    if (!n) return ZJ_EINVAL;
    k_fillf<<<nb256(N_PARAMS), 256, 0, (hipStream_t)stream>>>(N_PARAMS, 0.f, n->G);
    HIPCHK(hipGetLastError());
    return ZJ_OK;
}

int zj_mpnn_adam(zj_mpnn* n, float lr, float beta1, float beta2, float eps, void* stream) {
    if (!n || !(lr > 0)) return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    n->t += 1;
    const float c1 = 1.f - powf(beta1, (float)n->t), c2 = 1.f - powf(beta2, (float)n->t);
    k_adam<<<nb256(N_PARAMS), 256, 0, st>>>(N_PARAMS, n->P, n->G, n->M, n->V, lr, beta1, beta2,
                                            eps, c1, c2);
    HIPCHK(hipGetLastError());
    return cast_weights(n, st);
}

int zj_mpnn_forward(zj_mpnn* n, int32_t B, int32_t nx, int32_t ny, const float* r,
                    const float* diag, const float* cE, const float* cN, float* z, int32_t add,
                    int32_t slot, void* stream) {
    if (!n || B < 1 || nx < 2 || ny < 2 || (nx & 1) || (ny & 1) || !r || !diag || !cE || !cN ||
        !z || slot < 0 || slot > 2)
        return ZJ_EINVAL;
    return forward(n, make_grid(B, nx, ny), r, diag, cE, cN, z, add, slot, nullptr,
                   (hipStream_t)stream);
}

int zj_mpnn_time_forward(zj_mpnn* n, int32_t nx, int32_t ny, const float* r, const float* diag,
                         const float* cE, const float* cN, float* z, int32_t reps,
                         float* ms_out, void* stream) {
    if (!n || nx < 2 || ny < 2 || (nx & 1) || (ny & 1) || !r || !diag || !cE || !cN || !z ||
        reps < 1 || !ms_out)
        return ZJ_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const Grid g = make_grid(1, nx, ny);
    n->phi_key = nullptr;
    if (int rc = forward(n, g, r, diag, cE, cN, z, 0, 2, nullptr, st)) return rc;  // features
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int k = 0; k < reps; ++k)
        if (int rc = forward(n, g, r, diag, cE, cN, z, 0, 2, nullptr, st)) return rc;
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = ms / (float)reps;
    return ZJ_OK;
}

int zj_mpnn_backward(zj_mpnn* n, int32_t B, int32_t nx, int32_t ny, const float* dz,
                     const float* diag, float* dr, int32_t slot, void* stream) {
    if (!n || B < 1 || nx < 2 || ny < 2 || (nx & 1) || (ny & 1) || !dz || !diag || slot < 0 ||
        slot > 1)
        return ZJ_EINVAL;
    return backward(n, make_grid(B, nx, ny), dz, diag, dr, slot, (hipStream_t)stream);
}

}  // extern "C"
