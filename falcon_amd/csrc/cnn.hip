// Learned V-cycle correction network on MFMA (gfx950): forward (fused, all
// three layers in LDS), data and weight gradients (implicit GEMM on
// v_mfma_f32_16x16x32_bf16 / 16x16x16_bf16), mean-squared-error loss and Adam.
//
// MFMA operand images (64 lanes, lane l = 16 g + n):
//   16x16x32: A lane l holds A[n][8g .. 8g+7], B lane l holds B[8g .. 8g+7][n]
//   16x16x16: A lane l holds A[n][4g .. 4g+3], B lane l holds B[4g .. 4g+3][n]
//   D (16 x 16): lane l holds D[4g .. 4g+3][n]
// A convolution is computed as D = W^T X: rows = output channels, columns =
// 16 voxels of one k-line, K = (tap, input channel) pairs, so each lane ends
// with 4 consecutive output channels of one voxel (one 8-byte store of
// channel-fastest activations).
//
// Layouts: images x/out [img][4096] float; activations and their gradients
// [img][4096][16] bfloat16 (channel fastest); voxel v = (i*16 + j)*16 + k.
// LDS planes: x [18][18] with a zero halo; forward activations [16][16][16]
// (k-shifts by DPP, out-of-block lines skipped); gradient kernels [18][18][16]
// with a zero halo.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "cnn.hpp"

namespace falcon {

#define CCHECK(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int PL = 18 * 18;                 // plane with halo
constexpr int OFF_W1 = 0, OFF_B1 = 432, OFF_W2 = 448, OFF_B2 = 7360, OFF_W3 = 7376, OFF_B3 = 7808;
// packed operand images of the 16x16x32 form, 64 lanes x 8 bf16 per MFMA step:
// P1 layer-1 forward (1 step), Q3 layer-3 data gradient (1), Q2 layer-2 data gradient (14)
constexpr int PK_P1 = 0, PK_Q3 = 64, PK_Q2 = 128, PK_TOTAL = PK_Q2 + 14 * 64;

// forward operand images (bf16 elements, lane-major):
//   L1  [3 di][64] x 4: 16x16x16 A, rows co, K = the 9 (dj,dk) taps of plane offset di
//   L2A [9 (di,dj)][64] x 8: 16x16x32 A, rows co, K = (dk = -1: ci 0..15, dk = +1: ci 0..15)
//   L2B [9 (di,dj)][64] x 4: 16x16x16 A, rows co, K = ci of tap dk = 0
//   L3  [3 di][64] x 4: 16x16x16 A, rows r = 4 (dj+1) + (dk+1) (r % 4 < 3, r < 12), K = ci
constexpr int PF_L1 = 0, PF_L2A = PF_L1 + 3 * 64 * 4, PF_L2B = PF_L2A + 9 * 64 * 8, PF_L3 = PF_L2B + 9 * 64 * 4,
              PF_TOTAL = PF_L3 + 3 * 64 * 4;
// fused forward LDS carve (uint16 elements): h1 and h2 rings of 3 planes, each
// [18 j][4 channel groups][18 k][4 c] with zero halo rows and voxels (a lane's
// 8-byte access is consecutive to its neighbour's: no bank conflicts); x ring
// of 3 planes [18][18] with a zero halo
constexpr int HGRP = 18 * 4, HROW = 4 * HGRP, HPL = 18 * HROW;
constexpr int kFwdWaves = 8;                // wavefronts per image (16 / kFwdWaves k-lines each)
constexpr int F_H1 = 0, F_H2 = 3 * HPL, F_X = 6 * HPL, F_TOTAL = F_X + 3 * PL;
// single-layer LDS carve: ring of 3 input planes + zero plane
constexpr int S_IN = 0, S_Z = 3 * PL * 16, S_TOTAL = 4 * PL * 16;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u = u + 0x7fffu + ((u >> 16) & 1u);   // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ f4 mfma(us8 a, us8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
                                                   0);
}
// 16x16x16 form (K = 16 input channels of one tap): lane l = 16g + n holds
// A[n][4g .. 4g+3] and B[4g .. 4g+3][n]; D as above.
__device__ __forceinline__ f4 mfma16(uint2 a, uint2 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a), __builtin_bit_cast(s4, b), c, 0, 0, 0);
}
// lane k of each 16-lane row takes lane k-1 (shr) or k+1 (shl); 0 at the row ends:
// a shift along the k-line with the block's zero padding
__device__ __forceinline__ uint32_t shr1(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false); }
__device__ __forceinline__ uint32_t shl1(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xf, 0xf, false); }
__device__ __forceinline__ float shr1f(float v) { return __uint_as_float(shr1(__float_as_uint(v))); }
__device__ __forceinline__ float shl1f(float v) { return __uint_as_float(shl1(__float_as_uint(v))); }
__device__ __forceinline__ void tap_off(int t, int &di, int &dj, int &dk) {
    di = t / 9 - 1;
    dj = (t / 3) % 3 - 1;
    dk = t % 3 - 1;
}
__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
    return make_uint2((uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16), (uint32_t)f2bf(c) | ((uint32_t)f2bf(d) << 16));
}
__device__ __forceinline__ float relu(float a) { return a > 0.0f ? a : 0.0f; }

// ------------------------------------------------------------------ packing
__global__ void __launch_bounds__(256) k_cnn_pack(const float *__restrict__ p, uint16_t *__restrict__ pk) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PK_TOTAL * 8) return;
    const int e = q & 7, entry = q >> 3, lane = entry & 63, row = lane & 15, g = lane >> 4;
    float w = 0.0f;
    if (entry < PK_Q3) {                                   // P1: k = tap, row = co
        const int t = 8 * g + e;
        if (t < 27) w = p[OFF_W1 + t * 16 + row];
    } else if (entry < PK_Q2) {                            // Q3: k = tap (flipped), row = ci of layer 3
        const int t = 8 * g + e;
        if (t < 27) w = p[OFF_W3 + (26 - t) * 16 + row];
    } else {                                               // Q2: k = (t flipped, co), row = ci of layer 2
        const int k = 32 * ((entry - PK_Q2) >> 6) + 8 * g + e, t = k >> 4, co = k & 15;
        if (t < 27) w = p[OFF_W2 + ((26 - t) * 16 + row) * 16 + co];
    }
    pk[q] = f2bf(w);
}

__global__ void __launch_bounds__(256) k_cnn_packf(const float *__restrict__ p, uint16_t *__restrict__ pk) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PF_TOTAL) return;
    float w = 0.0f;
    if (q < PF_L2A) {                                      // L1: W1[t = 9 di + s][co]
        const int e = q & 3, entry = q >> 2, di = entry >> 6, lane = entry & 63, co = lane & 15, gg = lane >> 4;
        const int s9 = 4 * gg + e;
        if (s9 < 9) w = p[OFF_W1 + (9 * di + s9) * 16 + co];
    } else if (q < PF_L2B) {                               // L2A: W2[t][ci][co], dk = -1 / +1
        const int r = q - PF_L2A, e = r & 7, entry = r >> 3, pr = entry >> 6, lane = entry & 63, co = lane & 15,
                  gg = lane >> 4;
        const int dk = gg < 2 ? 0 : 2, ci = 8 * (gg & 1) + e, t = 3 * pr + dk;
        w = p[OFF_W2 + (t * 16 + ci) * 16 + co];
    } else if (q < PF_L3) {                                // L2B: dk = 0
        const int r = q - PF_L2B, e = r & 3, entry = r >> 2, pr = entry >> 6, lane = entry & 63, co = lane & 15,
                  gg = lane >> 4;
        w = p[OFF_W2 + ((3 * pr + 1) * 16 + 4 * gg + e) * 16 + co];
    } else {                                               // L3: W3[t][ci]
        const int r = q - PF_L3, e = r & 3, entry = r >> 2, di = entry >> 6, lane = entry & 63, row = lane & 15,
                  gg = lane >> 4;
        const int dj = row >> 2, dk = row & 3;
        if (dj < 3 && dk < 3) w = p[OFF_W3 + (9 * di + 3 * dj + dk) * 16 + 4 * gg + e];
    }
    pk[q] = f2bf(w);
}

// ------------------------------------------------------------------ fused forward
// One workgroup per image; march over the 16 i-planes keeping rings of 3 planes
// of x, h1 and h2 in LDS: step n loads x plane n+1, computes h1 plane n, h2
// plane n-1 and the output plane n-2.  Wave w owns the k-lines j = 4w .. 4w+3
// of every plane; an MFMA column is one voxel of a k-line.
//   layer 1: per plane offset di one 16x16x16 MFMA, K = the 9 (dj, dk) taps;
//   layer 2: per input line (plane, row) one 16x16x32 MFMA over the dk = -1 and
//     dk = +1 taps and one 16x16x16 over dk = 0, each operand read once from LDS
//     (the halo in k makes the shifted reads plain loads) and used for the up
//     to 3 output lines the row touches;
//   layer 3: per input line one 16x16x16 MFMA whose rows are the 9 (dj, dk)
//     taps (accumulated over di), combined by DPP shifts (dk) and lane shuffles
//     (dj) into the output line.
// STORE also writes h1/h2 (channel-fastest) for the backward pass.
struct HeatIO {                 // solver form of the forward (HEAT = true)
    const double *scale;       // per system, stride scale_stride
    int scale_stride, nb;
    long long istride;         // cells per system of c
    double *c;                 // FP64 output, [system][block][i][j][k]
    const int *active;         // systems still iterating (others are skipped)
};

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
// 4 floats -> 4 bfloat16 (round to nearest even, v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint2 cvt4(float a, float b, float c, float d) {
    bf16x4 v;
    v[0] = (__bf16)a; v[1] = (__bf16)b; v[2] = (__bf16)c; v[3] = (__bf16)d;
    return __builtin_bit_cast(uint2, v);
}
__device__ __forceinline__ uint16_t cvt1(float a) { return __builtin_bit_cast(uint16_t, (__bf16)a); }
__device__ __forceinline__ f4 mfma32(uint4 a, uint4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
                                                   0);
}

template <bool STORE, bool HEAT, int NW>
__global__ void __launch_bounds__(64 * NW) k_cnn_fwd(const float *__restrict__ x, float *__restrict__ out,
                                                 const uint16_t *__restrict__ pf, const float *__restrict__ prm,
                                                 uint16_t *__restrict__ h1g, uint16_t *__restrict__ h2g, HeatIO hio) {
    __shared__ __align__(16) uint16_t sm[F_TOTAL];
    const long long img = blockIdx.x;
    double sc = 1.0;
    float inv_sc = 1.0f;
    if (HEAT) {
        if (!hio.active[img / hio.nb]) return;
        sc = hio.scale[(img / hio.nb) * hio.scale_stride];
        inv_sc = sc > 0.0 ? (float)(1.0 / sc) : 0.0f;
    }
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, n = lane & 15;
    // wavefront index as a scalar: the row-range tests below are uniform
    // branches, so the MFMA accumulators stay in place across them
    constexpr int LPW = 16 / NW;            // k-lines per wavefront
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j0 = LPW * wave;
    for (int q = tid; q < F_TOTAL / 2; q += 64 * NW) reinterpret_cast<uint32_t *>(sm)[q] = 0u;

    uint2 a1[3], a2b[9], a3[3];
    uint4 a2a[9];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        a1[d] = reinterpret_cast<const uint2 *>(pf + PF_L1)[d * 64 + lane];
        a3[d] = reinterpret_cast<const uint2 *>(pf + PF_L3)[d * 64 + lane];
    }
#pragma unroll
    for (int pr = 0; pr < 9; pr++) {
        a2a[pr] = reinterpret_cast<const uint4 *>(pf + PF_L2A)[pr * 64 + lane];
        a2b[pr] = reinterpret_cast<const uint2 *>(pf + PF_L2B)[pr * 64 + lane];
    }
    float bb1[4], bb2[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        bb1[r] = prm[OFF_B1 + 4 * g + r];
        bb2[r] = prm[OFF_B2 + 4 * g + r];
    }
    const float bb3 = prm[OFF_B3];
    // layer-1 gather: K slot s = 4g + e is tap (dj, dk) = (s / 3, s % 3) - 1 at
    // offset (s/3) * 18 + s % 3 + n from the output row in a halo'd x plane
    int xo[4];
    bool xv[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int s9 = 4 * g + e;
        xv[e] = s9 < 9;
        xo[e] = xv[e] ? (s9 / 3) * 18 + s9 % 3 + n : n;
    }
    // per-lane LDS offsets in an h plane row: the k-shifted 8 channels
    // 8(g&1) .. +7 (dk = -1 for g < 2, +1 for g >= 2) as two 4-channel groups,
    // and the 4 channels 4g .. 4g+3 at k
    const int oA = (2 * (g & 1)) * HGRP + (g < 2 ? n : n + 2) * 4;
    const int oB = g * HGRP + (n + 1) * 4;
    const float *X = x + img * 4096;
    __syncthreads();

    // Planes -1 and 16 (the block's zero padding in i) are ring slots holding
    // zeros: slot 2 is not written before plane -1 is read, and plane 16 of x,
    // h1 and h2 is written as zeros one step after plane 15.  So every layer
    // runs its MFMAs over 3 planes without branches.
    // x planes are fetched from HBM one step before they are stored to LDS,
    // so the load latency overlaps a step of MFMA work
    const int xj = (tid >> 4) & 15, xk = tid & 15;
    const bool xl = tid < 256;              // the first 256 lanes move x
    auto fetch_x = [&](int q) -> float { return q < 16 ? X[(q * 16 + xj) * 16 + xk] : 0.0f; };
    auto store_x = [&](int q, float v) {
        sm[F_X + (q % 3) * PL + (xj + 1) * 18 + xk + 1] = cvt1(HEAT ? v * inv_sc : v);
    };

    if (xl) store_x(0, fetch_x(0));
    float xnext = xl ? fetch_x(1) : 0.0f;
    for (int step = 0; step <= 17; step++) {
        if (xl && step + 1 <= 16) {
            store_x(step + 1, xnext);
            if (step + 2 <= 16) xnext = fetch_x(step + 2);
        }
        __syncthreads();
        if (step == 16) {                       // h1 plane 16 = 0
#pragma unroll
            for (int jj = 0; jj < LPW; jj++)
                *reinterpret_cast<uint2 *>(sm + F_H1 + (16 % 3) * HPL + (j0 + jj + 1) * HROW + oB) = make_uint2(0u, 0u);
        }
        if (step <= 15) {                       // h1 plane step
            const int p = step;
            f4 acc[LPW];
#pragma unroll
            for (int jj = 0; jj < LPW; jj++) acc[jj] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int di = 0; di < 3; di++) {
                const int q = p + di - 1;
                const uint16_t *XP_ = sm + F_X + ((q + 3) % 3) * PL;
#pragma unroll
                for (int jj = 0; jj < LPW; jj++) {
                    const uint16_t *row = XP_ + (j0 + jj) * 18;
                    uint32_t v[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const uint32_t t = row[xo[e]];
                        v[e] = xv[e] ? t : 0u;
                    }
                    acc[jj] = mfma16(a1[di], make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16)), acc[jj]);
                }
            }
#pragma unroll
            for (int jj = 0; jj < LPW; jj++) {
                const int j = j0 + jj;
                const uint2 o = cvt4(relu(acc[jj][0] + bb1[0]), relu(acc[jj][1] + bb1[1]), relu(acc[jj][2] + bb1[2]),
                                     relu(acc[jj][3] + bb1[3]));
                *reinterpret_cast<uint2 *>(sm + F_H1 + (p % 3) * HPL + (j + 1) * HROW + oB) = o;
                if (STORE) *reinterpret_cast<uint2 *>(h1g + ((img * 4096 + (p * 16 + j) * 16 + n) * 16 + 4 * g)) = o;
            }
        }
        __syncthreads();
        if (step == 17) {                       // h2 plane 16 = 0
#pragma unroll
            for (int jj = 0; jj < LPW; jj++)
                *reinterpret_cast<uint2 *>(sm + F_H2 + (16 % 3) * HPL + (j0 + jj + 1) * HROW + oB) = make_uint2(0u, 0u);
        }
        if (step >= 1 && step <= 16) {          // h2 plane step-1
            const int p = step - 1;
            // the 16x16x32 and 16x16x16 products accumulate in separate chains
            // (an accumulator is never shared by MFMAs of different shapes)
            f4 acc[LPW], acc16[LPW];
#pragma unroll
            for (int jj = 0; jj < LPW; jj++) {
                acc[jj] = f4{0.0f, 0.0f, 0.0f, 0.0f};
                acc16[jj] = f4{0.0f, 0.0f, 0.0f, 0.0f};
            }
#pragma unroll
            for (int di = 0; di < 3; di++) {
                const int q = p + di - 1;
                const uint16_t *P = sm + F_H1 + ((q + 3) % 3) * HPL;
                // input rows j0 - 1 .. j0 + LPW (halo rows are zero); each row's
                // operands feed the output lines within one row of it
#pragma unroll
                for (int ro = 0; ro < LPW + 2; ro++) {
                    const uint16_t *row = P + (j0 + ro) * HROW;
                    const uint2 bA0 = *reinterpret_cast<const uint2 *>(row + oA);
                    const uint2 bA1 = *reinterpret_cast<const uint2 *>(row + oA + HGRP);
                    const uint4 bA = make_uint4(bA0.x, bA0.y, bA1.x, bA1.y);
                    const uint2 bB = *reinterpret_cast<const uint2 *>(row + oB);
#pragma unroll
                    for (int jj = 0; jj < LPW; jj++) {
                        const int dj = ro - 1 - jj;
                        if (dj < -1 || dj > 1) continue;
                        const int pr = di * 3 + dj + 1;
                        acc[jj] = mfma32(a2a[pr], bA, acc[jj]);
                        acc16[jj] = mfma16(a2b[pr], bB, acc16[jj]);
                    }
                }
            }
#pragma unroll
            for (int jj = 0; jj < LPW; jj++) {
                const int j = j0 + jj;
                const f4 a = acc[jj] + acc16[jj];
                const uint2 o = cvt4(relu(a[0] + bb2[0]), relu(a[1] + bb2[1]), relu(a[2] + bb2[2]), relu(a[3] + bb2[3]));
                *reinterpret_cast<uint2 *>(sm + F_H2 + (p % 3) * HPL + (j + 1) * HROW + oB) = o;
                if (STORE) *reinterpret_cast<uint2 *>(h2g + ((img * 4096 + (p * 16 + j) * 16 + n) * 16 + 4 * g)) = o;
            }
        }
        __syncthreads();
        if (step >= 2) {                        // output plane step-2
            const int p = step - 2;
            f4 D[LPW + 2];
#pragma unroll
            for (int ro = 0; ro < LPW + 2; ro++) D[ro] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int di = 0; di < 3; di++) {
                const int q = p + di - 1;
                const uint16_t *P = sm + F_H2 + ((q + 3) % 3) * HPL;
#pragma unroll
                for (int ro = 0; ro < LPW + 2; ro++)
                    D[ro] = mfma16(a3[di], *reinterpret_cast<const uint2 *>(P + (j0 + ro) * HROW + oB), D[ro]);
            }
            // lane group g of D[ro] holds the taps dj = g - 1 at input voxel n,
            // elements dk + 1: t = their sum at output voxel n (DPP for dk)
            float t[LPW + 2];
#pragma unroll
            for (int ro = 0; ro < LPW + 2; ro++) t[ro] = shr1f(D[ro][0]) + D[ro][1] + shl1f(D[ro][2]);
#pragma unroll
            for (int jj = 0; jj < LPW; jj++) {
                // output row j0 + jj: dj = -1 from row jj (group 0), dj = 0 from
                // row jj + 1 (group 1), dj = +1 from row jj + 2 (group 2)
                const float o = (t[jj] + __shfl(t[jj + 1], 16 + n, 64)) + __shfl(t[jj + 2], 32 + n, 64);
                if (g == 0) {
                    if (HEAT) {
                        const long long sys = img / hio.nb, b = img - sys * hio.nb;
                        hio.c[sys * hio.istride + b * 4096 + (p * 16 + j0 + jj) * 16 + n] = sc * (double)(o + bb3);
                    } else {
                        out[img * 4096 + (p * 16 + j0 + jj) * 16 + n] = o + bb3;
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------ data gradients
// gout[v][c] = (mask[v][c] > 0) ? sum_{t,ci} Wq[t][ci][c] gin[v + off_t][ci] : 0, stored as bfloat16.
// CIN = 1: gin is the float output gradient (layer 3, weights Q3); CIN = 16: gin is
// the bfloat16 layer-2 output gradient (weights Q2).
template <int CIN>
__global__ void __launch_bounds__(256) k_cnn_dgrad(const void *__restrict__ gin_, const uint16_t *__restrict__ mask,
                                                   uint16_t *__restrict__ gout, const us8 *__restrict__ pk) {
    __shared__ __align__(16) uint16_t sm[S_TOTAL];
    const long long img = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, kk = lane & 15;
    constexpr int S = CIN == 1 ? 1 : 14;
    constexpr int STRIDE = CIN == 1 ? 1 : 16;   // elements per voxel in an LDS plane
    for (int q = tid; q < S_TOTAL / 2; q += 256) reinterpret_cast<uint32_t *>(sm)[q] = 0u;
    us8 w[S];
#pragma unroll
    for (int s = 0; s < S; s++) w[s] = pk[(CIN == 1 ? PK_Q3 : PK_Q2) + s * 64 + lane];
    __syncthreads();

    auto load = [&](int q) {
        const int j = tid >> 4, k = tid & 15;
        const long long v = img * 4096 + (q * 16 + j) * 16 + k;
        uint16_t *dst = sm + S_IN + (q % 3) * PL * STRIDE + ((j + 1) * 18 + k + 1) * STRIDE;
        if (CIN == 1) {
            dst[0] = f2bf(reinterpret_cast<const float *>(gin_)[v]);
        } else {
            const uint4 *src = reinterpret_cast<const uint4 *>(reinterpret_cast<const uint16_t *>(gin_) + v * 16);
            reinterpret_cast<uint4 *>(dst)[0] = src[0];
            reinterpret_cast<uint4 *>(dst)[1] = src[1];
        }
    };
    auto plane = [&](int q) -> const uint16_t * {
        return (q < 0 || q > 15) ? sm + S_Z : sm + S_IN + (q % 3) * PL * STRIDE;
    };

    load(0);
    for (int p = 0; p <= 15; p++) {
        if (p + 1 <= 15) load(p + 1);
        __syncthreads();
#pragma unroll
        for (int jl = 0; jl < 4; jl++) {
            const int j = wave + 4 * jl;
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            if (CIN == 1) {
                us8 b;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int t = 8 * g + e;
                    uint16_t val = 0;
                    if (t < 27) {
                        int di, dj, dk;
                        tap_off(t, di, dj, dk);
                        val = plane(p + di)[(j + dj + 1) * 18 + kk + dk + 1];
                    }
                    b[e] = val;
                }
                acc = mfma(w[0], b, acc);
            } else {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const int t = 2 * s + (g >> 1);
                    const uint16_t *src = sm + S_Z;
                    if (t < 27) {
                        int di, dj, dk;
                        tap_off(t, di, dj, dk);
                        src = plane(p + di) + ((j + dj + 1) * 18 + kk + dk + 1) * 16 + 8 * (g & 1);
                    }
                    acc = mfma(w[s], *reinterpret_cast<const us8 *>(src), acc);
                }
            }
            const long long o = (img * 4096 + (p * 16 + j) * 16 + kk) * 16 + 4 * g;
            const uint2 mk = *reinterpret_cast<const uint2 *>(mask + o);
            const float m0 = bf2f((uint16_t)(mk.x & 0xffffu)), m1 = bf2f((uint16_t)(mk.x >> 16)),
                        m2 = bf2f((uint16_t)(mk.y & 0xffffu)), m3 = bf2f((uint16_t)(mk.y >> 16));
            *reinterpret_cast<uint2 *>(gout + o) = pack4(m0 > 0.0f ? acc[0] : 0.0f, m1 > 0.0f ? acc[1] : 0.0f,
                                                         m2 > 0.0f ? acc[2] : 0.0f, m3 > 0.0f ? acc[3] : 0.0f);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ weight gradients
// Per image, into part[img][kCnnParams] (the parameter layout):
//   L = 3: dW3[t][ci] = sum_v h2[v+t][ci] g3[v]      (in = h2, g = g3 float)
//   L = 2: dW2[t][ci][co] = sum_v h1[v+t][ci] g2[v][co]
//   L = 1: dW1[t][co] = sum_v x[v+t] g1[v][co]        (in = x float)
// and the bias gradient of the layer, sum_v g[v][co].
// D[co][n] = sum over 32-voxel chunks of A[co][voxel] B[voxel][n]; n = ci
// (L = 2, 3; taps spread over the 4 waves) or n = tap (L = 1; wave w < 2 holds
// taps 16w .. 16w+15).
template <int L>
__global__ void __launch_bounds__(256) k_cnn_wgrad(const void *__restrict__ in_, const void *__restrict__ g_,
                                                   float *__restrict__ part) {
    constexpr int ISTR = L == 1 ? 1 : 16;       // input elements per voxel
    constexpr int GCH = L == 3 ? 1 : 16;        // gradient channels
    __shared__ __align__(16) uint16_t sIn[4 * PL * ISTR];   // ring of 3 + zero plane
    __shared__ __align__(16) uint16_t sG[256 * GCH];        // gradient plane, [j][k][c]
    __shared__ float sBias[16][16];
    const long long img = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, n = lane & 15;
    for (int q = tid; q < 4 * PL * ISTR / 2; q += 256) reinterpret_cast<uint32_t *>(sIn)[q] = 0u;
    const uint16_t *zp = sIn + 3 * PL * ISTR;

    auto load_in = [&](int q) {
        const int j = tid >> 4, k = tid & 15;
        const long long v = img * 4096 + (q * 16 + j) * 16 + k;
        uint16_t *dst = sIn + (q % 3) * PL * ISTR + ((j + 1) * 18 + k + 1) * ISTR;
        if (L == 1) {
            dst[0] = f2bf(reinterpret_cast<const float *>(in_)[v]);
        } else {
            const uint4 *src = reinterpret_cast<const uint4 *>(reinterpret_cast<const uint16_t *>(in_) + v * 16);
            reinterpret_cast<uint4 *>(dst)[0] = src[0];
            reinterpret_cast<uint4 *>(dst)[1] = src[1];
        }
    };
    auto plane = [&](int q) -> const uint16_t * { return (q < 0 || q > 15) ? zp : sIn + (q % 3) * PL * ISTR; };

    f4 acc[7];
#pragma unroll
    for (int a = 0; a < 7; a++) acc[a] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    float bias_acc = 0.0f;    // thread tid: channel tid & 15, voxels of its plane rows (tid >> 4)
    __syncthreads();
    load_in(0);
    for (int p = 0; p <= 15; p++) {
        if (p + 1 <= 15) load_in(p + 1);
        {   // gradient plane p
            const int j = tid >> 4, k = tid & 15;
            const long long v = img * 4096 + (p * 16 + j) * 16 + k;
            if (L == 3) {
                sG[tid] = f2bf(reinterpret_cast<const float *>(g_)[v]);
            } else {
                const uint4 *src = reinterpret_cast<const uint4 *>(reinterpret_cast<const uint16_t *>(g_) + v * 16);
                reinterpret_cast<uint4 *>(sG + tid * 16)[0] = src[0];
                reinterpret_cast<uint4 *>(sG + tid * 16)[1] = src[1];
            }
        }
        __syncthreads();
        {   // bias: thread (c = tid & 15, row = tid >> 4) sums its row of 16 voxels
            const int c = tid & 15, jr = tid >> 4;
            if (L == 3) {
                if (c == 0)
                    for (int k = 0; k < 16; k++) bias_acc += reinterpret_cast<const float *>(g_)[img * 4096 + (p * 16 + jr) * 16 + k];
            } else {
                for (int k = 0; k < 16; k++) bias_acc += bf2f(sG[(jr * 16 + k) * 16 + c]);
            }
        }
#pragma unroll 1
        for (int ch = 0; ch < 8; ch++) {          // 32-voxel chunk: lines 2ch, 2ch+1
            const int jj = 2 * ch + (g >> 1), k0 = 8 * (g & 1);
            us8 a;                                 // A[co = n][voxels 8g..8g+7]
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int vv = jj * 16 + k0 + e;
                a[e] = (L == 3) ? (n == 0 ? sG[vv] : (uint16_t)0) : sG[vv * 16 + n];
            }
            if (L == 1) {
                if (wave < 2) {
                    const int t = 16 * wave + n;
                    us8 b;
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        uint16_t val = 0;
                        if (t < 27) {
                            int di, dj, dk;
                            tap_off(t, di, dj, dk);
                            val = plane(p + di)[(jj + dj + 1) * 18 + k0 + e + dk + 1];
                        }
                        b[e] = val;
                    }
                    acc[0] = mfma(a, b, acc[0]);
                }
            } else {
#pragma unroll
                for (int a7 = 0; a7 < 7; a7++) {
                    const int t = wave + 4 * a7;
                    if (t < 27) {
                        int di, dj, dk;
                        tap_off(t, di, dj, dk);
                        const uint16_t *P = plane(p + di);
                        us8 b;
#pragma unroll
                        for (int e = 0; e < 8; e++) b[e] = P[((jj + dj + 1) * 18 + k0 + e + dk + 1) * 16 + n];
                        acc[a7] = mfma(a, b, acc[a7]);
                    }
                }
            }
        }
        __syncthreads();
    }
    float *P = part + img * (long long)kCnnParams;
    // D[co = 4g + r][n]
    if (L == 1) {
        if (wave < 2) {
            const int t = 16 * wave + n;
            if (t < 27)
                for (int r = 0; r < 4; r++) P[OFF_W1 + t * 16 + 4 * g + r] = acc[0][r];
        }
    } else {
        for (int a7 = 0; a7 < 7; a7++) {
            const int t = wave + 4 * a7;
            if (t >= 27) continue;
            if (L == 2) {
                for (int r = 0; r < 4; r++) P[OFF_W2 + (t * 16 + n) * 16 + 4 * g + r] = acc[a7][r];
            } else if (g == 0) {
                P[OFF_W3 + t * 16 + n] = acc[a7][0];
            }
        }
    }
    sBias[tid >> 4][tid & 15] = bias_acc;
    __syncthreads();
    if (tid < 16) {
        float s = 0.0f;
        for (int r = 0; r < 16; r++) s += sBias[r][tid];
        if (L == 1) P[OFF_B1 + tid] = s;
        else if (L == 2) P[OFF_B2 + tid] = s;
        else if (tid == 0) P[OFF_B3] = s;
    }
}

// grad[q] = sum over images of part[img][q], fixed order, FP64
__global__ void __launch_bounds__(256) k_cnn_reduce(const float *__restrict__ part, int nimg, double *__restrict__ grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kCnnParams) return;
    double s = 0.0;
    for (int i = 0; i < nimg; i++) s += (double)part[(long long)i * kCnnParams + q];
    grad[q] = s;
}

// g3 = 2 (out - y) / N ; loss partials sum (out - y)^2
__global__ void __launch_bounds__(256) k_cnn_loss(const float *__restrict__ out, const float *__restrict__ y, long long n,
                                                  float *__restrict__ g3, double *__restrict__ part) {
    __shared__ double red[4];
    double acc = 0.0;
    const float sc = (float)(2.0 / (double)n);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = out[i] - y[i];
        g3[i] = sc * d;
        acc += (double)d * (double)d;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Gradient-weighted error per 16^3 image, d = out - y:
//   L = alpha sum_cells d^2 + sum over the faces inside the image w_a (d_a - d_b)^2,
// w_a the weight of the face's axis (image axes i, j, k = r, theta, phi): an
// energy-like norm of the diffusion operator (mass term alpha, face couplings
// w) that weighs the error's variation instead of its offset.
// g3 = (dL/d out) * sc; loss partials per workgroup.  One workgroup walks
// images blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ void __launch_bounds__(256) k_cnn_loss_h1(const float *__restrict__ out, const float *__restrict__ y,
                                                     int nimg, float alpha, float wi, float wj, float wk, float sc,
                                                     float *__restrict__ g3, double *__restrict__ part) {
    __shared__ float dl[4096];
    __shared__ double red[4];
    double acc = 0.0;
    for (int img = blockIdx.x; img < nimg; img += gridDim.x) {
        const long long o = (long long)img * 4096;
        for (int c = threadIdx.x; c < 4096; c += 256) dl[c] = out[o + c] - y[o + c];
        __syncthreads();
        for (int c = threadIdx.x; c < 4096; c += 256) {
            const int i = c >> 8, j = (c >> 4) & 15, k = c & 15;
            const float d = dl[c];
            float g = alpha * d;
            double l = (double)alpha * (double)d * (double)d;
            // each face enters the loss once (from its lower cell) and the
            // gradient of both cells
            if (i > 0) g += wi * (d - dl[c - 256]);
            if (i < 15) { const float e = d - dl[c + 256]; g += wi * e; l += (double)wi * ((double)e * (double)e); }
            if (j > 0) g += wj * (d - dl[c - 16]);
            if (j < 15) { const float e = d - dl[c + 16]; g += wj * e; l += (double)wj * ((double)e * (double)e); }
            if (k > 0) g += wk * (d - dl[c - 1]);
            if (k < 15) { const float e = d - dl[c + 1]; g += wk * e; l += (double)wk * ((double)e * (double)e); }
            g3[o + c] = sc * (2.0f * g);
            acc += l;
        }
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Scale-free form of the same norm, per image: with A = alpha I + sum_a w_a
// (the face-difference operator along axis a), the fraction of e's A-norm that
// the best multiple of c leaves,
//   L = 1 - <c,Ae>^2 / (<c,Ac> <e,Ae>),
// which is what the FCG step with its A-optimal scale a rewards.  dL/dc =
// -2 <c,Ae> / (<c,Ac><e,Ae>) (Ae - (<c,Ae>/<c,Ac>) Ac); g3 = dL/dc * sc.
// Images with <c,Ac> or <e,Ae> zero contribute nothing.
__global__ void __launch_bounds__(256) k_cnn_loss_cos(const float *__restrict__ out, const float *__restrict__ y,
                                                      int nimg, float alpha, float wi, float wj, float wk, float sc,
                                                      float *__restrict__ g3, double *__restrict__ part) {
    __shared__ float cl[4096], el[4096];
    __shared__ double red[3][4];
    double lacc = 0.0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int img = blockIdx.x; img < nimg; img += gridDim.x) {
        const long long o = (long long)img * 4096;
        for (int c = threadIdx.x; c < 4096; c += 256) {
            cl[c] = out[o + c];
            el[c] = y[o + c];
        }
        __syncthreads();
        // A v at cell c, v in LDS
        auto apply = [&](const float *v, int c) -> double {
            const int i = c >> 8, j = (c >> 4) & 15, k = c & 15;
            const double x = v[c];
            double a = (double)alpha * x;
            if (i > 0) a += (double)wi * (x - v[c - 256]);
            if (i < 15) a += (double)wi * (x - v[c + 256]);
            if (j > 0) a += (double)wj * (x - v[c - 16]);
            if (j < 15) a += (double)wj * (x - v[c + 16]);
            if (k > 0) a += (double)wk * (x - v[c - 1]);
            if (k < 15) a += (double)wk * (x - v[c + 1]);
            return a;
        };
        double s_ce = 0.0, s_cc = 0.0, s_ee = 0.0;
        for (int c = threadIdx.x; c < 4096; c += 256) {
            const double ae = apply(el, c), ac = apply(cl, c);
            s_ce += (double)cl[c] * ae;
            s_cc += (double)cl[c] * ac;
            s_ee += (double)el[c] * ae;
        }
        for (int off = 32; off > 0; off >>= 1) {
            s_ce += __shfl_down(s_ce, off, 64);
            s_cc += __shfl_down(s_cc, off, 64);
            s_ee += __shfl_down(s_ee, off, 64);
        }
        if (lane == 0) { red[0][wv] = s_ce; red[1][wv] = s_cc; red[2][wv] = s_ee; }
        __syncthreads();
        const double ce = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        const double cc = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
        const double ee = ((red[2][0] + red[2][1]) + red[2][2]) + red[2][3];
        const bool ok = cc > 0.0 && ee > 0.0;
        const double f = ok ? -2.0 * ce / (cc * ee) : 0.0, r = ok ? ce / cc : 0.0;
        for (int c = threadIdx.x; c < 4096; c += 256) {
            const double g = ok ? f * (apply(el, c) - r * apply(cl, c)) : 0.0;
            g3[o + c] = (float)(g * (double)sc);
        }
        if (threadIdx.x == 0 && ok) lacc += 1.0 - ce * ce / (cc * ee);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = lacc;
}

__global__ void k_cnn_sum(const double *__restrict__ part, int n, double scale, double *__restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += part[i];
    *out = s * scale;
}

__global__ void __launch_bounds__(256) k_cnn_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                                  const double *__restrict__ grad, float lr, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kCnnParams) return;
    const float g = (float)grad[q];
    const float mm = 0.9f * m[q] + 0.1f * g, vv = 0.999f * v[q] + 0.001f * g * g;
    m[q] = mm;
    v[q] = vv;
    p[q] = p[q] - lr * (mm * c1) / (sqrtf(vv * c2) + 1e-8f);
}

// batch gather: dst[b][4096] = src[idx[b]][4096]
__global__ void __launch_bounds__(256) k_cnn_gather(const float *__restrict__ src, const long long *__restrict__ idx,
                                                    float *__restrict__ dst) {
    const long long b = blockIdx.x;
    const long long s = idx[b];
    for (int v = threadIdx.x; v < 4096; v += 256) dst[b * 4096 + v] = src[s * 4096 + v];
}

}  // namespace

// ------------------------------------------------------------------ host class

CnnModel::CnnModel() {
    CCHECK(hipMalloc(&params_, sizeof(float) * kCnnParams));
    CCHECK(hipMalloc(&adam_m_, sizeof(float) * kCnnParams));
    CCHECK(hipMalloc(&adam_v_, sizeof(float) * kCnnParams));
    CCHECK(hipMalloc(&grad_, sizeof(double) * kCnnParams));
    CCHECK(hipMalloc(&packed_, sizeof(uint16_t) * (PK_TOTAL * 8 + PF_TOTAL)));
    CCHECK(hipMalloc(&loss_part_, sizeof(double) * 1025));
    init(12345);
}

CnnModel::~CnnModel() {
    for (void *p : {(void *)params_, (void *)adam_m_, (void *)adam_v_, (void *)grad_, packed_, (void *)loss_part_,
                    (void *)h1_, (void *)h2_, (void *)g2_, (void *)g1_, (void *)out_, (void *)g3_, (void *)part_,
                    (void *)xb_, (void *)yb_})
        if (p) (void)hipFree(p);
}

void CnnModel::init(uint64_t seed) {
    std::mt19937_64 rng(seed);
    std::vector<float> p(kCnnParams, 0.0f);
    auto fill = [&](int off, int n, double std) {
        std::normal_distribution<double> nd(0.0, std);
        for (int i = 0; i < n; i++) p[off + i] = (float)nd(rng);
    };
    fill(OFF_W1, 432, std::sqrt(2.0 / 27.0));
    fill(OFF_W2, 6912, std::sqrt(2.0 / 432.0));
    fill(OFF_W3, 432, std::sqrt(1.0 / 432.0));
    set_params(p.data());
}

void CnnModel::set_params(const float *p) {
    CCHECK(hipMemcpy(params_, p, sizeof(float) * kCnnParams, hipMemcpyHostToDevice));
    std::vector<float> z(kCnnParams, 0.0f);
    CCHECK(hipMemcpy(adam_m_, z.data(), sizeof(float) * kCnnParams, hipMemcpyHostToDevice));
    CCHECK(hipMemcpy(adam_v_, z.data(), sizeof(float) * kCnnParams, hipMemcpyHostToDevice));
    adam_t_ = 0;
    packed_valid_ = false;
}

void CnnModel::get_params(float *p) const {
    CCHECK(hipDeviceSynchronize());
    CCHECK(hipMemcpy(p, params_, sizeof(float) * kCnnParams, hipMemcpyDeviceToHost));
}

void CnnModel::pack(hipStream_t stream) {
    if (packed_valid_) return;
    hipLaunchKernelGGL(k_cnn_pack, dim3((PK_TOTAL * 8 + 255) / 256), dim3(256), 0, stream, (const float *)params_,
                       (uint16_t *)packed_);
    hipLaunchKernelGGL(k_cnn_packf, dim3((PF_TOTAL + 255) / 256), dim3(256), 0, stream, (const float *)params_,
                       (uint16_t *)packed_ + PK_TOTAL * 8);
    packed_valid_ = true;
}

void CnnModel::ensure_batch(int nimg) {
    if (nimg <= max_batch_) return;
    for (void *p : {(void *)h1_, (void *)h2_, (void *)g2_, (void *)g1_, (void *)out_, (void *)g3_, (void *)part_,
                    (void *)xb_, (void *)yb_})
        if (p) CCHECK(hipFree(p));
    const size_t n = (size_t)nimg;
    CCHECK(hipMalloc(&h1_, n * 4096 * 16 * sizeof(uint16_t)));
    CCHECK(hipMalloc(&h2_, n * 4096 * 16 * sizeof(uint16_t)));
    CCHECK(hipMalloc(&g2_, n * 4096 * 16 * sizeof(uint16_t)));
    CCHECK(hipMalloc(&g1_, n * 4096 * 16 * sizeof(uint16_t)));
    CCHECK(hipMalloc(&out_, n * 4096 * sizeof(float)));
    CCHECK(hipMalloc(&g3_, n * 4096 * sizeof(float)));
    CCHECK(hipMalloc(&xb_, n * 4096 * sizeof(float)));
    CCHECK(hipMalloc(&yb_, n * 4096 * sizeof(float)));
    CCHECK(hipMalloc(&part_, n * kCnnParams * sizeof(float)));
    max_batch_ = nimg;
}

void CnnModel::infer(const float *x, float *out, long long nimg, hipStream_t stream) {
    if (nimg <= 0) return;
    pack(stream);
    hipLaunchKernelGGL((k_cnn_fwd<false, false, kFwdWaves>), dim3((unsigned)nimg), dim3(64 * kFwdWaves), 0, stream, x, out,
                       (const uint16_t *)packed_ + PK_TOTAL * 8, (const float *)params_,
                       (uint16_t *)nullptr, (uint16_t *)nullptr, HeatIO{});
}

void CnnModel::infer_heat(const float *x, const double *scale, int scale_stride, int nb, long long istride, double *c,
                          long long nimg, const int *active, hipStream_t stream) {
    if (nimg <= 0) return;
    pack(stream);
    HeatIO hio{scale, scale_stride, nb, istride, c, active};
    hipLaunchKernelGGL((k_cnn_fwd<false, true, kFwdWaves>), dim3((unsigned)nimg), dim3(64 * kFwdWaves), 0, stream, x, (float *)nullptr,
                       (const uint16_t *)packed_ + PK_TOTAL * 8, (const float *)params_,
                       (uint16_t *)nullptr, (uint16_t *)nullptr, hio);
}

void CnnModel::backward(const float *x, int nimg, hipStream_t stream) {
    hipLaunchKernelGGL((k_cnn_dgrad<1>), dim3(nimg), dim3(256), 0, stream, (const void *)g3_, (const uint16_t *)h2_, g2_,
                       (const us8 *)packed_);
    hipLaunchKernelGGL((k_cnn_dgrad<16>), dim3(nimg), dim3(256), 0, stream, (const void *)g2_, (const uint16_t *)h1_, g1_,
                       (const us8 *)packed_);
    hipLaunchKernelGGL((k_cnn_wgrad<3>), dim3(nimg), dim3(256), 0, stream, (const void *)h2_, (const void *)g3_, part_);
    hipLaunchKernelGGL((k_cnn_wgrad<2>), dim3(nimg), dim3(256), 0, stream, (const void *)h1_, (const void *)g2_, part_);
    hipLaunchKernelGGL((k_cnn_wgrad<1>), dim3(nimg), dim3(256), 0, stream, (const void *)x, (const void *)g1_, part_);
    hipLaunchKernelGGL(k_cnn_reduce, dim3((kCnnParams + 255) / 256), dim3(256), 0, stream, (const float *)part_, nimg,
                       grad_);
}

void CnnModel::gradient(const float *x, const float *gout, int nimg, double *grad_host, hipStream_t stream) {
    ensure_batch(nimg);
    pack(stream);
    hipLaunchKernelGGL((k_cnn_fwd<true, false, kFwdWaves>), dim3(nimg), dim3(64 * kFwdWaves), 0, stream, x, out_,
                       (const uint16_t *)packed_ + PK_TOTAL * 8, (const float *)params_, h1_, h2_, HeatIO{});
    CCHECK(hipMemcpyAsync(g3_, gout, sizeof(float) * 4096 * (size_t)nimg, hipMemcpyDeviceToDevice, stream));
    backward(x, nimg, stream);
    CCHECK(hipStreamSynchronize(stream));
    CCHECK(hipMemcpy(grad_host, grad_, sizeof(double) * kCnnParams, hipMemcpyDeviceToHost));
}

double CnnModel::train_step(const float *x, const float *y, int nimg, double lr, hipStream_t stream) {
    ensure_batch(nimg);
    pack(stream);
    hipLaunchKernelGGL((k_cnn_fwd<true, false, kFwdWaves>), dim3(nimg), dim3(64 * kFwdWaves), 0, stream, x, out_,
                       (const uint16_t *)packed_ + PK_TOTAL * 8, (const float *)params_, h1_, h2_, HeatIO{});
    const long long n = 4096LL * nimg;
    if (h1_alpha_ >= 0.0 && loss_cos_) {
        // mean over the images of the scale-free loss
        const int nwg = nimg < 1024 ? nimg : 1024;
        hipLaunchKernelGGL(k_cnn_loss_cos, dim3(nwg), dim3(256), 0, stream, (const float *)out_, y, nimg,
                           (float)h1_alpha_, (float)h1_w_[0], (float)h1_w_[1], (float)h1_w_[2],
                           (float)(1.0 / (double)nimg), g3_, loss_part_);
        hipLaunchKernelGGL(k_cnn_sum, dim3(1), dim3(64), 0, stream, (const double *)loss_part_, nwg,
                           1.0 / (double)nimg, loss_part_ + 1024);
    } else if (h1_alpha_ >= 0.0) {
        const int nwg = nimg < 1024 ? nimg : 1024;
        hipLaunchKernelGGL(k_cnn_loss_h1, dim3(nwg), dim3(256), 0, stream, (const float *)out_, y, nimg,
                           (float)h1_alpha_, (float)h1_w_[0], (float)h1_w_[1], (float)h1_w_[2],
                           (float)(1.0 / (double)n), g3_, loss_part_);
        hipLaunchKernelGGL(k_cnn_sum, dim3(1), dim3(64), 0, stream, (const double *)loss_part_, nwg, 1.0 / (double)n,
                           loss_part_ + 1024);
    } else {
        hipLaunchKernelGGL(k_cnn_loss, dim3(1024), dim3(256), 0, stream, (const float *)out_, y, n, g3_, loss_part_);
        hipLaunchKernelGGL(k_cnn_sum, dim3(1), dim3(64), 0, stream, (const double *)loss_part_, 1024, 1.0 / (double)n,
                           loss_part_ + 1024);
    }
    backward(x, nimg, stream);
    adam_t_++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)adam_t_)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)adam_t_)));
    hipLaunchKernelGGL(k_cnn_adam, dim3((kCnnParams + 255) / 256), dim3(256), 0, stream, params_, adam_m_, adam_v_,
                       (const double *)grad_, (float)lr, c1, c2);
    packed_valid_ = false;
    double loss = 0.0;
    CCHECK(hipMemcpyAsync(&loss, loss_part_ + 1024, sizeof(double), hipMemcpyDeviceToHost, stream));
    CCHECK(hipStreamSynchronize(stream));
    return loss;
}

double CnnModel::train(const float *x, const float *y, long long nimg, int epochs, int batch, double lr, uint64_t seed,
                       hipStream_t stream) {
    if (nimg <= 0 || epochs <= 0) return 0.0;
    batch = (int)std::min<long long>(batch, nimg);
    ensure_batch(batch);
    std::mt19937_64 rng(seed);
    std::vector<long long> idx(nimg);
    long long *didx = nullptr;
    CCHECK(hipMalloc(&didx, sizeof(long long) * nimg));
    double last = 0.0;
    try {
        for (int ep = 0; ep < epochs; ep++) {
            // learning rate decays linearly from lr to lr/10 over the epochs
            const double lr_ep = epochs > 1 ? lr * (1.0 - 0.9 * (double)ep / (double)(epochs - 1)) : lr;
            for (long long i = 0; i < nimg; i++) idx[i] = i;
            std::shuffle(idx.begin(), idx.end(), rng);
            CCHECK(hipMemcpy(didx, idx.data(), sizeof(long long) * nimg, hipMemcpyHostToDevice));
            double tot = 0.0;
            long long seen = 0;
            for (long long b0 = 0; b0 + batch <= nimg; b0 += batch) {
                hipLaunchKernelGGL(k_cnn_gather, dim3(batch), dim3(256), 0, stream, x, (const long long *)(didx + b0), xb_);
                hipLaunchKernelGGL(k_cnn_gather, dim3(batch), dim3(256), 0, stream, y, (const long long *)(didx + b0), yb_);
                tot += train_step(xb_, yb_, batch, lr_ep, stream) * batch;
                seen += batch;
            }
            last = seen ? tot / (double)seen : 0.0;
        }
    } catch (...) {
        (void)hipFree(didx);
        throw;
    }
    CCHECK(hipFree(didx));
    return last;
}

}  // namespace falcon
