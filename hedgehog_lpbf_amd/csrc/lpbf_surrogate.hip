// SPDX-License-Identifier: 0BSD
// Melt-pool surrogate kernels for gfx950 (include/lpbf_surrogate.h).
//
// Activations are NHWC bf16 with 32 channels per pixel (64 bytes), so one
// 16-byte load gives a lane the 8 consecutive channels an MFMA operand needs.
//
//   k_conv_mfma  conv3x3 32 -> 32 as an implicit GEMM on
//                v_mfma_f32_16x16x32_bf16: per tap, D[co][px] += W[co][ci] *
//                X[px][ci]; a wavefront owns 64 pixels of one row and all 32
//                output channels (8 accumulator tiles), weights stay in
//                registers for all 9 taps.  Epilogue: bias + gelu (forward,
//                z and a stored) or the gelu' factor of the previous layer
//                (backward data, weights flipped and transposed).
//   k_dw_mfma    weight gradient: dW[t][co][ci] = sum_px dz[px][co] *
//                a[px + off(t)][ci]; a workgroup stages one row of dz and the
//                three rows of a it touches in LDS (72-byte pixel pitch, no
//                bank conflicts for the transposed reads) and reduces over its
//                rows with MFMAs whose K is 32 pixels; one extra MFMA against
//                ones gives the bias gradient.  Per-workgroup partials are
//                summed in a fixed order (deterministic).
//   layer 6      (32 -> 1) and its gradients are VALU kernels: one output
//                channel is no GEMM.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstddef>

#include "lpbf_common.hpp"
#include "lpbf_kernels.h"
#include "lpbf_surrogate.h"

namespace lpbf_amd {
namespace {

constexpr int kW = LPBF_SG_WIN;        // 128
constexpr int kC = LPBF_SG_CH;         // 32
constexpr int kPix = kW * kW;
constexpr int kT = 256;
constexpr int kPitch = 36;             // bf16 per pixel in LDS (72 bytes)
constexpr int kDwSlots = 10;           // 9 taps + bias

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// Parameter vector offsets (FP32 master layout).
constexpr int kL1W = 9 * 4 * kC, kL1 = kL1W + kC;           // 1184
constexpr int kLW = 9 * kC * kC, kLn = kLW + kC;            // 9248
constexpr int kL6 = kL1 + 4 * kLn;                          // 38176
static_assert(kL6 + 9 * kC + 1 == LPBF_SG_NPARAM, "parameter layout");

// Pack layout (bytes).
constexpr size_t kWsz = 9 * kC * kC;                        // bf16 per layer
constexpr size_t kOffWF = 0;                                // [5][9][co][ci]
constexpr size_t kOffWB = kOffWF + 5 * kWsz * 2;            // [5][9][out][in], flipped
constexpr size_t kOffB = kOffWB + 5 * kWsz * 2;             // [5][32] fp32
constexpr size_t kOffW6 = kOffB + 5 * kC * 4;               // [9][32] fp32 (bf16 values)
constexpr size_t kOffB6 = kOffW6 + 9 * kC * 4;              // fp32
constexpr size_t kPackBytes = (kOffB6 + 4 + 255) / 256 * 256;

__device__ __host__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline float bf2f(uint16_t h) {
    return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16);
}

// Hardware round-to-nearest-even conversion (v_cvt_pk_bf16_f32).
__device__ inline uint32_t pack_bf2(float a, float b) {
    return uint32_t(__builtin_bit_cast(uint16_t, static_cast<__bf16>(a))) |
           (uint32_t(__builtin_bit_cast(uint16_t, static_cast<__bf16>(b))) << 16);
}

// gelu (tanh form): 0.5 z (1 + tanh(u)), u = c (z + 0.044715 z^3), c = sqrt(2/pi),
// evaluated as z s with s = 0.5 (1 + tanh(u)) = 1 / (1 + 2^(-2 log2(e) u)): one
// exponential, one reciprocal and four other operations (the exponent
// saturates to inf or 0, so s goes to 0 or 1 without NaNs).
constexpr float kGeluC = 0.7978845608028654f;               // sqrt(2/pi)
constexpr float kGeluE1 = -2.0f * 1.4426950408889634f * kGeluC;   // -2 log2(e) c
constexpr float kGeluE3 = kGeluE1 * 0.044715f;
__device__ inline float gelu(float z) {
    const float e = z * fmaf(z * z, kGeluE3, kGeluE1);
    return z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(e));
}
// tanh(u) = 1 - 2 / (1 + e^{2u}) with the hardware exp and reciprocal: a few
// ulp of FP32, far below the bf16 rounding that follows (saturates to +-1).
// The derivative in this form measured faster in the backward-data conv than
// the sigmoid form above.
__device__ inline float fast_tanh(float u) {
    return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __expf(2.f * u));
}
__device__ inline float gelu_grad(float z) {
    const float t = fast_tanh(kGeluC * (z + 0.044715f * z * z * z));
    return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * kGeluC * (1.f + 3.f * 0.044715f * z * z);
}

__device__ inline bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// ---- pack ------------------------------------------------------------------

__global__ void k_pack(const float* __restrict__ p, uint8_t* __restrict__ pack) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // over 5*9*32*32
    uint16_t* wf = reinterpret_cast<uint16_t*>(pack + kOffWF);
    uint16_t* wb = reinterpret_cast<uint16_t*>(pack + kOffWB);
    if (i < 5 * static_cast<int>(kWsz)) {
        const int l = i / kWsz, r = i % kWsz;
        const int t = r / (kC * kC), co = (r / kC) % kC, ci = r % kC;
        float w;
        if (l == 0) w = ci < 4 ? p[(t * 4 + ci) * kC + co] : 0.f;
        else w = p[kL1 + (l - 1) * kLn + (t * kC + ci) * kC + co];
        wf[i] = f2bf(w);
        // backward data of layer l: out channel = ci of the forward, in = co,
        // tap flipped (the adjoint reads dz at p - off(t))
        const int tb = 8 - t;
        const int out = co, in = ci;   // reuse indices: entry [tb][out][in]
        float wbv;
        if (l == 0) wbv = out < 4 ? p[(t * 4 + out) * kC + in] : 0.f;
        else wbv = p[kL1 + (l - 1) * kLn + (t * kC + out) * kC + in];
        wb[static_cast<size_t>(l) * kWsz + (tb * kC + out) * kC + in] = f2bf(wbv);
    }
    if (i < 5 * kC) {
        const int l = i / kC, c = i % kC;
        reinterpret_cast<float*>(pack + kOffB)[i] =
            l == 0 ? p[kL1W + c] : p[kL1 + (l - 1) * kLn + kLW + c];
    }
    if (i < 9 * kC)
        reinterpret_cast<float*>(pack + kOffW6)[i] = bf2f(f2bf(p[kL6 + i]));
    if (i == 0) reinterpret_cast<float*>(pack + kOffB6)[0] = p[kL6 + 9 * kC];
}

// ---- conv 32 -> 32 on MFMA --------------------------------------------------

// kModeFwd: forward, activation a only (inference); kModeFwdZ: forward,
// pre-activation z and a (training); kModeBwd: backward data.
enum { kModeFwd = 0, kModeBwd = 1, kModeFwdZ = 2 };

// X, out0, out1, zprev: [batch][128][128][32] bf16.  W: [9][32 out][32 in].
// A workgroup owns a strip of `strip` output rows of one window and walks it
// one row at a time; wavefront w takes pixels 32w .. 32w+31 of the row, all
// 32 output channels (2 x 2 accumulator tiles; the weights of the nine taps
// stay in registers).  The input rows live in a six-slot LDS ring (130
// pixels with the zero border, 64 bytes each): three hold the rows the
// current output row reads, three stream in from HBM for the rows ahead
// (global_load_lds_dwordx4: lane l of a wavefront lands 16 bytes at the slot
// offset + 16 l, so one instruction moves 16 pixels; nothing passes through
// registers).  An input row is read once per strip (plus one halo row each
// side), and its load has three rows of MFMAs to land in.  Each output tile
// accumulates its nine taps in the same order whatever the strip height.
constexpr int kConvCols = kW + 2;
constexpr int kSlot = kConvCols * 4;            // uint4 per ring slot
constexpr int kRing = 6;
__device__ __forceinline__ int ring_slot(int yy) { return (yy + kRing) % kRing; }   // yy >= -1

typedef __attribute__((address_space(3))) uint4 lds_uint4_t;

// 16 bytes per lane from global memory straight into LDS: lane l writes at
// LDS address lds_base + 16 l (lds_base wave-uniform, in M0).  Written as
// asm so that the compiler, which would otherwise wait for the transfer
// before every later LDS read (it cannot tell the ring slots apart), leaves
// the wait to the explicit s_waitcnt before the barrier that publishes the
// slot.  The instruction has no register result, and vector-memory
// operations complete in issue order, so an operation the compiler does not
// know about only makes its own vmcnt waits stricter.
__device__ __forceinline__ void lds_dma16(const uint4* gptr, uint32_t lds_base) {
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
                 :
                 : "v"(gptr), "s"(lds_base)
                 : "memory", "m0");
}

// Row yy of the window into its ring slot: wavefront w moves pixels
// 32w .. 32w+31 (two instructions, kDmaOps); a row outside the window is
// written as zeros (no memory operation).  Wave-uniform.
constexpr int kDmaOps = 2;
__device__ __forceinline__ void conv_row_to_lds(const uint4* __restrict__ Xv, uint4* s_x, int yy,
                                                int wave, int lane) {
    uint4* slot = s_x + ring_slot(yy) * kSlot + 4;   // pixel 0
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int px0 = wave * 32 + h * 16;
        if (yy >= 0 && yy < kW) {
            const uint32_t base = static_cast<uint32_t>(
                reinterpret_cast<uintptr_t>((lds_uint4_t*)(slot + px0 * 4)));
            lds_dma16(Xv + (yy * kW + px0) * 4 + lane, __builtin_amdgcn_readfirstlane(base));
        } else {
            slot[px0 * 4 + lane] = make_uint4(0, 0, 0, 0);
        }
    }
}

// Waits until at most n vector-memory operations of this wavefront are
// outstanding (n an immediate).
template <int N>
__device__ __forceinline__ void vm_wait() {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

template <int MODE>
__global__ __launch_bounds__(kT) void k_conv_mfma(const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ W,
                                                  const float* __restrict__ bias,
                                                  const uint16_t* __restrict__ zprev,
                                                  uint16_t* __restrict__ out0,
                                                  uint16_t* __restrict__ out1, int batch,
                                                  int strip) {
    // vector-memory operations a wavefront issues per output row after that
    // row's transfer: the stores of one or two output tensors (four each),
    // or the backward form's four z loads and four stores
    constexpr int kRowOps = MODE == kModeFwd ? 4 : 8;
    __shared__ __attribute__((aligned(16))) uint4 s_x[kRing * kSlot];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nstrip = kW / strip;
    const int img = blockIdx.x / nstrip, ys = (blockIdx.x % nstrip) * strip;
    if (img >= batch) return;
    const int lr = lane & 15, lg = lane >> 4;
    const int ye = ys + strip;                       // one past the strip's last row
    const uint4* Xv = reinterpret_cast<const uint4*>(X) + static_cast<size_t>(img) * kPix * 4;

    // the zero border pixels of every slot (never overwritten); the first
    // row's three input rows and the two after them
    if (threadIdx.x < 8 * kRing) {
        const int sl = threadIdx.x >> 3, e = threadIdx.x & 7;
        s_x[sl * kSlot + (e < 4 ? e : kSlot - 8 + e)] = make_uint4(0, 0, 0, 0);
    }
    for (int yy = ys - 1; yy <= min(ys + 3, ye); ++yy) conv_row_to_lds(Xv, s_x, yy, wave, lane);

    bf16x8 w[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
            w[t][ct] = as_bf8(*reinterpret_cast<const uint4*>(
                W + (static_cast<size_t>(t) * kC + ct * 16 + lr) * kC + 8 * lg));
    const int x0 = wave * 32;
    float bv[2][4];   // this lane's output channels' bias (forward)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[ct][i] = MODE != kModeBwd ? bias[ct * 16 + 4 * lg + i] : 0.f;
    vm_wait<0>();
    __syncthreads();

    for (int row = ys; row < ye; ++row) {
        // the input row of output row + 3, into the slot row - 2 vacated
        const bool dma = row + 3 < ye;               // workgroup-uniform
        if (dma) conv_row_to_lds(Xv, s_x, row + 4, wave, lane);
        // backward: the previous layer's z at this row's pixels, read before
        // the MFMAs so that its latency hides behind them
        uint2 zp[2][2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
                zp[ct][pt] = MODE != kModeBwd
                                 ? make_uint2(0, 0)
                                 : *reinterpret_cast<const uint2*>(
                                       zprev + (static_cast<size_t>(img) * kPix + row * kW + x0 +
                                                pt * 16 + lr) * kC + ct * 16 + 4 * lg);

        f32x4 acc[2][2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) acc[ct][pt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t % 3;   // input row row - 1 + dy, column px + dx
            const int sl = ring_slot(row - 1 + dy);
            bf16x8 xv[2];
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
                xv[pt] = as_bf8(s_x[sl * kSlot + (x0 + pt * 16 + lr + dx) * 4 + lg]);
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int pt = 0; pt < 2; ++pt)
                    acc[ct][pt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[t][ct], xv[pt],
                                                                          acc[ct][pt], 0, 0, 0);
        }

        // lane holds channels co0..co0+3 of pixel x0 + pt*16 + lr
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const int co0 = ct * 16 + 4 * lg;
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                const size_t off =
                    (static_cast<size_t>(img) * kPix + row * kW + x0 + pt * 16 + lr) * kC + co0;
                if (MODE != kModeBwd) {
                    float z[4], a[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        z[i] = acc[ct][pt][i] + bv[ct][i];
                        a[i] = gelu(z[i]);
                    }
                    if (MODE == kModeFwdZ)
                        *reinterpret_cast<uint2*>(out0 + off) =
                            make_uint2(pack_bf2(z[0], z[1]), pack_bf2(z[2], z[3]));
                    *reinterpret_cast<uint2*>(out1 + off) =
                        make_uint2(pack_bf2(a[0], a[1]), pack_bf2(a[2], a[3]));
                } else {
                    const uint2 zq = zp[ct][pt];
                    const uint16_t zs[4] = {uint16_t(zq.x & 0xffff), uint16_t(zq.x >> 16),
                                            uint16_t(zq.y & 0xffff), uint16_t(zq.y >> 16)};
                    float d[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) d[i] = acc[ct][pt][i] * gelu_grad(bf2f(zs[i]));
                    *reinterpret_cast<uint2*>(out0 + off) =
                        make_uint2(pack_bf2(d[0], d[1]), pack_bf2(d[2], d[3]));
                }
            }
        }
        // output row + 1 reads input row + 2, whose transfer was issued two
        // rows ago (row - 2 + 4; the first two rows' come from the
        // prologue, already waited for).  Operations issued after it: this
        // row's and the previous row's transfers (if any) and the three
        // rows' other operations; vector memory completes in issue order.
        // A row past the window's last is written as zeros with no memory
        // operation, so it does not count (counting it would let the wait
        // pass before the transfer it guards has landed).
        if (row + 1 < ye && row - 2 >= ys) {
            const bool dma_mem = dma && row + 4 < kW;                 // this row's transfer
            const bool dma_prev = row + 2 < ye && row + 3 < kW;       // row - 1's
            if (dma_mem && dma_prev) vm_wait<3 * kRowOps + 2 * kDmaOps>();
            else if (dma_mem || dma_prev) vm_wait<3 * kRowOps + kDmaOps>();
            else vm_wait<3 * kRowOps>();
        }
        // after the barrier every wavefront's part of input row + 2 is in
        // its slot, and every wavefront has finished with row - 1's slot
        __syncthreads();
    }
}

// ---- weight gradient on MFMA ------------------------------------------------

// partial[wg][slot][co][ci], slot 0..8 = taps, 9 = bias (ci column 0..15 of
// the it == 0 waves; every column holds the same sum).
__global__ __launch_bounds__(kT) void k_dw_mfma(const uint16_t* __restrict__ dz,
                                                const uint16_t* __restrict__ A,
                                                float* __restrict__ partial, int rows,
                                                int rows_per_wg) {
    __shared__ __attribute__((aligned(16))) uint16_t s_dz[kW * kPitch];
    __shared__ __attribute__((aligned(16))) uint16_t s_a[3 * (kW + 2) * kPitch];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ct = wave & 1, it = wave >> 1;
    const int lr = lane & 15, lg = lane >> 4;
    const int r0 = blockIdx.x * rows_per_wg;
    const int r1 = min(r0 + rows_per_wg, rows);

    f32x4 acc[kDwSlots];
#pragma unroll
    for (int s = 0; s < kDwSlots; ++s) acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t one = 0x3f80;
    const uint4 ones4 = make_uint4(one | (uint32_t(one) << 16), one | (uint32_t(one) << 16),
                                   one | (uint32_t(one) << 16), one | (uint32_t(one) << 16));
    const bf16x8 ones = as_bf8(ones4);

    for (int gr = r0; gr < r1; ++gr) {
        const int img = gr / kW, y = gr % kW;
        // all loads of the row in flight before the LDS stores
        constexpr int kDz = kW * 4, kA = 3 * (kW + 2) * 4;
        constexpr int kItD = (kDz + kT - 1) / kT, kItA = (kA + kT - 1) / kT;
        uint4 vd[kItD], va[kItA];
#pragma unroll
        for (int it = 0; it < kItD; ++it) {
            const int q = tid + it * kT, px = q >> 2, c = q & 3;
            vd[it] = q < kDz ? *reinterpret_cast<const uint4*>(
                                   dz + (static_cast<size_t>(img) * kPix + y * kW + px) * kC + c * 8)
                             : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int it = 0; it < kItA; ++it) {
            const int q = tid + it * kT;
            const int rr = q / ((kW + 2) * 4), rem = q % ((kW + 2) * 4);
            const int px = rem / 4 - 1, c = rem & 3;
            const int yy = y + rr - 1;
            va[it] = make_uint4(0, 0, 0, 0);
            if (q < kA && yy >= 0 && yy < kW && px >= 0 && px < kW)
                va[it] = *reinterpret_cast<const uint4*>(
                    A + (static_cast<size_t>(img) * kPix + yy * kW + px) * kC + c * 8);
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kItD; ++it) {
            const int q = tid + it * kT, px = q >> 2, c = q & 3;
            if (q < kDz) {
                uint2* d = reinterpret_cast<uint2*>(s_dz + px * kPitch + c * 8);
                d[0] = make_uint2(vd[it].x, vd[it].y);
                d[1] = make_uint2(vd[it].z, vd[it].w);
            }
        }
#pragma unroll
        for (int it = 0; it < kItA; ++it) {
            const int q = tid + it * kT;
            if (q < kA) {
                const int rr = q / ((kW + 2) * 4), rem = q % ((kW + 2) * 4);
                const int px = rem / 4 - 1, c = rem & 3;
                uint2* d = reinterpret_cast<uint2*>(s_a + (rr * (kW + 2) + px + 1) * kPitch + c * 8);
                d[0] = make_uint2(va[it].x, va[it].y);
                d[1] = make_uint2(va[it].z, va[it].w);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kb = 0; kb < kW / 32; ++kb) {
            bf16x8 a;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint16_t h = s_dz[(kb * 32 + 8 * lg + i) * kPitch + ct * 16 + lr];
                a[i] = __builtin_bit_cast(__bf16, h);
            }
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int dy = t / 3, dx = t % 3;   // row in s_a, column shift + 1
                bf16x8 b;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const uint16_t h =
                        s_a[(dy * (kW + 2) + kb * 32 + 8 * lg + i + dx) * kPitch + it * 16 + lr];
                    b[i] = __builtin_bit_cast(__bf16, h);
                }
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
            if (it == 0) acc[9] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, ones, acc[9], 0, 0, 0);
        }
    }
    float* P = partial + static_cast<size_t>(blockIdx.x) * kDwSlots * kC * kC;
#pragma unroll
    for (int s = 0; s < kDwSlots; ++s) {
        if (s == 9 && it != 0) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            P[(s * kC + ct * 16 + 4 * lg + i) * kC + it * 16 + lr] = acc[s][i];
    }
}

// Sum the per-workgroup partials (fixed order) and scatter into the FP32
// gradient vector of layer l (0-based, 0..4).  A block takes 32 entries;
// thread (o, g) sums workgroups g, g+8, ..., then a fixed tree over g.
__global__ __launch_bounds__(kT) void k_dw_reduce(const float* __restrict__ partial, int nwg,
                                                  int l, float* __restrict__ grad) {
    __shared__ float s[8][32];
    const int o = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int j = blockIdx.x * 32 + o;                     // slot*1024 + co*32 + ci
    const size_t stride = static_cast<size_t>(kDwSlots) * kC * kC;
    float a = 0.f;
    for (int w = g; w < nwg; w += 8) a += partial[w * stride + j];
    s[g][o] = a;
    __syncthreads();
    if (g != 0) return;
    const float sum = ((s[0][o] + s[1][o]) + (s[2][o] + s[3][o])) +
                      ((s[4][o] + s[5][o]) + (s[6][o] + s[7][o]));
    const int sl = j / (kC * kC), co = (j / kC) % kC, ci = j % kC;
    if (sl == 9 && ci != 0) return;
    if (l == 0 && sl < 9 && ci >= 4) return;
    if (l == 0) {
        if (sl < 9) grad[(sl * 4 + ci) * kC + co] = sum;
        else grad[kL1W + co] = sum;
    } else {
        float* gp = grad + kL1 + (l - 1) * kLn;
        if (sl < 9) gp[(sl * kC + ci) * kC + co] = sum;
        else gp[kLW + co] = sum;
    }
}

// ---- layer 6 (32 -> 1), VALU ------------------------------------------------

// y = b6 + sum_t sum_ci a5[p + off(t)][ci] W6[t][ci]; with a target also the
// loss gradient dy = 2 (y - target) / n_total and per-workgroup sums of
// (y-t)^2.  A workgroup owns a strip of `strip` rows (even) of one window
// and takes two rows at a time, one pixel per thread: the four a5 rows the
// pair reads sit in LDS (staged once, 64 bytes per pixel, zero border), and
// the two rows the next pair adds are loaded into registers while this pair
// is summed.  Each pixel sums its 288 products in the order t, channel.
__global__ __launch_bounds__(kT) void k_out_fwd(const uint16_t* __restrict__ a5,
                                                const uint8_t* __restrict__ pack,
                                                const float* __restrict__ target, float* __restrict__ y,
                                                float* __restrict__ dy, double* __restrict__ lpart,
                                                int batch, int strip) {
    constexpr int kCols = kW + 2, kSlot = kCols * 4;
    __shared__ __attribute__((aligned(16))) uint4 s_a[4 * kSlot];   // ring of four rows
    const float* W6 = reinterpret_cast<const float*>(pack + kOffW6);
    const float b6 = reinterpret_cast<const float*>(pack + kOffB6)[0];
    const long long n = static_cast<long long>(batch) * kPix;
    const int nstrip = kW / strip;
    const int img = blockIdx.x / nstrip, ys = (blockIdx.x % nstrip) * strip;
    const uint4* Av = reinterpret_cast<const uint4*>(a5) + static_cast<size_t>(img) * kPix * 4;
    auto slot_of = [](int yy) { return (yy + 4) & 3; };
    // rows y0 .. y0+nrow-1 into registers / into their slots
    auto fetch = [&](uint4* v, int nit, int y0, int nrow) {
        for (int it = 0; it < nit; ++it) {
            const int q = threadIdx.x + it * kT;
            const int rr = q / kSlot, rem = q % kSlot;
            const int px = rem / 4 - 1, c = rem & 3, yy = y0 + rr;
            v[it] = make_uint4(0, 0, 0, 0);
            if (rr < nrow && yy >= 0 && yy < kW && px >= 0 && px < kW)
                v[it] = Av[(yy * kW + px) * 4 + c];
        }
    };
    auto stash = [&](const uint4* v, int nit, int y0, int nrow) {
        for (int it = 0; it < nit; ++it) {
            const int q = threadIdx.x + it * kT;
            const int rr = q / kSlot, rem = q % kSlot;
            if (rr < nrow) s_a[slot_of(y0 + rr) * kSlot + rem] = v[it];
        }
    };
    double d2 = 0.0;
    if (img < batch) {
        {
            constexpr int kIt = (4 * kSlot + kT - 1) / kT;
            uint4 v[kIt];
            fetch(v, kIt, ys - 1, 4);
            stash(v, kIt, ys - 1, 4);
        }
        __syncthreads();
        const int lrow = threadIdx.x / kW, px = threadIdx.x % kW;
        constexpr int kNext = (2 * kSlot + kT - 1) / kT;
        for (int r0 = ys; r0 < ys + strip; r0 += 2) {
            const bool more = r0 + 2 < ys + strip;   // workgroup-uniform
            uint4 nv[kNext];
            if (more) fetch(nv, kNext, r0 + 3, 2);
            const int py = r0 + lrow;
            float acc = b6;
            for (int t = 0; t < 9; ++t) {
                // pixel px + dx - 1 of row py + dy - 1 is LDS column px + dx
                const uint4* src =
                    s_a + slot_of(py + t / 3 - 1) * kSlot + (px + t % 3) * 4;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint4 v = src[c];
                    const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        acc = fmaf(bf2f(uint16_t(u[e] & 0xffff)), W6[t * kC + c * 8 + 2 * e], acc);
                        acc = fmaf(bf2f(uint16_t(u[e] >> 16)), W6[t * kC + c * 8 + 2 * e + 1], acc);
                    }
                }
            }
            const long long p = static_cast<long long>(img) * kPix + py * kW + px;
            y[p] = acc;
            if (target) {
                const float d = acc - target[p];
                dy[p] = 2.f * d / static_cast<float>(n);
                d2 += static_cast<double>(d) * d;
            }
            if (more) {
                __syncthreads();   // the pair's reads of rows r0-1 and r0 are done
                stash(nv, kNext, r0 + 3, 2);
                __syncthreads();
            }
        }
    }
    if (target) {
        __shared__ double s[kT / 64];
        d2 = wave_sum(d2);
        if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = d2;
        __syncthreads();
        if (threadIdx.x == 0) lpart[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
    }
}

// The same sums with one pixel per thread and its nine neighbours read from
// global memory: for small batches, whose few strips would leave most CUs
// idle (the two-level run's single window).  Per-block sums of (y-t)^2.
__global__ __launch_bounds__(kT) void k_out_fwd_px(const uint16_t* __restrict__ a5,
                                                const uint8_t* __restrict__ pack,
                                                const float* __restrict__ target, float* __restrict__ y,
                                                float* __restrict__ dy, double* __restrict__ lpart,
                                                int batch) {
    const float* W6 = reinterpret_cast<const float*>(pack + kOffW6);
    const float b6 = reinterpret_cast<const float*>(pack + kOffB6)[0];
    const long long n = static_cast<long long>(batch) * kPix;
    const long long p = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    double d2 = 0.0;
    if (p < n) {
        const int img = static_cast<int>(p / kPix), r = static_cast<int>(p % kPix);
        const int py = r / kW, px = r % kW;
        float acc = b6;
        for (int t = 0; t < 9; ++t) {
            const int yy = py + t / 3 - 1, xx = px + t % 3 - 1;
            if (yy < 0 || yy >= kW || xx < 0 || xx >= kW) continue;
            const uint4* src = reinterpret_cast<const uint4*>(
                a5 + (static_cast<size_t>(img) * kPix + yy * kW + xx) * kC);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint4 v = src[c];
                const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    acc = fmaf(bf2f(uint16_t(u[e] & 0xffff)), W6[t * kC + c * 8 + 2 * e], acc);
                    acc = fmaf(bf2f(uint16_t(u[e] >> 16)), W6[t * kC + c * 8 + 2 * e + 1], acc);
                }
            }
        }
        y[p] = acc;
        if (target) {
            const float d = acc - target[p];
            dy[p] = 2.f * d / static_cast<float>(n);
            d2 = static_cast<double>(d) * d;
        }
    }
    if (target) {
        __shared__ double s[kT / 64];
        d2 = wave_sum(d2);
        if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = d2;
        __syncthreads();
        if (threadIdx.x == 0) lpart[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
    }
}

// One workgroup: strided per-thread sums, then a fixed tree (deterministic).
__global__ __launch_bounds__(kT) void k_loss_final(const double* __restrict__ lpart, int nb,
                                                   long long n, double* __restrict__ loss) {
    __shared__ double s[kT];
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += kT) a += lpart[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int h = kT / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = s[0] / static_cast<double>(n);
}

// dz5[p][ci] = bf16( (sum_t dy[p - off(t)] W6[t][ci]) * gelu'(z5[p][ci]) )
__global__ __launch_bounds__(kT) void k_out_bwd_data(const float* __restrict__ dy,
                                                     const uint8_t* __restrict__ pack,
                                                     const uint16_t* __restrict__ z5,
                                                     uint16_t* __restrict__ dz5, int batch) {
    const float* W6 = reinterpret_cast<const float*>(pack + kOffW6);
    const long long n = static_cast<long long>(batch) * kPix;
    const long long p = static_cast<long long>(blockIdx.x) * kT + threadIdx.x;
    if (p >= n) return;
    const int img = static_cast<int>(p / kPix), r = static_cast<int>(p % kPix);
    const int py = r / kW, px = r % kW;
    float g[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int yy = py - (t / 3 - 1), xx = px - (t % 3 - 1);
        g[t] = (yy >= 0 && yy < kW && xx >= 0 && xx < kW)
                   ? dy[static_cast<size_t>(img) * kPix + yy * kW + xx] : 0.f;
    }
    const uint4* zs = reinterpret_cast<const uint4*>(z5 + static_cast<size_t>(p) * kC);
    uint4* out = reinterpret_cast<uint4*>(dz5 + static_cast<size_t>(p) * kC);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint4 zv = zs[c];
        const uint32_t zu[4] = {zv.x, zv.y, zv.z, zv.w};
        uint32_t ou[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint16_t h[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int ci = c * 8 + 2 * e + k;
                float da = 0.f;
#pragma unroll
                for (int t = 0; t < 9; ++t) da = fmaf(g[t], W6[t * kC + ci], da);
                const uint16_t zb = k ? uint16_t(zu[e] >> 16) : uint16_t(zu[e] & 0xffff);
                h[k] = f2bf(da * gelu_grad(bf2f(zb)));
            }
            ou[e] = h[0] | (uint32_t(h[1]) << 16);
        }
        out[c] = make_uint4(ou[0], ou[1], ou[2], ou[3]);
    }
}

// Per workgroup over its rows: dW6[t][ci] = sum a5[p+off(t)][ci] dy[p] and
// db6 = sum dy (slot 288).  partial[wg][289].  Thread (g, cp) owns channel
// pair 2cp, 2cp+1 and pixels 8g .. 8g+7 of each row: per row it reads ten
// pixels (two of halo) of each of the three staged a5 rows as channel pairs
// and its eight dy values, and accumulates all nine taps (144 FMAs), so all
// 256 threads work.  The next row's a5 rows and dy are loaded into
// registers while the current row is summed.  The 16 pixel groups are added
// in a fixed order at the end (deterministic).
__global__ __launch_bounds__(kT) void k_out_dw(const uint16_t* __restrict__ a5,
                                               const float* __restrict__ dy,
                                               float* __restrict__ partial, int rows,
                                               int rows_per_wg) {
    constexpr int kCols = kW + 2, kA = 3 * kCols * 4, kItA = (kA + kT - 1) / kT;
    __shared__ float s_dy[kW];
    __shared__ __attribute__((aligned(16))) uint16_t s_a[3 * kCols * kPitch];
    __shared__ float s_red[16][9][kC];
    const int tid = threadIdx.x;
    const int cp = tid & 15, g = tid >> 4;           // channel pair, pixel group
    const int r0 = blockIdx.x * rows_per_wg, r1 = min(r0 + rows_per_wg, rows);
    float acc[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t][0] = acc[t][1] = 0.f;
    float accb = 0.f;

    // a5 rows y-1 .. y+1 of row gr (zero outside the window) and its dy
    auto fetch = [&](int gr, uint4* va, float& dyv) {
        const int img = gr / kW, y = gr % kW;
        dyv = tid < kW ? dy[static_cast<size_t>(img) * kPix + y * kW + tid] : 0.f;
#pragma unroll
        for (int it = 0; it < kItA; ++it) {
            const int q = tid + it * kT;
            const int rr = q / (kCols * 4), rem = q % (kCols * 4);
            const int px = rem / 4 - 1, c = rem & 3;
            const int yy = y + rr - 1;
            va[it] = make_uint4(0, 0, 0, 0);
            if (q < kA && yy >= 0 && yy < kW && px >= 0 && px < kW)
                va[it] = *reinterpret_cast<const uint4*>(
                    a5 + (static_cast<size_t>(img) * kPix + yy * kW + px) * kC + c * 8);
        }
    };
    uint4 va[kItA];
    float dyv = 0.f;
    if (r0 < r1) fetch(r0, va, dyv);
    for (int gr = r0; gr < r1; ++gr) {
        __syncthreads();   // the previous row's reads are done
        if (tid < kW) s_dy[tid] = dyv;
#pragma unroll
        for (int it = 0; it < kItA; ++it) {
            const int q = tid + it * kT;
            if (q < kA) {
                const int rr = q / (kCols * 4), rem = q % (kCols * 4);
                const int col = rem / 4, c = rem & 3;
                uint2* d = reinterpret_cast<uint2*>(s_a + (rr * kCols + col) * kPitch + c * 8);
                d[0] = make_uint2(va[it].x, va[it].y);
                d[1] = make_uint2(va[it].z, va[it].w);
            }
        }
        __syncthreads();
        if (gr + 1 < r1) fetch(gr + 1, va, dyv);    // in flight while this row is summed
        float d[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] = s_dy[8 * g + j];
        if (cp == 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) accb += d[j];
        }
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
            // LDS column c holds pixel c - 1: pixel 8g + j + dx - 1 is column 8g + j + dx
            float a0[10], a1[10];
#pragma unroll
            for (int j = 0; j < 10; ++j) {
                const uint32_t v = *reinterpret_cast<const uint32_t*>(
                    s_a + (rr * kCols + 8 * g + j) * kPitch + 2 * cp);
                a0[j] = bf2f(uint16_t(v & 0xffff));
                a1[j] = bf2f(uint16_t(v >> 16));
            }
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    acc[rr * 3 + dx][0] = fmaf(a0[j + dx], d[j], acc[rr * 3 + dx][0]);
                    acc[rr * 3 + dx][1] = fmaf(a1[j + dx], d[j], acc[rr * 3 + dx][1]);
                }
        }
    }
    // the 16 pixel groups in a fixed order
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        s_red[g][t][2 * cp] = acc[t][0];
        s_red[g][t][2 * cp + 1] = acc[t][1];
    }
    __shared__ float s_b[16];
    if (cp == 0) s_b[g] = accb;
    __syncthreads();
    float* P = partial + static_cast<size_t>(blockIdx.x) * (9 * kC + 1);
    for (int e = tid; e < 9 * kC; e += kT) {
        const int t = e / kC, ci = e % kC;
        float sum = 0.f;
        for (int q = 0; q < 16; ++q) sum += s_red[q][t][ci];
        P[e] = sum;
    }
    if (tid == 0) {
        float sb = 0.f;
        for (int q = 0; q < 16; ++q) sb += s_b[q];
        P[9 * kC] = sb;
    }
}

__global__ void k_out_dw_reduce(const float* __restrict__ partial, int nwg,
                                float* __restrict__ grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > 9 * kC) return;
    const size_t stride = 9 * kC + 1;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int w = 0;
    for (; w + 4 <= nwg; w += 4) {
        s0 += partial[(w + 0) * stride + j];
        s1 += partial[(w + 1) * stride + j];
        s2 += partial[(w + 2) * stride + j];
        s3 += partial[(w + 3) * stride + j];
    }
    for (; w < nwg; ++w) s0 += partial[w * stride + j];
    grad[kL6 + j] = (s0 + s1) + (s2 + s3);
}

// ---- Adam ---------------------------------------------------------------------

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace
}  // namespace lpbf_amd

using namespace lpbf_amd;

namespace {

// Workspace: z1..z5 (act[0..4], training only), a1..a5 (act[5..9]), and for
// training two dz buffers, dy, the prediction, loss and gradient partials.
struct Offsets {
    size_t act[10], dz[2], dy, ypred, lpart, part, bytes;
};

constexpr size_t kAct = static_cast<size_t>(kPix) * kC * 2;   // bytes per window per tensor

int dw_nwg(int rows) { return rows < 512 ? rows : 512; }

Offsets offsets(int batch, bool training) {
    Offsets o{};
    size_t p = 0;
    const size_t A = kAct * batch;
    for (int i = 0; i < 10; ++i) {
        if (i < 5 && !training) { o.act[i] = SIZE_MAX; continue; }
        o.act[i] = p;
        p += A;
    }
    if (training) {
        o.dz[0] = p; p += A;
        o.dz[1] = p; p += A;
        o.dy = p; p += static_cast<size_t>(batch) * kPix * 4;
        o.ypred = p; p += static_cast<size_t>(batch) * kPix * 4;
        const size_t nb = (static_cast<size_t>(batch) * kPix + kT - 1) / kT;
        o.lpart = p; p += (nb * 8 + 255) / 256 * 256;
        o.part = p; p += static_cast<size_t>(dw_nwg(batch * kW)) * kDwSlots * kC * kC * 4;
    }
    o.bytes = p;
    return o;
}

struct Work {
    uint16_t* act[10];
    uint16_t* dz[2];
    float* dy;
    float* ypred;
    double* lpart;
    float* part;
};

Work layout(void* base, int batch, bool training) {
    const Offsets o = offsets(batch, training);
    uint8_t* b = static_cast<uint8_t*>(base);
    Work w{};
    for (int i = 0; i < 10; ++i)
        w.act[i] = o.act[i] == SIZE_MAX ? nullptr : reinterpret_cast<uint16_t*>(b + o.act[i]);
    if (training) {
        w.dz[0] = reinterpret_cast<uint16_t*>(b + o.dz[0]);
        w.dz[1] = reinterpret_cast<uint16_t*>(b + o.dz[1]);
        w.dy = reinterpret_cast<float*>(b + o.dy);
        w.ypred = reinterpret_cast<float*>(b + o.ypred);
        w.lpart = reinterpret_cast<double*>(b + o.lpart);
        w.part = reinterpret_cast<float*>(b + o.part);
    }
    return w;
}

bool batch_ok(int b) { return b >= 1 && b <= 4096; }

// Strip height of the conv launches: one row per workgroup while that
// gives at most ~1k workgroups (small batches are latency-bound and want
// every CU), taller strips beyond (fewer halo rows, loads overlapped with
// more MFMAs); a power of two dividing 128.
int conv_strip(int batch) {
    int s = 1;
    while (s < kW && static_cast<long long>(batch) * (kW / s) > 1024) s *= 2;
    return s;
}

// The output layer: strips of at least two rows (like conv_strip) once the
// batch fills the chip with them, one pixel per thread below.  Both sum each
// pixel's products in the same order.  Returns the number of loss partials.
int out_fwd(const uint16_t* a5, const uint8_t* pack, const float* target, float* y, float* dy,
            double* lpart, int batch, hipStream_t s) {
    const int strip = conv_strip(batch);
    if (strip == 1) {
        const int nb = static_cast<int>((static_cast<long long>(batch) * kPix + kT - 1) / kT);
        hipLaunchKernelGGL(k_out_fwd_px, dim3(nb), dim3(kT), 0, s, a5, pack, target, y, dy, lpart,
                           batch);
        return nb;
    }
    const int nw = batch * (kW / strip);
    hipLaunchKernelGGL(k_out_fwd, dim3(nw), dim3(kT), 0, s, a5, pack, target, y, dy, lpart, batch,
                       strip);
    return nw;
}

void conv_fwd(const uint16_t* X, const uint8_t* pack, int l, uint16_t* z, uint16_t* a, int batch,
              hipStream_t s) {
    const int strip = conv_strip(batch);
    const uint16_t* Wl =
        reinterpret_cast<const uint16_t*>(pack + kOffWF) + static_cast<size_t>(l) * kWsz;
    const float* bl = reinterpret_cast<const float*>(pack + kOffB) + l * kC;
    if (z)
        hipLaunchKernelGGL(k_conv_mfma<kModeFwdZ>, dim3(batch * (kW / strip)), dim3(kT), 0, s, X,
                           Wl, bl, static_cast<const uint16_t*>(nullptr), z, a, batch, strip);
    else
        hipLaunchKernelGGL(k_conv_mfma<kModeFwd>, dim3(batch * (kW / strip)), dim3(kT), 0, s, X,
                           Wl, bl, static_cast<const uint16_t*>(nullptr), z, a, batch, strip);
}

void forward_impl(const uint8_t* pack, const uint16_t* x0, int batch, Work const& w, hipStream_t s) {
    const uint16_t* in = x0;
    for (int l = 0; l < 5; ++l) {
        conv_fwd(in, pack, l, w.act[l], w.act[5 + l], batch, s);
        in = w.act[5 + l];
    }
}

}  // namespace

extern "C" {

size_t lpbf_sg_pack_bytes(void) { return kPackBytes; }

size_t lpbf_sg_work_bytes(int batch, int training) {
    if (!batch_ok(batch)) return 0;
    // sizes only: lay out from a null base
    return offsets(batch, training != 0).bytes;
}

void lpbf_sg_pack(const float* d_params, void* d_pack, hipStream_t stream) {
    if (!d_params || !d_pack) { record_error(LPBF_ERR_INVALID_ARGUMENT); return; }
    const int n = 5 * static_cast<int>(kWsz);
    hipLaunchKernelGGL(k_pack, dim3((n + kT - 1) / kT), dim3(kT), 0, stream, d_params,
                       static_cast<uint8_t*>(d_pack));
    check_launch();
}

void lpbf_sg_forward(const void* d_pack, const void* d_x0, int batch, void* d_work, float* d_y,
                     hipStream_t stream) {
    if (!d_pack || !d_x0 || !d_work || !d_y || !batch_ok(batch)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Work w = layout(d_work, batch, false);
    const uint8_t* pack = static_cast<const uint8_t*>(d_pack);
    forward_impl(pack, static_cast<const uint16_t*>(d_x0), batch, w, stream);
    out_fwd(w.act[9], pack, nullptr, d_y, nullptr, nullptr, batch, stream);
    check_launch();
}

void lpbf_sg_loss_grad(const void* d_pack, const void* d_x0, const float* d_target, int batch,
                       void* d_work, float* d_grad, double* d_loss, float* d_y,
                       hipStream_t stream) {
    if (!d_pack || !d_x0 || !d_target || !d_work || !d_grad || !d_loss || !batch_ok(batch)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Work w = layout(d_work, batch, true);
    const uint8_t* pack = static_cast<const uint8_t*>(d_pack);
    const uint16_t* x0 = static_cast<const uint16_t*>(d_x0);
    forward_impl(pack, x0, batch, w, stream);
    const long long n = static_cast<long long>(batch) * kPix;
    const int nb = static_cast<int>((n + kT - 1) / kT);
    float* y = d_y ? d_y : w.ypred;
    const int nwo = out_fwd(w.act[9], pack, d_target, y, w.dy, w.lpart, batch, stream);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(kT), 0, stream, w.lpart, nwo, n, d_loss);
    const int rows = batch * kW;
    const int nwg = dw_nwg(rows), per = (rows + nwg - 1) / nwg;
    // layer 6: weight gradient and dz5
    hipLaunchKernelGGL(k_out_dw, dim3(nwg), dim3(kT), 0, stream, w.act[9], w.dy, w.part, rows, per);
    hipLaunchKernelGGL(k_out_dw_reduce, dim3((9 * kC + 1 + 63) / 64), dim3(64), 0, stream, w.part,
                       nwg, d_grad);
    hipLaunchKernelGGL(k_out_bwd_data, dim3(nb), dim3(kT), 0, stream, w.dy, pack, w.act[4],
                       w.dz[0], batch);
    // layers 5..1
    int cur = 0;
    for (int l = 4; l >= 0; --l) {
        const uint16_t* a_prev = l > 0 ? w.act[5 + l - 1] : x0;
        hipLaunchKernelGGL(k_dw_mfma, dim3(nwg), dim3(kT), 0, stream, w.dz[cur], a_prev, w.part,
                           rows, per);
        hipLaunchKernelGGL(k_dw_reduce, dim3(kDwSlots * kC * kC / 32), dim3(kT), 0, stream,
                           w.part, nwg, l, d_grad);
        if (l > 0) {
            const int strip = conv_strip(batch);
            hipLaunchKernelGGL(k_conv_mfma<kModeBwd>, dim3(batch * (kW / strip)), dim3(kT), 0,
                               stream, w.dz[cur],
                               reinterpret_cast<const uint16_t*>(pack + kOffWB) +
                                   static_cast<size_t>(l) * kWsz,
                               static_cast<const float*>(nullptr), w.act[l - 1], w.dz[cur ^ 1],
                               static_cast<uint16_t*>(nullptr), batch, strip);
            cur ^= 1;
        }
    }
    check_launch();
}

void lpbf_sg_adam(float* d_params, const float* d_grad, float* d_m, float* d_v, int n, int t,
                  float lr, float beta1, float beta2, float eps, hipStream_t stream) {
    if (!d_params || !d_grad || !d_m || !d_v || n < 1 || t < 1) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const float c1 = 1.f - powf(beta1, static_cast<float>(t));
    const float c2 = 1.f - powf(beta2, static_cast<float>(t));
    hipLaunchKernelGGL(k_adam, dim3((n + kT - 1) / kT), dim3(kT), 0, stream, d_params, d_grad, d_m,
                       d_v, n, lr, beta1, beta2, eps, c1, c2);
    check_launch();
}


}  // extern "C"
