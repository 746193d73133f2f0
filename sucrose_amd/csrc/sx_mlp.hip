// Learned warm start for the ECP contact-parameter solve (no counterpart in
// the reference; it accelerates the lambda solve of force.f90:1046-1200).
//
// Network: 7 -> 64 -> 64 -> 64 -> 2, SiLU on the hidden layers.  Inputs are
// the unit separation in particle i's body frame and the relative rotation
// conj(q_i) q_j (w >= 0); outputs are lambda* and the contact distance over
// the pair's largest axis, 2 max(a).  Both outputs are independent of |r|.
//
// All products run on the matrix cores as v_mfma_f32_16x16x16_bf16 with FP32
// accumulation, 16 pairs per tile, in the transposed form Z^T = W A^T: the
// weights are the A operand (16 output units x 16 input units per tile), the
// activations of 16 pairs the B operand and the result lands in the layout a
// B operand needs.  Activations pass between layers through LDS (one wave
// per workgroup, so a workgroup barrier is only an ordering point).  FP32
// master weights live in global memory; the BF16 operand tiles are rebuilt
// from them after each optimiser step, pre-arranged so that every lane loads
// its 4 values of a tile with one 8-byte load:
//   forward tile (mo, ki):  lane l holds W[16mo + l%16][16ki + 4(l/16) + j]
//   backward tile (mi, ko): lane l holds W[16ko + 4(l/16) + j][16mi + l%16]
// Inference uses v_mfma_f32_16x16x32_bf16 for the 64-wide layers: its B
// operand takes 8 inputs per lane, which two result tiles (2ki, 2ki+1)
// provide as hidden units 32ki + 4(l/16) + j and 32ki + 16 + 4(l/16) + j
// (j < 4); the inference weight tiles use that same order of the inputs:
//   inference tile (mo, ki): lane l holds W[16mo + l%16][h(ki, l/16, e)],
//   h(ki, g, e) = 32ki + 4g + e (e < 4), 32ki + 16 + 4g + e - 4 (e >= 4)
// Training computes the loss gradient per workgroup into a private partial
// (weight gradients accumulate in LDS in a fixed order), a second kernel sums
// the partials in block order and applies Adam: the result is deterministic.
#include "sx_ecp.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int NPARAM = 8962;
constexpr int OFF_W1 = 0, OFF_B1 = 448, OFF_W2 = 512, OFF_B2 = 4608, OFF_W3 = 4672, OFF_B3 = 8768, OFF_W4 = 8832,
              OFF_B4 = 8960;
// per-workgroup gradient partial: gW1 (64 x 8: inputs 7 + 1 zero column),
// gW2, gW3 (64 x 64), gW4 (2 x 64), gb1..gb3 (64), gb4 (2), loss (1), pad
constexpr int G_W1 = 0, G_W2 = 512, G_W3 = 4608, G_W4 = 8704, G_B1 = 8832, G_B2 = 8896, G_B3 = 8960, G_B4 = 9024,
              G_LOSS = 9026, G_N = 9028;
// operand tiles (64 lanes x 4 BF16 each)
constexpr int F_L1 = 0, F_L2 = 4, F_L3 = 20, F_L4 = 36, F_N = 40;
constexpr int T_L2 = 0, T_L3 = 16, T_L4 = 32, T_N = 36;
constexpr int GROUPS_PER_BLOCK = 4;   // 4 x 16 = 64 samples per training workgroup

// inference tiles of the 64-wide products for v_mfma_f32_16x16x32_bf16
// (layers 2, 3: 4 x 2 tiles, layer 4: 1 x 2)
constexpr int I_L2 = 0, I_L3 = 8, I_L4 = 16, I_N = 18;

typedef short bf16x4 __attribute__((__vector_size__(4 * sizeof(short))));
typedef short short8 __attribute__((__vector_size__(8 * sizeof(short))));
typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));

}  // namespace

struct MlpState {
    float *p = nullptr, *m = nullptr, *v = nullptr;   // FP32 master parameters, Adam moments
    bf16x4 *wf = nullptr, *wt = nullptr;              // forward / backward operand tiles
    bf16x8 *wi = nullptr;                             // inference tiles (K = 32)
    float *partial = nullptr;                         // [blocks][G_N]
    int partial_blocks = 0;
    float *loss = nullptr;                            // per-step loss trace (device)
    long loss_cap = 0;
    long adam_t = 0;
    // training samples: X [cap][8] (7 inputs + 0), T [cap][2]
    float *X = nullptr, *T = nullptr;
    long ns = 0, cap = 0;
    long *d_ns = nullptr;
    // per work-list entry predictions of the current force call
    float *lam0 = nullptr;
    long lam0_cap = 0;
    // scratch for host-driven evaluation
    float *io = nullptr;
    long io_cap = 0;
    int *idx = nullptr;
    long idx_cap = 0;
};

static MlpState *mlp_of(sx_ctx *c)
{
    if (!c->mlp) throw sx_error{SX_ERR_STATE, "lambda* network not initialised (sx_mlp_init)"};
    return (MlpState *)c->mlp;
}

// ---------------------------------------------------------------------------
// device building blocks
// ---------------------------------------------------------------------------
// round to BF16, nearest with ties to even (one v_cvt_pk_bf16_f32)
__device__ __forceinline__ short to_bf16(float f)
{
    __bf16 b = (__bf16)f;
    return __builtin_bit_cast(short, b);
}

// SiLU and its derivative with the hardware exp and reciprocal: their ~1 ulp
// error is far below the BF16 step the results are rounded to
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }

__device__ __forceinline__ float dsilu(float z)
{
    float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

// out[n][16mo + r] = sum_k W[16mo + r][k] * act(in[n][k]) + bias, for the 16
// pairs n of a group; MT x KT tiles.  act: 0 identity, 1 SiLU of a stored
// pre-activation.  nbias: number of outputs that carry a bias.
__device__ __forceinline__ void dense(const bf16x4 *tiles, int MT, int KT, const float *in, int in_ld, int act,
                                      float *out, int out_ld, const float *bias, int nbias, int lane)
{
    const int n = lane & 15, g = lane >> 4;
    for (int mo = 0; mo < MT; ++mo) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ki = 0; ki < KT; ++ki) {
            bf16x4 a = tiles[(mo * KT + ki) * 64 + lane];
            const float *src = in + n * in_ld + 16 * ki + 4 * g;
            bf16x4 b;
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = to_bf16(act ? silu(src[j]) : src[j]);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int m = 16 * mo + 4 * g + j;
            out[n * out_ld + m] = acc[j] + ((bias && m < nbias) ? bias[m] : 0.0f);
        }
    }
}

// gW[o][k] += sum_n dz[n][o] * act(h[n][k]) over the 16 pairs of a group
// (pairs along the product's inner dimension); only o < rows, k < cols kept
__device__ __forceinline__ void wgrad(const float *dz, int dz_ld, int MT, const float *h, int h_ld, int act, int KT,
                                      float *gW, int gW_ld, int rows, int cols, int lane)
{
    const int n = lane & 15, g = lane >> 4;
    for (int mo = 0; mo < MT; ++mo)
        for (int ki = 0; ki < KT; ++ki) {
            bf16x4 a, b;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a[j] = to_bf16(dz[(4 * g + j) * dz_ld + 16 * mo + n]);
                float hv = h[(4 * g + j) * h_ld + 16 * ki + n];
                b[j] = to_bf16(act ? silu(hv) : hv);
            }
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int o = 16 * mo + 4 * g + j, k = 16 * ki + n;
                if (o < rows && k < cols) gW[o * gW_ld + k] += acc[j];
            }
        }
}

// network inputs of pair (i, j); RIJ = r_i - r_j minimum-imaged (FP64), the
// rest in FP32, which is far finer than the BF16 the network rounds them to
__device__ __forceinline__ void mlp_features(const DState &s, int i, int j, const double RIJ[3], float x[7])
{
    float r0 = (float)RIJ[0], r1 = (float)RIJ[1], r2 = (float)RIJ[2];
    float inv = rsqrtf(r0 * r0 + r1 * r1 + r2 * r2);
    float rh0 = r0 * inv, rh1 = r1 * inv, rh2 = r2 * inv;
    float w = (float)s.f[SX_QW][i], qx = (float)s.f[SX_QX][i], qy = (float)s.f[SX_QY][i], qz = (float)s.f[SX_QZ][i];
    // U = R(q)^T rows applied to rhat (force.f90:1427-1435)
    float u1 = (w * w + qx * qx - qy * qy - qz * qz) * rh0 + 2.0f * (qx * qy + w * qz) * rh1 +
               2.0f * (qx * qz - w * qy) * rh2;
    float u2 = 2.0f * (qx * qy - w * qz) * rh0 + (w * w - qx * qx + qy * qy - qz * qz) * rh1 +
               2.0f * (qy * qz + w * qx) * rh2;
    float u3 = 2.0f * (qx * qz + w * qy) * rh0 + 2.0f * (qy * qz - w * qx) * rh1 +
               (w * w - qx * qx - qy * qy + qz * qz) * rh2;
    float aw = w, ax = -qx, ay = -qy, az = -qz;
    float bw = (float)s.f[SX_QW][j], bx = (float)s.f[SX_QX][j], by = (float)s.f[SX_QY][j], bz = (float)s.f[SX_QZ][j];
    float rw = aw * bw - ax * bx - ay * by - az * bz;
    float rx = aw * bx + ax * bw + ay * bz - az * by;
    float ry = aw * by - ax * bz + ay * bw + az * bx;
    float rz = aw * bz + ax * by - ay * bx + az * bw;
    float sg = rw < 0.0f ? -1.0f : 1.0f;
    x[0] = u1; x[1] = u2; x[2] = u3;
    x[3] = sg * rw; x[4] = sg * rx; x[5] = sg * ry; x[6] = sg * rz;
}

// forward pass of one 16-pair group whose inputs sit in X[16][16]; leaves the
// pre-activations in Z1..Z3 [16][64] and the outputs in Y[16][16]
__device__ __forceinline__ void forward_group(const bf16x4 *wf, const float *p, const float *X, float *Z1, float *Z2,
                                              float *Z3, float *Y, int lane)
{
    dense(wf + F_L1 * 64, 4, 1, X, 16, 0, Z1, 64, p + OFF_B1, 64, lane);
    __syncthreads();
    dense(wf + F_L2 * 64, 4, 4, Z1, 64, 1, Z2, 64, p + OFF_B2, 64, lane);
    __syncthreads();
    dense(wf + F_L3 * 64, 4, 4, Z2, 64, 1, Z3, 64, p + OFF_B3, 64, lane);
    __syncthreads();
    dense(wf + F_L4 * 64, 1, 4, Z3, 64, 1, Y, 16, p + OFF_B4, 2, lane);
    __syncthreads();
}

// Inference form of the same forward pass with the activations kept in
// registers: a product's result tile (lane l: rows 4(l/16)..+3 of pair l%16)
// is the operand a next-layer tile needs (two result tiles feed one K = 32
// operand, in the order of the inference tiles), so bias, SiLU and the BF16
// rounding happen in place and no LDS or barrier is involved.  w1: layer-1
// tiles (K = 16), wi: inference tiles; bias: b1, b2, b3 (64 each) and b4 (2)
// at bias + {0, 64, 128, 192}; xb: this lane's layer-1 operand (inputs
// 4(l/16)..+3 of pair l%16).  Returns output rows 4(l/16)..+3 of pair l%16.
__device__ __forceinline__ f32x4 forward_regs(const bf16x4 *w1, const bf16x8 *wi, const float *bias, bf16x4 xb,
                                              int lane)
{
    const int g = lane >> 4;
    bf16x4 h[4], h2[4];
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[mo * 64 + lane], xb, acc, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) h[mo][j] = to_bf16(silu(acc[j] + bias[16 * mo + 4 * g + j]));
    }
    // the K = 32 operand: two result tiles joined as one 8-element vector (a
    // whole-vector shuffle and bit cast; assembling a BF16 vector element by
    // element from 16-bit integers is lowered wrongly by the compiler)
    auto operand = [](const bf16x4 &lo, const bf16x4 &hi) {
        short8 s8 = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
        return __builtin_bit_cast(bf16x8, s8);
    };
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const bf16x8 *tiles = wi + (layer == 0 ? I_L2 : I_L3) * 64;
        const float *b = bias + 64 * (layer + 1);
        bf16x8 in0 = operand(h[0], h[1]), in1 = operand(h[2], h[3]);
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tiles[(mo * 2 + 0) * 64 + lane], in0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tiles[(mo * 2 + 1) * 64 + lane], in1, acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j) h2[mo][j] = to_bf16(silu(acc[j] + b[16 * mo + 4 * g + j]));
        }
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) h[mo] = h2[mo];
    }
    f32x4 y = {0.f, 0.f, 0.f, 0.f};
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wi[(I_L4 + 0) * 64 + lane], operand(h[0], h[1]), y, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wi[(I_L4 + 1) * 64 + lane], operand(h[2], h[3]), y, 0, 0, 0);
    if (g == 0) {
        y[0] += bias[192];
        y[1] += bias[193];
    }
    return y;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// operand tiles from the FP32 master parameters (padding entries stay zero)
__global__ void k_mlp_pack(const float *p, bf16x4 *wf, bf16x4 *wt, bf16x8 *wi)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NPARAM) return;
    int layer, m, k;
    if (q < OFF_B1) { layer = 0; m = q / 7; k = q % 7; }
    else if (q >= OFF_W2 && q < OFF_B2) { layer = 1; m = (q - OFF_W2) / 64; k = (q - OFF_W2) % 64; }
    else if (q >= OFF_W3 && q < OFF_B3) { layer = 2; m = (q - OFF_W3) / 64; k = (q - OFF_W3) % 64; }
    else if (q >= OFF_W4 && q < OFF_B4) { layer = 3; m = (q - OFF_W4) / 64; k = (q - OFF_W4) % 64; }
    else return;   // biases are read in FP32
    short v = to_bf16(p[q]);
    const int fbase[4] = {F_L1, F_L2, F_L3, F_L4};
    const int KT = layer == 0 ? 1 : 4;
    int tile = fbase[layer] + (m / 16) * KT + k / 16;
    int lane = (m % 16) + 16 * ((k % 16) / 4);
    ((short *)wf)[(tile * 64 + lane) * 4 + (k % 4)] = v;
    if (layer >= 1) {
        const int ibase[4] = {0, I_L2, I_L3, I_L4};
        int ki = k / 32, rr = k % 32, g, e;
        if (rr < 16) { g = rr / 4; e = rr % 4; } else { g = (rr - 16) / 4; e = 4 + (rr - 16) % 4; }
        int it = ibase[layer] + (m / 16) * 2 + ki;
        int il = (m % 16) + 16 * g;
        ((short *)wi)[(it * 64 + il) * 8 + e] = v;   // one 2-byte store per parameter
    }
    if (layer >= 1) {
        const int tbase[4] = {0, T_L2, T_L3, T_L4};
        const int KTo = layer == 3 ? 1 : 4;
        int t = tbase[layer] + (k / 16) * KTo + m / 16;
        int tl = (k % 16) + 16 * ((m % 16) / 4);
        ((short *)wt)[(t * 64 + tl) * 4 + (m % 4)] = v;
    }
}

// Loss gradient of one minibatch: workgroup b takes 64 samples; sample
// indices come from idx (test batches, -1 = padding) or are drawn with
// Philox from (sample slot, step, seed).  Writes the workgroup's gradient
// partial (and loss) to partial[b].
__global__ void __launch_bounds__(64) k_mlp_grad(const float *p, const bf16x4 *wf, const bf16x4 *wt, const float *X,
                                                 const float *T, const long *ns, const int *idx, long nidx, int step,
                                                 unsigned seed, float w0, float w1, float inv_batch, float *partial)
{
    __shared__ float sX[16 * 16], sZ1[16 * 64], sZ2[16 * 64], sZ3[16 * 64], sY[16 * 16];
    __shared__ float sA[16 * 64], sB[16 * 64];
    __shared__ float gW1[64 * 8], gW2[64 * 64], gW3[64 * 64], gW4[2 * 64];
    const int lane = threadIdx.x;
    for (int q = lane; q < 64 * 8; q += 64) gW1[q] = 0.f;
    for (int q = lane; q < 64 * 64; q += 64) { gW2[q] = 0.f; gW3[q] = 0.f; }
    for (int q = lane; q < 2 * 64; q += 64) gW4[q] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gb4 = 0.f, lsum = 0.f;
    const long nsamp = *ns;
    __syncthreads();
    for (int grp = 0; grp < GROUPS_PER_BLOCK; ++grp) {
        // ---- inputs and targets of the group ----
        const int n = lane & 15, g = lane >> 4;
        long slot = (long)blockIdx.x * (16 * GROUPS_PER_BLOCK) + grp * 16 + n;
        long sidx = -1;
        if (idx) {
            if (slot < nidx) sidx = idx[slot];
        } else if (nsamp > 0) {
            uint32_t c[4] = {(uint32_t)slot, (uint32_t)step, seed, 0x4D4C5031u};
            sx_philox(c);
            sidx = (long)(((uint64_t)c[0] << 32 | c[1]) % (uint64_t)nsamp);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int k = 4 * g + j;
            sX[n * 16 + k] = (sidx >= 0 && k < 7) ? X[sidx * 8 + k] : 0.f;
        }
        __syncthreads();
        forward_group(wf, p, sX, sZ1, sZ2, sZ3, sY, lane);
        // ---- loss and dY (reuse sY rows: columns 0,1 become dY, rest 0) ----
        if (lane < 16) {
            float d0 = 0.f, d1 = 0.f;
            if (sidx >= 0) {
                float e0 = sY[n * 16 + 0] - T[sidx * 2 + 0];
                float e1 = sY[n * 16 + 1] - T[sidx * 2 + 1];
                lsum += w0 * e0 * e0 + w1 * e1 * e1;
                d0 = 2.f * w0 * e0 * inv_batch;
                d1 = 2.f * w1 * e1 * inv_batch;
            }
            for (int k = 0; k < 16; ++k) sY[n * 16 + k] = 0.f;
            sY[n * 16 + 0] = d0;
            sY[n * 16 + 1] = d1;
        }
        __syncthreads();
        // ---- layer 4 ----
        wgrad(sY, 16, 1, sZ3, 64, 1, 4, gW4, 64, 2, 64, lane);
        if (lane < 2) for (int q = 0; q < 16; ++q) gb4 += sY[q * 16 + lane];
        dense(wt + T_L4 * 64, 4, 1, sY, 16, 0, sA, 64, nullptr, 0, lane);   // dH3
        __syncthreads();
        for (int q = lane; q < 16 * 64; q += 64) sA[q] = sA[q] * dsilu(sZ3[q]);   // dZ3
        __syncthreads();
        // ---- layer 3 ----
        wgrad(sA, 64, 4, sZ2, 64, 1, 4, gW3, 64, 64, 64, lane);
        for (int q = 0; q < 16; ++q) gb3 += sA[q * 64 + lane];
        dense(wt + T_L3 * 64, 4, 4, sA, 64, 0, sB, 64, nullptr, 0, lane);   // dH2
        __syncthreads();
        for (int q = lane; q < 16 * 64; q += 64) sB[q] = sB[q] * dsilu(sZ2[q]);   // dZ2
        __syncthreads();
        // ---- layer 2 ----
        wgrad(sB, 64, 4, sZ1, 64, 1, 4, gW2, 64, 64, 64, lane);
        for (int q = 0; q < 16; ++q) gb2 += sB[q * 64 + lane];
        dense(wt + T_L2 * 64, 4, 4, sB, 64, 0, sA, 64, nullptr, 0, lane);   // dH1
        __syncthreads();
        for (int q = lane; q < 16 * 64; q += 64) sA[q] = sA[q] * dsilu(sZ1[q]);   // dZ1
        __syncthreads();
        // ---- layer 1 ----
        wgrad(sA, 64, 4, sX, 16, 0, 1, gW1, 8, 64, 8, lane);
        for (int q = 0; q < 16; ++q) gb1 += sA[q * 64 + lane];
        __syncthreads();
    }
    // ---- partial out ----
    float *out = partial + (long)blockIdx.x * G_N;
    for (int q = lane; q < 64 * 8; q += 64) out[G_W1 + q] = gW1[q];
    for (int q = lane; q < 64 * 64; q += 64) { out[G_W2 + q] = gW2[q]; out[G_W3 + q] = gW3[q]; }
    for (int q = lane; q < 2 * 64; q += 64) out[G_W4 + q] = gW4[q];
    out[G_B1 + lane] = gb1;
    out[G_B2 + lane] = gb2;
    out[G_B3 + lane] = gb3;
    if (lane < 2) out[G_B4 + lane] = gb4;
    for (int off = 8; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (lane == 0) out[G_LOSS] = lsum * inv_batch;
}

// gradient partial position of flat parameter q
__device__ __forceinline__ int grad_slot(int q)
{
    if (q < OFF_B1) return G_W1 + (q / 7) * 8 + q % 7;
    if (q < OFF_W2) return G_B1 + (q - OFF_B1);
    if (q < OFF_B2) return G_W2 + (q - OFF_W2);
    if (q < OFF_W3) return G_B2 + (q - OFF_B2);
    if (q < OFF_B3) return G_W3 + (q - OFF_W3);
    if (q < OFF_W4) return G_B3 + (q - OFF_B3);
    if (q < OFF_B4) return G_W4 + (q - OFF_W4);
    return G_B4 + (q - OFF_B4);
}

// sums the workgroup partials in order; with lr > 0 applies one Adam step
// (Kingma & Ba, Algorithm 1); grad (may be NULL) receives the summed gradient
__global__ void k_mlp_adam(float *p, float *m, float *v, const float *partial, int nblocks, float lr, float b1c,
                           float b2c, float *grad, float *loss_out)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0 && loss_out) {
        float l = 0.f;
        for (int b = 0; b < nblocks; ++b) l += partial[(long)b * G_N + G_LOSS];
        *loss_out = l;
    }
    if (q >= NPARAM) return;
    int s = grad_slot(q);
    float g = 0.f;
    for (int b = 0; b < nblocks; ++b) g += partial[(long)b * G_N + s];
    if (grad) grad[q] = g;
    if (lr > 0.f) {
        const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
        float mq = beta1 * m[q] + (1.f - beta1) * g;
        float vq = beta2 * v[q] + (1.f - beta2) * g * g;
        m[q] = mq;
        v[q] = vq;
        float mh = mq / b1c, vh = vq / b2c;
        p[q] = p[q] - lr * mh / (sqrtf(vh) + eps);
    }
}

// network outputs for host-given inputs X [n][8] -> Y [n][2]; one wave per
// 16 inputs (the inference path: activations in registers)
__global__ void __launch_bounds__(256) k_mlp_forward(const float *p, const bf16x4 *wf, const bf16x8 *wi, const float *X,
                                                     long n, float *Y)
{
    __shared__ float sp[3 * 64 + 2];
    for (int q = threadIdx.x; q < 64; q += blockDim.x) {
        sp[q] = p[OFF_B1 + q];
        sp[64 + q] = p[OFF_B2 + q];
        sp[128 + q] = p[OFF_B3 + q];
    }
    if (threadIdx.x < 2) sp[192 + threadIdx.x] = p[OFF_B4 + threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
    for (long base = wave * 16; base < n; base += nwaves * 16) {
        long t = base + r;
        bf16x4 xb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int k = 4 * g + j;
            xb[j] = to_bf16((t < n && k < 7) ? X[t * 8 + k] : 0.f);
        }
        f32x4 y = forward_regs(wf + F_L1 * 64, wi, sp, xb, lane);
        if (g == 0 && t < n) {
            Y[t * 2 + 0] = y[0];
            Y[t * 2 + 1] = y[1];
        }
    }
}

// warm starts for the compacted ECP work list of the current force call: one
// workgroup of 4 waves per segment, the waves take its groups of 16 live
// pairs in turn; operand tiles and biases are staged in LDS once per
// workgroup; each input is computed by the lanes that hold it as an operand
__global__ void __launch_bounds__(256) k_mlp_lambda0(DState s, const float *p, const bf16x4 *wf, const bf16x8 *wi,
                                                     const int *half_i, const int *half_j, const int *work,
                                                     const int *segcount, const int *nnz_half, float *lam0)
{
    const int cnt = segcount[blockIdx.x];
    if (cnt == 0) return;
    __shared__ bf16x4 s1[4 * 64];
    __shared__ bf16x8 si[I_N * 64];
    __shared__ float sp[3 * 64 + 2];
    for (int q = threadIdx.x; q < 4 * 64; q += blockDim.x) s1[q] = wf[F_L1 * 64 + q];
    for (int q = threadIdx.x; q < I_N * 64; q += blockDim.x) si[q] = wi[q];
    if (threadIdx.x < 64) {
        sp[threadIdx.x] = p[OFF_B1 + threadIdx.x];
        sp[64 + threadIdx.x] = p[OFF_B2 + threadIdx.x];
        sp[128 + threadIdx.x] = p[OFF_B3 + threadIdx.x];
        if (threadIdx.x < 2) sp[192 + threadIdx.x] = p[OFF_B4 + threadIdx.x];
    }
    __syncthreads();
    long start, end;
    ecp_segment(*nnz_half, blockIdx.x, start, end);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int base = wave * 16; base < cnt; base += 64) {
        int u = base + r;
        float x[7] = {0, 0, 0, 0, 0, 0, 0};
        if (u < cnt && g < 2) {
            int k = work[start + u];
            int a = half_i[k], b = half_j[k];
            double RIJ[3];
            pair_rij(s, a, b, RIJ);
            mlp_features(s, a, b, RIJ, x);
        }
        bf16x4 xb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int k = 4 * g + j;
            xb[j] = to_bf16(k < 7 ? x[k] : 0.f);
        }
        f32x4 y = forward_regs(s1, si, sp, xb, lane);
        if (g == 0 && u < cnt) lam0[start + u] = fminf(fmaxf(y[0], 0.02f), 0.98f);
    }
    (void)end;
}

// forward_regs for NG independent groups of 16 pairs at once: each weight
// tile is loaded once and its NG products issue back to back, so the matrix
// cores work on NG independent chains.  Each group's operations are those of
// forward_regs, so its outputs are the same.
template <int NG>
__device__ __forceinline__ void forward_regs_n(const bf16x4 *w1, const bf16x8 *wi, const float *bias,
                                               const bf16x4 (&xb)[NG], int lane, f32x4 (&y)[NG])
{
    const int g = lane >> 4;
    bf16x4 h[NG][4], h2[NG][4];
    auto operand = [](const bf16x4 &lo, const bf16x4 &hi) {
        short8 s8 = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
        return __builtin_bit_cast(bf16x8, s8);
    };
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) {
        const bf16x4 wt = w1[mo * 64 + lane];
        f32x4 acc[NG];
#pragma unroll
        for (int n = 0; n < NG; ++n) {
            acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
            acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wt, xb[n], acc[n], 0, 0, 0);
        }
#pragma unroll
        for (int n = 0; n < NG; ++n)
#pragma unroll
            for (int j = 0; j < 4; ++j) h[n][mo][j] = to_bf16(silu(acc[n][j] + bias[16 * mo + 4 * g + j]));
    }
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const bf16x8 *tiles = wi + (layer == 0 ? I_L2 : I_L3) * 64;
        const float *b = bias + 64 * (layer + 1);
        bf16x8 in0[NG], in1[NG];
#pragma unroll
        for (int n = 0; n < NG; ++n) {
            in0[n] = operand(h[n][0], h[n][1]);
            in1[n] = operand(h[n][2], h[n][3]);
        }
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) {
            const bf16x8 t0 = tiles[(mo * 2 + 0) * 64 + lane], t1 = tiles[(mo * 2 + 1) * 64 + lane];
            f32x4 acc[NG];
#pragma unroll
            for (int n = 0; n < NG; ++n) {
                acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
                acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(t0, in0[n], acc[n], 0, 0, 0);
            }
#pragma unroll
            for (int n = 0; n < NG; ++n) acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(t1, in1[n], acc[n], 0, 0, 0);
#pragma unroll
            for (int n = 0; n < NG; ++n)
#pragma unroll
                for (int j = 0; j < 4; ++j) h2[n][mo][j] = to_bf16(silu(acc[n][j] + b[16 * mo + 4 * g + j]));
        }
#pragma unroll
        for (int n = 0; n < NG; ++n)
#pragma unroll
            for (int mo = 0; mo < 4; ++mo) h[n][mo] = h2[n][mo];
    }
    const bf16x8 o0 = wi[(I_L4 + 0) * 64 + lane], o1 = wi[(I_L4 + 1) * 64 + lane];
#pragma unroll
    for (int n = 0; n < NG; ++n) {
        y[n] = f32x4{0.f, 0.f, 0.f, 0.f};
        y[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o0, operand(h[n][0], h[n][1]), y[n], 0, 0, 0);
    }
#pragma unroll
    for (int n = 0; n < NG; ++n) {
        y[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o1, operand(h[n][2], h[n][3]), y[n], 0, 0, 0);
        if (g == 0) {
            y[n][0] += bias[192];
            y[n][1] += bias[193];
        }
    }
}

// The learned solve stage in one kernel: one wavefront per work-list segment
// (as the plain solve stage);
// for each round of 64 pairs the network runs on four groups of 16 (its
// outputs are this kernel's warm starts, clamped to [0.02, 0.98] as in
// k_mlp_lambda0), the warm starts pass through LDS to the lanes that own the
// pairs, and every lane then finishes its pair as the plain solve stage does.
// stats (may be NULL): [0] solved, [1] iterations, [2] fallbacks, [3] pairs
// proven out of range by the warm start.
template <int VARIANT>
__global__ void __launch_bounds__(64) k_ecp_solve_learned(DState s, const double *pp, const float *p, const bf16x4 *wf,
                                                          const bf16x8 *wi, const int *half_i, const int *half_j,
                                                          const int *work, const int *segcount, const int *nnz_half,
                                                          EcpPar par, double *res, unsigned char *live,
                                                          long *stats, EcpSlots sl)
{
    const int cnt = segcount[blockIdx.x];
    if (cnt == 0) return;
    __shared__ float sp[3 * 64 + 2];
    __shared__ float l0s[64];
    __shared__ float xs[64 * 8];   // features of the round's 64 pairs, one row per pair
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    // operand tiles are read from global memory (each lane its own 8 or 16
    // bytes per tile; the 20 KB of tiles stay in the caches), biases from LDS
    const bf16x4 *s1 = wf + F_L1 * 64;
    const bf16x8 *si = wi;
    sp[lane] = p[OFF_B1 + lane];
    sp[64 + lane] = p[OFF_B2 + lane];
    sp[128 + lane] = p[OFF_B3 + lane];
    if (lane < 2) sp[192 + lane] = p[OFF_B4 + lane];
    __syncthreads();
    long start, end;
    ecp_segment(*nnz_half, blockIdx.x, start, end);
    long my_iters = 0, my_solved = 0, my_fb = 0, my_skip = 0;
    for (int base = 0; base < cnt; base += 64) {
        // every lane the features of its own pair (one round of loads), then
        // the network reads them from LDS in the operand layout
        const int u = base + lane;
        int k = 0, a = 0, b = 0;
        {
            float x[7] = {0, 0, 0, 0, 0, 0, 0};
            if (u < cnt) {
                k = work[start + u];
                a = half_i[k];
                b = half_j[k];
                double RIJ[3];
                pair_rij(s, a, b, RIJ);
                mlp_features(s, a, b, RIJ, x);
            }
#pragma unroll
            for (int j = 0; j < 7; ++j) xs[lane * 8 + j] = x[j];
            xs[lane * 8 + 7] = 0.f;
        }
        __syncthreads();
        {
            // the round's four groups of 16 pairs through the network together
            bf16x4 xb[4];
            f32x4 y[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int j = 0; j < 4; ++j) xb[q][j] = to_bf16(xs[(16 * q + r) * 8 + 4 * g + j]);
                if (g >= 2) xb[q] = bf16x4{0, 0, 0, 0};
            }
            forward_regs_n<4>(s1, si, sp, xb, lane, y);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (g == 0 && base + 16 * q < cnt) l0s[16 * q + r] = fminf(fmaxf(y[q][0], 0.02f), 0.98f);
        }
        __syncthreads();
        if (u < cnt) {
            int iters = 0, fb = 0;
            if (ecp_pair_solve<VARIANT>(s, pp, 0L, k, a, b, SX_ECP_LEARNED, true, (double)l0s[lane], par, res,
                                        live, sl, iters, fb)) {
                ++my_skip;
            } else {
                my_iters += iters;
                my_solved += 1;
                my_fb += fb;
            }
        }
        __syncthreads();
    }
    (void)end;
    if (!stats) return;
    for (int off = 32; off > 0; off >>= 1) {
        my_iters += __shfl_xor(my_iters, off, 64);
        my_solved += __shfl_xor(my_solved, off, 64);
        my_fb += __shfl_xor(my_fb, off, 64);
        my_skip += __shfl_xor(my_skip, off, 64);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&stats[0], (unsigned long long)my_solved);
        atomicAdd((unsigned long long *)&stats[1], (unsigned long long)my_iters);
        if (my_fb) atomicAdd((unsigned long long *)&stats[2], (unsigned long long)my_fb);
        if (my_skip) atomicAdd((unsigned long long *)&stats[3], (unsigned long long)my_skip);
    }
}

// training samples from the current work list: inputs, converged lambda*
// (Newton to 1e-12) and the contact distance over 2 max(a); slots are
// reserved per wave with one atomic on the sample counter
__global__ void __launch_bounds__(64) k_mlp_sample(DState s, const double *pp, const int *half_i, const int *half_j,
                                                   const int *work, const int *segcount, const int *nnz_half,
                                                   float *X, float *T, long *ns, long cap)
{
    long start, end;
    ecp_segment(*nnz_half, blockIdx.x, start, end);
    const int cnt = segcount[blockIdx.x];
    __shared__ long base;
    if (threadIdx.x == 0) base = cnt ? (long)atomicAdd((unsigned long long *)ns, (unsigned long long)cnt) : 0;
    __syncthreads();
    for (int u = threadIdx.x; u < cnt; u += blockDim.x) {
        long slot = base + u;
        if (slot >= cap) break;
        int k = work[start + u];
        int a = half_i[k], b = half_j[k];
        double RIJ[3];
        pair_rij(s, a, b, RIJ);
        float x[7];
        mlp_features(s, a, b, RIJ, x);
        Sym3 A = load_sym(pp, 0, PP_AR, a), B = load_sym(pp, 0, PP_AR, b);
        int it = 0, ok = 0;
        double lam = ecp_lambda_newton(A, B, RIJ, ECP_PSI, 1e-12, it, ok);
        double m = sqrt(RIJ[0] * RIJ[0] + RIJ[1] * RIJ[1] + RIJ[2] * RIJ[2]);
        double rh[3] = {RIJ[0] / m, RIJ[1] / m, RIJ[2] / m};
        double kv[3];
        sx_chol_solve((1 - lam) * A.a11 + lam * B.a11, (1 - lam) * A.a21 + lam * B.a21, (1 - lam) * A.a31 + lam * B.a31,
                      (1 - lam) * A.a22 + lam * B.a22, (1 - lam) * A.a32 + lam * B.a32, (1 - lam) * A.a33 + lam * B.a33,
                      rh, kv);
        double sig = 1.0 / sqrt(lam * (1 - lam) * (rh[0] * kv[0] + rh[1] * kv[1] + rh[2] * kv[2]));
        double amax = fmax(PPX(pp, PP_AMAX, a), PPX(pp, PP_AMAX, b));
        for (int q = 0; q < 7; ++q) X[slot * 8 + q] = x[q];
        X[slot * 8 + 7] = 0.f;
        T[slot * 2 + 0] = (float)lam;
        T[slot * 2 + 1] = (float)(sig / (2.0 * amax));
    }
    (void)end;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static void pack(sx_ctx *c, MlpState *M)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(sx_blocks(NPARAM, 256)), dim3(256), 0, c->stream, M->p, M->wf, M->wt, M->wi);
}

static void ensure_partial(MlpState *M, int blocks)
{
    if (blocks > M->partial_blocks) {
        if (M->partial) HIPCHK(hipFree(M->partial));
        HIPCHK(hipMalloc(&M->partial, sizeof(float) * G_N * (size_t)blocks));
        M->partial_blocks = blocks;
    }
}

static void ensure_io(MlpState *M, long floats)
{
    if (floats > M->io_cap) {
        if (M->io) HIPCHK(hipFree(M->io));
        HIPCHK(hipMalloc(&M->io, sizeof(float) * (size_t)floats));
        M->io_cap = floats;
    }
}

void sx_ecp_mlp_free(sx_ctx *c)
{
    MlpState *M = (MlpState *)c->mlp;
    if (!M) return;
    for (void *q : {(void *)M->p, (void *)M->m, (void *)M->v, (void *)M->wf, (void *)M->wt, (void *)M->wi,
                    (void *)M->partial,
                    (void *)M->loss, (void *)M->X, (void *)M->T, (void *)M->d_ns, (void *)M->lam0, (void *)M->io,
                    (void *)M->idx})
        if (q) hipFree(q);
    delete M;
    c->mlp = nullptr;
}

// predictions for the work list of the force call in progress
const float *sx_ecp_mlp_predict(sx_ctx *c)
{
    MlpState *M = mlp_of(c);
    long need = c->nl.cap_half;
    if (need > M->lam0_cap) {
        if (M->lam0) HIPCHK(hipFree(M->lam0));
        HIPCHK(hipMalloc(&M->lam0, sizeof(float) * (size_t)need));
        M->lam0_cap = need;
        c->have_graph = false;
    }
    hipLaunchKernelGGL(k_mlp_lambda0, dim3(ECP_NSEG), dim3(256), 0, c->stream, c->dstate(), M->p, M->wf, M->wi,
                       c->nl.half_i, c->nl.half_j, c->d_work, c->d_segcount, c->d_flag + 2, M->lam0);
    return M->lam0;
}

// the learned solve stage over the current work list (results, live flags
// and the apply stage's slots as the plain solve stage)
void sx_ecp_solve_learned(sx_ctx *c, int variant, const EcpPar &par, long *stats)
{
    const EcpSlots sl{c->d_ecp_slot};
    MlpState *M = mlp_of(c);
    NList &nl = c->nl;
#define SX_LEARNED(V)                                                                                               \
    hipLaunchKernelGGL(k_ecp_solve_learned<V>, dim3(ECP_NSEG), dim3(64), 0, c->stream, c->dstate(), c->d_ecp_pp,  \
                       M->p, M->wf, M->wi, nl.half_i, nl.half_j, c->d_work, c->d_segcount, c->d_flag + 2, par,      \
                       c->d_pairbuf, c->d_live, stats, sl)
    switch (variant) {
    case 0: SX_LEARNED(0); break;
    case 1: SX_LEARNED(1); break;
    case 2: SX_LEARNED(2); break;
    case 3: SX_LEARNED(3); break;
    default: SX_LEARNED(4); break;
    }
#undef SX_LEARNED
}

extern "C" int sx_mlp_init(sx_ctx *c, uint32_t seed)
{
    return guard(c, [&] {
        if (!c->mlp) {
            MlpState *M = new MlpState();
            c->mlp = M;
            HIPCHK(hipMalloc(&M->p, sizeof(float) * NPARAM));
            HIPCHK(hipMalloc(&M->m, sizeof(float) * NPARAM));
            HIPCHK(hipMalloc(&M->v, sizeof(float) * NPARAM));
            HIPCHK(hipMalloc(&M->wf, sizeof(bf16x4) * 64 * F_N));
            HIPCHK(hipMalloc(&M->wt, sizeof(bf16x4) * 64 * T_N));
            HIPCHK(hipMalloc(&M->wi, sizeof(bf16x8) * 64 * I_N));
            HIPCHK(hipMemsetAsync(M->wi, 0, sizeof(bf16x8) * 64 * I_N, c->stream));
            HIPCHK(hipMalloc(&M->d_ns, sizeof(long)));
            HIPCHK(hipMemsetAsync(M->wf, 0, sizeof(bf16x4) * 64 * F_N, c->stream));
            HIPCHK(hipMemsetAsync(M->wt, 0, sizeof(bf16x4) * 64 * T_N, c->stream));
            HIPCHK(hipMemsetAsync(M->d_ns, 0, sizeof(long), c->stream));
        }
        MlpState *M = (MlpState *)c->mlp;
        // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases, from
        // Philox words of (parameter index, seed)
        std::vector<float> h(NPARAM);
        for (int q = 0; q < NPARAM; ++q) {
            int fan_in = (q < OFF_W2) ? 7 : 64;
            uint32_t ctr[4] = {(uint32_t)q, seed, 0x4D4C5049u, 0u}, out[4];
            sx_philox4x32_10(ctr, out);
            double u = out[0] * (1.0 / 4294967296.0);
            h[q] = (float)((2.0 * u - 1.0) / std::sqrt((double)fan_in));
        }
        HIPCHK(hipMemcpyAsync(M->p, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(M->m, 0, sizeof(float) * NPARAM, c->stream));
        HIPCHK(hipMemsetAsync(M->v, 0, sizeof(float) * NPARAM, c->stream));
        M->adam_t = 0;
        pack(c, M);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int sx_mlp_set_params(sx_ctx *c, const float *params)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        if (!params) throw sx_error{SX_ERR_ARG, "params is NULL"};
        HIPCHK(hipMemcpyAsync(M->p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(M->m, 0, sizeof(float) * NPARAM, c->stream));
        HIPCHK(hipMemsetAsync(M->v, 0, sizeof(float) * NPARAM, c->stream));
        M->adam_t = 0;
        pack(c, M);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int sx_mlp_get_params(sx_ctx *c, float *params)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        HIPCHK(hipMemcpyAsync(params, M->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int sx_mlp_num_params(void) { return NPARAM; }

extern "C" int sx_mlp_forward(sx_ctx *c, long n, const float *X7, float *Y2)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        if (n <= 0 || !X7 || !Y2) throw sx_error{SX_ERR_ARG, "bad forward arguments"};
        std::vector<float> x8((size_t)n * 8, 0.f);
        for (long t = 0; t < n; ++t)
            for (int k = 0; k < 7; ++k) x8[t * 8 + k] = X7[t * 7 + k];
        ensure_io(M, n * 10);
        HIPCHK(hipMemcpyAsync(M->io, x8.data(), sizeof(float) * 8 * n, hipMemcpyHostToDevice, c->stream));
        int blocks = (int)std::min<long>((n + 63) / 64, 4096);
        hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, c->stream, M->p, M->wf, M->wi, M->io, n,
                           M->io + 8 * n);
        HIPCHK(hipMemcpyAsync(Y2, M->io + 8 * n, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

// loss and gradient of an explicit batch (host arrays X [n][7], T [n][2])
extern "C" int sx_mlp_loss_grad(sx_ctx *c, long n, const float *X7, const float *T2, const float w[2], float *loss,
                                float *grad)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        if (n <= 0 || !X7 || !T2 || !w || !grad) throw sx_error{SX_ERR_ARG, "bad loss_grad arguments"};
        std::vector<float> buf((size_t)n * 10, 0.f);
        for (long t = 0; t < n; ++t) {
            for (int k = 0; k < 7; ++k) buf[t * 8 + k] = X7[t * 7 + k];
            buf[8 * n + t * 2] = T2[t * 2];
            buf[8 * n + t * 2 + 1] = T2[t * 2 + 1];
        }
        ensure_io(M, n * 10 + NPARAM + 1);
        if (n > M->idx_cap) {
            if (M->idx) HIPCHK(hipFree(M->idx));
            HIPCHK(hipMalloc(&M->idx, sizeof(int) * (size_t)n));
            M->idx_cap = n;
        }
        std::vector<int> idx((size_t)n);
        for (long t = 0; t < n; ++t) idx[t] = (int)t;
        HIPCHK(hipMemcpyAsync(M->io, buf.data(), sizeof(float) * 10 * n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(M->idx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
        int blocks = (int)((n + 16 * GROUPS_PER_BLOCK - 1) / (16 * GROUPS_PER_BLOCK));
        ensure_partial(M, blocks);
        hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64), 0, c->stream, M->p, M->wf, M->wt, M->io, M->io + 8 * n,
                           (const long *)M->d_ns, M->idx, n, 0, 0u, w[0], w[1], 1.0f / (float)n, M->partial);
        float *d_grad = M->io + 10 * n;
        hipLaunchKernelGGL(k_mlp_adam, dim3(sx_blocks(NPARAM, 256)), dim3(256), 0, c->stream, M->p, M->m, M->v,
                           M->partial, blocks, 0.0f, 1.0f, 1.0f, d_grad, d_grad + NPARAM);
        HIPCHK(hipMemcpyAsync(grad, d_grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->stream));
        if (loss) HIPCHK(hipMemcpyAsync(loss, d_grad + NPARAM, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

void sx_ecp_prepare(sx_ctx *c, int kind, const double *p, int np);

extern "C" int sx_mlp_collect(sx_ctx *c, int kind, const double *params, int nparams, long capacity, long *total)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        if (capacity > M->cap) {
            float *X = nullptr, *T = nullptr;
            HIPCHK(hipMalloc(&X, sizeof(float) * 8 * (size_t)capacity));
            HIPCHK(hipMalloc(&T, sizeof(float) * 2 * (size_t)capacity));
            if (M->ns > 0) {
                HIPCHK(hipMemcpyAsync(X, M->X, sizeof(float) * 8 * M->ns, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(T, M->T, sizeof(float) * 2 * M->ns, hipMemcpyDeviceToDevice, c->stream));
            }
            HIPCHK(hipStreamSynchronize(c->stream));
            if (M->X) HIPCHK(hipFree(M->X));
            if (M->T) HIPCHK(hipFree(M->T));
            M->X = X;
            M->T = T;
            M->cap = capacity;
        }
        sx_ecp_prepare(c, kind, params, nparams);
        hipLaunchKernelGGL(k_mlp_sample, dim3(ECP_NSEG), dim3(64), 0, c->stream, c->dstate(), c->d_ecp_pp,
                           c->nl.half_i, c->nl.half_j, c->d_work, c->d_segcount, c->d_flag + 2, M->X, M->T, M->d_ns,
                           M->cap);
        long ns = 0;
        HIPCHK(hipMemcpyAsync(&ns, M->d_ns, sizeof(long), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (ns > M->cap) {
            ns = M->cap;
            HIPCHK(hipMemcpyAsync(M->d_ns, &ns, sizeof(long), hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        M->ns = ns;
        if (total) *total = ns;
        sx_check_device_error(c);
    });
}

extern "C" int sx_mlp_samples(sx_ctx *c, long first, long n, float *X7, float *T2)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        if (first < 0 || n < 0 || first + n > M->ns) throw sx_error{SX_ERR_ARG, "sample range out of bounds"};
        std::vector<float> x((size_t)n * 8);
        if (n) {
            HIPCHK(hipMemcpyAsync(x.data(), M->X + first * 8, sizeof(float) * 8 * n, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(T2, M->T + first * 2, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        for (long t = 0; t < n; ++t)
            for (int k = 0; k < 7; ++k) X7[t * 7 + k] = x[t * 8 + k];
    });
}

extern "C" int sx_mlp_clear_samples(sx_ctx *c)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        M->ns = 0;
        HIPCHK(hipMemsetAsync(M->d_ns, 0, sizeof(long), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

// steps Adam steps on minibatches of `batch` samples (a multiple of 64)
// drawn from the collected samples; loss_trace (may be NULL) gets each
// step's minibatch loss
extern "C" int sx_mlp_train(sx_ctx *c, int steps, int batch, double lr, const float w[2], uint32_t seed,
                            float *loss_trace)
{
    return guard(c, [&] {
        MlpState *M = mlp_of(c);
        if (steps <= 0 || batch <= 0 || batch % (16 * GROUPS_PER_BLOCK) || !w)
            throw sx_error{SX_ERR_ARG, "batch must be a positive multiple of 64"};
        if (M->ns <= 0) throw sx_error{SX_ERR_STATE, "no training samples (sx_mlp_collect)"};
        int blocks = batch / (16 * GROUPS_PER_BLOCK);
        ensure_partial(M, blocks);
        if (steps > M->loss_cap) {
            if (M->loss) HIPCHK(hipFree(M->loss));
            HIPCHK(hipMalloc(&M->loss, sizeof(float) * (size_t)steps));
            M->loss_cap = steps;
        }
        for (int t = 0; t < steps; ++t) {
            M->adam_t += 1;
            float b1c = (float)(1.0 - std::pow(0.9, (double)M->adam_t));
            float b2c = (float)(1.0 - std::pow(0.999, (double)M->adam_t));
            hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64), 0, c->stream, M->p, M->wf, M->wt, M->X, M->T,
                               (const long *)M->d_ns, (const int *)nullptr, 0L, (int)M->adam_t, seed, w[0], w[1],
                               1.0f / (float)batch, M->partial);
            hipLaunchKernelGGL(k_mlp_adam, dim3(sx_blocks(NPARAM, 256)), dim3(256), 0, c->stream, M->p, M->m, M->v,
                               M->partial, blocks, (float)lr, b1c, b2c, (float *)nullptr, M->loss + t);
            pack(c, M);
        }
        if (loss_trace)
            HIPCHK(hipMemcpyAsync(loss_trace, M->loss, sizeof(float) * steps, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        sx_check_device_error(c);
    });
}
