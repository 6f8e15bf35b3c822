// ipf_learned.hip -- the learned return-branch predictor: a 7 -> 64 -> 64 -> 6
// MLP (ReLU, BF16 weights and activations as MFMA operands, FP32
// accumulation) trained on the GPU on pairs the exact return map emits.
//
//   inputs  x = (s1/c, s2/c, s3/c, ln c, phi, psi, alpha): trial principal
//           stresses (descending, from the spectral decomposition of the
//           trial stress), material, accumulated plastic variable;
//   outputs 5 branch logits (elastic, main plane, right edge, left edge,
//           apex) and one regression output for the plastic increment
//           (1e3 * d_alpha of the exact return).
//
// Training: loss = mean cross-entropy of the branch + 0.1 * mean squared
// error of the increment.  One wavefront handles 16 samples at a time with
// __builtin_amdgcn_mfma_f32_16x16x16bf16_1k (16 x 16 x 16 tiles); every
// wavefront accumulates its weight gradients in registers and writes them
// to its own slot, a second kernel sums the slots in a fixed order (the
// gradient is reproducible) and applies Adam to FP32 master weights, then
// refreshes their BF16 copies.
//
// MFMA 16x16x16 operand layout (64 lanes, 4 values per lane):
//   A (16 x 16, row m, column k):  lane = m + 16 * (k / 4), value k % 4
//   B (16 x 16, row k, column n):  lane = n + 16 * (k / 4), value k % 4
//   C (16 x 16, row m, column n):  lane = n + 16 * (m / 4), value m % 4
#include "ipf_kernels.h"
#include "ipf_learned.h"

namespace ipf {

namespace {

constexpr int kWave = 64;
using namespace mlp;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ short f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(((uint32_t)(unsigned short)h) << 16); }

// ReLU on the bit pattern -- the int32 maximum with 0 keeps every
// non-negative float and maps negative ones and -0 to +0: one v_max_i32,
// where fmaxf compiles to two v_max_f32 (the first quiets a signalling NaN);
// inference 15 % faster.  A NaN with the sign bit clear passes through (its
// predictions are garbage either way; they only order work).  The derivative
// is 0 at and below 0.
__device__ __forceinline__ float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }
__device__ __forceinline__ float relu_grad(float x) { return x > 0.f ? 1.f : 0.f; }

// C += A (16 x K) B (K x 16); a(m, k), b(k, n) return BF16 bits.
template <class FA, class FB>
__device__ __forceinline__ f32x4 mma(FA a, FB b, int K, f32x4 c) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            av[i] = a(r, k0 + kq + i);
            bv[i] = b(k0 + kq + i, r);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, c, 0, 0, 0);
    }
    return c;
}

// Per-wavefront LDS of the forward / backward pass over 16 samples.
struct WaveLDS {
    short X[16][INP];       // inputs (BF16)
    float Z1[16][H];        // pre-activations (FP32)
    float Z2[16][H];
    short H1[16][H];        // activations (BF16)
    short H2[16][H];
    float Z3[16][OUTP];     // outputs
    short D3[16][OUTP];     // dL/dZ3 (BF16)
    short D2[16][H];        // dL/dZ2
    short D1[16][H];        // dL/dZ1
    float L[16];            // per-sample loss
};

struct WeightsLDS {
    short W1[INP][H];
    short W2[H][H];
    short W3[H][OUTP];
    float b1[H], b2[H], b3[OUTP];
};

__device__ void load_weights(WeightsLDS& w, const short* __restrict__ wbf, const float* __restrict__ par) {
    for (int i = threadIdx.x; i < INP * H; i += blockDim.x) (&w.W1[0][0])[i] = wbf[OFF_W1 + i];
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) (&w.W2[0][0])[i] = wbf[OFF_W2 + i];
    for (int i = threadIdx.x; i < H * OUTP; i += blockDim.x) (&w.W3[0][0])[i] = wbf[OFF_W3 + i];
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        w.b1[i] = par[OFF_B1 + i];
        w.b2[i] = par[OFF_B2 + i];
    }
    for (int i = threadIdx.x; i < OUTP; i += blockDim.x) w.b3[i] = par[OFF_B3 + i];
}

// forward pass of the wave's 16 samples starting at s0 (rows >= n are zero)
__device__ void forward16(const WeightsLDS& w, WaveLDS& s, const float* __restrict__ feat, int64_t s0, int64_t n) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    for (int t = lane; t < 16 * INP; t += kWave) {
        const int m = t / INP, k = t - m * INP;
        const int64_t q = s0 + m;
        s.X[m][k] = (k < IN && q < n) ? f2bf(feat[q * FEAT_STRIDE + k]) : (short)0;
    }
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < H / 16; nt++) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        c = mma([&](int m, int k) { return s.X[m][k]; }, [&](int k, int j) { return w.W1[k][nt * 16 + j]; }, INP, c);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float z = c[i] + w.b1[nt * 16 + r];
            s.Z1[kq + i][nt * 16 + r] = z;
            s.H1[kq + i][nt * 16 + r] = f2bf(relu(z));
        }
    }
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < H / 16; nt++) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        c = mma([&](int m, int k) { return s.H1[m][k]; }, [&](int k, int j) { return w.W2[k][nt * 16 + j]; }, H, c);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float z = c[i] + w.b2[nt * 16 + r];
            s.Z2[kq + i][nt * 16 + r] = z;
            s.H2[kq + i][nt * 16 + r] = f2bf(relu(z));
        }
    }
    __syncthreads();
    {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        c = mma([&](int m, int k) { return s.H2[m][k]; }, [&](int k, int j) { return w.W3[k][j]; }, H, c);
#pragma unroll
        for (int i = 0; i < 4; i++) s.Z3[kq + i][r] = c[i] + w.b3[r];
    }
    __syncthreads();
}

// One wavefront per 16-sample tile (grid-stride); 4 wavefronts per block.
// grad_slots[wave][NPAR] receives the wave's summed gradient of the mean
// loss over n samples, loss_slots[wave] its summed loss.
__global__ void __launch_bounds__(4 * kWave)
k_mlp_grad(const short* __restrict__ wbf, const float* __restrict__ par, const float* __restrict__ feat,
           const int* __restrict__ label, const float* __restrict__ target, int64_t n, float* __restrict__ grad_slots,
           double* __restrict__ loss_slots) {
    __shared__ WeightsLDS w;
    __shared__ WaveLDS wl[4];
    load_weights(w, wbf, par);
    __syncthreads();
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    WaveLDS& s = wl[wave];
    const int64_t gw = (int64_t)blockIdx.x * 4 + wave;
    const int64_t nw = (int64_t)gridDim.x * 4;
    const int64_t ntile = (n + 15) / 16;
    const float inv_n = 1.f / (float)n;
    f32x4 g1[H / 16], g2[H / 16][H / 16], g3[H / 16];
#pragma unroll
    for (int a = 0; a < H / 16; a++) {
        g1[a] = g3[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < H / 16; b++) g2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f;  // lane j: bias column j
    double loss = 0.0;
    // every wavefront runs the same number of rounds, so the block barriers
    // inside forward16 / the backward are reached by all four
    const int64_t rounds = (ntile + nw - 1) / nw;
    for (int64_t rd = 0; rd < rounds; rd++) {
        const int64_t tile = rd * nw + gw;
        const int64_t s0 = tile * 16;
        forward16(w, s, feat, s0, n);
        // loss and dL/dZ3 per sample (lanes 0..15 own one sample each)
        if (lane < 16) {
            const int64_t q = s0 + lane;
            float d[OUTP];
#pragma unroll
            for (int j = 0; j < OUTP; j++) d[j] = 0.f;
            float l = 0.f;
            if (q < n) {
                float zmax = s.Z3[lane][0];
#pragma unroll
                for (int j = 1; j < NBR; j++) zmax = fmaxf(zmax, s.Z3[lane][j]);
                float e[NBR], se = 0.f;
#pragma unroll
                for (int j = 0; j < NBR; j++) {
                    e[j] = expf(s.Z3[lane][j] - zmax);
                    se += e[j];
                }
                const int y = label[q];
                l = -(s.Z3[lane][y] - zmax - logf(se));
#pragma unroll
                for (int j = 0; j < NBR; j++) d[j] = (e[j] / se - (j == y ? 1.f : 0.f)) * inv_n;
                const float diff = s.Z3[lane][NBR] - target[q];
                l += REG_WEIGHT * diff * diff;
                d[NBR] = 2.f * REG_WEIGHT * diff * inv_n;
            }
            s.L[lane] = l;
#pragma unroll
            for (int j = 0; j < OUTP; j++) s.D3[lane][j] = f2bf(d[j]);
        }
        __syncthreads();
        if (lane == 0)
            for (int m = 0; m < 16; m++) loss += s.L[m];
        // layer 3: dW3 += H2^T D3, db3, dZ2 = (D3 W3^T) relu'(Z2)
#pragma unroll
        for (int mt = 0; mt < H / 16; mt++)
            g3[mt] = mma([&](int m, int k) { return s.H2[k][mt * 16 + m]; }, [&](int k, int j) { return s.D3[k][j]; }, 16,
                         g3[mt]);
        if (lane < OUTP)
            for (int m = 0; m < 16; m++) gb3 += bf2f(s.D3[m][lane]);
#pragma unroll
        for (int nt = 0; nt < H / 16; nt++) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
            c = mma([&](int m, int k) { return s.D3[m][k]; }, [&](int k, int j) { return w.W3[nt * 16 + j][k]; }, OUTP,
                    c);
#pragma unroll
            for (int i = 0; i < 4; i++) s.D2[kq + i][nt * 16 + r] = f2bf(c[i] * relu_grad(s.Z2[kq + i][nt * 16 + r]));
        }
        __syncthreads();
        // layer 2: dW2 += H1^T D2, db2, dZ1 = (D2 W2^T) relu'(Z1)
#pragma unroll
        for (int mt = 0; mt < H / 16; mt++)
#pragma unroll
            for (int nt = 0; nt < H / 16; nt++)
                g2[mt][nt] = mma([&](int m, int k) { return s.H1[k][mt * 16 + m]; },
                                 [&](int k, int j) { return s.D2[k][nt * 16 + j]; }, 16, g2[mt][nt]);
        for (int m = 0; m < 16; m++) gb2 += bf2f(s.D2[m][lane]);
#pragma unroll
        for (int nt = 0; nt < H / 16; nt++) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
            c = mma([&](int m, int k) { return s.D2[m][k]; }, [&](int k, int j) { return w.W2[nt * 16 + j][k]; }, H, c);
#pragma unroll
            for (int i = 0; i < 4; i++) s.D1[kq + i][nt * 16 + r] = f2bf(c[i] * relu_grad(s.Z1[kq + i][nt * 16 + r]));
        }
        __syncthreads();
        // layer 1: dW1 += X^T D1, db1
#pragma unroll
        for (int nt = 0; nt < H / 16; nt++)
            g1[nt] = mma([&](int m, int k) { return s.X[k][m]; }, [&](int k, int j) { return s.D1[k][nt * 16 + j]; }, 16,
                         g1[nt]);
        for (int m = 0; m < 16; m++) gb1 += bf2f(s.D1[m][lane]);
        __syncthreads();
    }
    // the wave's slot: C layout element (kq + i, r) of each 16 x 16 tile
    float* G = grad_slots + gw * (int64_t)NPAR;
#pragma unroll
    for (int nt = 0; nt < H / 16; nt++)
#pragma unroll
        for (int i = 0; i < 4; i++) G[OFF_W1 + (kq + i) * H + nt * 16 + r] = g1[nt][i];
#pragma unroll
    for (int mt = 0; mt < H / 16; mt++)
#pragma unroll
        for (int nt = 0; nt < H / 16; nt++)
#pragma unroll
            for (int i = 0; i < 4; i++) G[OFF_W2 + (mt * 16 + kq + i) * H + nt * 16 + r] = g2[mt][nt][i];
#pragma unroll
    for (int mt = 0; mt < H / 16; mt++)
#pragma unroll
        for (int i = 0; i < 4; i++) G[OFF_W3 + (mt * 16 + kq + i) * OUTP + r] = g3[mt][i];
    G[OFF_B1 + lane] = gb1;
    G[OFF_B2 + lane] = gb2;
    if (lane < OUTP) G[OFF_B3 + lane] = gb3;
    if (lane == 0) loss_slots[gw] = loss;
}

// Parameters a padded slot never trains: W1 rows >= IN, W3 / b3 columns >= OUT.
__device__ __forceinline__ bool trainable(int p) {
    if (p >= OFF_W1 && p < OFF_B1) return (p - OFF_W1) / H < IN;
    if (p >= OFF_W3 && p < OFF_B3) return (p - OFF_W3) % OUTP < OUT;
    if (p >= OFF_B3) return p - OFF_B3 < OUT;
    return true;
}

// grad[p] = sum of the slots in slot order; optionally one Adam step.
__global__ void k_mlp_reduce(const float* __restrict__ slots, int nslots, float* __restrict__ grad,
                             float* __restrict__ par, float* __restrict__ m1, float* __restrict__ m2,
                             short* __restrict__ wbf, int adam, float lr, float b1, float b2, float eps, float bc1,
                             float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPAR) return;
    float g = 0.f;
    for (int w = 0; w < nslots; w++) g += slots[(int64_t)w * NPAR + p];
    if (!trainable(p)) g = 0.f;
    grad[p] = g;
    if (!adam) return;
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    float x = par[p];
    if (trainable(p)) x -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
    par[p] = x;
    wbf[p] = f2bf(x);
}

__global__ void k_mlp_refresh(const float* __restrict__ par, short* __restrict__ wbf) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPAR) wbf[p] = f2bf(par[p]);
}

typedef __bf16 bf16x4_hw __attribute__((ext_vector_type(4)));

// Four FP32 values to BF16 with the hardware conversion (v_cvt_pk_bf16_f32:
// round to nearest even, f2bf's result for every finite value).
__device__ __forceinline__ bf16x4 to_bf16x4(f32x4 v) {
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bf16x4_hw));
}

// C += A B over 32 k with gfx950's 16 x 16 x 32 BF16 MFMA, the operands two
// 16 x 16 x 16 fragments each: A = (a0 | a1), B = (b0 | b1) per lane.  A lane
// group's element e of A and of B address the same k, so the pairs (a0, b0)
// and (a1, b1) -- k steps t and t + 1 of the K = 16 form -- are summed
// together in one instruction instead of two.
__device__ __forceinline__ f32x4 mfma_k32(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c) {
    typedef short s16x8 __attribute__((ext_vector_type(8)));
    typedef __bf16 bf16x8_hw __attribute__((ext_vector_type(8)));
    const s16x8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    const s16x8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_hw, a), __builtin_bit_cast(bf16x8_hw, b), c,
                                                   0, 0, 0);
}

__device__ __forceinline__ f32x4 load_feat(const float* __restrict__ feat, int64_t q, int64_t n, int g) {
    f32x4 f = {0.f, 0.f, 0.f, 0.f};
    if (g < 2 && q < n) f = *reinterpret_cast<const f32x4*>(feat + q * FEAT_STRIDE + 4 * g);
    return f;
}

// Inference over n samples, 16 per wavefront step, every layer computed
// transposed (Z^T = W^T A^T): then the C fragment of one layer's 16 x 16
// output tile -- lane = sample + 16 (hidden / 4), value hidden % 4 -- is
// exactly the B fragment of the next layer's K step over those 16 hidden
// units, so activations stay in registers from input to output, and the
// weights sit in registers as A fragments for the whole grid-stride loop
// (W1^T 4 tiles, W2^T 16, W3^T 4: 48 VGPRs).  Biases start the
// accumulators; the hidden layers' biases are read from LDS per group (32
// VGPRs fewer; 3 waves per SIMD without spills -- a bound of 4 spilled and
// ran slower).  The hidden layers' K = 64 products run as two K = 32 MFMAs
// each (mfma_k32).  Inputs are read as one 16-byte vector per lane
// (features 4g..4g+3 of the lane's sample, g < 2; reading them three groups
// ahead measured 5 % slower), outputs written by the lanes that hold them.
__global__ void __launch_bounds__(4 * kWave)
k_mlp_predict(const short* __restrict__ wbf, const float* __restrict__ par, const float* __restrict__ feat, int64_t n,
              int* __restrict__ branch, float* __restrict__ out) {
    __shared__ f32x4 bias[2 * H / 4];  // b1, b2 in groups of 4 units
    const int lane = threadIdx.x & 63;
    const int s = lane & 15, g = lane >> 4;
    bf16x4 a1[H / 16], a2[H / 16][H / 16], a3[H / 16];
    f32x4 c3;
#pragma unroll
    for (int t = 0; t < H / 16; t++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = 4 * g + i;
            a1[t][i] = wbf[OFF_W1 + k * H + 16 * t + s];
            a3[t][i] = wbf[OFF_W3 + (16 * t + k) * OUTP + s];
#pragma unroll
            for (int t2 = 0; t2 < H / 16; t2++) a2[t2][t][i] = wbf[OFF_W2 + (16 * t + k) * H + 16 * t2 + s];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) c3[i] = par[OFF_B3 + 4 * g + i];
    for (int j = threadIdx.x; j < 2 * H / 4; j += blockDim.x) {
        const int o = j < H / 4 ? OFF_B1 + 4 * j : OFF_B2 + 4 * (j - H / 4);
        bias[j] = f32x4{par[o], par[o + 1], par[o + 2], par[o + 3]};
    }
    __syncthreads();
    const int64_t step = (int64_t)gridDim.x * 4 * 16;
    int64_t s0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    for (; s0 < n; s0 += step) {
        const int64_t q = s0 + s;
        const bf16x4 xb = to_bf16x4(load_feat(feat, q, n, g));
        bf16x4 h1[H / 16];
#pragma unroll
        for (int t = 0; t < H / 16; t++) {
            f32x4 z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[t], xb, bias[4 * t + g], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++) z[i] = relu(z[i]);
            h1[t] = to_bf16x4(z);
        }
        bf16x4 h2[H / 16];
#pragma unroll
        for (int t2 = 0; t2 < H / 16; t2++) {
            f32x4 z = bias[H / 4 + 4 * t2 + g];
#pragma unroll
            for (int t = 0; t < H / 16; t += 2) z = mfma_k32(a2[t2][t], a2[t2][t + 1], h1[t], h1[t + 1], z);
#pragma unroll
            for (int i = 0; i < 4; i++) z[i] = relu(z[i]);
            h2[t2] = to_bf16x4(z);
        }
        f32x4 o = c3;
#pragma unroll
        for (int t = 0; t < H / 16; t += 2) o = mfma_k32(a3[t], a3[t + 1], h2[t], h2[t + 1], o);
        // lane (s, g) holds outputs 4g..4g+3 of sample s; logit 4 sits in lane s + 16
        const float o4 = __shfl(o[0], s + 16, kWave);
        if (q < n) {
            if (g == 0) {
                if (branch) {
                    int best = 0;
                    float bv = o[0];
#pragma unroll
                    for (int j = 1; j < 4; j++)
                        if (o[j] > bv) {
                            bv = o[j];
                            best = j;
                        }
                    if (o4 > bv) best = 4;
                    branch[q] = best;
                }
                if (out)
#pragma unroll
                    for (int j = 0; j < 4; j++) out[q * OUT + j] = o[j];
            } else if (g == 1 && out) {
                out[q * OUT + 4] = o[0];
                out[q * OUT + 5] = o[1];
            }
        }
    }
}

// Features and labels of every point from the exact return map at the trial
// state (strain of u, committed plastic strain and alpha); nothing is
// committed.  One wavefront per tile of a group, as k_sigma.
__global__ void __launch_bounds__(kWave)
k_point_features(GroupDev g, int tp, const double* __restrict__ strain, const int64_t* __restrict__ orig,
                 PointState st, float* __restrict__ feat, int* __restrict__ label, float* __restrict__ target) {
    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x;
    if (lane >= tp) return;
    const int64_t P = g.pt0 + tile * tp + lane;
    const int64_t o = orig[P];
    if (o < 0) return;
    const MatConst& m = tile_material(g, tile);
    const double e3[3] = {strain[3 * o], strain[3 * o + 1], strain[3 * o + 2]};
    double ep[4] = {st.ep0[P], st.ep1[P], st.ep2[P], st.ep3[P]};
    const double alpha0 = st.alpha[P];
    // trial stress and its principal values (point_sigma's first half)
    double el[6];
    el[XX] = e3[0] - ep[0];
    el[XY] = e3[1] - ep[1];
    el[XZ] = 0.0;
    el[YY] = e3[2] - ep[2];
    el[YZ] = 0.0;
    el[ZZ] = 0 - ep[3];
    double s[6];
    s[XX] = m.lam2mu * el[XX] + m.lambda * el[YY] + m.lambda * el[ZZ];
    s[YY] = m.lambda * el[XX] + m.lam2mu * el[YY] + m.lambda * el[ZZ];
    s[ZZ] = m.lambda * el[XX] + m.lambda * el[YY] + m.lam2mu * el[ZZ];
    s[XY] = m.mu * el[XY];
    s[XZ] = m.mu * el[XZ];
    s[YZ] = m.mu * el[YZ];
    double ev[3], vec[9];
    spectral(s, ev, vec);
    double sp[3];
    int mt, br;
    double hard = 0.0;
    project<false>(ev, sp, mt, br, hard, m, nullptr);
    const double c = m.phi_c / (2. * m.cosphi);
    float* f = feat + o * FEAT_STRIDE;
    f[0] = (float)(ev[0] / c);
    f[1] = (float)(ev[1] / c);
    f[2] = (float)(ev[2] / c);
    f[3] = (float)log(c);
    f[4] = (float)asin(m.sinphi);
    f[5] = (float)asin(m.sinpsi);
    f[6] = (float)(alpha0 * 1e3);
    f[7] = 0.f;
    label[o] = br;
    target[o] = (float)(hard * 1e3);
}

}  // namespace

void launch_point_features(int shape, const GroupDev& g, const double* strain, const int64_t* orig,
                           const PointState& st, float* feat, int* label, float* target, hipStream_t s) {
    if (g.ntiles == 0) return;
    hipLaunchKernelGGL(k_point_features, dim3((unsigned)g.ntiles), dim3(kWave), 0, s, g, shape_tp(shape), strain,
                       orig, st, feat, label, target);
}

namespace mlp {

void launch_grad(const short* wbf, const float* par, const float* feat, const int* label, const float* target,
                 int64_t n, int nblocks, float* slots, double* loss_slots, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(4 * kWave), 0, s, wbf, par, feat, label, target, n, slots,
                       loss_slots);
}

void launch_reduce(const float* slots, int nslots, float* grad, float* par, float* m1, float* m2, short* wbf,
                   bool adam, float lr, float b1, float b2, float eps, float bc1, float bc2, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPAR + 255) / 256), dim3(256), 0, s, slots, nslots, grad, par, m1, m2, wbf,
                       adam ? 1 : 0, lr, b1, b2, eps, bc1, bc2);
}

void launch_refresh(const float* par, short* wbf, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_refresh, dim3((NPAR + 255) / 256), dim3(256), 0, s, par, wbf);
}

void launch_predict(const short* wbf, const float* par, const float* feat, int64_t n, int nblocks, int* branch,
                    float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_predict, dim3(nblocks), dim3(4 * kWave), 0, s, wbf, par, feat, n, branch, out);
}

}  // namespace mlp

}  // namespace ipf

namespace ipf {

namespace {

constexpr int kSortBlock = IPF_SORT_BLOCK;

// counting sort of keys in [0, NK): per-block histograms
template <int NK>
__global__ void __launch_bounds__(kSortBlock) k_key_hist(const int* __restrict__ key, int64_t n, int* __restrict__ hist) {
    __shared__ int h[NK];
    if (threadIdx.x < NK) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kSortBlock + threadIdx.x;
    if (i < n) {
        const int k = key[i];
        atomicAdd(&h[(k >= 0 && k < NK) ? k : NK - 1], 1);
    }
    __syncthreads();
    if (threadIdx.x < NK) hist[(int64_t)blockIdx.x * NK + threadIdx.x] = h[threadIdx.x];
}

// exclusive offsets: key-major, then block order.  One block of 1024
// threads; thread t owns a contiguous chunk of the block histograms.
constexpr int kScanThreads = 1024;
template <int NK>
__global__ void __launch_bounds__(kScanThreads) k_key_scan(int* __restrict__ hist, int nblocks) {
    __shared__ long long part[kScanThreads][NK];
    __shared__ long long base[NK];
    const int t = threadIdx.x;
    const int chunk = (nblocks + kScanThreads - 1) / kScanThreads;
    const int b0 = min(nblocks, t * chunk), b1 = min(nblocks, b0 + chunk);
    long long sum[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) sum[k] = 0;
    for (int b = b0; b < b1; b++)
#pragma unroll
        for (int k = 0; k < NK; k++) sum[k] += hist[(int64_t)b * NK + k];
#pragma unroll
    for (int k = 0; k < NK; k++) part[t][k] = sum[k];
    __syncthreads();
    if (t < NK) {  // per key: exclusive prefix over the chunks, then the key totals
        long long run = 0;
        for (int i = 0; i < kScanThreads; i++) {
            const long long c = part[i][t];
            part[i][t] = run;
            run += c;
        }
        base[t] = run;
    }
    __syncthreads();
    if (t == 0) {
        long long acc = 0;
        for (int k = 0; k < NK; k++) {
            const long long c = base[k];
            base[k] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NK; k++) {
        long long run = base[k] + part[t][k];
        for (int b = b0; b < b1; b++) {
            const int c = hist[(int64_t)b * NK + k];
            hist[(int64_t)b * NK + k] = (int)run;
            run += c;
        }
    }
}

// stable scatter: rank of an element among the same keys before it in its
// block, from wavefront ballots and per-wave counts
template <int NK>
__global__ void __launch_bounds__(kSortBlock) k_key_scatter(const int* __restrict__ key, int64_t n,
                                                           const int* __restrict__ offs, int* __restrict__ order) {
    constexpr int NW = kSortBlock / kWave;
    __shared__ int wc[NW][NK];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t i = (int64_t)blockIdx.x * kSortBlock + threadIdx.x;
    int k = -1;
    if (i < n) {
        k = key[i];
        if (k < 0 || k >= NK) k = NK - 1;
    }
    int rank_in_wave = 0;
#pragma unroll
    for (int j = 0; j < NK; j++) {
        const unsigned long long b = __ballot(k == j);
        if (lane == 0) wc[wave][j] = __popcll(b);
        if (k == j) rank_in_wave = __popcll(b & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (i < n) {
        int before = 0;
        for (int w = 0; w < wave; w++) before += wc[w][k];
        order[offs[(int64_t)blockIdx.x * NK + k] + before + rank_in_wave] = (int)i;
    }
}

// The return map of the points in `order` (reference point indexes), one
// point per lane; state and material looked up per point.
__global__ void __launch_bounds__(256)
k_sigma_order(int64_t n, const int* __restrict__ order, const int64_t* __restrict__ new_of,
              const int* __restrict__ pmat, const MatConst* __restrict__ mats, const double* __restrict__ strain,
              double* __restrict__ wsigma, PointState st, int commit) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t o = order[t];
    const int64_t P = new_of[o];
    const MatConst& m = mats[pmat[P]];
    const double e3[3] = {strain[3 * o], strain[3 * o + 1], strain[3 * o + 2]};
    double ep[4] = {st.ep0[P], st.ep1[P], st.ep2[P], st.ep3[P]};
    double alpha = st.alpha[P];
    int mt, br;
    double ws[3];
    point_sigma<false>(e3, ep, alpha, mt, br, m, st.w[P], ws, nullptr);
    wsigma[3 * o + 0] = ws[0];
    wsigma[3 * o + 1] = ws[1];
    wsigma[3 * o + 2] = ws[2];
    if (commit) {
        st.ep0[P] = ep[0];
        st.ep1[P] = ep[1];
        st.ep2[P] = ep[2];
        st.ep3[P] = ep[3];
        st.alpha[P] = alpha;
        st.mtype[P] = (signed char)mt;
        st.branch[P] = (signed char)br;
    }
}

}  // namespace

int key_sort_blocks(int64_t n) { return (int)((n + kSortBlock - 1) / kSortBlock); }

void launch_branch_order(const int* key, int64_t n, int* hist, int* order, hipStream_t s) {
    if (n == 0) return;
    const int nb = key_sort_blocks(n);
    hipLaunchKernelGGL(k_key_hist<IPF_NBRANCH>, dim3(nb), dim3(kSortBlock), 0, s, key, n, hist);
    hipLaunchKernelGGL(k_key_scan<IPF_NBRANCH>, dim3(1), dim3(kScanThreads), 0, s, hist, nb);
    hipLaunchKernelGGL(k_key_scatter<IPF_NBRANCH>, dim3(nb), dim3(kSortBlock), 0, s, key, n, hist, order);
}

void launch_sigma_order(int64_t n, const int* order, const int64_t* new_of, const int* pmat, const MatConst* mats,
                        const double* strain, double* wsigma, const PointState& st, bool commit, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_sigma_order, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, order, new_of, pmat, mats,
                       strain, wsigma, st, commit ? 1 : 0);
}

}  // namespace ipf
