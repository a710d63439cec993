// Learned warm start for the secondary-field solve.
//
// A per-(node, source) MLP 32 -> 128 -> 128 -> 128 -> 1 with GELU (tanh form)
// predicts the relative secondary potential t = u_s / (rho_s p) (p: the unit
// half-space primary of the source); the block solver then starts from
// X0 = t_hat rho_s p instead of 0.  Weights are BF16 on the matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation), biases and the master copy
// FP32; training is Adam on mean squared error, with hand-written forward,
// activation-gradient (MFMA) and weight-gradient kernels.  The converged
// FP64 residual still decides when the solve stops, so the network only moves
// iteration counts.
#include "internal.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>

namespace pgert {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int NF = 32, NH = 128;
// packed FP32 parameter layout
constexpr int OW1 = 0, OB1 = OW1 + NF * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NH, NPARAM = OB4 + 1;
constexpr float C_GELU = 0.7978845608028654f;  // sqrt(2/pi)

__device__ __host__ inline unsigned short f2bf(float f) {
    unsigned int u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) {
    unsigned int u = ((unsigned int)h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
// tanh(y) = 1 - 2 / (e^{2y} + 1) with the hardware exponential and
// reciprocal (relative error ~1e-7 in e^{2y}; absolute error ~1e-7 in the
// result, far below the BF16 activations it feeds); exact limits +-1
__device__ inline float fast_tanh(float y) {
    const float e = __expf(2.0f * y);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ inline float gelu(float x) {
    return 0.5f * x * (1.0f + fast_tanh(C_GELU * (x + 0.044715f * x * x * x)));
}
__device__ inline float gelu_grad(float x) {
    float t = fast_tanh(C_GELU * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * C_GELU * (1.0f + 3.0f * 0.044715f * x * x);
}

// ------------------------------------------------------------------ features
struct FeatArgs {
    int n, nE, dim;
    const double * xyz;
    const double * elecPos;
    const double * prim;   // [n][nE] of this wavenumber
    const double * rhoS;   // [nE]
    const float * lr0;     // node log10 rho (1-ring mean)
    const float * s1;
    const float * s2;
    const float * s3;
    double surfaceZ;
    float L;
};

__device__ void features(const FeatArgs & a, int node, int e, float * f) {
    const double * x = a.xyz + 3 * (size_t)node;
    const double * s = a.elecPos + 3 * (size_t)e;
    const float L = a.L;
    float d[3] = {(float)(x[0] - s[0]), (float)(x[1] - s[1]), (float)(x[2] - s[2])};
    float r = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const int up = a.dim - 1;
    float dm[3] = {d[0], d[1], d[2]};
    double zs = (a.surfaceZ == -DBL_MAX) ? 0.0 : a.surfaceZ;
    dm[up] = (float)(x[up] - (2.0 * zs - s[up]));
    float rm = sqrtf(dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2]);
    float p = (float)a.prim[(size_t)node * a.nE + e];
    float lrs = log10f((float)a.rhoS[e]);
    float l0 = a.lr0[node], l1 = a.s1[node], l2 = a.s2[node], l3 = a.s3[node];
    float pn = 6.2831853f * L * p;
    float ex = expf(-r / L), rr = r / (r + L);
    float ir = 1.0f / fmaxf(r, 1e-9f);
    f[0] = d[0] / L; f[1] = d[1] / L; f[2] = d[2] / L;
    f[3] = r / L;
    f[4] = logf(r / L + 1e-3f);
    f[5] = rm / L;
    f[6] = (float)((zs - x[up]) / L);
    f[7] = pn;
    f[8] = logf(fabsf(pn) + 1e-6f);
    f[9] = lrs - 2.0f;
    f[10] = l0 - lrs; f[11] = l1 - lrs; f[12] = l2 - lrs; f[13] = l3 - lrs;
    f[14] = l0 - 2.0f; f[15] = l1 - 2.0f; f[16] = l2 - 2.0f; f[17] = l3 - 2.0f;
    f[18] = l2 - l1; f[19] = l3 - l2;
    f[20] = ex; f[21] = rr;
    f[22] = d[0] * ir; f[23] = d[1] * ir; f[24] = d[2] * ir;
    f[25] = (float)(s[0] / L); f[26] = (float)(s[1] / L); f[27] = (float)(s[2] / L);
    f[28] = f[11] * ex; f[29] = f[12] * rr; f[30] = f[13] * rr;
    f[31] = 1.0f;
}

// node log10 rho = mean over the cells containing the node: the diagonal
// slot's contributions are exactly those cells
__global__ void k_node_logrho(int n, const int * __restrict__ diagSlot, const int * __restrict__ cptr,
                              const int * __restrict__ ccell, const double * __restrict__ rho,
                              float * __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = diagSlot[i];
    double acc = 0.0;
    int cnt = cptr[s + 1] - cptr[s];
    for (int e = cptr[s]; e < cptr[s + 1]; ++e) acc += log10(rho[ccell[e]]);
    out[i] = (float)(acc / (cnt > 0 ? cnt : 1));
}

// graph smoothing: mean over the CSR row (node and its neighbours)
__global__ void k_smooth(int n, const int * __restrict__ rowptr, const int * __restrict__ colidx,
                         const float * __restrict__ in, float * __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.f;
    for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) acc += in[colidx[q]];
    out[i] = acc / (float)(rowptr[i + 1] - rowptr[i]);
}

// features + relative-secondary targets of selected samples (index = node * nE + e)
__global__ void k_collect(long count, const long * __restrict__ idx, FeatArgs a, const double * __restrict__ U,
                          float * __restrict__ feat, float * __restrict__ target) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    long s = idx[t];
    int node = (int)(s / a.nE), e = (int)(s % a.nE);
    features(a, node, e, feat + t * NF);
    if (target) {
        double p = a.prim[s] * a.rhoS[e];
        double u = U[s];
        target[t] = (float)((u - p) / p);
    }
}

// ------------------------------------------------------------------ MLP forward
// One wavefront = 16 samples.  MFMA 16x16x16 (bf16) operand layouts:
//   A (16 x 16, row-major M x K): lane l holds A[l % 16][4 (l / 16) + j], j = 0..3
//   B (16 x 16, K x N):           lane l holds B[4 (l / 16) + j][l % 16]
//   D (16 x 16, M x N):           lane l holds D[4 (l / 16) + j][l % 16]
// Weights are kept transposed (W^T[out][in], bf16) so a B fragment is 4
// consecutive bf16.  Hidden activations go through LDS (bf16) to turn the D
// layout of one layer into the A layout of the next.
struct MlpW {
    const unsigned short * w1t;  // [NH][NF]
    const unsigned short * w2t;  // [NH][NH]
    const unsigned short * w3t;  // [NH][NH]
    const unsigned short * w4;   // [NH]
    const float * prm;           // FP32 master (biases)
};

// four consecutive bf16 (8-byte aligned: every fragment offset is a multiple of 4 elements)
__device__ inline v4s load4(const unsigned short * p) { return *reinterpret_cast<const v4s *>(p); }

// Network weights staged in LDS once per workgroup (the forward kernels run
// 8 wavefronts over many 16-sample tiles): W^T rows padded by WPAD bf16 so the
// 16 rows of a B fragment fall into different LDS banks.
constexpr int WPAD = 8, KW1 = NF + WPAD, KW2 = NH + WPAD;
constexpr int MLP_WAVES = 8;
constexpr size_t MLP_W_LDS = (size_t)(NH * KW1 + 2 * NH * KW2) * 2 + 3 * NH * 4 + NH * 2;

struct MlpL {
    const unsigned short *w1t, *w2t, *w3t, *w4;
    const float *b1, *b2, *b3;
    float b4;
};

__device__ MlpL stage_weights(const MlpW & w, unsigned char * smem) {
    unsigned short * s1 = (unsigned short *)smem;
    unsigned short * s2 = s1 + NH * KW1;
    unsigned short * s3 = s2 + NH * KW2;
    float * bb = (float *)(s3 + NH * KW2);
    unsigned short * s4 = (unsigned short *)(bb + 3 * NH);
    const auto g1 = glob(w.w1t), g2 = glob(w.w2t), g3 = glob(w.w3t), g4 = glob(w.w4);
    const auto gp = glob(w.prm);
    for (int t = threadIdx.x; t < NH * NF; t += blockDim.x) s1[(t / NF) * KW1 + t % NF] = g1[t];
    for (int t = threadIdx.x; t < NH * NH; t += blockDim.x) {
        s2[(t / NH) * KW2 + t % NH] = g2[t];
        s3[(t / NH) * KW2 + t % NH] = g3[t];
    }
    for (int t = threadIdx.x; t < NH; t += blockDim.x) {
        bb[t] = gp[OB1 + t];
        bb[NH + t] = gp[OB2 + t];
        bb[2 * NH + t] = gp[OB3 + t];
        s4[t] = g4[t];
    }
    __syncthreads();
    return MlpL{s1, s2, s3, s4, bb, bb + NH, bb + 2 * NH, gp[OB4]};
}

// dense layer: act (16 x K bf16 in LDS, row stride K) x W^T (LDS, row stride KW)
// -> 16 x NH, + bias, optional GELU; writes bf16 activations to `out` (16 x NH)
// and, when zsave is given, the FP32 pre-activations
template <int K, int KW>
__device__ void dense(const unsigned short * act, const unsigned short * wt, const float * bias, bool act_gelu,
                      unsigned short * out, float * zsave, int lane) {
    const int r16 = lane % 16, q = lane / 16;
    v4f acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < K / 16; ++s) {
        v4s a = load4(act + r16 * K + 16 * s + 4 * q);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            v4s b = load4(wt + (16 * t + r16) * KW + 16 * s + 4 * q);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int col = 16 * t + r16;
        const float bc = bias[col];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * q + j;
            float z = acc[t][j] + bc;
            if (zsave) zsave[row * NH + col] = z;
            out[row * NH + col] = f2bf(act_gelu ? gelu(z) : z);
        }
    }
}

// Forward of a 16-sample tile whose bf16 inputs are in x0 (16 x NF, LDS).
// h1/h2 scratch are 16 x NH bf16 LDS buffers.  Returns y for sample lane (lanes 0..15).
__device__ float mlp_tile(const MlpL & w, unsigned short * x0, unsigned short * h1, unsigned short * h2, int lane,
                          float * z1, float * z2, float * z3, unsigned short * a1, unsigned short * a2,
                          unsigned short * a3) {
    dense<NF, KW1>(x0, w.w1t, w.b1, true, h1, z1, lane);
    __builtin_amdgcn_wave_barrier();
    dense<NH, KW2>(h1, w.w2t, w.b2, true, h2, z2, lane);
    __builtin_amdgcn_wave_barrier();
    if (a1)
        for (int t = lane; t < 16 * NH; t += 64) a1[t] = h1[t];
    dense<NH, KW2>(h2, w.w3t, w.b3, true, h1, z3, lane);
    __builtin_amdgcn_wave_barrier();
    if (a2)
        for (int t = lane; t < 16 * NH; t += 64) a2[t] = h2[t];
    if (a3)
        for (int t = lane; t < 16 * NH; t += 64) a3[t] = h1[t];
    float y = 0.f;
    if (lane < 16) {
        for (int c = 0; c < NH; ++c) y += bf2f(h1[lane * NH + c]) * bf2f(w.w4[c]);
        y += w.b4;
    }
    return y;
}

// Inference over every (node, source) sample of one wavenumber: features on
// the fly, MLP, and X0 = t_hat rho_s p written in the solver's node-major layout.
// A wavefront takes 64 samples at a time: every lane computes one sample's
// features (kept in registers), then the four 16-sample tiles go through the
// MLP, the lanes of tile t handing their features over through LDS.
__global__ void __launch_bounds__(64 * MLP_WAVES) k_mlp_warmstart(long nsamp, FeatArgs a, MlpW w,
                                                                  double * __restrict__ X0) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ __attribute__((aligned(16))) unsigned short xs[MLP_WAVES][16 * NF];
    __shared__ __attribute__((aligned(16))) unsigned short hA[MLP_WAVES][16 * NH], hB[MLP_WAVES][16 * NH];
    const MlpL wl = stage_weights(w, smem);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const long ngroup = (nsamp + 63) / 64;
    for (long grp = (long)blockIdx.x * MLP_WAVES + wave; grp < ngroup; grp += (long)gridDim.x * MLP_WAVES) {
        const long s0 = grp * 64;
        float f[NF];
        {
            const long s = s0 + lane;
            if (s < nsamp) features(a, (int)(s / a.nE), (int)(s % a.nE), f);
            else
                for (int q = 0; q < NF; ++q) f[q] = 0.f;
        }
        for (int t = 0; t < 4; ++t) {
            if (lane / 16 == t)
#pragma unroll
                for (int q = 0; q < NF; ++q) xs[wave][(lane % 16) * NF + q] = f2bf(f[q]);
            __builtin_amdgcn_wave_barrier();
            const float y = mlp_tile(wl, xs[wave], hA[wave], hB[wave], lane, nullptr, nullptr, nullptr,
                                     nullptr, nullptr, nullptr);
            if (lane < 16) {
                const long s = s0 + 16 * t + lane;
                if (s < nsamp) {
                    const int e = (int)(s % a.nE);
                    X0[s] = (double)y * a.rhoS[e] * a.prim[s];
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Forward on given features (training / tests); optionally stores the
// activations needed by the backward pass.
__global__ void __launch_bounds__(64 * MLP_WAVES) k_mlp_forward(int nsamp, const float * __restrict__ feat, MlpW w,
                                                                float * __restrict__ y, float * __restrict__ Z,
                                                                unsigned short * __restrict__ Aact) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ __attribute__((aligned(16))) unsigned short xs[MLP_WAVES][16 * NF];
    __shared__ __attribute__((aligned(16))) unsigned short hA[MLP_WAVES][16 * NH], hB[MLP_WAVES][16 * NH];
    const MlpL wl = stage_weights(w, smem);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int ntile = (nsamp + 15) / 16;
    for (int tile = blockIdx.x * MLP_WAVES + wave; tile < ntile; tile += gridDim.x * MLP_WAVES) {
        const int s0 = tile * 16;
        for (int t = lane; t < 16 * NF; t += 64) {
            int s = s0 + t / NF;
            xs[wave][t] = f2bf(s < nsamp ? feat[(size_t)s0 * NF + t] : 0.f);
        }
        __builtin_amdgcn_wave_barrier();
        // Z: [3][nsamp][NH] FP32 pre-activations; Aact: [4][nsamp][NH] bf16
        // inputs of each layer (layer 0 uses the first NF columns)
        const bool keep = Z != nullptr && s0 + 16 <= nsamp;
        float yy = mlp_tile(wl, xs[wave], hA[wave], hB[wave], lane,
                            keep ? Z + (size_t)s0 * NH : nullptr,
                            keep ? Z + ((size_t)nsamp + s0) * NH : nullptr,
                            keep ? Z + ((size_t)2 * nsamp + s0) * NH : nullptr,
                            keep ? Aact + ((size_t)nsamp + s0) * NH : nullptr,
                            keep ? Aact + ((size_t)2 * nsamp + s0) * NH : nullptr,
                            keep ? Aact + ((size_t)3 * nsamp + s0) * NH : nullptr);
        if (keep)
            for (int t = lane; t < 16 * NF; t += 64) Aact[(size_t)s0 * NH + (t / NF) * NH + t % NF] = xs[wave][t];
        if (lane < 16 && s0 + lane < nsamp) y[s0 + lane] = yy;
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------ backward
// G: [3][nsamp][NH] FP32 gradients w.r.t. the pre-activations z1, z2, z3;
// g4: [nsamp] dL/dy.  gh_l = gz_{l+1} W_{l+1}^T on the matrix cores with the
// bf16 weights in their [in][out] layout: a B fragment (k = out, n = in) is
// W[in][out .. out + 3], 4 consecutive bf16 of the row-major master layout.
__global__ void __launch_bounds__(256) k_mlp_backward(int nsamp, const float * __restrict__ y,
                                                      const float * __restrict__ target,
                                                      const float * __restrict__ Z,
                                                      const unsigned short * __restrict__ w2,  // [in][out] bf16
                                                      const unsigned short * __restrict__ w3,
                                                      const unsigned short * __restrict__ w4,
                                                      float * __restrict__ G, float * __restrict__ g4) {
    __shared__ __attribute__((aligned(16))) unsigned short gs[4][16 * NH];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int r16 = lane % 16, q = lane / 16;
    const int ntile = nsamp / 16;  // the caller passes a multiple of 16
    const float invN = 1.0f / (float)nsamp;
    for (int tile = blockIdx.x * 4 + wave; tile < ntile; tile += gridDim.x * 4) {
        const int s0 = tile * 16;
        // layer 4: g = (y - t)/N ; gz3 = g w4 * gelu'(z3)
        for (int t = lane; t < 16 * NH; t += 64) {
            int r = t / NH, c = t % NH;
            int s = s0 + r;
            float g = (y[s] - target[s]) * invN;
            float gz = g * bf2f(w4[c]) * gelu_grad(Z[((size_t)2 * nsamp + s) * NH + c]);
            G[((size_t)2 * nsamp + s) * NH + c] = gz;
            gs[wave][t] = f2bf(gz);
        }
        if (lane < 16) g4[s0 + lane] = (y[s0 + lane] - target[s0 + lane]) * invN;
        __builtin_amdgcn_wave_barrier();
        for (int layer = 1; layer >= 0; --layer) {
            const unsigned short * W = layer == 1 ? w3 : w2;  // gh_{layer+1} = gz_{layer+2} W^T
            v4f acc[8];
            for (int t = 0; t < 8; ++t) acc[t] = (v4f){0.f, 0.f, 0.f, 0.f};
            for (int s = 0; s < NH / 16; ++s) {
                v4s a = load4(gs[wave] + r16 * NH + 16 * s + 4 * q);
                for (int t = 0; t < 8; ++t) {
                    v4s b = load4(W + (size_t)(16 * t + r16) * NH + 16 * s + 4 * q);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
                }
            }
            __builtin_amdgcn_wave_barrier();
            for (int t = 0; t < 8; ++t) {
                const int col = 16 * t + r16;
                for (int j = 0; j < 4; ++j) {
                    const int row = 4 * q + j;
                    const size_t o = ((size_t)layer * nsamp + s0 + row) * NH + col;
                    float gz = acc[t][j] * gelu_grad(Z[o]);
                    G[o] = gz;
                    gs[wave][row * NH + col] = f2bf(gz);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Weight gradients dW_l[i][j] = sum_s a_l[s][i] gz_l[s][j] (FP32), split over
// KSPLIT sample ranges; part[(layer, split)][i][j]; bias grads in column NH.
constexpr int KSPLIT = 64;
__global__ void __launch_bounds__(256) k_mlp_wgrad(int nsamp, const unsigned short * __restrict__ Aact,
                                                   const float * __restrict__ G, const float * __restrict__ g4,
                                                   float * __restrict__ part, int layer0) {
    // blockIdx.x: output row block of 8 (in-features), blockIdx.y: split, blockIdx.z: layer - layer0
    const int layer = layer0 + blockIdx.z;
    const int nin = layer == 0 ? NF : NH;
    const int nout = layer == 3 ? 1 : NH;
    const int i0 = blockIdx.x * 8;
    if (i0 >= nin + 1) return;
    const int s_lo = (int)((long)nsamp * blockIdx.y / KSPLIT), s_hi = (int)((long)nsamp * (blockIdx.y + 1) / KSPLIT);
    const int j = threadIdx.x % 128, ii = threadIdx.x / 128;  // 2 rows per pass
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    if (j < nout) {
        for (int s = s_lo; s < s_hi; ++s) {
            float gz = layer == 3 ? g4[s] : G[((size_t)layer * nsamp + s) * NH + j];
            for (int r = 0; r < 4; ++r) {
                int i = i0 + ii + 2 * r;
                float av = (i < nin) ? bf2f(Aact[((size_t)layer * nsamp + s) * NH + i]) : (i == nin ? 1.0f : 0.0f);
                acc[r] += av * gz;
            }
        }
    }
    float * P = part + ((size_t)layer * KSPLIT + blockIdx.y) * (NH + 1) * NH;
    for (int r = 0; r < 4; ++r) {
        int i = i0 + ii + 2 * r;
        if (i <= nin && j < nout) P[(size_t)i * NH + j] = acc[r];
    }
}

// Weight gradients of layers 0..2 on the matrix cores (FP32 MFMA 16x16x4:
// A fragment = activations^T (in-feature x sample), B fragment = gz (sample x
// out-feature), K = samples).  blockIdx.x: split of the samples, blockIdx.y:
// layer.  Chunks of 64 samples are staged in LDS (activations widened to
// FP32, the bias column as 1.0); each wavefront owns a quarter of the 16 x 16
// output tiles.  Same partial layout as k_mlp_wgrad.
constexpr int WG_CHUNK = 64, WG_LD = 144, WG_TILES = 18;  // 9 x 8 tiles / 4 wavefronts
typedef float v4f_ __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_mlp_wgrad_mfma(int nsamp, const unsigned short * __restrict__ Aact,
                                                        const float * __restrict__ G, float * __restrict__ part) {
    __shared__ float As[WG_CHUNK * WG_LD], Gs[WG_CHUNK * WG_LD];
    const int layer = blockIdx.y;
    const int nin = layer == 0 ? NF : NH;
    const int TI = (nin + 1 + 15) / 16, ntile = TI * 8;
    const int s_lo = (int)((long)nsamp * blockIdx.x / KSPLIT), s_hi = (int)((long)nsamp * (blockIdx.x + 1) / KSPLIT);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    v4f_ acc[WG_TILES];
#pragma unroll
    for (int k = 0; k < WG_TILES; ++k) acc[k] = (v4f_){0.f, 0.f, 0.f, 0.f};
    const auto A = glob(Aact) + (size_t)layer * nsamp * NH;
    const auto Gl = glob(G) + (size_t)layer * nsamp * NH;
    for (int c0 = s_lo; c0 < s_hi; c0 += WG_CHUNK) {
        const int cn = min(WG_CHUNK, s_hi - c0);
        __syncthreads();
        for (int t = threadIdx.x; t < WG_CHUNK * NH; t += 256) {
            const int r = t / NH, col = t % NH;
            const bool in = r < cn;
            const int s = c0 + (in ? r : 0);
            const float a = bf2f(A[(size_t)s * NH + col]);
            const float g = Gl[(size_t)s * NH + col];
            As[r * WG_LD + col] = !in ? 0.f : col < nin ? a : col == nin ? 1.f : 0.f;  // column nin: bias
            Gs[r * WG_LD + col] = in ? g : 0.f;
        }
        if (nin == NH)  // the bias column lies beyond the loaded columns
            for (int r = threadIdx.x; r < WG_CHUNK; r += 256) As[r * WG_LD + NH] = r < cn ? 1.f : 0.f;
        __syncthreads();
#pragma unroll 4
        for (int ks = 0; ks < WG_CHUNK / 4; ++ks) {
            const int kr = 4 * ks + lane / 16;
#pragma unroll
            for (int k = 0; k < WG_TILES; ++k) {
                const int tile = wave * WG_TILES + k;
                if (tile < ntile) {
                    const int ti = tile / 8, tj = tile % 8;
                    const float a = As[kr * WG_LD + ti * 16 + lane % 16];
                    const float b = Gs[kr * WG_LD + tj * 16 + lane % 16];
                    acc[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[k], 0, 0, 0);
                }
            }
        }
    }
    float * P = part + ((size_t)layer * KSPLIT + blockIdx.x) * (NH + 1) * NH;
#pragma unroll
    for (int k = 0; k < WG_TILES; ++k) {
        const int tile = wave * WG_TILES + k;
        if (tile < ntile) {
            const int ti = tile / 8, tj = tile % 8;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = ti * 16 + 4 * (lane / 16) + r, j = tj * 16 + lane % 16;
                if (i <= nin) P[(size_t)i * NH + j] = acc[k][r];
            }
        }
    }
}

// sum the splits into the packed gradient vector, then Adam
__global__ void k_mlp_gather_grad(const float * __restrict__ part, float * __restrict__ grad) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NPARAM) return;
    int layer, i, j;
    if (t < OB1) { layer = 0; i = t / NH; j = t % NH; }
    else if (t < OW2) { layer = 0; i = NF; j = t - OB1; }
    else if (t < OB2) { layer = 1; i = (t - OW2) / NH; j = (t - OW2) % NH; }
    else if (t < OW3) { layer = 1; i = NH; j = t - OB2; }
    else if (t < OB3) { layer = 2; i = (t - OW3) / NH; j = (t - OW3) % NH; }
    else if (t < OW4) { layer = 2; i = NH; j = t - OB3; }
    else if (t < OB4) { layer = 3; i = t - OW4; j = 0; }
    else { layer = 3; i = NH; j = 0; }
    float s = 0.f;
    for (int k = 0; k < KSPLIT; ++k) s += part[(((size_t)layer * KSPLIT + k) * (NH + 1) + i) * NH + j];
    grad[t] = s;
}

__global__ void k_adam(const float * __restrict__ grad, float * __restrict__ prm, float * __restrict__ m,
                       float * __restrict__ v, float lr, float b1t, float b2t) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float g = grad[t];
    float mm = b1 * m[t] + (1 - b1) * g;
    float vv = b2 * v[t] + (1 - b2) * g * g;
    m[t] = mm;
    v[t] = vv;
    prm[t] -= lr * (mm / (1 - b1t)) / (sqrtf(vv / (1 - b2t)) + eps);
}

// bf16 copies: transposed for the forward, [in][out] for the backward
__global__ void k_mlp_pack(const float * __restrict__ prm, unsigned short * __restrict__ w1t,
                           unsigned short * __restrict__ w2t, unsigned short * __restrict__ w3t,
                           unsigned short * __restrict__ w2, unsigned short * __restrict__ w3,
                           unsigned short * __restrict__ w4) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < NF * NH) {
        int i = t / NH, o = t % NH;
        w1t[o * NF + i] = f2bf(prm[OW1 + t]);
    }
    if (t < NH * NH) {
        int i = t / NH, o = t % NH;
        unsigned short a = f2bf(prm[OW2 + t]), b = f2bf(prm[OW3 + t]);
        w2t[o * NH + i] = a;
        w3t[o * NH + i] = b;
        w2[t] = a;
        w3[t] = b;
    }
    if (t < NH) w4[t] = f2bf(prm[OW4 + t]);
}

}  // namespace

// ------------------------------------------------------------------ host side
struct MlpState {
    DBuf<float> prm, m, v, grad, part;
    DBuf<unsigned short> w1t, w2t, w3t, w2, w3, w4;
    DBuf<float> lr0, s1, s2, s3, tmp, tmp2;
    DBuf<float> y, Z, G, g4;
    DBuf<unsigned short> A;
    int step = 0;
    bool ready = false;
    float L = 1.0f;
};

static MlpState & state(Ctx & c) {
    if (!c.mlp) c.mlp = std::make_shared<MlpState>();
    return *c.mlp;
}

int mlp_param_count() { return NPARAM; }

static MlpW weights(MlpState & m) {
    static bool attr = false;  // the forward kernels stage the weights in dynamic LDS
    if (!attr) {
        PG_CHECK(hipFuncSetAttribute((const void *)k_mlp_warmstart, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)MLP_W_LDS));
        PG_CHECK(hipFuncSetAttribute((const void *)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)MLP_W_LDS));
        attr = true;
    }
    return MlpW{m.w1t.p, m.w2t.p, m.w3t.p, m.w4.p, m.prm.p};
}

void mlp_set_params(Ctx & c, const float * host) {
    MlpState & m = state(c);
    m.prm.upload(host, NPARAM, c.stream);
    m.m.alloc(NPARAM);
    m.v.alloc(NPARAM);
    PG_CHECK(hipMemsetAsync(m.m.p, 0, m.m.bytes(), c.stream));
    PG_CHECK(hipMemsetAsync(m.v.p, 0, m.v.bytes(), c.stream));
    m.grad.alloc(NPARAM);
    m.w1t.alloc(NF * NH);
    m.w2t.alloc(NH * NH);
    m.w3t.alloc(NH * NH);
    m.w2.alloc(NH * NH);
    m.w3.alloc(NH * NH);
    m.w4.alloc(NH);
    hipLaunchKernelGGL(k_mlp_pack, dim3(div_up(NH * NH, 256)), dim3(256), 0, c.stream, m.prm.p, m.w1t.p, m.w2t.p,
                       m.w3t.p, m.w2.p, m.w3.p, m.w4.p);
    PG_CHECK(hipGetLastError());
    m.step = 0;
    m.ready = true;
}

void mlp_get_params(Ctx & c, float * host) {
    MlpState & m = state(c);
    PG_REQUIRE(m.ready, "mlp: no parameters");
    PG_CHECK(hipMemcpyAsync(host, m.prm.p, NPARAM * 4, hipMemcpyDeviceToHost, c.stream));
    PG_CHECK(hipStreamSynchronize(c.stream));
}

// node log-rho fields of the current model (c.rho) and the feature scale
static FeatArgs feat_args(Ctx & c, int kk) {
    MlpState & m = state(c);
    const int n = c.n;
    m.lr0.alloc(n);
    m.s1.alloc(n);
    m.s2.alloc(n);
    m.s3.alloc(n);
    m.tmp.alloc(n);
    hipLaunchKernelGGL(k_node_logrho, dim3(div_up(n, 256)), dim3(256), 0, c.stream, n, c.diagSlot.p, c.cptr.p,
                       c.ccell.p, c.rho.p, m.lr0.p);
    // s1 / s2 / s3: 2 / 8 / 32 smoothing sweeps of the node field
    m.tmp2.alloc(n);
    const float * src = m.lr0.p;
    for (int it = 1; it <= 32; ++it) {
        float * dst = it == 2 ? m.s1.p : it == 8 ? m.s2.p : it == 32 ? m.s3.p : (src == m.tmp.p ? m.tmp2.p : m.tmp.p);
        hipLaunchKernelGGL(k_smooth, dim3(div_up(n, 256)), dim3(256), 0, c.stream, n, c.rowptr.p, c.colidx.p, src,
                           dst);
        src = dst;
    }
    PG_CHECK(hipGetLastError());
    FeatArgs a;
    a.n = n;
    a.nE = c.nE;
    a.dim = c.dim;
    a.xyz = c.xyz.p;
    a.elecPos = c.elecPos.p;
    a.prim = c.prim.p + (size_t)kk * n * c.nE;
    a.rhoS = c.rhoS.p;
    a.lr0 = m.lr0.p;
    a.s1 = m.s1.p;
    a.s2 = m.s2.p;
    a.s3 = m.s3.p;
    a.surfaceZ = c.surfaceZ;
    a.L = c.featureScale;
    return a;
}

void mlp_warm_start(Ctx & c) {
    MlpState & m = state(c);
    PG_REQUIRE(m.ready, "learned warm start requested but no network parameters are set");
    for (int kk = 0; kk < c.nK; ++kk) {
        FeatArgs a = feat_args(c, kk);
        long nsamp = (long)c.n * c.nE;
        hipLaunchKernelGGL(k_mlp_warmstart, dim3(std::max(1L, std::min(512L, (nsamp + 511) / 512))),
                           dim3(64 * MLP_WAVES), MLP_W_LDS, c.stream, nsamp, a, weights(m),
                           c.X.p + (size_t)kk * c.n * c.nE);
        PG_CHECK(hipGetLastError());
    }
}

void mlp_collect(Ctx & c, int kk, const long * idx, long count, float * feat, float * target) {
    PG_REQUIRE(c.haveU, "mlp collect: no forward result");
    FeatArgs a = feat_args(c, kk);
    hipLaunchKernelGGL(k_collect, dim3(div_up(count, 256)), dim3(256), 0, c.stream, count, idx, a,
                       c.U.p + (size_t)kk * c.n * c.nE, feat, target);
    PG_CHECK(hipGetLastError());
}

void mlp_predict(Ctx & c, const float * feat, int nsamp, float * out) {
    MlpState & m = state(c);
    PG_REQUIRE(m.ready, "mlp: no parameters");
    hipLaunchKernelGGL(k_mlp_forward, dim3(std::max(1, std::min(512, div_up(nsamp, 16 * MLP_WAVES)))),
                       dim3(64 * MLP_WAVES), MLP_W_LDS, c.stream, nsamp, feat, weights(m), out, (float *)nullptr,
                       (unsigned short *)nullptr);
    PG_CHECK(hipGetLastError());
}

// loss and gradient of a batch (nsamp multiple of 16); apply = 1 takes an Adam step
float mlp_train_step(Ctx & c, const float * feat, const float * target, int nsamp, float lr, bool apply,
                     float * grad_host) {
    MlpState & m = state(c);
    PG_REQUIRE(m.ready, "mlp: no parameters");
    PG_REQUIRE(nsamp > 0 && nsamp % 16 == 0, "mlp: batch must be a positive multiple of 16");
    m.y.alloc(nsamp);
    m.Z.alloc((size_t)3 * nsamp * NH);
    m.G.alloc((size_t)3 * nsamp * NH);
    m.g4.alloc(nsamp);
    m.A.alloc((size_t)4 * nsamp * NH);
    m.part.alloc((size_t)4 * KSPLIT * (NH + 1) * NH);
    PG_CHECK(hipMemsetAsync(m.part.p, 0, m.part.bytes(), c.stream));
    const int grid = std::max(1, std::min(1024, div_up(nsamp, 64)));
    hipLaunchKernelGGL(k_mlp_forward, dim3(std::max(1, std::min(512, div_up(nsamp, 16 * MLP_WAVES)))),
                       dim3(64 * MLP_WAVES), MLP_W_LDS, c.stream, nsamp, feat, weights(m), m.y.p, m.Z.p, m.A.p);
    hipLaunchKernelGGL(k_mlp_backward, dim3(grid), dim3(256), 0, c.stream, nsamp, m.y.p, target, m.Z.p, m.w2.p,
                       m.w3.p, m.w4.p, m.G.p, m.g4.p);
    hipLaunchKernelGGL(k_mlp_wgrad_mfma, dim3(KSPLIT, 3), dim3(256), 0, c.stream, nsamp, m.A.p, m.G.p, m.part.p);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((NH + 1 + 7) / 8, KSPLIT, 1), dim3(256), 0, c.stream, nsamp, m.A.p, m.G.p,
                       m.g4.p, m.part.p, 3);
    hipLaunchKernelGGL(k_mlp_gather_grad, dim3(div_up(NPARAM, 256)), dim3(256), 0, c.stream, m.part.p, m.grad.p);
    PG_CHECK(hipGetLastError());
    if (apply) {
        ++m.step;
        float b1t = powf(0.9f, (float)m.step), b2t = powf(0.999f, (float)m.step);
        hipLaunchKernelGGL(k_adam, dim3(div_up(NPARAM, 256)), dim3(256), 0, c.stream, m.grad.p, m.prm.p, m.m.p,
                           m.v.p, lr, b1t, b2t);
        hipLaunchKernelGGL(k_mlp_pack, dim3(div_up(NH * NH, 256)), dim3(256), 0, c.stream, m.prm.p, m.w1t.p,
                           m.w2t.p, m.w3t.p, m.w2.p, m.w3.p, m.w4.p);
        PG_CHECK(hipGetLastError());
    }
    if (grad_host) PG_CHECK(hipMemcpyAsync(grad_host, m.grad.p, NPARAM * 4, hipMemcpyDeviceToHost, c.stream));
    // loss = 0.5 mean (y - t)^2, from g4 = (y - t)/N
    std::vector<float> g((size_t)nsamp);
    PG_CHECK(hipMemcpyAsync(g.data(), m.g4.p, (size_t)nsamp * 4, hipMemcpyDeviceToHost, c.stream));
    PG_CHECK(hipStreamSynchronize(c.stream));
    double loss = 0.0;
    for (float v : g) loss += (double)v * v;
    return (float)(0.5 * loss * nsamp);
}

}  // namespace pgert
