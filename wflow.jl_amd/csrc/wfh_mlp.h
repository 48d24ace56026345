// wfh_mlp.h - the learned Newton warm start: an 8-64-64-1 perceptron with BF16 weights and
// activations and FP32 accumulation, evaluated and trained with gfx950 matrix-core instructions
// (v_mfma_f32_32x32x16_bf16), one wave per 64 samples. Included by wflow_hip.hip.
//
// Network (unit indices 0..63 of both hidden layers are the matrix-core row indices below):
//   z1 = W1 x + b1,  h1 = bf16(relu(z1))          W1: [64][8]  (x: 8 normalised features, bf16)
//   z2 = W2 h1 + b2, h2 = relu(z2)                 W2: [64][64]
//   y  = w3 . h2 + b3                              w3: [64] (FP32)
// Training minimises the mean squared error with Adam on FP32 master weights; the BF16 operand
// copies are re-derived after every update. The backward pass rounds the output gradient of
// layer 2 and of layer 1 to BF16 where they enter a matrix-core product.
//
// Matrix-core lane layouts (checked on the device by tests/test_gpu_mlp.py):
//   A (32x16): lane l holds row l % 32, k = 8 (l / 32) + j, j = 0..7
//   B (16x32): lane l holds col l % 32, k = 8 (l / 32) + j
//   D (32x32): lane l, register r holds row 8 (r / 4) + 4 (l / 32) + r % 4, col l % 32
// Samples are columns: lane l owns sample l; a column block c holds samples 32 c .. 32 c + 31.
// A layer's D tile can be the next layer's B operand without moving data between lanes when the
// next layer's k positions are the D rows in register order: k position 16 kb + 8 h + j of
// in-block kb (0, 1) of a 32-unit block is unit 8 (2 kb + j / 4) + 4 h + j % 4 of that block.

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MLP_IN = 8, MLP_H = 64;
// flat parameter vector: W1 [64][8], b1 [64], W2 [64][64], b2 [64], w3 [64], b3
constexpr int MLP_OW1 = 0, MLP_OB1 = MLP_OW1 + MLP_H * MLP_IN, MLP_OW2 = MLP_OB1 + MLP_H,
              MLP_OB2 = MLP_OW2 + MLP_H * MLP_H, MLP_OW3 = MLP_OB2 + MLP_H, MLP_OB3 = MLP_OW3 + MLP_H,
              MLP_NPARAM = MLP_OB3 + 1;

// in-unit of k position (kb2 = 0..3 over 64 units, lane half h, j = 0..7)
__host__ __device__ inline int mlp_kunit(int kb2, int h, int j) {
    return 32 * (kb2 >> 1) + 8 * (2 * (kb2 & 1) + (j >> 2)) + 4 * h + (j & 3);
}
// unit held by lane half h, register r of a D tile of row block rb
__host__ __device__ inline int mlp_dunit(int rb, int h, int r) { return 32 * rb + 8 * (r >> 2) + 4 * h + (r & 3); }

// Operand copies of the weights (device), in the order one lane reads them.
struct MlpOps {
    bf16x8 a1[2][64];        // layer 1, row block hb: lane < 32 W1[32 hb + lane][0..7], lane >= 32 zero
    bf16x8 a2[2][4][64];     // layer 2, row block ob, k block kb2: W2[32 ob + l % 32][kunit(kb2, l / 32, j)]
    bf16x8 a2t[2][4][64];    // W2 transposed (backward): W2[kunit(kb2, l / 32, j)][32 ub + l % 32]
    float b1[2][2][16];      // [row block][lane half][register]: b1 of the unit in that D slot
    float b2[2][2][16];
    float w3[2][2][16];
    float b3;
    float x_mean[MLP_IN], x_inv_std[MLP_IN];  // feature normalisation
    float y_mean, y_std;                      // target scaling
};

__device__ __forceinline__ bf16x8 bf8_zero() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
    return z;
}

__device__ __forceinline__ bf16x8 shfl_bf8(const bf16x8 v, const int src) {
    union {
        bf16x8 b;
        int w[4];
    } u;
    u.b = v;
#pragma unroll
    for (int k = 0; k < 4; ++k) u.w[k] = __shfl(u.w[k], src);
    return u.b;
}

// Forward pass of one wave: lane l's normalised features x (8 floats) -> network output
// (normalised target space). With T != nullptr the intermediate tiles are kept for the backward
// pass: z1/z2 pre-activations (with bias) per column block.
struct MlpTape {
    f32x16 z1[2][2];  // [column block][row block]
    f32x16 z2[2][2];
    bf16x8 xb;        // own features (bf16)
};

__device__ __forceinline__ float mlp_forward(const MlpOps& W, const float* x, MlpTape* T) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    bf16x8 xb;
#pragma unroll
    for (int j = 0; j < 8; ++j) xb[j] = (__bf16)((x[j] - W.x_mean[j]) * W.x_inv_std[j]);
    if (T) T->xb = xb;
    const bf16x8 xhi = shfl_bf8(xb, (lane & 31) + 32);
    float part[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const bf16x8 b1op = h ? bf8_zero() : (c == 0 ? xb : xhi);
        bf16x8 b2op[4];
#pragma unroll
        for (int hb = 0; hb < 2; ++hb) {
            f32x16 z = {};
            z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a1[hb][lane], b1op, z, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) z[r] += W.b1[hb][h][r];
            if (T) T->z1[c][hb] = z;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int j = 0; j < 8; ++j) b2op[2 * hb + kb][j] = (__bf16)fmaxf(z[8 * kb + j], 0.0f);
        }
        float acc = 0.0f;
#pragma unroll
        for (int ob = 0; ob < 2; ++ob) {
            f32x16 z = {};
#pragma unroll
            for (int kb2 = 0; kb2 < 4; ++kb2)
                z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a2[ob][kb2][lane], b2op[kb2], z, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                z[r] += W.b2[ob][h][r];
                acc += fmaxf(z[r], 0.0f) * W.w3[ob][h][r];
            }
            if (T) T->z2[c][ob] = z;
        }
        part[c] = acc + __shfl_xor(acc, 32);
    }
    return (h ? part[1] : part[0]) + W.b3;
}

// The forward pass as a call of its own, with the weights read from global memory (L2): for
// callers that consult the network rarely, so that their own register allocation and LDS are
// not shaped by it.
__device__ __noinline__ float mlp_forward_call(const MlpOps* W, float x0, float x1, float x2, float x3, float x4,
                                               float x5, float x6, float x7) {
    const float x[8] = {x0, x1, x2, x3, x4, x5, x6, x7};
    return mlp_forward(*W, x, nullptr);
}

// ---- training ----------------------------------------------------------------------------
// One wave per block: `tiles` tiles of 64 samples from x [n][8], t [n] (normalised targets),
// starting at tile blockIdx.x * tiles; partial gradients of the mean squared error over n_total
// samples into part[blockIdx.x][MLP_NPARAM]; loss sum into loss[blockIdx.x].
__global__ void __launch_bounds__(64) k_mlp_grad(const MlpOps* __restrict__ Wp, const float* __restrict__ x,
                                                 const float* __restrict__ t, long long n, int tiles, float inv_n,
                                                 float* __restrict__ part, float* __restrict__ loss) {
    __shared__ __attribute__((aligned(16))) __bf16 G[MLP_H][64];  // [unit][sample]: bf16 layer output gradients
    __shared__ __attribute__((aligned(16))) __bf16 H[MLP_H][64];  // [unit][sample]: h1
    __shared__ __attribute__((aligned(16))) __bf16 XT[8][64];     // [feature][sample]
    __shared__ float U[3][MLP_H];  // per-unit sums of the block: db1, db2, dw3
    const MlpOps& W = *Wp;
    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    f32x16 gw2[2][2], gw1[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        gw1[a] = f32x16{};
#pragma unroll
        for (int b = 0; b < 2; ++b) gw2[a][b] = f32x16{};
    }
    for (int k = lane; k < 3 * MLP_H; k += 64) (&U[0][0])[k] = 0.0f;
    float gb3 = 0.0f, lsum = 0.0f;
    __syncthreads();
    for (int tt = 0; tt < tiles; ++tt) {
        const long long s = ((long long)blockIdx.x * tiles + tt) * 64 + lane;
        const bool valid = s < n;
        float xv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) xv[j] = valid ? x[s * 8 + j] : W.x_mean[j];
        const float tv = valid ? t[s] : 0.0f;
        MlpTape T;
        const float y = mlp_forward(W, xv, &T);
        const float err = valid ? y - tv : 0.0f;
        lsum += err * err;
        const float dy = 2.0f * err * inv_n;
        gb3 += dy;
#pragma unroll
        for (int j = 0; j < 8; ++j) XT[j][lane] = T.xb[j];
        // h1 tile into LDS [unit][sample]
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int hb = 0; hb < 2; ++hb)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    H[mlp_dunit(hb, h, r)][32 * c + col] = (__bf16)fmaxf(T.z1[c][hb][r], 0.0f);
        // layer 3 and layer 2 output gradients (g2 replaces z2 in the tape)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const float dyc = __shfl(dy, 32 * c + col);
#pragma unroll
            for (int ob = 0; ob < 2; ++ob)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float z = T.z2[c][ob][r];
                    const int u = mlp_dunit(ob, h, r);
                    atomicAdd(&U[2][u], fmaxf(z, 0.0f) * dyc);
                    const float g = z > 0.0f ? W.w3[ob][h][r] * dyc : 0.0f;
                    T.z2[c][ob][r] = g;
                    atomicAdd(&U[1][u], g);
                    G[u][32 * c + col] = (__bf16)g;
                }
        }
        __syncthreads();
        // dW2 += G h1^T over the tile's samples (k = samples)
#pragma unroll
        for (int vb = 0; vb < 2; ++vb)
#pragma unroll
            for (int ub = 0; ub < 2; ++ub)
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&G[32 * vb + col][16 * kb + 8 * h]);
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&H[32 * ub + col][16 * kb + 8 * h]);
                    gw2[vb][ub] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, gw2[vb][ub], 0, 0, 0);
                }
        __syncthreads();
        // dh1 = W2^T g2 (B from the g2 tiles in register order), dz1 = dh1 (z1 > 0)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            bf16x8 bg[4];
#pragma unroll
            for (int kb2 = 0; kb2 < 4; ++kb2)
#pragma unroll
                for (int j = 0; j < 8; ++j) bg[kb2][j] = (__bf16)T.z2[c][kb2 >> 1][8 * (kb2 & 1) + j];
#pragma unroll
            for (int ub = 0; ub < 2; ++ub) {
                f32x16 d = {};
#pragma unroll
                for (int kb2 = 0; kb2 < 4; ++kb2)
                    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a2t[ub][kb2][lane], bg[kb2], d, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float g = T.z1[c][ub][r] > 0.0f ? d[r] : 0.0f;
                    const int u = mlp_dunit(ub, h, r);
                    atomicAdd(&U[0][u], g);
                    G[u][32 * c + col] = (__bf16)g;
                }
            }
        }
        __syncthreads();
        // dW1 += dz1 x^T (B columns beyond the 8 features are zero)
#pragma unroll
        for (int ub = 0; ub < 2; ++ub)
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&G[32 * ub + col][16 * kb + 8 * h]);
                const bf16x8 b = col < 8 ? *reinterpret_cast<const bf16x8*>(&XT[col][16 * kb + 8 * h]) : bf8_zero();
                gw1[ub] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, gw1[ub], 0, 0, 0);
            }
        __syncthreads();
    }
#pragma unroll
    for (int mask = 1; mask < 64; mask <<= 1) {
        gb3 += __shfl_xor(gb3, mask);
        lsum += __shfl_xor(lsum, mask);
    }
    float* P = part + (size_t)blockIdx.x * MLP_NPARAM;
    P[MLP_OB1 + lane] = U[0][lane];
    P[MLP_OB2 + lane] = U[1][lane];
    P[MLP_OW3 + lane] = U[2][lane];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int u = mlp_dunit(a, h, r);
            if (col < 8) P[MLP_OW1 + u * MLP_IN + col] = gw1[a][r];
#pragma unroll
            for (int ub = 0; ub < 2; ++ub) P[MLP_OW2 + u * MLP_H + 32 * ub + col] = gw2[a][ub][r];
        }
    if (lane == 0) {
        P[MLP_OB3] = gb3;
        loss[blockIdx.x] = lsum;
    }
}

// Sum of the per-block partial gradients, then one Adam step on the FP32 master parameters
// (adam_t = the step number, 1-based; lr = 0 leaves the parameters and moments unchanged and only
// sums the gradient into grad).
__global__ void k_mlp_adam(const float* __restrict__ part, int nblocks, float* __restrict__ p, float* __restrict__ m1,
                           float* __restrict__ m2, float* __restrict__ grad, float lr, int adam_t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    float g = 0.0f;
    for (int b = 0; b < nblocks; ++b) g += part[(size_t)b * MLP_NPARAM + i];
    grad[i] = g;
    if (lr == 0.0f) return;
    const float be1 = 0.9f, be2 = 0.999f;
    m1[i] = be1 * m1[i] + (1.0f - be1) * g;
    m2[i] = be2 * m2[i] + (1.0f - be2) * g * g;
    const float mh = m1[i] / (1.0f - powf(be1, (float)adam_t));
    const float vh = m2[i] / (1.0f - powf(be2, (float)adam_t));
    p[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Operand copies from the master parameters (one block of 256 threads).
__global__ void k_mlp_layout(const float* __restrict__ p, MlpOps* __restrict__ W) {
    for (int idx = threadIdx.x; idx < 2 * 64; idx += blockDim.x) {  // a1
        const int hb = idx / 64, l = idx % 64;
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            v[j] = (__bf16)(l < 32 ? p[MLP_OW1 + (32 * hb + l) * MLP_IN + j] : 0.0f);
        W->a1[hb][l] = v;
    }
    for (int idx = threadIdx.x; idx < 2 * 4 * 64; idx += blockDim.x) {  // a2, a2t
        const int ob = idx / 256, kb2 = (idx / 64) % 4, l = idx % 64;
        bf16x8 v, vt;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int u = mlp_kunit(kb2, l >> 5, j);
            v[j] = (__bf16)p[MLP_OW2 + (32 * ob + (l & 31)) * MLP_H + u];
            vt[j] = (__bf16)p[MLP_OW2 + u * MLP_H + 32 * ob + (l & 31)];
        }
        W->a2[ob][kb2][l] = v;
        W->a2t[ob][kb2][l] = vt;
    }
    for (int idx = threadIdx.x; idx < 2 * 2 * 16; idx += blockDim.x) {
        const int rb = idx / 32, hh = (idx / 16) % 2, r = idx % 16;
        const int u = mlp_dunit(rb, hh, r);
        W->b1[rb][hh][r] = p[MLP_OB1 + u];
        W->b2[rb][hh][r] = p[MLP_OB2 + u];
        W->w3[rb][hh][r] = p[MLP_OW3 + u];
    }
    if (threadIdx.x == 0) W->b3 = p[MLP_OB3];
}

// Network outputs (normalised target space) for n samples x [n][8].
__global__ void __launch_bounds__(256) k_mlp_predict(const MlpOps* __restrict__ Wp, const float* __restrict__ x,
                                                     long long n, float* __restrict__ y) {
    __shared__ MlpOps W;
    {
        const int* src = reinterpret_cast<const int*>(Wp);
        int* dst = reinterpret_cast<int*>(&W);
        for (int k = threadIdx.x; k < (int)(sizeof(MlpOps) / 4); k += blockDim.x) dst[k] = src[k];
    }
    __syncthreads();
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float xv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) xv[j] = s < n ? x[s * 8 + j] : W.x_mean[j];
    const float out = mlp_forward(W, xv, nullptr);
    if (s < n) y[s] = out;
}

}  // namespace

struct wfh_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *p = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;  // master parameters, Adam moments
    MlpOps* ops = nullptr;
    MlpOps host_ops{};
    float* part = nullptr;
    float* loss = nullptr;
    int part_blocks = 0;
    int adam_t = 0;
    std::string err;
};

namespace {
int32_t mfail(wfh_mlp* x, int32_t code, const std::string& msg) {
    if (x) x->err = msg;
    return code;
}
#define MHIP(x, call)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) return mfail((x), WFH_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

int32_t mlp_relayout(wfh_mlp* x) {
    // normalisation lives in the host copy; the kernel fills the operand part
    MHIP(x, hipMemcpyAsync(x->ops, &x->host_ops, sizeof(MlpOps), hipMemcpyHostToDevice, x->stream));
    hipLaunchKernelGGL(k_mlp_layout, dim3(1), dim3(256), 0, x->stream, x->p, x->ops);
    MHIP(x, hipGetLastError());
    return WFH_OK;
}

int32_t mlp_ensure_part(wfh_mlp* x, int blocks) {
    if (blocks <= x->part_blocks) return WFH_OK;
    if (x->part) hipFree(x->part);
    if (x->loss) hipFree(x->loss);
    x->part = nullptr;
    x->loss = nullptr;
    MHIP(x, hipMalloc(&x->part, (size_t)blocks * MLP_NPARAM * sizeof(float)));
    MHIP(x, hipMalloc(&x->loss, (size_t)blocks * sizeof(float)));
    x->part_blocks = blocks;
    return WFH_OK;
}
}  // namespace

extern "C" int32_t wfh_mlp_create(int32_t device, uint64_t seed, wfh_mlp** out) {
    if (!out) return WFH_ERR_ARG;
    *out = nullptr;
    wfh_mlp* x = new wfh_mlp();
    x->device = device;
    auto bail = [&](int32_t code, const std::string& m) {
        g_create_error = m;
        wfh_mlp_destroy(x);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&x->stream) != hipSuccess)
        return bail(WFH_ERR_HIP, "mlp: device / stream");
    if (hipMalloc(&x->p, MLP_NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&x->m1, MLP_NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&x->m2, MLP_NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&x->grad, MLP_NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&x->ops, sizeof(MlpOps)) != hipSuccess)
        return bail(WFH_ERR_HIP, "mlp: hipMalloc");
    // He-uniform initialisation from a 64-bit LCG (reproducible for a seed)
    std::vector<float> p(MLP_NPARAM, 0.0f), z(MLP_NPARAM, 0.0f);
    uint64_t st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    for (int i = 0; i < MLP_H * MLP_IN; ++i) p[MLP_OW1 + i] = (float)(uni() * std::sqrt(6.0 / MLP_IN));
    for (int i = 0; i < MLP_H * MLP_H; ++i) p[MLP_OW2 + i] = (float)(uni() * std::sqrt(6.0 / MLP_H));
    for (int i = 0; i < MLP_H; ++i) p[MLP_OW3 + i] = (float)(uni() * std::sqrt(3.0 / MLP_H));
    if (hipMemcpy(x->p, p.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(x->m1, z.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(x->m2, z.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bail(WFH_ERR_HIP, "mlp: upload");
    for (int j = 0; j < MLP_IN; ++j) {
        x->host_ops.x_mean[j] = 0.0f;
        x->host_ops.x_inv_std[j] = 1.0f;
    }
    x->host_ops.y_mean = 0.0f;
    x->host_ops.y_std = 1.0f;
    if (mlp_relayout(x) || hipStreamSynchronize(x->stream) != hipSuccess) return bail(WFH_ERR_HIP, "mlp: layout");
    *out = x;
    return WFH_OK;
}

extern "C" int32_t wfh_mlp_set_normalization(wfh_mlp* x, const float* x_mean, const float* x_std, float y_mean,
                                             float y_std) {
    if (!x || !x_mean || !x_std) return WFH_ERR_ARG;
    for (int j = 0; j < MLP_IN; ++j) {
        if (!(x_std[j] > 0.0f) || !std::isfinite(x_mean[j])) return mfail(x, WFH_ERR_ARG, "bad normalisation");
        x->host_ops.x_mean[j] = x_mean[j];
        x->host_ops.x_inv_std[j] = 1.0f / x_std[j];
    }
    if (!(y_std > 0.0f) || !std::isfinite(y_mean)) return mfail(x, WFH_ERR_ARG, "bad target scaling");
    x->host_ops.y_mean = y_mean;
    x->host_ops.y_std = y_std;
    MHIP(x, hipSetDevice(x->device));
    int32_t rc = mlp_relayout(x);
    if (rc) return rc;
    MHIP(x, hipStreamSynchronize(x->stream));
    return WFH_OK;
}

// params: [MLP_NPARAM] host floats in the flat order W1, b1, W2, b2, w3, b3
extern "C" int32_t wfh_mlp_set_params(wfh_mlp* x, const float* params) {
    if (!x || !params) return WFH_ERR_ARG;
    MHIP(x, hipSetDevice(x->device));
    MHIP(x, hipMemcpy(x->p, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    int32_t rc = mlp_relayout(x);
    if (rc) return rc;
    MHIP(x, hipStreamSynchronize(x->stream));
    return WFH_OK;
}

extern "C" int32_t wfh_mlp_get_params(wfh_mlp* x, float* params) {
    if (!x || !params) return WFH_ERR_ARG;
    MHIP(x, hipSetDevice(x->device));
    MHIP(x, hipStreamSynchronize(x->stream));
    MHIP(x, hipMemcpy(params, x->p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return WFH_OK;
}

// One pass over n samples (device arrays x [n][8], t [n] normalised targets) in minibatches of
// `batch` samples (a multiple of 64), one Adam step each (lr > 0), or with lr = 0 one gradient
// of the whole set (no update). Returns the mean squared error of the pass in *loss and the
// last minibatch's gradient (the whole set's with lr = 0) in grad (host, [MLP_NPARAM]) if given.
extern "C" int32_t wfh_mlp_train(wfh_mlp* x, const float* xs, const float* ts, int64_t n, int64_t batch, float lr,
                                 double* loss, float* grad) {
    if (!x || !xs || !ts || n <= 0) return WFH_ERR_ARG;
    if (lr == 0.0f) batch = ((n + 63) / 64) * 64;
    if (batch < 64 || batch % 64) return mfail(x, WFH_ERR_ARG, "batch must be a positive multiple of 64");
    MHIP(x, hipSetDevice(x->device));
    const int tiles_per_block = 4;
    double lsum = 0.0;
    std::vector<float> lbuf;
    for (int64_t s0 = 0; s0 < n; s0 += batch) {
        const int64_t nb = std::min<int64_t>(batch, n - s0);
        const int64_t ntiles = (nb + 63) / 64;
        const int blocks = (int)((ntiles + tiles_per_block - 1) / tiles_per_block);
        int32_t rc = mlp_ensure_part(x, blocks);
        if (rc) return rc;
        hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64), 0, x->stream, (const MlpOps*)x->ops, xs + s0 * 8,
                           ts + s0, (long long)nb, tiles_per_block, 1.0f / (float)nb, x->part, x->loss);
        if (lr != 0.0f) ++x->adam_t;
        hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, x->stream, (const float*)x->part,
                           blocks, x->p, x->m1, x->m2, x->grad, lr, x->adam_t);
        if (lr != 0.0f) hipLaunchKernelGGL(k_mlp_layout, dim3(1), dim3(256), 0, x->stream, (const float*)x->p, x->ops);
        MHIP(x, hipGetLastError());
        lbuf.resize(blocks);
        MHIP(x, hipMemcpyAsync(lbuf.data(), x->loss, blocks * sizeof(float), hipMemcpyDeviceToHost, x->stream));
        MHIP(x, hipStreamSynchronize(x->stream));
        for (float v : lbuf) lsum += v;
    }
    if (loss) *loss = lsum / (double)n;
    if (grad) MHIP(x, hipMemcpy(grad, x->grad, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return WFH_OK;
}

// Network outputs (normalised target space) for device arrays x [n][8] -> y [n].
extern "C" int32_t wfh_mlp_predict(wfh_mlp* x, const float* xs, int64_t n, float* ys) {
    if (!x || !xs || !ys || n < 0) return WFH_ERR_ARG;
    MHIP(x, hipSetDevice(x->device));
    if (n > 0)
        hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, x->stream,
                           (const MlpOps*)x->ops, xs, (long long)n, ys);
    MHIP(x, hipGetLastError());
    MHIP(x, hipStreamSynchronize(x->stream));
    return WFH_OK;
}

extern "C" const char* wfh_mlp_last_error(const wfh_mlp* x) { return x ? x->err.c_str() : g_create_error.c_str(); }

extern "C" void wfh_mlp_destroy(wfh_mlp* x) {
    if (!x) return;
    if (x->stream) hipStreamSynchronize(x->stream);
    for (void* q : {(void*)x->p, (void*)x->m1, (void*)x->m2, (void*)x->grad, (void*)x->ops, (void*)x->part,
                    (void*)x->loss})
        if (q) hipFree(q);
    if (x->stream) hipStreamDestroy(x->stream);
    delete x;
}
