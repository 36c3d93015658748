// mlp.hip — the learned light proposal on gfx950 (no reference counterpart).
//
// A 128 -> 128 -> 128 -> 128 -> K ReLU MLP maps per-pixel features (position, normal and view
// direction with 6 sin/cos octaves, 114 values) to a softmax over K light clusters. The proposal
// mixes it with the reference's source distribution (alias table over light power):
//   q(l | x) = (1 - mix) * pi_x(c(l)) * p_within(l) + mix * p_src(l),   mix = 0.1,
// so q > 0 wherever p_src > 0 and RIS stays unbiased. Weights are bf16 (fp32 master copy, Adam);
// every product runs on v_mfma_f32_32x32x16_bf16 with fp32 accumulation.
//
// Kernels:
//   k_propose      fused inference per 2x2 pixel quad: features -> 4 MFMA layers -> softmax ->
//                  M candidates (l, q) for each of its pixels (K <= 256)
//   k_propose_wide the same for K up to 4096: the output layer streamed in 128-cluster blocks
//   k_features     features of listed pixels (bf16), for training and tests
//   k_gemm         strided C = A*B with bias / ReLU / ReLU-gate epilogues (training fwd/bwd),
//                  or split-K fp32 partials for the weight gradients
//   k_softmax_ce   mean cross-entropy of the logits against target clusters, and dlogits
//   k_colsum_partial, k_sum_partials   bias gradients and the split-K reduction, in fixed order
//   k_adam         Adam step on fp32 master weights, refreshes the bf16 copy
//   k_mlp_init     seeded He-uniform initialisation
#include "../../include/restir_abi.h"
#include "carve.h"
#include "kernels.h"
#include "mlp.h"

namespace rm {

using rt::f3;

// ------------------------------------------------------------------------------ features ---
// Writes the 128 bf16 features of one surface point. The scalars are the position normalised
// to the scene box, the normal and the unit view direction; raw position and normal first, then
// sin/cos(2^k * pi * s) for k = 0..5 (double-angle recurrences from k = 0), zero padding last.
__device__ __forceinline__ void make_features(f3 pos, f3 n, f3 cam, const float *lo, const float *hi,
                                              uint16_t *dst) {
    float s[9];
    const float pp[3] = {pos.x, pos.y, pos.z};
    for (int k = 0; k < 3; ++k) {
        const float c = 0.5f * (lo[k] + hi[k]), h = 0.5f * (hi[k] - lo[k]);
        s[k] = (pp[k] - c) / h;
    }
    s[3] = n.x;
    s[4] = n.y;
    s[5] = n.z;
    f3 v = cam - pos;
    const float vl = sqrtf(rt::dot(v, v));
    v = vl > 0.0f ? v / vl : f3{0.0f, 0.0f, 0.0f};
    s[6] = v.x;
    s[7] = v.y;
    s[8] = v.z;
    // dst is 4-byte aligned and every pair below starts at an even index: one packed
    // conversion and one 32-bit store per pair
    uint32_t *d2 = reinterpret_cast<uint32_t *>(dst);
    for (int k = 0; k < 3; ++k) d2[k] = f2bf2(s[2 * k], s[2 * k + 1]);
    for (int i = 0; i < 9; ++i) {
        float sn, cs;
        sincosf(3.14159265358979f * s[i], &sn, &cs);
        for (int f = 0; f < 6; ++f) {
            d2[3 + 6 * i + f] = f2bf2(sn, cs);
            const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
            sn = s2;
            cs = c2;
        }
    }
    for (int k = kFeatures / 2; k < kIn / 2; ++k) d2[k] = 0u;
}

struct FeatArgs {
    rk::GView g;
    f3 cam;
    float lo[3], hi[3];
};

__device__ __forceinline__ bool pixel_features(const FeatArgs &f, uint32_t p, uint16_t *dst) {
    const f3 pos = f3{f.g.pos[p], f.g.pos[f.g.P + p], f.g.pos[2 * f.g.P + p]};
    const f3 n = rk::load_normal(f.g, p);
    make_features(pos, n, f.cam, f.lo, f.hi, dst);
    return rt::dot(n, n) != 0.0f;
}

__global__ __launch_bounds__(256) void k_features(const FeatArgs f, const uint32_t *pixels, uint32_t n,
                                                  uint16_t *out) {
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    if (b >= n) return;
    uint16_t *dst = out + static_cast<size_t>(b) * kIn;
    if (pixels[b] < f.g.W * f.g.H) {
        pixel_features(f, pixels[b], dst);
    } else {
        for (int k = 0; k < kIn; ++k) dst[k] = 0;
    }
}

// -------------------------------------------------------------------------------- GEMM ------
struct GemmArgs {
    const uint16_t *A;
    int64_t sam, sak;
    const uint16_t *B;
    int64_t sbk, sbn;
    int M, N, K;
    const float *bias;     // per output column, or null
    int relu;              // clamp at 0
    const uint16_t *gate;  // out *= (gate(m, n) > 0), or null
    int64_t sgm, sgn;
    float *Cf;             // fp32 output, or null
    uint16_t *Cb;          // bf16 output, or null
    int64_t scm, scn;
    int kChunk;            // split-K: workgroup z covers K range [z*kChunk, (z+1)*kChunk) and
    int64_t scz;           // writes its fp32 partial at Cf + z*scz (no bias/relu/gate); 0 = off
};

// 64x64 output tile per 256-lane workgroup; each wave owns a 32x32 quarter. K advances in steps
// of 32 through LDS tiles As[64][32] and Bt[64][32] (B staged transposed so every fragment is
// eight consecutive bf16). Requires M, N multiples of 64, K a multiple of 32 and 16-byte aligned
// operands whose strides are multiples of 8 elements.
__global__ __launch_bounds__(256) void k_gemm(const GemmArgs g) {
    __shared__ __align__(16) uint16_t As[64][40];
    __shared__ __align__(16) uint16_t Bt[64][40];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kBeg = g.kChunk ? static_cast<int>(blockIdx.z) * g.kChunk : 0;
    const int kEnd = g.kChunk ? kBeg + g.kChunk : g.K;
    // staging order: consecutive lanes read consecutive addresses of whichever operand index is
    // contiguous in memory
    const bool aRowFast = g.sak != 1, bRowFast = g.sbn == 1;
    f32x16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    // 16-byte loads where one operand index is contiguous (every launch here): a lane loads 8
    // consecutive k of one row, or 8 consecutive rows of one k and scatters them to LDS
    const bool vecA = g.sak == 1 || g.sam == 1, vecB = g.sbk == 1 || g.sbn == 1;
    for (int k0 = kBeg; k0 < kEnd; k0 += 32) {
        if (vecA) {
            if (g.sak == 1) {
                const int r = tid >> 2, c = (tid & 3) * 8;
                *reinterpret_cast<uint4 *>(&As[r][c]) = *reinterpret_cast<const uint4 *>(&g.A[(m0 + r) * g.sam + k0 + c]);
            } else {
                const int kk = tid >> 3, r = (tid & 7) * 8;
                const uint4 v = *reinterpret_cast<const uint4 *>(&g.A[(k0 + kk) * g.sak + m0 + r]);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    As[r + 2 * i][kk] = static_cast<uint16_t>(w[i] & 0xffffu);
                    As[r + 2 * i + 1][kk] = static_cast<uint16_t>(w[i] >> 16);
                }
            }
        }
        if (vecB) {
            if (g.sbk == 1) {
                const int r = tid >> 2, c = (tid & 3) * 8;
                *reinterpret_cast<uint4 *>(&Bt[r][c]) = *reinterpret_cast<const uint4 *>(&g.B[k0 + c + (n0 + r) * g.sbn]);
            } else {
                const int kk = tid >> 3, r = (tid & 7) * 8;
                const uint4 v = *reinterpret_cast<const uint4 *>(&g.B[(k0 + kk) * g.sbk + n0 + r]);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    Bt[r + 2 * i][kk] = static_cast<uint16_t>(w[i] & 0xffffu);
                    Bt[r + 2 * i + 1][kk] = static_cast<uint16_t>(w[i] >> 16);
                }
            }
        }
        if (!vecA || !vecB) {
            for (int e = tid; e < 64 * 32; e += 256) {
                if (!vecA) {
                    const int ra = aRowFast ? (e & 63) : (e >> 5), ca = aRowFast ? (e >> 6) : (e & 31);
                    As[ra][ca] = g.A[(m0 + ra) * g.sam + (k0 + ca) * g.sak];
                }
                if (!vecB) {
                    const int rb = bRowFast ? (e & 63) : (e >> 5), cb = bRowFast ? (e >> 6) : (e & 31);
                    Bt[rb][cb] = g.B[(k0 + cb) * g.sbk + (n0 + rb) * g.sbn];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 32; ks += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm + (lane & 31)][ks + 8 * (lane >> 5)]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bt[wn + (lane & 31)][ks + 8 * (lane >> 5)]);
            acc = mfma32(a, b, acc);
        }
        __syncthreads();
    }
    const int col = n0 + wn + (lane & 31);
    if (g.kChunk) {
        float *P = g.Cf + blockIdx.z * g.scz;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            P[row * g.scm + col * g.scn] = acc[i];
        }
        return;
    }
    const float bias = g.bias ? g.bias[col] : 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        float v = acc[i] + bias;
        if (g.relu) v = v > 0.0f ? v : 0.0f;
        if (g.gate && !(bf2f(g.gate[row * g.sgm + col * g.sgn]) > 0.0f)) v = 0.0f;
        if (g.Cf) g.Cf[row * g.scm + col * g.scn] = v;
        if (g.Cb) g.Cb[row * g.scm + col * g.scn] = f2bf(v);
    }
}

// ----------------------------------------------------------------------------- loss/grad ----
// dZ = (softmax(z) - onehot(target)) / n_valid for sample b with mask, 0 otherwise (bf16), by the
// 64 lanes of one wave (lanes stride over the K logits, reductions across the wave); returns the
// sample's loss divided by n_valid (0 for a sample that does not train).
__device__ __forceinline__ float softmax_ce_row(const float *Z, int K, const int32_t *target, const uint32_t *mask,
                                                int n, const uint32_t *nValid, uint16_t *dZ, int b, int lane) {
    if (b >= n) return 0.0f;
    const float *z = Z + static_cast<size_t>(b) * K;
    uint16_t *d = dZ + static_cast<size_t>(b) * K;
    const int t = target[b];
    if (!mask[b] || t < 0 || t >= K) {  // an out-of-range target does not train
        for (int c = lane; c < K; c += 64) d[c] = 0;
        return 0.0f;
    }
    const float inv = 1.0f / static_cast<float>(*nValid);
    float mx = -INFINITY;
    for (int c = lane; c < K; c += 64) mx = fmaxf(mx, z[c]);
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    float se = 0.0f;
    for (int c = lane; c < K; c += 64) se += expf(z[c] - mx);
    for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off);
    const float rse = 1.0f / se;
    for (int c = lane; c < K; c += 64) d[c] = f2bf((expf(z[c] - mx) * rse - (c == t ? 1.0f : 0.0f)) * inv);
    return -(z[t] - mx - logf(se)) * inv;
}

// The mean cross-entropy's gradient for 4 samples per workgroup (softmax_ce_row); the loss of the
// workgroup's samples, added in sample order, goes to lossPart[blockIdx.x] (k_sum_partials then
// adds the workgroups' sums in order).
__global__ __launch_bounds__(256) void k_softmax_ce(const float *Z, int K, const int32_t *target,
                                                    const uint32_t *mask, int n,
                                                    const uint32_t *nValid, uint16_t *dZ,
                                                    float *lossPart) {
    __shared__ float wl[4];
    const int w = static_cast<int>(threadIdx.x >> 6);
    const int lane = static_cast<int>(threadIdx.x & 63);
    const float sampleLoss = softmax_ce_row(Z, K, target, mask, n, nValid, dZ, blockIdx.x * 4 + w, lane);
    if (lane == 0) wl[w] = sampleLoss;
    __syncthreads();
    if (threadIdx.x == 0) lossPart[blockIdx.x] = ((wl[0] + wl[1]) + wl[2]) + wl[3];
}

__global__ __launch_bounds__(256) void k_count_mask(const uint32_t *mask, uint32_t n, uint32_t *count) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const bool on = i < n && mask[i] != 0u;
    const uint64_t b = __ballot(on);
    if ((threadIdx.x & 63u) == 0u && b) atomicAdd(count, static_cast<uint32_t>(__popcll(b)));
}

__global__ void k_copy_f32(const float *src, float *dst) { *dst = *src; }

// Column sums of a bf16 [rows][cols] matrix in two deterministic steps: workgroup (x, y) sums
// rows [y*chunk, (y+1)*chunk) of columns [64x, 64x+64) into part[y][col]; k_sum_partials adds
// the partials of each column in order.
__global__ __launch_bounds__(256) void k_colsum_partial(const uint16_t *G, int cols, int chunk,
                                                        float *part) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + static_cast<int>(threadIdx.x & 63);
    const int r0 = blockIdx.y * chunk, q = static_cast<int>(threadIdx.x >> 6);
    float sum = 0.0f;
    for (int r = r0 + q; r < r0 + chunk; r += 4) sum += bf2f(G[static_cast<size_t>(r) * cols + c]);
    red[q][threadIdx.x & 63] = sum;
    __syncthreads();
    if (q == 0) part[static_cast<size_t>(blockIdx.y) * cols + c] = ((red[0][c & 63] + red[1][c & 63]) + red[2][c & 63]) + red[3][c & 63];
}

// Up to 9 reductions in one launch: for segment j, out_j[i] = sum over s < nParts_j of
// part_j[s * count_j + i], in order of s. Lane t of the grid takes element t of the segments laid
// end to end.
struct SumSegments {
    const float *part[9];
    float *out[9];
    int nParts[9];
    int64_t count[9];
    int n;
};
__global__ __launch_bounds__(256) void k_sum_partials(const SumSegments a) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    int j = 0;
    while (j < a.n && i >= a.count[j]) i -= a.count[j++];
    if (j >= a.n) return;
    const float *part = a.part[j];
    const int64_t count = a.count[j];
    float sum = 0.0f;
    // unrolled so that the loads of 8 parts are in flight together; the additions stay in order
#pragma unroll 8
    for (int s = 0; s < a.nParts[j]; ++s) sum += part[s * count + i];
    a.out[j][i] = sum;
}

// --------------------------------------------------------------------------------- Adam -----
struct AdamArgs {
    float *w;          // fp32 master weights (nW), then biases (nB) in `b`
    float *b;
    uint16_t *wbf;     // bf16 copy of the weights
    const float *gw, *gb;
    float *m, *v;      // nW + nB each
    uint32_t nW, nB;
    float lr, b1, b2, eps;
    double b1d, b2d;               // b1, b2 in double for the bias corrections
    uint32_t stepHost;             // step t, unless stepDev is set
    const uint32_t *stepDev;
};

// b^t by binary exponentiation in double (the same products in the same order wherever t is
// read), for the bias corrections 1 - b1^t and 1 - b2^t.
__host__ __device__ inline double pow_step(double b, uint32_t t) {
    double r = 1.0;
    while (t) {
        if (t & 1u) r *= b;
        b *= b;
        t >>= 1;
    }
    return r;
}

// Advances the device step count by one (lane 0; a vector atomic).
__global__ void k_adam_count(uint32_t *step) {
    if (threadIdx.x == 0) atomicAdd(step, 1u);
}

// Adam step; skipped when the batch held no training sample (*nValid == 0; nValid null: always).
__global__ __launch_bounds__(256) void k_adam_guarded(const AdamArgs a, const uint32_t *nValid) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.nW + a.nB || (nValid && *nValid == 0u)) return;
    const bool isW = i < a.nW;
    const float g = isW ? a.gw[i] : a.gb[i - a.nW];
    const uint32_t t = a.stepDev ? *a.stepDev : a.stepHost;
    const float c1 = static_cast<float>(1.0 - pow_step(a.b1d, t));
    const float c2 = static_cast<float>(1.0 - pow_step(a.b2d, t));
    const float m = a.b1 * a.m[i] + (1.0f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.0f - a.b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    const float step = a.lr * (m / c1) / (sqrtf(v / c2) + a.eps);
    if (isW) {
        const float w = a.w[i] - step;
        a.w[i] = w;
        a.wbf[i] = f2bf(w);
    } else {
        a.b[i - a.nW] -= step;
    }
}

// He-uniform weights U(-sqrt(6/fan_in), sqrt(6/fan_in)) from PCG32(seed, element index); zero
// biases except the output layer, which starts at `prior` (log cluster power shares) so the
// untrained proposal already samples clusters by power. Adam moments start at zero.
__global__ __launch_bounds__(256) void k_mlp_init(float *w, uint16_t *wbf, float *b, float *m, float *v,
                                                  uint32_t nW, uint32_t nB, uint32_t K, const float *prior,
                                                  uint64_t seed) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < nW) {
        rt::Pcg r = rt::pcg_seed(seed, i);
        const float u = rt::pcg_float(r) * 2.0f - 1.0f;
        const float val = u * sqrtf(6.0f / static_cast<float>(kIn));
        w[i] = val;
        wbf[i] = f2bf(val);
    }
    if (i < nB) b[i] = (i >= 3u * kHidden) ? prior[i - 3u * kHidden] : 0.0f;
    if (i < nW + nB) {
        m[i] = 0.0f;
        v[i] = 0.0f;
    }
    (void)K;
}

// ------------------------------------------------------------------ training-sample gather --
struct GatherArgs {
    FeatArgs f;
    const restir_reservoir *res;
    const int32_t *lightCluster;
    int32_t nLights;  // lights the clusters cover (the alias table's index space)
    int32_t triBase;  // index of triangle light 0 in that space, -1 if it has none
    uint32_t npx, n, seed;
    uint16_t *X;
    int32_t *target;
    uint32_t *mask, *nValid;
    uint32_t *pixels;
};

// Sample b takes a pseudo-random pixel; it trains (mask = 1) if that pixel's final reservoir
// holds a light of the clustered index space with W > 0, with the light's cluster as the target
// (a reservoir names triangle light t as -1 - t).
__global__ __launch_bounds__(256) void k_train_gather(const GatherArgs a) {
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    if (b >= a.n) return;
    rt::Pcg r = rt::pcg_seed(a.seed, b);
    const uint32_t p = rt::pcg_next(r) % a.npx;
    a.pixels[b] = p;
    const bool surface = pixel_features(a.f, p, a.X + static_cast<size_t>(b) * kIn);
    const float4 q2 = reinterpret_cast<const float4 *>(a.res + p)[2];  // lightIndex, pHat, sumW, w
    const int32_t li = __float_as_int(q2.x);
    const int32_t j = li >= 0 ? li : (a.triBase >= 0 ? a.triBase + (-1 - li) : -1);
    const bool ok = surface && j >= 0 && j < a.nLights && q2.w > 0.0f && q2.z > 0.0f;
    a.mask[b] = ok ? 1u : 0u;
    a.target[b] = ok ? a.lightCluster[j] : 0;
    if (ok) atomicAdd(a.nValid, 1u);
}

// ------------------------------------------------------------------- fused inference -------
struct ProposeArgs {
    FeatArgs f;
    const uint16_t *wbf;   // W1..W4 bf16 [out][in]
    const float *bias;
    int K;
    uint32_t npx, M, seed;
    float mix;             // weight of the source distribution
    // proposal tables
    const float4 *alias;   // global alias table (aliasTableColumn)
    int32_t nAlias;
    const int32_t *lightCluster;
    const float *lightPWithin;
    const int32_t *clusterBegin, *clusterCount, *clusterLights;
    const float4 *clusterAlias;  // per-cluster alias columns (alias = local index)
    const float4 *clusterFused;  // per slot: {prob, light, alias light, p_within}, {alias p_within, p_src, alias p_src, 0}
    int32_t *candLight;
    float *candQ;
};

// pixel quads per workgroup of k_propose_wide: 128 quads, 256 lanes and 69.6 KB of LDS, so that
// two workgroups share a CU and one's barriers and weight loads overlap the other's work
constexpr int kTilePx = 128;
constexpr int kRow = kIn + 8;                // padded LDS row (bf16) against bank conflicts

__device__ __forceinline__ void stage_weights(uint16_t *wbuf, const uint16_t *src, int rows) {
    // rows x 128 bf16 from global into LDS rows of kRow, 16 bytes per lane per step
    for (int e = threadIdx.x; e < rows * (kIn / 8); e += blockDim.x) {
        const int r = e / (kIn / 8), c = (e % (kIn / 8)) * 8;
        *reinterpret_cast<uint4 *>(&wbuf[r * kRow + c]) = *reinterpret_cast<const uint4 *>(&src[r * kIn + c]);
    }
}

template <int NB>
__device__ __forceinline__ void layer_mfma(const uint16_t *act, const uint16_t *wbuf, const float *bias,
                                           int wave, int lane, f32x16 (&acc)[NB]) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const float bv = bias[32 * nb + (lane & 31)];
        for (int i = 0; i < 16; ++i) acc[nb][i] = bv;
    }
#pragma unroll
    for (int s = 0; s < kIn / 16; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&act[(32 * wave + (lane & 31)) * kRow + 16 * s + 8 * (lane >> 5)]);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&wbuf[(32 * nb + (lane & 31)) * kRow + 16 * s + 8 * (lane >> 5)]);
            acc[nb] = mfma32(a, b, acc[nb]);
        }
    }
}

__device__ __forceinline__ float4 unpack_h4(uint2 v) {  // four fp16 -> float
    const float2 lo = __half22float2(__builtin_bit_cast(__half2, v.x));
    const float2 hi = __half22float2(__builtin_bit_cast(__half2, v.y));
    return make_float4(lo.x, lo.y, hi.x, hi.y);
}

// The pixel whose features stand for quad (qx, qy): its first surface pixel in the order
// (0,0), (1,0), (0,1), (1,1) inside the image, else its top-left pixel.
__device__ __forceinline__ uint32_t quad_pixel(const rk::GView &g, uint32_t qx, uint32_t qy) {
    const uint32_t x0 = 2 * qx, y0 = 2 * qy;
    for (int j = 0; j < 4; ++j) {
        const uint32_t x = x0 + (j & 1), y = y0 + (j >> 1);
        if (x >= g.W || y >= g.H) continue;
        const uint32_t p = y * g.W + x;
        const f3 nq = rk::load_normal(g, p);
        if ((nq.x * nq.x + nq.y * nq.y) + nq.z * nq.z != 0.0f) return p;
    }
    return y0 * g.W + x0;
}

// Features of the workgroup's TILE quads (2 * TILE lanes, a wave per 32 quads) and the three
// hidden layers: leaves H3 in `act` (bf16 rows of kRow) and the first `outRows` rows of W4 staged
// in `wbuf`. Ends with a barrier.
template <int TILE>
__device__ __forceinline__ void propose_hidden(const ProposeArgs &a, uint16_t *act, uint16_t *wbuf, uint32_t q0,
                                               uint32_t QW, uint32_t QH, int outRows) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // features (lanes 0..TILE-1) while the other TILE lanes stage W1
    if (tid < TILE) {
        const uint32_t qd = q0 + tid;
        uint16_t *row = act + tid * kRow;
        if (qd < QW * QH) pixel_features(a.f, quad_pixel(a.f.g, qd % QW, qd / QW), row);
        else for (int k = 0; k < kIn; ++k) row[k] = 0;
    } else {
        for (int e = tid - TILE; e < kHidden * (kIn / 8); e += TILE) {
            const int r = e / (kIn / 8), c = (e % (kIn / 8)) * 8;
            *reinterpret_cast<uint4 *>(&wbuf[r * kRow + c]) = *reinterpret_cast<const uint4 *>(&a.wbf[r * kIn + c]);
        }
    }
    __syncthreads();
    for (int L = 0; L < 3; ++L) {
        f32x16 acc[4];
        layer_mfma<4>(act, wbuf, a.bias + b_off(L), wave, lane, acc);
        __syncthreads();
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int row = 32 * wave + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const float v = acc[nb][i];
                act[row * kRow + 32 * nb + (lane & 31)] = f2bf(v > 0.0f ? v : 0.0f);
            }
        stage_weights(wbuf, a.wbf + w_off(L + 1), L + 1 < 3 ? kHidden : outRows);
        __syncthreads();
    }
}

// One workgroup of 256 lanes (4 waves) proposes for kProposeTile = 128 pixel quads (2x2 pixels
// each): one MLP evaluation per quad, each pixel drawing its own candidates from the quad's
// distribution with its own random stream (q is that distribution, so RIS stays unbiased). LDS
// holds the activations [128][136] bf16 and 128 rows of one layer's weights (W4 of K = 256 in two
// halves); after the last layer the same LDS holds the fp16 softmax numerators [128][K + 4].
// 69.6 KB of LDS lets two workgroups share a CU, so one's barriers and loads overlap the other's
// work.
constexpr int kProposeTile = 128;
template <int KC>
__global__ __launch_bounds__(2 * kProposeTile) void k_propose(const ProposeArgs a) {
    constexpr int T = kProposeTile;
    __shared__ __align__(16) uint16_t lds[2 * T * kRow];
    static_assert(T * (KC + 4) <= 2 * T * kRow, "softmax numerators fit the LDS tile");
    uint16_t *act = lds;
    uint16_t *wbuf = lds + T * kRow;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // one MLP row per 2x2 pixel quad: the quad's first surface pixel in the order (0,0), (1,0),
    // (0,1), (1,1) stands for it (quad_pixel)
    const uint32_t W = a.f.g.W, H = a.f.g.H, QW = (W + 1) / 2, QH = (H + 1) / 2;
    const uint32_t q0 = blockIdx.x * T;
    propose_hidden<T>(a, act, wbuf, q0, QW, QH, KC < 128 ? KC : 128);
    {
        f32x16 acc[KC / 32];
        if constexpr (KC <= 128) {
            layer_mfma<KC / 32>(act, wbuf, a.bias + b_off(3), wave, lane, acc);
        } else {
            static_assert(KC == 256, "W4 in two halves of 128 rows");
            f32x16 lo[4], hi[4];
            layer_mfma<4>(act, wbuf, a.bias + b_off(3), wave, lane, lo);
            __syncthreads();
            stage_weights(wbuf, a.wbf + w_off(3) + 128 * kIn, 128);
            __syncthreads();
            layer_mfma<4>(act, wbuf, a.bias + b_off(3) + 128, wave, lane, hi);
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                acc[nb] = lo[nb];
                acc[nb + 4] = hi[nb];
            }
        }
        __syncthreads();
        // softmax numerators e = exp(z - rowmax) as fp16 rows of KC + 4 halves: the row maximum
        // is reduced across the 32 lanes holding a row's columns; the padding keeps rows 8-byte
        // aligned and spreads the per-pixel scans below over the banks (lane t reads dwords
        // 130 t + c/2, c/2 + 1)
        __half *e = reinterpret_cast<__half *>(lds);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            float m = acc[0][i];
#pragma unroll
            for (int nb = 1; nb < KC / 32; ++nb) m = fmaxf(m, acc[nb][i]);
#pragma unroll
            for (int off = 16; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            const int row = 32 * wave + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
#pragma unroll
            for (int nb = 0; nb < KC / 32; ++nb)
                e[row * (KC + 4) + 32 * nb + (lane & 31)] = __float2half(__expf(acc[nb][i] - m));
        }
        __syncthreads();
    }
    // two lanes per quad, each drawing for one row of its 2x2 pixels
    const uint32_t qd = q0 + (tid >> 1);
    if (qd >= QW * QH) return;
    const uint32_t qx = qd % QW, qy = qd / QW;
    // the cluster pmf is e[c] / se with se the sequential sum of the stored fp16 numerators; the
    // inverse-CDF scan below accumulates the same values in the same order, so the sampled
    // distribution and the q written for each candidate agree
    // one 8-byte LDS read yields 4 numerators. Cluster blocks of 32: block sums in registers
    // (sequential inside a block, blocks in order), so a draw picks its block with KC/32
    // compares and scans at most 32 numerators of that block
    const uint2 *row = reinterpret_cast<const uint2 *>(reinterpret_cast<const __half *>(lds) + (tid >> 1) * (KC + 4));
    constexpr int kBlocks = KC / 32;
    // the two lanes of a quad read the same row: each sums every other block (block 2j + lane
    // parity), and the partner's sums come by a lane swap; se then adds the blocks in order
    static_assert(kBlocks % 2 == 0, "block pairs");
    const bool odd = (tid & 1) != 0;
    float bs[kBlocks];
#pragma unroll
    for (int j = 0; j < kBlocks / 2; ++j) {
        const int b = 2 * j + (odd ? 1 : 0);
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 f = unpack_h4(row[8 * b + q]);
            sum += f.x;
            sum += f.y;
            sum += f.z;
            sum += f.w;
        }
        const float other = __shfl_xor(sum, 1);
        bs[2 * j] = odd ? other : sum;
        bs[2 * j + 1] = odd ? sum : other;
    }
    float se = 0.0f;
#pragma unroll
    for (int b = 0; b < kBlocks; ++b) se += bs[b];
    const float rse = 1.0f / se;
    // cluster of the draw t in [0, se): the first c whose running sum exceeds t (numerators of
    // zero are never chosen); rounding overflow falls to the last cluster of the block
    auto pick_cluster = [&](float t) {
        int blk = kBlocks - 1;
        float base = 0.0f, cum = 0.0f;
        bool found = false;
#pragma unroll
        for (int b = 0; b < kBlocks; ++b) {
            const float nx = cum + bs[b];
            if (!found && (t < nx || b == kBlocks - 1)) {
                blk = b;
                base = cum;
                found = true;
            }
            cum = nx;
        }
        int c = 32 * blk + 31;
        float run = base;
        bool hit = false;
        // the block's 32 numerators in one batch of LDS reads, then the same sequential sum
        uint2 raw[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) raw[q] = row[8 * blk + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 f = unpack_h4(raw[q]);
            const float e4[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                run += e4[k];
                if (!hit && t < run) {
                    c = 32 * blk + 4 * q + k;
                    hit = true;
                }
            }
        }
        return c;
    };
    const int M = static_cast<int>(a.M);
    const FeatArgs &f = a.f;
    const uint32_t py = 2 * qy + static_cast<uint32_t>(tid & 1);
    for (uint32_t px = 2 * qx; px < 2 * qx + 2; ++px) {
    if (px >= W || py >= H) continue;
    const uint32_t p = py * W + px;
    const f3 n = rk::load_normal(f.g, p);
    if (rt::dot(n, n) == 0.0f) continue;  // background: Pass A draws no candidates there
    rt::Pcg rng = rt::pcg_seed(a.seed, p);
    // Candidates in groups of 8, each step's loads issued together (no load waits on another
    // candidate's). Per candidate: one draw for the mixture, one for the cluster (MLP
    // candidates), then two for the alias column.
    for (int j0 = 0; j0 < M; j0 += 8) {
        float r1[8], r2[8], pw[8], ps[8];
        int32_t cl[8], col[8], cnt[8], beg[8], l[8];
        float4 t[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const int j = j0 + jj;
            cl[jj] = -2;  // beyond M
            if (j < M) {
                const bool src = rt::pcg_float(rng) < a.mix;
                cl[jj] = src ? -1 : pick_cluster(rt::pcg_float(rng) * se);
                r1[jj] = rt::pcg_float(rng);
                r2[jj] = rt::pcg_float(rng);
            }
        }
        // step 1: source candidates read their alias column, MLP candidates their cluster
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            if (cl[jj] == -1) {
                int32_t c = static_cast<int32_t>(static_cast<float>(a.nAlias) * r1[jj]);
                col[jj] = c < a.nAlias - 1 ? c : a.nAlias - 1;
                t[jj] = a.alias[col[jj]];
            } else if (cl[jj] >= 0) {
                cnt[jj] = a.clusterCount[cl[jj]];
                beg[jj] = a.clusterBegin[cl[jj]];
            }
        }
        // step 2: MLP candidates read the fused record of the drawn slot of the cluster's own
        // alias table: the column's and its alias's light, p_within and p_src in 32 bytes
        float4 t2[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            if (cl[jj] >= 0 && cnt[jj] > 0) {
                int32_t c = static_cast<int32_t>(static_cast<float>(cnt[jj]) * r1[jj]);
                col[jj] = c < cnt[jj] - 1 ? c : cnt[jj] - 1;
                const float4 *rec = a.clusterFused + 2 * static_cast<size_t>(beg[jj] + col[jj]);
                t[jj] = rec[0];
                t2[jj] = rec[1];
            }
        }
        // step 3: the light; its probability under the table that drew it comes with the column
        // (oriProb of the column or of its alias, as the alias table stores both), and an MLP
        // candidate's p_src with its record
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            l[jj] = -1;
            if (cl[jj] == -1) {
                const bool keep = t[jj].x > r2[jj];
                l[jj] = keep ? col[jj] : __float_as_int(t[jj].y);
                ps[jj] = keep ? t[jj].z : t[jj].w;
            } else if (cl[jj] >= 0 && cnt[jj] > 0) {
                const bool keep = t[jj].x > r2[jj];
                l[jj] = __float_as_int(keep ? t[jj].y : t[jj].z);
                pw[jj] = keep ? t[jj].w : t2[jj].x;
                ps[jj] = keep ? t2[jj].y : t2[jj].z;
            }
        }
        // step 4: a source candidate's cluster and p_within
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            if (l[jj] >= 0 && cl[jj] == -1) {
                cl[jj] = a.lightCluster[l[jj]];
                pw[jj] = a.lightPWithin[l[jj]];
            }
        }
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const int j = j0 + jj;
            if (j >= M) continue;
            float q = 0.0f;
            if (l[jj] >= 0) {
                const int c = cl[jj];
                const float pi = __half2float(reinterpret_cast<const __half *>(row)[c]) * rse;
                q = (1.0f - a.mix) * pi * pw[jj] + a.mix * ps[jj];
            }
            a.candLight[static_cast<size_t>(j) * a.npx + p] = l[jj];
            a.candQ[static_cast<size_t>(j) * a.npx + p] = q;
        }
    }
    }  // pixels of this lane
}

// ----------------------------------------------------------- fused inference, many clusters --
// k_propose keeps all K softmax numerators of a quad in LDS, which has room for K <= 256. For
// larger K (a multiple of 128, at most 4096) the output layer streams through LDS in column
// blocks of kSuper = 128 clusters, three times over the same H3 rows:
//   1. the row maxima of the logits;
//   2. the numerators e = fp16(exp(z - max)) and their row sum se, accumulated as in k_propose:
//      blocks of 32 numerators summed in order, block sums added to se in cluster order;
//   3. the same numerators again: each of the lane's pending draws t = u * se selects the first
//      block whose running sum exceeds t (the last block if rounding leaves none) and the first
//      cluster in it whose running sum from the block's start exceeds t (the block's last if
//      none), exactly k_propose's pick_cluster over all K / 32 blocks; every candidate picks
//      up the numerator of its cluster, whether drawn from the MLP or the cluster of a light
//      drawn from the source table.
// Then each pixel's random stream is replayed to draw the lights and write (l, q) as k_propose
// does, with the same draws in the same order. At most kWideCand candidates per pixel.
constexpr int kERow = kSuper + 4;  // fp16 numerators per LDS row

__device__ __forceinline__ void super_block_sums(const uint2 *row, float (&bs)[kSuper / 32]) {
#pragma unroll
    for (int b = 0; b < kSuper / 32; ++b) {
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 f = unpack_h4(row[8 * b + q]);
            sum += f.x;
            sum += f.y;
            sum += f.z;
            sum += f.w;
        }
        bs[b] = sum;
    }
}

__global__ __launch_bounds__(2 * kTilePx) void k_propose_wide(const ProposeArgs a) {
    __shared__ __align__(16) uint16_t lds[2 * kTilePx * kRow];
    static_assert(kTilePx * kERow <= kTilePx * kRow, "a column block's numerators fit the weight buffer");
    static_assert(kTilePx % 32 == 0, "32 quads per wave in the MFMA layers, two lanes per quad");
    uint16_t *act = lds;
    uint16_t *wbuf = lds + kTilePx * kRow;
    __half *etile = reinterpret_cast<__half *>(wbuf);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t W = a.f.g.W, H = a.f.g.H, QW = (W + 1) / 2, QH = (H + 1) / 2;
    const uint32_t q0 = blockIdx.x * kTilePx;
    const int nSuper = a.K / kSuper;
    const uint16_t *w4 = a.wbf + w_off(3);
    const float *b4 = a.bias + b_off(3);
    propose_hidden<kTilePx>(a, act, wbuf, q0, QW, QH, kSuper);

    // 1. row maxima: each lane covers 16 rows x 4 columns of every block, then the 32 lanes
    // sharing a row reduce
    float rmax[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) rmax[i] = -INFINITY;
    for (int sb = 0; sb < nSuper; ++sb) {
        if (sb > 0) {
            stage_weights(wbuf, w4 + static_cast<size_t>(sb) * kSuper * kIn, kSuper);
            __syncthreads();
        }
        f32x16 acc[4];
        layer_mfma<4>(act, wbuf, b4 + sb * kSuper, wave, lane, acc);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) rmax[i] = fmaxf(rmax[i], acc[nb][i]);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i)
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) rmax[i] = fmaxf(rmax[i], __shfl_xor(rmax[i], off));

    // one column block's numerators into the LDS tile (which overlays the staged weights)
    auto numerators = [&](int sb) {
        stage_weights(wbuf, w4 + static_cast<size_t>(sb) * kSuper * kIn, kSuper);
        __syncthreads();
        f32x16 acc[4];
        layer_mfma<4>(act, wbuf, b4 + sb * kSuper, wave, lane, acc);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = 32 * wave + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                etile[row * kERow + 32 * nb + (lane & 31)] = __float2half(__expf(acc[nb][i] - rmax[i]));
        }
        __syncthreads();
    };

    // lane pair per quad from here on, each lane drawing for one row of its 2x2 pixels
    const uint32_t qd = q0 + (tid >> 1);
    const bool quadOk = qd < QW * QH;
    const uint32_t qx = quadOk ? qd % QW : 0, qy = quadOk ? qd / QW : 0;
    const uint32_t py = 2 * qy + static_cast<uint32_t>(tid & 1);
    const uint2 *erow = reinterpret_cast<const uint2 *>(etile + (tid >> 1) * kERow);
    const __half *erowh = etile + (tid >> 1) * kERow;

    // 2. the row sum
    float se = 0.0f;
    for (int sb = 0; sb < nSuper; ++sb) {
        numerators(sb);
        float bs[kSuper / 32];
        super_block_sums(erow, bs);
#pragma unroll
        for (int b = 0; b < kSuper / 32; ++b) se += bs[b];
        __syncthreads();  // the next block's weights overwrite the tile
    }
    const float rse = 1.0f / se;

    // the lane's candidates: slot k = s * kWideCand + j for pixel s of its row, candidate j.
    // MLP draws carry their target t; source draws the cluster of the drawn light.
    const int M = static_cast<int>(a.M);
    const FeatArgs &f = a.f;
    uint32_t pix[2];
    bool pixOk[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const uint32_t px = 2 * qx + s;
        pix[s] = py * W + px;
        pixOk[s] = false;
        if (quadOk && px < W && py < H) {
            const uint32_t p = pix[s];
            const f3 n = rk::load_normal(f.g, p);
            pixOk[s] = rt::dot(n, n) != 0.0f;  // background: Pass A draws no candidates there
        }
    }
    float tk[2 * kWideCand], ek[2 * kWideCand];
    int32_t ck[2 * kWideCand];
    uint32_t pend = 0u;  // MLP draws whose cluster is not found yet
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        rt::Pcg rng = rt::pcg_seed(a.seed, pix[s]);
#pragma unroll
        for (int j = 0; j < kWideCand; ++j) {
            const int k = s * kWideCand + j;
            tk[k] = 0.0f;
            ek[k] = 0.0f;
            ck[k] = -1;
            if (!pixOk[s] || j >= M) continue;
            const bool src = rt::pcg_float(rng) < a.mix;
            if (!src) {
                tk[k] = rt::pcg_float(rng) * se;
                ck[k] = a.K - 1;
                pend |= 1u << k;
            }
            const float r1 = rt::pcg_float(rng), r2 = rt::pcg_float(rng);
            if (src) {
                int32_t c = static_cast<int32_t>(static_cast<float>(a.nAlias) * r1);
                c = c < a.nAlias - 1 ? c : a.nAlias - 1;
                const float4 t4 = a.alias[c];
                const int32_t l = t4.x > r2 ? c : __float_as_int(t4.y);
                ck[k] = a.lightCluster[l];
            }
        }
    }

    // 3. find the drawn clusters and pick up every candidate's numerator
    float cum = 0.0f;
    for (int sb = 0; sb < nSuper; ++sb) {
        numerators(sb);
        float bs[kSuper / 32];
        super_block_sums(erow, bs);
        float nx[kSuper / 32];
        float c2 = cum;
#pragma unroll
        for (int b = 0; b < kSuper / 32; ++b) {
            c2 += bs[b];
            nx[b] = c2;
        }
        const bool last = sb == nSuper - 1;
        const int c0 = sb * kSuper;
#pragma unroll
        for (int k = 0; k < 2 * kWideCand; ++k) {
            if (pend & (1u << k)) {
                if (tk[k] < nx[kSuper / 32 - 1] || last) {
                    int blk = kSuper / 32 - 1;
                    float base = nx[kSuper / 32 - 2];
                    bool found = false;
#pragma unroll
                    for (int b = 0; b < kSuper / 32; ++b) {
                        if (!found && (tk[k] < nx[b] || b == kSuper / 32 - 1)) {
                            blk = b;
                            base = b == 0 ? cum : nx[b - 1];
                            found = true;
                        }
                    }
                    int c = 32 * blk + 31;
                    float run = base;
                    bool hit = false;
                    for (int q = 0; q < 8 && !hit; ++q) {
                        const float4 e4 = unpack_h4(erow[8 * blk + q]);
                        const float ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk) {
                            run += ev[kk];
                            if (!hit && tk[k] < run) {
                                c = 32 * blk + 4 * q + kk;
                                hit = true;
                            }
                        }
                    }
                    ck[k] = c0 + c;
                    ek[k] = __half2float(erowh[c]);
                    pend &= ~(1u << k);
                }
            } else if (ck[k] >= c0 && ck[k] < c0 + kSuper) {
                ek[k] = __half2float(erowh[ck[k] - c0]);
            }
        }
        cum = nx[kSuper / 32 - 1];
        __syncthreads();
    }

    // replay each pixel's stream: the lights and their q, as k_propose draws them
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (!pixOk[s]) continue;
        const uint32_t p = pix[s];
        rt::Pcg rng = rt::pcg_seed(a.seed, p);
#pragma unroll
        for (int j = 0; j < kWideCand; ++j) {
            if (j >= M) continue;
            const int k = s * kWideCand + j;
            const bool src = rt::pcg_float(rng) < a.mix;
            if (!src) (void)rt::pcg_float(rng);
            const float r1 = rt::pcg_float(rng), r2 = rt::pcg_float(rng);
            int32_t l = -1;
            float pw = 0.0f, ps = 0.0f;
            if (src) {
                int32_t c = static_cast<int32_t>(static_cast<float>(a.nAlias) * r1);
                c = c < a.nAlias - 1 ? c : a.nAlias - 1;
                const float4 t4 = a.alias[c];
                const bool keep = t4.x > r2;
                l = keep ? c : __float_as_int(t4.y);
                ps = keep ? t4.z : t4.w;
                pw = a.lightPWithin[l];
            } else {
                const int32_t cnt = a.clusterCount[ck[k]], beg = a.clusterBegin[ck[k]];
                if (cnt > 0) {
                    int32_t c = static_cast<int32_t>(static_cast<float>(cnt) * r1);
                    c = c < cnt - 1 ? c : cnt - 1;
                    const float4 t4 = a.clusterAlias[beg + c];
                    const bool keep = t4.x > r2;
                    l = a.clusterLights[beg + (keep ? c : __float_as_int(t4.y))];
                    pw = keep ? t4.z : t4.w;
                    ps = a.alias[l].z;
                }
            }
            const float q = l >= 0 ? (1.0f - a.mix) * (ek[k] * rse) * pw + a.mix * ps : 0.0f;
            a.candLight[static_cast<size_t>(j) * a.npx + p] = l;
            a.candQ[static_cast<size_t>(j) * a.npx + p] = q;
        }
    }
}

}  // namespace rm

// ----------------------------------------------------------------------------- launchers ----
namespace rk {

using rm::GemmArgs;
using rm::kHidden;
using rm::kIn;
using rt::f3;

static void gemm(const GemmArgs &g, hipStream_t s) {
    RK_LAUNCH("k_gemm", rm::k_gemm, dim3(g.N / 64, g.M / 64), dim3(256), 0, s, g);
}
static GemmArgs gemm_args(int M, int N, int K) {
    GemmArgs g{};
    g.M = M;
    g.N = N;
    g.K = K;
    return g;
}

// Training tensors carved from one scratch allocation.
struct MlpScratch {
    uint16_t *X, *H[3], *dA, *dB, *dZ;
    float *Z, *gw, *gb, *loss;
    float *part;      // split-K and column-sum partials
    float *lossPart;  // per-workgroup loss sums of k_softmax_ce
    int32_t *target;
    uint32_t *mask, *pixels, *nValid;
};
// split-K of the weight-gradient GEMMs (reduction over the n samples): at most kMaxSplit parts.
// Every layer keeps its own partials (weights: kMaxSplit x rows x kIn, biases: n/64 chunks x rows)
// so that one launch reduces them all after the backward pass.
constexpr int kMaxSplit = 16;
static size_t rows_before(int l) { return static_cast<size_t>(l) * kHidden; }  // output rows of layers < l
static size_t part_floats(int K, uint32_t n) {
    const size_t rows = 3 * static_cast<size_t>(kHidden) + static_cast<size_t>(K);
    return static_cast<size_t>(kMaxSplit) * rows * kIn + static_cast<size_t>(n / 64) * rows;
}
// parts for an M x N (x 64-tiles) gradient over n samples: about 256 workgroups, each K range a
// multiple of 32
static int split_for(int M, int N, int n) {
    const int tiles = (M / 64) * (N / 64);
    int sp = 1;
    while (sp < kMaxSplit && tiles * sp < 256 && (n % (2 * sp * 32)) == 0) sp *= 2;
    return sp;
}
static MlpScratch carve_mlp(void *base, uint32_t n, int K, size_t *bytes = nullptr) {
    Carver c(base);
    const size_t nn = n, kk = static_cast<size_t>(K);
    MlpScratch m;
    m.X = c.take<uint16_t>(2 * nn * kIn);
    for (int l = 0; l < 3; ++l) m.H[l] = c.take<uint16_t>(2 * nn * kHidden);
    m.dA = c.take<uint16_t>(2 * nn * kHidden);
    m.dB = c.take<uint16_t>(2 * nn * kHidden);
    m.dZ = c.take<uint16_t>(2 * nn * kk);
    m.Z = c.take<float>(4 * nn * kk);
    m.target = c.take<int32_t>(4 * nn);
    m.mask = c.take<uint32_t>(4 * nn);
    m.pixels = c.take<uint32_t>(4 * nn);
    m.nValid = c.take<uint32_t>(256);
    m.loss = c.take<float>(256);
    m.gw = c.take<float>(4 * rm::n_weights(K));
    m.gb = c.take<float>(4 * rm::n_bias(K));
    m.part = c.take<float>(4 * part_floats(K, n));
    m.lossPart = c.take<float>(4 * (nn / 4));
    if (bytes) *bytes = c.end();
    return m;
}
size_t mlp_scratch_bytes(uint32_t n, int K) {
    size_t bytes = 0;
    carve_mlp(nullptr, n, K, &bytes);
    return bytes;
}

// forward through the GEMM path, keeping the hidden activations for the backward pass
static void mlp_fwd(const restir_mlp &m, const uint16_t *X, uint32_t n, const MlpScratch &t, float *logits,
                    hipStream_t s) {
    const int K = static_cast<int>(m.n_clusters);
    const uint16_t *in = X;
    for (int l = 0; l < 4; ++l) {
        const int N = l < 3 ? kHidden : K;
        GemmArgs g = gemm_args(static_cast<int>(n), N, kIn);
        g.A = in;
        g.sam = kIn;
        g.sak = 1;
        g.B = m.weights_bf16 + rm::w_off(l);  // B(k, o) = W[o][k]
        g.sbk = 1;
        g.sbn = kIn;
        g.bias = m.bias + rm::b_off(l);
        g.scm = N;
        g.scn = 1;
        if (l < 3) {
            g.relu = 1;
            g.Cb = t.H[l];
            in = t.H[l];
        } else {
            g.Cf = logits;
        }
        gemm(g, s);
    }
}

hipError_t mlp_forward(const restir_mlp &m, const uint16_t *X, uint32_t n, float *logits, void *scratch,
                       hipStream_t s) {
    const MlpScratch t = carve_mlp(scratch, n, static_cast<int>(m.n_clusters));
    mlp_fwd(m, X, n, t, logits, s);
    return hipGetLastError();
}

// mean cross-entropy and its gradient; gw/gb/loss are written (not accumulated)
static void mlp_loss_grad_impl(const restir_mlp &m, const uint16_t *X, const int32_t *target,
                               const uint32_t *mask, uint32_t n, const MlpScratch &t, float *gw, float *gb,
                               float *loss, bool countMask, hipStream_t s) {
    const int K = static_cast<int>(m.n_clusters);
    const int nn = static_cast<int>(n);
    if (countMask) {
        RK_LAUNCH("k_zero_u32", k_zero_u32, dim3(1), dim3(64), 0, s, t.nValid, 1u);
        RK_LAUNCH("k_count_mask", rm::k_count_mask, dim3((n + 255) / 256), dim3(256), 0, s, mask, n, t.nValid);
    }
    mlp_fwd(m, X, n, t, t.Z, s);
    RK_LAUNCH("k_softmax_ce", rm::k_softmax_ce, dim3((n + 3) / 4), dim3(256), 0, s, t.Z, K, target, mask, nn,
                       t.nValid, t.dZ, t.lossPart);
    // layer 4: gW4[o][k] = sum_b dZ[b][o] H3[b][k]
    const uint16_t *dY = t.dZ;
    int N = K;
    rm::SumSegments sums{};
    // the loss: the workgroups' sums of k_softmax_ce, in order
    sums.part[0] = t.lossPart;
    sums.nParts[0] = (nn + 3) / 4;
    sums.count[0] = 1;
    sums.out[0] = loss;
    sums.n = 1;
    int64_t sumCount = 1;
    const int chunks = nn / 64;
    float *partB = t.part + static_cast<size_t>(kMaxSplit) * (3 * static_cast<size_t>(kHidden) + K) * kIn;
    for (int l = 3; l >= 0; --l) {
        float *pw = t.part + static_cast<size_t>(kMaxSplit) * rows_before(l) * kIn;
        float *pb = partB + static_cast<size_t>(chunks) * rows_before(l);
        const uint16_t *in = l > 0 ? t.H[l - 1] : X;
        // split-K partials over the samples, then an in-order reduction (deterministic)
        const int sp = split_for(N, kIn, nn);
        GemmArgs gwA = gemm_args(N, kIn, nn);
        gwA.A = dY;      // A(o, b) = dY[b][o]
        gwA.sam = 1;
        gwA.sak = N;
        gwA.B = in;      // B(b, k) = in[b][k]
        gwA.sbk = kIn;
        gwA.sbn = 1;
        gwA.Cf = pw;
        gwA.scm = kIn;
        gwA.scn = 1;
        gwA.kChunk = nn / sp;
        gwA.scz = static_cast<int64_t>(N) * kIn;
        RK_LAUNCH("k_gemm_splitk", rm::k_gemm, dim3(kIn / 64, N / 64, sp), dim3(256), 0, s, gwA);
        // bias gradients: column sums of dY over 64-row chunks (added in order below)
        RK_LAUNCH("k_colsum_partial", rm::k_colsum_partial, dim3(N / 64, chunks), dim3(256), 0, s, dY, N, 64, pb);
        const int seg = sums.n;
        sums.part[seg] = pw;
        sums.nParts[seg] = sp;
        sums.count[seg] = static_cast<int64_t>(N) * kIn;
        sums.out[seg] = gw + rm::w_off(l);
        sums.part[seg + 1] = pb;
        sums.nParts[seg + 1] = chunks;
        sums.count[seg + 1] = N;
        sums.out[seg + 1] = gb + rm::b_off(l);
        sums.n = seg + 2;
        sumCount += sums.count[seg] + sums.count[seg + 1];
        if (l == 0) break;
        // dIn[b][k] = sum_o dY[b][o] W[o][k], gated by the ReLU of the layer below
        uint16_t *dIn = (dY == t.dA) ? t.dB : t.dA;
        GemmArgs gd = gemm_args(nn, kIn, N);
        gd.A = dY;
        gd.sam = N;
        gd.sak = 1;
        gd.B = m.weights_bf16 + rm::w_off(l);  // B(o, k) = W[o][k]
        gd.sbk = kIn;
        gd.sbn = 1;
        gd.gate = t.H[l - 1];
        gd.sgm = kIn;
        gd.sgn = 1;
        gd.Cb = dIn;
        gd.scm = kIn;
        gd.scn = 1;
        gemm(gd, s);
        dY = dIn;
        N = kHidden;
    }
    // the split-K and bias partials of all four layers, each summed in a fixed order
    RK_LAUNCH("k_sum_partials", rm::k_sum_partials, dim3(static_cast<uint32_t>((sumCount + 255) / 256)), dim3(256), 0,
              s, sums);
}

hipError_t mlp_loss_grad(const restir_mlp &m, const uint16_t *X, const int32_t *target, const uint32_t *mask,
                         uint32_t n, float *gw, float *gb, float *loss, void *scratch, hipStream_t s) {
    const MlpScratch t = carve_mlp(scratch, n, static_cast<int>(m.n_clusters));
    mlp_loss_grad_impl(m, X, target, mask, n, t, gw, gb, loss, true, s);
    return hipGetLastError();
}

hipError_t mlp_init(const restir_mlp &m, uint64_t seed, const float *prior, hipStream_t s) {
    const uint32_t K = m.n_clusters;
    const uint32_t nW = static_cast<uint32_t>(rm::n_weights(K)), nB = static_cast<uint32_t>(rm::n_bias(K));
    RK_LAUNCH("k_mlp_init", rm::k_mlp_init, dim3((nW + nB + 255) / 256), dim3(256), 0, s, m.weights, m.weights_bf16,
                       m.bias, m.adam_m, m.adam_v, nW, nB, K, prior, seed);
    return hipGetLastError();
}

hipError_t mlp_features(const GView &g, f3 cam, const float lo[3], const float hi[3], const uint32_t *pixels,
                        uint32_t n, uint16_t *out, hipStream_t s) {
    rm::FeatArgs f{g, cam, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
    RK_LAUNCH("k_features", rm::k_features, dim3((n + 255) / 256), dim3(256), 0, s, f, pixels, n, out);
    return hipGetLastError();
}

hipError_t mlp_train(restir_mlp &m, const GView &g, f3 cam, const float lo[3], const float hi[3],
                     const restir_reservoir *res, const int32_t *lightCluster, int32_t nLights,
                     int32_t triBase, uint32_t n, float lr,
                     uint32_t seed, float *lossOut, void *scratch, hipStream_t s) {
    const int K = static_cast<int>(m.n_clusters);
    const MlpScratch t = carve_mlp(scratch, n, K);
    rm::GatherArgs ga{};
    ga.f = rm::FeatArgs{g, cam, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
    ga.res = res;
    ga.lightCluster = lightCluster;
    ga.nLights = nLights;
    ga.triBase = triBase;
    ga.npx = g.W * g.H;
    ga.n = n;
    ga.seed = seed;
    ga.X = t.X;
    ga.target = t.target;
    ga.mask = t.mask;
    ga.nValid = t.nValid;
    ga.pixels = t.pixels;
    RK_LAUNCH("k_zero_u32", k_zero_u32, dim3(1), dim3(64), 0, s, t.nValid, 1u);
    RK_LAUNCH("k_train_gather", rm::k_train_gather, dim3((n + 255) / 256), dim3(256), 0, s, ga);
    mlp_loss_grad_impl(m, t.X, t.target, t.mask, n, t, t.gw, t.gb, t.loss, false, s);
    const hipError_t ea = mlp_adam(m, t.gw, t.gb, t.nValid, lr, s);
    if (ea != hipSuccess) return ea;
    if (lossOut) RK_LAUNCH("k_copy_f32", rm::k_copy_f32, dim3(1), dim3(1), 0, s, t.loss, lossOut);
    return hipGetLastError();
}

// One Adam step (beta 0.9 / 0.999, eps 1e-8) from the gradients gw, gb: advances the step count
// (host, and the device copy when set) and updates the fp32 master weights, the moments and the
// bf16 copy. nValid (device, may be null): the step is skipped when it reads 0.
hipError_t mlp_adam(restir_mlp &m, const float *gw, const float *gb, const uint32_t *nValid, float lr,
                    hipStream_t s) {
    const int K = static_cast<int>(m.n_clusters);
    m.adam_step += 1;
    const double b1 = 0.9, b2 = 0.999;
    rm::AdamArgs a{};
    a.w = m.weights;
    a.b = m.bias;
    a.wbf = m.weights_bf16;
    a.gw = gw;
    a.gb = gb;
    a.m = m.adam_m;
    a.v = m.adam_v;
    a.nW = static_cast<uint32_t>(rm::n_weights(K));
    a.nB = static_cast<uint32_t>(rm::n_bias(K));
    a.lr = lr;
    a.b1 = static_cast<float>(b1);
    a.b2 = static_cast<float>(b2);
    a.eps = 1e-8f;
    a.b1d = b1;
    a.b2d = b2;
    a.stepHost = m.adam_step;
    a.stepDev = m.adam_step_device;
    if (m.adam_step_device)
        RK_LAUNCH("k_adam_count", rm::k_adam_count, dim3(1), dim3(64), 0, s, m.adam_step_device);
    RK_LAUNCH("k_adam_guarded", rm::k_adam_guarded, dim3((a.nW + a.nB + 255) / 256), dim3(256), 0, s, a, nValid);
    return hipGetLastError();
}

hipError_t mlp_propose(const restir_mlp &m, const GView &g, f3 cam, const float lo[3], const float hi[3],
                       const SceneView &sc, const int32_t *lightCluster, const float *lightPWithin,
                       const int32_t *clusterBegin, const int32_t *clusterCount, const int32_t *clusterLights,
                       const restir_alias_table_column *clusterAlias, const float4 *clusterFused, uint32_t M,
                       float mix, uint32_t seed, int32_t *candLight, float *candQ, hipStream_t s) {
    rm::ProposeArgs a{};
    a.f = rm::FeatArgs{g, cam, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
    a.wbf = m.weights_bf16;
    a.bias = m.bias;
    a.K = static_cast<int>(m.n_clusters);
    a.npx = g.W * g.H;
    a.M = M;
    a.seed = seed;
    a.mix = mix;
    a.alias = sc.alias;
    a.nAlias = sc.nAlias;
    a.lightCluster = lightCluster;
    a.lightPWithin = lightPWithin;
    a.clusterBegin = clusterBegin;
    a.clusterCount = clusterCount;
    a.clusterLights = clusterLights;
    a.clusterAlias = reinterpret_cast<const float4 *>(clusterAlias);
    a.clusterFused = clusterFused;
    a.candLight = candLight;
    a.candQ = candQ;
    const uint32_t quads = ((g.W + 1) / 2) * ((g.H + 1) / 2);
    const dim3 grid((quads + rm::kProposeTile - 1) / rm::kProposeTile), block(2 * rm::kProposeTile);
    const dim3 gridW((quads + rm::kTilePx - 1) / rm::kTilePx), blockW(2 * rm::kTilePx);
    switch (a.K) {
        case 64: RK_LAUNCH("k_propose", (rm::k_propose<64>), grid, block, 0, s, a); break;
        case 128: RK_LAUNCH("k_propose", (rm::k_propose<128>), grid, block, 0, s, a); break;
        case 256: RK_LAUNCH("k_propose", (rm::k_propose<256>), grid, block, 0, s, a); break;
        default: RK_LAUNCH("k_propose_wide", rm::k_propose_wide, gridW, blockW, 0, s, a); break;
    }
    return hipGetLastError();
}

}  // namespace rk
