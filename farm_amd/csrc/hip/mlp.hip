// The learned part: pair features and an 80-128-128-2 MLP (intersect logit,
// log refinement cost) trained and evaluated on the GPU.
//
// GEMMs run on MFMA (v_mfma_f32_16x16x16_bf16): operands are rounded to
// bf16 when staged into LDS (round to nearest even), products accumulate in
// fp32.  One generic kernel covers the forward products, the transposed
// backward products and the batch reductions of the weight gradients
// (split-K with atomic accumulation).  The reference has no learned
// component; the features come from the same raster windows as the filter
// (Ideal::intersect, Ideal.cpp:1329-1357).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "handle.h"

namespace farm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
    return (short)(u >> 16);
}

// ------------------------------------------------------------------------
// Pair features (one lane per candidate pair).  decided: 1 = an IN pixel
// overlap (the exact answer is true), 0 = no BORDER x BORDER pixel pair and
// no IN overlap (the exact answer is false), -1 = needs refinement.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_features(DevSet S, const uint32_t *__restrict__ pa_ids,
                                                   const uint32_t *__restrict__ pb_ids, uint32_t n,
                                                   float *__restrict__ feat, int8_t *__restrict__ decided) {
    __shared__ uint32_t hist_all[128][64];  // 32 KB: per-lane 32-bit bins (no wrap on huge pairs)
    uint32_t *hist = hist_all[threadIdx.x];
    for (int i = 0; i < 64; i++) hist[i] = 0;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const farm_raster_hdr A = S.hdr[pa_ids[k]];
    const farm_raster_hdr B = S.hdr[pb_ids[k]];
    const uint8_t *sa = S.status + A.pix_start;
    const uint8_t *sb = S.status + B.pix_start;
    const uint32_t *ea = S.pixedges + A.pix_start;
    const uint32_t *eb = S.pixedges + B.pix_start;
    const Window wa = window_of(A, B.mbr[0], B.mbr[1], B.mbr[2], B.mbr[3]);
    const double area_a = A.step[0] * A.step[1], area_b = B.step[0] * B.step[1];
    const double amax = fmax(area_a, area_b);
    uint32_t n_in = 0, n_bb = 0, wa_border = 0, wa_in = 0, wa_total = 0;
    double sum_cost = 0, sum_ea = 0, sum_eb = 0, max_prob = 0, sum_prob = 0;
    for (int x = wa.x0; x <= wa.x1; x++) {
        for (int y = wa.y0; y <= wa.y1; y++) {
            const int pa = y * A.dimx + x;
            const uint8_t fa = sa[pa];
            const int st_a = show(fa);
            wa_total++;
            if (st_a == 0) continue;
            if (st_a == 1) wa_border++;
            else wa_in++;
            double lx, ly, hx, hy;
            pixel_box(A, x, y, lx, ly, hx, hy);
            const Window wb = window_of(B, lx, ly, hx, hy);
            for (int u = wb.x0; u <= wb.x1; u++)
                for (int v = wb.y0; v <= wb.y1; v++) {
                    const int pb = v * B.dimx + u;
                    const uint8_t fb = sb[pb];
                    const int st_b = show(fb);
                    if (st_b == 0) continue;
                    if (st_a == 2 || st_b == 2) {
                        n_in++;
                        continue;
                    }
                    n_bb++;
                    const double xa = apx_area_cc(S.fill[A.pix_start + pa], area_a, S.cc);
                    const double xb = apx_area_cc(S.fill[B.pix_start + pb], area_b, S.cc);
                    const double prob = (xa + xb) / amax;
                    max_prob = fmax(max_prob, prob);
                    sum_prob += prob;
                    const double p = (xa / area_a) * (xb / area_b);
                    hist[min(63, (int)(p * 64.0))]++;
                    sum_cost += (double)ea[pa] * (double)eb[pb];
                    sum_ea += ea[pa];
                    sum_eb += eb[pb];
                }
        }
    }
    float *f = feat + (size_t)k * kNFeat;
    const float inv = n_bb ? 1.0f / n_bb : 0.f;
    for (int i = 0; i < 64; i++) f[i] = hist[i] * inv;
    const double ix = fmax(0.0, fmin(A.mbr[2], B.mbr[2]) - fmax(A.mbr[0], B.mbr[0]));
    const double iy = fmax(0.0, fmin(A.mbr[3], B.mbr[3]) - fmax(A.mbr[1], B.mbr[1]));
    const double ma = (A.mbr[2] - A.mbr[0]) * (A.mbr[3] - A.mbr[1]);
    const double mb = (B.mbr[2] - B.mbr[0]) * (B.mbr[3] - B.mbr[1]);
    f[64] = log1pf((float)n_in);
    f[65] = log1pf((float)n_bb);
    f[66] = log1pf((float)A.nverts);
    f[67] = log1pf((float)B.nverts);
    f[68] = (float)(ix * iy / fmax(fmin(ma, mb), 1e-30));
    f[69] = (float)log1p(sum_cost);
    f[70] = (float)max_prob;
    f[71] = n_bb ? (float)(sum_prob / n_bb) : 0.f;
    f[72] = (float)log1p(sum_ea);
    f[73] = (float)log1p(sum_eb);
    f[74] = (float)log(area_a / area_b);
    f[75] = wa_total ? (float)wa_border / wa_total : 0.f;
    f[76] = wa_total ? (float)wa_in / wa_total : 0.f;
    f[77] = f[78] = f[79] = 0.f;
    decided[k] = n_in ? 1 : (n_bb ? -1 : 0);
}

// x[i, :] = (x[i, :] - mean) * inv_std
__global__ __launch_bounds__(256) void k_normalize(float *__restrict__ x, uint64_t n, const float *__restrict__ norm) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kNFeat) return;
    const int c = (int)(i % kNFeat);
    x[i] = (x[i] - norm[c]) * norm[kNFeat + c];
}

// ------------------------------------------------------------------------
// C (M x N, row stride ldc) = op(A) op(B) [+ bias] [relu], or accumulated
// atomically (split-K over blockIdx.z).  A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn].  Block tile 64 x 64, 4 waves x (16 rows x 4
// MFMA tiles of 16 x 16), K step 16.
// MFMA 16x16x16 operand layout (lane l): A row l%16, k 4*(l/16)+j;
// B col l%16, k 4*(l/16)+j; D row 4*(l/16)+j, col l%16.
// ------------------------------------------------------------------------
enum { GEMM_BIAS = 1, GEMM_RELU = 2, GEMM_ATOMIC = 4 };

__global__ __launch_bounds__(256) void k_gemm(const float *__restrict__ A, int64_t sam, int64_t sak,
                                               const float *__restrict__ B, int64_t sbk, int64_t sbn,
                                               float *__restrict__ C, int64_t ldc, int M, int N, int K, int kchunk,
                                               const float *__restrict__ bias, int flags) {
    __shared__ short As[64][20];
    __shared__ short Bs[16][68];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    f32x4 acc[4];
    for (int i = 0; i < 4; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        for (int i = 0; i < 4; i++) {
            const int idx = t + 256 * i;
            const int r = idx >> 4, c = idx & 15;
            const int gm = m0 + r, gk = k0 + c;
            As[r][c] = (gm < M && gk < ke) ? f2bf(A[(int64_t)gm * sam + (int64_t)gk * sak]) : (short)0;
            const int rb = idx >> 6, cb = idx & 63;
            const int bk = k0 + rb, bn = n0 + cb;
            Bs[rb][cb] = (bk < ke && bn < N) ? f2bf(B[(int64_t)bk * sbk + (int64_t)bn * sbn]) : (short)0;
        }
        __syncthreads();
        bf16x4 a;
        const int ar = 16 * w + (l & 15), ak = 4 * (l >> 4);
        for (int j = 0; j < 4; j++) a[j] = As[ar][ak + j];
        for (int tt = 0; tt < 4; tt++) {
            bf16x4 b;
            for (int j = 0; j < 4; j++) b[j] = Bs[ak + j][16 * tt + (l & 15)];
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tt], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int tt = 0; tt < 4; tt++) {
        for (int j = 0; j < 4; j++) {
            const int row = m0 + 16 * w + 4 * (l >> 4) + j;
            const int col = n0 + 16 * tt + (l & 15);
            if (row >= M || col >= N) continue;
            float v = acc[tt][j];
            float *dst = C + (int64_t)row * ldc + col;
            if (flags & GEMM_ATOMIC) {
                atomicAdd(dst, v);
            } else {
                if (flags & GEMM_BIAS) v += bias[col];
                if (flags & GEMM_RELU) v = fmaxf(v, 0.f);
                *dst = v;
            }
        }
    }
}

// ------------------------------------------------------------------------
// Inference in one pass: normalisation and the three layers of 64 pairs at a
// time, the weights staged in LDS once per workgroup, the activations never
// leaving it.  The operands are rounded to bf16 exactly where k_gemm rounds
// them and every output takes the same MFMA sequence over K (16 at a time,
// ascending), so the logits equal the three-launch path's bit for bit.
// Wave w owns rows 16w .. 16w + 15 of the tile.
// ------------------------------------------------------------------------
constexpr int kInferRows = 64;

template <int K, bool RELU>
__device__ __forceinline__ void infer_layer(const short (*A)[K], const short (*B)[kHidden], const float *bias,
                                            short (*H)[kHidden], int w, int l) {
    f32x4 acc[kHidden / 16];
#pragma unroll
    for (int tt = 0; tt < kHidden / 16; tt++) acc[tt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int ar = 16 * w + (l & 15), ak = 4 * (l >> 4);
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a;
#pragma unroll
        for (int j = 0; j < 4; j++) a[j] = A[ar][k0 + ak + j];
#pragma unroll
        for (int tt = 0; tt < kHidden / 16; tt++) {
            bf16x4 b;
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = B[k0 + ak + j][16 * tt + (l & 15)];
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int tt = 0; tt < kHidden / 16; tt++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int col = 16 * tt + (l & 15);
            float v = acc[tt][j] + bias[col];
            if (RELU) v = fmaxf(v, 0.f);
            H[16 * w + 4 * (l >> 4) + j][col] = f2bf(v);
        }
}

__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ feat, const float *__restrict__ norm,
                                                    const float *__restrict__ W, uint64_t n,
                                                    float *__restrict__ pred) {
    __shared__ short sW1[kNFeat][kHidden];
    __shared__ short sW2[kHidden][kHidden];
    __shared__ short sW3[kHidden][16];  // columns >= kOut are zero
    __shared__ float sb[2 * kHidden + kOut];
    __shared__ short sX[kInferRows][kNFeat];
    __shared__ short sH1[kInferRows][kHidden];
    __shared__ short sH2[kInferRows][kHidden];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    for (int i = t; i < kNFeat * kHidden; i += 256) sW1[i / kHidden][i % kHidden] = f2bf(W[kOffW1 + i]);
    for (int i = t; i < kHidden * kHidden; i += 256) sW2[i / kHidden][i % kHidden] = f2bf(W[kOffW2 + i]);
    for (int i = t; i < kHidden * 16; i += 256) {
        const int k = i / 16, c = i % 16;
        sW3[k][c] = c < kOut ? f2bf(W[kOffW3 + k * kOut + c]) : (short)0;
    }
    for (int i = t; i < kHidden; i += 256) {
        sb[i] = W[kOffB1 + i];
        sb[kHidden + i] = W[kOffB2 + i];
    }
    if (t < kOut) sb[2 * kHidden + t] = W[kOffB3 + t];
    __syncthreads();
    for (uint64_t r0 = (uint64_t)blockIdx.x * kInferRows; r0 < n; r0 += (uint64_t)gridDim.x * kInferRows) {
        for (int i = t; i < kInferRows * kNFeat; i += 256) {
            const int r = i / kNFeat, c = i % kNFeat;
            float v = 0.f;
            if (r0 + r < n) v = (feat[(r0 + r) * kNFeat + c] - norm[c]) * norm[kNFeat + c];
            sX[r][c] = f2bf(v);
        }
        __syncthreads();
        infer_layer<kNFeat, true>(sX, sW1, sb, sH1, w, l);
        __syncthreads();
        infer_layer<kHidden, true>(sH1, sW2, sb + kHidden, sH2, w, l);
        __syncthreads();
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        const int ar = 16 * w + (l & 15), ak = 4 * (l >> 4);
#pragma unroll
        for (int k0 = 0; k0 < kHidden; k0 += 16) {
            bf16x4 a, b;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                a[j] = sH2[ar][k0 + ak + j];
                b[j] = sW3[k0 + ak + j][l & 15];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        const int col = l & 15;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t row = r0 + 16 * w + 4 * (l >> 4) + j;
            if (col < kOut && row < n) pred[2 * row + col] = acc[j] + sb[2 * kHidden + col];
        }
        __syncthreads();
    }
}

// loss = mean(BCE-with-logits(z0, y)) + lambda * mean((z1 - c)^2), and dL/dz
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ z, const float *__restrict__ y,
                                               const float *__restrict__ c, int M, float lambda,
                                               float *__restrict__ dz, float *__restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float L = 0.f;
    if (i < M) {
        const float z0 = z[2 * i], z1 = z[2 * i + 1];
        const float s = 1.f / (1.f + expf(-z0));
        dz[2 * i] = (s - y[i]) / M;
        dz[2 * i + 1] = 2.f * lambda * (z1 - c[i]) / M;
        L = (fmaxf(z0, 0.f) - z0 * y[i] + log1pf(expf(-fabsf(z0)))) / M + lambda * (z1 - c[i]) * (z1 - c[i]) / M;
    }
    for (int o = 32; o > 0; o >>= 1) L += __shfl_xor(L, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, L);
}

// d[i] *= (h[i] > 0)   (ReLU backward through the stored activation)
__global__ __launch_bounds__(256) void k_relu_back(float *__restrict__ d, const float *__restrict__ h, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(h[i] > 0.f)) d[i] = 0.f;
}

// out[c] += sum_r d[r * N + c]
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ d, int M, int N, float *__restrict__ out) {
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < M; r += blockDim.x) s += d[(int64_t)r * N + c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&out[c], part[0] + part[1] + part[2] + part[3]);
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                                               float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                                               float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_gather_rows(const float *__restrict__ src, const uint32_t *__restrict__ idx,
                                                      int rows, int cols, float *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    dst[i] = src[(int64_t)idx[r] * cols + c];
}

__global__ __launch_bounds__(256) void k_gather_vec(const float *__restrict__ src, const uint32_t *__restrict__ idx,
                                                     int rows, float *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) dst[i] = src[idx[i]];
}

}  // namespace farm

using namespace farm;

namespace {

int gemm(farm_handle h, const float *A, int64_t sam, int64_t sak, const float *B, int64_t sbk, int64_t sbn, float *C,
         int64_t ldc, int M, int N, int K, const float *bias, int flags, int ksplit = 1) {
    if (M <= 0 || N <= 0 || K <= 0) return FARM_OK;
    int kchunk = K;
    if (ksplit > 1) kchunk = ((K + ksplit - 1) / ksplit + 15) / 16 * 16;
    const int kz = (K + kchunk - 1) / kchunk;
    dim3 grid((M + 63) / 64, (N + 63) / 64, kz);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, h->stream, A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, kchunk, bias,
                       flags);
    HIPCHK(h, hipGetLastError());
    return FARM_OK;
}

int mlp_buffers(farm_handle h, int M) {
    int rc;
    if ((rc = ensure(h, h->mlp_x, 4ull * M * kNFeat))) return rc;
    if ((rc = ensure(h, h->mlp_h1, 4ull * M * kHidden))) return rc;
    if ((rc = ensure(h, h->mlp_h2, 4ull * M * kHidden))) return rc;
    if ((rc = ensure(h, h->mlp_z3, 4ull * M * kOut))) return rc;
    if ((rc = ensure(h, h->mlp_d3, 4ull * M * kOut))) return rc;
    if ((rc = ensure(h, h->mlp_dh2, 4ull * M * kHidden))) return rc;
    if ((rc = ensure(h, h->mlp_dh1, 4ull * M * kHidden))) return rc;
    if ((rc = ensure(h, h->mlp_y, 4ull * M))) return rc;
    if ((rc = ensure(h, h->mlp_c, 4ull * M))) return rc;
    if ((rc = ensure(h, h->mlp_loss, 64))) return rc;
    return FARM_OK;
}

// forward of M rows of X (device, normalised): h1, h2, z3 in the handle
int mlp_forward_dev(farm_handle h, const float *X, int M) {
    int rc;
    const float *W = P<float>(h->mlp_w);
    if ((rc = gemm(h, X, kNFeat, 1, W + kOffW1, kHidden, 1, P<float>(h->mlp_h1), kHidden, M, kHidden, kNFeat,
                   W + kOffB1, GEMM_BIAS | GEMM_RELU)))
        return rc;
    if ((rc = gemm(h, P<float>(h->mlp_h1), kHidden, 1, W + kOffW2, kHidden, 1, P<float>(h->mlp_h2), kHidden, M,
                   kHidden, kHidden, W + kOffB2, GEMM_BIAS | GEMM_RELU)))
        return rc;
    return gemm(h, P<float>(h->mlp_h2), kHidden, 1, W + kOffW3, kOut, 1, P<float>(h->mlp_z3), kOut, M, kOut, kHidden,
                W + kOffB3, GEMM_BIAS);
}

// loss and gradients (into mlp_g, zeroed here) of M rows; y, c device
int mlp_backward_dev(farm_handle h, const float *X, const float *y, const float *c, int M, float lambda) {
    int rc;
    const float *W = P<float>(h->mlp_w);
    float *G = P<float>(h->mlp_g);
    if ((rc = fill_u32(h, h->mlp_g.p, kNParams, 0))) return rc;
    if ((rc = fill_u32(h, h->mlp_loss.p, 1, 0))) return rc;
    hipLaunchKernelGGL(k_loss, dim3(blocks_for(M)), dim3(256), 0, h->stream, (const float *)P<float>(h->mlp_z3), y, c,
                       M, lambda, P<float>(h->mlp_d3), P<float>(h->mlp_loss));
    HIPCHK(h, hipGetLastError());
    const int ks = max(1, M / 1024);
    float *d3 = P<float>(h->mlp_d3), *h2 = P<float>(h->mlp_h2), *h1 = P<float>(h->mlp_h1);
    float *dh2 = P<float>(h->mlp_dh2), *dh1 = P<float>(h->mlp_dh1);
    // layer 3: dW3 = H2^T dZ3, db3 = colsum dZ3, dH2 = dZ3 W3^T
    if ((rc = gemm(h, h2, 1, kHidden, d3, kOut, 1, G + kOffW3, kOut, kHidden, kOut, M, nullptr, GEMM_ATOMIC, ks)))
        return rc;
    hipLaunchKernelGGL(k_colsum, dim3(kOut), dim3(256), 0, h->stream, (const float *)d3, M, kOut, G + kOffB3);
    HIPCHK(h, hipGetLastError());
    if ((rc = gemm(h, d3, kOut, 1, W + kOffW3, 1, kOut, dh2, kHidden, M, kHidden, kOut, nullptr, 0))) return rc;
    hipLaunchKernelGGL(k_relu_back, dim3(blocks_for((uint64_t)M * kHidden)), dim3(256), 0, h->stream, dh2,
                       (const float *)h2, (uint64_t)M * kHidden);
    HIPCHK(h, hipGetLastError());
    // layer 2
    if ((rc = gemm(h, h1, 1, kHidden, dh2, kHidden, 1, G + kOffW2, kHidden, kHidden, kHidden, M, nullptr,
                   GEMM_ATOMIC, ks)))
        return rc;
    hipLaunchKernelGGL(k_colsum, dim3(kHidden), dim3(256), 0, h->stream, (const float *)dh2, M, kHidden, G + kOffB2);
    HIPCHK(h, hipGetLastError());
    if ((rc = gemm(h, dh2, kHidden, 1, W + kOffW2, 1, kHidden, dh1, kHidden, M, kHidden, kHidden, nullptr, 0)))
        return rc;
    hipLaunchKernelGGL(k_relu_back, dim3(blocks_for((uint64_t)M * kHidden)), dim3(256), 0, h->stream, dh1,
                       (const float *)h1, (uint64_t)M * kHidden);
    HIPCHK(h, hipGetLastError());
    // layer 1
    if ((rc = gemm(h, X, 1, kNFeat, dh1, kHidden, 1, G + kOffW1, kHidden, kNFeat, kHidden, M, nullptr, GEMM_ATOMIC,
                   ks)))
        return rc;
    hipLaunchKernelGGL(k_colsum, dim3(kHidden), dim3(256), 0, h->stream, (const float *)dh1, M, kHidden, G + kOffB1);
    HIPCHK(h, hipGetLastError());
    return FARM_OK;
}

int require_mlp(farm_handle h) {
    if (!h) return FARM_EINVAL;
    if (!h->mlp_ready) return fail(h, FARM_ESTATE, "MLP parameters not set (farm_mlp_set_params)");
    return FARM_OK;
}

}  // namespace

extern "C" {

int farm_mlp_set_params(farm_handle h, const float *params, uint64_t n, const float *norm) {
    if (!h || !params || n != (uint64_t)kNParams)
        return fail(h, FARM_EINVAL, "expected %d parameters", kNParams);
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = upload(h, h->mlp_w, params, 4ull * kNParams))) return rc;
    if ((rc = ensure(h, h->mlp_m, 4ull * kNParams))) return rc;
    if ((rc = ensure(h, h->mlp_v, 4ull * kNParams))) return rc;
    if ((rc = ensure(h, h->mlp_g, 4ull * kNParams))) return rc;
    if ((rc = fill_u32(h, h->mlp_m.p, kNParams, 0))) return rc;
    if ((rc = fill_u32(h, h->mlp_v.p, kNParams, 0))) return rc;
    std::vector<float> nm(2 * kNFeat);
    for (int i = 0; i < kNFeat; i++) {
        nm[i] = norm ? norm[i] : 0.f;
        nm[kNFeat + i] = norm ? norm[kNFeat + i] : 1.f;
    }
    if ((rc = upload(h, h->mlp_norm, nm.data(), 4ull * nm.size()))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->adam_t = 0;
    h->mlp_ready = true;
    return FARM_OK;
}

int farm_mlp_get_params(farm_handle h, float *params, uint64_t n) {
    int rc = require_mlp(h);
    if (rc) return rc;
    if (n != (uint64_t)kNParams) return fail(h, FARM_EINVAL, "expected %d parameters", kNParams);
    HIPCHK(h, hipMemcpy(params, h->mlp_w.p, 4ull * kNParams, hipMemcpyDeviceToHost));
    return FARM_OK;
}

int farm_mlp_forward(farm_handle h, const float *X, uint64_t n, float *out) {
    int rc = require_mlp(h);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->device));
    const int chunk = 1 << 16;
    for (uint64_t s = 0; s < n; s += chunk) {
        const int M = (int)std::min<uint64_t>(chunk, n - s);
        if ((rc = mlp_buffers(h, M))) return rc;
        if ((rc = upload(h, h->mlp_x, X + s * kNFeat, 4ull * M * kNFeat))) return rc;
        if ((rc = mlp_forward_dev(h, P<float>(h->mlp_x), M))) return rc;
        HIPCHK(h, hipMemcpyAsync(out + s * kOut, h->mlp_z3.p, 4ull * M * kOut, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return FARM_OK;
}

int farm_mlp_loss_grad(farm_handle h, const float *X, const float *y, const float *c, uint64_t n, float lambda,
                       float *loss, float *grad) {
    int rc = require_mlp(h);
    if (rc) return rc;
    if (n == 0 || n > (1u << 24)) return fail(h, FARM_EINVAL, "batch of %llu rows", (unsigned long long)n);
    HIPCHK(h, hipSetDevice(h->device));
    const int M = (int)n;
    if ((rc = mlp_buffers(h, M))) return rc;
    if ((rc = upload(h, h->mlp_x, X, 4ull * M * kNFeat))) return rc;
    if ((rc = upload(h, h->mlp_y, y, 4ull * M))) return rc;
    if ((rc = upload(h, h->mlp_c, c, 4ull * M))) return rc;
    if ((rc = mlp_forward_dev(h, P<float>(h->mlp_x), M))) return rc;
    if ((rc = mlp_backward_dev(h, P<float>(h->mlp_x), P<float>(h->mlp_y), P<float>(h->mlp_c), M, lambda))) return rc;
    if (loss) HIPCHK(h, hipMemcpyAsync(loss, h->mlp_loss.p, 4, hipMemcpyDeviceToHost, h->stream));
    if (grad) HIPCHK(h, hipMemcpyAsync(grad, h->mlp_g.p, 4ull * kNParams, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}

int farm_mlp_train(farm_handle h, const float *X, const float *y, const float *c, uint64_t n, int epochs, int batch,
                   float lr, float lambda, uint64_t seed, float *epoch_loss) {
    int rc = require_mlp(h);
    if (rc) return rc;
    if (n == 0 || batch <= 0 || epochs <= 0) return fail(h, FARM_EINVAL, "bad training arguments");
    HIPCHK(h, hipSetDevice(h->device));
    DevBuf dX, dy, dc, dperm, bx;
    auto cleanup = [&]() {
        free_buf(dX);
        free_buf(dy);
        free_buf(dc);
        free_buf(dperm);
        free_buf(bx);
    };
    if ((rc = upload(h, dX, X, 4ull * n * kNFeat)) || (rc = upload(h, dy, y, 4ull * n)) ||
        (rc = upload(h, dc, c, 4ull * n)) || (rc = ensure(h, dperm, 4ull * n)) ||
        (rc = ensure(h, bx, 4ull * batch * kNFeat)) || (rc = mlp_buffers(h, batch))) {
        cleanup();
        return rc;
    }
    std::vector<uint32_t> perm(n);
    uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
    for (int ep = 0; ep < epochs; ep++) {
        for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
        for (uint64_t i = n - 1; i > 0; i--) {  // Fisher-Yates, PCG-style LCG
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            const uint64_t j = (s >> 33) % (i + 1);
            std::swap(perm[i], perm[j]);
        }
        if ((rc = upload(h, dperm, perm.data(), 4ull * n))) {
            cleanup();
            return rc;
        }
        double ep_loss = 0;
        int nb = 0;
        for (uint64_t b0 = 0; b0 + batch <= n || (b0 < n && nb == 0); b0 += batch) {
            const int M = (int)std::min<uint64_t>(batch, n - b0);
            const uint32_t *ix = P<uint32_t>(dperm) + b0;
            hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for((uint64_t)M * kNFeat)), dim3(256), 0, h->stream,
                               (const float *)P<float>(dX), ix, M, kNFeat, P<float>(bx));
            hipLaunchKernelGGL(k_gather_vec, dim3(blocks_for(M)), dim3(256), 0, h->stream, (const float *)P<float>(dy),
                               ix, M, P<float>(h->mlp_y));
            hipLaunchKernelGGL(k_gather_vec, dim3(blocks_for(M)), dim3(256), 0, h->stream, (const float *)P<float>(dc),
                               ix, M, P<float>(h->mlp_c));
            if ((rc = mlp_forward_dev(h, P<float>(bx), M)) ||
                (rc = mlp_backward_dev(h, P<float>(bx), P<float>(h->mlp_y), P<float>(h->mlp_c), M, lambda))) {
                cleanup();
                return rc;
            }
            h->adam_t++;
            const float c1 = 1.f - powf(0.9f, (float)h->adam_t), c2 = 1.f - powf(0.999f, (float)h->adam_t);
            hipLaunchKernelGGL(k_adam, dim3(blocks_for(kNParams)), dim3(256), 0, h->stream, P<float>(h->mlp_w),
                               (const float *)P<float>(h->mlp_g), P<float>(h->mlp_m), P<float>(h->mlp_v), kNParams,
                               lr, 0.9f, 0.999f, 1e-8f, c1, c2);
            float L = 0;
            hipError_t e = hipMemcpyAsync(&L, h->mlp_loss.p, 4, hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) {
                cleanup();
                return fail(h, FARM_EDEVICE, "training step failed: %s", hipGetErrorString(e));
            }
            ep_loss += L;
            nb++;
        }
        if (epoch_loss) epoch_loss[ep] = (float)(ep_loss / std::max(nb, 1));
    }
    cleanup();
    return FARM_OK;
}

// features of pairs [begin, end) on the device (normalised copy in h->feat
// when the MLP norm is set); optional host copies of raw features / flags
int farm_pair_features(farm_handle h, uint64_t begin, uint64_t end, float *feat_out, int8_t *decided_out) {
    if (!h) return FARM_EINVAL;
    if (h->npolys <= 0) return fail(h, FARM_ESTATE, "no polygons uploaded");
    if (begin > end || end > h->npairs) return fail(h, FARM_EINVAL, "bad pair window");
    const uint64_t n = end - begin;
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = ensure(h, h->feat, 4ull * (n ? n : 1) * kNFeat))) return rc;
    if ((rc = ensure(h, h->decided, n ? n : 1))) return rc;
    if (n) {
        hipLaunchKernelGGL(k_features, dim3(blocks_for(n, 128)), dim3(128), 0, h->stream, devset(h),
                           (const uint32_t *)P<uint32_t>(h->pa) + begin, (const uint32_t *)P<uint32_t>(h->pb) + begin,
                           (uint32_t)n, P<float>(h->feat), P<int8_t>(h->decided));
        HIPCHK(h, hipGetLastError());
        if (feat_out)
            HIPCHK(h, hipMemcpyAsync(feat_out, h->feat.p, 4ull * n * kNFeat, hipMemcpyDeviceToHost, h->stream));
        if (decided_out) HIPCHK(h, hipMemcpyAsync(decided_out, h->decided.p, n, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}

// MLP predictions (logit, log-cost) of pairs [begin, end): features, norm,
// forward on the device; results stay in h->pred (2 floats per pair) and are
// copied to `out` when given
int farm_mlp_predict_pairs(farm_handle h, uint64_t begin, uint64_t end, float *out) {
    int rc = require_mlp(h);
    if (rc) return rc;
    if ((rc = farm_pair_features(h, begin, end, nullptr, nullptr))) return rc;
    const uint64_t n = end - begin;
    if ((rc = ensure(h, h->pred, 8ull * (n ? n : 1)))) return rc;
    if (!n) return FARM_OK;
    // normalisation and the three layers in one pass (the logits of the
    // three-launch path, bit for bit)
    const unsigned grid = (unsigned)std::min<uint64_t>((n + kInferRows - 1) / kInferRows, 2ull * h->num_cus);
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(256), 0, h->stream, (const float *)P<float>(h->feat),
                       (const float *)P<float>(h->mlp_norm), (const float *)P<float>(h->mlp_w), n, P<float>(h->pred));
    HIPCHK(h, hipGetLastError());
    if (out) HIPCHK(h, hipMemcpyAsync(out, h->pred.p, 8ull * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FARM_OK;
}

}  // extern "C"
