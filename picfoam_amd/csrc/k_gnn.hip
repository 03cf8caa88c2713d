// Learned initial guess for the Poisson PCG: a message-passing network over the cell
// graph (cells coupled by the Laplacian), hidden width 64, BF16 weights and activations,
// FP32 accumulation on the MFMA units (v_mfma_f32_16x16x16_bf16), trained on the device
// with its own backward kernels and Adam.
//
//   x_i   = [r_i / s_r, b_i / s_b, phi_prev_i / s_phi]     (r = b - A phi_prev)
//   h0_i  = relu(W_enc^T x_i + b_enc)
//   layer : m_i = mean_{j in N(i)} h_j ;  z_i = W^T [h_i ; m_i] + b ;  h_i += relu(z_i)
//   out_i = w_dec . h_i + b_dec
//   guess = phi_prev + out * s_r / dbar                     (dbar = mean diagonal of A)
//   loss  = e^T A e / (d^T A d),  e = out - d,  d = (phi* - phi_prev) dbar / s_r
// The loss is the A-norm error of the guess relative to the warm start (1 = no gain);
// CG minimises the same norm.
#include <cmath>
#include <vector>

#include "device.h"
#include "gnn.h"

namespace picmi {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ double block_sum_d(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- features / target
__global__ void k_sumsq3(int n, const double* a, const double* b, const double* c, double* partial) {
    __shared__ double sh[4];
    double s0 = 0, s1 = 0, s2 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        s0 += a[i] * a[i]; s1 += b[i] * b[i]; s2 += c[i] * c[i];
    }
    s0 = block_sum_d(s0, sh); s1 = block_sum_d(s1, sh); s2 = block_sum_d(s2, sh);
    if (threadIdx.x == 0) { partial[3 * blockIdx.x] = s0; partial[3 * blockIdx.x + 1] = s1; partial[3 * blockIdx.x + 2] = s2; }
}

__global__ void k_reduce_n(const double* partial, int nb, int nv, double* out) {
    const int k = threadIdx.x;
    if (k >= nv) return;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += partial[(size_t)b * nv + k];
    out[k] = acc;
}

// scal: [0] sum r^2, [1] sum b^2, [2] sum phi^2 -> x (n x 3 f32)
__global__ void k_features(int n, const double* r, const double* b, const double* phi, const double* scal, float* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sr = sqrt(scal[0] / n), sb = sqrt(scal[1] / n), sp = sqrt(scal[2] / n);
    x[3 * i] = (float)(sr > 0 ? r[i] / sr : 0.0);
    x[3 * i + 1] = (float)(sb > 0 ? b[i] / sb : 0.0);
    x[3 * i + 2] = (float)(sp > 0 ? phi[i] / sp : 0.0);
}

// target d = (phi* - phi_prev) dbar / s_r
__global__ void k_target(int n, const double* phi_star, const double* phi_prev, const double* scal, double dbar,
                         float* t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sr = sqrt(scal[0] / n);
    t[i] = (float)(sr > 0 ? (phi_star[i] - phi_prev[i]) * dbar / sr : 0.0);
}

// guess = phi_prev + out * s_r / dbar
__global__ void k_guess(int n, const float* out, const double* phi_prev, const double* scal, double dbar,
                        double* guess) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sr = sqrt(scal[0] / n);
    guess[i] = phi_prev[i] + (double)out[i] * sr / dbar;
}

// ---------------------------------------------------------------- forward
__global__ void k_encode(int n, const float* __restrict__ x, const float* __restrict__ W /*[3][64]*/,
                         const float* __restrict__ b, uint16_t* __restrict__ z, uint16_t* __restrict__ h) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * 64) return;
    const int i = (int)(t >> 6), c = (int)(t & 63);
    const float v = b[c] + x[3 * i] * W[c] + x[3 * i + 1] * W[64 + c] + x[3 * i + 2] * W[128 + c];
    if (z) z[t] = f2bf(v);
    h[t] = f2bf(fmaxf(v, 0.f));
}

// m_i = mean_{j in N(i)} h_j  (one thread per (node, 4 channels))
__global__ void k_mean(int n, int K, const int32_t* __restrict__ nb, const uint16_t* __restrict__ h,
                       uint16_t* __restrict__ m) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * 16) return;
    const int i = (int)(t >> 4), c0 = (int)(t & 15) * 4;
    float a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int deg = 0;
    for (int k = 0; k < K; ++k) {
        const int j = nb[(int64_t)i * K + k];
        if (j < 0) continue;
        const uint64_t v = *(const uint64_t*)(h + (int64_t)j * 64 + c0);
        a0 += bf2f((uint16_t)v); a1 += bf2f((uint16_t)(v >> 16)); a2 += bf2f((uint16_t)(v >> 32)); a3 += bf2f((uint16_t)(v >> 48));
        ++deg;
    }
    const float s = deg ? 1.f / deg : 0.f;
    const uint64_t o = (uint64_t)f2bf(a0 * s) | ((uint64_t)f2bf(a1 * s) << 16) | ((uint64_t)f2bf(a2 * s) << 32) |
                       ((uint64_t)f2bf(a3 * s) << 48);
    *(uint64_t*)(m + (int64_t)i * 64 + c0) = o;
}

// z = [h | m] W + b ; h_out = h + relu(z).  MFMA 16x16x16 BF16 on tiles of 16 nodes.
// Fragment layout (v_mfma_f32_16x16x16_bf16): lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and D[4(l/16)+r][l%16], r = 0..3.
// Persistent waves: a wave loads its 32 B fragments of W (8 k-blocks x 4 column tiles) once
// and keeps them in registers for every tile it takes, instead of reading W (16 KiB) again
// per tile.  The tile's pre-activations z (FP32) go through LDS, so that h, z and h_out
// move as 16-byte pieces, 16 consecutive channels of one node per lane.  Every block runs
// the same number of tile rounds, so its barriers are reached by all of its waves.
constexpr int kLayerBlocks = 1024;
__global__ void __launch_bounds__(256) k_layer_fwd(int n, const uint16_t* __restrict__ h, const uint16_t* __restrict__ m,
                                                   const uint16_t* __restrict__ WT /*[64][128]*/,
                                                   const float* __restrict__ b, uint16_t* __restrict__ z,
                                                   uint16_t* __restrict__ hout) {
    __shared__ float tile[4][16][65];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    s4 bv[8][4];
#pragma unroll
    for (int kb = 0; kb < 8; ++kb)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) bv[kb][ct] = *(const s4*)(WT + (ct * 16 + row) * 128 + kb * 16 + kq * 4);
    float bias[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) bias[ct] = b[ct * 16 + row];
    float(*tw)[65] = tile[wave];
    const int tiles = (n + 15) / 16;
    const int rounds = (tiles + 4 * (int)gridDim.x - 1) / (4 * (int)gridDim.x);
    const int ln = lane >> 2, q = lane & 3;   // epilogue: node ln of the tile, channels 16 q ..
    for (int rd = 0; rd < rounds; ++rd) {
        const int t = (rd * (int)gridDim.x + (int)blockIdx.x) * 4 + wave;
        const int n0 = t * 16;
        if (t < tiles) {
            const int node = n0 + row;
            f4 acc[4];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[ct] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) {
                const uint16_t* src = kb < 4 ? h : m;
                const int k0 = (kb & 3) * 16 + kq * 4;
                s4 a = (s4){0, 0, 0, 0};
                if (node < n) a = *(const s4*)(src + (int64_t)node * 64 + k0);
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) acc[ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv[kb][ct], acc[ct], 0, 0, 0);
            }
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) tw[kq * 4 + r][ct * 16 + row] = acc[ct][r] + bias[ct];
        }
        __syncthreads();
        const int nd = n0 + ln;
        if (t < tiles && nd < n) {
            const int64_t o = (int64_t)nd * 64 + q * 16;
            const uint4 h0 = *(const uint4*)(h + o), h1 = *(const uint4*)(h + o + 8);
            const uint32_t hw[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
            uint32_t zo[8], ho[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float z0 = tw[ln][q * 16 + 2 * c], z1 = tw[ln][q * 16 + 2 * c + 1];
                zo[c] = (uint32_t)f2bf(z0) | ((uint32_t)f2bf(z1) << 16);
                const float o0 = bf2f((uint16_t)(hw[c] & 0xffffu)) + fmaxf(z0, 0.f);
                const float o1 = bf2f((uint16_t)(hw[c] >> 16)) + fmaxf(z1, 0.f);
                ho[c] = (uint32_t)f2bf(o0) | ((uint32_t)f2bf(o1) << 16);
            }
            *(uint4*)(hout + o) = make_uint4(ho[0], ho[1], ho[2], ho[3]);
            *(uint4*)(hout + o + 8) = make_uint4(ho[4], ho[5], ho[6], ho[7]);
            if (z) {
                *(uint4*)(z + o) = make_uint4(zo[0], zo[1], zo[2], zo[3]);
                *(uint4*)(z + o + 8) = make_uint4(zo[4], zo[5], zo[6], zo[7]);
            }
        }
        __syncthreads();
    }
}

__global__ void k_decode(int n, const uint16_t* __restrict__ h, const float* __restrict__ w, const float* bdec,
                         float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = bdec[0];
    for (int c = 0; c < 64; ++c) s += w[c] * bf2f(h[(int64_t)i * 64 + c]);
    out[i] = s;
}

// ---------------------------------------------------------------- loss
// e = out - t ; q = A e ; partial = e.q  (A in ELL, FP64)
__global__ void k_loss_q(int n, int K, const int32_t* nb, const double* a, const double* diag, const float* out,
                         const float* t, double* q, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ei = (double)out[i] - (double)t[i];
        double s = diag[i] * ei;
        for (int k = 0; k < K; ++k) {
            const int j = nb[(int64_t)i * K + k];
            if (j >= 0) s -= a[(int64_t)i * K + k] * ((double)out[j] - (double)t[j]);
        }
        q[i] = s;
        acc += ei * s;
    }
    acc = block_sum_d(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// q2 = A q ; partial = q.q   (the residual A e of the guess is q)
__global__ void k_loss_res(int n, int K, const int32_t* nb, const double* a, const double* diag, const double* q,
                           double* q2, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = diag[i] * q[i];
        for (int k = 0; k < K; ++k) {
            const int j = nb[(int64_t)i * K + k];
            if (j >= 0) s -= a[(int64_t)i * K + k] * q[j];
        }
        q2[i] = s;
        acc += q[i] * q[i];
    }
    acc = block_sum_d(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// q <- wa q + wr q2  (dloss/dout / 2)
__global__ void k_combine(int n, double* q, const double* q2, double wa, double wr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = wa * q[i] + wr * q2[i];
}

// ---------------------------------------------------------------- backward
// dout = 2 q / norm ; g = dout w_dec (f32) ; partials of dw_dec, db_dec
__global__ void __launch_bounds__(256) k_decode_bwd(int n, const double* q, double norm, const uint16_t* h,
                                                    const float* w, float* g, float* partial /*[blocks][65]*/) {
    __shared__ float sh[65][4];
    float acc[65];
    for (int c = 0; c < 65; ++c) acc[c] = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = (float)(2.0 * q[i] / norm);
        for (int c = 0; c < 64; ++c) {
            acc[c] += d * bf2f(h[(int64_t)i * 64 + c]);
            g[(int64_t)i * 64 + c] = d * w[c];
        }
        acc[64] += d;
    }
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int c = 0; c < 65; ++c) {
        float v = acc[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (l == 0) sh[c][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 65)
        partial[(int64_t)blockIdx.x * 65 + threadIdx.x] =
            sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
}

// dz = g * (z > 0) as BF16, partial db
__global__ void __launch_bounds__(256) k_dz(int n, const float* __restrict__ g, const uint16_t* __restrict__ z,
                                            uint16_t* __restrict__ dz, float* partial /*[blocks][64]*/) {
    __shared__ float sh[256];
    const int c = threadIdx.x & 63;
    float acc = 0.f;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)n * 64;
         t += (int64_t)gridDim.x * blockDim.x) {
        const float v = bf2f(z[t]) > 0.f ? g[t] : 0.f;
        dz[t] = f2bf(v);
        acc += v;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < 64)
        partial[(int64_t)blockIdx.x * 64 + c] = sh[c] + sh[c + 64] + sh[c + 128] + sh[c + 192];
}

// dX = dz W^T (Wf stored [128][64]); g += dX[:, :64]; dm = dX[:, 64:] / deg
__global__ void __launch_bounds__(256) k_dx(int n, int K, const int32_t* __restrict__ nb, const uint16_t* __restrict__ dz,
                                            const uint16_t* __restrict__ Wf, float* __restrict__ g,
                                            float* __restrict__ dm) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n0 = (blockIdx.x * 4 + wave) * 16;
    if (n0 >= n) return;
    const int row = lane & 15, kq = lane >> 4;
    const int node = n0 + row;
    f4 acc[8];
    for (int ct = 0; ct < 8; ++ct) acc[ct] = (f4){0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < 4; ++kb) {
        s4 a = (s4){0, 0, 0, 0};
        if (node < n) a = *(const s4*)(dz + (int64_t)node * 64 + kb * 16 + kq * 4);
        for (int ct = 0; ct < 8; ++ct) {
            const int j = ct * 16 + row;
            const s4 bv = *(const s4*)(Wf + j * 64 + kb * 16 + kq * 4);
            acc[ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[ct], 0, 0, 0);
        }
    }
    for (int r = 0; r < 4; ++r) {
        const int nd = n0 + kq * 4 + r;
        if (nd >= n) continue;
        int deg = 0;
        for (int k = 0; k < K; ++k) deg += nb[(int64_t)nd * K + k] >= 0;
        const float s = deg ? 1.f / deg : 0.f;
        for (int ct = 0; ct < 8; ++ct) {
            const int j = ct * 16 + row;
            if (j < 64) g[(int64_t)nd * 64 + j] += acc[ct][r];
            else dm[(int64_t)nd * 64 + (j - 64)] = acc[ct][r] * s;
        }
    }
}

// g_j += sum_{i in N(j)} dm_i  (the adjacency is symmetric)
__global__ void k_mean_bwd(int n, int K, const int32_t* __restrict__ nb, const float* __restrict__ dm,
                           float* __restrict__ g) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * 64) return;
    const int j = (int)(t >> 6), c = (int)(t & 63);
    float s = 0.f;
    for (int k = 0; k < K; ++k) {
        const int i = nb[(int64_t)j * K + k];
        if (i >= 0) s += dm[(int64_t)i * 64 + c];
    }
    g[t] += s;
}

// dW = [h | m]^T dz over node chunks of 64 (split-K); partial [blocks][128][64].
// Tiles staged in LDS; each wave owns 8 of the 32 16x16 output tiles.
__global__ void __launch_bounds__(256) k_dw(int n, const uint16_t* __restrict__ h, const uint16_t* __restrict__ m,
                                            const uint16_t* __restrict__ dz, float* __restrict__ partial) {
    __shared__ uint16_t X[64][128 + 8];
    __shared__ uint16_t D[64][64 + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    f4 acc[8];
    for (int q = 0; q < 8; ++q) acc[q] = (f4){0.f, 0.f, 0.f, 0.f};
    for (int base = blockIdx.x * 64; base < n; base += gridDim.x * 64) {
        for (int t = threadIdx.x; t < 64 * 16; t += 256) {
            const int r = t >> 4, c4 = (t & 15) * 4;
            const int nd = base + r;
            uint64_t vh = 0, vm = 0, vd = 0;
            if (nd < n) {
                vh = *(const uint64_t*)(h + (int64_t)nd * 64 + c4);
                vm = *(const uint64_t*)(m + (int64_t)nd * 64 + c4);
                vd = *(const uint64_t*)(dz + (int64_t)nd * 64 + c4);
            }
            for (int u = 0; u < 4; ++u) {
                X[r][c4 + u] = (uint16_t)(vh >> (16 * u));
                X[r][64 + c4 + u] = (uint16_t)(vm >> (16 * u));
                D[r][c4 + u] = (uint16_t)(vd >> (16 * u));
            }
        }
        __syncthreads();
        for (int kb = 0; kb < 4; ++kb) {   // 16 nodes per k-step
            for (int q = 0; q < 8; ++q) {
                const int tile = wave * 8 + q;      // 0..31
                const int it = tile >> 2, ct = tile & 3;
                s4 a, bv;
                for (int u = 0; u < 4; ++u) {
                    a[u] = (short)X[kb * 16 + kq * 4 + u][it * 16 + row];
                    bv[u] = (short)D[kb * 16 + kq * 4 + u][ct * 16 + row];
                }
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[q], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    for (int q = 0; q < 8; ++q) {
        const int tile = wave * 8 + q;
        const int it = tile >> 2, ct = tile & 3;
        for (int r = 0; r < 4; ++r) {
            const int i = it * 16 + kq * 4 + r, j = ct * 16 + row;
            partial[(int64_t)blockIdx.x * 8192 + i * 64 + j] = acc[q][r];
        }
    }
}

// encoder backward: d0 = g (z0 > 0); partial of dW_enc [3][64] and db_enc [64]
__global__ void __launch_bounds__(256) k_encode_bwd(int n, const float* __restrict__ g, const uint16_t* __restrict__ z0,
                                                    const float* __restrict__ x, float* partial /*[blocks][256]*/) {
    __shared__ float sh[4][256];
    const int c = threadIdx.x & 63;
    float a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)n * 64;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t >> 6);
        const float d = bf2f(z0[t]) > 0.f ? g[t] : 0.f;
        a0 += x[3 * i] * d; a1 += x[3 * i + 1] * d; a2 += x[3 * i + 2] * d; a3 += d;
    }
    sh[0][threadIdx.x] = a0; sh[1][threadIdx.x] = a1; sh[2][threadIdx.x] = a2; sh[3][threadIdx.x] = a3;
    __syncthreads();
    if (threadIdx.x < 64)
        for (int v = 0; v < 4; ++v)
            partial[(int64_t)blockIdx.x * 256 + v * 64 + c] = sh[v][c] + sh[v][c + 64] + sh[v][c + 128] + sh[v][c + 192];
}

// sum partial[b][k] over blocks into out[k] (fixed order)
__global__ void k_reduce_f(const float* partial, int nb, int nv, float* out, float scale) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nv) return;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += partial[(int64_t)b * nv + k];
    out[k] = (float)(acc * scale);
}

// Adam on the FP32 master parameters; refresh the BF16 copies of the layer weights
__global__ void k_adam(int np, float* theta, const float* grad, float* m1, float* m2, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const float gk = grad[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk;
    const float v = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = v;
    theta[k] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// layer weights W[128][64] (f32, inside theta) -> Wf [128][64] and WT [64][128] BF16
__global__ void k_wcopy(const float* W, uint16_t* Wf, uint16_t* WT) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8192) return;
    const int r = t >> 6, c = t & 63;
    const uint16_t v = f2bf(W[t]);
    Wf[t] = v;
    WT[c * 128 + r] = v;
}

}  // namespace

// ====================================================================== host side

static int nblk(int64_t work, int cap = 1024) {
    int64_t b = (work + 255) / 256;
    if (b < 1) b = 1;
    return (int)(b > cap ? cap : b);
}

int Gnn::n_params(int L) { return 3 * 64 + 64 + L * (128 * 64 + 64) + 64 + 1; }

hipError_t Gnn::alloc(int n_, int K_, int L_, hipStream_t s) {
    n = n_; K = K_; L = L_; stream = s;
    np = n_params(L);
    auto A = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
        hipMemsetAsync(p, 0, bytes ? bytes : 8, s);
        bufs.push_back(p);
        return p;
    };
    theta = (float*)A(sizeof(float) * np);
    grad = (float*)A(sizeof(float) * np);
    m1 = (float*)A(sizeof(float) * np);
    m2 = (float*)A(sizeof(float) * np);
    wbf = (uint16_t*)A(sizeof(uint16_t) * 2 * 8192 * L);
    x = (float*)A(sizeof(float) * 3 * (size_t)n);
    h = (uint16_t*)A(sizeof(uint16_t) * 64 * (size_t)n * (L + 1));
    mm = (uint16_t*)A(sizeof(uint16_t) * 64 * (size_t)n * L);
    z = (uint16_t*)A(sizeof(uint16_t) * 64 * (size_t)n * (L + 1));
    out = (float*)A(sizeof(float) * (size_t)n);
    g = (float*)A(sizeof(float) * 64 * (size_t)n);
    dm = (float*)A(sizeof(float) * 64 * (size_t)n);
    dz = (uint16_t*)A(sizeof(uint16_t) * 64 * (size_t)n);
    q = (double*)A(sizeof(double) * (size_t)n);
    q2 = (double*)A(sizeof(double) * (size_t)n);
    red = (double*)A(sizeof(double) * 4096);
    scal = (double*)A(sizeof(double) * 16);
    part = (float*)A(sizeof(float) * 1024 * 8192);
    for (void* p : bufs)
        if (!p) return hipErrorOutOfMemory;
    return hipGetLastError();
}

void Gnn::release() {
    for (void* p : bufs) hipFree(p);
    bufs.clear();
}

void Gnn::refresh_weights() {
    const float* lay = theta + 3 * 64 + 64;
    for (int l = 0; l < L; ++l) {
        const float* W = lay + l * (8192 + 64);
        hipLaunchKernelGGL(k_wcopy, dim3(32), dim3(256), 0, stream, W, wbf + (size_t)l * 2 * 8192,
                           wbf + (size_t)l * 2 * 8192 + 8192);
    }
}

hipError_t Gnn::features(const double* r, const double* b, const double* phi) {
    const int nb = nblk(n, 512);
    hipLaunchKernelGGL(k_sumsq3, dim3(nb), dim3(256), 0, stream, n, r, b, phi, red);
    hipLaunchKernelGGL(k_reduce_n, dim3(1), dim3(64), 0, stream, red, nb, 3, scal);
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, stream, n, r, b, phi, scal, x);
    return hipGetLastError();
}

hipError_t Gnn::target(const double* phi_star, const double* phi_prev, float* t) {
    hipLaunchKernelGGL(k_target, dim3((n + 255) / 256), dim3(256), 0, stream, n, phi_star, phi_prev, scal, dbar, t);
    return hipGetLastError();
}

hipError_t Gnn::guess(const double* phi_prev, double* out_phi) {
    hipLaunchKernelGGL(k_guess, dim3((n + 255) / 256), dim3(256), 0, stream, n, out, phi_prev, scal, dbar, out_phi);
    return hipGetLastError();
}

hipError_t Gnn::forward(const float* xin, bool train) {
    cur_x = xin;
    const float* Wenc = theta;
    const float* benc = theta + 192;
    const float* lay = theta + 256;
    const float* wdec = lay + L * (8192 + 64);
    const size_t NH = (size_t)n * 64;
    hipLaunchKernelGGL(k_encode, dim3((unsigned)((NH + 255) / 256)), dim3(256), 0, stream, n, xin, Wenc, benc,
                       train ? z : nullptr, h);
    const int tiles = (n + 15) / 16;
    for (int l = 0; l < L; ++l) {
        const uint16_t* hl = h + (train ? (size_t)l * NH : (l % 2) * NH);
        uint16_t* hn = h + (train ? (size_t)(l + 1) * NH : ((l + 1) % 2) * NH);
        uint16_t* ml = mm + (train ? (size_t)l * NH : 0);
        hipLaunchKernelGGL(k_mean, dim3((unsigned)((n * 16 + 255) / 256)), dim3(256), 0, stream, n, K, nb, hl, ml);
        const float* bl = lay + l * (8192 + 64) + 8192;
        hipLaunchKernelGGL(k_layer_fwd, dim3(std::min((tiles + 3) / 4, kLayerBlocks)), dim3(256), 0, stream, n, hl, ml,
                           wbf + (size_t)l * 2 * 8192 + 8192, bl, train ? z + (size_t)(l + 1) * NH : nullptr, hn);
    }
    const uint16_t* hL = h + (train ? (size_t)L * NH : (L % 2) * NH);
    hipLaunchKernelGGL(k_decode, dim3((n + 255) / 256), dim3(256), 0, stream, n, hL, wdec, wdec + 64, out);
    return hipGetLastError();
}

hipError_t Gnn::loss_backward(const float* t, double norm, const int32_t* anb, const double* aa, const double* ad,
                              int aK, double* loss_out) {
    return loss_backward2(t, norm, 0.0, anb, aa, ad, aK, loss_out);
}

hipError_t Gnn::loss_backward2(const float* t, double norm_a, double norm_r, const int32_t* anb, const double* aa,
                               const double* ad, int aK, double* loss_out) {
    const size_t NH = (size_t)n * 64;
    const int nbq = nblk(n, 512);
    // A-norm part: e^T A e / norm_a ; residual part: |A e|^2 / norm_r
    const double wa = (loss_kind == 1) ? 0.0 : 1.0 / norm_a;
    const double wr = (loss_kind == 0 || norm_r <= 0.0) ? 0.0 : 1.0 / norm_r;
    hipLaunchKernelGGL(k_loss_q, dim3(nbq), dim3(256), 0, stream, n, aK, anb, aa, ad, out, t, q, red);
    hipLaunchKernelGGL(k_reduce_n, dim3(1), dim3(64), 0, stream, red, nbq, 1, scal + 4);
    if (wr != 0.0) {
        hipLaunchKernelGGL(k_loss_res, dim3(nbq), dim3(256), 0, stream, n, aK, anb, aa, ad, q, q2, red);
        hipLaunchKernelGGL(k_reduce_n, dim3(1), dim3(64), 0, stream, red, nbq, 1, scal + 6);
    }
    hipLaunchKernelGGL(k_combine, dim3((n + 255) / 256), dim3(256), 0, stream, n, q, q2, wa, wr);
    const double norm = 1.0;
    const float* lay = theta + 256;
    const float* wdec = lay + L * (8192 + 64);
    float* glay = grad + 256;
    float* gdec = glay + L * (8192 + 64);
    int nb = nblk(n, 1024);
    hipLaunchKernelGGL(k_decode_bwd, dim3(nb), dim3(256), 0, stream, n, q, norm, h + (size_t)L * NH, wdec, g, part);
    hipLaunchKernelGGL(k_reduce_f, dim3(1), dim3(128), 0, stream, part, nb, 65, gdec, 1.f);
    const int tiles = (n + 15) / 16;
    for (int l = L - 1; l >= 0; --l) {
        nb = nblk(NH, 1024);
        float* gb = glay + l * (8192 + 64) + 8192;
        float* gW = glay + l * (8192 + 64);
        hipLaunchKernelGGL(k_dz, dim3(nb), dim3(256), 0, stream, n, g, z + (size_t)(l + 1) * NH, dz, part);
        hipLaunchKernelGGL(k_reduce_f, dim3(1), dim3(64), 0, stream, part, nb, 64, gb, 1.f);
        const int nbw = (int)std::min<int64_t>(1024, ((int64_t)n + 63) / 64);
        hipLaunchKernelGGL(k_dw, dim3(nbw), dim3(256), 0, stream, n, h + (size_t)l * NH, mm + (size_t)l * NH, dz, part);
        hipLaunchKernelGGL(k_reduce_f, dim3(32), dim3(256), 0, stream, part, nbw, 8192, gW, 1.f);
        hipLaunchKernelGGL(k_dx, dim3((tiles + 3) / 4), dim3(256), 0, stream, n, K, nb_ptr(), dz, wbf + (size_t)l * 2 * 8192,
                           g, dm);
        hipLaunchKernelGGL(k_mean_bwd, dim3((unsigned)((NH + 255) / 256)), dim3(256), 0, stream, n, K, nb_ptr(), dm, g);
    }
    nb = nblk(NH, 1024);
    hipLaunchKernelGGL(k_encode_bwd, dim3(nb), dim3(256), 0, stream, n, g, z, cur_x, part);
    hipLaunchKernelGGL(k_reduce_f, dim3(1), dim3(256), 0, stream, part, nb, 256, grad, 1.f);
    if (loss_out) {
        double v[3] = {0.0, 0.0, 0.0};
        hipMemcpyAsync(v, scal + 4, sizeof(double) * 3, hipMemcpyDeviceToHost, stream);
        hipStreamSynchronize(stream);
        *loss_out = wa * v[0] + wr * v[2];
    }
    return hipGetLastError();
}

hipError_t Gnn::adam(float lr, int64_t t) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), dim3(256), 0, stream, np, theta, grad, m1, m2, lr, b1, b2,
                       1e-8f, c1, c2);
    refresh_weights();
    return hipGetLastError();
}

}  // namespace picmi

namespace picmi {
namespace {
__global__ void k_anorm(int n, int K, const int32_t* nb, const double* a, const double* diag, const float* t,
                        double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = diag[i] * (double)t[i];
        for (int k = 0; k < K; ++k) {
            const int j = nb[(int64_t)i * K + k];
            if (j >= 0) s -= a[(int64_t)i * K + k] * (double)t[j];
        }
        acc += (double)t[i] * s;
    }
    acc = block_sum_d(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}
}  // namespace

double gnn_rnorm(Gnn& G, const float* t, const int32_t* anb, const double* aa, const double* ad, int aK) {
    // |A t|^2 : q = A t (via the loss kernel with out = t, target 0), then q.q
    const int nb = nblk(G.n, 512);
    hipMemsetAsync(G.dm, 0, sizeof(float) * G.n, G.stream);
    hipLaunchKernelGGL(k_loss_q, dim3(nb), dim3(256), 0, G.stream, G.n, aK, anb, aa, ad, t, G.dm, G.q, G.red);
    hipLaunchKernelGGL(k_loss_res, dim3(nb), dim3(256), 0, G.stream, G.n, aK, anb, aa, ad, G.q, G.q2, G.red);
    hipLaunchKernelGGL(k_reduce_n, dim3(1), dim3(64), 0, G.stream, G.red, nb, 1, G.scal + 7);
    double v = 0.0;
    hipMemcpyAsync(&v, G.scal + 7, sizeof(double), hipMemcpyDeviceToHost, G.stream);
    hipStreamSynchronize(G.stream);
    return v;
}

double gnn_anorm(Gnn& G, const float* t, const int32_t* anb, const double* aa, const double* ad, int aK) {
    const int nb = nblk(G.n, 512);
    hipLaunchKernelGGL(k_anorm, dim3(nb), dim3(256), 0, G.stream, G.n, aK, anb, aa, ad, t, G.red);
    hipLaunchKernelGGL(k_reduce_n, dim3(1), dim3(64), 0, G.stream, G.red, nb, 1, G.scal + 5);
    double v = 0.0;
    hipMemcpyAsync(&v, G.scal + 5, sizeof(double), hipMemcpyDeviceToHost, G.stream);
    hipStreamSynchronize(G.stream);
    return v;
}
}  // namespace picmi
