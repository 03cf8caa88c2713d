// Poisson solve of ElectroStaticSolver::solveFields (ElectroStaticSolver.C:52-71):
// A phi = rho V / eps0 with the FV Laplacian, solved by PCG with a DIC (incomplete
// Cholesky, zero fill) preconditioner in the greedy multicolour ordering.  The solve runs
// on the system renumbered by colour (each colour a contiguous range, coefficients
// column-major), so each colour's forward and backward sweep is one streaming kernel.
// One PCG iteration is 2 + 2 (ncol - 1) streaming kernels plus three one-block kernels
// that finish alpha, r.r (the convergence flag) and beta on the device: the
// search-direction update is fused into the matrix product, the colour-0 sweep into the
// vector update and the dot products into the sweeps; the host enqueues iterations in
// batches and reads the convergence flag once per batch.  Mode 0 runs everything in FP64.  Mode 1 runs FP32
// inner PCG (FP32 vectors and coefficients, FP64 dot products) inside FP64 iterative
// refinement; the refinement residual b - A phi and the stopping test are FP64.
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "device.h"
#include "kernels.h"

namespace picmi {

namespace {
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
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

// partial[block] = sum a.b
template <class T>
__global__ void __launch_bounds__(256) k_dot(int n, const T* __restrict__ a, const T* __restrict__ b, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// scal[dst] = sum(partial)  (fixed order: deterministic)
__global__ void k_reduce(const double* partial, int nb, double* scal, int dst) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partial[i];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) scal[dst] = acc;
}

// r = b - A x (FP64), partial = r.r
__global__ void __launch_bounds__(256) k_residual(int n, int K, const int32_t* __restrict__ nb,
                                                  const double* __restrict__ a, const double* __restrict__ diag,
                                                  const double* __restrict__ x, const double* __restrict__ b,
                                                  double* __restrict__ r, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = diag[i] * x[i];
        for (int k = 0; k < K; ++k) {
            const int j = nb[(int64_t)i * K + k];
            if (j >= 0) s -= a[(int64_t)i * K + k] * x[j];
        }
        const double ri = b[i] - s;
        r[i] = ri;
        acc += ri * ri;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_rhs(int n, const double* rho, const double* V, const double* bsrc, const double* corr, double* b,
                      double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v = rho[i] * V[i] / kEps0 + bsrc[i];
        if (corr) v = v + corr[i];
        b[i] = v;
        acc += v;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_sub_mean(int n, double* v, const double* scal, int slot) {
    const double m = scal[slot] / n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] -= m;
}

__global__ void k_sum_partial(int n, const double* v, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += v[i];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_d2f(int n, const double* s, float* d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = (float)s[i];
}
__global__ void k_add_f2d(int n, const float* s, double* d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] += (double)s[i];
}
template <class T>
__global__ void k_zero(int n, T* d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = (T)0;
}

int nblk(int n) {
    int b = (n + 255) / 256;
    return b < 1 ? 1 : (b > kRedBlocks ? kRedBlocks : b);
}
__device__ __forceinline__ int nblk_d(int n) {
    int b = (n + 255) / 256;
    return b < 1 ? 1 : (b > kRedBlocks ? kRedBlocks : b);
}

// scalar slots of W.scal
enum : int { kRR = 5, kBB = 6, kMean = 7, kDone = 8, kConvIt = 9, kTol2 = 10, kRzPar = 11 /* and 12 */ };
// reduction segments of W.red (kRedBlocks doubles each): 0 p.q, 1 r.r, 2 spare, 3 + c r.z of colour c
constexpr int kSegPq = 0, kSegRr = 1, kSegRz0 = 3;

// fixed-order sum of nb partials by one block (thread 0 gets the result)
__device__ __forceinline__ double sum_partials(const double* partial, int nb, double* sh) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) t += partial[b];
    return block_sum(t, sh);
}

// The scalars of an iteration (alpha, beta, the convergence test) are finished inside the
// grid kernel that needs them: every block sums the previous kernel's partials itself, in
// the same fixed order, so all blocks hold the same value bit for bit and no one-block
// finishing launch (about 4 us plus its gap, three per iteration) sits between the grid
// kernels.  A "last block" finish inside the producing kernel would need a device-scope
// fence per block, which writes back the XCD's L2.  Block 0 stores what the host or a later
// kernel reads: r.z by iteration parity (the kernel that stores it still reads the previous
// one), r.r and the convergence flag.

// sum_partials given to every thread of the block
__device__ __forceinline__ double all_sum(const double* partial, int nb, double* sh) {
    const double t = sum_partials(partial, nb, sh);
    if (threadIdx.x == 0) sh[4] = t;
    __syncthreads();
    const double v = sh[4];
    __syncthreads();
    return v;
}

// r.r of the last update against the tolerance: true once converged (block 0 then records
// the iteration for the host); r.r is kept for the host either way
__device__ __forceinline__ bool converged(double* scal, const double* red, int nb_rr, int it, double* sh) {
    const double rr = all_sum(red + kSegRr * kRedBlocks, nb_rr, sh);
    const bool done = rr <= scal[kTol2];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[kRR] = rr;
        if (done) { scal[kDone] = 1.0; scal[kConvIt] = (double)it; }
    }
    return done;
}

__global__ void k_pcg_reset(double* scal, double tol2) {
    if (threadIdx.x == 0) {
        scal[kDone] = 0.0; scal[kConvIt] = -1.0; scal[kTol2] = tol2;
    }
}

// iteration it (from 1): r.z of the current z (every colour's partials), beta = r.z /
// r.z of the previous iteration (0 at it = 1); p_new = z + beta p_old, q = A p_new
// (neighbours' p formed on the fly), p.q partials.  rr_test (one colour, so no sweep kernel
// follows the update): the convergence test of iteration it - 1 comes first.
template <class T>
__global__ void __launch_bounds__(256) k_pcg_spmv(int n, int K, const int32_t* __restrict__ nbT,
                                                  const T* __restrict__ aT, const T* __restrict__ diag,
                                                  const T* __restrict__ z, const T* __restrict__ pold,
                                                  T* __restrict__ pnew, T* __restrict__ q, double* scal, double* red,
                                                  int it, int ncol, const int32_t* __restrict__ col_start,
                                                  int rr_test) {
    __shared__ double sh[8];
    if (scal[kDone] != 0.0) return;
    if (rr_test && it > 1 && converged(scal, red, gridDim.x, it - 1, sh)) return;
    double rz = 0.0;
    for (int c = 0; c < ncol; ++c)
        rz += all_sum(red + (size_t)(kSegRz0 + c) * kRedBlocks, nblk_d(col_start[c + 1] - col_start[c]), sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[kRzPar + (it & 1)] = rz;
    const T beta = (T)(it == 1 ? 0.0 : rz / scal[kRzPar + ((it - 1) & 1)]);
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T pi = z[i] + beta * pold[i];
        pnew[i] = pi;
        T s = diag[i] * pi;
        for (int k = 0; k < K; ++k) {
            const int j = nbT[(size_t)k * n + i];
            if (j >= 0) s -= aT[(size_t)k * n + i] * (z[j] + beta * pold[j]);
        }
        q[i] = s;
        acc += (double)pi * (double)s;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) red[kSegPq * kRedBlocks + blockIdx.x] = acc;
}

// x += alpha p, r -= alpha q, r.r partials; z = r / rD on the first colour (its whole
// forward sweep); with one colour also the r.z partials
template <class T>
__global__ void __launch_bounds__(256) k_pcg_update(int n, int c0end, int rz_here, T* __restrict__ x, T* __restrict__ r,
                                                    const T* __restrict__ p, const T* __restrict__ q,
                                                    const T* __restrict__ rD, T* __restrict__ z, double* scal,
                                                    double* red, int it) {
    __shared__ double sh[8];
    if (scal[kDone] != 0.0) return;
    // alpha = r.z / p.q of this iteration (p.q from the spmv's partials, same grid)
    const T alpha = (T)(scal[kRzPar + (it & 1)] / all_sum(red + kSegPq * kRedBlocks, gridDim.x, sh));
    double acc = 0.0, acc2 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc += (double)ri * (double)ri;
        if (i < c0end) {
            const T zi = ri / rD[i];
            z[i] = zi;
            acc2 += (double)ri * (double)zi;
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) red[kSegRr * kRedBlocks + blockIdx.x] = acc;
    if (rz_here) {
        acc2 = block_sum(acc2, sh);
        if (threadIdx.x == 0) red[kSegRz0 * kRedBlocks + blockIdx.x] = acc2;
    }
}

// start of a solve: z = r / rD on the first colour, p_old = 0; with one colour r.z partials
template <class T>
__global__ void __launch_bounds__(256) k_pcg_zc0(int n, int c0end, int rz_here, const T* __restrict__ r,
                                                 const T* __restrict__ rD, T* __restrict__ z, T* __restrict__ pold,
                                                 double* red) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        pold[i] = (T)0;
        if (i < c0end) {
            const T zi = r[i] / rD[i];
            z[i] = zi;
            acc += (double)r[i] * (double)zi;
        }
    }
    if (!rz_here) return;
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) red[kSegRz0 * kRedBlocks + blockIdx.x] = acc;
}

// DIC forward sweep of colour range [c0, c1): z_i = (r_i + sum_{j < c0} a_ij z_j) / rD_i;
// the last colour's z is final there, so its r.z partials are stored (segment seg).  The
// first sweep after an update (it > 0) makes the convergence test of iteration it on the
// update's nb_rr partials.
template <class T>
__global__ void __launch_bounds__(256) k_pcg_fwd(int n, int c0, int c1, int K, const int32_t* __restrict__ nbT,
                                                 const T* __restrict__ aT, const T* __restrict__ rD,
                                                 const T* __restrict__ r, T* __restrict__ z, double* scal, double* red,
                                                 int seg, int it, int nb_rr) {
    __shared__ double sh[8];
    if (scal[kDone] != 0.0) return;
    if (it > 0 && converged(scal, red, nb_rr, it, sh)) return;
    double acc = 0.0;
    for (int i = c0 + blockIdx.x * blockDim.x + threadIdx.x; i < c1; i += gridDim.x * blockDim.x) {
        T s = r[i];
        for (int k = 0; k < K; ++k) {
            const int j = nbT[(size_t)k * n + i];
            if (j >= 0 && j < c0) s += aT[(size_t)k * n + i] * z[j];
        }
        const T zi = s / rD[i];
        z[i] = zi;
        acc += (double)r[i] * (double)zi;
    }
    if (seg < 0) return;
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) red[(size_t)(kSegRz0 + seg) * kRedBlocks + blockIdx.x] = acc;
}

// DIC backward sweep of colour range [c0, c1): z_i += (sum_{j >= c1} a_ij z_j) / rD_i and
// its r.z partials (z of the colour is final there)
template <class T>
__global__ void __launch_bounds__(256) k_pcg_bwd(int n, int c0, int c1, int K, const int32_t* __restrict__ nbT,
                                                 const T* __restrict__ aT, const T* __restrict__ rD,
                                                 const T* __restrict__ r, T* __restrict__ z, double* scal, double* red,
                                                 int seg) {
    __shared__ double sh[8];
    if (scal[kDone] != 0.0) return;
    double acc = 0.0;
    for (int i = c0 + blockIdx.x * blockDim.x + threadIdx.x; i < c1; i += gridDim.x * blockDim.x) {
        T s = 0;
        for (int k = 0; k < K; ++k) {
            const int j = nbT[(size_t)k * n + i];
            if (j >= c1) s += aT[(size_t)k * n + i] * z[j];
        }
        const T zi = z[i] + s / rD[i];
        z[i] = zi;
        acc += (double)r[i] * (double)zi;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) red[(size_t)(kSegRz0 + seg) * kRedBlocks + blockIdx.x] = acc;
}

// residual of the colour-ordered FP64 system: r = b - A x, partial r.r
__global__ void __launch_bounds__(256) k_residual_p(int n, int K, const int32_t* __restrict__ nbT,
                                                    const double* __restrict__ aT, const double* __restrict__ diag,
                                                    const double* __restrict__ x, const double* __restrict__ b,
                                                    double* __restrict__ r, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = diag[i] * x[i];
        for (int k = 0; k < K; ++k) {
            const int j = nbT[(size_t)k * n + i];
            if (j >= 0) s -= aT[(size_t)k * n + i] * x[j];
        }
        const double ri = b[i] - s;
        r[i] = ri;
        acc += ri * ri;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_gather(int n, const int32_t* __restrict__ perm, const double* __restrict__ src,
                         double* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[perm[i]];
}
__global__ void k_scatter_back(int n, const int32_t* __restrict__ perm, const double* __restrict__ src,
                               double* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[perm[i]] = src[i];
}

// fixed-order sum of partial into scal[slot]
void reduce_to(const double* partial, int nb, double* scal, int slot, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, partial, nb, scal, slot);
}

template <class T>
struct Sys {
    const T* aT;
    const T* diag;
    const T* rD;
};

// Preconditioned CG on the colour-ordered system A x = r (x holds the start, r its
// residual).  Iterations are enqueued in batches; the host reads the device convergence
// flag after each batch.  Returns the iteration count of convergence or -1.
template <class T>
int pcg_fused(const PoissonDev& P, const Sys<T>& S, PoissonWork& W, T* x, T* r, T* z, T* p, T* p2, T* q,
              double tol_abs, int max_iter, hipStream_t s, int* iters_done, double* last_rr) {
    const int n = P.n, K = P.K, ncol = P.ncol, nb = nblk(n);
    const int* cs = P.col_start_h;
    const int c0end = cs[1];
    const int rz_here = ncol == 1 ? 1 : 0;
    hipLaunchKernelGGL(k_pcg_reset, dim3(1), dim3(64), 0, s, W.scal, tol_abs * tol_abs);
    // z = M^-1 r and its r.z partials (the first spmv forms beta = 0); p_old = 0
    hipLaunchKernelGGL((k_pcg_zc0<T>), dim3(nb), dim3(256), 0, s, n, c0end, rz_here, r, S.rD, z, p2, W.red);
    auto sweeps = [&](int it) {
        for (int c = 1; c < ncol; ++c) {
            const int a0 = cs[c], cnt = cs[c + 1] - a0;
            hipLaunchKernelGGL((k_pcg_fwd<T>), dim3(nblk(cnt)), dim3(256), 0, s, n, a0, cs[c + 1], K, P.nbT, S.aT,
                               S.rD, r, z, W.scal, W.red, c == ncol - 1 ? c : -1, c == 1 ? it : 0, nb);
        }
        for (int c = ncol - 2; c >= 0; --c) {
            const int a0 = cs[c], cnt = cs[c + 1] - a0;
            hipLaunchKernelGGL((k_pcg_bwd<T>), dim3(nblk(cnt)), dim3(256), 0, s, n, a0, cs[c + 1], K, P.nbT, S.aT,
                               S.rD, r, z, W.scal, W.red, c);
        }
    };
    sweeps(0);
    T* pa = p;    // written this iteration
    T* pb = p2;   // previous direction
    int it = 0, batch = 8, conv = -1;
    double h[3] = {0.0, -1.0, 0.0};
    while (it < max_iter) {
        const int B = std::min(batch, max_iter - it);
        for (int k = 0; k < B; ++k) {
            ++it;
            hipLaunchKernelGGL((k_pcg_spmv<T>), dim3(nb), dim3(256), 0, s, n, K, P.nbT, S.aT, S.diag, z, pb, pa, q,
                               W.scal, W.red, it, ncol, P.col_start_d, rz_here);
            hipLaunchKernelGGL((k_pcg_update<T>), dim3(nb), dim3(256), 0, s, n, c0end, rz_here, x, r, pa, q, S.rD, z,
                               W.scal, W.red, it);
            sweeps(it);
            std::swap(pa, pb);
        }
        hipMemcpyAsync(h, W.scal + kDone, sizeof(double) * 2, hipMemcpyDeviceToHost, s);
        hipMemcpyAsync(h + 2, W.scal + kRR, sizeof(double), hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
        *last_rr = h[2];
        if (h[0] != 0.0) { conv = (int)h[1]; break; }
        batch = std::min(batch * 2, 32);
    }
    *iters_done = conv >= 0 ? conv : max_iter;
    return conv;
}

}  // namespace

hipError_t poisson_rhs(const PoissonDev& P, const double* rho, const double* corr, double* b, bool singular,
                       double* scratch, hipStream_t s) {
    const int n = P.n, nb = nblk(n);
    double* partial = scratch;
    double* scal = scratch + kRedBlocks;
    hipLaunchKernelGGL(k_rhs, dim3(nb), dim3(256), 0, s, n, rho, P.V, P.bsrc, corr, b, partial);
    if (singular) {
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, partial, nb, scal, 0);
        hipLaunchKernelGGL(k_sub_mean, dim3(nb), dim3(256), 0, s, n, b, scal, 0);
    }
    return hipGetLastError();
}

hipError_t poisson_residual(const PoissonDev& P, PoissonWork& W, const double* x, const double* b, double* r,
                            hipStream_t s) {
    const int nb = nblk(P.n);
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, P.n, P.K, P.nb, P.a, P.diag, x, b, r, W.red);
    return hipGetLastError();
}

__global__ void k_extrapolate(double* __restrict__ phi, double* __restrict__ old, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t = phi[i];
    phi[i] = 2.0 * t - old[i];
    old[i] = t;
}

hipError_t poisson_extrapolate(double* phi, double* phi_old, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_extrapolate, dim3((n + 255) / 256), dim3(256), 0, s, phi, phi_old, n);
    return hipGetLastError();
}

int poisson_solve(const PoissonDev& P, PoissonWork& W, double* phi, const double* b, double tol, int max_iter,
                  int mode, bool singular, int* iters, int* refinements, double* rel, hipStream_t s) {
    const int n = P.n, nb = nblk(n);
    double* scal = W.scal;
    // colour order: x = phi[perm], b = b[perm]
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, s, n, P.corder, phi, W.xp);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, s, n, P.corder, b, W.bp);
    hipLaunchKernelGGL((k_dot<double>), dim3(nb), dim3(256), 0, s, n, W.bp, W.bp, W.red);
    reduce_to(W.red, nb, scal, kBB, s);
    auto residual = [&](double* rr_host) {
        hipLaunchKernelGGL(k_residual_p, dim3(nb), dim3(256), 0, s, n, P.K, P.nbT, P.aT, P.diag_p, W.xp, W.bp, W.r,
                           W.red);
        reduce_to(W.red, nb, scal, kRR, s);
        double h[2] = {0.0, 0.0};
        hipMemcpyAsync(h, scal + kRR, sizeof(double) * 2, hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
        *rr_host = h[0];
        return h[1];
    };
    double rr0 = 0.0;
    const double bb = residual(&rr0);
    double bn = std::sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    const double tol_abs = tol * bn;
    int total = 0, outer = 0;
    double last_rr = rr0;
    int status = 0;
    if (mode == 0) {
        if (std::sqrt(rr0) > tol_abs) {
            Sys<double> S{P.aT, P.diag_p, P.rD_p};
            int done = 0;
            const int conv = pcg_fused<double>(P, S, W, W.xp, W.r, W.z, W.p, W.p2, W.q, tol_abs, max_iter, s, &done,
                                               &last_rr);
            total = done;
            if (conv < 0) status = -1;
        }
    } else {
        // FP64 iterative refinement around FP32 PCG
        const double inner_rel = 1e-4;
        Sys<float> S{P.aT_f, P.diag_pf, P.rD_pf};
        double rr = rr0;
        for (outer = 0; outer < 50; ++outer) {
            if (outer > 0) residual(&rr);
            last_rr = rr;
            if (std::sqrt(rr) <= tol_abs) break;
            if (total >= max_iter) { status = -1; break; }
            hipLaunchKernelGGL(k_d2f, dim3(nb), dim3(256), 0, s, n, W.r, W.rf);
            hipLaunchKernelGGL((k_zero<float>), dim3(nb), dim3(256), 0, s, n, W.xf);
            const double itol = std::max(inner_rel * std::sqrt(rr), 0.5 * tol_abs);
            int done = 0;
            double irr = 0.0;
            pcg_fused<float>(P, S, W, W.xf, W.rf, W.zf, W.pf, W.pf2, W.qf, itol, max_iter - total, s, &done, &irr);
            total += done;
            hipLaunchKernelGGL(k_add_f2d, dim3(nb), dim3(256), 0, s, n, W.xf, W.xp);
        }
    }
    if (mode == 0 && total > 0) residual(&last_rr);  // the true FP64 residual
    hipLaunchKernelGGL(k_scatter_back, dim3(nb), dim3(256), 0, s, n, P.corder, W.xp, phi);
    if (singular) {
        hipLaunchKernelGGL(k_sum_partial, dim3(nb), dim3(256), 0, s, n, phi, W.red);
        reduce_to(W.red, nb, scal, kMean, s);
        hipLaunchKernelGGL(k_sub_mean, dim3(nb), dim3(256), 0, s, n, phi, scal, kMean);
    }
    *iters = total;
    *refinements = outer;
    *rel = std::sqrt(last_rr) / bn;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return -2;
    return status;
}

}  // namespace picmi
