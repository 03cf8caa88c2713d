// Poisson solve of ElectroStaticSolver::solveFields (ElectroStaticSolver.C:52-71):
// A phi = rho V / eps0 with the FV Laplacian in ELL form (K couplings per cell), solved
// by PCG with a DIC (incomplete Cholesky, zero fill) preconditioner in the greedy
// multicolour ordering, so each colour's forward and backward sweep is one parallel
// kernel.  Mode 0 runs everything in FP64.  Mode 1 runs FP32 inner PCG sweeps (FP32
// vectors and coefficients, FP64 dot products) inside FP64 iterative refinement; the
// refinement residual b - A phi and the stopping test are FP64.
#include <cmath>
#include <vector>

#include "device.h"
#include "kernels.h"

namespace picmi {

namespace {
constexpr int kRedBlocks = 512;

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

// y = A x, partial[block] = sum x.y
template <class T>
__global__ void __launch_bounds__(256) k_spmv_dot(int n, int K, const int32_t* __restrict__ nb,
                                                  const T* __restrict__ a, const T* __restrict__ diag,
                                                  const T* __restrict__ x, T* __restrict__ y, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T s = diag[i] * x[i];
        for (int k = 0; k < K; ++k) {
            const int j = nb[(int64_t)i * K + k];
            if (j >= 0) s -= a[(int64_t)i * K + k] * x[j];
        }
        y[i] = s;
        acc += (double)x[i] * (double)s;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0 && partial) partial[blockIdx.x] = acc;
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

// scalar slots: 0 rz, 1 pq, 2 alpha, 3 rz_new, 4 beta, 5 rr, 6 bb
__global__ void k_alpha(double* scal) {
    if (threadIdx.x == 0) scal[2] = scal[0] / scal[1];
}
__global__ void k_beta(double* scal) {
    if (threadIdx.x == 0) {
        scal[4] = scal[3] / scal[0];
        scal[0] = scal[3];
    }
}

// x += alpha p, r -= alpha q, partial = r.r
template <class T>
__global__ void __launch_bounds__(256) k_update(int n, const double* scal, T* __restrict__ x, T* __restrict__ r,
                                                const T* __restrict__ p, const T* __restrict__ q, double* partial) {
    __shared__ double sh[4];
    const T alpha = (T)scal[2];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc += (double)ri * (double)ri;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <class T>
__global__ void k_xpby(int n, const double* scal, const T* __restrict__ z, T* __restrict__ p) {
    const T beta = (T)scal[4];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

template <class T>
__global__ void k_copy(int n, const T* __restrict__ s, T* __restrict__ d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = s[i];
}

// DIC forward sweep of one colour: z_i = (r_i + sum_{lower colours} a_ij z_j) / rD_i
template <class T>
__global__ void k_dic_fwd(int cnt, const int32_t* __restrict__ cells, int col, int K, const int32_t* __restrict__ nb,
                          const T* __restrict__ a, const int32_t* __restrict__ colour, const T* __restrict__ rD,
                          const T* __restrict__ r, T* __restrict__ z) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = cells[t];
    T s = r[i];
    for (int k = 0; k < K; ++k) {
        const int j = nb[(int64_t)i * K + k];
        if (j >= 0 && colour[j] < col) s += a[(int64_t)i * K + k] * z[j];
    }
    z[i] = s / rD[i];
}

// DIC backward sweep of one colour: z_i += (sum_{higher colours} a_ij z_j) / rD_i
template <class T>
__global__ void k_dic_bwd(int cnt, const int32_t* __restrict__ cells, int col, int K, const int32_t* __restrict__ nb,
                          const T* __restrict__ a, const int32_t* __restrict__ colour, const T* __restrict__ rD,
                          T* __restrict__ z) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = cells[t];
    T s = 0;
    for (int k = 0; k < K; ++k) {
        const int j = nb[(int64_t)i * K + k];
        if (j >= 0 && colour[j] > col) s += a[(int64_t)i * K + k] * z[j];
    }
    z[i] = z[i] + s / rD[i];
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

__global__ void k_rhs(int n, const double* rho, const double* V, const double* bsrc, double* b, double* partial) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = rho[i] * V[i] / kEps0 + bsrc[i];
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

template <class T>
struct Ops {
    const PoissonDev& P;
    hipStream_t s;
    double* red;
    double* scal;
    const T* a;
    const T* diag;
    const T* rD;
    void precond(const T* r, T* z) const {
        const int K = P.K;
        for (int c = 0; c < P.ncol; ++c) {
            const int c0 = P.col_start_h[c], cnt = P.col_start_h[c + 1] - c0;
            if (cnt > 0)
                hipLaunchKernelGGL((k_dic_fwd<T>), dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, P.corder + c0, c, K,
                                   P.nb, a, P.colour, rD, r, z);
        }
        for (int c = P.ncol - 2; c >= 0; --c) {
            const int c0 = P.col_start_h[c], cnt = P.col_start_h[c + 1] - c0;
            if (cnt > 0)
                hipLaunchKernelGGL((k_dic_bwd<T>), dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, P.corder + c0, c, K,
                                   P.nb, a, P.colour, rD, z);
        }
    }
    void dot(const T* x, const T* y, int slot) const {
        const int nb = nblk(P.n);
        hipLaunchKernelGGL((k_dot<T>), dim3(nb), dim3(256), 0, s, P.n, x, y, red);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, red, nb, scal, slot);
    }
};

// Preconditioned CG on A x = r0 with x starting at the current content of x.
// Stops when sqrt(r.r) <= tol_abs; checks the device residual every `check` iterations
// (the per-iteration norms are recorded so the first converged iteration is reported).
template <class T>
int pcg(const Ops<T>& O, T* x, T* r, T* z, T* p, T* q, double tol_abs, int max_iter, double* rr_hist,
        int* converged_at, double* last_rr) {
    const PoissonDev& P = O.P;
    const int n = P.n, nb = nblk(n);
    hipStream_t s = O.s;
    O.precond(r, z);
    hipLaunchKernelGGL((k_copy<T>), dim3(nb), dim3(256), 0, s, n, z, p);
    O.dot(r, z, 0);
    const int check = 8;
    int it = 0;
    *converged_at = -1;
    std::vector<double> hist(check);
    for (it = 1; it <= max_iter; ++it) {
        hipLaunchKernelGGL((k_spmv_dot<T>), dim3(nb), dim3(256), 0, s, n, P.K, P.nb, O.a, O.diag, p, q, O.red);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, O.red, nb, O.scal, 1);
        hipLaunchKernelGGL(k_alpha, dim3(1), dim3(64), 0, s, O.scal);
        hipLaunchKernelGGL((k_update<T>), dim3(nb), dim3(256), 0, s, n, O.scal, x, r, p, q, O.red);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, O.red, nb, rr_hist, (it - 1) % check);
        if (it % check == 0 || it == max_iter) {
            hipMemcpyAsync(hist.data(), rr_hist, sizeof(double) * check, hipMemcpyDeviceToHost, s);
            hipStreamSynchronize(s);
            const int upto = ((it - 1) % check) + 1;
            for (int k = 0; k < upto; ++k)
                if (std::sqrt(hist[k]) <= tol_abs) { *converged_at = it - upto + k + 1; *last_rr = hist[k]; break; }
            if (*converged_at < 0) *last_rr = hist[upto - 1];
            if (*converged_at >= 0) break;
        }
        O.precond(r, z);
        O.dot(r, z, 3);
        hipLaunchKernelGGL(k_beta, dim3(1), dim3(64), 0, s, O.scal);
        hipLaunchKernelGGL((k_xpby<T>), dim3(nb), dim3(256), 0, s, n, O.scal, z, p);
    }
    return *converged_at >= 0 ? *converged_at : max_iter;
}

}  // namespace

hipError_t poisson_rhs(const PoissonDev& P, const double* rho, double* b, bool singular, double* scratch,
                       hipStream_t s) {
    const int n = P.n, nb = nblk(n);
    double* partial = scratch;
    double* scal = scratch + kRedBlocks;
    hipLaunchKernelGGL(k_rhs, dim3(nb), dim3(256), 0, s, n, rho, P.V, P.bsrc, b, partial);
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
    double* hist = W.scal + 16;
    // ||b||
    {
        Ops<double> O{P, s, W.red, scal, P.a, P.diag, P.rD};
        O.dot(b, b, 6);
    }
    double bb = 0.0;
    hipMemcpyAsync(&bb, scal + 6, sizeof(double), hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    double bn = std::sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    const double tol_abs = tol * bn;
    int total = 0, outer = 0;
    double last_rr = 0.0;
    int status = 0;
    if (mode == 0) {
        hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, n, P.K, P.nb, P.a, P.diag, phi, b, W.r, W.red);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, W.red, nb, scal, 5);
        double rr0 = 0.0;
        hipMemcpyAsync(&rr0, scal + 5, sizeof(double), hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
        last_rr = rr0;
        if (std::sqrt(rr0) > tol_abs) {
            Ops<double> O{P, s, W.red, scal, P.a, P.diag, P.rD};
            int conv = -1;
            total = pcg<double>(O, phi, W.r, W.z, W.p, W.q, tol_abs, max_iter, hist, &conv, &last_rr);
            if (conv < 0) status = -1;
        }
    } else {
        // FP64 iterative refinement around FP32 PCG
        const double inner_rel = 1e-4;
        Ops<float> O{P, s, W.red, scal, P.a_f, P.diag_f, P.rD_f};
        for (outer = 0; outer < 50; ++outer) {
            hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, n, P.K, P.nb, P.a, P.diag, phi, b, W.r, W.red);
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, W.red, nb, scal, 5);
            double rr = 0.0;
            hipMemcpyAsync(&rr, scal + 5, sizeof(double), hipMemcpyDeviceToHost, s);
            hipStreamSynchronize(s);
            last_rr = rr;
            if (std::sqrt(rr) <= tol_abs) break;
            if (total >= max_iter) { status = -1; break; }
            hipLaunchKernelGGL(k_d2f, dim3(nb), dim3(256), 0, s, n, W.r, W.rf);
            hipLaunchKernelGGL((k_zero<float>), dim3(nb), dim3(256), 0, s, n, W.xf);
            int conv = -1;
            double irr = 0.0;
            const double itol = std::max(inner_rel * std::sqrt(rr), 0.5 * tol_abs);
            total += pcg<float>(O, W.xf, W.rf, W.zf, W.pf, W.qf, itol, max_iter - total, hist, &conv, &irr);
            hipLaunchKernelGGL(k_add_f2d, dim3(nb), dim3(256), 0, s, n, W.xf, phi);
        }
    }
    if (singular) {
        hipLaunchKernelGGL(k_sum_partial, dim3(nb), dim3(256), 0, s, n, phi, W.red);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, W.red, nb, scal, 7);
        hipLaunchKernelGGL(k_sub_mean, dim3(nb), dim3(256), 0, s, n, phi, scal, 7);
    }
    if (mode == 0 && total > 0) {
        // report the true FP64 residual
        hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, n, P.K, P.nb, P.a, P.diag, phi, b, W.r, W.red);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, W.red, nb, scal, 5);
        hipMemcpyAsync(&last_rr, scal + 5, sizeof(double), hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
    }
    *iters = total;
    *refinements = outer;
    *rel = std::sqrt(last_rr) / bn;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return -2;
    return status;
}

}  // namespace picmi
