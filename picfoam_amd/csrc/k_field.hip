// Grid-side kernels of libpicmi: exclusive scan of cell counts, charge-density
// finalisation, Gauss gradient, cell->point interpolation, diagnostics reductions.
#include "device.h"
#include "kernels.h"

namespace picmi {

// ---------------------------------------------------------------- scan

// block scan of 1024 counts (exclusive, into out) and the block's total into totals[block]
__device__ __forceinline__ void k_scan_block_body(const int* __restrict__ in, int* __restrict__ out, int n,
                                                  int* __restrict__ totals) {
    __shared__ int s[kScanBlock];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + tid;
    int v = (i < n) ? in[i] : 0;
    s[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        int add = (tid >= off) ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += add;
        __syncthreads();
    }
    if (i < n) out[i] = s[tid] - v;
    if (tid == kScanBlock - 1) totals[blockIdx.x] = s[tid];
}

// exclusive scan of nb block totals in place by one block; returns the grand total
__device__ __forceinline__ int k_scan_totals_body(int* totals, int nb) {
    __shared__ int s[kScanBlock];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += kScanBlock) {
        const int tid = threadIdx.x;
        int v = (base + tid < nb) ? totals[base + tid] : 0;
        s[tid] = v;
        __syncthreads();
        for (int off = 1; off < kScanBlock; off <<= 1) {
            int add = (tid >= off) ? s[tid - off] : 0;
            __syncthreads();
            s[tid] += add;
            __syncthreads();
        }
        if (base + tid < nb) totals[base + tid] = carry + s[tid] - v;
        __syncthreads();
        if (tid == kScanBlock - 1) carry += s[tid];
        __syncthreads();
    }
    return carry;
}

// Exclusive scans of the cell counts of every species in three launches (block scans,
// scan of the block totals, offsets added), blockIdx.y the species; the totals pass also
// stores each species' grand total (its live parcel count after the sort).
__global__ void __launch_bounds__(kScanBlock) k_scan_cells_block(ScanSet a) {
    const int sp = blockIdx.y;
    const int nb = (a.n + kScanBlock - 1) / kScanBlock;
    k_scan_block_body(a.in[sp], a.out[sp], a.n, a.tmp + sp * (nb + 1));
}

__global__ void __launch_bounds__(kScanBlock) k_scan_cells_totals(ScanSet a) {
    const int sp = blockIdx.x;
    const int nb = (a.n + kScanBlock - 1) / kScanBlock;
    const int total = k_scan_totals_body(a.tmp + sp * (nb + 1), nb);
    if (threadIdx.x == 0) *a.total[sp] = total;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_cells_add(ScanSet a) {
    const int sp = blockIdx.y;
    const int nb = (a.n + kScanBlock - 1) / kScanBlock;
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    if (i < a.n) a.out[sp][i] += a.tmp[sp * (nb + 1) + blockIdx.x];
}

// Meshes of up to kScanOne cells: one block of 1024 threads per species scans the counts in
// a single launch.  The counts are taken in kScanRounds rounds of 8192; in round r thread t
// holds counts [8192 r + 8 t, 8192 r + 8 t + 8), so a wave's loads cover 2 KiB without gaps,
// and every round's loads are in flight together.  The thread sums of all rounds are
// scanned over the wave with shuffles and over the 16 waves through LDS.  A count at index
// kScanOne (n = kScanOne + 1: the zero past the last cell) is taken by the last thread.
constexpr int kScanChunk = 8;
constexpr int kScanRounds = 4;
constexpr int kScanRound = kScanChunk * 1024;
constexpr int kScanOne = kScanRounds * kScanRound;

__global__ void __launch_bounds__(1024) k_scan_cells_one(ScanSet a) {
    __shared__ int wsum[kScanRounds][16];
    const int sp = blockIdx.x;
    const int* __restrict__ in = a.in[sp];
    int* __restrict__ out = a.out[sp];
    const int n = a.n;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int v[kScanRounds][kScanChunk];
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {
        const int b = r * kScanRound + t * kScanChunk;
        if (b + kScanChunk <= n) {
            const int4 x = *(const int4*)(in + b), y = *(const int4*)(in + b + 4);
            v[r][0] = x.x; v[r][1] = x.y; v[r][2] = x.z; v[r][3] = x.w;
            v[r][4] = y.x; v[r][5] = y.y; v[r][6] = y.z; v[r][7] = y.w;
        } else {
#pragma unroll
            for (int k = 0; k < kScanChunk; ++k) v[r][k] = b + k < n ? in[b + k] : 0;
        }
    }
    int sum[kScanRounds], incl[kScanRounds];
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {
        sum[r] = 0;
#pragma unroll
        for (int k = 0; k < kScanChunk; ++k) sum[r] += v[r][k];
        incl[r] = sum[r];
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int r = 0; r < kScanRounds; ++r) {
            const int y = __shfl_up(incl[r], off);
            if (lane >= off) incl[r] += y;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int r = 0; r < kScanRounds; ++r) wsum[r][w] = incl[r];
    }
    __syncthreads();
    int carry = 0;
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) {
        int before = 0, total = 0;
        for (int k = 0; k < 16; ++k) {
            const int x = wsum[r][k];
            before += k < w ? x : 0;
            total += x;
        }
        int run = carry + before + incl[r] - sum[r];
        const int b = r * kScanRound + t * kScanChunk;
        if (b + kScanChunk <= n) {
            int4 x, y;
            x.x = run; run += v[r][0];
            x.y = run; run += v[r][1];
            x.z = run; run += v[r][2];
            x.w = run; run += v[r][3];
            y.x = run; run += v[r][4];
            y.y = run; run += v[r][5];
            y.z = run; run += v[r][6];
            y.w = run;
            *(int4*)(out + b) = x;
            *(int4*)(out + b + 4) = y;
        } else {
#pragma unroll
            for (int k = 0; k < kScanChunk; ++k) {
                if (b + k < n) out[b + k] = run;
                run += v[r][k];
            }
        }
        carry += total;
    }
    if (t == 1023) {
        if (n > kScanOne) {
            out[kScanOne] = carry;
            carry += in[kScanOne];
        }
        *a.total[sp] = carry;
    }
}

size_t scan_cells_temp_bytes(int n) { return sizeof(int) * kMaxSpecies * (size_t)((n + kScanBlock - 1) / kScanBlock + 1); }

hipError_t launch_scan_cells(const ScanSet& a, int n_species, hipStream_t s) {
    if (a.n <= 0 || n_species <= 0) return hipSuccess;
    if (a.n <= kScanOne + 1) {
        hipLaunchKernelGGL(k_scan_cells_one, dim3(n_species), dim3(1024), 0, s, a);
        return hipGetLastError();
    }
    const int nb = (a.n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_cells_block, dim3(nb, n_species), dim3(kScanBlock), 0, s, a);
    hipLaunchKernelGGL(k_scan_cells_totals, dim3(n_species), dim3(kScanBlock), 0, s, a);
    hipLaunchKernelGGL(k_scan_cells_add, dim3(nb, n_species), dim3(kScanBlock), 0, s, a);
    return hipGetLastError();
}

// zeroes the listed ranges of 4-byte words in one launch (the per-step counters)
__global__ void __launch_bounds__(256) k_zero(ZeroSet z) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int r = 0; r < z.k; ++r) {
        uint32_t* p = (uint32_t*)z.p[r];
        for (int64_t i = t; i < z.words[r]; i += stride) p[i] = 0u;
    }
}

hipError_t launch_zero(const ZeroSet& z, hipStream_t s) {
    int64_t most = 0;
    for (int r = 0; r < z.k; ++r) most = z.words[r] > most ? z.words[r] : most;
    if (most <= 0) return hipSuccess;
    const int64_t nb = (most + 255) / 256;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)(nb < 1024 ? nb : 1024)), dim3(256), 0, s, z);
    return hipGetLastError();
}

// ---------------------------------------------------------------- charge density
// CellAverage.C:66-68 and VolumeWeighting.C:127-164 from the int64 accumulators;
// CellAverage sums are count * llrint(q w/Q0 2^42) per species, the same integers the
// per-parcel accumulation produces.
__global__ void k_deposit_finish(DepositFinishArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_cells) return;
    long long acc;
    if (a.weighting == kCellAverage) {
        acc = 0;
        for (int s = 0; s < a.n_species; ++s)
            if (a.counts[s]) acc += (long long)a.counts[s][c] * a.whole[s];
    } else {
        acc = (long long)a.acc_cell[c];
    }
    const double V = a.V[c];
    double r = ((double)acc * a.q42) / V;
    if (a.weighting == kVolumeWeighting)
        for (int k = a.cpo[c]; k < a.cpo[c + 1]; ++k)
            r = r + a.vw_w[k] * ((double)(long long)a.acc_point[a.cp_master[k]] * a.q42) / V;
    if (a.space_charge) r = r + a.space_charge[c];
    a.rho[c] = r;
}

hipError_t launch_deposit_finish(const DepositFinishArgs& a, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_deposit_finish, dim3((a.n_cells + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- Gauss gradient
// fvc::grad Gauss linear (ElectroStaticSolver.C:68): grad_c = (1/V) sum_f S_f phi_f, faces
// in ascending order; E = -grad.  The face table is column-major, so a wave's loads of one
// slot are contiguous.
__global__ void __launch_bounds__(256) k_grad(GradArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_cells) return;
    const size_t n = (size_t)a.n_cells, kn = (size_t)a.K * n;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const double ps = a.phi[c];
    for (int k = 0; k < a.K; ++k) {
        const size_t e = (size_t)k * n + c;
        const int src = a.src[e];
        if (src == kGradUnused) break;
        double pf;
        if (src >= 0) pf = a.w[e] * ps + a.val[e] * a.phi[src];
        else if (src == kGradFixed) pf = a.val[e];
        else pf = ps;
        gx = gx + a.S[e] * pf;
        gy = gy + a.S[kn + e] * pf;
        gz = gz + a.S[2 * kn + e] * pf;
    }
    const double V = a.V[c];
    const double ex = -(gx / V), ey = -(gy / V), ez = -(gz / V);
    a.E[3 * c] = ex; a.E[3 * c + 1] = ey; a.E[3 * c + 2] = ez;
    if (a.E_f) { a.E_f[3 * c] = (float)ex; a.E_f[3 * c + 1] = (float)ey; a.E_f[3 * c + 2] = (float)ez; }
}

hipError_t launch_grad(const GradArgs& a, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grad, dim3((a.n_cells + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// The explicit part of fvm::laplacian with "Gauss linear corrected" (gaussLaplacianScheme,
// correctedSnGrad::correction): per cell, the sum over its internal and coupled faces of
// |Sf| k . (w grad_self + (1 - w) grad_other), k the face's nonOrthCorrectionVector out of the
// cell.  a.E holds -grad of the potential the solve starts from, so the sum is negated.
__global__ void __launch_bounds__(256) k_nonorth_corr(GradArgs a, double* __restrict__ corr) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_cells) return;
    const size_t n = (size_t)a.n_cells, kn = (size_t)a.K * n;
    const double* mg = a.E;
    const double sx = mg[3 * c], sy = mg[3 * c + 1], sz = mg[3 * c + 2];
    double acc = 0.0;
    for (int k = 0; k < a.K; ++k) {
        const size_t e = (size_t)k * n + c;
        const int o = a.src[e];
        if (o == kGradUnused) break;
        if (o < 0) continue;
        const double w = a.w[e], v = a.val[e];
        const double gx = w * sx + v * mg[3 * o], gy = w * sy + v * mg[3 * o + 1], gz = w * sz + v * mg[3 * o + 2];
        acc = acc + (a.gk[e] * gx + a.gk[kn + e] * gy + a.gk[2 * kn + e] * gz);
    }
    corr[c] = -acc;
}

hipError_t launch_nonorth_corr(const GradArgs& a, double* corr, hipStream_t s) {
    if (a.n_cells <= 0 || !a.gk) return hipSuccess;
    hipLaunchKernelGGL(k_nonorth_corr, dim3((a.n_cells + 255) / 256), dim3(256), 0, s, a, corr);
    return hipGetLastError();
}

// volPointInterpolation: inverse-distance weighted cell values summed over every cell of
// the coupled point group (cyclic images share the master's value).
__global__ void k_point_interp(PointInterpArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_points) return;
    const int m = a.pmaster[p];
    double ex = 0.0, ey = 0.0, ez = 0.0;
    for (int k = a.pco[m]; k < a.pco[m + 1]; ++k) {
        const int c = a.pc_cell[k];
        const double w = a.pc_w[k];
        ex = ex + w * a.Ecell[3 * c];
        ey = ey + w * a.Ecell[3 * c + 1];
        ez = ez + w * a.Ecell[3 * c + 2];
    }
    a.Epoint[3 * p] = ex; a.Epoint[3 * p + 1] = ey; a.Epoint[3 * p + 2] = ez;
    if (a.Epoint_f) { a.Epoint_f[3 * p] = (float)ex; a.Epoint_f[3 * p + 1] = (float)ey; a.Epoint_f[3 * p + 2] = (float)ez; }
}

hipError_t launch_point_interp(const PointInterpArgs& a, hipStream_t s) {
    if (a.n_points <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_point_interp, dim3((a.n_points + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void k_to_float(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

hipError_t launch_to_float(const double* src, float* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- diagnostics
// per species: [0] parcels, [1..3] sum U, [4] sum U.U, [5] sum (gamma - 1), [6..8] sum of
// the squared components, [9..11] sum gamma U (the relativistic momentum per unit mass)
constexpr int kDiagVals = 12;
template <class UT>
__global__ void __launch_bounds__(256) k_diag_partial(PartSoA<UT> p, int64_t n, double* partial) {
    __shared__ double s[kDiagVals][4];
    double v[kDiagVals] = {};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (p.tet[i] < 0) continue;
        const double ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
        const double xx = ux * ux, yy = uy * uy, zz = uz * uz;
        const double u2 = xx + yy + zz;
        const double b = sqrt(u2) / kC;
        const double g = 1.0 / sqrt(1.0 - b * b);
        v[0] += 1.0; v[1] += ux; v[2] += uy; v[3] += uz; v[4] += u2;
        v[5] += g - 1.0;
        v[6] += xx; v[7] += yy; v[8] += zz;
        v[9] += g * ux; v[10] += g * uy; v[11] += g * uz;
    }
    for (int k = 0; k < kDiagVals; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < kDiagVals; ++k) s[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x < kDiagVals) {
        const int k = threadIdx.x;
        partial[(int64_t)blockIdx.x * kDiagVals + k] = (s[k][0] + s[k][1]) + (s[k][2] + s[k][3]);
    }
}

__global__ void k_diag_final(const double* partial, int nb, int nv, double* out) {
    const int k = threadIdx.x;
    if (k >= nv) return;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += partial[(int64_t)b * nv + k];
    out[k] = acc;
}

template <class UT>
static hipError_t launch_diag_t(const PartSoA<UT>& p, int64_t n, double* partial, int nblocks, double* out,
                                hipStream_t s) {
    hipLaunchKernelGGL((k_diag_partial<UT>), dim3(nblocks), dim3(256), 0, s, p, n, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, partial, nblocks, kDiagVals, out);
    return hipGetLastError();
}
hipError_t launch_diag(const PartSoA<double>& p, int64_t n, double* partial, int nblocks, double* out,
                       hipStream_t s) {
    return launch_diag_t(p, n, partial, nblocks, out, s);
}
hipError_t launch_diag(const PartSoA<float>& p, int64_t n, double* partial, int nblocks, double* out,
                       hipStream_t s) {
    return launch_diag_t(p, n, partial, nblocks, out, s);
}

__global__ void __launch_bounds__(256) k_energy_partial(const double* E, const double* V, int n, double* partial) {
    __shared__ double s[256];
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double ex = E[3 * i], ey = E[3 * i + 1], ez = E[3 * i + 2];
        v += (ex * ex + ey * ey + ez * ez) * V[i];
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

hipError_t launch_field_energy(const double* E, const double* V, int n, double* partial, int nblocks, double* out,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_energy_partial, dim3(nblocks), dim3(256), 0, s, E, V, n, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, partial, nblocks, 1, out);
    return hipGetLastError();
}

}  // namespace picmi

namespace picmi {

// ---------------------------------------------------------------- cell moments
// PICCloud::calculateFields (PICCloud.C:587-695) per cell for one species, from the
// cell-sorted arrays: one wave per cell orders the cell's parcels by id in LDS (so the
// sums do not depend on the order the sort left them in) and reduces
//   [0] rhoN = sum w / V, [1] rhoM = sum w m / V, [2] picRhoN = parcel count,
//   [3] linearKE = sum 0.5 w m U.U / V, [4..6] momentum = sum w m U / V,
//   [7..9] j = sum w q U / V, [10] rhoCharge = sum w q / V
template <class UT>
__global__ void __launch_bounds__(64) k_moments(MomentArgs<UT> a) {
    extern __shared__ uint64_t lds_m[];
    uint64_t* key = lds_m;
    int* val = (int*)(lds_m + a.maxp);
    const int cell = blockIdx.x;
    const int f0 = a.start[cell], n = a.start[cell + 1] - f0;
    double* o = a.out + (size_t)cell * kMomentVals;
    if (n > a.maxp) {
        if (threadIdx.x == 0) atomicOr(a.flags, kFlagCellOverflow);
        return;
    }
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int k = threadIdx.x; k < np2; k += blockDim.x) {
        key[k] = (k < n) ? (((uint64_t)a.p.id[f0 + k] << 32) | (uint32_t)k) : ~0ull;
        val[k] = (k < n) ? f0 + k : -1;
    }
    __syncthreads();
    for (int kk = 2; kk <= np2; kk <<= 1)
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < np2; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & kk) == 0;
                    const uint64_t x = key[i], y = key[l];
                    if ((x > y) == up) {
                        key[i] = y; key[l] = x;
                        const int t = val[i]; val[i] = val[l]; val[l] = t;
                    }
                }
            }
            __syncthreads();
        }
    double v[kMomentVals] = {};
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int i = val[k];
        const double ux = a.p.ux[i], uy = a.p.uy[i], uz = a.p.uz[i];
        v[0] += 1.0;
        v[3] += ux * ux + uy * uy + uz * uz;
        v[4] += ux; v[5] += uy; v[6] += uz;
    }
    for (int q = 0; q < kMomentVals; ++q)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
    if (threadIdx.x == 0) {
        const double iV = 1.0 / a.V[cell];
        const double w = a.weight, m = a.mass, qc = a.charge;
        o[0] += v[0] * w * iV;
        o[1] += v[0] * w * m * iV;
        o[2] += v[0];
        o[3] += 0.5 * w * m * v[3] * iV;
        for (int d = 0; d < 3; ++d) {
            o[4 + d] += w * m * v[4 + d] * iV;
            o[7 + d] += w * qc * v[4 + d] * iV;
        }
        o[10] += v[0] * w * qc * iV;
    }
}

template <class UT>
static hipError_t launch_moments_t(const MomentArgs<UT>& a, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    const size_t lds = (size_t)a.maxp * (sizeof(uint64_t) + sizeof(int));
    hipLaunchKernelGGL((k_moments<UT>), dim3(a.n_cells), dim3(64), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_moments(const MomentArgs<double>& a, hipStream_t s) { return launch_moments_t(a, s); }
hipError_t launch_moments(const MomentArgs<float>& a, hipStream_t s) { return launch_moments_t(a, s); }

// the time sums of the picFields averages (the fieldAverage means picFields.C:103-130 reads):
// acc[c][0..5] += rhoN, rhoM, linearKE, momentum of the cell moments m[c]
__global__ void k_fields_accumulate(const double* __restrict__ m, double* __restrict__ acc, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double* r = m + (size_t)c * kMomentVals;
    double* a = acc + (size_t)c * kFieldAvgVals;
    a[0] += r[0];
    a[1] += r[1];
    a[2] += r[3];
    a[3] += r[4];
    a[4] += r[5];
    a[5] += r[6];
}
hipError_t launch_fields_accumulate(const double* moments, double* acc, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fields_accumulate, dim3((n + 255) / 256), dim3(256), 0, s, moments, acc, n);
    return hipGetLastError();
}

}  // namespace picmi
