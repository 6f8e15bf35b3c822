// ipf_solver.hip -- vector kernels of the preconditioned conjugate gradient
// used for the Newton steps (the reference solves with a direct LDL^T,
// main.cpp:283-285; its unused TPZCudaCalls::SolveCG, TPZCudaCalls.cu:213-290,
// is a plain CG).  Dot products are FP64 with a fixed reduction order (block
// partials in order, then one block), so repeated solves are bit-identical.
// CG scalars live in a small device array so an iteration needs no host sync:
//   scal[0] rz = r.z    scal[1] pq = p.q    scal[2] rz_new
//   scal[3] rr = r.r    scal[4] alpha      scal[5] beta
#include "ipf_kernels.h"

namespace ipf {

namespace {

constexpr int kDotThreads = 256;
constexpr int kDotMaxBlocks = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kDotThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

__global__ void __launch_bounds__(kDotThreads)
k_dot_partial(int64_t n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ partial) {
    __shared__ double sh[kDotThreads];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kDotThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kDotThreads)
        s = s + x[i] * y[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kDotThreads)
k_dot_final(int nb, const double* __restrict__ partial, double* __restrict__ out) {
    __shared__ double sh[kDotThreads];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kDotThreads) s = s + partial[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) *out = t;
}

// The preconditioner works on DoF pairs (i, i + 1), i even: z_i = pa_i r_i +
// pb_i r_(i+1), z_(i+1) = pa_(i+1) r_(i+1) + pb_(i+1) r_i -- the inverse of a
// node's 2x2 diagonal block of the tangent, or Jacobi (pb = 0) where no block
// applies.  One thread owns a pair, so each z sees its partner's new r.
__global__ void k_cg_init(int64_t n, const signed char* __restrict__ mask, const double* __restrict__ b,
                          double* __restrict__ r, const double* __restrict__ pa, const double* __restrict__ pb,
                          double* __restrict__ z, double* __restrict__ p) {
    const int64_t i0 = 2 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i0 >= n) return;
    const int64_t i1 = i0 + 1;
    const bool has1 = i1 < n;
    const double r0 = mask[i0] ? 0.0 : b[i0];
    const double r1 = has1 && !mask[i1] ? b[i1] : 0.0;
    const double z0 = pa[i0] * r0 + pb[i0] * r1;
    r[i0] = r0;
    z[i0] = z0;
    p[i0] = z0;
    if (has1) {
        const double z1 = pa[i1] * r1 + pb[i1] * r0;
        r[i1] = r1;
        z[i1] = z1;
        p[i1] = z1;
    }
}

// pa, pb from the tangent's diagonal and (aligned pairs) the symmetric
// coupling offd of a node's x and y DoFs: the 2x2 block's inverse where both
// DoFs are free and the block is positive definite, else 1 / diagonal on a
// free DoF and 0 on a fixed one.
__global__ void k_block_jacobi(int64_t n, const signed char* __restrict__ mask, const double* __restrict__ diag,
                               const double* __restrict__ offd, double* __restrict__ pa, double* __restrict__ pb) {
    const int64_t i0 = 2 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i0 >= n) return;
    const int64_t i1 = i0 + 1;
    const bool has1 = i1 < n;
    const double d0 = diag[i0], d1 = has1 ? diag[i1] : 0.0;
    const bool f0 = !mask[i0] && d0 != 0.0, f1 = has1 && !mask[i1] && d1 != 0.0;
    const double c = (offd && has1) ? offd[i0] : 0.0;
    const double det = d0 * d1 - c * c;
    if (offd && f0 && f1 && d0 > 0.0 && d1 > 0.0 && det > 0.0) {
        pa[i0] = d1 / det;
        pa[i1] = d0 / det;
        pb[i0] = -c / det;
        pb[i1] = -c / det;
    } else {
        pa[i0] = f0 ? 1.0 / d0 : 0.0;
        pb[i0] = 0.0;
        if (has1) {
            pa[i1] = f1 ? 1.0 / d1 : 0.0;
            pb[i1] = 0.0;
        }
    }
}

__global__ void k_cg_mask(int64_t n, const signed char* __restrict__ mask, double* __restrict__ q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && mask[i]) q[i] = 0.0;
}

__global__ void k_cg_update(int64_t n, const double* __restrict__ scal, const double* __restrict__ p,
                            const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
                            const double* __restrict__ pa, const double* __restrict__ pb, double* __restrict__ z) {
    const int64_t i0 = 2 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i0 >= n) return;
    const int64_t i1 = i0 + 1;
    const bool has1 = i1 < n;
    const double alpha = scal[4];
    x[i0] = x[i0] + alpha * p[i0];
    const double r0 = r[i0] - alpha * q[i0];
    r[i0] = r0;
    double r1 = 0.0;
    if (has1) {
        x[i1] = x[i1] + alpha * p[i1];
        r1 = r[i1] - alpha * q[i1];
        r[i1] = r1;
        z[i1] = pa[i1] * r1 + pb[i1] * r0;
    }
    z[i0] = pa[i0] * r0 + pb[i0] * r1;
}

// q = 0 on fixed DoFs, and the block partials of p.q (k_dot_partial's order)
__global__ void __launch_bounds__(kDotThreads)
k_cg_mask_dot(int64_t n, const signed char* __restrict__ mask, const double* __restrict__ p, double* __restrict__ q,
              double* __restrict__ partial) {
    __shared__ double sh[kDotThreads];
    double s = 0.0;
    const int64_t S = (int64_t)gridDim.x * kDotThreads;
    int64_t i = (int64_t)blockIdx.x * kDotThreads + threadIdx.x;
    // batches of U grid strides: every load of the batch in flight at once,
    // then the products added in stride order (the order of the loop below)
    constexpr int U = 4;
    for (; i + (U - 1) * S < n; i += U * S) {
        double pv[U], qv[U];
        signed char mv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            pv[u] = p[i + u * S];
            qv[u] = q[i + u * S];
            mv[u] = mask[i + u * S];
        }
        asm volatile("" ::: "memory");  // the loads above are issued before any use
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (mv[u]) {
                qv[u] = 0.0;
                q[i + u * S] = 0.0;
            }
            s = s + pv[u] * qv[u];
        }
    }
    for (; i < n; i += S) {
        double qi = q[i];
        if (mask[i]) {
            qi = 0.0;
            q[i] = 0.0;
        }
        s = s + p[i] * qi;
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// k_cg_update and the block partials of r.z
__global__ void __launch_bounds__(kDotThreads)
k_cg_update_dot(int64_t n, const double* __restrict__ scal, const double* __restrict__ p,
                const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
                const double* __restrict__ pa, const double* __restrict__ pb, double* __restrict__ z,
                double* __restrict__ partial) {
    __shared__ double sh[kDotThreads];
    const double alpha = scal[4];
    double s = 0.0;
    // one DoF pair per thread and step (see k_cg_init); batches of U whole
    // pairs with every load in flight at once, then the same arithmetic and
    // the same order of the r.z terms as the loop below
    const int64_t S = (int64_t)gridDim.x * kDotThreads;
    int64_t k = (int64_t)blockIdx.x * kDotThreads + threadIdx.x;
    constexpr int U = 2;
    for (; 2 * (k + (U - 1) * S) + 1 < n; k += U * S) {
        double xv[2 * U], pv[2 * U], qv[2 * U], rv[2 * U], av[2 * U], bv[2 * U];
#pragma unroll
        for (int j = 0; j < 2 * U; j++) {
            const int64_t i = 2 * (k + (j >> 1) * S) + (j & 1);
            xv[j] = x[i];
            pv[j] = p[i];
            qv[j] = q[i];
            rv[j] = r[i];
            av[j] = pa[i];
            bv[j] = pb[i];
        }
        asm volatile("" ::: "memory");  // the loads above are issued before any use
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t i0 = 2 * (k + u * S), i1 = i0 + 1;
            const int j0 = 2 * u, j1 = j0 + 1;
            x[i0] = xv[j0] + alpha * pv[j0];
            const double r0 = rv[j0] - alpha * qv[j0];
            r[i0] = r0;
            x[i1] = xv[j1] + alpha * pv[j1];
            const double r1 = rv[j1] - alpha * qv[j1];
            r[i1] = r1;
            const double z0 = av[j0] * r0 + bv[j0] * r1;
            z[i0] = z0;
            s = s + r0 * z0;
            const double z1 = av[j1] * r1 + bv[j1] * r0;
            z[i1] = z1;
            s = s + r1 * z1;
        }
    }
    for (; 2 * k < n; k += S) {
        const int64_t i0 = 2 * k, i1 = i0 + 1;
        const bool has1 = i1 < n;
        x[i0] = x[i0] + alpha * p[i0];
        const double r0 = r[i0] - alpha * q[i0];
        r[i0] = r0;
        double r1 = 0.0;
        if (has1) {
            x[i1] = x[i1] + alpha * p[i1];
            r1 = r[i1] - alpha * q[i1];
            r[i1] = r1;
        }
        const double z0 = pa[i0] * r0 + pb[i0] * r1;
        z[i0] = z0;
        s = s + r0 * z0;
        if (has1) {
            const double z1 = pa[i1] * r1 + pb[i1] * r0;
            z[i1] = z1;
            s = s + r1 * z1;
        }
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = z + beta p
__global__ void k_cg_direction(int64_t n, const double* __restrict__ scal, const double* __restrict__ z,
                               double* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = z[i] + scal[5] * p[i];
}

// The block partials summed as k_dot_final does, then the CG scalars
// (scal: rz, pq, rz_new, ., alpha, beta).  stage 0: pq = the sum, alpha =
// rz / pq;  stage 1: rz_new = the sum, beta = rz_new / rz, rz = rz_new.
__global__ void __launch_bounds__(kDotThreads)
k_cg_final(int nb, const double* __restrict__ partial, double* __restrict__ scal, int stage) {
    __shared__ double sh[kDotThreads];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kDotThreads) s = s + partial[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x != 0) return;
    if (stage == 0) {
        scal[1] = t;
        scal[4] = t != 0.0 ? scal[0] / t : 0.0;
    } else {
        const double rz = scal[0];
        scal[2] = t;
        scal[5] = rz != 0.0 ? t / rz : 0.0;
        scal[0] = t;
    }
}

__global__ void k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

__global__ void k_scale(int64_t n, double a, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = a * x[i];
}

__global__ void k_mask_copy(int64_t n, const signed char* __restrict__ mask, const double* __restrict__ src,
                            double* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = mask[i] ? 0.0 : src[i];
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// Sum of squares of the free entries of x in chunk c = [begin[c], end[c]),
// fixed reduction order.
__global__ void __launch_bounds__(kDotThreads)
k_chunk_sumsq(const double* __restrict__ x, const signed char* __restrict__ mask, const int64_t* __restrict__ begin,
              const int64_t* __restrict__ end, double* __restrict__ partial) {
    __shared__ double sh[kDotThreads];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int64_t i = begin[c] + threadIdx.x; i < end[c]; i += kDotThreads)
        if (!mask[i]) s = s + x[i] * x[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partial[c] = t;
}

// Segment k's chunks [first[k], first[k+1]) summed in order.
__global__ void k_segment_sum(const double* __restrict__ partial, const int* __restrict__ first, int nseg,
                              double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nseg) return;
    double s = 0.0;
    for (int c = first[k]; c < first[k + 1]; c++) s = s + partial[c];
    out[k] = s;
}

}  // namespace

void launch_segment_sumsq(const double* x, const signed char* mask, const int64_t* chunk_begin,
                          const int64_t* chunk_end, int nchunk, double* partial, const int* seg_first_chunk, int nseg,
                          double* out, hipStream_t s) {
    if (nchunk > 0)
        hipLaunchKernelGGL(k_chunk_sumsq, dim3(nchunk), dim3(kDotThreads), 0, s, x, mask, chunk_begin, chunk_end,
                           partial);
    if (nseg > 0)
        hipLaunchKernelGGL(k_segment_sum, dim3((nseg + 255) / 256), dim3(256), 0, s, partial, seg_first_chunk, nseg,
                           out);
}

void launch_cg_mask_dot(int64_t n, const signed char* mask, const double* p, double* q, double* partial, int nb,
                        double* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_mask_dot, dim3(nb), dim3(kDotThreads), 0, s, n, mask, p, q, partial);
    hipLaunchKernelGGL(k_cg_final, dim3(1), dim3(kDotThreads), 0, s, nb, partial, scal, 0);
}

void launch_cg_update_dot(int64_t n, double* scal, const double* p, const double* q, double* x, double* r,
                          const double* pa, const double* pb, double* z, double* partial, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_update_dot, dim3(nb), dim3(kDotThreads), 0, s, n, scal, p, q, x, r, pa, pb, z, partial);
    hipLaunchKernelGGL(k_cg_final, dim3(1), dim3(kDotThreads), 0, s, nb, partial, scal, 1);
}

int dot_blocks(int64_t n) {
    int64_t b = (n + kDotThreads - 1) / kDotThreads;
    if (b > kDotMaxBlocks) b = kDotMaxBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

void launch_dot(int64_t n, const double* x, const double* y, double* partial, int nblocks, double* out,
                hipStream_t s) {
    hipLaunchKernelGGL(k_dot_partial, dim3(nblocks), dim3(kDotThreads), 0, s, n, x, y, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kDotThreads), 0, s, nblocks, partial, out);
}

void launch_cg_init(int64_t n, const signed char* mask, const double* b, double* r, const double* pa,
                    const double* pb, double* z, double* p, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_cg_init, grid1((n + 1) / 2), dim3(256), 0, s, n, mask, b, r, pa, pb, z, p);
}

void launch_block_jacobi(int64_t n, const signed char* mask, const double* diag, const double* offd, double* pa,
                         double* pb, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_block_jacobi, grid1((n + 1) / 2), dim3(256), 0, s, n, mask, diag, offd, pa, pb);
}

void launch_cg_mask(int64_t n, const signed char* mask, double* q, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_cg_mask, grid1(n), dim3(256), 0, s, n, mask, q);
}

void launch_cg_update(int64_t n, const double* scal, const double* p, const double* q, double* x, double* r,
                      const double* pa, const double* pb, double* z, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_cg_update, grid1((n + 1) / 2), dim3(256), 0, s, n, scal, p, q, x, r, pa, pb, z);
}

void launch_cg_direction(int64_t n, const double* scal, const double* z, double* p, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_cg_direction, grid1(n), dim3(256), 0, s, n, scal, z, p);
}


void launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_axpy, grid1(n), dim3(256), 0, s, n, a, x, y);
}

void launch_scale(int64_t n, double a, double* x, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_scale, grid1(n), dim3(256), 0, s, n, a, x);
}

void launch_mask_copy(int64_t n, const signed char* mask, const double* src, double* dst, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_mask_copy, grid1(n), dim3(256), 0, s, n, mask, src, dst);
}

}  // namespace ipf
