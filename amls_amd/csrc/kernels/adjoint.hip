// Eigenvalue sensitivities with respect to per-element stiffness and density
// scales.  With K = sum_e s_e P_e^T K0 P_e, M = sum_e r_e P_e^T M0 P_e and an
// M-normalised eigenpair (lambda_j, x_j):
//     d lambda_j / d s_e =            x_je^T K0 x_je
//     d lambda_j / d r_e = -lambda_j  x_je^T M0 x_je
// where x_je is x_j gathered on the element's dofs (zero on constrained
// dofs).  No counterpart in the reference's sources.
//
// One 64-lane workgroup per (element, block of 64 modes): the gathered
// element vectors sit in LDS, one column per lane, and each lane evaluates
// both quadratic forms over the symmetric element matrices, whose entries
// are wave-uniform loads.  Cost per (element, mode): npe^2 + npe FMAs per form.
#include "common.hpp"
#include "launch.hpp"

namespace amls {
namespace dev {

constexpr int ADJ_MAX_NPE = 96;

__global__ void __launch_bounds__(64) adjoint_kernel(int64_t n, int64_t nelem, int npe,
                                                     const int64_t* __restrict__ conn,
                                                     const double* __restrict__ K0,
                                                     const double* __restrict__ M0, int nev,
                                                     const double* __restrict__ theta,
                                                     const double* __restrict__ Z,
                                                     double* __restrict__ gs,
                                                     double* __restrict__ gr) {
    __shared__ double xs[ADJ_MAX_NPE * 64];
    __shared__ int64_t dofs[ADJ_MAX_NPE];
    const int64_t e = blockIdx.x;
    const int lane = threadIdx.x;
    const int j = blockIdx.y * 64 + lane;
    for (int i = lane; i < npe; i += 64) dofs[i] = conn[e * npe + i];
    __syncthreads();
    const bool on = j < nev;
    for (int i = 0; i < npe; ++i) {
        const int64_t g = dofs[i];
        xs[i * 64 + lane] = (on && g >= 0 && g < n) ? Z[g + (int64_t)j * n] : 0.0;
    }
    __syncthreads();
    if (!on) return;
    // x^T A x = sum_i x_i (A_ii x_i + 2 sum_{k<i} A_ik x_k), A symmetric
    double qk = 0.0, qm = 0.0;
    for (int i = 0; i < npe; ++i) {
        const double xi = xs[i * 64 + lane];
        double sk = 0.0, sm = 0.0;
        const double* kr = K0 + (int64_t)i * npe;   // column i (= row i)
        const double* mr = M0 + (int64_t)i * npe;
        for (int k = 0; k < i; ++k) {
            const double xk = xs[k * 64 + lane];
            sk += kr[k] * xk;
            sm += mr[k] * xk;
        }
        qk += xi * (kr[i] * xi + 2.0 * sk);
        qm += xi * (mr[i] * xi + 2.0 * sm);
    }
    if (gs) gs[e + (int64_t)j * nelem] = qk;
    if (gr) gr[e + (int64_t)j * nelem] = -theta[j] * qm;
}

}  // namespace dev

int adjoint_max_npe() { return dev::ADJ_MAX_NPE; }

void launch_adjoint(int64_t n, int64_t nelem, int npe, const int64_t* conn, const double* K0,
                    const double* M0, int nev, const double* theta, const double* Z, double* gs,
                    double* gr, hipStream_t s) {
    if (nelem <= 0 || nev <= 0) return;
    hipLaunchKernelGGL(dev::adjoint_kernel, dim3((unsigned)nelem, (unsigned)((nev + 63) / 64)),
                       dim3(64), 0, s, n, nelem, npe, conn, K0, M0, nev, theta, Z, gs, gr);
}

}  // namespace amls
