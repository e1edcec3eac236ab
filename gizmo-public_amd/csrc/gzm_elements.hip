// gzm_elements.hip -- the device Riemann solver on independent face states,
// one state per thread: the same inline functions the force kernels call
// (riemann_solve, riemann_kt_body), for element-level parity checks of the
// branches that particle sets rarely reach (KT, Roe and PVRS estimates, the
// exact solver's vacuum exits).
#include "grid_device.h"
#include "gzm_internal.h"

namespace gzm {
namespace {

template <typename T>
__global__ __launch_bounds__(kBlock) void riemann_states_kernel(int mode, GammaConsts G, int64_t n, const double *L,
                                                                const double *R, const double *normal,
                                                                const double *limiter, double *out, int *mask)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    RState<T> SL, SR;
    SL.rho = (T)L[5 * t + 0]; SL.p = (T)L[5 * t + 1];
    SR.rho = (T)R[5 * t + 0]; SR.p = (T)R[5 * t + 1];
    T nv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        SL.v[k] = (T)L[5 * t + 2 + k];
        SR.v[k] = (T)R[5 * t + 2 + k];
        nv[k] = (T)normal[3 * t + k];
    }
    ROut<T> o = {0, 0, {0, 0, 0}, 0};
    int m;
    if (mode == 1) {
        // the KT/PWK flux alone, from the face-frame quantities of riemann_solve
        using M = Real<T>;
        const T cL = M::root(M::div(T(G.G0) * SL.p, SL.rho));
        const T cR = M::root(M::div(T(G.G0) * SR.p, SR.rho));
        const T uL = M::div(SL.p, (T(G.G9) * SL.rho));
        const T uR = M::div(SR.p, (T(G.G9) * SR.rho));
        const T hL = M::div(SL.p, SL.rho) + uL + T(0.5) * (SL.v[0] * SL.v[0] + SL.v[1] * SL.v[1] + SL.v[2] * SL.v[2]);
        const T hR = M::div(SR.p, SR.rho) + uR + T(0.5) * (SR.v[0] * SR.v[0] + SR.v[1] * SR.v[1] + SR.v[2] * SR.v[2]);
        const T vL = SL.v[0] * nv[0] + SL.v[1] * nv[1] + SL.v[2] * nv[2];
        const T vR = SR.v[0] * nv[0] + SR.v[1] * nv[1] + SR.v[2] * nv[2];
        o = riemann_kt_body<T>(SL.rho, SL.p, SL.v[0], SL.v[1], SL.v[2], SR.rho, SR.p, SR.v[0], SR.v[1], SR.v[2], nv[0],
                               nv[1], nv[2], vL, vR, cL, cR, hL, hR);
        m = 2;
    } else {
        int err = 0;
        m = riemann_solve<T, false, true>(SL, SR, nv, (T)limiter[t], G, o, err);
        if (err) m = -err;
    }
    double *y = out + 6 * t;
    y[0] = (double)o.P_M; y[1] = (double)o.S_M;
    y[2] = (double)o.Fv[0]; y[3] = (double)o.Fv[1]; y[4] = (double)o.Fv[2];
    y[5] = (double)o.Fp;
    mask[t] = m;
}

}  // namespace

void launch_riemann_states(hipStream_t s, int mode, int precision, const GammaConsts &G, int64_t n, const double *L,
                           const double *R, const double *normal, const double *limiter, double *out, int *mask)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
    if (precision == 0)
        hipLaunchKernelGGL(riemann_states_kernel<double>, grid, dim3(kBlock), 0, s, mode, G, n, L, R, normal, limiter,
                           out, mask);
    else
        hipLaunchKernelGGL(riemann_states_kernel<float>, grid, dim3(kBlock), 0, s, mode, G, n, L, R, normal, limiter,
                           out, mask);
}

}  // namespace gzm
