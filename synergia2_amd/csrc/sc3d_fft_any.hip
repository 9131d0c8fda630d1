// sc3d_fft_any.hip -- the Hockney convolution passes for the transform lengths
// the tuned passes of sc3d_fft.hip do not take (those: powers of two, rows of
// gx <= 256 and columns of Dy, Dz <= 512): any length up to 1024, as the
// reference's FFTW plans take any shape (distributed_fft3d_fftw.cc:67-71).
//
// Same inputs, outputs and intermediate layouts as the tuned passes, so the
// two kinds mix per axis (a 48x64x128 grid runs these x passes and the tuned
// y and z passes).  The 1-D transforms are Stockham passes in LDS computed
// one output per thread: output (j/Ns)*Ns*R + j%Ns + r*Ns of a radix-R pass
// is sum_q x[j + q L/R] w_L^(q (k L/(Ns R) + r L/R)), k = j % Ns -- the
// twiddle and the length-R DFT in one table lookup per term, so one routine
// serves every factor, primes included.  Real rows are transformed as full
// complex lines (no packing).  Slower than the tuned passes; the same
// transforms to rounding.
#include "sc3d_green.h"
#include "sc3d_internal.h"

namespace sc3d {

namespace {

// Complex elements of the transform lines of one block (all its lines).
constexpr int kAnyElems = 2048;

// LDS buffer of K line groups of up to kAnyMaxLen: room for the padding
// slot of every line (stride L + 1).
template <int K> struct AnyBuf {
    static constexpr int elems = (kAnyElems > K * kAnyMaxLen ? kAnyElems : K * kAnyMaxLen) + 64;
    static constexpr int per = (elems + kThreads - 1) / kThreads;
};

// lines per block for K groups of length L
inline int any_lines(int L, int K, int cap = 64)
{
    int nl = kAnyElems / (K * L);
    if (nl < 1) nl = 1;
    return nl < cap ? nl : cap;
}

template <typename T2, bool INV, int PER>
__device__ void any_fft(T2* buf, const T2* __restrict__ tw, const AnyPlan& pl, int nlines,
                        int stride)
{
    using F = decltype(T2{}.x);
    const int L = pl.L;
    const int tot = L * nlines;
    int Ns = 1;
    for (int s = 0; s < pl.np; ++s) {
        const int R = pl.R[s];
        const int LR = L / R;
        const int tstep = L / (Ns * R);
        T2 out[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int o = threadIdx.x + q * kThreads;
            if (o < tot) {
                const int line = o / L, oo = o - line * L;
                const int k = oo % Ns, rest = oo / Ns;
                const int r = rest % R, jb = rest / R;
                const T2* lb = buf + line * stride + jb * Ns + k;
                const int base = k * tstep + r * LR;  // < L
                F ax = F(0), ay = F(0);
                int e = 0;
                for (int t = 0; t < R; ++t) {
                    const T2 x = lb[t * LR];
                    const T2 w = tw[e];
                    const F wy = INV ? -w.y : w.y;
                    ax += x.x * w.x - x.y * wy;
                    ay += x.x * wy + x.y * w.x;
                    e += base;
                    if (e >= L) e -= L;
                }
                out[q].x = ax;
                out[q].y = ay;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int o = threadIdx.x + q * kThreads;
            if (o < tot) {
                const int line = o / L, oo = o - line * L;
                buf[line * stride + oo] = out[q];
            }
        }
        __syncthreads();
        Ns *= R;
    }
}

template <typename T2, typename R>
__device__ __forceinline__ T2 times_minus_i(T2 v, R s)  // v * (-i s)
{
    T2 o;
    o.x = s * v.y;
    o.y = -(s * v.x);
    return o;
}

// sin(2 pi k / L) from the length-L table (-Im of e^{-2 pi i k/L}), exactly
// 0 at k = 0 and k = L/2
template <typename T2>
__device__ __forceinline__ auto any_sin(const T2* __restrict__ tw, int k, int L)
{
    return (k == 0 || 2 * k == L) ? decltype(tw[0].y)(0) : -tw[k].y;
}

}  // namespace

// ---------------------------------------------------------------------------
// x pass, forward: rho rows (z < gz, y < gy; gx reals, zero to Dx) -> S rows
// kx = 0..gx.  Precision form: unit-peak scaling on load, as the tuned pass.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kThreads) void ga_x_r2c(const GridJob* __restrict__ jobs,
                                                     const typename cplx<R>::type* __restrict__ tw,
                                                     AnyPlan pl, int nl, int nrows, int gx, int gy,
                                                     int Dy)
{
    using T2 = typename cplx<R>::type;
    __shared__ T2 buf[AnyBuf<1>::elems];
    const GridJob& J = jobs[blockIdx.y];
    const R* __restrict__ rho = static_cast<const R*>(J.rho);
    T2* __restrict__ S = static_cast<T2*>(J.S);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const int L = pl.L, st = L + 1, Hx = gx + 1;
    const int r0 = blockIdx.x * nl;
    const float peak = __int_as_float(dom->rho_peak);
    const bool unit_peak = sizeof(R) == sizeof(float) && peak > 0.0f;
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int line = e / L, x = e - line * L;
        const int r = r0 + line;
        T2 v;
        v.x = R(0);
        v.y = R(0);
        if (x < gx && r < nrows) {
            R a = rho[int64_t(r) * gx + x];
            v.x = unit_peak ? R(a / peak) : a;
        }
        buf[line * st + x] = v;
    }
    __syncthreads();
    any_fft<T2, false, AnyBuf<1>::per>(buf, tw, pl, nl, st);
    for (int e = threadIdx.x; e < nl * Hx; e += kThreads) {
        const int line = e / Hx, k = e - line * Hx;
        const int r = r0 + line;
        if (r >= nrows) continue;
        const int z = r / gy, y = r - z * gy;
        S[(int64_t(z) * Dy + y) * Hx + k] = buf[line * st + k];
    }
}

// x pass, inverse (C2R): S rows -> potential rows x < gx.  The imaginary
// parts of the DC and Nyquist terms are ignored (FFTW c2r); the line is the
// Hermitian extension, the real part of its inverse is kept.
template <typename R>
__global__ __launch_bounds__(kThreads) void ga_x_c2r(const GridJob* __restrict__ jobs,
                                                     const typename cplx<R>::type* __restrict__ tw,
                                                     AnyPlan pl, int nl, int nrows, int gx, int gy,
                                                     int Dy)
{
    using T2 = typename cplx<R>::type;
    __shared__ T2 buf[AnyBuf<1>::elems];
    const GridJob& J = jobs[blockIdx.y];
    const T2* __restrict__ S = static_cast<const T2*>(J.S);
    R* __restrict__ phi = static_cast<R*>(J.phi);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const int L = pl.L, st = L + 1, Hx = gx + 1;
    const int r0 = blockIdx.x * nl;
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int line = e / L, k = e - line * L;
        const int r = r0 + line;
        T2 v;
        v.x = R(0);
        v.y = R(0);
        if (r < nrows) {
            const int z = r / gy, y = r - z * gy;
            const T2* row = S + (int64_t(z) * Dy + y) * Hx;
            if (k <= gx) {
                v = row[k];
                if (k == 0 || k == gx) v.y = R(0);
            } else {
                const T2 c = row[L - k];
                v.x = c.x;
                v.y = -c.y;
            }
        }
        buf[line * st + k] = v;
    }
    __syncthreads();
    any_fft<T2, true, AnyBuf<1>::per>(buf, tw, pl, nl, st);
    for (int e = threadIdx.x; e < nl * gx; e += kThreads) {
        const int line = e / gx, x = e - line * gx;
        const int r = r0 + line;
        if (r < nrows) phi[int64_t(r) * gx + x] = buf[line * st + x].x;
    }
}

// Column passes over S, as k_cols_reg: line l -> (plane = l / lpp, c = l % lpp),
// element p at S[plane*plane_stride + p*pstride + c]; loads p < nload,
// stores p < nstore.  MODE 0 forward, 1 inverse, 2 forward x Green's
// spectrum, inverse (the z pass).
// over: transform this buffer instead of the job's S, and only where h
// changed (the g2 correction spectrum, any_chat).
template <typename R, int MODE>
__global__ __launch_bounds__(kThreads) void ga_cols(const GridJob* __restrict__ jobs,
                                                    const typename cplx<R>::type* __restrict__ tw,
                                                    AnyPlan pl, int nl, int nlines, int lpp,
                                                    int64_t plane_stride, int64_t pstride,
                                                    int nload, int nstore, int gy, int Hx,
                                                    typename cplx<R>::type* over)
{
    using T2 = typename cplx<R>::type;
    __shared__ T2 buf[AnyBuf<1>::elems];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S = over ? over : static_cast<T2*>(J.S);
    const R* __restrict__ ghat = static_cast<const R*>(J.ghat);
    const T2* __restrict__ chat = static_cast<const T2*>(J.chat);
    const DomainRec* dom = J.dom;
    if (dom->error || (over && !dom->ghat_dirty)) return;
    const int L = pl.L, st = L + 1;
    const int l0 = blockIdx.x * nl;
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int l = l0 + line;
        T2 v;
        v.x = R(0);
        v.y = R(0);
        if (p < nload && l < nlines) {
            const int plane = l / lpp, c = l - plane * lpp;
            v = S[plane * plane_stride + p * pstride + c];
        }
        buf[line * st + p] = v;
    }
    __syncthreads();
    any_fft<T2, MODE == 1, AnyBuf<1>::per>(buf, tw, pl, nl, st);
    if constexpr (MODE == 2) {
        const int Dy2 = lpp / Hx;
        for (int e = threadIdx.x; e < nl * L; e += kThreads) {
            const int p = e / nl, line = e - p * nl;
            const int l = l0 + line;
            if (l >= nlines) continue;
            const int ky = l / Hx, kx = l - ky * Hx;
            const int my = ky <= Dy2 / 2 ? ky : Dy2 - ky;
            const int mz = p <= L / 2 ? p : L - p;
            const R gv = ghat[(int64_t(mz) * (gy + 1) + my) * Hx + kx];
            const T2 u = buf[line * st + p];
            T2 v;
            v.x = u.x * gv;
            v.y = u.y * gv;
            if (chat) {
                const T2 c = chat[(int64_t(p) * Dy2 + ky) * Hx + kx];
                v.x += u.x * c.x - u.y * c.y;
                v.y += u.x * c.y + u.y * c.x;
            }
            buf[line * st + p] = v;
        }
        __syncthreads();
        any_fft<T2, true, AnyBuf<1>::per>(buf, tw, pl, nl, st);
    }
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int l = l0 + line;
        if (p < nstore && l < nlines) {
            const int plane = l / lpp, c = l - plane * lpp;
            S[plane * plane_stride + p * pstride + c] = buf[line * st + p];
        }
    }
}

// ---------------------------------------------------------------------------
// Precision form: the field straight from the spectrum, as k_z_field_reg /
// k_y_field_reg / k_x_field (D_a(k) = -i sin(2 pi k / D_a) / h_a).
// z : per (ky, kx) column: forward, x Ghat, inverse of T and of T*D_z,
//     keep z < gz: T -> S, T_z -> S2
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kThreads) void ga_z_field(const GridJob* __restrict__ jobs,
                                                       const typename cplx<R>::type* __restrict__ tw,
                                                       AnyPlan pl, int nl, int nlines, int gz,
                                                       int gy, int Hx)
{
    using T2 = typename cplx<R>::type;
    __shared__ T2 buf[AnyBuf<2>::elems];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S = static_cast<T2*>(J.S);
    T2* __restrict__ S2 = static_cast<T2*>(J.S2);
    const R* __restrict__ ghat = static_cast<const R*>(J.ghat);
    const T2* __restrict__ chat = static_cast<const T2*>(J.chat);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihz = R(dom->inv_h[2]);
    const int L = pl.L, st = L + 1;
    const int l0 = blockIdx.x * nl;
    const int Dy = nlines / Hx;
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int l = l0 + line;
        T2 v;
        v.x = R(0);
        v.y = R(0);
        if (p < gz && l < nlines) v = S[int64_t(p) * nlines + l];
        buf[line * st + p] = v;
    }
    __syncthreads();
    any_fft<T2, false, AnyBuf<2>::per>(buf, tw, pl, nl, st);
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int l = l0 + line;
        R gv = R(0);
        T2 c;
        c.x = R(0);
        c.y = R(0);
        if (l < nlines) {
            const int ky = l / Hx, kx = l - ky * Hx;
            const int my = ky <= Dy / 2 ? ky : Dy - ky;
            const int mz = p <= L / 2 ? p : L - p;
            gv = ghat[(int64_t(mz) * (gy + 1) + my) * Hx + kx];
            if (chat) c = chat[(int64_t(p) * Dy + ky) * Hx + kx];
        }
        const T2 u = buf[line * st + p];
        T2 v;
        v.x = u.x * gv + (u.x * c.x - u.y * c.y);
        v.y = u.y * gv + (u.x * c.y + u.y * c.x);
        buf[line * st + p] = v;
        buf[(nl + line) * st + p] = times_minus_i(v, any_sin(tw, p, L) * ihz);
    }
    __syncthreads();
    any_fft<T2, true, AnyBuf<2>::per>(buf, tw, pl, 2 * nl, st);
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int l = l0 + line;
        if (p < gz && l < nlines) {
            S[int64_t(p) * nlines + l] = buf[line * st + p];
            S2[int64_t(p) * nlines + l] = buf[(nl + line) * st + p];
        }
    }
}

// y': per (z < gz, kx) column: inverse of T, T*D_y and T_z, keep y < gy:
// S rows [0, gy) <- T, S rows [gy, Dy) <- T*D_y, S2 rows [0, gy) <- T_z.
template <typename R>
__global__ __launch_bounds__(kThreads) void ga_y_field(const GridJob* __restrict__ jobs,
                                                       const typename cplx<R>::type* __restrict__ tw,
                                                       AnyPlan pl, int nl, int nlines, int gy,
                                                       int Hx)
{
    using T2 = typename cplx<R>::type;
    __shared__ T2 buf[AnyBuf<3>::elems];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S = static_cast<T2*>(J.S);
    T2* __restrict__ S2 = static_cast<T2*>(J.S2);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihy = R(dom->inv_h[1]);
    const int L = pl.L, st = L + 1;
    const int l0 = blockIdx.x * nl;
    const int64_t plane = int64_t(L) * Hx;
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int l = l0 + line;
        T2 v, vz;
        v.x = R(0);
        v.y = R(0);
        vz = v;
        if (l < nlines) {
            const int z = l / Hx, c = l - z * Hx;
            const int64_t at = z * plane + int64_t(p) * Hx + c;
            v = S[at];
            vz = S2[at];
        }
        buf[line * st + p] = v;
        buf[(nl + line) * st + p] = times_minus_i(v, any_sin(tw, p, L) * ihy);
        buf[(2 * nl + line) * st + p] = vz;
    }
    __syncthreads();
    any_fft<T2, true, AnyBuf<3>::per>(buf, tw, pl, 3 * nl, st);
    for (int e = threadIdx.x; e < nl * gy; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int l = l0 + line;
        if (l < nlines) {
            const int z = l / Hx, c = l - z * Hx;
            const int64_t at = z * plane + int64_t(p) * Hx + c;
            S[at] = buf[line * st + p];
            S[at + int64_t(gy) * Hx] = buf[(nl + line) * st + p];
            S2[at] = buf[(2 * nl + line) * st + p];
        }
    }
}

// x': per (z < gz, y < gy) row: C2R of (S row) * D_x, of S row gy + y and of
// the S2 row -> E (x, y, z, 0) on interior nodes, 0 on the boundary.
template <typename R>
__global__ __launch_bounds__(kThreads) void ga_x_field(const GridJob* __restrict__ jobs,
                                                       const typename cplx<R>::type* __restrict__ tw,
                                                       AnyPlan pl, int nl, int gx, int gy, int gz,
                                                       int Dy)
{
    using T2 = typename cplx<R>::type;
    using T4 = typename vec4_of<R>::type;
    __shared__ T2 buf[AnyBuf<3>::elems];
    const GridJob& J = jobs[blockIdx.y];
    const T2* __restrict__ S = static_cast<const T2*>(J.S);
    const T2* __restrict__ S2 = static_cast<const T2*>(J.S2);
    T4* __restrict__ E = static_cast<T4*>(J.E);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihx = R(dom->inv_h[0]);
    const int L = pl.L, st = L + 1, Hx = gx + 1;
    const int nrows = gz * gy;
    const int r0 = blockIdx.x * nl;
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int line = e / L, k = e - line * L;
        const int r = r0 + line;
        const int z = r / gy, y = r - z * gy;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            T2 v;
            v.x = R(0);
            v.y = R(0);
            if (r < nrows) {
                const T2* row = f == 2 ? S2 + (int64_t(z) * Dy + y) * Hx
                                       : S + (int64_t(z) * Dy + y + (f == 1 ? gy : 0)) * Hx;
                const int kk = k <= gx ? k : L - k;
                v = row[kk];
                if (kk == 0 || kk == gx) v.y = R(0);
                if (f == 0) v = times_minus_i(v, any_sin(tw, kk, L) * ihx);
                if (k > gx) v.y = -v.y;
            }
            buf[(f * nl + line) * st + k] = v;
        }
    }
    __syncthreads();
    any_fft<T2, true, AnyBuf<3>::per>(buf, tw, pl, 3 * nl, st);
    for (int e = threadIdx.x; e < nl * gx; e += kThreads) {
        const int line = e / gx, x = e - line * gx;
        const int r = r0 + line;
        if (r >= nrows) continue;
        const int z = r / gy, y = r - z * gy;
        T4 o;
        o.x = R(0);
        o.y = R(0);
        o.z = R(0);
        o.w = R(0);
        if (x > 0 && x < gx - 1 && y > 0 && y < gy - 1 && z > 0 && z < gz - 1) {
            o.x = buf[line * st + x].x;
            o.y = buf[(nl + line) * st + x].x;
            o.z = buf[(2 * nl + line) * st + x].x;
        }
        E[int64_t(r) * gx + x] = o;
    }
}

// ---------------------------------------------------------------------------
// Green's spectrum, as k_green_x / k_green_cols: rows (Z <= gz, Y <= gy) of
// the doubled grid generated (FP64 values, rounded once to T) and
// transformed along x, keeping the real kx = 0..gx (the row is even);
// then real-even column transforms along y (AXIS 1) and z (AXIS 2) in place
// over the octant, keeping p <= L/2.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void ga_green_x(const GridJob* __restrict__ jobs,
                                                       const typename cplx<T>::type* __restrict__ tw,
                                                       AnyPlan pl, int nl, int gx, int gy, int gz,
                                                       int green_fn)
{
    using T2 = typename cplx<T>::type;
    __shared__ T2 buf[AnyBuf<1>::elems];
    const GridJob& J = jobs[blockIdx.y];
    T* __restrict__ ghat = sizeof(T) == sizeof(double) ? reinterpret_cast<T*>(J.ghat64)
                                                         : static_cast<T*>(J.ghat);
    const DomainRec* dom = J.dom;
    if (dom->error || !dom->ghat_dirty) return;
    const int L = pl.L, st = L + 1, Hx = gx + 1;
    const int nrows = (gz + 1) * (gy + 1);
    const int r0 = blockIdx.x * nl;
    const double hx = dom->h[0], hy = dom->h[1], hz = dom->h[2];
    const double G000 = green_g000(green_fn, hx, hy, hz);
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int line = e / L, X = e - line * L;
        const int r = r0 + line;
        double v = 0.0;
        if (r < nrows) {
            const int Z = r / (gy + 1), Y = r - Z * (gy + 1);
            const int mx = X <= gx ? X : L - X;
            v = green_value(green_fn, mx * hx, Y * hy, Z * hz, hz, G000);
        }
        T2 c;
        c.x = T(v);
        c.y = T(0);
        buf[line * st + X] = c;
    }
    __syncthreads();
    any_fft<T2, false, AnyBuf<1>::per>(buf, tw, pl, nl, st);
    for (int e = threadIdx.x; e < nl * Hx; e += kThreads) {
        const int line = e / Hx, k = e - line * Hx;
        const int r = r0 + line;
        if (r < nrows) ghat[int64_t(r) * Hx + k] = buf[line * st + k].x;
    }
}

template <typename T, int AXIS>
__global__ __launch_bounds__(kThreads) void ga_green_cols(const GridJob* __restrict__ jobs,
                                                          const typename cplx<T>::type* __restrict__ tw,
                                                          AnyPlan pl, int nl, int ncols, int gy,
                                                          int Hx)
{
    using T2 = typename cplx<T>::type;
    __shared__ T2 buf[AnyBuf<1>::elems];
    const GridJob& J = jobs[blockIdx.y];
    T* __restrict__ ghat = sizeof(T) == sizeof(double) ? reinterpret_cast<T*>(J.ghat64)
                                                         : static_cast<T*>(J.ghat);
    const DomainRec* dom = J.dom;
    if (dom->error || !dom->ghat_dirty) return;
    const int L = pl.L, st = L + 1;
    const int c0 = blockIdx.x * nl;
    auto addr = [&](int c, int p) -> int64_t {
        const int plane = c / Hx, kx = c - plane * Hx;
        if (AXIS == 1) return (int64_t(plane) * (gy + 1) + p) * Hx + kx;
        return (int64_t(p) * (gy + 1) + plane) * Hx + kx;
    };
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int c = c0 + line;
        T2 v;
        v.x = c < ncols ? ghat[addr(c, p <= L / 2 ? p : L - p)] : T(0);
        v.y = T(0);
        buf[line * st + p] = v;
    }
    __syncthreads();
    any_fft<T2, false, AnyBuf<1>::per>(buf, tw, pl, nl, st);
    for (int e = threadIdx.x; e < nl * (L / 2 + 1); e += kThreads) {
        const int p = e / nl, line = e - p * nl;
        const int c = c0 + line;
        if (c < ncols) ghat[addr(c, p)] = buf[line * st + p].x;
    }
}

// The g2 correction's x pass: every row (Z < Dz, Y < Dy) of the real
// C[Dz][Dy][Dx], full length, -> chat rows kx = 0..gx; where h changed.
template <typename R>
__global__ __launch_bounds__(kThreads) void ga_chat_x(const GridJob* __restrict__ jobs,
                                                      const typename cplx<R>::type* __restrict__ tw,
                                                      AnyPlan pl, int nl, int nrows, int Hx,
                                                      const R* __restrict__ C,
                                                      typename cplx<R>::type* __restrict__ chat)
{
    using T2 = typename cplx<R>::type;
    __shared__ T2 buf[AnyBuf<1>::elems];
    const DomainRec* dom = jobs[0].dom;
    if (dom->error || !dom->ghat_dirty) return;
    const int L = pl.L, st = L + 1;
    const int r0 = blockIdx.x * nl;
    for (int e = threadIdx.x; e < nl * L; e += kThreads) {
        const int line = e / L, x = e - line * L;
        const int r = r0 + line;
        T2 v;
        v.x = r < nrows ? C[int64_t(r) * L + x] : R(0);
        v.y = R(0);
        buf[line * st + x] = v;
    }
    __syncthreads();
    any_fft<T2, false, AnyBuf<1>::per>(buf, tw, pl, nl, st);
    for (int e = threadIdx.x; e < nl * Hx; e += kThreads) {
        const int line = e / Hx, k = e - line * Hx;
        const int r = r0 + line;
        if (r < nrows) chat[int64_t(r) * Hx + k] = buf[line * st + k];
    }
}

// ---------------------------------------------------------------------------
// host: plans, twiddles, launches
// ---------------------------------------------------------------------------
AnyPlan any_plan(int L)
{
    AnyPlan p{};
    p.L = L;
    int m = L;
    while (m % 4 == 0 && p.np < kAnyMaxPasses) {
        p.R[p.np++] = 4;
        m /= 4;
    }
    for (int f = 2; m > 1 && p.np < kAnyMaxPasses; ++f)
        while (m % f == 0 && p.np < kAnyMaxPasses) {
            p.R[p.np++] = f;
            m /= f;
        }
    if (m != 1) p.np = 0;  // not reachable for L <= kAnyMaxLen
    return p;
}

namespace {
inline int nblk(int64_t items, int per) { return int((items + per - 1) / per); }
}  // namespace

template <typename R>
void any_green_x(const Geometry& g, int green_fn, const GridJob* jobs, int nj,
                 const typename cplx<R>::type* tw, hipStream_t st)
{
    const int nl = any_lines(g.Dx, 1);
    const int rows = (g.gz + 1) * (g.gy + 1);
    hipLaunchKernelGGL(ga_green_x<R>, dim3(nblk(rows, nl), nj), dim3(kThreads), 0, st, jobs,
                       tw + g.tw_any[0], g.plan[0], nl, g.gx, g.gy, g.gz, green_fn);
}

template <typename R>
void any_green_cols(const Geometry& g, int axis, const GridJob* jobs, int nj,
                    const typename cplx<R>::type* tw, hipStream_t st)
{
    const int L = axis == 1 ? g.Dy : g.Dz;
    const int nl = any_lines(L, 1, 16);
    const int cols = (axis == 1 ? (g.gz + 1) : (g.gy + 1)) * g.Hx;
    if (axis == 1)
        hipLaunchKernelGGL((ga_green_cols<R, 1>), dim3(nblk(cols, nl), nj), dim3(kThreads), 0, st,
                           jobs, tw + g.tw_any[1], g.plan[1], nl, cols, g.gy, g.Hx);
    else
        hipLaunchKernelGGL((ga_green_cols<R, 2>), dim3(nblk(cols, nl), nj), dim3(kThreads), 0, st,
                           jobs, tw + g.tw_any[2], g.plan[2], nl, cols, g.gy, g.Hx);
}

template <typename R>
void any_x_pass(const Geometry& g, bool inverse, const GridJob* jobs, int nj,
                const typename cplx<R>::type* tw, hipStream_t st)
{
    const int nl = any_lines(g.Dx, 1);
    const int rows = g.gz * g.gy;
    if (!inverse)
        hipLaunchKernelGGL(ga_x_r2c<R>, dim3(nblk(rows, nl), nj), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[0], g.plan[0], nl, rows, g.gx, g.gy, g.Dy);
    else
        hipLaunchKernelGGL(ga_x_c2r<R>, dim3(nblk(rows, nl), nj), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[0], g.plan[0], nl, rows, g.gx, g.gy, g.Dy);
}

// which: 0 y forward, 1 y inverse, 2 z convolution (the SolveX column passes)
template <typename R>
void any_cols(const Geometry& g, int which, const GridJob* jobs, int nj,
              const typename cplx<R>::type* tw, hipStream_t st)
{
    if (which == 2) {
        const int lines = g.Dy * g.Hx;
        const int nl = any_lines(g.Dz, 1, 16);
        hipLaunchKernelGGL((ga_cols<R, 2>), dim3(nblk(lines, nl), nj), dim3(kThreads), 0, st,
                           jobs, tw + g.tw_any[2], g.plan[2], nl, lines, lines, int64_t(0),
                           int64_t(lines), g.gz, g.gz, g.gy, g.Hx, nullptr);
        return;
    }
    const int lines = g.gz * g.Hx;
    const int nl = any_lines(g.Dy, 1, 16);
    const int64_t plane = int64_t(g.Dy) * g.Hx;
    if (which == 0)
        hipLaunchKernelGGL((ga_cols<R, 0>), dim3(nblk(lines, nl), nj), dim3(kThreads), 0, st,
                           jobs, tw + g.tw_any[1], g.plan[1], nl, lines, g.Hx, plane,
                           int64_t(g.Hx), g.gy, g.Dy, g.gy, g.Hx, nullptr);
    else
        hipLaunchKernelGGL((ga_cols<R, 1>), dim3(nblk(lines, nl), nj), dim3(kThreads), 0, st,
                           jobs, tw + g.tw_any[1], g.plan[1], nl, lines, g.Hx, plane,
                           int64_t(g.Hx), g.Dy, g.gy, g.gy, g.Hx, nullptr);
}

// which: 0 z field, 1 y field, 2 x field (the precision form's field passes)
template <typename R>
void any_field(const Geometry& g, int which, const GridJob* jobs, int nj,
               const typename cplx<R>::type* tw, hipStream_t st)
{
    if (which == 0) {
        const int lines = g.Dy * g.Hx;
        const int nl = any_lines(g.Dz, 2, 16);
        hipLaunchKernelGGL(ga_z_field<R>, dim3(nblk(lines, nl), nj), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[2], g.plan[2], nl, lines, g.gz, g.gy, g.Hx);
    } else if (which == 1) {
        const int lines = g.gz * g.Hx;
        const int nl = any_lines(g.Dy, 3, 16);
        hipLaunchKernelGGL(ga_y_field<R>, dim3(nblk(lines, nl), nj), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[1], g.plan[1], nl, lines, g.gy, g.Hx);
    } else {
        const int rows = g.gz * g.gy;
        const int nl = any_lines(g.Dx, 3);
        hipLaunchKernelGGL(ga_x_field<R>, dim3(nblk(rows, nl), nj), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[0], g.plan[0], nl, g.gx, g.gy, g.gz, g.Dy);
    }
}

template <typename R>
void any_chat(const Geometry& g, const void* C, const GridJob* jobs, void* chat_v,
              const typename cplx<R>::type* tw, hipStream_t st)
{
    using T2 = typename cplx<R>::type;
    T2* chat = static_cast<T2*>(chat_v);
    // x: every row of the doubled grid, full length Dx
    {
        const int nl = any_lines(g.Dx, 1);
        const int rows = g.Dz * g.Dy;
        hipLaunchKernelGGL(ga_chat_x<R>, dim3(nblk(rows, nl)), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[0], g.plan[0], nl, rows, g.Hx,
                           static_cast<const R*>(C), chat);
    }
    // y: columns of Dy in every plane z < Dz
    {
        const int lines = g.Dz * g.Hx;
        const int nl = any_lines(g.Dy, 1, 16);
        hipLaunchKernelGGL((ga_cols<R, 0>), dim3(nblk(lines, nl)), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[1], g.plan[1], nl, lines, g.Hx,
                           int64_t(g.Dy) * g.Hx, int64_t(g.Hx), g.Dy, g.Dy, g.gy, g.Hx, chat);
    }
    // z: columns of Dz for every (ky, kx)
    {
        const int lines = g.Dy * g.Hx;
        const int nl = any_lines(g.Dz, 1, 16);
        hipLaunchKernelGGL((ga_cols<R, 0>), dim3(nblk(lines, nl)), dim3(kThreads), 0, st, jobs,
                           tw + g.tw_any[2], g.plan[2], nl, lines, lines, int64_t(0),
                           int64_t(lines), g.Dz, g.Dz, g.gy, g.Hx, chat);
    }
}

#define SC3D_ANY_INST(R)                                                                       \
    template void any_green_x<R>(const Geometry&, int, const GridJob*, int,                    \
                                 const cplx<R>::type*, hipStream_t);                           \
    template void any_green_cols<R>(const Geometry&, int, const GridJob*, int,                 \
                                    const cplx<R>::type*, hipStream_t);                        \
    template void any_x_pass<R>(const Geometry&, bool, const GridJob*, int,                    \
                                const cplx<R>::type*, hipStream_t);                            \
    template void any_cols<R>(const Geometry&, int, const GridJob*, int, const cplx<R>::type*, \
                              hipStream_t);                                                    \
    template void any_chat<R>(const Geometry&, const void*, const GridJob*, void*,            \
                              const cplx<R>::type*, hipStream_t);
SC3D_ANY_INST(double)
SC3D_ANY_INST(float)
template void any_field<float>(const Geometry&, int, const GridJob*, int, const float2*,
                               hipStream_t);
#undef SC3D_ANY_INST

}  // namespace sc3d
