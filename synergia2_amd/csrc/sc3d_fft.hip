// sc3d_fft.hip -- the doubled-grid Hockney convolution on MI355X.
//
// Replaces the three unpruned 3-D FFTs + complex multiply of the reference
// (Distributed_fft3d::transform / inv_transform over (Dz, Dy, Dx) and
// alg_cplx_multiplier, space_charge_3d_open_hockney.cc:641-676) and the
// Green's function evaluation + its FFT (alg_g2_linear / alg_g2_pointlike,
// :136-348, 591-639) with hand-written LDS Stockham passes:
//
//   x  : R2C of each non-zero rho row (y<gy, z<gz): real row of Dx = 2gx with
//        only x<gx non-zero, packed as gx complex, one length-gx FFT and the
//        real-split post-process                        -> S[z][y][kx]
//   y  : forward length-Dy FFT of each (z<gz, kx) column, rows y>=gy are
//        zero and are never read                        -> S (all Dy rows)
//   z  : per (ky, kx) column: forward length-Dz FFT with planes z>=gz zero,
//        multiply by the real Green's spectrum, inverse FFT, keep z<gz only
//   y' : inverse length-Dy FFT per (z<gz, kx), keep y<gy
//   x' : C2R per (z<gz, y<gy) row, keep x<gx             -> phi octant
//
// Only the gz planes z<gz of the spectrum are ever stored (S is a quarter
// of the reference's padded doubled array).  The convention is the
// reference's (FFTW): forward sign -1, both directions unnormalised, so phi
// equals the reference's phi2 on the octant x<gx, y<gy, z<gz.
//
// The Green's function on the doubled grid is real and even in X, Y and Z
// (mirror writes of alg_g2_*), so its spectrum is real and even in ky, kz:
// it is computed from the (gx+1)(gy+1)(gz+1) octant with three real-even
// passes and stored as that octant only.
#include <type_traits>

#include "sc3d_green.h"
#include "sc3d_internal.h"

namespace sc3d {

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// LDS layout of the transform lines: line l, element p at
// l * stride + p + p/8.  One pad slot every 8 elements spreads the radix-8
// pass that writes with a stride of 8 elements over the banks; the odd
// line stride spreads the column loads that write one element per line.
template <int L> struct LdsLine {
    static constexpr int stride = L + L / 8 + 1;
    static __device__ __forceinline__ int pad(int p) { return p + (p >> 3); }
    static __device__ __forceinline__ int at(int line, int p) { return line * stride + pad(p); }
};

bool tuned_row_length(int gx) { return gx >= 8 && gx <= 256 && (gx & (gx - 1)) == 0; }
bool tuned_col_length(int L) { return L >= 8 && L <= 512 && (L & (L - 1)) == 0; }

// Lines per block: at least 8 so that strided passes move >=128 B runs, and
// at least 1024 elements so every thread owns one radix-4 butterfly.
template <int L> struct Lines {
    static constexpr int value = (1024 / L) > 8 ? (1024 / L) : 8;
};

// Half as many lines (>= 4, >= 512 elements) where more, smaller blocks pay:
// the x passes of the solve and of the field (rows are contiguous, so the
// line count does not set the run width) and the Green's column passes.
// These launches are latency-bound and at 64x64x128 held two to four blocks
// per CU.  Measured (us per kernel): at 64x64x128 FP64, x R2C 6.9 -> 5.8,
// x C2R 5.6 -> 4.9, the two Green's column passes 17.5 -> 16.3; at
// 128x128x256 FP32, the field x pass 50.4 -> 43.3.  k_green_x keeps Lines
// (its FP32 pass at 128x128x256: Green's phase 112 -> 119 us with half).
template <int L> struct HalfLines {
    static constexpr int value = (512 / L) > 4 ? (512 / L) : 4;
};
// The solve's x R2C / C2R passes: half only for rows shorter than 128 (the
// grids with few rows; at 128x128x256 the R2C pass is 18.5 us with Lines,
// 21.6 with half, its 32768 rows giving 4096 blocks already).
template <int M> struct XLines {
    static constexpr int value = M >= 128 ? Lines<M>::value : HalfLines<M>::value;
};

// Column passes: enough lines per block that each strided load or store
// moves a 128-byte run (16 FP32 or 8 FP64 complex values).
template <int L, typename T2> struct ColLines {
    static constexpr int runs = 128 / int(sizeof(T2));
    static constexpr int value = Lines<L>::value > runs ? Lines<L>::value : runs;
};

// Lines per block, from a preferred count, halved (down to 2) while the
// block's LDS (K lines per transform line plus the twiddle table) exceeds
// the cap: the passes are latency-bound at low occupancy, and gfx950 has
// 160 KiB of LDS per CU.
// (Measured at 64x64x128 and 128x128x256: 32 KiB for FP64 lines (24 KiB
// the same, 48 KiB slower); 32 KiB for FP32 lines (field passes 273 us vs
// 280 us at 48 KiB, 303 us at 64 KiB).  16 KiB for both: solve 57 -> 62 us
// at 64x64x128, 265 -> 291 us at 128x128x256.)
#ifndef SC3D_LDS_CAP_FP64
#define SC3D_LDS_CAP_FP64 32768
#endif
#ifndef SC3D_LDS_CAP_FP32
#define SC3D_LDS_CAP_FP32 32768
#endif
template <int L, int K, typename T2, int PREF> struct CappedLines {
    static constexpr size_t cap = sizeof(T2) >= 16 ? SC3D_LDS_CAP_FP64 : SC3D_LDS_CAP_FP32;
    static constexpr size_t bytes(int nl)
    {
        return (size_t(K) * nl * LdsLine<L>::stride + L) * sizeof(T2);
    }
    static constexpr int fit(int nl) { return (nl > 2 && bytes(nl) > cap) ? fit(nl / 2) : nl; }
    static constexpr int value = fit(PREF);
};
// column passes holding K lines per transform line
template <int L, int K, typename T2>
using FieldLines = CappedLines<L, K, T2, ColLines<L, T2>::value>;
// row passes holding K lines per row (rows are contiguous: no run width)
template <int M, int K, typename T2>
using RowLines = CappedLines<M, K, T2, HalfLines<M>::value>;
// the Green's function column passes (two real columns per line)
template <int L, typename T2>
using GreenLines = CappedLines<L, 1, T2, HalfLines<L>::value>;

template <int R, bool INV, typename T2>
__device__ __forceinline__ void dft_small(T2* v)
{
    if constexpr (R == 8) {
        // y_k = E_(k mod 4) + w8^k O_(k mod 4), E/O = DFT4 of the even/odd
        // inputs, w8 = e^(-i pi/4) (inverse: conjugate), w8^(k+4) = -w8^k
        T2 e[4] = {v[0], v[2], v[4], v[6]};
        T2 o[4] = {v[1], v[3], v[5], v[7]};
        dft_small<4, INV>(e);
        dft_small<4, INV>(o);
        using F = decltype(v[0].x);
        const F c = F(0.70710678118654752440);
        T2 t1, t2, t3;
        if (!INV) {
            t1.x = (o[1].x + o[1].y) * c; t1.y = (o[1].y - o[1].x) * c;  // (1 - i) c
            t2.x = o[2].y;                t2.y = -o[2].x;                // -i
            t3.x = (o[3].y - o[3].x) * c; t3.y = -(o[3].x + o[3].y) * c; // (-1 - i) c
        } else {
            t1.x = (o[1].x - o[1].y) * c; t1.y = (o[1].x + o[1].y) * c;  // (1 + i) c
            t2.x = -o[2].y;               t2.y = o[2].x;                 // i
            t3.x = -(o[3].x + o[3].y) * c; t3.y = (o[3].x - o[3].y) * c; // (-1 + i) c
        }
        v[0].x = e[0].x + o[0].x; v[0].y = e[0].y + o[0].y;
        v[4].x = e[0].x - o[0].x; v[4].y = e[0].y - o[0].y;
        v[1].x = e[1].x + t1.x;   v[1].y = e[1].y + t1.y;
        v[5].x = e[1].x - t1.x;   v[5].y = e[1].y - t1.y;
        v[2].x = e[2].x + t2.x;   v[2].y = e[2].y + t2.y;
        v[6].x = e[2].x - t2.x;   v[6].y = e[2].y - t2.y;
        v[3].x = e[3].x + t3.x;   v[3].y = e[3].y + t3.y;
        v[7].x = e[3].x - t3.x;   v[7].y = e[3].y - t3.y;
    } else if constexpr (R == 2) {
        T2 a = v[0], b = v[1];
        v[0].x = a.x + b.x; v[0].y = a.y + b.y;
        v[1].x = a.x - b.x; v[1].y = a.y - b.y;
    } else {
        T2 a, b, c, d;
        a.x = v[0].x + v[2].x; a.y = v[0].y + v[2].y;
        b.x = v[0].x - v[2].x; b.y = v[0].y - v[2].y;
        c.x = v[1].x + v[3].x; c.y = v[1].y + v[3].y;
        d.x = v[1].x - v[3].x; d.y = v[1].y - v[3].y;
        v[0].x = a.x + c.x; v[0].y = a.y + c.y;
        v[2].x = a.x - c.x; v[2].y = a.y - c.y;
        if (!INV) {  // y1 = b - i d, y3 = b + i d
            v[1].x = b.x + d.y; v[1].y = b.y - d.x;
            v[3].x = b.x - d.y; v[3].y = b.y + d.x;
        } else {     // y1 = b + i d, y3 = b - i d
            v[1].x = b.x - d.y; v[1].y = b.y + d.x;
            v[3].x = b.x + d.y; v[3].y = b.y - d.x;
        }
    }
}

// One Stockham autosort pass of radix R on NL lines of length L in LDS
// (Govindaraju et al., SC'08 formulation): butterfly j reads j + r*L/R,
// twiddles by w_L^{r*(j mod Ns)*L/(Ns*R)}, writes (j/Ns)*Ns*R + j mod Ns +
// r*Ns.  All reads of the pass complete before any write (in place).
template <typename T2, int R, int L, int NL, bool INV>
__device__ __forceinline__ void stockham_pass(T2* buf, const T2* __restrict__ tw,
                                              int lns)
{
    constexpr int NB = L / R;
    constexpr int TOT = NB * NL;
    constexpr int BPT = (TOT + kThreads - 1) / kThreads;
    constexpr int LL = ilog2(L), LR = ilog2(R);
    const int Ns = 1 << lns;
    T2 v[BPT][R];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        int b = threadIdx.x + q * kThreads;
        if (TOT % kThreads == 0 || b < TOT) {
            int line = b / NB, j = b % NB;
            const T2* lb = buf + line * LdsLine<L>::stride;
            int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = lb[LdsLine<L>::pad(j + r * NB)];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                T2 w = tw[(r * k) << (LL - lns - LR)];
                if (INV) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
            dft_small<R, INV>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        int b = threadIdx.x + q * kThreads;
        if (TOT % kThreads == 0 || b < TOT) {
            int line = b / NB, j = b % NB;
            T2* lb = buf + line * LdsLine<L>::stride;
            int k = j & (Ns - 1);
            int d = ((j >> lns) << (lns + LR)) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lb[LdsLine<L>::pad(d + r * Ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// NL unnormalised length-L FFTs in LDS; forward sign -1 (INV: +1).  Radix-8
// passes, then one radix-4 or radix-2 pass for the remaining factor.  tw is
// the length-L twiddle table, normally the block's LDS copy (load_twiddles).
// The caller has synchronised after filling buf.
template <typename T2, int L, int NL, bool INV>
__device__ __forceinline__ void lds_fft(T2* buf, const T2* __restrict__ tw)
{
    constexpr int LOG = ilog2(L);
    int lns = 0;
#pragma unroll
    for (int p = 0; p < LOG / 3; ++p, lns += 3)
        stockham_pass<T2, 8, L, NL, INV>(buf, tw, lns);
    if constexpr (LOG % 3 == 2) stockham_pass<T2, 4, L, NL, INV>(buf, tw, lns);
    if constexpr (LOG % 3 == 1) stockham_pass<T2, 2, L, NL, INV>(buf, tw, lns);
}

// The block's LDS copy of a length-L twiddle table (filled before the
// __syncthreads that precedes the first lds_fft).
template <typename T2, int L>
__device__ __forceinline__ void load_twiddles(T2* twl, const T2* __restrict__ tw)
{
    for (int e = threadIdx.x; e < L; e += kThreads) twl[e] = tw[e];
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx.x mod 8),
// each with its own L2.  Column passes read runs of a few lines per block,
// and neighbouring blocks share 128-byte lines, so the block index is
// remapped to give every XCD a contiguous range of line groups.  The
// launchers round the grid up to a multiple of 8 (the extra groups find no
// lines and store nothing).
#ifndef SC3D_XCD_REMAP
#define SC3D_XCD_REMAP 1
#endif
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_group()
{
#if SC3D_XCD_REMAP
    const int per = int(gridDim.x) / kXcds;
    return int(blockIdx.x % kXcds) * per + int(blockIdx.x / kXcds);
#else
    return int(blockIdx.x);
#endif
}

// ---------------------------------------------------------------------------
// x pass, forward: real rows of the rho octant -> half spectrum rows of S.
// rows r in [0, gz*gy); input row r is rho + r*gx (gx = M reals); the
// doubled row has Dx = 2M reals, so packed complex z[n] = (x[2n], x[2n+1])
// is non-zero only for n < M/2.
// ---------------------------------------------------------------------------
template <typename R, int M>
__global__ __launch_bounds__(kThreads) void k_x_r2c(const GridJob* __restrict__ jobs,
                                                    const typename cplx<R>::type* __restrict__ twM,
                                                    const typename cplx<R>::type* __restrict__ tw2M,
                                                    int nrows, int gy, int Dy)
{
    using T2 = typename cplx<R>::type;
    constexpr int NL = XLines<M>::value;
    __shared__ T2 buf[NL * LdsLine<M>::stride];
    __shared__ T2 twl[M];
    const GridJob& J = jobs[blockIdx.y];
    const R* __restrict__ rho = static_cast<const R*>(J.rho);
    T2* __restrict__ S = static_cast<T2*>(J.S);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const int r0 = blockIdx.x * NL;
    const int Hx = M + 1;
    // precision form: the unit-weight density is scaled to unit peak here,
    // before any transform arithmetic (the peak returns at the kick)
    const float peak = __int_as_float(dom->rho_peak);
    const bool unit_peak = sizeof(R) == sizeof(float) && peak > 0.0f;
    // load: NL consecutive rows are NL*M contiguous reals = NL*M/2 complex
    const T2* src = reinterpret_cast<const T2*>(rho + int64_t(r0) * M);
    for (int e = threadIdx.x; e < NL * M; e += kThreads) {
        int line = e / M, n = e % M;
        T2 v;
        if (n < M / 2 && r0 + line < nrows) {
            v = src[line * (M / 2) + n];
            if (unit_peak) {
                v.x = R(v.x / peak);
                v.y = R(v.y / peak);
            }
        } else {
            v.x = R(0); v.y = R(0);
        }
        buf[LdsLine<M>::at(line, n)] = v;
    }
    load_twiddles<T2, M>(twl, twM);
    __syncthreads();
    lds_fft<T2, M, NL, false>(buf, twl);
    // post-process: X[k] = 1/2 (A + (-i) w^k B), A = Z[k] + conj Z[M-k],
    // B = Z[k] - conj Z[M-k], w = e^{-2 pi i / 2M}; k = 0..M.
    for (int e = threadIdx.x; e < NL * Hx; e += kThreads) {
        int line = e / Hx, k = e % Hx;
        int r = r0 + line;
        if (r >= nrows) continue;
        T2 zk = buf[LdsLine<M>::at(line, (k == M ? 0 : k))];
        T2 zm = buf[LdsLine<M>::at(line, (k == 0 ? 0 : M - k))];
        T2 A, B;
        A.x = zk.x + zm.x; A.y = zk.y - zm.y;
        B.x = zk.x - zm.x; B.y = zk.y + zm.y;
        T2 Cc = cmul(tw2M[k], B);
        T2 X;
        X.x = R(0.5) * (A.x + Cc.y);
        X.y = R(0.5) * (A.y - Cc.x);
        int z = r / gy, y = r - z * gy;
        S[(int64_t(z) * Dy + y) * Hx + k] = X;
    }
}

// ---------------------------------------------------------------------------
// x pass, inverse (C2R): rows r of S (z<gz, y<gy) -> phi octant rows (gx=M).
// The imaginary parts of the DC and Nyquist terms are ignored (FFTW c2r).
// Z'[k] = A + i conj(w^k) B, A = X[k] + conj X[M-k], B = X[k] - conj X[M-k];
// z = inverse FFT_M(Z'); x[2n] = Re z[n], x[2n+1] = Im z[n]; keep n < M/2.
// ---------------------------------------------------------------------------
template <typename R, int M>
__global__ __launch_bounds__(kThreads) void k_x_c2r(const GridJob* __restrict__ jobs,
                                                    const typename cplx<R>::type* __restrict__ twM,
                                                    const typename cplx<R>::type* __restrict__ tw2M,
                                                    int nrows, int gy, int Dy)
{
    using T2 = typename cplx<R>::type;
    constexpr int NL = XLines<M>::value;
    __shared__ T2 buf[NL * LdsLine<M>::stride];
    __shared__ T2 twl[M];
    const GridJob& J = jobs[blockIdx.y];
    const T2* __restrict__ S = static_cast<const T2*>(J.S);
    R* __restrict__ phi = static_cast<R*>(J.phi);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const int r0 = blockIdx.x * NL;
    const int Hx = M + 1;
    for (int e = threadIdx.x; e < NL * M; e += kThreads) {
        int line = e / M, k = e % M;
        int r = r0 + line;
        T2 v;
        if (r < nrows) {
            int z = r / gy, y = r - z * gy;
            const T2* row = S + (int64_t(z) * Dy + y) * Hx;
            T2 xk = row[k];
            T2 xm = row[M - k];
            if (k == 0) { xk.y = R(0); xm.y = R(0); }
            T2 A, B;
            A.x = xk.x + xm.x; A.y = xk.y - xm.y;
            B.x = xk.x - xm.x; B.y = xk.y + xm.y;
            T2 Cc = cmul_conj(B, tw2M[k]);  // conj(w^k) * B
            v.x = A.x - Cc.y;
            v.y = A.y + Cc.x;
        } else {
            v.x = R(0); v.y = R(0);
        }
        buf[LdsLine<M>::at(line, k)] = v;
    }
    load_twiddles<T2, M>(twl, twM);
    __syncthreads();
    lds_fft<T2, M, NL, true>(buf, twl);
    T2* dst = reinterpret_cast<T2*>(phi + int64_t(r0) * M);
    for (int e = threadIdx.x; e < NL * (M / 2); e += kThreads) {
        int line = e / (M / 2), n = e % (M / 2);
        if (r0 + line < nrows) dst[line * (M / 2) + n] = buf[LdsLine<M>::at(line, n)];
    }
}

// ---------------------------------------------------------------------------
// Column passes over S.  Line l -> (plane = l / lpp, c = l % lpp), element p
// at S[plane*plane_stride + p*pstride + c].  Loads p < nload (rest zero),
// stores p < nstore.  MODE 0: forward; 1: inverse; 2: forward, multiply by
// the real Green's spectrum, inverse (the z pass).
// ---------------------------------------------------------------------------
template <typename R, int L, int MODE>
__global__ __launch_bounds__(kThreads) void k_cols(const GridJob* __restrict__ jobs,
                                                   const typename cplx<R>::type* __restrict__ tw,
                                                   int nlines, int lpp, int64_t plane_stride,
                                                   int64_t pstride, int nload, int nstore, int gy,
                                                   int Hx)
{
    using T2 = typename cplx<R>::type;
    constexpr int NL = FieldLines<L, 1, T2>::value;
    __shared__ T2 buf[NL * LdsLine<L>::stride];
    __shared__ T2 twl[L];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S = static_cast<T2*>(J.S);
    const R* __restrict__ ghat = static_cast<const R*>(J.ghat);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const int l0 = xcd_group() * NL;
    for (int e = threadIdx.x; e < NL * L; e += kThreads) {
        int p = e / NL, line = e % NL;
        int l = l0 + line;
        T2 v;
        if (p < nload && l < nlines) {
            int plane = l / lpp, c = l - plane * lpp;
            v = S[plane * plane_stride + p * pstride + c];
        } else {
            v.x = R(0); v.y = R(0);
        }
        buf[LdsLine<L>::at(line, p)] = v;
    }
    load_twiddles<T2, L>(twl, tw);
    __syncthreads();
    if constexpr (MODE == 1) {
        lds_fft<T2, L, NL, true>(buf, twl);
    } else {
        lds_fft<T2, L, NL, false>(buf, twl);
    }
    if constexpr (MODE == 2) {
        // line l = ky*Hx + kx over the full Dy rows; element p = kz.
        const int Dy2 = lpp / Hx;  // Dy
        for (int e = threadIdx.x; e < NL * L; e += kThreads) {
            int p = e / NL, line = e % NL;
            int l = l0 + line;
            if (l >= nlines) continue;
            int ky = l / Hx, kx = l - ky * Hx;
            int my = ky <= Dy2 / 2 ? ky : Dy2 - ky;
            int mz = p <= L / 2 ? p : L - p;
            R gv = ghat[(int64_t(mz) * (gy + 1) + my) * Hx + kx];
            T2 v = buf[LdsLine<L>::at(line, p)];
            v.x = v.x * gv;
            v.y = v.y * gv;
            buf[LdsLine<L>::at(line, p)] = v;
        }
        __syncthreads();
        lds_fft<T2, L, NL, true>(buf, twl);
    }
    for (int e = threadIdx.x; e < NL * L; e += kThreads) {
        int p = e / NL, line = e % NL;
        int l = l0 + line;
        if (p < nstore && l < nlines) {
            int plane = l / lpp, c = l - plane * lpp;
            S[plane * plane_stride + p * pstride + c] = buf[LdsLine<L>::at(line, p)];
        }
    }
}

// ---------------------------------------------------------------------------
// Precision form: the field straight from the spectrum.  The reference's
// centred difference E_a[i] = -(phi[i+1] - phi[i-1]) * (0.5/h_a) on interior
// nodes only reaches octant values, so on the doubled (periodic) grid it is
// exactly the spectral multiplier
//     D_a(k) = -i sin(2 pi k / D_a) / h_a,     E_a = IDFT(phi^ * D_a).
// In FP32 the three fields are transformed back from the spectrum instead of
// differencing a rounded potential: the inverse transforms' rounding is then
// relative to the field itself and not to the much larger potential, which
// keeps the field ~30x closer to the FP64 one.  sin(2 pi k / L) is -Im of
// the twiddle e^{-2 pi i k / L} of the length-L table.
//   z : per (ky, kx) column: forward, x Ghat, inverse of T and of T*D_z,
//       keep z < gz: T -> S, T_z -> S2
//   y': per (z, kx) column: inverse of T, T*D_y and T_z, keep y < gy:
//       S rows [0, gy) <- T, S rows [gy, Dy) <- T*D_y, S2 rows [0, gy) <- T_z
//   x': per (z, y) row: C2R of (S row)*D_x, S row gy+y, S2 row -> E (AoS,
//       interior nodes; boundary nodes 0, as the reference leaves them)
// ---------------------------------------------------------------------------


// v * (-i s) = (s v.y, -s v.x)
template <typename T2, typename R>
__device__ __forceinline__ T2 mul_minus_i(T2 v, R s)
{
    T2 o;
    o.x = s * v.y;
    o.y = -(s * v.x);
    return o;
}

// sin(2 pi k / L) from the length-L twiddle table, exactly 0 at k = 0, L/2
template <typename T2>
__device__ __forceinline__ auto sin_2pi(const T2* __restrict__ tw, int k, int L)
{
    return (k == 0 || 2 * k == L) ? decltype(tw[0].y)(0) : -tw[k].y;
}

// The z pass holds one LDS line per column: the Green-multiplied spectrum
// stays in registers (each thread keeps the elements it multiplied) for the
// second inverse transform, and the Green values are fetched before the
// forward transform so that both global loads are in flight together.
// Lines per block: at most SC3D_ZF_LINES (measured at 128x128x256: 4 lines
// 287 us, 8 lines 323 us, 16 lines 424 us for the FP32 field passes; the
// pass is latency-bound and more lines cost occupancy; later, with the
// current passes, 2 lines 287 us against 266 with 4: 16-byte runs).
#ifndef SC3D_ZF_LINES
#define SC3D_ZF_LINES 4
#endif
template <int L, typename T2>
struct ZFieldLines {
    static constexpr int capped = FieldLines<L, 1, T2>::value;
    static constexpr int value = capped < SC3D_ZF_LINES ? capped : SC3D_ZF_LINES;
};

template <typename R, int L>
__global__ __launch_bounds__(kThreads) void k_z_field(const GridJob* __restrict__ jobs,
                                                      const typename cplx<R>::type* __restrict__ tw,
                                                      int nlines, int gz, int gy, int Hx)
{
    using T2 = typename cplx<R>::type;
    constexpr int NL = ZFieldLines<L, T2>::value;
    constexpr int PER = (NL * L + kThreads - 1) / kThreads;
    __shared__ T2 buf[NL * LdsLine<L>::stride];
    __shared__ T2 twl[L];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S = static_cast<T2*>(J.S);
    T2* __restrict__ S2 = static_cast<T2*>(J.S2);
    const R* __restrict__ ghat = static_cast<const R*>(J.ghat);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihz = R(dom->inv_h[2]);
    const int l0 = xcd_group() * NL;
    const int Dy = nlines / Hx;
    // Green values of the elements this thread multiplies (element e =
    // threadIdx.x + q * kThreads: kz = e / NL, line = e % NL)
    R gv[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * kThreads;
        const int p = e / NL, l = l0 + e % NL;
        gv[q] = R(0);
        if (e < NL * L && l < nlines) {
            const int ky = l / Hx, kx = l - ky * Hx;
            const int my = ky <= Dy / 2 ? ky : Dy - ky;
            const int mz = p <= L / 2 ? p : L - p;
            gv[q] = ghat[(int64_t(mz) * (gy + 1) + my) * Hx + kx];
        }
    }
    for (int e = threadIdx.x; e < NL * L; e += kThreads) {
        int p = e / NL, line = e % NL;
        int l = l0 + line;
        T2 v;
        if (p < gz && l < nlines) {
            v = S[int64_t(p) * nlines + l];
        } else {
            v.x = R(0); v.y = R(0);
        }
        buf[LdsLine<L>::at(line, p)] = v;
    }
    load_twiddles<T2, L>(twl, tw);
    __syncthreads();
    lds_fft<T2, L, NL, false>(buf, twl);
    T2 keep[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * kThreads;
        if (e < NL * L) {
            const int p = e / NL, line = e % NL;
            T2 v = buf[LdsLine<L>::at(line, p)];
            v.x = v.x * gv[q];
            v.y = v.y * gv[q];
            buf[LdsLine<L>::at(line, p)] = v;
            keep[q] = v;
        }
    }
    __syncthreads();
    lds_fft<T2, L, NL, true>(buf, twl);
    for (int e = threadIdx.x; e < NL * L; e += kThreads) {
        int p = e / NL, line = e % NL;
        int l = l0 + line;
        if (p < gz && l < nlines) S[int64_t(p) * nlines + l] = buf[LdsLine<L>::at(line, p)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * kThreads;
        if (e < NL * L) {
            const int p = e / NL, line = e % NL;
            buf[LdsLine<L>::at(line, p)] = mul_minus_i(keep[q], sin_2pi(twl, p, L) * ihz);
        }
    }
    __syncthreads();
    lds_fft<T2, L, NL, true>(buf, twl);
    for (int e = threadIdx.x; e < NL * L; e += kThreads) {
        int p = e / NL, line = e % NL;
        int l = l0 + line;
        if (p < gz && l < nlines) S2[int64_t(p) * nlines + l] = buf[LdsLine<L>::at(line, p)];
    }
}

template <typename R, int L>
__global__ __launch_bounds__(kThreads) void k_y_field(const GridJob* __restrict__ jobs,
                                                      const typename cplx<R>::type* __restrict__ tw,
                                                      int nlines, int gy, int Hx)
{
    using T2 = typename cplx<R>::type;
    constexpr int NL = FieldLines<L, 3, T2>::value;
    __shared__ T2 buf[3 * NL * LdsLine<L>::stride];
    __shared__ T2 twl[L];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S = static_cast<T2*>(J.S);
    T2* __restrict__ S2 = static_cast<T2*>(J.S2);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihy = R(dom->inv_h[1]);
    const int l0 = xcd_group() * NL;
    const int64_t plane = int64_t(L) * Hx;  // L = Dy rows of Hx
    for (int e = threadIdx.x; e < NL * L; e += kThreads) {
        int p = e / NL, line = e % NL;
        int l = l0 + line;
        T2 v, vz;
        if (l < nlines) {
            int z = l / Hx, c = l - z * Hx;
            int64_t at = z * plane + int64_t(p) * Hx + c;
            v = S[at];
            vz = S2[at];
        } else {
            v.x = R(0); v.y = R(0);
            vz = v;
        }
        buf[LdsLine<L>::at(line, p)] = v;
        buf[LdsLine<L>::at((NL + line), p)] = mul_minus_i(v, sin_2pi(tw, p, L) * ihy);
        buf[LdsLine<L>::at((2 * NL + line), p)] = vz;
    }
    load_twiddles<T2, L>(twl, tw);
    __syncthreads();
    lds_fft<T2, L, 3 * NL, true>(buf, twl);
    for (int e = threadIdx.x; e < NL * gy; e += kThreads) {
        int p = e / NL, line = e % NL;
        int l = l0 + line;
        if (l < nlines) {
            int z = l / Hx, c = l - z * Hx;
            int64_t at = z * plane + int64_t(p) * Hx + c;
            S[at] = buf[LdsLine<L>::at(line, p)];
            S[at + int64_t(gy) * Hx] = buf[LdsLine<L>::at((NL + line), p)];
            S2[at] = buf[LdsLine<L>::at((2 * NL + line), p)];
        }
    }
}

template <typename R, int M>
__global__ __launch_bounds__(kThreads) void k_x_field(
    const GridJob* __restrict__ jobs, const typename cplx<R>::type* __restrict__ twM,
    const typename cplx<R>::type* __restrict__ tw2M, int gy, int gz, int Dy)
{
    using T2 = typename cplx<R>::type;
    using T4 = typename vec4_of<R>::type;
    constexpr int NL = RowLines<M, 3, T2>::value;
    __shared__ T2 buf[3 * NL * LdsLine<M>::stride];
    __shared__ T2 twl[M];
    const GridJob& J = jobs[blockIdx.y];
    const T2* __restrict__ S = static_cast<const T2*>(J.S);
    const T2* __restrict__ S2 = static_cast<const T2*>(J.S2);
    T4* __restrict__ E = static_cast<T4*>(J.E);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihx = R(dom->inv_h[0]);
    const int nrows = gz * gy;
    const int r0 = blockIdx.x * NL;
    const int Hx = M + 1;
    for (int e = threadIdx.x; e < NL * M; e += kThreads) {
        int line = e / M, k = e % M;
        int r = r0 + line;
        int z = r / gy, y = r - z * gy;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            T2 v;
            if (r < nrows) {
                const T2* row = f == 2 ? S2 + (int64_t(z) * Dy + y) * Hx
                                       : S + (int64_t(z) * Dy + y + (f == 1 ? gy : 0)) * Hx;
                T2 xk = row[k];
                T2 xm = row[M - k];
                if (k == 0) { xk.y = R(0); xm.y = R(0); }
                if (f == 0) {  // x derivative, D_x(k) for k and M-k (2M = Dx)
                    xk = mul_minus_i(xk, sin_2pi(tw2M, k, 2 * M) * ihx);
                    xm = mul_minus_i(xm, sin_2pi(tw2M, M - k, 2 * M) * ihx);
                }
                T2 A, B;
                A.x = xk.x + xm.x; A.y = xk.y - xm.y;
                B.x = xk.x - xm.x; B.y = xk.y + xm.y;
                T2 Cc = cmul_conj(B, tw2M[k]);
                v.x = A.x - Cc.y;
                v.y = A.y + Cc.x;
            } else {
                v.x = R(0); v.y = R(0);
            }
            buf[LdsLine<M>::at((f * NL + line), k)] = v;
        }
    }
    load_twiddles<T2, M>(twl, twM);
    __syncthreads();
    lds_fft<T2, M, 3 * NL, true>(buf, twl);
    for (int e = threadIdx.x; e < NL * M; e += kThreads) {
        int line = e / M, x = e % M;
        int r = r0 + line;
        if (r >= nrows) continue;
        int z = r / gy, y = r - z * gy;
        T4 o;
        o.x = R(0); o.y = R(0); o.z = R(0); o.w = R(0);
        if (x > 0 && x < M - 1 && y > 0 && y < gy - 1 && z > 0 && z < gz - 1) {
            const int n = x >> 1;
            T2 a = buf[LdsLine<M>::at(line, n)];
            T2 b = buf[LdsLine<M>::at((NL + line), n)];
            T2 c = buf[LdsLine<M>::at((2 * NL + line), n)];
            o.x = (x & 1) ? a.y : a.x;
            o.y = (x & 1) ? b.y : b.x;
            o.z = (x & 1) ? c.y : c.x;
        }
        E[int64_t(r) * M + x] = o;
    }
}

// ---------------------------------------------------------------------------
// Green's spectrum (space_charge_3d_open_hockney.cc:136-348).  The value at
// the doubled-grid point (X, Y, Z) is f(x, y, z) with x = min(X, Dx-X)*hx
// etc. (sc3d_green.h), the value the reference's mirror writes leave there
// wherever its index decomposition is exact (sc3d_quirk.hip patches the
// few points of the shapes where it is not).
// ---------------------------------------------------------------------------
// The spectrum is transformed in T: FP64 in the FP64 form, FP32 in the
// precision form (whose convolution reads it in FP32).  The function values
// are always evaluated in FP64 (2 s0 - sn - sp cancels) and rounded once.
template <typename T> __device__ __forceinline__ T* green_buffer(const GridJob& J)
{
    if constexpr (sizeof(T) == sizeof(double))
        return J.ghat64;
    else
        return static_cast<float*>(J.ghat);
}

// G1: rows (Z<=gz, Y<=gy) of the doubled grid: generate, R2C along x, keep
// the real part of kx = 0..M (the row is even, its spectrum real).
template <typename T, int M>
__global__ __launch_bounds__(kThreads) void k_green_x(
    const GridJob* __restrict__ jobs, const typename cplx<T>::type* __restrict__ twM,
    const typename cplx<T>::type* __restrict__ tw2M, int gy, int gz, int green_fn)
{
    using T2 = typename cplx<T>::type;
    constexpr int NL = Lines<M>::value;
    constexpr int Dx = 2 * M;
    __shared__ T2 buf[NL * LdsLine<M>::stride];
    __shared__ T2 twl[M];
    __shared__ T row[NL * (M + 1)];
    const GridJob& J = jobs[blockIdx.y];
    T* __restrict__ ghat = green_buffer<T>(J);
    const DomainRec* dom = J.dom;
    if (dom->error || !dom->ghat_dirty) return;
    const int nrows = (gz + 1) * (gy + 1);
    const int r0 = blockIdx.x * NL;
    const double hx = dom->h[0], hy = dom->h[1], hz = dom->h[2];
    const double G000 = green_g000(green_fn, hx, hy, hz);
    for (int e = threadIdx.x; e < NL * (M + 1); e += kThreads) {
        int line = e / (M + 1), X = e % (M + 1);
        int r = r0 + line;
        double v = 0.0;
        if (r < nrows) {
            int Z = r / (gy + 1), Y = r - Z * (gy + 1);
            double x = X * hx, y = Y * hy, z = Z * hz;
            v = green_value(green_fn, x, y, z, hz, G000);
        }
        row[e] = T(v);
    }
    load_twiddles<T2, M>(twl, twM);
    __syncthreads();
    for (int e = threadIdx.x; e < NL * M; e += kThreads) {
        int line = e / M, n = e % M;
        int X0 = 2 * n, X1 = 2 * n + 1;
        int m0 = X0 <= M ? X0 : Dx - X0;
        int m1 = X1 <= M ? X1 : Dx - X1;
        T2 v;
        v.x = row[line * (M + 1) + m0];
        v.y = row[line * (M + 1) + m1];
        buf[LdsLine<M>::at(line, n)] = v;
    }
    __syncthreads();
    lds_fft<T2, M, NL, false>(buf, twl);
    const int Hx = M + 1;
    for (int e = threadIdx.x; e < NL * Hx; e += kThreads) {
        int line = e / Hx, k = e % Hx;
        int r = r0 + line;
        if (r >= nrows) continue;
        T2 zk = buf[LdsLine<M>::at(line, (k == M ? 0 : k))];
        T2 zm = buf[LdsLine<M>::at(line, (k == 0 ? 0 : M - k))];
        T2 A, B;
        A.x = zk.x + zm.x; A.y = zk.y - zm.y;
        B.x = zk.x - zm.x; B.y = zk.y + zm.y;
        T2 Cc = cmul(tw2M[k], B);
        ghat[int64_t(r) * Hx + k] = T(0.5) * (A.x + Cc.y);
    }
}

// G2 / G3: real-even column transforms in place over the octant buffer.
// Real column c -> (plane = c / Hx, kx = c % Hx); element p of the doubled
// axis reads the stored octant index min(p, L-p); writes real parts
// p <= L/2.  The spectrum of a real even column is real, so two columns
// share one complex transform: columns 2l and 2l+1 (adjacent kx) go in as
// the real and imaginary parts of line l and come out as its real and
// imaginary parts.
// AXIS 1 (y): address (plane*(gy+1) + p)*Hx + kx, plane = Z <= gz.
// AXIS 2 (z): address (p*(gy+1) + plane)*Hx + kx, plane = ky <= gy.
template <typename T, int L, int AXIS>
__global__ __launch_bounds__(kThreads) void k_green_cols(
    const GridJob* __restrict__ jobs, const typename cplx<T>::type* __restrict__ tw, int ncols,
    int gy, int Hx)
{
    using T2 = typename cplx<T>::type;
    constexpr int NL = GreenLines<L, T2>::value;
    __shared__ T2 buf[NL * LdsLine<L>::stride];
    __shared__ T2 twl[L];
    const GridJob& J = jobs[blockIdx.y];
    T* __restrict__ ghat = green_buffer<T>(J);
    const DomainRec* dom = J.dom;
    if (dom->error || !dom->ghat_dirty) return;
    const int l0 = xcd_group() * NL;
    auto addr = [&](int c, int p) -> int64_t {
        int plane = c / Hx, kx = c - plane * Hx;
        if (AXIS == 1) return (int64_t(plane) * (gy + 1) + p) * Hx + kx;
        return (int64_t(p) * (gy + 1) + plane) * Hx + kx;
    };
    // element e: p = e / (2 NL), column q = e % (2 NL) of the block's 2 NL
    // consecutive columns (runs of 2 NL doubles along kx)
    for (int e = threadIdx.x; e < 2 * NL * L; e += kThreads) {
        int p = e / (2 * NL), q = e % (2 * NL);
        int c = 2 * l0 + q;
        T v = c < ncols ? ghat[addr(c, p <= L / 2 ? p : L - p)] : T(0);
        T2* slot = &buf[LdsLine<L>::at((q >> 1), p)];
        if (q & 1) slot->y = v; else slot->x = v;
    }
    load_twiddles<T2, L>(twl, tw);
    __syncthreads();
    lds_fft<T2, L, NL, false>(buf, twl);
    for (int e = threadIdx.x; e < 2 * NL * (L / 2 + 1); e += kThreads) {
        int p = e / (2 * NL), q = e % (2 * NL);
        int c = 2 * l0 + q;
        if (c < ncols) {
            const T2 z = buf[LdsLine<L>::at((q >> 1), p)];
            ghat[addr(c, p)] = (q & 1) ? z.y : z.x;
        }
    }
}

// ---------------------------------------------------------------------------
// dispatch on the runtime lengths
// ---------------------------------------------------------------------------
template <template <int> class F, typename... A>
void dispatch_len(int L, A&&... a)
{
    switch (L) {
    case 8: F<8>::run(a...); break;
    case 16: F<16>::run(a...); break;
    case 32: F<32>::run(a...); break;
    case 64: F<64>::run(a...); break;
    case 128: F<128>::run(a...); break;
    case 256: F<256>::run(a...); break;
    case 512: F<512>::run(a...); break;
    default: break;  // the launchers route other lengths to the generic passes
    }
}

inline int nblocks(int64_t items, int per) { return int((items + per - 1) / per); }
// grid of a column pass: a multiple of the XCD count (see xcd_group)
inline int col_blocks(int64_t items, int per)
{
    return (nblocks(items, per) + kXcds - 1) / kXcds * kXcds;
}

template <typename T> struct Green {
    using T2 = typename cplx<T>::type;
    template <int M> struct X {
        static void run(const Geometry& g, int green_fn, const GridJob* jobs, int nj,
                        const T2* tw, hipStream_t st)
        {
            int rows = (g.gz + 1) * (g.gy + 1);
            hipLaunchKernelGGL((k_green_x<T, M>), dim3(nblocks(rows, Lines<M>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(M),
                               tw + twiddle_offset(2 * M), g.gy, g.gz, green_fn);
        }
    };
    template <int L> struct Y {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int cols = (g.gz + 1) * g.Hx;
            hipLaunchKernelGGL((k_green_cols<T, L, 1>),
                               dim3(col_blocks(cols, 2 * GreenLines<L, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(L), cols, g.gy,
                               g.Hx);
        }
    };
    template <int L> struct Z {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int cols = (g.gy + 1) * g.Hx;
            hipLaunchKernelGGL((k_green_cols<T, L, 2>),
                               dim3(col_blocks(cols, 2 * GreenLines<L, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(L), cols, g.gy,
                               g.Hx);
        }
    };
};

template <typename R>
void launch_green(const Geometry& g, int green_fn, const GridJob* jobs, int nj,
                  const typename cplx<R>::type* tw, hipStream_t st)
{
    if (g.tuned[0]) dispatch_len<Green<R>::template X>(g.gx, g, green_fn, jobs, nj, tw, st);
    else any_green_x<R>(g, green_fn, jobs, nj, tw, st);
    if (g.tuned[1]) dispatch_len<Green<R>::template Y>(g.Dy, g, jobs, nj, tw, st);
    else any_green_cols<R>(g, 1, jobs, nj, tw, st);
    if (g.tuned[2]) dispatch_len<Green<R>::template Z>(g.Dz, g, jobs, nj, tw, st);
    else any_green_cols<R>(g, 2, jobs, nj, tw, st);
}

template <typename R> struct SolveX {
    using T2 = typename cplx<R>::type;
    template <int M> struct Fwd {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int rows = g.gz * g.gy;
            hipLaunchKernelGGL((k_x_r2c<R, M>), dim3(nblocks(rows, XLines<M>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(M),
                               tw + twiddle_offset(2 * M), rows, g.gy, g.Dy);
        }
    };
    template <int M> struct Inv {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int rows = g.gz * g.gy;
            hipLaunchKernelGGL((k_x_c2r<R, M>), dim3(nblocks(rows, XLines<M>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(M),
                               tw + twiddle_offset(2 * M), rows, g.gy, g.Dy);
        }
    };
    template <int L> struct YFwd {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.gz * g.Hx;
            hipLaunchKernelGGL((k_cols<R, L, 0>),
                               dim3(col_blocks(lines, FieldLines<L, 1, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(L), lines, g.Hx,
                               int64_t(g.Dy) * g.Hx, int64_t(g.Hx), g.gy, g.Dy, g.gy, g.Hx);
        }
    };
    template <int L> struct YInv {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.gz * g.Hx;
            hipLaunchKernelGGL((k_cols<R, L, 1>),
                               dim3(col_blocks(lines, FieldLines<L, 1, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(L), lines, g.Hx,
                               int64_t(g.Dy) * g.Hx, int64_t(g.Hx), g.Dy, g.gy, g.gy, g.Hx);
        }
    };
    template <int L> struct ZConv {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.Dy * g.Hx;
            hipLaunchKernelGGL((k_cols<R, L, 2>),
                               dim3(col_blocks(lines, FieldLines<L, 1, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(L), lines, lines,
                               int64_t(0), int64_t(lines), g.gz, g.gz, g.gy, g.Hx);
        }
    };
    template <int L> struct ZField {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.Dy * g.Hx;
            hipLaunchKernelGGL((k_z_field<R, L>),
                               dim3(col_blocks(lines, ZFieldLines<L, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(L), lines, g.gz,
                               g.gy, g.Hx);
        }
    };
    template <int L> struct YField {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.gz * g.Hx;
            hipLaunchKernelGGL((k_y_field<R, L>),
                               dim3(col_blocks(lines, FieldLines<L, 3, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(L), lines, g.gy,
                               g.Hx);
        }
    };
    template <int M> struct XField {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int rows = g.gz * g.gy;
            hipLaunchKernelGGL((k_x_field<R, M>),
                               dim3(nblocks(rows, RowLines<M, 3, T2>::value), nj),
                               dim3(kThreads), 0, st, jobs, tw + twiddle_offset(M),
                               tw + twiddle_offset(2 * M), g.gy, g.gz, g.Dy);
        }
    };
};

template <typename R>
void launch_solve(const Geometry& g, const GridJob* jobs, int nj,
                  const typename cplx<R>::type* tw, hipStream_t st)
{
    if (g.tuned[0]) dispatch_len<SolveX<R>::template Fwd>(g.gx, g, jobs, nj, tw, st);
    else any_x_pass<R>(g, false, jobs, nj, tw, st);
    if (g.tuned[1]) dispatch_len<SolveX<R>::template YFwd>(g.Dy, g, jobs, nj, tw, st);
    else any_cols<R>(g, 0, jobs, nj, tw, st);
    if (g.tuned[2]) dispatch_len<SolveX<R>::template ZConv>(g.Dz, g, jobs, nj, tw, st);
    else any_cols<R>(g, 2, jobs, nj, tw, st);
    if (g.tuned[1]) dispatch_len<SolveX<R>::template YInv>(g.Dy, g, jobs, nj, tw, st);
    else any_cols<R>(g, 1, jobs, nj, tw, st);
    if (g.tuned[0]) dispatch_len<SolveX<R>::template Inv>(g.gx, g, jobs, nj, tw, st);
    else any_x_pass<R>(g, true, jobs, nj, tw, st);
}

template <typename R>
void launch_solve_field(const Geometry& g, const GridJob* jobs, int nj,
                        const typename cplx<R>::type* tw, hipStream_t st)
{
    if (g.tuned[0]) dispatch_len<SolveX<R>::template Fwd>(g.gx, g, jobs, nj, tw, st);
    else any_x_pass<R>(g, false, jobs, nj, tw, st);
    if (g.tuned[1]) dispatch_len<SolveX<R>::template YFwd>(g.Dy, g, jobs, nj, tw, st);
    else any_cols<R>(g, 0, jobs, nj, tw, st);
    if (g.tuned[2]) dispatch_len<SolveX<R>::template ZField>(g.Dz, g, jobs, nj, tw, st);
    else any_field<R>(g, 0, jobs, nj, tw, st);
    if (g.tuned[1]) dispatch_len<SolveX<R>::template YField>(g.Dy, g, jobs, nj, tw, st);
    else any_field<R>(g, 1, jobs, nj, tw, st);
    if (g.tuned[0]) dispatch_len<SolveX<R>::template XField>(g.gx, g, jobs, nj, tw, st);
    else any_field<R>(g, 2, jobs, nj, tw, st);
}

template void launch_green<double>(const Geometry&, int, const GridJob*, int, const double2*,
                                   hipStream_t);
template void launch_green<float>(const Geometry&, int, const GridJob*, int, const float2*,
                                  hipStream_t);
template void launch_solve<double>(const Geometry&, const GridJob*, int, const double2*,
                                   hipStream_t);
template void launch_solve_field<float>(const Geometry&, const GridJob*, int, const float2*,
                                        hipStream_t);

}  // namespace sc3d
