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

// Lines per block of the LDS passes that hold whole line blocks (the
// precision form's x' pass, the Green's column passes), from a preferred
// count, halved (down to 2) while the block's LDS (K lines per transform
// line plus the twiddle table) exceeds the cap: the passes are
// latency-bound at low occupancy, and gfx950 has 160 KiB of LDS per CU.
// (Measured when every transform pass held its lines this way, at 64x64x128
// and 128x128x256: 32 KiB for FP64 lines (24 KiB the same, 48 KiB slower);
// 32 KiB for FP32 lines (field passes 273 us vs 280 us at 48 KiB, 303 us at
// 64 KiB).  16 KiB for both: solve 57 -> 62 us at 64x64x128, 265 -> 291 us
// at 128x128x256.)
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
// row passes holding K lines per row (rows are contiguous: no run width)
template <int M, int K, typename T2>
using RowLines = CappedLines<M, K, T2, HalfLines<M>::value>;
// the Green's function column passes (two real columns per line)
template <int L, typename T2>
using GreenLines = CappedLines<L, 1, T2, HalfLines<L>::value>;

// a * w^m, w = e^(-2 pi i / 16) (inverse: e^(+2 pi i / 16)), written with
// the cosine c and sine s of 2 pi m / 16 (fused multiply-adds: the build
// contracts nothing by itself)
template <bool INV, typename T2, typename F>
__device__ __forceinline__ T2 rot16(T2 a, F c, F s)
{
    T2 o;
    if (!INV) { o.x = fma(a.x, c, a.y * s); o.y = fma(a.y, c, -(a.x * s)); }
    else      { o.x = fma(a.x, c, -(a.y * s)); o.y = fma(a.y, c, a.x * s); }
    return o;
}

// a * w for a twiddle w (INV: a * conj(w)), fused multiply-adds
template <bool INV, typename T2>
__device__ __forceinline__ T2 twmul(T2 a, T2 w)
{
    T2 o;
    if (!INV) { o.x = fma(a.x, w.x, -(a.y * w.y)); o.y = fma(a.x, w.y, a.y * w.x); }
    else      { o.x = fma(a.x, w.x, a.y * w.y);    o.y = fma(a.y, w.x, -(a.x * w.y)); }
    return o;
}

template <int R, bool INV, typename T2>
__device__ __forceinline__ void dft_small(T2* v)
{
    if constexpr (R == 16) {
        // 16 = 4 x 4: inputs x[4 n1 + n2]; a length-4 DFT over n1 for each
        // n2 gives A[n2][k1], then A[n2][k1] w^(n2 k1), then a length-4 DFT
        // over n2 for each k1 gives X[k1 + 4 k2]
        using F = decltype(v[0].x);
        const F c1 = F(0.92387953251128675613), s1 = F(0.38268343236508977173);
        const F c2 = F(0.70710678118654752440);
        T2 a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = v[4 * n1 + n2];
            dft_small<4, INV>(a[n2]);
        }
        a[1][1] = rot16<INV>(a[1][1], c1, s1);   // w^1
        a[1][2] = rot16<INV>(a[1][2], c2, c2);   // w^2
        a[1][3] = rot16<INV>(a[1][3], s1, c1);   // w^3
        a[2][1] = rot16<INV>(a[2][1], c2, c2);   // w^2
        {                                        // w^4 = -i (inverse: +i)
            T2 t = a[2][2];
            if (!INV) { a[2][2].x = t.y;  a[2][2].y = -t.x; }
            else      { a[2][2].x = -t.y; a[2][2].y = t.x; }
        }
        a[2][3] = rot16<INV>(a[2][3], -c2, c2);  // w^6
        a[3][1] = rot16<INV>(a[3][1], s1, c1);   // w^3
        a[3][2] = rot16<INV>(a[3][2], -c2, c2);  // w^6
        a[3][3] = rot16<INV>(a[3][3], -c1, -s1); // w^9
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            T2 b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            dft_small<4, INV>(b);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
        }
    } else if constexpr (R == 8) {
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
// Register-resident column transforms.  A line of length L is held by
// TPL = L / E threads with E elements each (E = the line's largest radix);
// a radix-r pass is E / r butterflies per thread in registers, and only the
// exchanges between passes go through LDS (one line buffer per line).
// Thread j of a line owns the butterflies b = j + q * TPL (q < E / r) of a
// radix-r pass, whose inputs are the elements b + s * L / r; it writes its
// outputs to the Stockham positions (b / Ns) * Ns * r + b % Ns + s * Ns.  In
// the last pass (Ns * r = L) those are the input positions again, so a
// forward transform that ends with radix r and an inverse one that starts
// with it (the inverse runs the radices in reverse) hold the same elements
// in the same registers: the z pass multiplies by the Green's spectrum in
// registers and goes straight on with the inverse transform.
// ---------------------------------------------------------------------------
template <int L, int W> struct RegShape {
    // forward radices, first to last; W = 16: radix-16 passes where L allows
    // (two passes for 128 and 256), W = 8: radix-8 passes and one smaller
    static constexpr int np = (W == 16 ? L == 512 : L >= 128) ? 3 : (L <= 16 ? 1 : 2);
    static constexpr int r0 = (L == 16 || (W == 16 && (L == 128 || L == 256))) ? 16 : 8;
    static constexpr int r1 = L == 32 ? 4 : ((W == 16 && L == 256) ? 16 : (L >= 64 ? 8 : 1));
    static constexpr int r2 = L == 512 ? 8 : ((W == 8 && L == 128) ? 2 : ((W == 8 && L == 256) ? 4 : 1));
    static constexpr int E = r0;
    static constexpr int TPL = L / E;
    static constexpr int radix(int p, bool inv)
    {
        const int q = inv ? np - 1 - p : p;
        return q == 0 ? r0 : (q == 1 ? r1 : r2);
    }
    static constexpr int stage(int p, bool inv)
    {
        int ns = 1;
        for (int i = 0; i < p; ++i) ns *= radix(i, inv);
        return ns;
    }
    static_assert(L <= 16 || r0 * r1 * r2 == L, "radices multiply to L");
};
// The plan of a precision: radix 16 for FP64 lines; radix 8 for FP32 lines,
// twice the threads per line (measured at 128x128x256 FP32: y' pass 61.3 ->
// 50.1 us, y pass 22.2 -> 20.9 us; at 64x64x128 FP64 the radix-16 z pass is
// 16.6 us against 17.6)
template <int L, typename T2> using RegPlan = RegShape<L, (sizeof(T2) >= 16 ? 16 : 8)>;

template <typename T2, int L, int P, bool INV>
__device__ __forceinline__ void reg_pass(T2 (&v)[(RegPlan<L, T2>::E)], int j, const T2* __restrict__ tw)
{
    using S = RegPlan<L, T2>;
    constexpr int r = S::radix(P, INV), NS = S::stage(P, INV), Q = S::E / r;
    constexpr int SH = L / (NS * r);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (NS > 1) {
            const int k = (j + q * S::TPL) & (NS - 1);
#pragma unroll
            for (int s = 1; s < r; ++s) v[q * r + s] = twmul<INV>(v[q * r + s], tw[s * k * SH]);
        }
        dft_small<r, INV>(v + q * r);
    }
}

// Line buffers are L elements apart with no padding; element p of a line
// sits at p ^ sw, sw = line mod G, G = the elements in 128 bytes (at most
// L): the lanes of a wavefront that touch the same element of G
// neighbouring lines then hit G different bank groups.
template <int L, typename T2> struct RegSwz {
    static constexpr int G0 = 128 / int(sizeof(T2));
    static constexpr int G = G0 < L ? G0 : L;
};

// outputs of pass P -> the line buffer
template <typename T2, int L, int P, bool INV>
__device__ __forceinline__ void reg_to_lds(const T2 (&v)[(RegPlan<L, T2>::E)], T2* lb, int j, int sw)
{
    using S = RegPlan<L, T2>;
    constexpr int r = S::radix(P, INV), NS = S::stage(P, INV), Q = S::E / r;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = j + q * S::TPL;
        const int d = (b / NS) * NS * r + (b & (NS - 1));
#pragma unroll
        for (int s = 0; s < r; ++s) lb[(d + s * NS) ^ sw] = v[q * r + s];
    }
}

// inputs of pass P <- the line buffer
template <typename T2, int L, int P, bool INV>
__device__ __forceinline__ void reg_from_lds(T2 (&v)[(RegPlan<L, T2>::E)], const T2* lb, int j, int sw)
{
    using S = RegPlan<L, T2>;
    constexpr int r = S::radix(P, INV), Q = S::E / r, B = L / r;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = j + q * S::TPL;
#pragma unroll
        for (int s = 0; s < r; ++s) v[q * r + s] = lb[(b + s * B) ^ sw];
    }
}

// One unnormalised length-L transform of the line held in v (the inputs of
// pass 0 on entry, the outputs of the last pass on return).  lb is the
// line's LDS buffer and sw its swizzle, tw the length-L twiddle table (the
// block's LDS copy, filled before the call); busy: the block has used its
// line buffers before (a barrier then precedes the first write).  Every
// thread of the block calls it (it holds block barriers).
template <typename T2, int L, bool INV>
__device__ __forceinline__ void reg_fft(T2 (&v)[(RegPlan<L, T2>::E)], T2* lb, int j, int sw,
                                        const T2* __restrict__ tw, bool busy)
{
    using S = RegPlan<L, T2>;
    reg_pass<T2, L, 0, INV>(v, j, tw);
    if constexpr (S::np > 1) {
        if (busy) __syncthreads();
        reg_to_lds<T2, L, 0, INV>(v, lb, j, sw);
        __syncthreads();
        reg_from_lds<T2, L, 1, INV>(v, lb, j, sw);
        reg_pass<T2, L, 1, INV>(v, j, tw);
    }
    if constexpr (S::np > 2) {
        __syncthreads();
        reg_to_lds<T2, L, 1, INV>(v, lb, j, sw);
        __syncthreads();
        reg_from_lds<T2, L, 2, INV>(v, lb, j, sw);
        reg_pass<T2, L, 2, INV>(v, j, tw);
    }
}

// Lines per block of the register-resident column passes: 128-byte runs
// across the lines (8 FP64 or 16 FP32 complex values: neighbouring lines
// are contiguous), fewer where the block's LDS (line buffers plus the
// twiddle table) would pass the cap, and at least one wavefront of threads.
// (Measured at 64x64x128 FP64, z pass: 128-byte runs 16.6 us, 64-byte 17.9.)
#ifndef SC3D_REG_RUN_BYTES
#define SC3D_REG_RUN_BYTES 128
#endif
#ifndef SC3D_REG_LDS_CAP
#define SC3D_REG_LDS_CAP 40960
#endif
template <int L, typename T2> struct RegLines {
    static constexpr int TPL = RegPlan<L, T2>::TPL;
    static constexpr size_t bytes(int nl) { return size_t(nl + 1) * L * sizeof(T2); }
    static constexpr int fit(int nl) { return (nl > 1 && bytes(nl) > SC3D_REG_LDS_CAP) ? fit(nl / 2) : nl; }
    static constexpr int runs = SC3D_REG_RUN_BYTES / int(sizeof(T2));
    static constexpr int wave = 64 / TPL > 1 ? 64 / TPL : 1;
    static constexpr int value = fit(runs > wave ? runs : wave);
    static constexpr int threads = value * TPL;
};

// Column passes over S.  Line l -> (plane = l / lpp, c = l % lpp), element p
// at S[plane*plane_stride + p*pstride + c].  MODE 0: forward, loads the
// non-zero half p < L/2; 1: inverse, stores the kept half p < L/2; 2 (the z
// pass): forward, multiply by the real Green's spectrum, inverse, half
// loaded and half stored.
template <typename R, int L, int MODE>
__global__ __launch_bounds__((RegLines<L, typename cplx<R>::type>::threads)) void k_cols_reg(
    const GridJob* __restrict__ jobs, const typename cplx<R>::type* __restrict__ tw, int nlines,
    int lpp, int64_t plane_stride, int64_t pstride, int gy, int Hx)
{
    using T2 = typename cplx<R>::type;
    using S = RegPlan<L, T2>;
    constexpr int NL = RegLines<L, T2>::value;
    constexpr int E = S::E, TPL = S::TPL;
    __shared__ T2 buf[NL * L];
    __shared__ T2 twl[L];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S_ = static_cast<T2*>(J.S);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const int line = int(threadIdx.x) % NL, j = int(threadIdx.x) / NL;
    const int l = xcd_group() * NL + line;
    const bool live = l < nlines;
    int64_t base = 0;
    if (live) {
        const int plane = l / lpp, c = l - plane * lpp;
        base = plane * plane_stride + c;
    }
    T2* lb = buf + line * L;
    const int sw = line & (RegSwz<L, T2>::G - 1);
    for (int e = threadIdx.x; e < L; e += NL * TPL) twl[e] = tw[e];
    // MODE 2: the Green's values of the elements this thread multiplies (the
    // outputs of the last forward pass), fetched with the line's loads
    R gv[E];
    if constexpr (MODE == 2) {
        const R* __restrict__ ghat = static_cast<const R*>(J.ghat);
        constexpr int rl = S::radix(S::np - 1, false), Bl = L / rl, Ql = E / rl;
        const int Dy = lpp / Hx;
        const int ky = l / Hx, kx = l - ky * Hx;
        const int my = ky <= Dy / 2 ? ky : Dy - ky;
#pragma unroll
        for (int q = 0; q < Ql; ++q)
#pragma unroll
            for (int s = 0; s < rl; ++s) {
                const int p = j + q * TPL + s * Bl;
                const int mz = p <= L / 2 ? p : L - p;
                gv[q * rl + s] = live ? ghat[(int64_t(mz) * (gy + 1) + my) * Hx + kx] : R(0);
            }
    }
    // inputs of the first pass: elements b + s * L / r, b = j + q * TPL; the
    // forward passes start with r = E (one butterfly: elements j + s * TPL,
    // so p < L/2 is s < E/2)
    constexpr bool INV0 = MODE == 1;
    constexpr int rf = S::radix(0, INV0), Bf = L / rf, Qf = E / rf;
    static_assert(MODE == 1 || Qf == 1, "forward passes start with the widest radix");
    T2 v[E];
#pragma unroll
    for (int q = 0; q < Qf; ++q)
#pragma unroll
        for (int s = 0; s < rf; ++s) {
            const int p = j + q * TPL + s * Bf;
            const bool in = MODE == 1 || s < rf / 2;
            if (in && live) {
                v[q * rf + s] = S_[base + p * pstride];
            } else {
                v[q * rf + s].x = R(0);
                v[q * rf + s].y = R(0);
            }
        }
    reg_fft<T2, L, MODE == 1>(v, lb, j, sw, twl, false);
    if constexpr (MODE == 2) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            v[e].x = v[e].x * gv[e];
            v[e].y = v[e].y * gv[e];
        }
        reg_fft<T2, L, true>(v, lb, j, sw, twl, true);
    }
    // outputs of the last pass: elements b + s * L / r; the inverse passes
    // end with r = E (p < L/2 is s < E/2)
    constexpr bool INVL = MODE != 0;
    constexpr int rl = S::radix(S::np - 1, INVL), Bl = L / rl, Ql = E / rl;
    static_assert(MODE == 0 || Ql == 1, "inverse passes end with the widest radix");
    if (!live) return;
#pragma unroll
    for (int q = 0; q < Ql; ++q)
#pragma unroll
        for (int s = 0; s < rl; ++s) {
            const int p = j + q * TPL + s * Bl;
            if (MODE == 0 || s < rl / 2) S_[base + p * pstride] = v[q * rl + s];
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

// The z' and y' passes register-resident (the register layout of the
// column passes above).  z': forward, x Ghat in registers, a copy kept for
// T * D_z, the two inverse transforms one after the other through the same
// line buffer.  y': the three inverse transforms likewise.  Every load of a
// line precedes the barrier of its first exchange, so the in-place stores
// cannot overtake a load (L <= 16: one thread per line).
template <typename R, int L>
__global__ __launch_bounds__((RegLines<L, typename cplx<R>::type>::threads)) void k_z_field_reg(
    const GridJob* __restrict__ jobs, const typename cplx<R>::type* __restrict__ tw, int nlines,
    int gy, int Hx)
{
    using T2 = typename cplx<R>::type;
    using S = RegPlan<L, T2>;
    constexpr int NL = RegLines<L, T2>::value;
    constexpr int E = S::E, TPL = S::TPL;
    __shared__ T2 buf[NL * L];
    __shared__ T2 twl[L];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S1 = static_cast<T2*>(J.S);
    T2* __restrict__ S2 = static_cast<T2*>(J.S2);
    const R* __restrict__ ghat = static_cast<const R*>(J.ghat);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihz = R(dom->inv_h[2]);
    const int line = int(threadIdx.x) % NL, j = int(threadIdx.x) / NL;
    const int l = xcd_group() * NL + line;
    const bool live = l < nlines;
    T2* lb = buf + line * L;
    const int sw = line & (RegSwz<L, T2>::G - 1);
    for (int e = threadIdx.x; e < L; e += NL * TPL) twl[e] = tw[e];
    // Green's values of the outputs of the last forward pass
    constexpr int rl = S::radix(S::np - 1, false), Bl = L / rl, Ql = E / rl;
    const int Dy = nlines / Hx;
    const int ky = l / Hx, kx = l - ky * Hx;
    const int my = ky <= Dy / 2 ? ky : Dy - ky;
    R gv[E];
#pragma unroll
    for (int q = 0; q < Ql; ++q)
#pragma unroll
        for (int s = 0; s < rl; ++s) {
            const int p = j + q * TPL + s * Bl;
            const int mz = p <= L / 2 ? p : L - p;
            gv[q * rl + s] = live ? ghat[(int64_t(mz) * (gy + 1) + my) * Hx + kx] : R(0);
        }
    // z < gz = L/2 loaded: elements j + s * TPL, s < E/2
    T2 v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        if (s < E / 2 && live) {
            v[s] = S1[int64_t(j + s * TPL) * nlines + l];
        } else {
            v[s].x = R(0);
            v[s].y = R(0);
        }
    }
    reg_fft<T2, L, false>(v, lb, j, sw, twl, false);
    T2 keep[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        v[e].x = v[e].x * gv[e];
        v[e].y = v[e].y * gv[e];
        keep[e] = v[e];
    }
    reg_fft<T2, L, true>(v, lb, j, sw, twl, true);
    if (live) {
#pragma unroll
        for (int s = 0; s < E / 2; ++s) S1[int64_t(j + s * TPL) * nlines + l] = v[s];
    }
#pragma unroll
    for (int q = 0; q < Ql; ++q)
#pragma unroll
        for (int s = 0; s < rl; ++s) {
            const int p = j + q * TPL + s * Bl;
            v[q * rl + s] = mul_minus_i(keep[q * rl + s], sin_2pi(twl, p, L) * ihz);
        }
    reg_fft<T2, L, true>(v, lb, j, sw, twl, true);
    if (live) {
#pragma unroll
        for (int s = 0; s < E / 2; ++s) S2[int64_t(j + s * TPL) * nlines + l] = v[s];
    }
}

template <typename R, int L>
__global__ __launch_bounds__((RegLines<L, typename cplx<R>::type>::threads)) void k_y_field_reg(
    const GridJob* __restrict__ jobs, const typename cplx<R>::type* __restrict__ tw, int nlines,
    int gy, int Hx)
{
    using T2 = typename cplx<R>::type;
    using S = RegPlan<L, T2>;
    constexpr int NL = RegLines<L, T2>::value;
    constexpr int E = S::E, TPL = S::TPL;
    __shared__ T2 buf[NL * L];
    __shared__ T2 twl[L];
    const GridJob& J = jobs[blockIdx.y];
    T2* __restrict__ S1 = static_cast<T2*>(J.S);
    T2* __restrict__ S2 = static_cast<T2*>(J.S2);
    const DomainRec* dom = J.dom;
    if (dom->error) return;
    const R ihy = R(dom->inv_h[1]);
    const int line = int(threadIdx.x) % NL, j = int(threadIdx.x) / NL;
    const int l = xcd_group() * NL + line;
    const bool live = l < nlines;
    int64_t base = 0;
    if (live) {
        const int z = l / Hx, c = l - z * Hx;
        base = int64_t(z) * L * Hx + c;  // L = Dy rows of Hx
    }
    T2* lb = buf + line * L;
    const int sw = line & (RegSwz<L, T2>::G - 1);
    for (int e = threadIdx.x; e < L; e += NL * TPL) twl[e] = tw[e];
    // inputs of the first inverse pass (radix r0): elements b + s * L / r0
    constexpr int r0 = S::radix(0, true), B0 = L / r0, Q0 = E / r0;
    T2 a[E], b[E], c[E];
#pragma unroll
    for (int q = 0; q < Q0; ++q)
#pragma unroll
        for (int s = 0; s < r0; ++s) {
            const int p = j + q * TPL + s * B0, e = q * r0 + s;
            if (live) {
                a[e] = S1[base + int64_t(p) * Hx];
                c[e] = S2[base + int64_t(p) * Hx];
            } else {
                a[e].x = R(0); a[e].y = R(0);
                c[e] = a[e];
            }
            b[e] = mul_minus_i(a[e], sin_2pi(tw, p, L) * ihy);
        }
    // rows y < gy kept: the inverse passes end with radix E, elements
    // j + s * TPL, s < E/2
    reg_fft<T2, L, true>(a, lb, j, sw, twl, false);
    if (live) {
#pragma unroll
        for (int s = 0; s < E / 2; ++s) S1[base + int64_t(j + s * TPL) * Hx] = a[s];
    }
    reg_fft<T2, L, true>(b, lb, j, sw, twl, true);
    if (live) {
#pragma unroll
        for (int s = 0; s < E / 2; ++s)
            S1[base + int64_t(gy + j + s * TPL) * Hx] = b[s];
    }
    reg_fft<T2, L, true>(c, lb, j, sw, twl, true);
    if (live) {
#pragma unroll
        for (int s = 0; s < E / 2; ++s) S2[base + int64_t(j + s * TPL) * Hx] = c[s];
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
            hipLaunchKernelGGL((k_cols_reg<R, L, 0>),
                               dim3(col_blocks(lines, RegLines<L, T2>::value), nj),
                               dim3(RegLines<L, T2>::threads), 0, st, jobs, tw + twiddle_offset(L),
                               lines, g.Hx, int64_t(g.Dy) * g.Hx, int64_t(g.Hx), g.gy, g.Hx);
        }
    };
    template <int L> struct YInv {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.gz * g.Hx;
            hipLaunchKernelGGL((k_cols_reg<R, L, 1>),
                               dim3(col_blocks(lines, RegLines<L, T2>::value), nj),
                               dim3(RegLines<L, T2>::threads), 0, st, jobs, tw + twiddle_offset(L),
                               lines, g.Hx, int64_t(g.Dy) * g.Hx, int64_t(g.Hx), g.gy, g.Hx);
        }
    };
    template <int L> struct ZConv {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.Dy * g.Hx;
            hipLaunchKernelGGL((k_cols_reg<R, L, 2>),
                               dim3(col_blocks(lines, RegLines<L, T2>::value), nj),
                               dim3(RegLines<L, T2>::threads), 0, st, jobs, tw + twiddle_offset(L),
                               lines, lines, int64_t(0), int64_t(lines), g.gy, g.Hx);
        }
    };
    template <int L> struct ZField {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.Dy * g.Hx;
            hipLaunchKernelGGL((k_z_field_reg<R, L>),
                               dim3(col_blocks(lines, RegLines<L, T2>::value), nj),
                               dim3(RegLines<L, T2>::threads), 0, st, jobs, tw + twiddle_offset(L),
                               lines, g.gy, g.Hx);
        }
    };
    template <int L> struct YField {
        static void run(const Geometry& g, const GridJob* jobs, int nj, const T2* tw,
                        hipStream_t st)
        {
            int lines = g.gz * g.Hx;
            hipLaunchKernelGGL((k_y_field_reg<R, L>),
                               dim3(col_blocks(lines, RegLines<L, T2>::value), nj),
                               dim3(RegLines<L, T2>::threads), 0, st, jobs, tw + twiddle_offset(L),
                               lines, g.gy, g.Hx);
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
