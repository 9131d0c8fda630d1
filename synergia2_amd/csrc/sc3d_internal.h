// sc3d_internal.h -- host-side launch interface between the translation
// units of libsc3d_hip.so.  Not part of the C ABI.
#pragma once

#include "sc3d_common.h"

namespace sc3d {

// Geometry of one context: grid g, doubled grid D = 2g, half-spectrum width
// Hx = gx + 1.  Spectrum workspace S holds the gz planes z < gz of the
// doubled complex grid (Dy rows of Hx each); Ghat holds the real Green's
// spectrum for kz <= gz, ky <= gy, kx <= gx.
struct Geometry {
    int gx, gy, gz;
    int Dx, Dy, Dz;
    int Hx;
    int64_t G;        // gx*gy*gz
    int64_t S_elems;  // gz*Dy*Hx complex
    int64_t Ghat_elems;  // (gz+1)(gy+1)Hx real
};

// Twiddle tables e^{-2 pi i k / L}, k < L, for every power of two L <= 1024,
// stored back to back: table for L starts at element L - 1 (1 + 2 + 4 ...).
inline int64_t twiddle_offset(int L) { return L - 1; }
constexpr int kMaxLen = 1024;
constexpr int64_t kTwiddleElems = 2 * kMaxLen - 1;

bool fft_length_supported(int L);

template <typename R> struct vec4_of;
template <> struct vec4_of<double> { using type = double4; };
template <> struct vec4_of<float> { using type = float4; };

// Per-job pointers of the grid pipeline.  The grid kernels take a device
// table of these and serve job blockIdx.y, so one launch covers a batch of
// same-shape jobs (a context alone is a table of one).
struct GridJob {
    void* rho;        // R[G] density octant
    void* S;          // cplx<R>[S_elems] spectrum workspace
    void* S2;         // precision form: second spectrum workspace
    void* phi;        // FP64 form: R[G] potential octant
    void* E;          // vec4<R>[G] field (Ex, Ey, Ez, 0)
    double* ghat64;   // FP64 form: Green's spectrum octant (null in the FP32 form)
    void* ghat;       // the spectrum the convolution reads: FP32 octant, or ghat64
    DomainRec* dom;
};

// Green's spectrum (when the job's cell size changed), transformed in R
// (values evaluated in FP64).  All launches on `st`, asynchronous.
template <typename R>
void launch_green(const Geometry& g, int green_fn, const GridJob* jobs, int nj,
                  const typename cplx<R>::type* tw, hipStream_t st);

// FP64 form: rho -> phi (five passes, convolution fused into the z pass).
template <typename R>
void launch_solve(const Geometry& g, const GridJob* jobs, int nj,
                  const typename cplx<R>::type* tw, hipStream_t st);

// Precision form: the same convolution, with the field transformed straight
// from the spectrum (E = IDFT(phi^ * D_a), D_a = -i sin(2 pi k/D_a)/h_a) and
// written as (Ex, Ey, Ez, 0) per node.  No potential is formed.
template <typename R>
void launch_solve_field(const Geometry& g, const GridJob* jobs, int nj,
                        const typename cplx<R>::type* tw, hipStream_t st);

}  // namespace sc3d
