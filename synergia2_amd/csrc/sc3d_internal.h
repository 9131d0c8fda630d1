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

// Grid pipeline, FP64 or FP32 (R).  All launches on `st`, asynchronous.
template <typename R>
void launch_green(const Geometry& g, int green_fn, const DomainRec* dom,
                  const double2* tw64, double* ghat64, R* ghat_r, hipStream_t st);

template <typename R>
void launch_solve(const Geometry& g, const R* rho, typename cplx<R>::type* S,
                  const R* ghat, const typename cplx<R>::type* tw, R* phi,
                  const DomainRec* dom, hipStream_t st);

}  // namespace sc3d
