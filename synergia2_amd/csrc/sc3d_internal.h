// sc3d_internal.h -- host-side launch interface between the translation
// units of libsc3d_hip.so.  Not part of the C ABI.
#pragma once

#include "sc3d_common.h"

namespace sc3d {

// Geometry of one context: grid g, doubled grid D = 2g, half-spectrum width
// Hx = gx + 1.  Spectrum workspace S holds the gz planes z < gz of the
// doubled complex grid (Dy rows of Hx each); Ghat holds the real Green's
// spectrum for kz <= gz, ky <= gy, kx <= gx.
// Radix plan of a transform of any length L <= kAnyMaxLen for the generic
// passes (sc3d_fft_any.hip): factors 4, then primes ascending.
constexpr int kAnyMaxLen = 1024;
constexpr int kAnyMaxPasses = 16;
struct AnyPlan {
    int L;
    int np;
    int R[kAnyMaxPasses];
};
AnyPlan any_plan(int L);

struct Geometry {
    int gx, gy, gz;
    int Dx, Dy, Dz;
    int Hx;
    int64_t G;        // gx*gy*gz
    int64_t S_elems;  // gz*Dy*Hx complex
    int64_t Ghat_elems;  // (gz+1)(gy+1)Hx real
    // per axis (x: rows of gx / Dx, y: columns of Dy, z: columns of Dz):
    // the tuned power-of-two passes (sc3d_fft.hip) or the generic ones
    bool tuned[3];
    AnyPlan plan[3];     // generic axes: plan of Dx, Dy, Dz
    int64_t tw_any[3];   // generic axes: offset of the length-Dx/Dy/Dz table
    int64_t tw_elems;    // twiddle elements of the context (both tables)
};

// Twiddle tables e^{-2 pi i k / L}, k < L, for every power of two L <= 1024,
// stored back to back: table for L starts at element L - 1 (1 + 2 + 4 ...);
// a context appends the tables of its generic axes after them.
inline int64_t twiddle_offset(int L) { return L - 1; }
constexpr int kMaxLen = 1024;
constexpr int64_t kTwiddleElems = 2 * kMaxLen - 1;

// Axis lengths the tuned passes take: rows gx a power of two in [8, 256],
// columns Dy, Dz powers of two in [8, 512].
bool tuned_row_length(int gx);
bool tuned_col_length(int L);

// A point of the doubled-grid g2 where the reference's index decomposition
// (alg_g2_*, space_charge_3d_open_hockney.cc:170-176,272-278) leaves another
// value than the symmetric one: target t = (z, y, x) and the (iz, iy, ix)
// of the index that writes it last, s[0] < 0 when none does (it stays 0).
struct GreenPatch {
    int32_t t[3];
    int32_t s[3];
};
// Up to this many patch points the correction is added in direct space
// (cost: points x grid nodes per apply); beyond it, as the spectrum chat.
constexpr int kMaxGreenPatches = 512;

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
    // shapes where the reference's g2 index decomposition misplaces many
    // points: the full complex spectrum of (reference g2 - symmetric g2),
    // cplx<R>[Dz][Dy][Hx], added to the real Green's spectrum in the z pass
    // (null otherwise)
    void* chat;
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

// The generic passes (sc3d_fft_any.hip) of the axes whose lengths the tuned
// ones do not take: Green's x / y, z passes, the solve's x pass (inverse:
// C2R), its column passes (which: 0 y forward, 1 y inverse, 2 z
// convolution), the precision form's field passes (which: 0 z, 1 y, 2 x).
template <typename R>
void any_green_x(const Geometry& g, int green_fn, const GridJob* jobs, int nj,
                 const typename cplx<R>::type* tw, hipStream_t st);
template <typename R>
void any_green_cols(const Geometry& g, int axis, const GridJob* jobs, int nj,
                    const typename cplx<R>::type* tw, hipStream_t st);
template <typename R>
void any_x_pass(const Geometry& g, bool inverse, const GridJob* jobs, int nj,
                const typename cplx<R>::type* tw, hipStream_t st);
template <typename R>
void any_cols(const Geometry& g, int which, const GridJob* jobs, int nj,
              const typename cplx<R>::type* tw, hipStream_t st);
template <typename R>
void any_field(const Geometry& g, int which, const GridJob* jobs, int nj,
               const typename cplx<R>::type* tw, hipStream_t st);
// The spectrum of the g2 correction (real C[Dz][Dy][Dx], zero except at
// the patch points) into jobs' chat, where h changed: full-length x R2C,
// then complex y and z transforms over the whole doubled grid.
template <typename R>
void any_chat(const Geometry& g, const void* C, const GridJob* jobs, void* chat,
              const typename cplx<R>::type* tw, hipStream_t st);

}  // namespace sc3d
