// sc3d_oracle.cc -- CPU restatement of the reference 3D open-Hockney step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker of the HIP
// library (tests/) and the timed CPU baseline (bench.py cpu_baseline leg).
// The product never links, loads or calls it.
//
// Built by oracle/Makefile with -O2 -fopenmp -ffp-contract=off
// -fno-fast-math so that every floating-point operation is the plain IEEE
// operation written here (no FMA contraction), which is what makes the cell
// indices reproducible bit for bit.
//
// What pins each function (tests/test_oracle_*.py):
//   orc_tree_sums        -- independent numpy restatement of the fixed tree;
//                           math.fsum within 1e-15; exact on +-a pairs.
//   orc_compute_domain   -- closed form on symmetric fixtures; throw on
//                           negligible extent; hand-computed constants.
//   orc_cell_index/keys  -- integer closed forms on hand-built coordinates.
//   orc_stable_sort      -- numpy.argsort(kind="stable").
//   orc_fft1d/3d         -- numpy.fft golden vectors (powers of two and
//                           lengths 12, 30, 74; 3-D 12x10x6, 10x6x14),
//                           roundtrip, Parseval.
//   orc_green_fn2        -- independent numpy evaluation of the closed form,
//                           mirror symmetry, far field -> hz/r; at the shapes
//                           where the float-reciprocal index decomposition
//                           misplaces points, the closed form at each
//                           misplaced writer's coordinates (or 0).
//   orc_deposit          -- single-particle closed form, charge
//                           conservation, brute-force O(8N) numpy deposit.
//   convolution (apply)  -- brute-force direct sum on a tiny grid.
//   orc_force            -- linear/quadratic phi gives exact differences.
//   orc_kick             -- rod-of-charge physics on the x and y axes;
//                           Gaussian-sphere probes on each axis (closed-form
//                           field: dp/p, x', y' and no cross terms).
//   orc_moments          -- numpy mean / centred cov, exact Hadamard
//                           fixture, restated summation tree (bit-equal).
//   orc_philox4x32_10    -- Random123 known-answer vectors; numpy restatement.
//   orc_det_log/sincos2pi-- 200-bit arithmetic (mpmath), within 1 / 2 ulp.
//   orc_populate_gaussian-- the pieces above; sample moments vs L L^T.
//   orc_adjust_matrix    -- numpy cholesky / inverse (1e-14).
//   orc_adjust_moments   -- the adjusted bunch's mean and second moments
//                           equal the requested ones (closed form, 1e-13);
//                           reference's symmetric-check messages.
#include "sc3d_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

namespace {

// physical_constants.h:17-24 (PDG 2012) and :94-106
constexpr double kPi = 3.141592653589793;  // Kokkos::numbers::pi_v<double>
constexpr double kE = 1.602176565e-19;
constexpr double kC = 299792458.0;
constexpr double kMu0 = 4 * kPi * 1.0e-7;
constexpr double kEpsilon0 = 1.0 / (kC * kC * kMu0);

thread_local std::string g_err;

int padded_real(int s) { return 2 * (s / 2 + 1); }  // distributed_fft3d.h:18-22
int padded_cplx(int s) { return s / 2 + 1; }        // distributed_fft3d.h:24-28

// space_charge_3d_open_hockney.cc:94-109
double smallest_non_tiny(double val, double o1, double o2, double tiny)
{
    if (val > tiny) return val;
    if ((o1 > tiny) && (o2 > tiny)) return std::min(o1, o2);
    return std::max(o1, o2);
}

double wtime() { return omp_get_wtime(); }

}  // namespace

extern "C" const char* orc_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Bunch-extent sums (core_diagnostics.cc:139-153).  The reference's Kokkos
// parallel_reduce order is unspecified; this build fixes it to one tree,
// shared bit for bit with the device kernel:
//   chunk c covers particles [4096c, 4096c+4096);
//   lane t of the chunk adds, starting from +0.0 and in order j = 0..15, the
//   values of particle 4096c + 256j + t (skipped when masked or >= n);
//   the 256 lane sums are combined pairwise: s[t] = s[t] + s[t+w] for
//   w = 1,2,...,128 and t a multiple of 2w; the chunk result is s[0];
//   the chunk results P[c] are combined by the same lane/pairwise rule.
// ---------------------------------------------------------------------------
namespace {
constexpr int kTreeLanes = 256;
constexpr int kTreeItems = 16;
constexpr int64_t kTreeChunk = kTreeLanes * kTreeItems;

void pairwise256(double (*s)[6])
{
    for (int w = 1; w < kTreeLanes; w *= 2)
        for (int t = 0; t < kTreeLanes; t += 2 * w)
            for (int q = 0; q < 6; ++q) s[t][q] = s[t][q] + s[t + w][q];
}
}  // namespace

extern "C" void orc_tree_sums(const double* parts, int64_t stride, int64_t n,
                              const uint8_t* masks, double out[6])
{
    int64_t nch = (n + kTreeChunk - 1) / kTreeChunk;
    std::vector<double> P(static_cast<size_t>(nch) * 6);
    static thread_local double s[kTreeLanes][6];
    for (int64_t c = 0; c < nch; ++c) {
        for (int t = 0; t < kTreeLanes; ++t) {
            for (int q = 0; q < 6; ++q) s[t][q] = 0.0;
            for (int j = 0; j < kTreeItems; ++j) {
                int64_t i = c * kTreeChunk + int64_t(j) * kTreeLanes + t;
                if (i >= n || !masks[i]) continue;
                double x = parts[i], y = parts[i + 2 * stride],
                       z = parts[i + 4 * stride];
                s[t][0] += x;
                s[t][1] += y;
                s[t][2] += z;
                s[t][3] += x * x;
                s[t][4] += y * y;
                s[t][5] += z * z;
            }
        }
        pairwise256(s);
        for (int q = 0; q < 6; ++q) P[c * 6 + q] = s[0][q];
    }
    for (int t = 0; t < kTreeLanes; ++t) {
        for (int q = 0; q < 6; ++q) s[t][q] = 0.0;
        for (int64_t c = t; c < nch; c += kTreeLanes)
            for (int q = 0; q < 6; ++q) s[t][q] += P[c * 6 + q];
    }
    pairwise256(s);
    for (int q = 0; q < 6; ++q) out[q] = s[0][q];
}

// ---------------------------------------------------------------------------
// Domain (space_charge_3d_open_hockney.cc:376-397, 493-540;
// core_diagnostics.cc:454-459; rectangular_grid_domain.h:20-36) and the
// scalars derived from it: deposit weight0 (deposit.cc:868-870),
// get_normalization_force (.cc:707-735), kick factor (.cc:759-770).
// ---------------------------------------------------------------------------
extern "C" int orc_compute_domain(const orc_options* opt, const orc_bunch* b,
                                  double time_step, orc_domain* d)
{
    std::memset(d, 0, sizeof(*d));
    double size[3], offset[3];
    if (opt->domain_fixed) {
        for (int a = 0; a < 3; ++a) {
            size[a] = opt->size[a];
            offset[a] = opt->offset[a];
        }
    } else {
        orc_tree_sums(b->parts, b->stride, b->n_active, b->masks, d->sums);
        double total = b->total_num;
        for (int j = 0; j < 3; ++j) {
            d->mean[j] = d->sums[j] / total;
            d->stddev[j] =
                std::sqrt(d->sums[j + 3] / total - d->mean[j] * d->mean[j]);
        }
        const double tiny = 1.0e-10;
        if ((d->stddev[0] < tiny) && (d->stddev[1] < tiny) &&
            (d->stddev[2] < tiny)) {
            g_err = "Space_charge_3d_open_hockney_eigen::update_domain: "
                    "all three spatial dimensions have neglible extent";
            return -3;
        }
        offset[0] = d->mean[0];
        offset[1] = d->mean[1];
        offset[2] = d->mean[2];
        size[0] = opt->n_sigma * smallest_non_tiny(d->stddev[0], d->stddev[1],
                                                   d->stddev[2], tiny);
        size[1] = opt->n_sigma * smallest_non_tiny(d->stddev[1], d->stddev[0],
                                                   d->stddev[2], tiny);
        size[2] = opt->n_sigma * smallest_non_tiny(d->stddev[2], d->stddev[0],
                                                   d->stddev[1], tiny);
        if (opt->grid_entire_period) {
            offset[2] = 0.0;
            size[2] = opt->z_period;
        }
    }
    for (int a = 0; a < 3; ++a) {
        d->size[a] = size[a];
        d->offset[a] = offset[a];
        d->left[a] = offset[a] - size[a] / 2.0;
        d->h[a] = size[a] / (1.0 * opt->shape[a]);
    }
    const double* h = d->h;
    d->w0 = (b->real_num / b->total_num) * b->charge * kE / (h[0] * h[1] * h[2]);

    double norm = h[0] * h[1] * h[2];
    norm *= 1.0 / (4.0 * kPi * kEpsilon0);
    norm *= 1.0;
    if (opt->green_fn == 1) norm *= 1.0 / (h[2] * h[2]);
    norm *= 1.0 / (2 * opt->shape[0] * 2 * opt->shape[1] * 2 * opt->shape[2]);
    d->fn_norm = norm;

    double q = b->charge * kE;
    double unit_conversion = kC / (1e9 * kE);
    d->factor = opt->kick_scale * unit_conversion * q * time_step * d->fn_norm /
                (b->ref_p * b->ref_gamma * b->ref_gamma * b->ref_beta);
    return 0;
}

// collective/utils.h:9-24.  The guard keeps the int conversion defined for
// positions far outside the grid (or NaN); every such particle is outside the
// grid either way.
extern "C" int32_t orc_cell_index(double pos, double left, double inv_h,
                                  double* off)
{
    double s = (pos - left) * inv_h - 0.5;
    if (!(s > -1.0e9 && s < 1.0e9)) {
        *off = 0.0;
        return -1000000000;
    }
    int32_t idx = static_cast<int32_t>(std::floor(s));
    *off = s - idx;
    return idx;
}

// Build-side cell key (the library's sort order, not the reference's: the
// reference does not sort).  Cells in bricks of 4 per axis (1 on an axis
// whose length 4 does not divide).  Bricks on a serpentine path: row
// r = bz*nby + by' with by' = by, reversed (nby-1-by) in odd layers bz;
// bx' = bx, reversed (nbx-1-bx) in odd rows r.  Cells inside one brick in
// (z, y, x) order with x fastest:
//   k = (r*nbx + bx') * (Bx*By*Bz) + (lz*By + ly)*Bx + lx
// for active particles whose lower corner lies in [0, g-2]^3 (exactly the
// particles that the deposit and the kick touch), else G = gx*gy*gz.
static int32_t orc_brick_key(int32_t ix, int32_t iy, int32_t iz, const int32_t g[3])
{
    const int32_t B[3] = {g[0] % 4 == 0 ? 4 : 1, g[1] % 4 == 0 ? 4 : 1, g[2] % 4 == 0 ? 4 : 1};
    const int32_t nbx = g[0] / B[0], nby = g[1] / B[1];
    const int32_t bz = iz / B[2];
    int32_t by = iy / B[1], bx = ix / B[0];
    if (bz % 2 == 1) by = nby - 1 - by;
    const int32_t r = bz * nby + by;
    if (r % 2 == 1) bx = nbx - 1 - bx;
    const int32_t brick = r * nbx + bx;
    const int32_t cell = ((iz % B[2]) * B[1] + iy % B[1]) * B[0] + ix % B[0];
    return brick * (B[0] * B[1] * B[2]) + cell;
}

extern "C" void orc_cell_keys(const orc_bunch* b, const orc_domain* d,
                              const int32_t g[3], int32_t* keys)
{
    const double ih[3] = {1.0 / d->h[0], 1.0 / d->h[1], 1.0 / d->h[2]};
    const int32_t G = g[0] * g[1] * g[2];
    for (int64_t i = 0; i < b->n_active; ++i) {
        double o;
        int32_t ix = orc_cell_index(b->parts[i], d->left[0], ih[0], &o);
        int32_t iy = orc_cell_index(b->parts[i + 2 * b->stride], d->left[1],
                                    ih[1], &o);
        int32_t iz = orc_cell_index(b->parts[i + 4 * b->stride], d->left[2],
                                    ih[2], &o);
        bool in = b->masks[i] && ix >= 0 && ix < g[0] - 1 && iy >= 0 &&
                  iy < g[1] - 1 && iz >= 0 && iz < g[2] - 1;
        keys[i] = in ? orc_brick_key(ix, iy, iz, g) : G;
    }
}

extern "C" void orc_stable_sort(const int32_t* keys, int64_t n, int32_t* perm)
{
    std::vector<int32_t> p(n);
    std::iota(p.begin(), p.end(), 0);
    std::stable_sort(p.begin(), p.end(),
                     [keys](int32_t a, int32_t b) { return keys[a] < keys[b]; });
    std::memcpy(perm, p.data(), sizeof(int32_t) * n);
}

// ---------------------------------------------------------------------------
// FFT.  The reference calls FFTW 3.3.10 (fftw_mpi_plan_dft_r2c_3d /
// _c2r_3d on dims (Dz, Dy, Dx), distributed_fft3d_fftw.cc:67-71): forward
// sign -1, unnormalised both ways, padded real rows of 2*(Dx/2+1).  Here:
// iterative radix-2 Cooley-Tukey for powers of two, a direct O(n^2) DFT
// otherwise; c2r uses the Hermitian extension of each x row with the
// imaginary parts of the DC and Nyquist terms ignored, as FFTW does.
// ---------------------------------------------------------------------------
namespace {
// Twiddle table exp(-2 pi i k / n), k < n/2, one per length and thread.
const std::vector<double>& twiddles(int n)
{
    thread_local std::vector<std::vector<double>> cache(32);
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    std::vector<double>& t = cache[lg];
    if (t.empty()) {
        t.resize(n);  // n/2 complex entries
        for (int k = 0; k < n / 2; ++k) {
            double ang = -2.0 * kPi * double(k) / n;
            t[2 * k] = std::cos(ang);
            t[2 * k + 1] = std::sin(ang);
        }
    }
    return t;
}

void fft_pow2(double* a, int n, int sign)
{
    const std::vector<double>& tw = twiddles(n);
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(a[2 * i], a[2 * j]);
            std::swap(a[2 * i + 1], a[2 * j + 1]);
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len / 2;
        int step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                // twiddle exp(sign * 2 pi i k / len) = table[k*step]
                double wr = tw[2 * (k * step)];
                double wi = sign < 0 ? tw[2 * (k * step) + 1]
                                     : -tw[2 * (k * step) + 1];
                double* u = a + 2 * (i + k);
                double* v = a + 2 * (i + k + half);
                double tr = v[0] * wr - v[1] * wi;
                double ti = v[0] * wi + v[1] * wr;
                v[0] = u[0] - tr;
                v[1] = u[1] - ti;
                u[0] = u[0] + tr;
                u[1] = u[1] + ti;
            }
        }
    }
}

void dft_direct(double* a, int n, int sign)
{
    std::vector<double> out(2 * n);
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < n; ++j) {
            long long m = (static_cast<long long>(j) * k) % n;
            double ang = sign * 2.0 * kPi * double(m) / n;
            double wr = std::cos(ang), wi = std::sin(ang);
            sr += a[2 * j] * wr - a[2 * j + 1] * wi;
            si += a[2 * j] * wi + a[2 * j + 1] * wr;
        }
        out[2 * k] = sr;
        out[2 * k + 1] = si;
    }
    std::memcpy(a, out.data(), sizeof(double) * 2 * n);
}
}  // namespace

extern "C" void orc_fft1d(double* data, int n, int sign)
{
    if (n <= 1) return;
    if ((n & (n - 1)) == 0)
        fft_pow2(data, n, sign);
    else
        dft_direct(data, n, sign);
}

extern "C" void orc_fft3d_r2c(int nx, int ny, int nz, const double* in,
                              double* out, int nthreads)
{
    const int px = padded_real(nx), hx = padded_cplx(nx);
    const int64_t nrows = int64_t(ny) * nz;
    // x rows: complex FFT of the real row, keep nx/2+1 outputs.
    std::vector<double> tmp(size_t(nrows) * 2 * hx);
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<double> buf(2 * std::max(nx, std::max(ny, nz)));
#pragma omp for schedule(static)
        for (int64_t r = 0; r < nrows; ++r) {
            const double* row = in + r * px;
            for (int x = 0; x < nx; ++x) {
                buf[2 * x] = row[x];
                buf[2 * x + 1] = 0.0;
            }
            orc_fft1d(buf.data(), nx, -1);
            std::memcpy(&tmp[size_t(r) * 2 * hx], buf.data(),
                        sizeof(double) * 2 * hx);
        }
#pragma omp barrier
        // y columns
#pragma omp for schedule(static)
        for (int64_t zc = 0; zc < int64_t(nz) * hx; ++zc) {
            int64_t z = zc / hx, c = zc % hx;
            for (int y = 0; y < ny; ++y) {
                size_t idx = (size_t(z * ny + y) * hx + c) * 2;
                buf[2 * y] = tmp[idx];
                buf[2 * y + 1] = tmp[idx + 1];
            }
            orc_fft1d(buf.data(), ny, -1);
            for (int y = 0; y < ny; ++y) {
                size_t idx = (size_t(z * ny + y) * hx + c) * 2;
                tmp[idx] = buf[2 * y];
                tmp[idx + 1] = buf[2 * y + 1];
            }
        }
#pragma omp barrier
        // z columns
#pragma omp for schedule(static)
        for (int64_t yc = 0; yc < int64_t(ny) * hx; ++yc) {
            for (int z = 0; z < nz; ++z) {
                size_t idx = (size_t(z) * ny * hx + yc) * 2;
                buf[2 * z] = tmp[idx];
                buf[2 * z + 1] = tmp[idx + 1];
            }
            orc_fft1d(buf.data(), nz, -1);
            for (int z = 0; z < nz; ++z) {
                size_t idx = (size_t(z) * ny * hx + yc) * 2;
                tmp[idx] = buf[2 * z];
                tmp[idx + 1] = buf[2 * z + 1];
            }
        }
    }
    std::memcpy(out, tmp.data(), sizeof(double) * tmp.size());
}

extern "C" void orc_fft3d_c2r(int nx, int ny, int nz, const double* in,
                              double* out, int nthreads)
{
    const int px = padded_real(nx), hx = padded_cplx(nx);
    const int64_t nrows = int64_t(ny) * nz;
    std::vector<double> tmp(in, in + size_t(nrows) * 2 * hx);
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<double> buf(2 * std::max(nx, std::max(ny, nz)));
#pragma omp for schedule(static)
        for (int64_t yc = 0; yc < int64_t(ny) * hx; ++yc) {
            for (int z = 0; z < nz; ++z) {
                size_t idx = (size_t(z) * ny * hx + yc) * 2;
                buf[2 * z] = tmp[idx];
                buf[2 * z + 1] = tmp[idx + 1];
            }
            orc_fft1d(buf.data(), nz, +1);
            for (int z = 0; z < nz; ++z) {
                size_t idx = (size_t(z) * ny * hx + yc) * 2;
                tmp[idx] = buf[2 * z];
                tmp[idx + 1] = buf[2 * z + 1];
            }
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t zc = 0; zc < int64_t(nz) * hx; ++zc) {
            int64_t z = zc / hx, c = zc % hx;
            for (int y = 0; y < ny; ++y) {
                size_t idx = (size_t(z * ny + y) * hx + c) * 2;
                buf[2 * y] = tmp[idx];
                buf[2 * y + 1] = tmp[idx + 1];
            }
            orc_fft1d(buf.data(), ny, +1);
            for (int y = 0; y < ny; ++y) {
                size_t idx = (size_t(z * ny + y) * hx + c) * 2;
                tmp[idx] = buf[2 * y];
                tmp[idx + 1] = buf[2 * y + 1];
            }
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t r = 0; r < nrows; ++r) {
            const double* X = &tmp[size_t(r) * 2 * hx];
            for (int k = 0; k < hx; ++k) {
                buf[2 * k] = X[2 * k];
                buf[2 * k + 1] = X[2 * k + 1];
            }
            buf[1] = 0.0;
            if (nx % 2 == 0) buf[2 * (nx / 2) + 1] = 0.0;
            for (int k = hx; k < nx; ++k) {
                buf[2 * k] = X[2 * (nx - k)];
                buf[2 * k + 1] = -X[2 * (nx - k) + 1];
            }
            orc_fft1d(buf.data(), nx, +1);
            double* row = out + r * px;
            for (int x = 0; x < nx; ++x) row[x] = buf[2 * x];
            for (int x = nx; x < px; ++x) row[x] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// Green's functions on the padded doubled grid
// (space_charge_3d_open_hockney.cc:136-348, launched :591-639), including the
// reference's float-reciprocal index decomposition and mirror writes.
// ---------------------------------------------------------------------------
extern "C" void orc_green_fn2(int green_fn, const int32_t g[3],
                              const double h[3], double* g2)
{
    const int dgx = 2 * g[0], dgy = 2 * g[1], dgz = 2 * g[2];
    const int pdx = padded_real(dgx);
    const int64_t tot = int64_t(pdx) * dgy * dgz;
    std::fill(g2, g2 + tot, 0.0);
    const double hx = h[0], hy = h[1], hz = h[2];
    const int gx = g[0] + 1, gy = g[1] + 1;
    const double igxgy = 1.0 / (gx * gy);
    const double igx = 1.0 / gx;
    auto at = [&](int iz, int iy, int ix) -> double& {
        return g2[int64_t(iz) * pdx * dgy + int64_t(iy) * pdx + ix];
    };
    if (green_fn == 0) {
        // alg_g2_pointlike (:136-228)
        const double G000 = 2.8 / std::min(hx, std::min(hy, hz));
        const int n = gx * gy * (g[2] + 1);
        for (int i = 0; i < n; ++i) {
            int iz = i * igxgy;
            int iy = (i - iz * gx * gy) * igx;
            int ix = i - iz * gx * gy - iy * gx;
            double dx = ix * hx, dy = iy * hy, dz = iz * hz;
            double G;
            if (ix == 0 && iy == 0 && iz == 0)
                G = G000;
            else
                G = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
            int mix = dgx - ix;
            if (mix == dgx) mix = 0;
            int miy = dgy - iy;
            if (miy == dgy) miy = 0;
            int miz = dgz - iz;
            if (miz == dgz) miz = 0;
            at(iz, iy, ix) = G;
            at(iz, miy, ix) = G;
            at(iz, iy, mix) = G;
            at(iz, miy, mix) = G;
            at(miz, iy, ix) = G;
            at(miz, miy, ix) = G;
            at(miz, iy, mix) = G;
            at(miz, miy, mix) = G;
        }
        return;
    }
    // alg_g2_linear (:230-348)
    const int gx0 = g[0], gy0 = g[1], gz0 = g[2];
    double rr_h = hx * hx + hy * hy;
    double r1_h = std::sqrt(hx * hx + hy * hy + hz * hz);
    const double G000 = (2.0 / rr_h) *
                        (hz * r1_h + rr_h * std::log((hz + r1_h) / std::sqrt(rr_h)) -
                         hz * hz);
    const int n = gx * gy * dgz;
    for (int i = 0; i < n; ++i) {
        int iz = i * igxgy;
        int iy = (i - iz * gx * gy) * igx;
        int ix = i - iz * gx * gy - iy * gx;
        double z = (iz > gz0) ? (dgz - iz) * hz : iz * hz;
        double y = iy * hy;
        double x = ix * hx;
        double rr = x * x + y * y;
        double s_rr_0 = std::sqrt(rr + z * z);
        double s_rr_n = std::sqrt(rr + (z - hz) * (z - hz));
        double s_rr_p = std::sqrt(rr + (z + hz) * (z + hz));
        double G = 2.0 * s_rr_0 - s_rr_n - s_rr_p;
        double T1, T2, r1, r2;
        const double epsz = 1e-12 * hz;
        if (z < -hz) {
            r1 = (s_rr_n - z + hz) / (s_rr_0 - z);
            T1 = (hz - z) * std::log(r1);
            r2 = (s_rr_0 - z) / (s_rr_p - z - hz);
            T2 = (hz + z) * std::log(r2);
            G += T1 + T2;
        } else if (std::abs(z + hz) < epsz) {
            r1 = (s_rr_n - z + hz) / (s_rr_0 - z);
            T1 = (hz - z) * std::log(r1);
            G += T1;
        } else if (std::abs(z) < epsz) {
            if (std::abs(x) + std::abs(y) < 2. * epsz) {
                G += hz * G000;
            } else {
                r1 = (std::sqrt(hz * hz + rr) + hz) / std::sqrt(rr);
                G += 2.0 * hz * std::log(r1);
            }
        } else if (std::abs(z - hz) < epsz) {
            r1 = (s_rr_p + z + hz) / (s_rr_0 + z);
            T1 = (hz + z) * std::log(r1);
            G += T1;
        } else if (z > hz) {
            r1 = (s_rr_0 + z) / (s_rr_n + z - hz);
            T1 = (hz - z) * std::log(r1);
            r2 = (s_rr_p + z + hz) / (s_rr_0 + z);
            T2 = (hz + z) * std::log(r2);
            G += T1 + T2;
        }
        int miy = dgy - iy;
        int mix = dgx - ix;
        if (mix == dgx) mix = 0;
        if (miy == dgy) miy = 0;
        at(iz, iy, ix) = G;
        if (iy != gy0) {
            at(iz, miy, ix) = G;
            if (ix != gx0) at(iz, miy, mix) = G;
        }
        if (ix != gx0) at(iz, iy, mix) = G;
    }
}

// ---------------------------------------------------------------------------
// CIC deposit, CPU form (deposit.cc:852-1031): per-thread private G grids,
// static particle split, z-slab reduction into the padded doubled rho2.
// ---------------------------------------------------------------------------
namespace {
inline bool ingrid(int ix, int iy, int iz, int gx, int gy, int gz)
{
    return ix > 0 && ix < gx - 1 && iy > 0 && iy < gy - 1 && iz > 0 &&
           iz < gz - 1;  // deposit.cc:21-30
}
}  // namespace

extern "C" void orc_deposit(const orc_bunch* b, const orc_domain* d,
                            const int32_t g[3], double* rho2, int nthreads)
{
    const int gx = g[0], gy = g[1], gz = g[2];
    const int dx = padded_real(2 * gx), dy = 2 * gy, dz = 2 * gz;
    const int64_t nc = int64_t(gx) * gy * gz;
    std::fill(rho2, rho2 + int64_t(dx) * dy * dz, 0.0);
    const double w0 = d->w0;
    const double lx = d->left[0], ly = d->left[1], lz = d->left[2];
    const double ihx = 1.0 / d->h[0], ihy = 1.0 / d->h[1], ihz = 1.0 / d->h[2];
    const int64_t npart = b->n_active;
    const double* p = b->parts;
    const int64_t st = b->stride;
    std::vector<double> rl(size_t(nthreads) * nc);
#pragma omp parallel num_threads(nthreads)
    {
        int nt = omp_get_num_threads();
        int it = omp_get_thread_num();
        int64_t le = npart / nt;
        int64_t ps = it * le;
        int64_t pe = (it == nt - 1) ? npart : (it + 1) * le;
        double* r = rl.data() + size_t(it) * nc;
        std::memset(r, 0, sizeof(double) * nc);
        for (int64_t n = ps; n < pe; ++n) {
            if (!b->masks[n]) continue;
            double ox, oy, oz;
            int ix = orc_cell_index(p[n], lx, ihx, &ox);
            int iy = orc_cell_index(p[n + 2 * st], ly, ihy, &oy);
            int iz = orc_cell_index(p[n + 4 * st], lz, ihz, &oz);
            int64_t base = int64_t(iz) * gx * gy;
            if (ingrid(ix, iy, iz, gx, gy, gz))
                r[base + iy * gx + ix] += w0 * (1 - ox) * (1 - oy) * (1 - oz);
            if (ingrid(ix + 1, iy, iz, gx, gy, gz))
                r[base + iy * gx + ix + 1] += w0 * (ox) * (1 - oy) * (1 - oz);
            if (ingrid(ix, iy + 1, iz, gx, gy, gz))
                r[base + (iy + 1) * gx + ix] += w0 * (1 - ox) * (oy) * (1 - oz);
            if (ingrid(ix + 1, iy + 1, iz, gx, gy, gz))
                r[base + (iy + 1) * gx + ix + 1] += w0 * (ox) * (oy) * (1 - oz);
            base = int64_t(iz + 1) * gx * gy;
            if (ingrid(ix, iy, iz + 1, gx, gy, gz))
                r[base + iy * gx + ix] += w0 * (1 - ox) * (1 - oy) * (oz);
            if (ingrid(ix + 1, iy, iz + 1, gx, gy, gz))
                r[base + iy * gx + ix + 1] += w0 * (ox) * (1 - oy) * (oz);
            if (ingrid(ix, iy + 1, iz + 1, gx, gy, gz))
                r[base + (iy + 1) * gx + ix] += w0 * (1 - ox) * (oy) * (oz);
            if (ingrid(ix + 1, iy + 1, iz + 1, gx, gy, gz))
                r[base + (iy + 1) * gx + ix + 1] += w0 * (ox) * (oy) * (oz);
        }
#pragma omp barrier
        int zle = gz / nt;
        int zps = it * zle;
        int zpe = (it == nt - 1) ? gz : (zps + zle);
        for (int z = zps; z < zpe; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    double w = 0.0;
                    for (int t = 0; t < nt; ++t)
                        w += rl[size_t(t) * nc + (int64_t(z) * gx * gy + y * gx + x)];
                    rho2[int64_t(z) * dx * dy + int64_t(y) * dx + x] = w;
                }
    }
}

// ---------------------------------------------------------------------------
// E field, zyx::alg_force_extractor_doubled_domain
// (space_charge_3d_kernels.h:217-307): interior nodes only.
// ---------------------------------------------------------------------------
extern "C" void orc_force(const double* phi2, const int32_t g[3],
                          const double h[3], double* enx, double* eny,
                          double* enz)
{
    const int gx = g[0], gy = g[1], gz = g[2];
    const int dgx = padded_real(2 * gx), dgy = 2 * gy;
    const double ihx = 0.5 / h[0], ihy = 0.5 / h[1], ihz = 0.5 / h[2];
    const double igxgy = 1.0 / (gx * gy);
    const double igx = 1.0 / gx;
    const int n = gx * gy * gz;
    std::fill(enx, enx + n, 0.0);
    std::fill(eny, eny + n, 0.0);
    std::fill(enz, enz + n, 0.0);
    for (int i = 0; i < n; ++i) {
        int iz = i * igxgy;
        int iy = (i - iz * gx * gy) * igx;
        int ix = i - iz * gx * gy - iy * gx;
        if (ix == 0 || ix == gx - 1 || iy == 0 || iy == gy - 1 || iz == 0 ||
            iz == gz - 1)
            continue;
        int64_t r = int64_t(iz) * dgx * dgy + int64_t(iy) * dgx + (ix + 1);
        int64_t l = int64_t(iz) * dgx * dgy + int64_t(iy) * dgx + (ix - 1);
        enx[i] = -(phi2[r] - phi2[l]) * ihx;
        r = int64_t(iz) * dgx * dgy + int64_t(iy + 1) * dgx + ix;
        l = int64_t(iz) * dgx * dgy + int64_t(iy - 1) * dgx + ix;
        eny[i] = -(phi2[r] - phi2[l]) * ihy;
        r = int64_t(iz + 1) * dgx * dgy + int64_t(iy) * dgx + ix;
        l = int64_t(iz - 1) * dgx * dgy + int64_t(iy) * dgx + ix;
        enz[i] = -(phi2[r] - phi2[l]) * ihz;
    }
}

// ---------------------------------------------------------------------------
// Kick, zyx::alg_kicker (space_charge_3d_kernels.h:91-215).
// ---------------------------------------------------------------------------
extern "C" void orc_kick(orc_bunch* b, const double* enx, const double* eny,
                         const double* enz, const int32_t g[3],
                         const orc_domain* d)
{
    const int gx = g[0], gy = g[1], gz = g[2];
    const double ihx = 1.0 / d->h[0], ihy = 1.0 / d->h[1], ihz = 1.0 / d->h[2];
    const double lx = d->left[0], ly = d->left[1], lz = d->left[2];
    const double factor = d->factor, pref = b->ref_p, m = b->mass;
    double* p = b->parts;
    const int64_t st = b->stride;
    const int64_t n = b->n_active;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        if (!b->masks[i]) continue;
        double ox, oy, oz;
        int ix = orc_cell_index(p[i], lx, ihx, &ox);
        int iy = orc_cell_index(p[i + 2 * st], ly, ihy, &oy);
        int iz = orc_cell_index(p[i + 4 * st], lz, ihz, &oz);
        double aox = 1.0 - ox, aoy = 1.0 - oy, aoz = 1.0 - oz;
        if (!((ix >= 0 && ix < gx - 1) && (iy >= 0 && iy < gy - 1) &&
              (iz >= 0 && iz < gz - 1)))
            continue;
        auto interp = [&](const double* e) {
            int64_t base = int64_t(iz) * gx * gy + int64_t(iy) * gx + ix;
            double val = aoz * aoy * aox * e[base];
            val += aoz * aoy * ox * e[base + 1];
            val += aoz * oy * aox * e[base + gx];
            val += aoz * oy * ox * e[base + gx + 1];
            base = int64_t(iz + 1) * gx * gy + int64_t(iy) * gx + ix;
            val += oz * aoy * aox * e[base];
            val += oz * aoy * ox * e[base + 1];
            val += oz * oy * aox * e[base + gx];
            val += oz * oy * ox * e[base + gx + 1];
            return val;
        };
        double val = interp(enz);
        double pp = pref + p[i + 5 * st] * pref;
        double Eoc_i = std::sqrt(pp * pp + m * m);
        double Eoc_f = Eoc_i + factor * (-pref) * val;
        double dpop =
            (std::sqrt(Eoc_f * Eoc_f - m * m) - std::sqrt(Eoc_i * Eoc_i - m * m)) /
            pref;
        p[i + 5 * st] += dpop;
        val = interp(eny);
        p[i + 3 * st] += factor * val;
        val = interp(enx);
        p[i + 1 * st] += factor * val;
    }
}

// ---------------------------------------------------------------------------
// One operator application on one bunch (apply_bunch, .cc:422-453).
// ---------------------------------------------------------------------------
extern "C" int orc_apply(const orc_options* opt, orc_bunch* b, double time_step,
                         int nthreads, orc_debug* dbg)
{
    try {
        if (nthreads <= 0) nthreads = omp_get_max_threads();
        omp_set_num_threads(nthreads);
        const int32_t g[3] = {opt->shape[0], opt->shape[1], opt->shape[2]};
        const int dgx = 2 * g[0], dgy = 2 * g[1], dgz = 2 * g[2];
        const int pdx = padded_real(dgx);
        const int64_t tot = int64_t(pdx) * dgy * dgz;
        const int64_t G = int64_t(g[0]) * g[1] * g[2];
        double t0 = wtime(), t1;
        double secs[6] = {0, 0, 0, 0, 0, 0};

        orc_domain d;
        int rc = orc_compute_domain(opt, b, time_step, &d);
        if (rc) return rc;
        t1 = wtime(); secs[0] = t1 - t0; t0 = t1;

        if (opt->reorder || (dbg && (dbg->keys || dbg->perm))) {
            const int64_t n = b->n_active;
            std::vector<int32_t> keys(n), perm(n);
            orc_cell_keys(b, &d, g, keys.data());
            orc_stable_sort(keys.data(), n, perm.data());
            if (dbg && dbg->keys) std::memcpy(dbg->keys, keys.data(), sizeof(int32_t) * n);
            if (dbg && dbg->perm) std::memcpy(dbg->perm, perm.data(), sizeof(int32_t) * n);
            if (opt->reorder) {
                std::vector<double> col(n);
                for (int c = 0; c < 7; ++c) {
                    double* pc = b->parts + c * b->stride;
                    for (int64_t j = 0; j < n; ++j) col[j] = pc[perm[j]];
                    std::memcpy(pc, col.data(), sizeof(double) * n);
                }
                std::vector<uint8_t> m(n);
                for (int64_t j = 0; j < n; ++j) m[j] = b->masks[perm[j]];
                std::memcpy(b->masks, m.data(), n);
            }
            t0 = wtime();
        }

        std::vector<double> rho2(tot), g2(tot), phi2(tot);
        orc_deposit(b, &d, g, rho2.data(), nthreads);
        if (dbg && dbg->rho) {
            for (int z = 0; z < g[2]; ++z)
                for (int y = 0; y < g[1]; ++y)
                    for (int x = 0; x < g[0]; ++x)
                        dbg->rho[(int64_t(z) * g[1] + y) * g[0] + x] =
                            rho2[int64_t(z) * pdx * dgy + int64_t(y) * pdx + x];
        }
        if (dbg && dbg->rho2_full)
            std::memcpy(dbg->rho2_full, rho2.data(), sizeof(double) * tot);
        t1 = wtime(); secs[1] = t1 - t0; t0 = t1;

        if (opt->periodic_z) {
            g_err = opt->green_fn == 0
                        ? "Space_charge_3d_open_hockney::get_green_fn2_pointlike: "
                          "periodic_z not yet implemented"
                        : "Space_charge_3d_open_hockney::get_green_fn2_linear: "
                          "periodic_z not yet implemented";
            return -4;
        }
        orc_green_fn2(opt->green_fn, g, d.h, g2.data());
        if (dbg && dbg->g2_full)
            std::memcpy(dbg->g2_full, g2.data(), sizeof(double) * tot);
        t1 = wtime(); secs[2] = t1 - t0; t0 = t1;

        orc_fft3d_r2c(dgx, dgy, dgz, rho2.data(), rho2.data(), nthreads);
        orc_fft3d_r2c(dgx, dgy, dgz, g2.data(), g2.data(), nthreads);
        if (dbg && dbg->ghat_full)
            std::memcpy(dbg->ghat_full, g2.data(), sizeof(double) * tot);
        const int64_t ncplx = int64_t(dgz) * dgy * padded_cplx(dgx);
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int64_t i = 0; i < ncplx; ++i) {
            // alg_cplx_multiplier (.cc:350-372)
            const double ar = rho2[2 * i], ai = rho2[2 * i + 1];
            const double br = g2[2 * i], bi = g2[2 * i + 1];
            phi2[2 * i] = ar * br - ai * bi;
            phi2[2 * i + 1] = ar * bi + ai * br;
        }
        orc_fft3d_c2r(dgx, dgy, dgz, phi2.data(), phi2.data(), nthreads);
        t1 = wtime(); secs[3] = t1 - t0; t0 = t1;

        std::vector<double> ex(G), ey(G), ez(G);
        orc_force(phi2.data(), g, d.h, ex.data(), ey.data(), ez.data());
        t1 = wtime(); secs[4] = t1 - t0; t0 = t1;

        orc_kick(b, ex.data(), ey.data(), ez.data(), g, &d);
        t1 = wtime(); secs[5] = t1 - t0;

        if (dbg) {
            if (dbg->domain) *dbg->domain = d;
            for (int k = 0; k < 6; ++k) dbg->seconds[k] = secs[k];
            auto octant = [&](const std::vector<double>& src, double* dst) {
                for (int z = 0; z < g[2]; ++z)
                    for (int y = 0; y < g[1]; ++y)
                        for (int x = 0; x < g[0]; ++x)
                            dst[(int64_t(z) * g[1] + y) * g[0] + x] =
                                src[int64_t(z) * pdx * dgy + int64_t(y) * pdx + x];
            };
            if (dbg->phi) octant(phi2, dbg->phi);
            if (dbg->ex) std::memcpy(dbg->ex, ex.data(), sizeof(double) * G);
            if (dbg->ey) std::memcpy(dbg->ey, ey.data(), sizeof(double) * G);
            if (dbg->ez) std::memcpy(dbg->ez, ez.data(), sizeof(double) * G);
        }
        return 0;
    } catch (std::exception const& e) {
        g_err = e.what();
        return -1;
    }
}

// ---------------------------------------------------------------------------
// libFF maps: drift_unit (ff_algorithm.h:45-79), thin_quadrupole_unit
// (:928-936), yoshida6 (:1835-2008) as used by FF_quadrupole::apply
// (ff_quadrupole.h:438-468) and the design-reference cdt
// (ff_drift.h:68-87, ff_quadrupole.h:262-300).
// ---------------------------------------------------------------------------
namespace {
void ff_drift_unit(double& x, double xp, double& y, double yp, double& cdt, double dpop,
                   double length, double reference_momentum, double m, double reference_cdt)
{
    double uni = 1.0;
    double sig = double((0.0 < length) - (length < 0.0));
    double dp = dpop + uni;
    double inv_npz = uni / std::sqrt(dp * dp - xp * xp - yp * yp);
    double lxpr = xp * length * inv_npz;
    double lypr = yp * length * inv_npz;
    double D2 = lxpr * lxpr + length * length + lypr * lypr;
    double p = dp * reference_momentum;
    double E2 = p * p + m * m;
    double ibeta2 = E2 / (p * p);
    x = x + lxpr;
    y = y + lypr;
    cdt = cdt + sig * std::sqrt(D2 * ibeta2) - reference_cdt;
}

void ff_thin_quad(double x, double& xp, double y, double& yp, const double* kL)
{
    xp = xp - kL[0] * x + kL[1] * y;
    yp = yp + kL[0] * y + kL[1] * x;
}

void ff_yoshida6(double& x, double& xp, double& y, double& yp, double& cdt, double dpop,
                 double ref_p, double m, double step_ref_cdt, double step_length,
                 const double* ss, int steps)
{
    const double c1 = 0.79361246386112147294625603763;
    const double c2 = -0.206276584816439780698721319354;
    const double c3 = -0.118008867881292655922412133144;
    const double c4 = 0.236949573653050744373598734219;
    const double d1 = 1.58722492772224294589251207526;
    const double d2 = -1.99977809735512250728995471397;
    const double d3 = -1.82324266348482825773733634154;
    const double d4 = 2.29714181079092974648453380998;
    double sub = step_ref_cdt * 0.1;
    double k1[2] = {d1 * ss[0], d1 * ss[1]}, k2[2] = {d2 * ss[0], d2 * ss[1]};
    double k3[2] = {d3 * ss[0], d3 * ss[1]}, k4[2] = {d4 * ss[0], d4 * ss[1]};
    const double cs[10] = {c1, c2, c2, c3, c4, c4, c3, c2, c2, c1};
    const double* ks[9] = {k1, k2, k1, k3, k4, k3, k1, k2, k1};
    for (int i = 0; i < steps; ++i) {
        for (int d = 0; d < 10; ++d) {
            ff_drift_unit(x, xp, y, yp, cdt, dpop, cs[d] * step_length, ref_p, m, sub);
            if (d < 9) ff_thin_quad(x, xp, y, yp, ks[d]);
        }
    }
}

void ff_slice(const orc_slice& s, double& x, double& xp, double& y, double& yp, double& cdt,
              double dpop, double ref_p, double m)
{
    if (s.kind == 0) {
        ff_drift_unit(x, xp, y, yp, cdt, dpop, s.length, ref_p, m, s.ref_cdt);
    } else {
        int steps = s.yoshida_steps;
        double kn[2] = {s.k1 * (s.length / steps), s.k1s * (s.length / steps)};
        ff_yoshida6(x, xp, y, yp, cdt, dpop, ref_p, m, s.ref_cdt / steps, s.length / steps, kn,
                    steps);
    }
}
}  // namespace

extern "C" double orc_reference_cdt(const orc_slice* s, double ref_p, double mass)
{
    double x = 0, xp = 0, y = 0, yp = 0, cdt = 0;
    orc_slice z = *s;
    z.ref_cdt = 0.0;
    ff_slice(z, x, xp, y, yp, cdt, 0.0, ref_p, mass);
    return cdt;
}

extern "C" void orc_apply_maps(orc_bunch* b, const orc_slice* s, int nslices)
{
    double* p = b->parts;
    const int64_t st = b->stride;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < b->n_active; ++i) {
        if (!b->masks[i]) continue;
        for (int k = 0; k < nslices; ++k)
            ff_slice(s[k], p[i], p[i + st], p[i + 2 * st], p[i + 3 * st], p[i + 4 * st],
                     p[i + 5 * st], b->ref_p, b->mass);
    }
}

// ---------------------------------------------------------------------------
// Moments for diagnostics with the fixed tree (chunks of 4096, 16 per lane,
// pairwise lanes): mean (calculate_mean) then centred second moments
// (particle_reducer<mom2_tag>, calculate_mom2).
// ---------------------------------------------------------------------------
namespace {
template <int K, typename F>
void tree_generic(int64_t n, const uint8_t* masks, F&& add, double out[K])
{
    int64_t nch = (n + kTreeChunk - 1) / kTreeChunk;
    std::vector<double> P(size_t(nch) * K);
    std::vector<double> s(size_t(kTreeLanes) * K);
    auto pairwise = [&]() {
        for (int w = 1; w < kTreeLanes; w *= 2)
            for (int t = 0; t < kTreeLanes; t += 2 * w)
                for (int q = 0; q < K; ++q) s[t * K + q] = s[t * K + q] + s[(t + w) * K + q];
    };
    for (int64_t c = 0; c < nch; ++c) {
        std::fill(s.begin(), s.end(), 0.0);
        for (int t = 0; t < kTreeLanes; ++t)
            for (int j = 0; j < kTreeItems; ++j) {
                int64_t i = c * kTreeChunk + int64_t(j) * kTreeLanes + t;
                if (i >= n || !masks[i]) continue;
                add(i, &s[t * K]);
            }
        pairwise();
        for (int q = 0; q < K; ++q) P[c * K + q] = s[q];
    }
    std::fill(s.begin(), s.end(), 0.0);
    for (int t = 0; t < kTreeLanes; ++t)
        for (int64_t c = t; c < nch; c += kTreeLanes)
            for (int q = 0; q < K; ++q) s[t * K + q] += P[c * K + q];
    pairwise();
    for (int q = 0; q < K; ++q) out[q] = s[q];
}
}  // namespace

extern "C" void orc_moments(const double* p, int64_t st, int64_t n, const uint8_t* masks,
                            double total_num, double mean[6], double mom2[36])
{
    double s1[6];
    tree_generic<6>(n, masks, [&](int64_t i, double* acc) {
        for (int j = 0; j < 6; ++j) acc[j] += p[i + j * st];
    }, s1);
    for (int j = 0; j < 6; ++j) mean[j] = s1[j] / total_num;
    double s2[21];
    tree_generic<21>(n, masks, [&](int64_t i, double* acc) {
        int q = 0;
        for (int j = 0; j < 6; ++j) {
            double dj = p[i + j * st] - mean[j];
            for (int k = 0; k <= j; ++k) acc[q++] += dj * (p[i + k * st] - mean[k]);
        }
    }, s2);
    int q = 0;
    for (int j = 0; j < 6; ++j)
        for (int k = 0; k <= j; ++k) {
            mom2[j * 6 + k] = mom2[k * 6 + j] = s2[q] / total_num;
            ++q;
        }
}

// ---------------------------------------------------------------------------
// adjust_moments (populate.cc:30-61, adjust_moments_host in
// populate_host.cc:8-44): the particles' mean and second moments are made
// exactly `means` and `cov`:  A = G H^-1 with G, H the lower Cholesky
// factors of cov and of the bunch's second moments (Eigen LLT there), then
// every row p of the first n slots becomes (p - bunch mean) A^T + means.
// is_symmetric66 (populate.cc:12-27) with floating_point_equal, tolerance
// 1e-14 (utils/floating_point.h:9-18).  Cholesky-Banachiewicz row by row,
// sums in ascending k; H^-1 by forward substitution; (p - mean) A^T summed
// in ascending k.  Returns 0, -1 (not symmetric), -2 (not positive
// definite).
// ---------------------------------------------------------------------------
namespace {
bool fp_equal(double a, double b, double tol)
{
    if (std::abs(a) < tol) return std::abs(a - b) < tol;
    return std::abs((a - b) / a) < tol;
}

bool cholesky6(const double* C, double* L)
{
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = C[i * 6 + j];
            for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i * 6 + i] = std::sqrt(s);
            } else {
                L[i * 6 + j] = s / L[j * 6 + j];
            }
        }
    return true;
}

void lower_inverse6(const double* L, double* M)
{
    for (int i = 0; i < 36; ++i) M[i] = 0.0;
    for (int c = 0; c < 6; ++c) {
        M[c * 6 + c] = 1.0 / L[c * 6 + c];
        for (int i = c + 1; i < 6; ++i) {
            double s = 0.0;
            for (int k = c; k < i; ++k) s += L[i * 6 + k] * M[k * 6 + c];
            M[i * 6 + c] = -s / L[i * 6 + i];
        }
    }
}
}  // namespace

extern "C" int orc_adjust_matrix(const double cov[36], const double mom2[36], double A[36])
{
    for (int i = 0; i < 6; ++i)
        for (int j = i + 1; j < 6; ++j)
            if (!fp_equal(cov[i * 6 + j], cov[j * 6 + i], 1.0e-14)) return -1;
    double G[36], H[36], Hi[36];
    if (!cholesky6(cov, G) || !cholesky6(mom2, H)) return -2;
    lower_inverse6(H, Hi);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += G[i * 6 + k] * Hi[k * 6 + j];
            A[i * 6 + j] = s;
        }
    return 0;
}

extern "C" int orc_adjust_moments(double* p, int64_t st, int64_t n, const uint8_t* masks,
                                  double total_num, const double means[6],
                                  const double cov[36])
{
    double mean[6], mom2[36], A[36];
    orc_moments(p, st, n, masks, total_num, mean, mom2);
    if (int rc = orc_adjust_matrix(cov, mom2, A)) return rc;
    for (int64_t i = 0; i < n; ++i) {
        double d[6], o[6];
        for (int k = 0; k < 6; ++k) d[k] = p[i + k * st] - mean[k];
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += d[k] * A[j * 6 + k];
            o[j] = s + means[j];
        }
        for (int j = 0; j < 6; ++j) p[i + j * st] = o[j];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Gaussian populate of the test bunches (the reference draws with GSL's
// ranlxd2 and gsl_ran_gaussian in populate_6d, populate.cc:128-145, then
// applies the Cholesky factor of the covariance; the build replaces the
// generator by a counter-based one so that host and device produce the same
// bunch).  Published algorithms restated here:
//   Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//   easy as 1, 2, 3", SC'11): round (c0..c3) -> (hi(M1 c2) ^ c1 ^ k0,
//   lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0)), M0 = 0xD2511F53,
//   M1 = 0xCD9E8D57; key bumped by (0x9E3779B9, 0xBB67AE85) between rounds.
//   Counter (id lo, id hi, call, 0), key (seed lo, seed hi), calls 0..2.
//   Uniform: the top 53 bits m of (r0, r1) or (r2, r3), (m + 0.5) 2^-53.
//   Box-Muller (1958): g = sqrt(-2 log u1) (cos 2 pi u2, sin 2 pi u2), with
//   log and sin/cos evaluated from IEEE basic operations only:
//     log: u = m 2^e, m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f/(2+f),
//          log(1+f) = f - (f^2/2 - s (f^2/2 + R)), R = sum_{k=1..11}
//          2 s^2k/(2k+1) (the arrangement of fdlibm's e_log.c), plus e ln 2
//          with ln 2 split into a 32-bit head and its tail;
//     sin/cos: quadrant of 4u, angle folded to [0, pi/4], Taylor series to
//          x^19 / x^20, exact quadrant rotation.
//   Coordinates: x_r = mean_r + sum_{c<=r} L_rc g_c in that order (L packed
//   lower-triangular, row-major); column 6 = id0 + slot.
// ---------------------------------------------------------------------------
extern "C" void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = uint32_t(p1);
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = uint32_t(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

extern "C" double orc_det_log(double u)
{
    const double ln2_head = 0x1.62e42feep-1;
    const double ln2_tail = 0x1.a39ef35793c76p-33;
    int e = 0;
    double m = std::frexp(u, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {  // sqrt(1/2)
        m = m + m;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 2.0 / 23.0;
    for (int k = 10; k >= 1; --k) R = 2.0 / double(2 * k + 1) + z * R;
    R = z * R;
    const double hfsq = 0.5 * f * f;
    const double de = double(e);
    return de * ln2_head + (f - (hfsq - (s * (hfsq + R) + de * ln2_tail)));
}

extern "C" void orc_det_sincos2pi(double u, double* sn, double* cs)
{
    const double half_pi = 0x1.921fb54442d18p+0;
    const double t = 4.0 * u;
    const double q = std::floor(t);
    const double r = t - q;
    const bool fold = r > 0.5;
    const double x = (fold ? 1.0 - r : r) * half_pi;
    const double z = x * x;
    // (2k+1)! and (2k)! are exact doubles up to 20!
    double fact[21];
    fact[0] = 1.0;
    for (int k = 1; k <= 20; ++k) fact[k] = fact[k - 1] * double(k);
    double S = -1.0 / fact[19];
    for (int k = 8; k >= 1; --k) S = ((k % 2) ? -1.0 : 1.0) / fact[2 * k + 1] + z * S;
    const double sin_x = x + x * (z * S);
    double Cq = 1.0 / fact[20];
    for (int k = 9; k >= 1; --k) Cq = ((k % 2) ? -1.0 : 1.0) / fact[2 * k] + z * Cq;
    const double cos_x = 1.0 + z * Cq;
    const double sa = fold ? cos_x : sin_x, ca = fold ? sin_x : cos_x;
    switch (int(q) & 3) {
    case 0: *sn = sa; *cs = ca; break;
    case 1: *sn = ca; *cs = -sa; break;
    case 2: *sn = -sa; *cs = -ca; break;
    default: *sn = -ca; *cs = sa; break;
    }
}

extern "C" void orc_populate_gaussian(double* parts, int64_t stride, int64_t n, uint8_t* masks,
                                      const double mean[6], const double chol[21],
                                      uint64_t seed, int64_t id0)
{
    const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t id = uint64_t(id0 + i);
        double u[6];
        for (uint32_t call = 0; call < 3; ++call) {
            const uint32_t ctr[4] = {uint32_t(id), uint32_t(id >> 32), call, 0u};
            uint32_t r[4];
            orc_philox4x32_10(ctr, key, r);
            for (int h = 0; h < 2; ++h) {
                const uint64_t m = (uint64_t(r[2 * h]) << 21) | (r[2 * h + 1] >> 11);
                u[2 * call + h] = (double(m) + 0.5) * 0x1.0p-53;
            }
        }
        double g[6];
        for (int k = 0; k < 3; ++k) {
            const double rad = std::sqrt(-2.0 * orc_det_log(u[2 * k]));
            double sn, cs;
            orc_det_sincos2pi(u[2 * k + 1], &sn, &cs);
            g[2 * k] = rad * cs;
            g[2 * k + 1] = rad * sn;
        }
        for (int r = 0; r < 6; ++r) {
            double v = mean[r];
            for (int c = 0; c <= r; ++c) v += chol[r * (r + 1) / 2 + c] * g[c];
            parts[i + r * stride] = v;
        }
        parts[i + 6 * stride] = double(id0 + i);
        masks[i] = 1;
    }
}
