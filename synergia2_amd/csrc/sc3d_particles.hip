// sc3d_particles.hip -- particle-side kernels of the MI355X space-charge step:
// bunch extent, cell keys + stable radix sort + in-place reorder, CIC
// deposit, E field, gather-interpolate-kick, and the Gaussian populate.
#include "sc3d_particles.h"

namespace sc3d {

// ---------------------------------------------------------------------------
// Served form: opens one apply of a scheduled job (see StepCtl).  Schedule
// table layout: [first | between | last], turn_steps programs each; step j
// of the job is schedule entry k = j mod turn_steps; the maps before the
// kick run on the first step only, and the program fused after the kick is
// between[k] (the second half-step of k and the first of k+1) except after
// the last step (last[k]).
// ---------------------------------------------------------------------------
__global__ void k_step_begin(StepCtl* __restrict__ ctl)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    StepCtl c = *ctl;
    const int64_t j = c.applies;
    if (c.turn_steps > 0 && j >= c.total_steps) {
        c.done = 1;
        *ctl = c;
        return;
    }
    c.done = 0;
    if (c.turn_steps > 0) {
        const int k = int(j % c.turn_steps);
        c.cur = k;
        c.first_prog = j == 0 ? k : -1;
        c.after_prog = j < c.total_steps - 1 ? c.turn_steps + k : 2 * c.turn_steps + k;
    }
    c.applies = j + 1;
    *ctl = c;
}

// ---------------------------------------------------------------------------
// Bunch extent: the six sums of Core_diagnostics::calculate_spatial_mean_stddev
// (core_diagnostics.cc:139-153) in one fixed tree, identical for every launch
// configuration and device:
//   chunk c = particles [4096c, 4096c+4096); lane t adds, from +0.0 and in
//   order j = 0..15, particle 4096c + 256j + t (masked and >= n skipped);
//   lanes combine pairwise (s[t] += s[t+w], w = 1..128), first inside each
//   64-wide wavefront by xor butterflies, then across the four wavefronts;
//   chunk partials combine by the same rule in k_domain_finish.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_pairwise_sum(double v);

__device__ __forceinline__ void block_pairwise6(double (&s)[6], double (*ws)[6])
{
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = wave_pairwise_sum(s[q]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 6; ++q) ws[wave][q] = s[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 6; ++q) s[q] = (ws[0][q] + ws[1][q]) + (ws[2][q] + ws[3][q]);
}

template <int K>
__device__ __forceinline__ void block_pairwiseK(double (&s)[K], double (*ws)[K])
{
#pragma unroll
    for (int q = 0; q < K; ++q) s[q] = wave_pairwise_sum(s[q]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < K; ++q) ws[wave][q] = s[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < K; ++q) s[q] = (ws[0][q] + ws[1][q]) + (ws[2][q] + ws[3][q]);
}

// Moments for diagnostics (Core_diagnostics::calculate_mean / calculate_mom2,
// core_diagnostics.cc:126-198,202-229,318-362): MODE 0 sums the six
// coordinates; MODE 1 sums the 21 centred products (p_j - mean_j)(p_k -
// mean_k), k <= j, with the mean read from `mean`.  Same fixed tree as the
// bunch extent (chunks of 4096, 16 per lane, pairwise lanes).
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_moments_partial(const double* __restrict__ parts,
                                                              int64_t stride,
                                                              const uint8_t* __restrict__ masks,
                                                              int64_t n,
                                                              const double* __restrict__ mean,
                                                              double* __restrict__ P)
{
    constexpr int K = MODE == 0 ? 6 : 21;
    __shared__ double ws[4][K];
    double s[K];
#pragma unroll
    for (int q = 0; q < K; ++q) s[q] = 0.0;
    double m[6] = {0, 0, 0, 0, 0, 0};
    if (MODE == 1)
        for (int j = 0; j < 6; ++j) m[j] = mean[j];
    const int64_t c = blockIdx.x;
    for (int jj = 0; jj < 16; ++jj) {
        int64_t i = c * 4096 + int64_t(jj) * 256 + threadIdx.x;
        if (i < n && masks[i]) {
            double v[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) v[j] = parts[i + j * stride];
            if (MODE == 0) {
#pragma unroll
                for (int j = 0; j < 6; ++j) s[j] += v[j];
            } else {
                int q = 0;
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    const double dj = v[j] - m[j];
#pragma unroll
                    for (int k = 0; k <= j; ++k) s[q++] += dj * (v[k] - m[k]);
                }
            }
        }
    }
    block_pairwiseK<K>(s, ws);
    if (threadIdx.x == 0)
        for (int q = 0; q < K; ++q) P[c * K + q] = s[q];
}

// Combines the chunk partials by the same rule and divides by total_num.
template <int K>
__global__ __launch_bounds__(kThreads) void k_moments_finish(const double* __restrict__ P,
                                                             int64_t nch, double total_num,
                                                             double* __restrict__ out)
{
    __shared__ double ws[4][K];
    double s[K];
#pragma unroll
    for (int q = 0; q < K; ++q) s[q] = 0.0;
    for (int64_t c = threadIdx.x; c < nch; c += kThreads)
#pragma unroll
        for (int q = 0; q < K; ++q) s[q] += P[c * K + q];
    block_pairwiseK<K>(s, ws);
    if (threadIdx.x == 0)
        for (int q = 0; q < K; ++q) out[q] = s[q] / total_num;
}

// ---------------------------------------------------------------------------
// Tail counts (diagnostic of the halo comparison of long tracking): per
// plane j, active particles with |p_j - c_j| > lim_j.  Integer counts, so
// the result does not depend on the reduction order.
// ---------------------------------------------------------------------------
__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n)
{
    if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(kThreads) void k_tail_counts(const double* __restrict__ parts,
                                                          int64_t stride,
                                                          const uint8_t* __restrict__ masks,
                                                          int64_t n, TailArgs a,
                                                          unsigned long long* __restrict__ counts)
{
    unsigned int cnt[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads) {
        if (!masks[i]) continue;
#pragma unroll
        for (int j = 0; j < 6; ++j) cnt[j] += fabs(parts[i + j * stride] - a.c[j]) > a.lim[j];
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
#pragma unroll
        for (int w = 1; w < 64; w *= 2) cnt[j] += __shfl_xor(cnt[j], w, 64);
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (cnt[j]) atomicAdd(&counts[j], (unsigned long long)cnt[j]);
}

void launch_moments(const double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                    double total_num, double* P, double* out /* 6 + 21 */, hipStream_t st)
{
    int64_t nch = (n + 4095) / 4096;
    if (nch > 0)
        hipLaunchKernelGGL(k_moments_partial<0>, dim3(unsigned(nch)), dim3(kThreads), 0, st,
                           parts, stride, masks, n, (const double*)nullptr, P);
    hipLaunchKernelGGL(k_moments_finish<6>, dim3(1), dim3(kThreads), 0, st, P, nch, total_num,
                       out);
    if (nch > 0)
        hipLaunchKernelGGL(k_moments_partial<1>, dim3(unsigned(nch)), dim3(kThreads), 0, st,
                           parts, stride, masks, n, (const double*)out, P);
    hipLaunchKernelGGL(k_moments_finish<21>, dim3(1), dim3(kThreads), 0, st, P, nch, total_num,
                       out + 6);
}

// space_charge_3d_open_hockney.cc:94-109
__device__ double smallest_non_tiny(double val, double o1, double o2, double tiny)
{
    if (val > tiny) return val;
    if ((o1 > tiny) && (o2 > tiny)) return fmin(o1, o2);
    return fmax(o1, o2);
}

// Finishes the domain on the device (update_domain, .cc:493-540, with the
// host finish of core_diagnostics.cc:454-459 and Rectangular_grid_domain,
// rectangular_grid_domain.h:20-36) and derives weight0 (deposit.cc:608-610),
// the force normalisation (.cc:707-735) and the kick factor (.cc:759-770).
// Also decides whether the cached Green's spectrum is still valid (same h
// bits) or must be recomputed this step.
__global__ __launch_bounds__(kThreads) void k_domain_finish(const double* __restrict__ P,
                                                            int64_t nch, DomainFinishArgs a,
                                                            DomainRec* __restrict__ dom, Sched sc)
{
    __shared__ double ws[4][6];
    if (sc.ctl) {
        // served form: a finished job's applies do nothing; the kick time
        // step comes from the schedule
        if (sc.ctl->done) {
            if (threadIdx.x == 0) dom->error = kDoneMark;
            return;
        }
        if (sc.ctl->turn_steps > 0) a.time_step = sc.dt[sc.ctl->cur];
    }
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!a.fixed) {
        // lane t adds chunks t, t + 256, ... in order; eight chunks' loads are
        // issued before their adds
        constexpr int U = 8;
        for (int64_t c0 = threadIdx.x; c0 < nch; c0 += U * kThreads) {
            double v[U][6];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t c = c0 + int64_t(u) * kThreads;
                if (c < nch)
#pragma unroll
                    for (int q = 0; q < 6; ++q) v[u][q] = P[c * 6 + q];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (c0 + int64_t(u) * kThreads < nch)
#pragma unroll
                    for (int q = 0; q < 6; ++q) s[q] += v[u][q];
        }
        block_pairwise6(s, ws);
    }
    if (threadIdx.x != 0) return;
    DomainRec d = *dom;  // keeps the Green's cache fields
    double size[3], offset[3];
    d.error = 0;
    d.rho_peak = 0;
    if (a.fixed) {
        for (int k = 0; k < 3; ++k) {
            size[k] = a.size[k];
            offset[k] = a.offset[k];
            d.sums[k] = d.sums[k + 3] = 0.0;
            d.mean[k] = d.stddev[k] = 0.0;
        }
    } else {
        for (int q = 0; q < 6; ++q) d.sums[q] = s[q];
        for (int j = 0; j < 3; ++j) {
            d.mean[j] = s[j] / a.total_num;
            d.stddev[j] = sqrt(s[j + 3] / a.total_num - d.mean[j] * d.mean[j]);
        }
        const double tiny = 1.0e-10;
        if ((d.stddev[0] < tiny) && (d.stddev[1] < tiny) && (d.stddev[2] < tiny)) {
            d.error = SC3D_ERR_DOMAIN;
            d.ghat_dirty = 0;
            *dom = d;
            return;
        }
        offset[0] = d.mean[0];
        offset[1] = d.mean[1];
        offset[2] = d.mean[2];
        size[0] = a.n_sigma * smallest_non_tiny(d.stddev[0], d.stddev[1], d.stddev[2], tiny);
        size[1] = a.n_sigma * smallest_non_tiny(d.stddev[1], d.stddev[0], d.stddev[2], tiny);
        size[2] = a.n_sigma * smallest_non_tiny(d.stddev[2], d.stddev[0], d.stddev[1], tiny);
        if (a.grid_entire_period) {
            offset[2] = 0.0;
            size[2] = a.z_period;
        }
    }
    for (int k = 0; k < 3; ++k) {
        d.size[k] = size[k];
        d.offset[k] = offset[k];
        d.left[k] = offset[k] - size[k] / 2.0;
        d.h[k] = size[k] / (1.0 * a.shape[k]);
        d.inv_h[k] = 1.0 / d.h[k];
        d.half_inv_h[k] = 0.5 / d.h[k];
    }
    const double* h = d.h;
    d.w0 = (a.real_num / a.total_num) * a.charge * kE / (h[0] * h[1] * h[2]);
    double norm = h[0] * h[1] * h[2];
    norm *= 1.0 / (4.0 * kPi * kEpsilon0);
    norm *= 1.0;
    if (a.green_fn == SC3D_GREEN_LINEAR) norm *= 1.0 / (h[2] * h[2]);
    norm *= 1.0 / (2 * a.shape[0] * 2 * a.shape[1] * 2 * a.shape[2]);
    d.fn_norm = norm;
    double q = a.charge * kE;
    double unit_conversion = kC / (1e9 * kE);
    d.factor = a.kick_scale * unit_conversion * q * a.time_step * d.fn_norm /
               (a.ref_p * a.ref_gamma * a.ref_gamma * a.ref_beta);
    bool same = d.ghat_valid && __double_as_longlong(d.ghat_h[0]) == __double_as_longlong(h[0]) &&
                __double_as_longlong(d.ghat_h[1]) == __double_as_longlong(h[1]) &&
                __double_as_longlong(d.ghat_h[2]) == __double_as_longlong(h[2]);
    d.ghat_dirty = same ? 0 : 1;
    for (int k = 0; k < 3; ++k) d.ghat_h[k] = h[k];
    d.ghat_valid = 1;
    *dom = d;
}

// The chunk sums and the clearing of the density grid in one launch: every
// block clears its slice of the grid (16-byte stores; the grid of the
// previous apply was consumed by its solve) and writes its chunk partials.
// (Finishing in the last block to count itself in, behind device-scope
// fences, measured slower: 41 against 25 us per apply at 4e6 particles, the
// fence of each of the 977 blocks writing back its L2.)
__global__ __launch_bounds__(kThreads) void k_domain_partial(
    const double* __restrict__ parts, int64_t stride, const uint8_t* __restrict__ masks,
    int64_t n, double* __restrict__ P, const StepCtl* __restrict__ ctl, uint4* __restrict__ grid,
    int64_t grid16)
{
    __shared__ double ws[4][6];
    if (ctl && ctl->done) return;
    if (grid16 > 0) {
        const int64_t per = (grid16 + gridDim.x - 1) / gridDim.x;
        const int64_t g0 = int64_t(blockIdx.x) * per;
        const int64_t g1 = g0 + per < grid16 ? g0 + per : grid16;
        for (int64_t i = g0 + threadIdx.x; i < g1; i += kThreads) grid[i] = make_uint4(0, 0, 0, 0);
    }
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int64_t i0 = int64_t(blockIdx.x) * 4096 + threadIdx.x;
    // All 16 rounds' loads first (none waits for a mask), then the sums in the
    // fixed order (measured at 1.6e7 particles: 108 -> 103 us per step for the
    // domain).
    double x[16], y[16], z[16];
    uint8_t on[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int64_t i = i0 + int64_t(j) * 256;
        on[j] = 0;
        if (i < n) {
            on[j] = masks[i];  // tested only after every load is issued
            x[j] = parts[i];
            y[j] = parts[i + 2 * stride];
            z[j] = parts[i + 4 * stride];
        }
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (on[j]) {
            s[0] += x[j];
            s[1] += y[j];
            s[2] += z[j];
            s[3] += x[j] * x[j];
            s[4] += y[j] * y[j];
            s[5] += z[j] * z[j];
        }
    }
    block_pairwise6(s, ws);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) P[int64_t(blockIdx.x) * 6 + q] = s[q];
}

// ---------------------------------------------------------------------------
// Cell keys in brick order: the cells are grouped in bricks of 4 cells per
// axis (an axis whose length 4 does not divide keeps bricks of 1 there).
// The bricks follow a serpentine path, x fastest: rows of bricks run
// backwards in x on odd rows (counting rows r = bz*nby + by'), and the rows
// of a layer run backwards in y on odd layers bz (by' = nby-1-by), so
// consecutive bricks are always face neighbours; the cells of one brick are
// in (z, y, x) order:
//   k = (r*nbx + bx') * |brick| + (lz*By + ly)*Bx + lx,
// a bijection of the cells onto [0, G).  Active particles with lower corner
// in [0, g-2]^3 (the particles the deposit and the kick touch) get their
// cell's key, the rest G.  A window of consecutive particles of this order
// spans a compact box of cells: the deposit's and the kick's node tiles, and
// the reorder's reads, stay local.
// ---------------------------------------------------------------------------
constexpr int kBrickLog2 = 2;  // 4 cells per axis

// log2 of the brick length along an axis of g cells
__device__ __forceinline__ int brick_log2(int g)
{
    return (g & ((1 << kBrickLog2) - 1)) == 0 ? kBrickLog2 : 0;
}

__device__ __forceinline__ int32_t brick_key(int ix, int iy, int iz, int gx, int gy, int gz)
{
    const int sx = brick_log2(gx), sy = brick_log2(gy), sz = brick_log2(gz);
    const int nbx = gx >> sx, nby = gy >> sy, bz = iz >> sz;
    int by = iy >> sy, bx = ix >> sx;
    if (bz & 1) by = nby - 1 - by;  // layers of odd bz: rows backwards
    const int row = bz * nby + by;
    if (row & 1) bx = nbx - 1 - bx;  // odd rows: bricks backwards
    const int brick = row * nbx + bx;
    const int cell = (((iz & ((1 << sz) - 1)) << sy) + (iy & ((1 << sy) - 1))) << sx |
                     (ix & ((1 << sx) - 1));
    return (brick << (sx + sy + sz)) + cell;
}

__device__ __forceinline__ int32_t cell_key(const double* __restrict__ parts, int64_t stride,
                                            const uint8_t* __restrict__ masks, int64_t i, int gx,
                                            int gy, int gz, const DomainRec& d)
{
    double o;
    const int ix = cell_index(parts[i], d.left[0], d.inv_h[0], o);
    const int iy = cell_index(parts[i + 2 * stride], d.left[1], d.inv_h[1], o);
    const int iz = cell_index(parts[i + 4 * stride], d.left[2], d.inv_h[2], o);
    const bool in = masks[i] && ix >= 0 && ix < gx - 1 && iy >= 0 && iy < gy - 1 && iz >= 0 &&
                    iz < gz - 1;
    return in ? brick_key(ix, iy, iz, gx, gy, gz) : gx * gy * gz;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, index): digits of `width` <= 8 bits, tiles
// of 4096 items (256 lanes x 16 rounds).  In the scatter, equal digits inside
// a wavefront are matched with `width` ballots, so ranks need one LDS update
// per distinct digit per wavefront.
// ---------------------------------------------------------------------------
constexpr int kSortItems = 16;
constexpr int kSortTile = kThreads * kSortItems;
constexpr int kMaxDigitBits = 8;
constexpr int kMaxBuckets = 1 << kMaxDigitBits;

// lanes of this wavefront holding the same digit (valid lanes only)
__device__ __forceinline__ uint64_t match_digit(int d, bool valid, int width)
{
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < width; ++b) {
        const uint64_t m = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? m : ~m;
    }
    return peers;
}

#ifndef SC3D_XCD_PART
#define SC3D_XCD_PART 1
#endif

// Block index of this workgroup in an XCD-contiguous order.  Workgroups of
// a launch go to the 8 XCDs round-robin (workgroup b to XCD b % 8) and each
// XCD has its own L2; renumbered so that XCD x runs the x-th eighth of the
// blocks, neighbouring tiles (whose partial cache lines of the sort output,
// the digit counts and the reorder buffers meet) share one L2.  Used by the
// sort and reorder kernels (measured faster); the deposit and the kick keep
// the dispatch order (measured slower with it).  Grids are rounded up to
// multiples of 8 by the launchers; any other grid keeps the identity.
__device__ __forceinline__ unsigned part_block()
{
#if SC3D_XCD_PART
    const unsigned nb = gridDim.x, b = blockIdx.x;
    if ((nb & 7u) == 0u) return (b & 7u) * (nb >> 3) + (b >> 3);
#endif
    return blockIdx.x;
}

// The per-tile digit counts: entry (digit d, tile b), digit-major (the scan
// reads a digit's row contiguously; measured faster than tile-major rows,
// whose column reads in the scan cost more than the count pass's scattered
// stores save: sort phase 175 vs 183 us at 4e6 particles, 654 vs 683 us at
// 1.6e7).
__device__ __forceinline__ int64_t count_at(int d, int64_t b, int nblk)
{
    return int64_t(d) * nblk + b;
}

// Lane moves inside the register file (DPP, GFX9 controls) for the
// wavefront scans, instead of LDS permutes: rows of 16 lanes; a lane with no
// source lane, or in a row outside ROW_MASK, returns `old`.
constexpr int kDppRowShr = 0x110;  // + n: from lane - n of the same row
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;  // from lane + 1 / lane - 1
constexpr int kDppRowBcast15 = 0x142;  // lane 15 of each row to the next row
constexpr int kDppRowBcast31 = 0x143;  // lane 31 to rows 2 and 3
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_move(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_move(float old, float v)
{
    return __int_as_float(dpp_move<CTRL, ROW_MASK>(__float_as_int(old), __float_as_int(v)));
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_move(double old, double v)
{
    const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    const unsigned lo = unsigned(dpp_move<CTRL, ROW_MASK>(int(o), int(x)));
    const int hi = dpp_move<CTRL, ROW_MASK>(int(o >> 32), int(x >> 32));
    return __longlong_as_double((long long)(uint64_t(lo) | (uint64_t(unsigned(hi)) << 32)));
}

// The xor-butterfly sum of one double per lane (s[t] += s[t ^ w], w = 1, 2,
// .., 32) as lane 0 forms it, in every lane.  After the step of w every
// lane of a group of 2w holds the same value, so any lane of the partner
// group serves: w = 1, 2 by quad permutes, 4 by the half-row mirror, 8 by
// the row mirror (DPP moves), and the rows' totals a, b, c, d (lanes 0, 16,
// 32, 48) combine as lane 0 combines them, (a + b) + (c + d).  Every lane of
// the wavefront calls it.  (Against six ds_bpermute exchanges per value:
// domain phase 25.0 -> 24.5 us at 4e6, 103.3 -> 102.2 us at 1.6e7.)
constexpr int kDppQuadXor1 = 0xB1;       // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadXor2 = 0x4E;       // quad_perm [2, 3, 0, 1]
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const long long x = __double_as_longlong(v);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(x), lane));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(x >> 32), lane));
    return __longlong_as_double((long long)(uint64_t(lo) | (uint64_t(hi) << 32)));
}
__device__ __forceinline__ double wave_pairwise_sum(double v)
{
    v = v + dpp_move<kDppQuadXor1>(v, v);
    v = v + dpp_move<kDppQuadXor2>(v, v);
    v = v + dpp_move<kDppRowHalfMirror>(v, v);
    v = v + dpp_move<kDppRowMirror>(v, v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Exclusive prefix of one value per thread over the block (DPP lane moves
// inside each wavefront, one LDS exchange of the wavefront sums); every
// thread of the block calls it.  `wsum` holds kThreads/64 entries; callers
// separate two uses by a barrier.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* wsum, int32_t& total)
{
    constexpr int kWaves = kThreads / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // inclusive wavefront scan: inside rows of 16 lanes, then across rows
    int32_t x = v;
    x += dpp_move<kDppRowShr + 1>(0, x);
    x += dpp_move<kDppRowShr + 2>(0, x);
    x += dpp_move<kDppRowShr + 4>(0, x);
    x += dpp_move<kDppRowShr + 8>(0, x);
    x += dpp_move<kDppRowBcast15, 0xa>(0, x);
    x += dpp_move<kDppRowBcast31, 0xc>(0, x);
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const int32_t ws = wsum[w];
        off += w < wave ? ws : 0;
        tot += ws;
    }
    total = tot;
    return off + x - v;
}

// Histogram of one digit per tile.  Every lane first loads (or, with KEYS,
// computes from the particles and stores) its 16 keys, so all loads of the
// tile are in flight together; each wavefront then counts its quarter of
// the tile into its own LDS histogram, one update per distinct digit and
// round (ballot match: after the first pass equal digits cluster), and the
// four are summed.  KEYS fuses the key pass into pass 0.
template <bool KEYS>
__global__ __launch_bounds__(kThreads) void k_radix_count(
    const double* __restrict__ parts, int64_t stride, const uint8_t* __restrict__ masks, int gx,
    int gy, int gz, int32_t* __restrict__ keys, int64_t n, int shift, int width,
    int32_t* __restrict__ counts, int nblk, const DomainRec* __restrict__ dom)
{
    constexpr int kWaves = kThreads / 64;
    constexpr int kWaveItems = kSortTile / kWaves;
    __shared__ int32_t hist[kWaves][kMaxBuckets];
    if (dom->error) return;
    const unsigned blk = part_block();
    if (blk >= unsigned(nblk)) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nb = 1 << width;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) hist[w][t] = 0;
    const int64_t base = int64_t(blk) * kSortTile + int64_t(wave) * kWaveItems;
    int32_t kk[kSortItems];
    if (KEYS) {
        const DomainRec d = *dom;
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const int64_t i = base + r * 64 + lane;
            kk[r] = i < n ? cell_key(parts, stride, masks, i, gx, gy, gz, d) : -1;
        }
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const int64_t i = base + r * 64 + lane;
            if (i < n) keys[i] = kk[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const int64_t i = base + r * 64 + lane;
            kk[r] = i < n ? keys[i] : -1;  // keys are >= 0
        }
    }
    __syncthreads();
    const uint64_t lt_mask = (1ull << lane) - 1;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const bool valid = kk[r] >= 0;
        const int dg = (kk[r] >> shift) & (nb - 1);
        const uint64_t peers = match_digit(dg, valid, width);
        if (valid && __popcll(peers & lt_mask) == 0) hist[wave][dg] += __popcll(peers);
    }
    __syncthreads();
    if (t < nb) {
        int32_t c = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) c += hist[w][t];
        counts[count_at(t, blk, nblk)] = c;
    }
}

// One block per digit: exclusive scan of that digit's per-tile counts (in
// place) and the digit total.
__global__ __launch_bounds__(kThreads) void k_radix_scan(int32_t* __restrict__ counts, int nblk,
                                                         int32_t* __restrict__ totals,
                                                         const DomainRec* dom)
{
    __shared__ int32_t wsum[kThreads / 64];
    if (dom->error) return;
    const int d = blockIdx.x;
    const int per = (nblk + kThreads - 1) / kThreads;
    const int b0 = threadIdx.x * per;
    int32_t s = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nblk) s += counts[count_at(d, b0 + k, nblk)];
    int32_t total;
    int32_t run = block_exclusive_scan(s, wsum, total);
    for (int k = 0; k < per; ++k)
        if (b0 + k < nblk) {
            const int64_t at = count_at(d, b0 + k, nblk);
            const int32_t v = counts[at];
            counts[at] = run;
            run += v;
        }
    if (threadIdx.x == 0) totals[d] = total;
}

// Stable scatter of one tile (4096 items).  Wavefront w ranks the items
// [1024w, 1024w + 1024) of the tile in 16 rounds of 64 with wave-private
// digit counters (ballot match: rank among equal digits of the round, the
// lowest lane of each digit group advances the counter), with no block
// barrier inside the rounds.  The tile's digit counts come from the count
// pass (difference of consecutive exclusive prefixes).  Every item is then
// placed in LDS at its position in the tile's stable digit order, and the
// tile is written out in that order, so each digit's items of the tile form
// one contiguous run in the output (positions as a per-item scatter would
// give them).
__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const int32_t* __restrict__ keys_in, const int32_t* __restrict__ vals_in, int64_t n,
    int shift, int width, const int32_t* __restrict__ counts, const int32_t* __restrict__ totals,
    int nblk, int32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out,
    const DomainRec* dom)
{
    constexpr int kWaves = kThreads / 64;
    constexpr int kWaveItems = kSortTile / kWaves;
    __shared__ int32_t base_d[kMaxBuckets];   // output position of the tile's first digit-d item
    __shared__ int32_t start_d[kMaxBuckets];  // its position inside the tile
    __shared__ int32_t wave_off[kWaves][kMaxBuckets];
    __shared__ int32_t wsum[2][kWaves];
    __shared__ int32_t lkey[kSortTile];
    __shared__ int32_t lval[kSortTile];
    if (dom->error) return;
    const unsigned blk = part_block();
    if (blk >= unsigned(nblk)) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nb = 1 << width;  // <= kMaxBuckets == kThreads: thread t owns digit t
    const int64_t tile = int64_t(blk) * kSortTile;
    const int nvalid = int(n - tile < kSortTile ? n - tile : kSortTile);
    // digit totals -> exclusive prefix over digits (global digit bases);
    // this tile's digit counts -> local digit starts
    const int32_t tot = t < nb ? totals[t] : 0;
    int32_t cnt = 0, before = 0;
    if (t < nb) {
        before = counts[count_at(t, blk, nblk)];
        const int32_t next = blk + 1 < unsigned(nblk) ? counts[count_at(t, blk + 1, nblk)] : tot;
        cnt = next - before;
    }
    int32_t all_n, tile_n;
    const int32_t digit_base = block_exclusive_scan(tot, wsum[0], all_n);
    const int32_t local_start = block_exclusive_scan(cnt, wsum[1], tile_n);
    if (t < nb) {
        base_d[t] = digit_base + before;
        start_d[t] = local_start;
    }
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wave_off[w][t] = 0;
    __syncthreads();
    // ranking inside each wavefront's quarter of the tile
    const uint64_t lt_mask = (1ull << lane) - 1;
    int32_t kk[kSortItems], vv[kSortItems], rk[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = tile + int64_t(wave) * kWaveItems + r * 64 + lane;
        const bool valid = i < n;
        const int32_t key = valid ? keys_in[i] : 0;
        const int32_t val = valid ? (vals_in ? vals_in[i] : int32_t(i)) : 0;
        const int dg = (key >> shift) & (nb - 1);
        const uint64_t peers = match_digit(dg, valid, width);
        const int rank = __popcll(peers & lt_mask);
        // every lane of a digit group reads the counter before its lowest
        // lane advances it (LDS operations of a wavefront complete in order)
        const int32_t seen = wave_off[wave][dg];
        if (valid && rank == 0) wave_off[wave][dg] = seen + __popcll(peers);
        kk[r] = valid ? key : -1;
        vv[r] = val;
        rk[r] = seen + rank;
    }
    __syncthreads();
    if (t < nb) {  // exclusive prefix over wavefronts, per digit
        int32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const int32_t c = wave_off[w][t];
            wave_off[w][t] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (kk[r] < 0) continue;  // keys are >= 0; -1 marks lanes past n
        const int dg = (kk[r] >> shift) & (nb - 1);
        const int lpos = start_d[dg] + wave_off[wave][dg] + rk[r];
        if (lpos >= 0 && lpos < kSortTile) {  // always true for consistent counts
            lkey[lpos] = kk[r];
            lval[lpos] = vv[r];
        }
    }
    __syncthreads();
    for (int j = t; j < nvalid; j += kThreads) {
        const int32_t key = lkey[j];
        const int dg = (key >> shift) & (nb - 1);
        const int64_t pos = int64_t(base_d[dg]) + (j - start_d[dg]);
        if (pos >= 0 && pos < n) {  // always true for consistent counts
            keys_out[pos] = key;
            vals_out[pos] = lval[j];
        }
    }
}

// Reorder: scratch[j] = bunch[perm[j]] (7 columns + mask); the apply's deposit reads
// scratch and its kick writes it back to the bunch storage.
__global__ __launch_bounds__(kThreads) void k_gather_particles(
    const double* __restrict__ parts, int64_t stride, const uint8_t* __restrict__ masks,
    const int32_t* __restrict__ perm, int64_t n, double* __restrict__ scratch,
    uint8_t* __restrict__ smask, const DomainRec* dom)
{
    if (dom->error) return;
    for (int64_t j = int64_t(part_block()) * kThreads + threadIdx.x; j < n;
         j += int64_t(gridDim.x) * kThreads) {
        int64_t src = perm[j];
        if (src < 0 || src >= n) src = j;  // perm is a permutation of [0, n)
#pragma unroll
        for (int c = 0; c < 7; ++c) scratch[j + c * n] = parts[src + c * stride];
        smask[j] = masks[src];
    }
}

template <typename R>
__global__ void k_fill(R* __restrict__ p, int64_t n, R v)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// CIC deposit (sv_zyx_rho_reducer_non_periodic, deposit.cc:309-395, launched
// :588-629): 8 corners with weights w0*aox*aoy*aoz (left to right), a corner
// written only if 0 < idx < g-1 in all three axes (ingrid, :21-30).
// Target: the G octant of rho2 (zyx, gx fastest).
//
// Through the sort permutation (k_deposit, used when the storage is never
// reordered): the eight corner weights are summed over runs of equal cell
// keys before any global atomic.  Every lane takes K consecutive particles
// of the cell-sorted order and sums its own runs in order; a run that
// starts and ends inside the lane is emitted at once.  The lanes' trailing runs are
// then summed across lanes by one segmented inclusive scan (a segment
// continues through lanes that hold a single run of the same key), and the
// total reaching a lane whose leading run has that key is carried into it.
// Every run is emitted once, by the lane where it ends.  Any particle order
// is correct; a cell-sorted order gives about one run per cell and
// wavefront.  Lanes past n and particles that are masked out or off the
// grid break runs and add nothing.
//
// Precision form (FP32 grid): the cell index and offsets come from the FP64
// coordinates as above; the weights are unit charge weights aox*aoy*aoz
// (left to right) rounded to FP32, and all run sums are compensated FP32
// (TwoSum: a sum and its rounding error, combined error-free), so a run's
// total is its exact sum rounded once.  The physical weight w0 and the
// unit-peak scale return in FP64 at the kick.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Eight corner sums: FP64 plain, FP32 compensated.
template <typename R> struct CornerSums;
template <> struct CornerSums<double> {
    double w[8];
    __device__ __forceinline__ void add(const CornerSums& o)
    {
#pragma unroll
        for (int c = 0; c < 8; ++c) w[c] += o.w[c];
    }
    template <int CTRL, int ROW_MASK = 0xf> __device__ __forceinline__ CornerSums moved() const
    {
        CornerSums o;
#pragma unroll
        for (int c = 0; c < 8; ++c) o.w[c] = dpp_move<CTRL, ROW_MASK>(0.0, w[c]);
        return o;
    }
    __device__ __forceinline__ double total(int c) const { return w[c]; }
};
template <> struct CornerSums<float> {
    float w[8], e[8];
    __device__ __forceinline__ void add(const CornerSums& o)
    {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            float t, err;
            two_sum(w[c], o.w[c], t, err);
            w[c] = t;
            e[c] = (e[c] + o.e[c]) + err;
        }
    }
    template <int CTRL, int ROW_MASK = 0xf> __device__ __forceinline__ CornerSums moved() const
    {
        CornerSums o;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            o.w[c] = dpp_move<CTRL, ROW_MASK>(0.0f, w[c]);
            o.e[c] = dpp_move<CTRL, ROW_MASK>(0.0f, e[c]);
        }
        return o;
    }
    __device__ __forceinline__ float total(int c) const { return w[c] + e[c]; }
};

// The corner weights of one particle (ok: on the grid; else zeros).
template <typename R>
__device__ __forceinline__ CornerSums<R> corner_weights(bool ok, double ox, double oy, double oz,
                                                        double w0)
{
    CornerSums<R> a;
    if constexpr (sizeof(R) == sizeof(double)) {
        const double aox = 1.0 - ox, aoy = 1.0 - oy, aoz = 1.0 - oz;
        a.w[0] = w0 * aox * aoy * aoz;
        a.w[1] = w0 * ox * aoy * aoz;
        a.w[2] = w0 * aox * oy * aoz;
        a.w[3] = w0 * ox * oy * aoz;
        a.w[4] = w0 * aox * aoy * oz;
        a.w[5] = w0 * ox * aoy * oz;
        a.w[6] = w0 * aox * oy * oz;
        a.w[7] = w0 * ox * oy * oz;
    } else {
        const float fx = float(ox), fy = float(oy), fz = float(oz);
        const float ax = float(1.0 - ox), ay = float(1.0 - oy), az = float(1.0 - oz);
        a.w[0] = ax * ay * az;
        a.w[1] = fx * ay * az;
        a.w[2] = ax * fy * az;
        a.w[3] = fx * fy * az;
        a.w[4] = ax * ay * fz;
        a.w[5] = fx * ay * fz;
        a.w[6] = ax * fy * fz;
        a.w[7] = fx * fy * fz;
#pragma unroll
        for (int c = 0; c < 8; ++c) a.e[c] = 0.0f;
    }
    if (!ok) {
#pragma unroll
        for (int c = 0; c < 8; ++c) a.w[c] = R(0);
    }
    return a;
}

// Adds one run's sums to the interior corners of cell `key`.
template <typename R>
__device__ __forceinline__ void add_run(R* __restrict__ rho, int key, const CornerSums<R>& a,
                                        int gx, int gy, int gz)
{
    const int ix = key % gx, r = key / gx;
    const int iy = r % gy, iz = r / gy;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int cx = ix + (c & 1), cy = iy + ((c >> 1) & 1), cz = iz + (c >> 2);
        if (cx > 0 && cx < gx - 1 && cy > 0 && cy < gy - 1 && cz > 0 && cz < gz - 1)
            unsafeAtomicAdd(rho + (int64_t(cz) * gy + cy) * gx + cx, R(a.total(c)));
    }
}

// The runs of one lane's K consecutive particles of a cell-sorted order and
// the wavefront's combination (see above).  key[q] < 0 marks particles that
// add nothing; those keys must differ from each other and from the keys of
// every other lane of the wavefront.
// Run totals go to `emit(key, sums, q_last)`, q_last = the lane index of the
// run's last particle in this lane (K-1 for a run that goes on into the next
// lane or wavefront).
template <typename R> struct AtomicRuns {
    R* __restrict__ rho;
    int gx, gy, gz;
    __device__ __forceinline__ void operator()(int key, const CornerSums<R>& a, int) const
    {
        add_run<R>(rho, key, a, gx, gy, gz);
    }
};

// One step of the segmented scan: add the partial sum the lane move brings
// when it comes from the same segment.
template <int CTRL, int ROW_MASK, typename R>
__device__ __forceinline__ void scan_step(CornerSums<R>& cur, int seg)
{
    const int seg_from = dpp_move<CTRL, ROW_MASK>(-1, seg);
    const CornerSums<R> up = cur.template moved<CTRL, ROW_MASK>();
    if (seg_from == seg) cur.add(up);
}

template <typename R, int K, typename Emit>
__device__ __forceinline__ void deposit_runs(const int (&key)[K], const double (&ox)[K],
                                             const double (&oy)[K], const double (&oz)[K],
                                             double w0, const Emit& emit)
{
    const int lane = threadIdx.x & 63;
    const uint64_t le_mask = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1);
    // runs inside the lane: the leading run (head), the trailing run (cur at
    // the end); runs in between are added at once
    CornerSums<R> cur, head;
    int cur_key = 0, head_key = 0, cur_start = 0, head_last = K - 1;
    bool single = true;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const CornerSums<R> p = corner_weights<R>(key[q] >= 0, ox[q], oy[q], oz[q], w0);
        if (q == 0) {
            cur = p;
            cur_key = key[q];
        } else if (key[q] == cur_key) {
            cur.add(p);
        } else {
            if (cur_start == 0) {
                head = cur;
                head_key = cur_key;
                head_last = q - 1;
                single = false;
            } else if (cur_key >= 0) {
                emit(cur_key, cur, q - 1);
            }
            cur = p;
            cur_key = key[q];
            cur_start = q;
        }
    }
    if (single) head_key = cur_key;
    // trailing runs across lanes: segmented inclusive scan (every lane takes
    // part in every lane move: a move must never run under a divergent exec
    // mask).  seg >= 1 numbers the segments; a lane adds a moved partial sum
    // only when it comes from its own segment (-1 marks no source lane).
    const int prev_tail = dpp_move<kDppWaveShr1>(0, cur_key);
    const bool start = lane == 0 || !single || prev_tail != cur_key;
    const int seg = __popcll(__ballot(start) & le_mask);
    // within rows of 16 lanes: from lane - 1, - 2, - 4, - 8 of the row
    scan_step<kDppRowShr + 1, 0xf>(cur, seg);
    scan_step<kDppRowShr + 2, 0xf>(cur, seg);
    scan_step<kDppRowShr + 4, 0xf>(cur, seg);
    scan_step<kDppRowShr + 8, 0xf>(cur, seg);
    // across rows: lane 15 into row 1 and lane 47 into row 3, then lane 31
    // (by now the sum through lane 31) into rows 2 and 3
    scan_step<kDppRowBcast15, 0xa>(cur, seg);
    scan_step<kDppRowBcast31, 0xc>(cur, seg);
    // the total of the previous lane's trailing run continues this lane's
    // leading run when the keys agree
    const CornerSums<R> carry = cur.template moved<kDppWaveShr1>();
    const int next_head = dpp_move<kDppWaveShl1>(0, head_key);
    if (!single && head_key >= 0) {
        if (lane > 0 && prev_tail == head_key) head.add(carry);
        emit(head_key, head, head_last);
    }
    if (cur_key >= 0 && (lane == 63 || next_head != cur_key)) emit(cur_key, cur, K - 1);
}

// Deposit through the sort permutation of this step (every apply sorts;
// used when the bunch storage is never reordered, sort_period 0).
template <typename R, int K>
__global__ __launch_bounds__(kThreads) void k_deposit(const double* __restrict__ parts,
                                                      int64_t stride,
                                                      const uint8_t* __restrict__ masks,
                                                      const int32_t* __restrict__ perm_in,
                                                      int64_t n, int gx, int gy, int gz,
                                                      const DomainRec* __restrict__ dom,
                                                      R* __restrict__ rho)
{
    if (dom->error) return;
    const int32_t* __restrict__ perm = perm_in;
    const double lx = dom->left[0], ly = dom->left[1], lz = dom->left[2];
    const double ihx = dom->inv_h[0], ihy = dom->inv_h[1], ihz = dom->inv_h[2];
    const int lane = threadIdx.x & 63;
    // this lane's K consecutive positions of the cell-sorted order
    const int64_t j0 = (blockIdx.x * int64_t(kThreads) + threadIdx.x) * K;
    int64_t idx[K];
    if (perm && j0 + K <= n && K % 4 == 0) {
#pragma unroll
        for (int v = 0; v < K / 4; ++v) {
            const int4 p4 = *reinterpret_cast<const int4*>(perm + j0 + 4 * v);
            idx[4 * v] = p4.x;
            idx[4 * v + 1] = p4.y;
            idx[4 * v + 2] = p4.z;
            idx[4 * v + 3] = p4.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int64_t j = j0 + q;
            idx[q] = j < n ? (perm ? int64_t(perm[j]) : j) : -1;
        }
    }
    double X[K], Y[K], Z[K];
    bool M[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int64_t j = j0 + q;
        if (j < n && (idx[q] < 0 || idx[q] >= n)) idx[q] = j;  // perm is a permutation of [0, n)
        M[q] = false;
        X[q] = Y[q] = Z[q] = 0.0;
        if (idx[q] >= 0) {
            M[q] = masks[idx[q]] != 0;
            X[q] = parts[idx[q]];
            Y[q] = parts[idx[q] + 2 * stride];
            Z[q] = parts[idx[q] + 4 * stride];
        }
    }
    int key[K];
    double ox[K], oy[K], oz[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        ox[q] = oy[q] = oz[q] = 0.0;
        key[q] = -1 - (lane * K + q);  // off-grid particles: singletons
        if (M[q]) {
            const int ix = cell_index(X[q], lx, ihx, ox[q]);
            const int iy = cell_index(Y[q], ly, ihy, oy[q]);
            const int iz = cell_index(Z[q], lz, ihz, oz[q]);
            if (ix >= 0 && ix <= gx - 2 && iy >= 0 && iy <= gy - 2 && iz >= 0 && iz <= gz - 2)
                key[q] = (iz * gy + iy) * gx + ix;
        }
    }
    deposit_runs<R, K>(key, ox, oy, oz, dom->w0, AtomicRuns<R>{rho, gx, gy, gz});
}

// Deposit in windows of the bunch storage through an LDS node tile.  The
// storage is cell-sorted at every reorder (sort_period applies) and
// particles move a few cells per step, so kTileWindow consecutive slots stay
// a compact group of cells.  A block takes one window, each thread four
// consecutive slots; the block reduces the window's cell bounding box, clears
// that box of nodes in LDS (up to kTileNodes FP64 nodes), and every particle
// adds its eight corner weights there with LDS FP64 atomics -- the corners of
// a thread's consecutive particles in one cell summed in registers first,
// the corner order permuted by lane so that the lanes of one cell hit eight
// different nodes at a time.  The block then adds the box's non-zero
// interior nodes to rho with global atomics (no sort, no scan, no run list).
// A window whose box does not fit gives each wavefront a quarter of the
// tile for the box of its own 256 particles (no block barrier: LDS
// operations of one wavefront complete in order); a wavefront whose box does
// not fit either sums its corners in a direct-mapped FP64 cache of nodes in
// its quarter (a crowded node of a scattered storage then takes one add per
// wavefront), the rest going to rho directly.  Those wide
// windows lie at the two ends of the cell order (the tails of the bunch), so
// the block order is rotated to start the last eighth of the windows
// first and keep them off the end of the kernel.  The tile accumulates in
// FP64 for both precisions (the precision form's unit weights are rounded
// to FP32 once, at the global add).
// Measured against the windowed kernel it replaces (LDS radix sort of each
// window, lane runs, DPP segmented scan, LDS run list; PMC: 465 VALU
// instructions per particle), deposit phase per step over sort periods of
// 12: 96.1 -> 54.9 us at 4e6 particles FP64 (64x64x128), 333.5 -> 214.7 us
// at 1.6e7 FP32 (128x128x256).  Variants measured on the way (us at 4e6,
// over a sort period): one particle per lane-stride instead of four
// consecutive per thread 81-99; tiles of 2048 / 3072 nodes 93-105 / 83-93;
// 512-thread blocks (6144 / 8192 nodes) 61-102 / 60-75; without the
// per-wavefront fallback and the block-order rotation 68-78; the corner sums
// picked by an indexed load (through scratch) 107-134.  In the library
// (deposit phase per step, 4e6 FP64 / 1.6e7 FP32): 4080 nodes (32 KiB with
// the box, 5 blocks per CU) 56.5 / 224.7 against 55.7 / 215.5; the same at 5
// waves per SIMD (96 VGPRs, 32 bytes spilled) 60.0 / 228.8; 3060 nodes at 6
// waves 93.3 / 351.9; the last eighth rotated first 55.8 / 215.7; 2 or 8
// slots per thread (windows of 512 / 2048) 61.6 / 242.4 and 72.6 / 287.0
// against 54.8 / 215.3.  With brick-ordered keys, sort period 24: the last
// eighth / sixteenth / thirty-second rotated first, or none, 54.6-54.8 /
// 55.4-55.5 / 54.3-54.5 / 56.1-56.3 at 4e6 and 200.6-200.7 / 201.3-202.4 /
// 201.8-202.3 / 203.5 at 1.6e7; 3 slots per thread 57.4-57.6 / 212.2-212.5.
constexpr int kTileThreads = 256;
constexpr int kTilePer = 4;                             // consecutive slots per thread
constexpr int kTileWindow = kTileThreads * kTilePer;    // slots per block
constexpr int kTileNodes = 4096;                        // FP64 nodes of the LDS tile (32 KiB)
constexpr int kTileTailDiv = 8;                         // 1/8 of the windows rotated first
constexpr int kTileWaves = kTileThreads / 64;

// Wavefront min / max, the same value in every lane: prefix reductions
// inside each row of 16 lanes by DPP moves (a lane with no source keeps its
// own value), then the four row results read from lanes 15, 31, 47, 63.
// Every lane of the wavefront calls it.  (Measured against six ds_bpermute
// exchanges per value, brick-ordered keys, sort period 24, deposit / kick
// phase: 54.3-54.7 / 64.4-64.7 -> 51.9-52.2 / 61.8-62.1 us at 4e6 FP64,
// 200.6-200.9 / 237.4-238.3 -> 190.2-190.8 / 233.5-233.8 us at 1.6e7 FP32.)
__device__ __forceinline__ int wave_min(int v)
{
    v = min(v, dpp_move<kDppRowShr + 1>(v, v));
    v = min(v, dpp_move<kDppRowShr + 2>(v, v));
    v = min(v, dpp_move<kDppRowShr + 4>(v, v));
    v = min(v, dpp_move<kDppRowShr + 8>(v, v));
    return min(min(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
               min(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
}
__device__ __forceinline__ int wave_max(int v)
{
    v = max(v, dpp_move<kDppRowShr + 1>(v, v));
    v = max(v, dpp_move<kDppRowShr + 2>(v, v));
    v = max(v, dpp_move<kDppRowShr + 4>(v, v));
    v = max(v, dpp_move<kDppRowShr + 8>(v, v));
    return max(max(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
               max(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
}

// Corner weights of one particle in the reference's order, w0*wx*wy*wz left
// to right (deposit.cc:366-391); corner c = (x bit 0, y bit 1, z bit 2).
__device__ __forceinline__ void tile_weights(double w0, double ox, double oy, double oz,
                                             double (&w)[8])
{
    const double aox = 1.0 - ox, aoy = 1.0 - oy, aoz = 1.0 - oz;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const double wx = (c & 1) ? ox : aox, wy = (c & 2) ? oy : aoy, wz = (c & 4) ? oz : aoz;
        w[c] = w0 * wx * wy * wz;
    }
}

// Adds the eight corner sums of the cell at tile offset b0 (box of nx x ny
// x .. nodes): at step k corner k ^ (lane mod 8), so the lanes of one cell
// hit eight different nodes at a time.  The sums are permuted by three
// conditional register swaps (an indexed pick would go through scratch).
__device__ __forceinline__ void tile_add(double* tile, int b0, int nx, int ny, int lane,
                                         const double (&w)[8])
{
    double v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = w[c];
    const int r = lane & 7;
#pragma unroll
    for (int bit = 1; bit < 8; bit <<= 1) {
        const bool sw = r & bit;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (!(c & bit)) {
                const double a = v[c], b = v[c | bit];
                v[c] = sw ? b : a;
                v[c | bit] = sw ? a : b;
            }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int c = k ^ r;  // v[k] holds corner c
        const int off = ((c >> 2) * ny + ((c >> 1) & 1)) * nx + (c & 1);
        unsafeAtomicAdd(&tile[b0 + off], v[k]);
    }
}

// This thread's particles into a tile whose box starts at (bx, by, bz):
// consecutive particles of one cell summed first.
__device__ __forceinline__ void tile_deposit(double* tile, int bx, int by, int bz, int nx, int ny,
                                             int lane, double w0, const int (&ix)[kTilePer],
                                             const int (&iy)[kTilePer], const int (&iz)[kTilePer],
                                             const bool (&ok)[kTilePer],
                                             const double (&ox)[kTilePer],
                                             const double (&oy)[kTilePer],
                                             const double (&oz)[kTilePer])
{
    double s[8];
    int cur = -1;
#pragma unroll
    for (int q = 0; q < kTilePer; ++q) {
        if (!ok[q]) continue;
        const int b0 = ((iz[q] - bz) * ny + (iy[q] - by)) * nx + (ix[q] - bx);
        double w[8];
        tile_weights(w0, ox[q], oy[q], oz[q], w);
        if (b0 == cur) {
#pragma unroll
            for (int c = 0; c < 8; ++c) s[c] += w[c];
        } else {
            if (cur >= 0) tile_add(tile, cur, nx, ny, lane, s);
#pragma unroll
            for (int c = 0; c < 8; ++c) s[c] = w[c];
            cur = b0;
        }
    }
    if (cur >= 0) tile_add(tile, cur, nx, ny, lane, s);
}

// The non-zero interior nodes of a tile (box at (bx, by, bz), nx x ny x ..,
// `nodes` in all) into rho, node e of the thread's share first..step.
template <typename R>
__device__ __forceinline__ void tile_flush(const double* tile, int nodes, int first, int step,
                                           int bx, int by, int bz, int nx, int ny, int gx, int gy,
                                           int gz, R* __restrict__ rho)
{
    for (int e = first; e < nodes; e += step) {
        const double v = tile[e];
        if (v == 0.0) continue;
        const int cx = bx + e % nx, r = e / nx;
        const int cy = by + r % ny, cz = bz + r / ny;
        if (cx > 0 && cx < gx - 1 && cy > 0 && cy < gy - 1 && cz > 0 && cz < gz - 1)
            unsafeAtomicAdd(rho + (int64_t(cz) * gy + cy) * gx + cx, R(v));
    }
}

template <typename R>
__global__ __launch_bounds__(kTileThreads) void k_deposit_tile(
    const double* __restrict__ parts, int64_t stride, const uint8_t* __restrict__ masks,
    int64_t n, int gx, int gy, int gz, const DomainRec* __restrict__ dom, R* __restrict__ rho)
{
    __shared__ double tile[kTileNodes];
    __shared__ int box[kTileWaves][6];
    if (dom->error) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const double lx = dom->left[0], ly = dom->left[1], lz = dom->left[2];
    const double ihx = dom->inv_h[0], ihy = dom->inv_h[1], ihz = dom->inv_h[2];
    // precision form: unit weights (w0 and the unit-peak scale return at the kick)
    const double w0 = sizeof(R) == sizeof(double) ? dom->w0 : 1.0;
    // the last eighth of the windows (one tail of the bunch) first
    const int nb = int(gridDim.x), tail = nb / kTileTailDiv;
    const int blk = (int(blockIdx.x) + nb - tail) % nb;
    const int64_t base = int64_t(blk) * kTileWindow + int64_t(t) * kTilePer;
    int ix[kTilePer], iy[kTilePer], iz[kTilePer];
    double ox[kTilePer], oy[kTilePer], oz[kTilePer];
    bool ok[kTilePer];
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {-1, -1, -1};
#pragma unroll
    for (int q = 0; q < kTilePer; ++q) {
        const int64_t j = base + q;
        double x = 0.0, y = 0.0, z = 0.0;
        bool m = false;
        if (j < n) {
            m = masks[j] != 0;
            x = parts[j];
            y = parts[j + 2 * stride];
            z = parts[j + 4 * stride];
        }
        ix[q] = cell_index(x, lx, ihx, ox[q]);
        iy[q] = cell_index(y, ly, ihy, oy[q]);
        iz[q] = cell_index(z, lz, ihz, oz[q]);
        // a particle outside [0, g-2] in any axis has no interior corner
        ok[q] = m && ix[q] >= 0 && ix[q] <= gx - 2 && iy[q] >= 0 && iy[q] <= gy - 2 &&
                iz[q] >= 0 && iz[q] <= gz - 2;
        if (ok[q]) {
            lo[0] = min(lo[0], ix[q]); hi[0] = max(hi[0], ix[q]);
            lo[1] = min(lo[1], iy[q]); hi[1] = max(hi[1], iy[q]);
            lo[2] = min(lo[2], iz[q]); hi[2] = max(hi[2], iz[q]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            box[wave][a] = lo[a];
            box[wave][3 + a] = hi[a];
        }
    __syncthreads();
    int blo[3], bhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        blo[a] = box[0][a];
        bhi[a] = box[0][3 + a];
#pragma unroll
        for (int w = 1; w < kTileWaves; ++w) {
            blo[a] = min(blo[a], box[w][a]);
            bhi[a] = max(bhi[a], box[w][3 + a]);
        }
    }
    if (bhi[0] < 0) return;  // nothing of the window on the grid (block-uniform)
    // the box of cells lo..hi has (hi - lo + 2) nodes per axis
    const int nx = bhi[0] - blo[0] + 2, ny = bhi[1] - blo[1] + 2, nz = bhi[2] - blo[2] + 2;
    const int nodes = nx * ny * nz;
    if (nodes <= kTileNodes) {  // block-uniform
        for (int e = t; e < nodes; e += kTileThreads) tile[e] = 0.0;
        __syncthreads();
        tile_deposit(tile, blo[0], blo[1], blo[2], nx, ny, lane, w0, ix, iy, iz, ok, ox, oy, oz);
        __syncthreads();
        tile_flush<R>(tile, nodes, t, kTileThreads, blo[0], blo[1], blo[2], nx, ny, gx, gy, gz,
                      rho);
        return;
    }
    // the window's box does not fit: this wavefront's own box in its quarter
    constexpr int kWaveNodes = kTileNodes / kTileWaves;
    double* wtile = tile + wave * kWaveNodes;
    const int wnx = hi[0] - lo[0] + 2, wny = hi[1] - lo[1] + 2, wnz = hi[2] - lo[2] + 2;
    const int wn = wnx * wny * wnz;
    if (hi[0] >= 0 && wn <= kWaveNodes) {  // wavefront-uniform; no block barrier below
        for (int e = lane; e < wn; e += 64) wtile[e] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        tile_deposit(wtile, lo[0], lo[1], lo[2], wnx, wny, lane, w0, ix, iy, iz, ok, ox, oy, oz);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        tile_flush<R>(wtile, wn, lane, 64, lo[0], lo[1], lo[2], wnx, wny, gx, gy, gz, rho);
        return;
    }
    // scattered particles (the sparse tails, or a storage out of cell order):
    // the quarter becomes a direct-mapped cache of kCacheSlots nodes (key =
    // node index, claimed by compare-and-swap) summed in FP64; a corner whose
    // slot holds another node goes to rho directly
    constexpr int kCacheSlots = kWaveNodes / 2;
    double* cval = wtile;
    int* ckey = reinterpret_cast<int*>(wtile + kCacheSlots);
    for (int e = lane; e < kCacheSlots; e += 64) {
        cval[e] = 0.0;
        ckey[e] = -1;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < kTilePer; ++q) {
        if (!ok[q]) continue;
        double w[8];
        tile_weights(w0, ox[q], oy[q], oz[q], w);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int cx = ix[q] + (c & 1), cy = iy[q] + ((c >> 1) & 1), cz = iz[q] + (c >> 2);
            if (!(cx > 0 && cx < gx - 1 && cy > 0 && cy < gy - 1 && cz > 0 && cz < gz - 1))
                continue;
            const int node = (cz * gy + cy) * gx + cx;
            const int slot = int((uint32_t(node) * 2654435761u) >> 23) & (kCacheSlots - 1);
            const int held = atomicCAS(&ckey[slot], -1, node);
            if (held == -1 || held == node)
                unsafeAtomicAdd(&cval[slot], w[c]);
            else
                unsafeAtomicAdd(rho + node, R(w[c]));
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int e = lane; e < kCacheSlots; e += 64) {
        const int node = ckey[e];
        if (node >= 0) unsafeAtomicAdd(rho + node, R(cval[e]));
    }
}

// Peak of the unit-weight FP32 density (all entries >= 0, so the IEEE bit
// patterns order like the values): block max (wavefront butterflies, then
// LDS), one integer atomicMax per block into the domain record (cleared by
// k_domain_finish each apply); the grid is bounded so few atomics meet
// (measured at 128x128x256: 8.9 us per launch over 512 blocks, 14.8 us with
// 16-byte loads over 1024 blocks).
__global__ __launch_bounds__(kThreads) void k_rho_peak(const float* __restrict__ rho, int64_t G,
                                                       DomainRec* __restrict__ dom)
{
    __shared__ float wmax[kThreads / 64];
    if (dom->error) return;
    float m = 0.0f;
    // eight independent loads in flight per lane and round
    const int64_t T = int64_t(gridDim.x) * kThreads;
    int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x;
    for (; i + 7 * T < G; i += 8 * T) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = rho[i + u * T];
#pragma unroll
        for (int u = 0; u < 8; ++u) m = fmaxf(m, v[u]);
    }
    for (; i < G; i += T) m = fmaxf(m, rho[i]);
    m = __int_as_float(wave_max(__float_as_int(m)));  // m >= 0: the bits order like the values
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = wmax[0];
        for (int w = 1; w < kThreads / 64; ++w) b = fmaxf(b, wmax[w]);
        if (b > 0.0f) atomicMax(&dom->rho_peak, __float_as_int(b));
    }
}

// ---------------------------------------------------------------------------
// E field (zyx::alg_force_extractor_doubled_domain, space_charge_3d_kernels.h:
// 217-307): E = -(phi[+1] - phi[-1]) * (0.5/h) on interior nodes, 0 on the
// boundary.  Stored interleaved (Ex, Ey, Ez, 0) per node for the kick's
// gathers.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kThreads) void k_efield(const R* __restrict__ phi, int gx, int gy,
                                                     int gz, const DomainRec* __restrict__ dom,
                                                     typename vec4<R>::type* __restrict__ E)
{
    if (dom->error) return;
    const double ihx = dom->half_inv_h[0], ihy = dom->half_inv_h[1], ihz = dom->half_inv_h[2];
    const int64_t G = int64_t(gx) * gy * gz;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < G;
         i += int64_t(gridDim.x) * kThreads) {
        int ix = int(i % gx);
        int64_t r = i / gx;
        int iy = int(r % gy), iz = int(r / gy);
        typename vec4<R>::type e;
        e.x = R(0); e.y = R(0); e.z = R(0); e.w = R(0);
        if (!(ix == 0 || ix == gx - 1 || iy == 0 || iy == gy - 1 || iz == 0 || iz == gz - 1)) {
            const int64_t sy = gx, sz = int64_t(gx) * gy;
            e.x = R(-(double(phi[i + 1]) - double(phi[i - 1])) * ihx);
            e.y = R(-(double(phi[i + sy]) - double(phi[i - sy])) * ihy);
            e.z = R(-(double(phi[i + sz]) - double(phi[i - sz])) * ihz);
        }
        E[i] = e;
    }
}

// ---------------------------------------------------------------------------
// Kick (zyx::alg_kicker, space_charge_3d_kernels.h:91-215): kicked iff
// 0 <= idx < g-1 in all axes; trilinear Ez -> energy kick on dp/p, then Ey
// -> y', then Ex -> x', each interpolation summed in the reference's corner
// order.  Momenta accumulate in FP64 for every precision.
// MAPS: the map program that follows the kick in the split-operator
// sequence is applied to every active particle in the same pass (one read
// and one write of the six coordinates for kick + maps).
// COPY (the apply that reordered the bunch): the particles are read from the
// cell-sorted copy the reorder gathered (src: 7 columns of stride n, and its
// masks), and every particle's 7 columns and mask are written to the bunch
// storage -- the reorder's copy back, folded into the kick.
// ---------------------------------------------------------------------------
// Kick factor: the FP64 form's is the record's; the precision form's grid
// holds unit-weight charge scaled to unit peak, so w0 * peak returns here.
template <typename R>
__device__ __forceinline__ double kick_factor(const DomainRec& d)
{
    if (sizeof(R) == sizeof(double)) return d.factor;
    const float peak = __int_as_float(d.rho_peak);
    return d.factor * d.w0 * (peak > 0.0f ? double(peak) : 1.0);
}

template <typename R, bool MAPS, bool COPY>
__global__ __launch_bounds__(kThreads) void k_kick(double* __restrict__ parts, int64_t stride,
                                                   uint8_t* __restrict__ masks, int64_t n,
                                                   int gx, int gy, int gz,
                                                   const DomainRec* __restrict__ dom,
                                                   const typename vec4<R>::type* __restrict__ E,
                                                   double pref, double m, MapProgram prog,
                                                   Sched sc, const double* __restrict__ src,
                                                   const uint8_t* __restrict__ smasks)
{
    if (dom->error) return;
    // served form: the program fused after the kick comes from the schedule
    const MapProgram* __restrict__ sp = (MAPS && sc.progs) ? sc.progs + sc.ctl->after_prog
                                                            : nullptr;
    const double lx = dom->left[0], ly = dom->left[1], lz = dom->left[2];
    const double ihx = dom->inv_h[0], ihy = dom->inv_h[1], ihz = dom->inv_h[2];
    const double factor = kick_factor<R>(*dom);
    const double* __restrict__ rd = COPY ? src : parts;
    const int64_t rs = COPY ? n : stride;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads) {
        const uint8_t on = COPY ? smasks[i] : masks[i];
        if (!COPY && !on) continue;
        // all six coordinates are loaded up front (independent loads in flight)
        double x = rd[i], y = rd[i + 2 * rs], cdt = rd[i + 4 * rs];
        double xp = rd[i + rs], yp = rd[i + 3 * rs], dpop = rd[i + 5 * rs];
        double ox, oy, oz;
        const int ix = cell_index(x, lx, ihx, ox);
        const int iy = cell_index(y, ly, ihy, oy);
        const int iz = cell_index(cdt, lz, ihz, oz);
        const bool in = on && (ix >= 0 && ix < gx - 1) && (iy >= 0 && iy < gy - 1) &&
                        (iz >= 0 && iz < gz - 1);
        if (COPY) {
            masks[i] = on;
            parts[i + 6 * stride] = rd[i + 6 * rs];
        }
        if (!MAPS && !COPY && !in) continue;
        if (in) {
            const double aox = 1.0 - ox, aoy = 1.0 - oy, aoz = 1.0 - oz;
            const int64_t b0 = (int64_t(iz) * gy + iy) * gx + ix;
            const int64_t b1 = b0 + int64_t(gx) * gy;
            typename vec4<R>::type e[8] = {E[b0],  E[b0 + 1],  E[b0 + gx],  E[b0 + gx + 1],
                                           E[b1],  E[b1 + 1],  E[b1 + gx],  E[b1 + gx + 1]};
            const double w[8] = {aoz * aoy * aox, aoz * aoy * ox, aoz * oy * aox, aoz * oy * ox,
                                 oz * aoy * aox,  oz * aoy * ox,  oz * oy * aox,  oz * oy * ox};
            double vz = w[0] * double(e[0].z), vy = w[0] * double(e[0].y),
                   vx = w[0] * double(e[0].x);
#pragma unroll
            for (int c = 1; c < 8; ++c) {
                vz += w[c] * double(e[c].z);
                vy += w[c] * double(e[c].y);
                vx += w[c] * double(e[c].x);
            }
            const double p = pref + dpop * pref;
            const double Eoc_i = sqrt(p * p + m * m);
            const double Eoc_f = Eoc_i + factor * (-pref) * vz;
            dpop += (sqrt(Eoc_f * Eoc_f - m * m) - sqrt(Eoc_i * Eoc_i - m * m)) / pref;
            yp += factor * vy;
            xp += factor * vx;
        }
        if (MAPS && on) {
            if (sp) {
                for (int k = 0; k < sp->n; ++k)
                    apply_slice(sp->s[k], x, xp, y, yp, cdt, dpop, sp->ref_p, sp->mass);
            } else {
                for (int k = 0; k < prog.n; ++k)
                    apply_slice(prog.s[k], x, xp, y, yp, cdt, dpop, prog.ref_p, prog.mass);
            }
        }
        if (MAPS || COPY) {
            parts[i] = x;
            parts[i + 2 * stride] = y;
            parts[i + 4 * stride] = cdt;
        }
        parts[i + stride] = xp;
        parts[i + 3 * stride] = yp;
        parts[i + 5 * stride] = dpop;
    }
}

// The kick without maps over windows of the bunch storage (kKickPer
// particles per lane, lane-interleaved so every load is coalesced) with the
// field of each window's box of nodes staged in LDS: the storage is
// cell-sorted at every reorder, so a window's particles share few nodes, and
// the eight-corner gathers read LDS instead of going through the vector
// cache (a timing probe without the gathers ran the kick in 47 of its 72 us
// at 4e6 particles on 64x64x128).  A window whose box exceeds the tile (the
// sparse tails, a storage out of cell order) gathers from global memory as
// k_kick does.  Same arithmetic in the same order as k_kick: the tile holds
// copies of the field values.  Measured (kick us, 4e6 FP64 on 64x64x128 /
// 1.6e7 FP32 on 128x128x256, one run): k_kick 75.7 / 242; 2 particles per
// lane and 1024 nodes 69.4 / 224; 4 and 2048 nodes 70.5 / 240; 1 and 512
// nodes 73.9 / 234.  With the maps fused (k_kick<R, true>) the extra
// registers of the tile version cost more than the gathers save: the fused
// kick stays on k_kick.
constexpr int kKickPer = 2;
constexpr int kKickNodes = 1024;

template <typename R, bool COPY>
__global__ __launch_bounds__(kThreads) void k_kick_tile(
    double* __restrict__ parts, int64_t stride, uint8_t* __restrict__ masks, int64_t n, int gx,
    int gy, int gz, const DomainRec* __restrict__ dom, const typename vec4<R>::type* __restrict__ E,
    double pref, double m, const double* __restrict__ src, const uint8_t* __restrict__ smasks)
{
    __shared__ R tile[3][kKickNodes];
    __shared__ int box[kThreads / 64][6];
    if (dom->error) return;
    const double lx = dom->left[0], ly = dom->left[1], lz = dom->left[2];
    const double ihx = dom->inv_h[0], ihy = dom->inv_h[1], ihz = dom->inv_h[2];
    const double factor = kick_factor<R>(*dom);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t base = int64_t(blockIdx.x) * (kThreads * kKickPer) + t;
    const double* __restrict__ rd = COPY ? src : parts;
    const int64_t rs = COPY ? n : stride;
    double x[kKickPer], y[kKickPer], cdt[kKickPer], xp[kKickPer], yp[kKickPer], dpop[kKickPer];
    double ox[kKickPer], oy[kKickPer], oz[kKickPer];
    int ix[kKickPer], iy[kKickPer], iz[kKickPer];
    bool act[kKickPer], in[kKickPer];
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {-1, -1, -1};
#pragma unroll
    for (int q = 0; q < kKickPer; ++q) {
        const int64_t i = base + int64_t(q) * kThreads;
        act[q] = false;
        x[q] = y[q] = cdt[q] = xp[q] = yp[q] = dpop[q] = 0.0;
        if (i < n) {
            const uint8_t on = COPY ? smasks[i] : masks[i];
            act[q] = on != 0;
            x[q] = rd[i];
            y[q] = rd[i + 2 * rs];
            cdt[q] = rd[i + 4 * rs];
            xp[q] = rd[i + rs];
            yp[q] = rd[i + 3 * rs];
            dpop[q] = rd[i + 5 * rs];
            if (COPY) {
                masks[i] = on;
                parts[i] = x[q];
                parts[i + 2 * stride] = y[q];
                parts[i + 4 * stride] = cdt[q];
                parts[i + 6 * stride] = rd[i + 6 * rs];
            }
        }
        ix[q] = cell_index(x[q], lx, ihx, ox[q]);
        iy[q] = cell_index(y[q], ly, ihy, oy[q]);
        iz[q] = cell_index(cdt[q], lz, ihz, oz[q]);
        in[q] = act[q] && (ix[q] >= 0 && ix[q] < gx - 1) && (iy[q] >= 0 && iy[q] < gy - 1) &&
                (iz[q] >= 0 && iz[q] < gz - 1);
        if (in[q]) {
            lo[0] = min(lo[0], ix[q]); hi[0] = max(hi[0], ix[q]);
            lo[1] = min(lo[1], iy[q]); hi[1] = max(hi[1], iy[q]);
            lo[2] = min(lo[2], iz[q]); hi[2] = max(hi[2], iz[q]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            box[wave][a] = lo[a];
            box[wave][3 + a] = hi[a];
        }
    __syncthreads();
    int blo[3], bhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        blo[a] = box[0][a];
        bhi[a] = box[0][3 + a];
#pragma unroll
        for (int w = 1; w < kThreads / 64; ++w) {
            blo[a] = min(blo[a], box[w][a]);
            bhi[a] = max(bhi[a], box[w][3 + a]);
        }
    }
    // the box of cells lo..hi has (hi - lo + 2) nodes per axis
    const int nx = bhi[0] - blo[0] + 2, ny = bhi[1] - blo[1] + 2, nz = bhi[2] - blo[2] + 2;
    const bool staged = bhi[0] >= 0 && nx * ny * nz <= kKickNodes;  // block-uniform
    if (staged) {
        const int nodes = nx * ny * nz;
        for (int e = t; e < nodes; e += kThreads) {
            const int cx = e % nx, r = e / nx;
            const int cy = r % ny, cz = r / ny;
            const auto v = E[(int64_t(blo[2] + cz) * gy + blo[1] + cy) * gx + blo[0] + cx];
            tile[0][e] = v.x;
            tile[1][e] = v.y;
            tile[2][e] = v.z;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kKickPer; ++q) {
        if (COPY ? base + int64_t(q) * kThreads >= n : !in[q]) continue;
        if (in[q]) {
            const double aox = 1.0 - ox[q], aoy = 1.0 - oy[q], aoz = 1.0 - oz[q];
            R ex[8], ey[8], ez[8];
            if (staged) {
                const int b0 = ((iz[q] - blo[2]) * ny + iy[q] - blo[1]) * nx + ix[q] - blo[0];
                const int b1 = b0 + nx * ny;
                const int c[8] = {b0, b0 + 1, b0 + nx, b0 + nx + 1,
                                  b1, b1 + 1, b1 + nx, b1 + nx + 1};
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    ex[k] = tile[0][c[k]];
                    ey[k] = tile[1][c[k]];
                    ez[k] = tile[2][c[k]];
                }
            } else {
                const int64_t b0 = (int64_t(iz[q]) * gy + iy[q]) * gx + ix[q];
                const int64_t b1 = b0 + int64_t(gx) * gy;
                const int64_t c[8] = {b0, b0 + 1, b0 + gx, b0 + gx + 1,
                                      b1, b1 + 1, b1 + gx, b1 + gx + 1};
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const auto v = E[c[k]];
                    ex[k] = v.x;
                    ey[k] = v.y;
                    ez[k] = v.z;
                }
            }
            const double w[8] = {aoz * aoy * aox, aoz * aoy * ox[q], aoz * oy[q] * aox,
                                 aoz * oy[q] * ox[q], oz[q] * aoy * aox, oz[q] * aoy * ox[q],
                                 oz[q] * oy[q] * aox, oz[q] * oy[q] * ox[q]};
            double vz = w[0] * double(ez[0]), vy = w[0] * double(ey[0]), vx = w[0] * double(ex[0]);
#pragma unroll
            for (int k = 1; k < 8; ++k) {
                vz += w[k] * double(ez[k]);
                vy += w[k] * double(ey[k]);
                vx += w[k] * double(ex[k]);
            }
            const double p = pref + dpop[q] * pref;
            const double Eoc_i = sqrt(p * p + m * m);
            const double Eoc_f = Eoc_i + factor * (-pref) * vz;
            dpop[q] += (sqrt(Eoc_f * Eoc_f - m * m) - sqrt(Eoc_i * Eoc_i - m * m)) / pref;
            yp[q] += factor * vy;
            xp[q] += factor * vx;
        }
        const int64_t i = base + int64_t(q) * kThreads;
        parts[i + stride] = xp[q];
        parts[i + 3 * stride] = yp[q];
        parts[i + 5 * stride] = dpop[q];
    }
}

// ---------------------------------------------------------------------------
// Populate: Philox-4x32-10 (Salmon et al., SC'11) keyed by the 64-bit seed,
// counter (id_lo, id_hi, call, 0); 53-bit uniforms in (0,1); Box-Muller.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// (m + 0.5) 2^-53 for the top 53 bits m of (a, b): in (0, 1]; exactly 1.0
// only for m = 2^53 - 1 (the sum rounds up), which the transforms below take.
__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
    uint64_t m = (uint64_t(a) << 21) | (b >> 11);
    return (double(m) + 0.5) * 0x1.0p-53;
}

// Box-Muller's log and sin/cos(2 pi u) from IEEE basic operations only
// (+ - * / sqrt floor frexp, no FMA contraction in this library), so the
// same seed gives the same bunch bit for bit on any IEEE machine, whatever
// its libm.  Accuracy (measured against 200-bit arithmetic over 2e5 samples):
// log within 0.76 ulp, sin and cos within 2 ulp.
//
// log(u), u in (0, 1]: u = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1
// (exact), s = f / (2 + f), log(1 + f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2
// + R)) with R = sum_k 2 s^2k / (2k+1), k = 1..11 (|s| <= 0.1716, the
// truncation is below 1e-18 relative) -- the arrangement of fdlibm's e_log.c
// with Taylor coefficients; ln 2 split into a 32-bit head (e * head exact)
// and its tail.
__device__ __forceinline__ double det_log(double u)
{
    constexpr double kLn2Hi = 0x1.62e42feep-1;             // ln 2, top 32 bits
    constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;       // ln 2 - kLn2Hi
    constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;
    int e;
    double m = frexp(u, &e);
    if (m < kSqrtHalf) {
        m = m + m;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 2.0 / 23.0;
    R = 2.0 / 21.0 + z * R;
    R = 2.0 / 19.0 + z * R;
    R = 2.0 / 17.0 + z * R;
    R = 2.0 / 15.0 + z * R;
    R = 2.0 / 13.0 + z * R;
    R = 2.0 / 11.0 + z * R;
    R = 2.0 / 9.0 + z * R;
    R = 2.0 / 7.0 + z * R;
    R = 2.0 / 5.0 + z * R;
    R = 2.0 / 3.0 + z * R;
    R = z * R;
    const double hfsq = 0.5 * f * f;
    const double dk = double(e);
    return dk * kLn2Hi + (f - (hfsq - (s * (hfsq + R) + dk * kLn2Lo)));
}

// sin and cos of 2 pi u, u in (0, 1]: quadrant q = floor(4u), r = 4u - q
// (both exact), the angle folded to a = r or 1 - r <= 1/2 (exact), x = a pi/2
// <= pi/4 (one rounding), Taylor series of sin x to x^19 and cos x to x^20
// (truncation below 1e-19), then the exact quadrant rotation.
__device__ __forceinline__ void det_sincos2pi(double u, double* sn, double* cs)
{
    constexpr double kPi2 = 0x1.921fb54442d18p+0;  // pi / 2
    const double t = 4.0 * u;
    const double q = floor(t);
    const double r = t - q;
    const bool fold = r > 0.5;
    const double a = fold ? 1.0 - r : r;
    const double x = a * kPi2;
    const double z = x * x;
    double S = -1.0 / 121645100408832000.0;  // -1/19!
    S = 1.0 / 355687428096000.0 + z * S;     // 1/17!
    S = -1.0 / 1307674368000.0 + z * S;      // -1/15!
    S = 1.0 / 6227020800.0 + z * S;          // 1/13!
    S = -1.0 / 39916800.0 + z * S;           // -1/11!
    S = 1.0 / 362880.0 + z * S;              // 1/9!
    S = -1.0 / 5040.0 + z * S;               // -1/7!
    S = 1.0 / 120.0 + z * S;                 // 1/5!
    S = -1.0 / 6.0 + z * S;                  // -1/3!
    const double s0 = x + x * (z * S);
    double Cp = 1.0 / 2432902008176640000.0;  // 1/20!
    Cp = -1.0 / 6402373705728000.0 + z * Cp;  // -1/18!
    Cp = 1.0 / 20922789888000.0 + z * Cp;     // 1/16!
    Cp = -1.0 / 87178291200.0 + z * Cp;       // -1/14!
    Cp = 1.0 / 479001600.0 + z * Cp;          // 1/12!
    Cp = -1.0 / 3628800.0 + z * Cp;           // -1/10!
    Cp = 1.0 / 40320.0 + z * Cp;              // 1/8!
    Cp = -1.0 / 720.0 + z * Cp;               // -1/6!
    Cp = 1.0 / 24.0 + z * Cp;                 // 1/4!
    Cp = -1.0 / 2.0 + z * Cp;                 // -1/2!
    const double c0 = 1.0 + z * Cp;
    const double sq = fold ? c0 : s0;  // sin and cos of the angle in the quadrant
    const double cq = fold ? s0 : c0;
    switch (int(q) & 3) {
    case 0: *sn = sq; *cs = cq; break;
    case 1: *sn = cq; *cs = -sq; break;
    case 2: *sn = -sq; *cs = -cq; break;
    default: *sn = -cq; *cs = sq; break;
    }
}

__global__ __launch_bounds__(kThreads) void k_populate(double* __restrict__ parts, int64_t stride,
                                                       int64_t n, uint8_t* __restrict__ masks,
                                                       PopulateArgs a)
{
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads) {
        int64_t id = a.id0 + i;
        uint2 key = make_uint2(uint32_t(a.seed), uint32_t(a.seed >> 32));
        double u[6];
        for (int call = 0; call < 3; ++call) {
            uint4 r = philox4x32_10(
                make_uint4(uint32_t(id), uint32_t(uint64_t(id) >> 32), uint32_t(call), 0u), key);
            u[2 * call] = u53(r.x, r.y);
            u[2 * call + 1] = u53(r.z, r.w);
        }
        double g[6];
        for (int k = 0; k < 3; ++k) {
            double rad = sqrt(-2.0 * det_log(u[2 * k]));
            double s, c;
            det_sincos2pi(u[2 * k + 1], &s, &c);
            g[2 * k] = rad * c;
            g[2 * k + 1] = rad * s;
        }
        for (int r = 0; r < 6; ++r) {
            double v = a.mean[r];
            for (int c = 0; c <= r; ++c) v += a.chol[r * (r + 1) / 2 + c] * g[c];
            parts[i + r * stride] = v;
        }
        parts[i + 6 * stride] = double(id);
        masks[i] = 1;
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// Blocks for n items, at most cap, rounded up to a multiple of 8 (the
// XCD-contiguous block order of part_block needs one; the kernels launched
// on these grids let surplus blocks run empty).
static int grid_for(int64_t n, int cap = 8192)
{
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return int((b + 7) / 8 * 8);
}

static unsigned round8(int64_t b) { return unsigned((b + 7) / 8 * 8); }

void launch_tail_counts(const double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                        const TailArgs& a, unsigned long long* counts, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, counts, 6);
    if (n > 0)
        hipLaunchKernelGGL(k_tail_counts, dim3(grid_for(n, 2048)), dim3(kThreads), 0, st, parts,
                           stride, masks, n, a, counts);
}


__global__ __launch_bounds__(kThreads) void k_adjust(double* __restrict__ parts, int64_t stride,
                                                     int64_t n, AdjustArgs a)
{
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads) {
        double d[6], o[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) d[k] = parts[i + k * stride] - a.mean[k];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += d[k] * a.A[j * 6 + k];
            o[j] = s + a.means[j];
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) parts[i + j * stride] = o[j];
    }
}

void launch_adjust(double* parts, int64_t stride, int64_t n, const AdjustArgs& a, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_adjust, dim3(grid_for(n)), dim3(kThreads), 0, st, parts, stride, n,
                           a);
}

// The density grid (grid_bytes, a multiple of 16) is cleared here: by the
// chunk pass when its chunks are enough for slices of at most 32 KiB each,
// else by a fill of its own.
void launch_domain(const double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                   double* P, const DomainFinishArgs& a, DomainRec* dom, const Sched& sc,
                   void* grid, int64_t grid_bytes, hipStream_t st)
{
    const int64_t nch = (n + 4095) / 4096;
    const int64_t grid16 = grid_bytes / 16;
    const bool chunks = !a.fixed && nch > 0;
    const bool clear = chunks && grid_bytes <= nch * 32768;
    if (chunks)
        hipLaunchKernelGGL(k_domain_partial, dim3(unsigned(nch)), dim3(kThreads), 0, st, parts,
                           stride, masks, n, P, sc.ctl, static_cast<uint4*>(grid),
                           clear ? grid16 : int64_t(0));
    if (!clear)
        hipLaunchKernelGGL(k_fill<uint4>, dim3(grid_for(grid16, 2048)), dim3(kThreads), 0, st,
                           static_cast<uint4*>(grid), grid16, make_uint4(0, 0, 0, 0));
    hipLaunchKernelGGL(k_domain_finish, dim3(1), dim3(kThreads), 0, st, P, chunks ? nch : 0, a,
                       dom, sc);
}

void launch_step_begin(StepCtl* ctl, hipStream_t st)
{
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, st, ctl);
}

int sort_passes(int64_t G)
{
    int bits = 0;
    while ((int64_t(1) << bits) <= G) ++bits;  // keys in [0, G]
    return (bits + kMaxDigitBits - 1) / kMaxDigitBits;
}

static int sort_bits(int64_t G)
{
    int bits = 0;
    while ((int64_t(1) << bits) <= G) ++bits;
    return bits;
}

void launch_sort(const double* parts, int64_t stride, const uint8_t* masks, int64_t n, int gx,
                 int gy, int gz, const DomainRec* dom, SortBuffers& b, hipStream_t st)
{
    const int nblk = int((n + kSortTile - 1) / kSortTile);
    const int bits = sort_bits(int64_t(gx) * gy * gz);
    const int passes = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
    const int width = (bits + passes - 1) / passes;
    const int nb = 1 << width;
    const int32_t* kin = b.keys0;
    const int32_t* vin = nullptr;
    int32_t* kout[2] = {b.keysA, b.keysB};
    int32_t* vout[2] = {b.valsA, b.valsB};
    for (int p = 0; p < passes; ++p) {
        const int shift = width * p;
        if (p == 0)
            hipLaunchKernelGGL(k_radix_count<true>, dim3(round8(nblk)), dim3(kThreads), 0, st,
                               parts, stride, masks, gx, gy, gz, b.keys0, n, shift, width,
                               b.counts, nblk, dom);
        else
            hipLaunchKernelGGL(k_radix_count<false>, dim3(round8(nblk)), dim3(kThreads), 0, st,
                               (const double*)nullptr, int64_t(0), (const uint8_t*)nullptr, gx,
                               gy, gz, const_cast<int32_t*>(kin), n, shift, width, b.counts,
                               nblk, dom);
        hipLaunchKernelGGL(k_radix_scan, dim3(nb), dim3(kThreads), 0, st, b.counts, nblk,
                           b.totals, dom);
        hipLaunchKernelGGL(k_radix_scatter, dim3(round8(nblk)), dim3(kThreads), 0, st, kin, vin,
                           n, shift, width, b.counts, b.totals, nblk, kout[p & 1], vout[p & 1],
                           dom);
        kin = kout[p & 1];
        vin = vout[p & 1];
    }
    b.sorted_keys = const_cast<int32_t*>(kin);
    b.perm = const_cast<int32_t*>(vin);
}

void launch_reorder(double* parts, int64_t stride, uint8_t* masks, int64_t n, const int32_t* perm,
                    double* scratch, uint8_t* smask, const DomainRec* dom, hipStream_t st)
{
    // One particle per thread up to 1.7e7 particles: the gather is bound by the latency of its
    // random reads, so more waves in flight help (measured at 1.6e7 particles, one sort every
    // 12 steps: sort plus reorder 132 us per step on average with 65536 blocks, 138 with 8192).
    // The copy of the gathered bunch back into the storage is the kick's (k_kick*<.., COPY>):
    // until then the deposit reads the gathered copy.
    const dim3 grid(grid_for(n, 65536));
    hipLaunchKernelGGL(k_gather_particles, grid, dim3(kThreads), 0, st, parts, stride, masks,
                       perm, n, scratch, smask, dom);
}

// Particles per lane of the deposit: 4 on dense bunches, 2 on sparse ones.
// Measured (deposit phase, us): 4e6 particles on 64x64x128 (7.6 per cell)
// FP64 K=4 72, K=2 74, K=8 99; FP32 K=4 84, K=2 86.  1.6e7 on 128x128x256
// (3.8 per cell) FP32 K=4 317, K=2 264, K=8 421; FP64 K=4 324, K=2 284.
// With few particles per cell most lanes end runs inside their K particles,
// and those additions run under divergent masks.
template <typename R>
static void launch_deposit_k(const double* parts, int64_t stride, const uint8_t* masks,
                             const int32_t* perm, int64_t n, int gx, int gy, int gz,
                             DomainRec* dom, R* rho, hipStream_t st)
{
    const int64_t G = int64_t(gx) * gy * gz;
    const int K = n >= 6 * G ? 4 : 2;
    const int64_t nb = (n + int64_t(kThreads) * K - 1) / (int64_t(kThreads) * K);
    if (nb < 1) return;
    if (K == 4)
        hipLaunchKernelGGL((k_deposit<R, 4>), dim3(unsigned(nb)), dim3(kThreads), 0, st, parts,
                           stride, masks, perm, n, gx, gy, gz, dom, rho);
    else
        hipLaunchKernelGGL((k_deposit<R, 2>), dim3(unsigned(nb)), dim3(kThreads), 0, st, parts,
                           stride, masks, perm, n, gx, gy, gz, dom, rho);
}

template <typename R>
static void launch_deposit_local(const double* parts, int64_t stride, const uint8_t* masks,
                                 int64_t n, int gx, int gy, int gz, DomainRec* dom, R* rho,
                                 hipStream_t st)
{
    const int64_t nb = (n + kTileWindow - 1) / kTileWindow;
    if (nb < 1) return;
    hipLaunchKernelGGL(k_deposit_tile<R>, dim3(unsigned(nb)), dim3(kTileThreads), 0, st, parts,
                       stride, masks, n, gx, gy, gz, dom, rho);
}

template <typename R>
void launch_deposit(const double* parts, int64_t stride, const uint8_t* masks,
                    const int32_t* perm, bool local, int64_t n, int gx, int gy, int gz,
                    DomainRec* dom, R* rho, hipStream_t st)
{
    if (local)
        launch_deposit_local<R>(parts, stride, masks, n, gx, gy, gz, dom, rho, st);
    else
        launch_deposit_k<R>(parts, stride, masks, perm, n, gx, gy, gz, dom, rho, st);
}

template <>
void launch_deposit<float>(const double* parts, int64_t stride, const uint8_t* masks,
                           const int32_t* perm, bool local, int64_t n, int gx, int gy, int gz,
                           DomainRec* dom, float* rho, hipStream_t st)
{
    if (local)
        launch_deposit_local<float>(parts, stride, masks, n, gx, gy, gz, dom, rho, st);
    else
        launch_deposit_k<float>(parts, stride, masks, perm, n, gx, gy, gz, dom, rho, st);
    const int64_t G = int64_t(gx) * gy * gz;
    hipLaunchKernelGGL(k_rho_peak, dim3(grid_for(G / 8, 512)), dim3(kThreads), 0, st, rho, G, dom);
}

template <typename R>
void launch_efield(const R* phi, int gx, int gy, int gz, const DomainRec* dom,
                   typename vec4<R>::type* E, hipStream_t st)
{
    hipLaunchKernelGGL(k_efield<R>, dim3(grid_for(int64_t(gx) * gy * gz)), dim3(kThreads), 0, st,
                       phi, gx, gy, gz, dom, E);
}

template <typename R>
void launch_kick(double* parts, int64_t stride, uint8_t* masks, int64_t n, int gx, int gy, int gz,
                 const DomainRec* dom, const typename vec4<R>::type* E, double pref, double m,
                 const MapProgram& prog, const Sched& sc, const double* src,
                 const uint8_t* smasks, hipStream_t st)
{
    if (n <= 0) return;
    const bool copy = src != nullptr;
    if (prog.n == 0 && !sc.progs) {
        constexpr int win = kThreads * kKickPer;
        const dim3 grid(unsigned((n + win - 1) / win));
        if (copy)
            hipLaunchKernelGGL((k_kick_tile<R, true>), grid, dim3(kThreads), 0, st, parts, stride,
                               masks, n, gx, gy, gz, dom, E, pref, m, src, smasks);
        else
            hipLaunchKernelGGL((k_kick_tile<R, false>), grid, dim3(kThreads), 0, st, parts,
                               stride, masks, n, gx, gy, gz, dom, E, pref, m, src, smasks);
        return;
    }
    // kick + maps: up to 32768 blocks (measured at 1.6e7 particles: 240 us,
    // against 251 with 8192 blocks, 272 with 4096 and 247 with one particle
    // per thread, for the kick alone)
    const dim3 grid(grid_for(n, 32768));
    if (copy)
        hipLaunchKernelGGL((k_kick<R, true, true>), grid, dim3(kThreads), 0, st, parts, stride,
                           masks, n, gx, gy, gz, dom, E, pref, m, prog, sc, src, smasks);
    else
        hipLaunchKernelGGL((k_kick<R, true, false>), grid, dim3(kThreads), 0, st, parts, stride,
                           masks, n, gx, gy, gz, dom, E, pref, m, prog, sc, src, smasks);
}

void launch_populate(double* parts, int64_t stride, int64_t n, uint8_t* masks,
                     const PopulateArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(k_populate, dim3(grid_for(n)), dim3(kThreads), 0, st, parts, stride, n,
                       masks, a);
}

template void launch_deposit<double>(const double*, int64_t, const uint8_t*, const int32_t*,
                                     bool, int64_t, int, int, int, DomainRec*, double*,
                                     hipStream_t);
template void launch_efield<double>(const double*, int, int, int, const DomainRec*, double4*,
                                    hipStream_t);
template void launch_efield<float>(const float*, int, int, int, const DomainRec*, float4*,
                                   hipStream_t);
template void launch_kick<double>(double*, int64_t, uint8_t*, int64_t, int, int, int,
                                  const DomainRec*, const double4*, double, double,
                                  const MapProgram&, const Sched&, const double*, const uint8_t*,
                                  hipStream_t);
template void launch_kick<float>(double*, int64_t, uint8_t*, int64_t, int, int, int,
                                 const DomainRec*, const float4*, double, double,
                                 const MapProgram&, const Sched&, const double*, const uint8_t*,
                                 hipStream_t);

}  // namespace sc3d
