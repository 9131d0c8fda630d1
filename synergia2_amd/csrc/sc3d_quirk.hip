// sc3d_quirk.hip -- the reference's Green's function at the grid shapes where
// its flat-index decomposition misplaces points.
//
// alg_g2_pointlike and alg_g2_linear (space_charge_3d_open_hockney.cc:170-176,
// 272-278) split their flat index i with double reciprocals,
//   iz = int(i * (1.0 / (gx*gy))),  iy = int((i - iz*gx*gy) * (1.0 / gx)),
//   ix = i - iz*gx*gy - iy*gx            (gx = shape[0]+1, gy = shape[1]+1),
// and when an exact multiple i = m*gx*gy (or a row start k*gx) times the
// rounded reciprocal lands just below the integer, the point comes out one
// plane (row) short with iy = gy (ix = gx).  That i then writes its value
// and mirrors over correct points of the previous plane (row), and its own
// point keeps the zero g2 was cleared with.  At (16, 32, z) and (32, 16, z)
// this happens for a few planes; every configured shape decomposes exactly.
//
// The device computes the spectrum of the symmetric g2 (the value the exact
// decomposition gives everywhere: sc3d_fft.hip), so at those shapes the
// difference g2_ref - g2_sym is a set of points ("patches"), found here
// on the host once per shape by replaying the reference's writes in its
// serial order (ascending i; the last write to a point stands, as in a
// one-thread run).  Each step the patch values are evaluated with the
// step's h and their convolution with rho is added in direct space:
//   FP64 form:      phi(r) += Dx*Dy*Dz * sum_p d_p rho((r - T_p) mod D)
//                   (unnormalised, like the FFT product), before the field;
//   precision form: the centred difference of the same sum (unit-peak
//                   density, as the spectral field) is added to E.
// Among shapes of other lengths the misplacement can be massive (at gx = 48,
// 1.0/49 makes most row starts k*49 land one row short, in every plane:
// 29376 points at 48x48x96): beyond kMaxGreenPatches points the difference
// is transformed instead -- C[Dz][Dy][Dx] holds the patch values, its full
// complex spectrum chat is added to the real Green's spectrum in the z pass
// (all axes on the generic passes, sc3d_fft_any.hip).
#include <algorithm>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "sc3d_ctx.h"
#include "sc3d_green.h"

namespace sc3d {

namespace {

using P3 = std::tuple<int, int, int>;  // (z, y, x)

struct Decomp {
    int gx1, gy1, g0x, g0y, g0z, dgx, dgy, dgz, nz;
    bool linear;
    double igxgy, igx;

    P3 floatwise(int64_t i) const
    {
        int iz = int(double(i) * igxgy);
        int iy = int(double(i - int64_t(iz) * gx1 * gy1) * igx);
        int ix = int(i - int64_t(iz) * gx1 * gy1 - int64_t(iy) * gx1);
        return {iz, iy, ix};
    }
    P3 exact(int64_t i) const
    {
        int64_t pl = int64_t(gx1) * gy1;
        int iz = int(i / pl);
        int iy = int((i - iz * pl) / gx1);
        int ix = int(i - iz * pl - int64_t(iy) * gx1);
        return {iz, iy, ix};
    }
    int64_t index(int iz, int iy, int ix) const
    {
        return (int64_t(iz) * gy1 + iy) * gx1 + ix;
    }
    // the points one index writes (linear :328-346, pointlike :206-227)
    std::vector<P3> writes(const P3& p) const
    {
        auto [iz, iy, ix] = p;
        int mix = dgx - ix;
        if (mix == dgx) mix = 0;
        int miy = dgy - iy;
        if (miy == dgy) miy = 0;
        std::vector<P3> w;
        if (linear) {
            w.push_back({iz, iy, ix});
            if (iy != g0y) {
                w.push_back({iz, miy, ix});
                if (ix != g0x) w.push_back({iz, miy, mix});
            }
            if (ix != g0x) w.push_back({iz, iy, mix});
        } else {
            int miz = dgz - iz;
            if (miz == dgz) miz = 0;
            for (int a : {iz, miz})
                for (int b : {iy, miy})
                    for (int c : {ix, mix}) w.push_back({a, b, c});
        }
        return w;
    }
};

}  // namespace

int green_patches(int green_fn, const int shape[3], std::vector<GreenPatch>& out)
{
    out.clear();
    Decomp d;
    d.g0x = shape[0];
    d.g0y = shape[1];
    d.g0z = shape[2];
    d.gx1 = shape[0] + 1;
    d.gy1 = shape[1] + 1;
    d.dgx = 2 * shape[0];
    d.dgy = 2 * shape[1];
    d.dgz = 2 * shape[2];
    d.linear = green_fn == SC3D_GREEN_LINEAR;
    d.nz = d.linear ? d.dgz : shape[2] + 1;
    d.igxgy = 1.0 / (d.gx1 * d.gy1);
    d.igx = 1.0 / d.gx1;
    // Only exact multiples can decompose wrongly: for any other i the
    // product is at least 1/(gx*gy) (1/gx) above the integer below it, far
    // more than its rounding error (one ulp of a number below 2^31).
    std::map<int64_t, P3> quirk;  // i -> its (wrong) decomposition
    for (int m = 1; m < d.nz; ++m) {
        int64_t i = int64_t(m) * d.gx1 * d.gy1;
        P3 f = d.floatwise(i);
        if (f != d.exact(i)) quirk[i] = f;
    }
    for (int iz = 0; iz < d.nz; ++iz)
        for (int k = 1; k < d.gy1; ++k) {
            int64_t i = d.index(iz, k, 0);
            P3 f = d.floatwise(i);
            if (f != d.exact(i)) quirk[i] = f;
        }
    if (quirk.empty()) return 0;
    // every point a misplaced index writes or should have written, and the
    // last (largest) misplaced index writing each point
    std::set<P3> touched;
    std::map<P3, std::pair<int64_t, P3>> qlast;
    for (auto& [i, f] : quirk) {  // ascending i
        for (const P3& t : d.writes(f)) {
            touched.insert(t);
            qlast[t] = {i, f};
        }
        for (const P3& t : d.writes(d.exact(i))) touched.insert(t);
    }
    for (const P3& t : touched) {
        auto [Z, Y, X] = t;
        // the last writer in ascending i: exact-decomposition writers (not
        // themselves misplaced) and misplaced ones
        int64_t best = -1;
        P3 src{-1, -1, -1};
        std::vector<int> zs{Z}, ys{Y}, xs{X};
        if (d.dgy - Y < d.dgy) ys.push_back(d.dgy - Y);
        if (d.dgx - X < d.dgx) xs.push_back(d.dgx - X);
        if (!d.linear && d.dgz - Z < d.dgz) zs.push_back(d.dgz - Z);
        for (int iz : zs)
            for (int iy : ys)
                for (int ix : xs) {
                    if (iz < 0 || iz >= d.nz || iy < 0 || iy > d.g0y || ix < 0 || ix > d.g0x)
                        continue;
                    int64_t i = d.index(iz, iy, ix);
                    if (quirk.count(i)) continue;
                    P3 e{iz, iy, ix};
                    auto w = d.writes(e);
                    if (std::find(w.begin(), w.end(), t) != w.end() && i > best) {
                        best = i;
                        src = e;
                    }
                }
        bool exact_last = best >= 0;
        auto q = qlast.find(t);
        if (q != qlast.end() && q->second.first > best) {
            best = q->second.first;
            src = q->second.second;
            exact_last = false;
        }
        if (exact_last) continue;  // the symmetric value stands
        GreenPatch p;
        p.t[0] = Z; p.t[1] = Y; p.t[2] = X;
        p.s[0] = std::get<0>(src); p.s[1] = std::get<1>(src); p.s[2] = std::get<2>(src);
        out.push_back(p);
    }
    return int(out.size());
}

// ---------------------------------------------------------------------------
// The patch values of this step (d_p = value the reference leaves at T_p -
// symmetric value there), then their convolution with rho.
// ---------------------------------------------------------------------------
template <typename R, bool FIELD>
__global__ __launch_bounds__(kThreads) void k_green_patch(
    const GreenPatch* __restrict__ patches, int np, int gx, int gy, int gz, int green_fn,
    const DomainRec* __restrict__ dom, const R* __restrict__ rho, double* __restrict__ phi,
    typename vec4_of<R>::type* __restrict__ E)
{
    __shared__ double dlt[kMaxGreenPatches];
    __shared__ int tz[kMaxGreenPatches], ty[kMaxGreenPatches], tx[kMaxGreenPatches];
    if (dom->error) return;
    const double hx = dom->h[0], hy = dom->h[1], hz = dom->h[2];
    const int Dx = 2 * gx, Dy = 2 * gy, Dz = 2 * gz;
    const double G000 = green_g000(green_fn, hx, hy, hz);
    for (int p = threadIdx.x; p < np; p += kThreads) {
        const GreenPatch P = patches[p];
        const int Z = P.t[0], Y = P.t[1], X = P.t[2];
        double sym = green_value(green_fn, min(X, Dx - X) * hx, min(Y, Dy - Y) * hy,
                                 min(Z, Dz - Z) * hz, hz, G000);
        double val = 0.0;
        if (P.s[0] >= 0) {
            const int iz = P.s[0], iy = P.s[1], ix = P.s[2];
            // linear: z folded as alg_g2_linear does (:280); pointlike: iz <= gz
            const double z = (green_fn == SC3D_GREEN_LINEAR && iz > gz) ? (Dz - iz) * hz : iz * hz;
            val = green_value(green_fn, ix * hx, iy * hy, z, hz, G000);
        }
        dlt[p] = val - sym;
        tz[p] = Z;
        ty[p] = Y;
        tx[p] = X;
    }
    __syncthreads();
    const double ND = double(Dx) * double(Dy) * double(Dz);
    double scale = ND;
    if (sizeof(R) == sizeof(float)) {
        // the precision form's spectral field works on the unit-peak density
        const float peak = __int_as_float(dom->rho_peak);
        if (peak > 0.0f) scale = ND / double(peak);
    }
    // convolution of the patches with rho at octant node (z, y, x)
    auto conv = [&](int z, int y, int x) -> double {
        double acc = 0.0;
        for (int p = 0; p < np; ++p) {
            int sz = z - tz[p], sy = y - ty[p], sx = x - tx[p];
            if (sz < 0) sz += Dz;
            if (sy < 0) sy += Dy;
            if (sx < 0) sx += Dx;
            if (sz < gz && sy < gy && sx < gx)
                acc += dlt[p] * double(rho[(int64_t(sz) * gy + sy) * gx + sx]);
        }
        return scale * acc;
    };
    const int64_t G = int64_t(gx) * gy * gz;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < G;
         i += int64_t(gridDim.x) * kThreads) {
        const int x = int(i % gx);
        const int64_t r = i / gx;
        const int y = int(r % gy), z = int(r / gy);
        if (!FIELD) {
            phi[i] += conv(z, y, x);
        } else if (x > 0 && x < gx - 1 && y > 0 && y < gy - 1 && z > 0 && z < gz - 1) {
            // the reference's centred difference (space_charge_3d_kernels.h:
            // 262-300) of the patch potential, boundary nodes untouched
            const double* ih = dom->half_inv_h;
            typename vec4_of<R>::type e = E[i];
            e.x = R(double(e.x) - (conv(z, y, x + 1) - conv(z, y, x - 1)) * ih[0]);
            e.y = R(double(e.y) - (conv(z, y + 1, x) - conv(z, y - 1, x)) * ih[1]);
            e.z = R(double(e.z) - (conv(z + 1, y, x) - conv(z - 1, y, x)) * ih[2]);
            E[i] = e;
        }
    }
}

// Many patch points (spectral_patch): their values, C[T_p] = d_p, into the
// otherwise zero real array C[Dz][Dy][Dx] (cleared at create; the points
// never change), where h changed.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_patch_values(
    const GreenPatch* __restrict__ patches, int np, int gx, int gy, int gz, int green_fn,
    const DomainRec* __restrict__ dom, R* __restrict__ C)
{
    if (dom->error || !dom->ghat_dirty) return;
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= np) return;
    const double hx = dom->h[0], hy = dom->h[1], hz = dom->h[2];
    const int Dx = 2 * gx, Dy = 2 * gy, Dz = 2 * gz;
    const double G000 = green_g000(green_fn, hx, hy, hz);
    const GreenPatch P = patches[p];
    const int Z = P.t[0], Y = P.t[1], X = P.t[2];
    const double sym = green_value(green_fn, min(X, Dx - X) * hx, min(Y, Dy - Y) * hy,
                                   min(Z, Dz - Z) * hz, hz, G000);
    double val = 0.0;
    if (P.s[0] >= 0) {
        const int iz = P.s[0], iy = P.s[1], ix = P.s[2];
        const double z = (green_fn == SC3D_GREEN_LINEAR && iz > gz) ? (Dz - iz) * hz : iz * hz;
        val = green_value(green_fn, ix * hx, iy * hy, z, hz, G000);
    }
    C[(int64_t(Z) * Dy + Y) * Dx + X] = R(val - sym);
}

void launch_green_chat(const sc3d_ctx* c, hipStream_t st)
{
    if (!c->spectral_patch) return;
    const Geometry& g = c->g;
    const int blocks = (c->npatch + kThreads - 1) / kThreads;
    if (!c->fp32) {
        hipLaunchKernelGGL(k_patch_values<double>, dim3(blocks), dim3(kThreads), 0, st,
                           c->d_patch, c->npatch, g.gx, g.gy, g.gz, c->opt.green_fn, c->dom,
                           static_cast<double*>(c->Cbuf));
        any_chat<double>(g, c->Cbuf, c->d_gj, c->chat, c->tw64, st);
    } else {
        hipLaunchKernelGGL(k_patch_values<float>, dim3(blocks), dim3(kThreads), 0, st,
                           c->d_patch, c->npatch, g.gx, g.gy, g.gz, c->opt.green_fn, c->dom,
                           static_cast<float*>(c->Cbuf));
        any_chat<float>(g, c->Cbuf, c->d_gj, c->chat, c->tw32, st);
    }
}

void launch_green_patches(const sc3d_ctx* c, hipStream_t st)
{
    if (c->npatch == 0 || c->spectral_patch) return;
    const Geometry& g = c->g;
    const int64_t blocks = std::min<int64_t>((g.G + kThreads - 1) / kThreads, 2048);
    if (!c->fp32)
        hipLaunchKernelGGL((k_green_patch<double, false>), dim3(blocks), dim3(kThreads), 0, st,
                           c->d_patch, c->npatch, g.gx, g.gy, g.gz, c->opt.green_fn, c->dom,
                           static_cast<const double*>(c->rho), static_cast<double*>(c->phi),
                           nullptr);
    else
        hipLaunchKernelGGL((k_green_patch<float, true>), dim3(blocks), dim3(kThreads), 0, st,
                           c->d_patch, c->npatch, g.gx, g.gy, g.gz, c->opt.green_fn, c->dom,
                           static_cast<const float*>(c->rho), nullptr,
                           static_cast<float4*>(c->E));
}

}  // namespace sc3d

extern "C" int sc3d_green_patches(int32_t green_fn, const int32_t shape[3], int32_t* out,
                                  int32_t cap)
{
    if (!shape || shape[0] < 1 || shape[1] < 1 || shape[2] < 1 ||
        (green_fn != SC3D_GREEN_LINEAR && green_fn != SC3D_GREEN_POINTLIKE))
        return sc3d::fail(SC3D_ERR_ARG, "sc3d_green_patches: bad argument");
    const int s[3] = {shape[0], shape[1], shape[2]};
    std::vector<sc3d::GreenPatch> p;
    const int n = sc3d::green_patches(green_fn, s, p);
    if (out)
        for (int k = 0; k < n && k < cap; ++k)
            for (int j = 0; j < 3; ++j) {
                out[6 * k + j] = p[k].t[j];
                out[6 * k + 3 + j] = p[k].s[j];
            }
    return n;
}
