// sc3d_maps.hip -- fused map program (drift / quadrupole slices) over the
// bunch: each active particle is loaded once, carried through every slice in
// registers (FP64) and stored once.
#include <string>

#include "sc3d_common.h"
#include "sc3d_maps.h"
#include "sc3d_particles.h"

namespace sc3d {

__global__ __launch_bounds__(kThreads) void k_maps(double* __restrict__ parts, int64_t stride,
                                                   const uint8_t* __restrict__ masks, int64_t n,
                                                   MapProgram prog)
{
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads) {
        if (!masks[i]) continue;
        double x = parts[i], xp = parts[i + stride], y = parts[i + 2 * stride],
               yp = parts[i + 3 * stride], cdt = parts[i + 4 * stride];
        const double dpop = parts[i + 5 * stride];
        for (int k = 0; k < prog.n; ++k)
            apply_slice(prog.s[k], x, xp, y, yp, cdt, dpop, prog.ref_p, prog.mass);
        parts[i] = x;
        parts[i + stride] = xp;
        parts[i + 2 * stride] = y;
        parts[i + 3 * stride] = yp;
        parts[i + 4 * stride] = cdt;
    }
}

// Served form: the scheduled program before the kick (first step only).
__global__ __launch_bounds__(kThreads) void k_maps_sched(double* __restrict__ parts,
                                                         int64_t stride,
                                                         const uint8_t* __restrict__ masks,
                                                         int64_t n, const MapProgram* progs,
                                                         const StepCtl* ctl)
{
    if (ctl->done || ctl->first_prog < 0) return;
    const MapProgram& prog = progs[ctl->first_prog];
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads) {
        if (!masks[i]) continue;
        double x = parts[i], xp = parts[i + stride], y = parts[i + 2 * stride],
               yp = parts[i + 3 * stride], cdt = parts[i + 4 * stride];
        const double dpop = parts[i + 5 * stride];
        for (int k = 0; k < prog.n; ++k)
            apply_slice(prog.s[k], x, xp, y, yp, cdt, dpop, prog.ref_p, prog.mass);
        parts[i] = x;
        parts[i + stride] = xp;
        parts[i + 2 * stride] = y;
        parts[i + 3 * stride] = yp;
        parts[i + 4 * stride] = cdt;
    }
}

void launch_maps_sched(double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                       const Sched& sc, hipStream_t st)
{
    int64_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) return;
    hipLaunchKernelGGL(k_maps_sched, dim3(unsigned(blocks)), dim3(kThreads), 0, st, parts, stride,
                       masks, n, sc.progs, sc.ctl);
}

}  // namespace sc3d

extern "C" {

int sc3d_internal_fail(int code, const char* msg);

double sc3d_reference_cdt(const sc3d_map_slice* s, double ref_p, double mass)
{
    // on-axis design reference: state zero, dpop zero, reference cdt zero
    double x = 0.0, xp = 0.0, y = 0.0, yp = 0.0, cdt = 0.0;
    sc3d_map_slice z = *s;
    z.ref_cdt = 0.0;
    sc3d::apply_slice(z, x, xp, y, yp, cdt, 0.0, ref_p, mass);
    return cdt;
}

int sc3d_apply_maps(const sc3d_bunch* b, const sc3d_map_slice* slices, int nslices, void* stream)
{
    if (!b || (nslices > 0 && !slices) || nslices < 0 || nslices > SC3D_MAX_SLICES)
        return sc3d_internal_fail(SC3D_ERR_ARG, "sc3d_apply_maps: bad argument");
    for (int k = 0; k < nslices; ++k) {
        if (slices[k].kind != SC3D_MAP_DRIFT && slices[k].kind != SC3D_MAP_QUADRUPOLE)
            return sc3d_internal_fail(SC3D_ERR_ARG, "sc3d_apply_maps: unknown slice kind");
        if (slices[k].kind == SC3D_MAP_QUADRUPOLE && slices[k].yoshida_steps < 1)
            return sc3d_internal_fail(SC3D_ERR_ARG, "sc3d_apply_maps: yoshida_steps < 1");
    }
    const int64_t n = b->n_active;
    if (n <= 0 || nslices == 0) return SC3D_OK;
    if (!b->parts || !b->masks || b->stride < n)
        return sc3d_internal_fail(SC3D_ERR_ARG, "sc3d_apply_maps: bad bunch");
    sc3d::MapProgram prog{};
    for (int k = 0; k < nslices; ++k) prog.s[k] = slices[k];
    prog.n = nslices;
    prog.ref_p = b->ref_p;
    prog.mass = b->mass;
    int64_t blocks = (n + sc3d::kThreads - 1) / sc3d::kThreads;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(sc3d::k_maps, dim3(unsigned(blocks)), dim3(sc3d::kThreads), 0,
                       static_cast<hipStream_t>(stream), b->parts, b->stride, b->masks, n, prog);
    if (hipGetLastError() != hipSuccess)
        return sc3d_internal_fail(SC3D_ERR_HIP, "sc3d_apply_maps: launch failed");
    return SC3D_OK;
}

}  // extern "C"
