// sc3d_maps.h -- single-particle maps of the FODO channel (drift and
// quadrupole), host+device so the design reference particle's cdt is
// computed with the identical arithmetic.
//
//   drift_unit          <- FF_algorithm::drift_unit, libFF/ff_algorithm.h:45-79
//   thin_quadrupole_unit<- libFF/ff_algorithm.h:928-936
//   yoshida6            <- libFF/ff_algorithm.h:1835-2008 (10 drifts and 9
//                          thin kicks per step; default 4 steps per slice,
//                          ff_quadrupole.h:438-441)
//   reference cdt       <- drift_impl::get_reference_cdt (ff_drift.h:68-87),
//                          quad_impl::get_reference_cdt (ff_quadrupole.h:262-300)
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/sc3d_capi.h"

namespace sc3d {

__host__ __device__ inline void drift_unit(double& x, double xp, double& y, double yp,
                                           double& cdt, double dpop, double length,
                                           double reference_momentum, double m,
                                           double reference_cdt)
{
    const double uni = 1.0;
    const double sig = double((0.0 < length) - (length < 0.0));
    const double vl = length, vm = m, vrm = reference_momentum, vrc = reference_cdt;
    double dp = dpop + uni;
    double inv_npz = uni / sqrt(dp * dp - xp * xp - yp * yp);
    double lxpr = xp * vl * inv_npz;
    double lypr = yp * vl * inv_npz;
    double D2 = lxpr * lxpr + vl * vl + lypr * lypr;
    double p = dp * vrm;
    double E2 = p * p + vm * vm;
    double ibeta2 = E2 / (p * p);
    x = x + lxpr;
    y = y + lypr;
    cdt = cdt + sig * sqrt(D2 * ibeta2) - vrc;
}

__host__ __device__ inline void thin_quadrupole_unit(double x, double& xp, double y, double& yp,
                                                     const double* kL)
{
    const double vk0 = kL[0], vk1 = kL[1];
    xp = xp - vk0 * x + vk1 * y;
    yp = yp + vk0 * y + vk1 * x;
}

__host__ __device__ inline void yoshida6_quad(double& x, double& xp, double& y, double& yp,
                                              double& cdt, double dpop, double reference_momentum,
                                              double m, double step_reference_cdt,
                                              double step_length, const double* step_strength,
                                              int steps)
{
    const double c1 = 0.79361246386112147294625603763;
    const double c2 = -0.206276584816439780698721319354;
    const double c3 = -0.118008867881292655922412133144;
    const double c4 = 0.236949573653050744373598734219;
    const double d1 = 1.58722492772224294589251207526;
    const double d2 = -1.99977809735512250728995471397;
    const double d3 = -1.82324266348482825773733634154;
    const double d4 = 2.29714181079092974648453380998;
    const double sub = step_reference_cdt * 0.1;
    const double k1[2] = {d1 * step_strength[0], d1 * step_strength[1]};
    const double k2[2] = {d2 * step_strength[0], d2 * step_strength[1]};
    const double k3[2] = {d3 * step_strength[0], d3 * step_strength[1]};
    const double k4[2] = {d4 * step_strength[0], d4 * step_strength[1]};
    const double rp = reference_momentum;
    for (int i = 0; i < steps; ++i) {
        drift_unit(x, xp, y, yp, cdt, dpop, c1 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k1);
        drift_unit(x, xp, y, yp, cdt, dpop, c2 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k2);
        drift_unit(x, xp, y, yp, cdt, dpop, c2 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k1);
        drift_unit(x, xp, y, yp, cdt, dpop, c3 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k3);
        drift_unit(x, xp, y, yp, cdt, dpop, c4 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k4);
        drift_unit(x, xp, y, yp, cdt, dpop, c4 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k3);
        drift_unit(x, xp, y, yp, cdt, dpop, c3 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k1);
        drift_unit(x, xp, y, yp, cdt, dpop, c2 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k2);
        drift_unit(x, xp, y, yp, cdt, dpop, c2 * step_length, rp, m, sub);
        thin_quadrupole_unit(x, xp, y, yp, k1);
        drift_unit(x, xp, y, yp, cdt, dpop, c1 * step_length, rp, m, sub);
    }
}

// One slice applied to one particle.  For a quadrupole: per-step strength
// k*(length/steps) (ff_quadrupole.h:446), step length length/steps, step
// reference cdt ref_cdt/steps (ff_quadrupole.h:466-468).
__host__ __device__ inline void apply_slice(const sc3d_map_slice& s, double& x, double& xp,
                                            double& y, double& yp, double& cdt, double dpop,
                                            double ref_p, double m)
{
    if (s.kind == SC3D_MAP_DRIFT) {
        drift_unit(x, xp, y, yp, cdt, dpop, s.length, ref_p, m, s.ref_cdt);
    } else {
        const int steps = s.yoshida_steps;
        const double kn[2] = {s.k1 * (s.length / steps), s.k1s * (s.length / steps)};
        yoshida6_quad(x, xp, y, yp, cdt, dpop, ref_p, m, s.ref_cdt / steps, s.length / steps, kn,
                      steps);
    }
}

// Map program passed by value to kernels (no device allocation).
struct MapProgram {
    sc3d_map_slice s[SC3D_MAX_SLICES];
    int32_t n;
    double ref_p, mass;
};

}  // namespace sc3d
