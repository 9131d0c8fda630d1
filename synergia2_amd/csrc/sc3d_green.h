// sc3d_green.h -- the reference's Green's functions, evaluated pointwise on
// the device (space_charge_3d_open_hockney.cc:136-348): the integrated
// (linear) Green's function of alg_g2_linear and the point-charge one of
// alg_g2_pointlike, with their origin values G000.  Shared by the spectrum
// passes (sc3d_fft.hip) and the index-decomposition patches
// (sc3d_quirk.hip), so both evaluate the same arithmetic.
#pragma once

#include "sc3d_common.h"

namespace sc3d {

__device__ __forceinline__ double green_linear(double x, double y, double z, double hz, double G000)
{
    double rr = x * x + y * y;
    double s_rr_0 = sqrt(rr + z * z);
    double s_rr_n = sqrt(rr + (z - hz) * (z - hz));
    double s_rr_p = sqrt(rr + (z + hz) * (z + hz));
    double G = 2.0 * s_rr_0 - s_rr_n - s_rr_p;
    double T1, T2, r1, r2;
    const double epsz = 1e-12 * hz;
    // z >= 0 here, so of the reference's branches (:294-326) only these three
    // are reachable; the z < -hz and |z + hz| < epsz ones never are.
    if (fabs(z) < epsz) {
        if (fabs(x) + fabs(y) < 2. * epsz) {
            G += hz * G000;
        } else {
            r1 = (sqrt(hz * hz + rr) + hz) / sqrt(rr);
            G += 2.0 * hz * log(r1);
        }
    } else if (fabs(z - hz) < epsz) {
        r1 = (s_rr_p + z + hz) / (s_rr_0 + z);
        T1 = (hz + z) * log(r1);
        G += T1;
    } else if (z > hz) {
        r1 = (s_rr_0 + z) / (s_rr_n + z - hz);
        T1 = (hz - z) * log(r1);
        r2 = (s_rr_p + z + hz) / (s_rr_0 + z);
        T2 = (hz + z) * log(r2);
        G += T1 + T2;
    }
    return G;
}

// G000 of the two functions (space_charge_3d_open_hockney.cc:165,265-270).
__device__ __forceinline__ double green_g000(int green_fn, double hx, double hy, double hz)
{
    if (green_fn == SC3D_GREEN_LINEAR) {
        double rr = hx * hx + hy * hy;
        double r1 = sqrt(hx * hx + hy * hy + hz * hz);
        return (2.0 / rr) * (hz * r1 + rr * log((hz + r1) / sqrt(rr)) - hz * hz);
    }
    return 2.8 / fmin(hx, fmin(hy, hz));
}

// The value the reference writes for the point with coordinates (x, y, z)
// >= 0 (z already folded: (Dz - iz) * hz above the middle plane for the
// linear function); pointlike: G000 at the origin, 1/r elsewhere.
__device__ __forceinline__ double green_value(int green_fn, double x, double y, double z,
                                              double hz, double G000)
{
    if (green_fn == SC3D_GREEN_LINEAR) return green_linear(x, y, z, hz, G000);
    return (x == 0.0 && y == 0.0 && z == 0.0) ? G000 : 1.0 / sqrt(x * x + y * y + z * z);
}

}  // namespace sc3d
