// operator_driver.cc -- drives the C++ Space_charge_3d_open_hockney class the
// way the reference's own test does (test_space_charge_3d_open_hockney_mpi.cc:
// 54-68: build options, construct the operator, sc.apply(bsim, dt, logger)).
//
// usage: operator_driver in.bin out.bin gx gy gz real_num gamma step_length
//                        [green(0|1)] [periodic_z(0|1)]
// in.bin: int64 n, then 7*n doubles (column j contiguous), then n uint8 masks.
// out.bin: same layout after one apply, then the 38-double domain record.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "space_charge_3d_open_hockney.h"

int main(int argc, char** argv)
{
    if (argc < 9) {
        std::fprintf(stderr, "usage: %s in out gx gy gz real_num gamma step [green] [periodic]\n",
                     argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    int64_t n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1 || n <= 0) return 3;
    std::vector<double> parts(7 * size_t(n));
    std::vector<uint8_t> masks(n);
    if (std::fread(parts.data(), sizeof(double), parts.size(), f) != parts.size() ||
        std::fread(masks.data(), 1, masks.size(), f) != masks.size())
        return 3;
    std::fclose(f);
    const int gx = std::atoi(argv[3]), gy = std::atoi(argv[4]), gz = std::atoi(argv[5]);
    const double real_num = std::atof(argv[6]), gamma = std::atof(argv[7]);
    const double step_length = std::atof(argv[8]);
    const int green = argc > 9 ? std::atoi(argv[9]) : 1;
    const int periodic = argc > 10 ? std::atoi(argv[10]) : 0;

    auto ref = Reference_particle(pconstants::proton_charge, pconstants::mp, pconstants::mp * gamma);
    auto bsim = Bunch_simulator::create_single_bunch_simulator(ref, int(n), real_num);
    auto& bunch = bsim.get_bunch();
    std::memcpy(bunch.host_particles(), parts.data(), parts.size() * sizeof(double));
    std::memcpy(bunch.host_masks(), masks.data(), masks.size());
    bunch.checkin_particles();

    auto ops = Space_charge_3d_open_hockney_options(gx, gy, gz);
    ops.green_fn = green ? green_fn_t::linear : green_fn_t::pointlike;
    ops.periodic_z = periodic != 0;
    auto sc = Space_charge_3d_open_hockney(ops);
    auto logger = Logger(0, LoggerV::WARNING);
    const double time_step = step_length / (ref.get_beta() * pconstants::c);
    try {
        sc.apply(bsim, time_step, logger);
    } catch (std::runtime_error const& e) {
        std::printf("runtime_error: %s\n", e.what());
        return 10;
    }
    bunch.checkout_particles();
    sc3d_domain_record d = sc.domain();
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 4;
    std::fwrite(&n, sizeof(n), 1, o);
    std::fwrite(bunch.host_particles(), sizeof(double), 7 * size_t(n), o);
    std::fwrite(bunch.host_masks(), 1, size_t(n), o);
    std::fwrite(&d, sizeof(d), 1, o);
    std::fclose(o);
    std::printf("ok n=%lld factor=%.17g\n", (long long)n, d.factor);
    return 0;
}
