// space_charge_3d_open_hockney.cc -- the operator class over the C ABI.
// Reference: collective/space_charge_3d_open_hockney.cc:376-453.
#include "space_charge_3d_open_hockney.h"

#include <stdexcept>
#include <string>

namespace {
void check(int rc)
{
    if (rc != SC3D_OK) throw std::runtime_error(sc3d_last_error());
}
}  // namespace

Space_charge_3d_open_hockney::Space_charge_3d_open_hockney(
    Space_charge_3d_open_hockney_options const& ops)
    : Collective_operator("sc_3d_open_hockney", 1.0), options(ops), bunch_sim_id()
{
    sc3d_options_default(&copts, ops.shape[0], ops.shape[1], ops.shape[2]);
    for (int k = 0; k < 3; ++k) {
        copts.size[k] = ops.size[k];
        copts.offset[k] = ops.offset[k];
    }
    copts.green_fn = ops.green_fn == green_fn_t::linear ? SC3D_GREEN_LINEAR : SC3D_GREEN_POINTLIKE;
    copts.periodic_z = ops.periodic_z;
    copts.grid_entire_period = ops.grid_entire_period;
    copts.domain_fixed = ops.domain_fixed;
    copts.comm_group_size = ops.comm_group_size;
    copts.z_period = ops.z_period;
    copts.n_sigma = ops.n_sigma;
    copts.kick_scale = ops.kick_scale;
    copts.precision = ops.precision;
    copts.sort_period = ops.sort_period;
}

Space_charge_3d_open_hockney::~Space_charge_3d_open_hockney() { release(); }

void Space_charge_3d_open_hockney::release()
{
    for (auto& v : ctxs) {
        for (sc3d_ctx* c : v) sc3d_destroy(c);
        v.clear();
    }
}

// construct_workspaces (.cc:455-491): one device context per bunch.
void Space_charge_3d_open_hockney::construct_workspaces(Bunch_simulator const& sim)
{
    release();
    for (size_t t = 0; t < 2; ++t)
        for (size_t b = 0; b < sim[t].get_bunch_array_size(); ++b) {
            sc3d_ctx* c = nullptr;
            check(sc3d_create(&copts, &c));
            ctxs[t].push_back(c);
        }
}

// apply_impl (.cc:399-420)
void Space_charge_3d_open_hockney::apply_impl(Bunch_simulator& sim, double time_step,
                                              Logger& logger)
{
    logger << "    Space charge 3d open hockney\n";
    if (bunch_sim_id != sim.id()) {
        construct_workspaces(sim);
        bunch_sim_id = sim.id();
    }
    for (size_t t = 0; t < 2; ++t)
        for (size_t b = 0; b < sim[t].get_bunch_array_size(); ++b)
            apply_bunch(sim[t][b], ctxs[t][b], time_step, logger);
}

// apply_bunch (.cc:422-453); synchronous like the reference's fences.
void Space_charge_3d_open_hockney::apply_bunch(Bunch& bunch, sc3d_ctx* ctx, double time_step,
                                               Logger&)
{
    sc3d_bunch d{};
    d.parts = bunch.get_local_particles();
    d.stride = bunch.capacity();
    d.n_active = bunch.size();
    d.masks = bunch.get_local_particle_masks();
    d.total_num = bunch.get_total_num();
    d.real_num = bunch.get_real_num();
    d.charge = bunch.get_particle_charge();
    d.mass = bunch.get_mass();
    auto const& ref = bunch.get_reference_particle();
    d.ref_gamma = ref.get_gamma();
    d.ref_beta = ref.get_beta();
    d.ref_p = ref.get_momentum();
    check(sc3d_apply(ctx, &d, time_step, nullptr));
    check(sc3d_sync(ctx));
}

sc3d_domain_record Space_charge_3d_open_hockney::domain(size_t t, size_t b) const
{
    sc3d_domain_record r{};
    if (t < 2 && b < ctxs[t].size())
        check(sc3d_debug_get(ctxs[t][b], SC3D_DBG_DOMAIN, &r, sizeof(r)));
    return r;
}
