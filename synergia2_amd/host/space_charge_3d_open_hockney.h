// space_charge_3d_open_hockney.h -- drop-in for Synergia's
// Space_charge_3d_open_hockney (collective/space_charge_3d_open_hockney.h:
// 16-73): same constructor and options, same apply/apply_impl contract
// (synchronous, throws std::runtime_error with the reference's messages).
// The work runs on the current HIP device through the C ABI of
// include/sc3d_capi.h; one device context per bunch, rebuilt when the
// Bunch_simulator id changes (space_charge_3d_open_hockney.cc:409-412).
#pragma once

#include <vector>

#include "../../include/sc3d_capi.h"
#include "synergia_api.h"

class Space_charge_3d_open_hockney : public Collective_operator {
  public:
    explicit Space_charge_3d_open_hockney(Space_charge_3d_open_hockney_options const& ops);
    ~Space_charge_3d_open_hockney() override;

    Space_charge_3d_open_hockney(Space_charge_3d_open_hockney const&) = delete;
    Space_charge_3d_open_hockney& operator=(Space_charge_3d_open_hockney const&) = delete;

    // the last domain record of bunch (t, b), for diagnostics and tests
    sc3d_domain_record domain(size_t t = 0, size_t b = 0) const;

  private:
    void apply_impl(Bunch_simulator& simulator, double time_step, Logger& logger) override;
    void apply_bunch(Bunch& bunch, sc3d_ctx* ctx, double time_step, Logger& logger);
    void construct_workspaces(Bunch_simulator const& sim);
    void release();

    const Space_charge_3d_open_hockney_options options;
    sc3d_options copts;
    std::string bunch_sim_id;
    std::vector<sc3d_ctx*> ctxs[2];
};
