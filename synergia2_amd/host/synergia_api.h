// synergia_api.h -- the slice of Synergia's C++ interface that the 3D
// open-Hockney operator sees, restated for a device-resident bunch:
//
//   green_fn_t, Space_charge_3d_open_hockney_options
//       <- simulation/implemented_collective_options.h:67-126
//   Reference_particle (charge, mass, gamma, beta, momentum)
//       <- foundation/reference_particle.cc:273-287, four_momentum.cc:9-17
//   Bunch (7-column LayoutLeft particles + u8 masks, counts, accessors)
//       <- bunch/bunch.h:66-74,267-289,317,458-472, bunch_particles.h:24-43
//   Bunch_simulator (two trains of bunches, id())
//       <- simulation/bunch_simulator.h:186-189, bunch_simulator.cc:132-147
//   Logger <- utils/logger.h:10-86 (levels only)
//   Operator / Collective_operator
//       <- simulation/operator.h:15-77, collective_operator.h:7-20
//
// Particles live in device memory (hipMalloc); host mirrors exist for
// checkout/checkin as in the reference.  Nothing here is Kokkos.
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pconstants {
// physical_constants.h:17-24 (PDG 2012), :94-106
constexpr double mp = 0.938272046;
constexpr double proton_mass = 0.938272046;
constexpr double e = 1.602176565e-19;
constexpr double c = 299792458.0;
constexpr int proton_charge = 1;
}  // namespace pconstants

enum class green_fn_t { pointlike = 0, linear = 1 };

struct Space_charge_3d_open_hockney_options {
    std::array<int, 3> shape;
    std::array<int, 3> doubled_shape;
    std::array<double, 3> size;
    std::array<double, 3> offset;
    green_fn_t green_fn;
    bool periodic_z;
    double z_period;
    bool grid_entire_period;
    double n_sigma;
    double kick_scale;
    bool domain_fixed;
    int comm_group_size;
    // device-side extensions
    int precision = 0;    // 0: FP64, 1: FP32 grid / FP64 particles
    int sort_period = 24;  // reorder the bunch storage by cell every N applies

    Space_charge_3d_open_hockney_options(int gridx = 32, int gridy = 32, int gridz = 64)
        : shape{gridx, gridy, gridz}
        , doubled_shape{gridx * 2, gridy * 2, gridz * 2}
        , size{0.0, 0.0, 0.0}
        , offset{0.0, 0.0, 0.0}
        , green_fn(green_fn_t::linear)
        , periodic_z(false)
        , z_period(0.0)
        , grid_entire_period(false)
        , n_sigma(8.0)
        , kick_scale(1.0)
        , domain_fixed(false)
        , comm_group_size(4)
    {}

    // implemented_collective_options.h:97-109: the grid_entire_period branch
    // is overwritten by the two assignments after it.
    void set_fixed_domain(std::array<double, 3> off, std::array<double, 3> sz)
    {
        if (grid_entire_period) {
            this->offset[2] = 0.0;
            this->size[2] = z_period;
        }
        this->offset = off;
        this->size = sz;
        this->domain_fixed = true;
    }
};

enum class LoggerV { DEBUG, DINFO, WARNING, INFO, INFO_OPN, INFO_OPR, INFO_STEP, INFO_TURN };

class Logger {
  public:
    explicit Logger(int rank = 0, LoggerV level = LoggerV::INFO) : rank_(rank), level_(level) {}
    template <typename T>
    Logger& operator<<(T const& v)
    {
        if (rank_ == 0 && level_ <= LoggerV::INFO) std::cout << v;
        return *this;
    }
    Logger& operator()(LoggerV) { return *this; }

  private:
    int rank_;
    LoggerV level_;
};

class Reference_particle {
  public:
    Reference_particle(int charge, double mass, double total_energy)
        : charge_(charge), mass_(mass), gamma_(total_energy / mass)
    {
        if (gamma_ < 1.0) throw std::range_error("Four_momentum: gamma not >= 1.0");
        beta_ = std::sqrt(1.0 - 1.0 / (gamma_ * gamma_));
        momentum_ = gamma_ * beta_ * mass_;
    }
    int get_charge() const { return charge_; }
    double get_mass() const { return mass_; }
    double get_gamma() const { return gamma_; }
    double get_beta() const { return beta_; }
    double get_momentum() const { return momentum_; }

  private:
    int charge_;
    double mass_, gamma_, beta_, momentum_;
};

// A single-rank bunch with its particles resident on the current HIP device.
class Bunch {
  public:
    static constexpr int x = 0, xp = 1, y = 2, yp = 3, z = 4, zp = 5, cdt = 4, dpop = 5, id = 6;

    Bunch(Reference_particle const& ref, int total_num, double real_num)
        : ref_(ref), n_active_(total_num), capacity_(total_num), total_num_(total_num),
          real_num_(real_num)
    {
        size_t cap = size_t(capacity_ > 0 ? capacity_ : 1);
        if (hipMalloc(&parts_, 7 * cap * sizeof(double)) != hipSuccess ||
            hipMalloc(&masks_, cap) != hipSuccess)
            throw std::runtime_error("Bunch: device allocation failed");
        hparts_.assign(7 * cap, 0.0);
        hmasks_.assign(cap, 1);
        for (int i = 0; i < n_active_; ++i) hparts_[size_t(id) * cap + i] = i;
        checkin_particles();
    }
    Bunch(Bunch const&) = delete;
    Bunch& operator=(Bunch const&) = delete;
    ~Bunch()
    {
        (void)hipFree(parts_);
        (void)hipFree(masks_);
    }

    // host mirror, column j at host_particles() + j*capacity()
    double* host_particles() { return hparts_.data(); }
    uint8_t* host_masks() { return hmasks_.data(); }
    void checkin_particles()
    {
        if (hipMemcpy(parts_, hparts_.data(), hparts_.size() * sizeof(double),
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(masks_, hmasks_.data(), hmasks_.size(), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("Bunch::checkin_particles failed");
    }
    void checkout_particles()
    {
        if (hipMemcpy(hparts_.data(), parts_, hparts_.size() * sizeof(double),
                      hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hmasks_.data(), masks_, hmasks_.size(), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("Bunch::checkout_particles failed");
    }

    double* get_local_particles() { return parts_; }  // device
    uint8_t* get_local_particle_masks() { return masks_; }
    int size() const { return n_active_; }
    int capacity() const { return capacity_; }
    int get_total_num() const { return total_num_; }
    double get_real_num() const { return real_num_; }
    int get_particle_charge() const { return ref_.get_charge(); }
    double get_mass() const { return ref_.get_mass(); }
    Reference_particle const& get_reference_particle() const { return ref_; }

  private:
    Reference_particle ref_;
    int n_active_, capacity_, total_num_;
    double real_num_;
    double* parts_ = nullptr;
    uint8_t* masks_ = nullptr;
    std::vector<double> hparts_;
    std::vector<uint8_t> hmasks_;
};

class Bunch_train {
  public:
    size_t get_bunch_array_size() const { return bunches_.size(); }
    Bunch& operator[](size_t i) { return *bunches_[i]; }
    Bunch const& operator[](size_t i) const { return *bunches_[i]; }
    void add(std::unique_ptr<Bunch> b) { bunches_.push_back(std::move(b)); }

  private:
    std::vector<std::unique_ptr<Bunch>> bunches_;
};

class Bunch_simulator {
  public:
    // bunch_simulator.cc:132-147
    static Bunch_simulator create_single_bunch_simulator(Reference_particle const& ref,
                                                         int num_particles, double real_num)
    {
        Bunch_simulator s;
        s.trains_[0].add(std::make_unique<Bunch>(ref, num_particles, real_num));
        return s;
    }
    Bunch_train& operator[](size_t t) { return trains_[t]; }
    Bunch_train const& operator[](size_t t) const { return trains_[t]; }
    Bunch& get_bunch(size_t t = 0, size_t b = 0) { return trains_[t][b]; }
    std::string const& id() const { return id_; }  // bunch_simulator.h:186-189

  private:
    Bunch_simulator() : id_("bsim-" + std::to_string(next_id_++)) {}
    static inline std::atomic<long> next_id_{0};
    Bunch_train trains_[2];
    std::string id_;
};

class Operator {
  public:
    Operator(std::string const& name, std::string const& type, double time_fraction)
        : name_(name), type_(type), time_fraction_(time_fraction)
    {}
    virtual ~Operator() = default;
    // operator.h:62-65
    void apply(Bunch_simulator& simulator, double time_step, Logger& logger)
    {
        apply_impl(simulator, time_step * time_fraction_, logger);
    }
    std::string const& get_name() const { return name_; }
    std::string const& get_type() const { return type_; }
    double get_time_fraction() const { return time_fraction_; }

  private:
    virtual void apply_impl(Bunch_simulator& simulator, double time_step, Logger& logger) = 0;
    std::string name_, type_;
    double time_fraction_;
};

class Collective_operator : public Operator {
  public:
    Collective_operator(std::string const& name, double time_fraction)
        : Operator(name, "collective", time_fraction)
    {}
};
