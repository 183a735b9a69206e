// The opaque State behind the C API: a chain of images (spin systems) plus the
// methods currently running on them.  Mirrors the reference's data model
// (core/include/data/State.hpp:15-49, Spin_System.hpp:26-83, Spin_System_Chain.hpp).
#pragma once

#include "gpu_engine.hpp"
#include "host_core.hpp"
#include "mc.hpp"

#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

namespace spirit_hip
{

struct Image
{
    std::shared_ptr<Geometry> geometry;
    std::shared_ptr<Hamiltonian> hamiltonian;
    std::shared_ptr<LLGParameters> llg;
    std::shared_ptr<MCParameters> mc;
    std::mt19937 prng; // seeded with llg.seed, used by Configuration_Random (Parameters_Method_LLG.hpp:22-24)
    std::vector<double> spins;     // AoS host mirror, handed out by System_Get_Spin_Directions
    std::vector<double> eff_field; // AoS
    double E          = 0;
    std::vector<std::pair<std::string, double>> E_array; // energy contributions (System_Update_Data)
    double max_torque = 0;
    double Rx         = 0;
    // written by Simulation_Stop from any thread while a run polls them
    std::atomic<bool> iteration_allowed{ false };
    std::atomic<bool> singleshot_allowed{ false };
    std::recursive_mutex mtx;

    Image() = default;
    Image( const Image & o );
    int nos() const
    {
        return geometry->nos;
    }
};

enum GNEBImageType
{
    GNEB_NORMAL     = 0,
    GNEB_CLIMBING   = 1,
    GNEB_FALLING    = 2,
    GNEB_STATIONARY = 3
};

struct Chain
{
    std::vector<std::shared_ptr<Image>> images;
    std::vector<int> image_type;
    std::shared_ptr<GNEBParameters> gneb;
    std::vector<double> Rx, Rx_interpolated, E_interpolated;
    std::atomic<bool> iteration_allowed{ false };
    std::atomic<bool> singleshot_allowed{ false };
    std::recursive_mutex mtx;
    int noi() const
    {
        return (int)images.size();
    }
};

// A running LLG method (one image), core/src/engine/Method_LLG.cpp
struct MethodLLG
{
    std::shared_ptr<Image> image;
    std::unique_ptr<LLGRunner> runner;
    int solver             = 2;
    long iteration         = 0;
    long n_iterations      = 0;
    long n_iterations_log  = 0;
    long amortize          = 1;
    double picoseconds     = 0;
    double ips             = 0;
    bool device_newer      = false; // device spins newer than the host mirror
    std::chrono::steady_clock::time_point t_start;
    std::string solver_name;
    std::string starttime; // date and time of the start (output file tags)
    int idx_image = 0;
    // appended at every Save_Current (Method_LLG.cpp:313-315), handed out by Simulation_Run_Info
    std::vector<int> history_iteration;
    std::vector<float> history_max_torque, history_energy;
};

// A running Monte Carlo method (one image), core/src/engine/Method_MC.cpp: the image's device
// spins and Hamiltonian record (held by an LLG runner), the adaptive cone, the rejection
// partials and the step counter of the Metropolis kernels.
struct MethodMC
{
    std::shared_ptr<Image> image;
    std::unique_ptr<LLGRunner> runner;
    DeviceBuffer cone, part, step;
    McArgs A{};
    double cone0       = 0;
    long iteration     = 0;
    long n_iterations  = 0;
    double ips         = 0;
    std::chrono::steady_clock::time_point t_start;
};

// A running GNEB method on the chain, core/src/engine/Method_GNEB.cpp
struct MethodGNEB
{
    std::unique_ptr<GNEBRunner> runner;
    int solver            = 0;
    long iteration        = 0;
    long n_iterations     = 0;
    long n_iterations_log = 0;
    double ips            = 0;
    std::vector<double> max_torque_all;
    double max_torque  = 0;
    bool device_newer  = false; // device spins newer than the host mirrors
    std::chrono::steady_clock::time_point t_start;
    std::string solver_name;
};

} // namespace spirit_hip

struct State
{
    std::shared_ptr<spirit_hip::Chain> chain;
    std::shared_ptr<spirit_hip::Image> active_image;
    std::shared_ptr<spirit_hip::Image> clipboard_image;
    std::shared_ptr<std::vector<double>> clipboard_spins;
    int nos              = 0;
    int noi              = 0;
    int idx_active_image = 0;
    std::map<int, std::shared_ptr<spirit_hip::MethodLLG>> method_image;
    std::map<int, std::shared_ptr<spirit_hip::MethodMC>> method_mc; // single-shot Monte Carlo
    std::shared_ptr<spirit_hip::MethodGNEB> method_chain;
    std::string datetime;
    std::string config_file;
    bool quiet     = false;
    int host_sync  = 1; // 1: host spin mirror refreshed after every single-shot batch
};
