// Shared helpers of the C API translation units.
#pragma once

#include "../../include/Spirit/Chain.h"
#include "../../include/Spirit/Configurations.h"
#include "../../include/Spirit/Constants.h"
#include "../../include/Spirit/Geometry.h"
#include "../../include/Spirit/Hamiltonian.h"
#include "../../include/Spirit/Log.h"
#include "../../include/Spirit/Parameters_GNEB.h"
#include "../../include/Spirit/Parameters_LLG.h"
#include "../../include/Spirit/Quantities.h"
#include "../../include/Spirit/Simulation.h"
#include "../../include/Spirit/Spirit_HIP.h"
#include "../../include/Spirit/State.h"
#include "../../include/Spirit/System.h"
#include "../../include/Spirit/Transitions.h"
#include "../../include/Spirit/Version.h"
#include "state.hpp"

#include <stdexcept>
#include <string>

namespace spirit_hip
{

struct api_error : std::runtime_error
{
    int level;
    api_error( int lvl, const std::string & m ) : std::runtime_error( m ), level( lvl ) {}
};

void from_indices( State * state, int & idx_image, int & idx_chain, std::shared_ptr<Image> & image,
                   std::shared_ptr<Chain> & chain );
void handle_exception( int idx_image, int idx_chain );
std::shared_ptr<Image> make_image( const std::string & cfg );
void sync_image_host( State * state, int idx_image );
void sync_all_hosts( State * state );
void notify_image_changed( State * state, int idx_image, bool spins_changed, bool ham_changed, bool llg_changed );
bool stop_file_present();
// energy, effective field (store_field) and energy contributions of an image's host spins,
// evaluated on the GPU; per_spin (nullable) receives the six per-spin terms [6][nos]
double update_image_energy( Image & image, bool store_field, std::vector<double> * per_spin = nullptr );
std::vector<std::pair<std::string, int>> energy_contribution_list( const Image & image );
void llg_save_configuration( const Image & img, int idx_image, const std::string & starttime, long iteration,
                             long n_iterations, double max_torque, const std::string & solver_full_name,
                             bool initial, bool final );
// the energy files of an LLG save (and, spin-resolved, the per-spin energies, which update the
// image's energy contributions)
void llg_save_energy( Image & img, int idx_image, const std::string & starttime, long iteration, long n_iterations,
                      bool initial, bool final );
std::string current_date_time(); // "%Y-%m-%d_%H-%M-%S", local time

// GNEB (gneb_api.cpp)
void gneb_n_shot( State * state, int N );
void gneb_finalize( State * state );
void gneb_sync_host( State * state );
void gneb_notify_changed( State * state ); // GNEB parameters / image types changed on the host

} // namespace spirit_hip
