// Device-resident runners: DDI plan, batched LLG runner and GNEB chain runner.
#pragma once

#include "ddi.hpp"
#include "ddi_rs.hpp"
#include "gneb.hpp"
#include "host_core.hpp"
#include "llg.hpp"
#include "minimize.hpp"

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace spirit_hip
{

// one field of n vectors: AoS [x0,y0,z0,x1,...] <-> SoA [x...][y...][z...]
void aos_to_soa( const double * aos, double * soa, long long n );
void soa_to_aos( const double * soa, double * aos, long long n );
// the same conversions of n spins in device memory, on stream st
void device_aos_to_soa( const double * aos, double * soa, long long n, hipStream_t st );
void device_soa_to_aos( const double * soa, double * aos, long long n, hipStream_t st );

void hip_check( hipError_t e, const char * what );
// division by an invariant d without hardware division (device_types.hpp FastDiv)
FastDiv make_fast_div( int d );
#define HIPCHECK( x ) ::spirit_hip::hip_check( ( x ), #x )

// Owning device allocation.
class DeviceBuffer
{
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer( size_t bytes )
    {
        resize( bytes );
    }
    ~DeviceBuffer();
    DeviceBuffer( const DeviceBuffer & )             = delete;
    DeviceBuffer & operator=( const DeviceBuffer & ) = delete;
    void resize( size_t bytes );
    void free();
    template<typename T>
    T * as() const
    {
        return static_cast<T *>( p );
    }
    size_t size() const
    {
        return bytes;
    }

private:
    void * p     = nullptr;
    size_t bytes = 0;
};

// Radix plan + twiddle table of one FFT length.
struct FftLenHost
{
    FftLen dev{};
    DeviceBuffer W, TW2;
    void build( int n );
};

// One zero-padded FFT convolution plan (one image geometry).
class DdiPlan
{
public:
    // throws std::runtime_error for unsupported grids
    void build( const Geometry & g, const Hamiltonian & h, hipStream_t st );
    void field( const double * spins_soa, double * h_soa, int nos, hipStream_t st, hipEvent_t evc0 = nullptr,
                hipEvent_t evc1 = nullptr );
    const DdiDev & dev() const
    {
        return d;
    }
    size_t bytes() const;
    // how the kernel spectrum is stored (real / mirror-compressed / complex)
    const char * spectrum_format() const;
    const DhatSymmetry & symmetry() const
    {
        return sym;
    }

    // milliseconds of the chosen pass C kernel when the plan was built (0 if there was no choice)
    double pass_c_ms() const
    {
        return c_ms;
    }
    // the real-space methods (ddi_method cutoff: pairs within the radius, or the direct sum
    // for a negative radius) instead of the FFT convolution
    bool real_space() const
    {
        return rs.method != 0;
    }
    const DdiRsDev & rs_dev() const
    {
        return rs;
    }

private:
    void build_real_space( const Geometry & g, const Hamiltonian & h );
    void choose_pass_c( hipStream_t st );
    DdiRsDev rs{};
    DeviceBuffer b_pos, b_pairs;
    DdiDev d{};
    DhatSymmetry sym{};
    double c_ms = 0;
    FftLenHost La, Lb, Lc;
    DeviceBuffer X, Dh;
};

// Padded DDI grid, Hamiltonian_Heisenberg.cpp:1501-1530 (without the kissFFT-only length-1 workaround).
std::array<int, 3> ddi_padded_dims( const Geometry & g, const Hamiltonian & h );

// The POD image description a kernel reads, built on the host: H gets every field but the
// device pointers (exch, dmi, site_K are null; has_site_K says whether the Hamiltonian has
// per-site K_i, from h.site_K); returns the pair table (n_exch exchange pairs, then n_dmi DMI
// pairs) that H.exch / H.dmi must point to on the device.
std::vector<PairDev> build_ham_dev( const Geometry & g, const Hamiltonian & h, HamDev & H );
// The same, with the pair table written to pv (at most capacity entries; more throws) instead
// of a new vector; returns the number of pairs.
int fill_ham_dev( const Geometry & g, const Hamiltonian & h, HamDev & H, PairDev * pv, int capacity );

// Builds the POD image description a kernel reads (pair tables are uploaded to `pairs`).
struct HamUpload
{
    HamDev dev{};
    DeviceBuffer pairs, quads;
    DeviceBuffer site_K;
    uint64_t version     = ~0ull;
    uint64_t ddi_version = ~0ull;
    void upload( const Geometry & g, const Hamiltonian & h );
};

struct ImageSpec
{
    std::shared_ptr<Geometry> geometry;
    std::shared_ptr<Hamiltonian> hamiltonian;
    LLGParameters llg;
    double * spins_aos = nullptr; // host mirror (nos*3)
};

// Runs the LLG step loop of one or several images (a batched ensemble) on the GPU.
class LLGRunner
{
public:
    LLGRunner( const std::vector<ImageSpec> & images, int solver );
    ~LLGRunner();

    void upload_spins();   // host AoS -> device
    void download_spins(); // device -> host AoS
    // Update per-image LLG parameters (dt, damping, temperature, ...) without reallocation.
    void set_llg( int img, const LLGParameters & p );
    void refresh_hamiltonian(); // re-upload after parameter changes

    // Enqueue n steps.  observe_every > 0 computes energy/torque every that many steps
    // (always on the last step); with check_convergence an image stops on the device once its
    // max torque drops below force_convergence (the others continue).  Returns the largest
    // number of steps any image performed; steps_last_run(img) gives each image's.
    long run( long n, int observe_every, bool check_convergence, double force_convergence );
    long steps_last_run( int img ) const
    {
        return steps_last[img];
    }
    // an image that has stopped (converged, or ended by its caller) takes no further steps
    bool image_stopped( int img ) const
    {
        return img_stopped[img] != 0;
    }
    void set_image_stopped( int img, bool stopped )
    {
        img_stopped[img] = stopped ? 1 : 0;
    }
    // Energy, max torque and effective field of the current spins (no step taken).
    void initial_observe();

    double energy( int img ) const
    {
        return obs_host[2 * img];
    }
    double max_torque( int img ) const
    {
        return obs_host[2 * img + 1];
    }
    void download_effective_field( int img, double * aos );
    // the fields of several images at once (aos[k]: image k's AoS destination or nullptr)
    void download_effective_fields( const std::vector<double *> & aos );
    // gradient (AoS) and energy of the current device spins of image img
    double gradient( int img, double * grad_aos );
    // per-spin energy contributions of image img's current spins, out[t * nos + i] for the
    // N_ENERGY_TERMS_HOST terms Zeeman, anisotropy, cubic anisotropy, exchange, DMI, DDI, quadruplets
    void energy_terms( int img, double * out );
    // injected N(0,1) noise for the next `n_steps` steps, AoS per spin over all images
    void set_injected_noise( const double * noise, long n_steps );

    int n_images() const
    {
        return (int)specs.size();
    }
    long long total_spins() const
    {
        return n_total;
    }
    uint64_t step_counter = 0;
    hipStream_t stream = nullptr;
    double last_step_ms = 0;

    // raw device pointers (for the adjoint module and tests)
    double * dev_spins()
    {
        return b_s.as<double>();
    }

    // Kernel timing with HIP events on the runner's stream: "ddi_conv" (the convolution
    // pass of the DDI field) and "llg_stage" (the fused field + solver stage kernels).
    struct ProfileEntry
    {
        double ms  = 0;
        long count = 0;
    };
    void set_profiling( bool on );
    ProfileEntry profile( const std::string & name ) const;
    void reset_profile();

    // DDI plan of image img (nullptr without DDI)
    const DdiPlan * ddi_plan( int img ) const
    {
        return img >= 0 && img < (int)ddi.size() ? ddi[img].get() : nullptr;
    }

    // For methods built on this runner's images (GNEB): the device image table, the DDI
    // field buffer (nullptr without DDI) and the DDI field of a batched configuration x.
    const ImageDev * dev_images() const
    {
        return b_imgs.as<ImageDev>();
    }
    // host copy of image k's device description (its HamDev holds device pointers)
    const ImageDev & image_host( int k ) const
    {
        return imgs_host[k];
    }
    double * dev_ddi_field()
    {
        return any_ddi ? b_h.as<double>() : nullptr;
    }
    void compute_ddi_field( const double * x )
    {
        if( any_ddi )
            ddi_field( x );
    }
    // One Depondt / Heun step of a single-image runner from s_in to s_out with step index
    // `step`, recording the forces and virtual forces of both evaluations (adjoint tape).
    // The arithmetic is that of run(): same kernels, same thermal-noise counters.
    void tape_step( uint64_t step, double * s_in, double * s_out, double * F1, double * fv1, double * F2,
                    double * fv2 );

private:
    // the AoS device stage of the images with a host destination (nullptr: skipped) to those arrays
    void stage_to_host( const std::vector<double *> & dst );
    hipEvent_t next_event();
    void collect_profile();
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<std::pair<std::string, ProfileEntry>> prof;

    void step( bool obs, bool count, bool check, double conv );
    // the DDI fields and solver stages of one step (no observation)
    void launch_step( const StepArgs & A );
    // Graph of `len` steps reading the step index from b_step; with period > 0 every
    // period-th step is observed (finalize with check / conv), with period 0 none is.
    bool capture_step_graph( int len, int period, bool check, double conv );
    void release_step_graph();
    static constexpr int GRAPH_STEPS     = 16; // steps per graph (at least)
    static constexpr int GRAPH_MAX_STEPS = 64; // longest observation period captured whole
    hipGraphExec_t step_graph = nullptr;
    int graph_len = 0, graph_period = -1, graph_check = -1;
    double graph_conv = 0;
    bool dev_step = false; // steps take their index from b_step (and advance it)
    void ddi_field( const double * x, bool prof_events = false );
    void upload_images();

    std::vector<ImageSpec> specs;
    int solver;
    long long n_total = 0;
    int n_blocks      = 0;
    std::vector<ImageDev> imgs_host;
    std::vector<int> blk_img_host;
    std::vector<std::unique_ptr<HamUpload>> hams;
    std::vector<std::unique_ptr<DdiPlan>> ddi;
    DeviceBuffer b_imgs, b_blk, b_s, b_sp, b_sp2, b_aux, b_aux2, b_h, b_eff, b_efft, b_pE, b_pT, b_stop, b_obs,
        b_noise, b_stopstep, b_img_stop, b_nstop, b_step, b_xi;
    void alloc_thermal_field();
    // image k's LLGDev from its parameters, with its per-site temperatures (gradient) uploaded
    void apply_llg( size_t k );
    std::vector<std::unique_ptr<DeviceBuffer>> b_temp; // per image: T_i of a temperature gradient
    DeviceBuffer b_stage; // the images' AoS spins on their way to or from the host mirrors
    int stt_gradient_any = 0;                          // some image's STT uses the spin derivative
    void tune_stage_tiles();
    int tuned_row_tile = -1; // row_tile_log2 picked by tune_stage_tiles (-1: the default)
    std::vector<double> obs_host;
    std::vector<int> img_stopped; // host view of the per-image stop flags
    std::vector<long> steps_last; // steps each image advanced in the last run
    int stop_host = 0, nstop_host = 0;
    bool any_ddi = false;
    // b_eff holds the unprojected force F1 of the last observed step (LLG stages and the
    // initial observation), projected onto the spins when downloaded; the direct-minimisation
    // kernels store the projected field
    bool eff_raw = true;
    long noise_steps = 0, noise_pos = 0;
    // direct-minimisation solvers (VP, VP_OSO, LBFGS_OSO, LBFGS_Atlas): state of the run
    void setup_minimiser();
    DeviceBuffer b_mv, b_mprev, b_mf, b_mgrad, b_mgpr, b_msd, b_mq, b_ma3, b_mpart, b_msc;
    DeviceBuffer b_mda[MIN_MEMORY], b_mdg[MIN_MEMORY];
    MinArgs min_args{};
};

// Differentiable LLG run of one image (Depondt or Heun, fixed number of steps): forward with
// checkpoints, the discrete adjoint (dL/ds0 and dL/d parameters from dL/ds_N) by recomputing
// each checkpoint segment with a tape, and a tangent-linear (forward-mode) run.  Spin fields
// crossing this interface are device SoA arrays [x...][y...][z...].
class AdjointRunner
{
public:
    AdjointRunner( const Geometry & g, const Hamiltonian & h, const LLGParameters & llg, int solver, int n_steps,
                   int checkpoint_every );
    ~AdjointRunner();
    int n_exchange_shells() const
    {
        return n_exch_shells;
    }
    int n_dmi_shells() const
    {
        return n_dmi_sh;
    }
    int n_aniso() const
    {
        return (int)ham->aniso_idx.size();
    }
    int nos() const
    {
        return geom->nos;
    }
    // parameter values (nullptr keeps): J per exchange shell, D per DMI shell, B [T], K per basis
    // anisotropy entry, K_i per site (along the per-site axis)
    void set_parameters( const double * J, const double * D, const double * B, const double * K,
                         const double * K_site );
    void get_parameters( double * J, double * D, double * B, double * K, double * K_site ) const;
    void forward( const double * s0, double * sN );
    // dL/ds_N -> dL/ds_0 and parameter gradients (host arrays; dK_site is a device array or nullptr)
    void backward( const double * dL_dsN, double * dL_ds0, double * dJ, double * dD, double * dB, double * dK,
                   double * dK_site );
    // tangent-linear run along (ds0, dJ, dD, dB, dK, dK_site): sN and dsN
    void tangent( const double * s0, const double * ds0, const double * dJ, const double * dD, const double * dB,
                  const double * dK, const double * dK_site, double * sN, double * dsN );
    hipStream_t stream() const
    {
        return run->stream;
    }

private:
    void upload_variants();
    std::shared_ptr<Geometry> geom;
    std::shared_ptr<Hamiltonian> ham; // private copy: the parameters being differentiated
    LLGParameters llg;
    int solver, n_steps, ckpt;
    int n_exch_shells = 0, n_dmi_sh = 0;
    std::vector<double> host_spins;
    std::unique_ptr<LLGRunner> run;
    HamUpload hess_up, delta_up;
    DeviceBuffer b_ckpt, b_tape_s, b_F1, b_fv1, b_F2, b_fv2, b_sp, b_lam, b_ms, b_msp, b_mfv1, b_mg, b_pg, b_hd,
        b_dsp, b_ds, b_dsn, b_site_zero;
    bool tape_ok = false;
};

// Runs Method_GNEB (core/src/engine/Method_GNEB.cpp) on a chain of images with equal spin
// counts: every force evaluation, tangent, spring and solver update of all images runs on
// the GPU, with the chain-level scalars kept on the device.
class GNEBRunner
{
public:
    GNEBRunner( const std::vector<ImageSpec> & images, const GNEBParameters & par, const std::vector<int> & types,
                int solver );
    ~GNEBRunner();

    // Enqueue up to n iterations (Method::Iterate with n_iterations_amortize = 1); with
    // check_convergence the run stops on the device once the chain's max torque is below the
    // force convergence.  Returns the number of iterations done.
    long run( long n, bool check_convergence );
    // 0 running, 1 converged, 2 stopped because two neighbouring images coincide
    int status();

    struct ChainData
    {
        std::vector<double> E, Rx, dE_dRx, torque;
        double max_torque = 0;
    };
    // energies, reaction coordinates, slopes and torques of the last force evaluation
    ChainData chain_data();
    void download_spins();                      // to the images' host mirrors
    void download_effective_field( int img, double * aos );
    // the fields of several images at once (aos[k]: image k's AoS destination or nullptr)
    void download_effective_fields( const std::vector<double *> & aos );
    void upload_spins();
    void set_types( const std::vector<int> & types );
    void set_parameters( const GNEBParameters & par );
    double * dev_spins()
    {
        return field->dev_spins();
    }
    hipStream_t stream() const
    {
        return field->stream;
    }

private:
    void calc_force( double * x, bool vp );
    void iteration();
    std::unique_ptr<LLGRunner> field; // images, Hamiltonians, DDI plans, the chain's spins
    std::vector<ImageSpec> specs;
    int solver = 0, noi = 0, nos = 0, bpi = 0;
    GnebArgs A{};
    DeviceBuffer b_eff, b_Fg, b_tau, b_Ft, b_ftr, b_fv, b_fv0, b_v, b_fprev, b_pred, b_tmp, b_part, b_ch, b_types,
        b_stop, b_iter, b_count, b_isum;
    // VP_OSO, LBFGS_OSO, LBFGS_Atlas on the chain's forces F_total
    void setup_minimiser();
    DeviceBuffer b_mv, b_mprev, b_mgrad, b_mgpr, b_msd, b_mq, b_ma3, b_mpart, b_msc;
    DeviceBuffer b_mda[MIN_MEMORY], b_mdg[MIN_MEMORY];
    MinArgs min_args{};
};

} // namespace spirit_hip
