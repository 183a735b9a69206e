/* Extensions of the C API that exist only in the HIP engine.  None of these
 * has a counterpart in the reference; each comment says which callers use it. */
#pragma once
#ifndef SPIRIT_HIP_EXTENSIONS_H
#define SPIRIT_HIP_EXTENSIONS_H
#include "Spirit_Defines.h"

#include "Simulation.h"

/* For ensembles of independent replicas (uneven lattice sizes, per-replica fields and
 * temperatures): runs Simulation_LLG_Start(solver_type, n_iterations, n_iterations_log) on the
 * active image of each of the n_states states, all advanced together -- one launch per solver
 * stage for every replica.  Each replica keeps its own Hamiltonian, LLG parameters, convergence
 * test and iteration count; the results land in each state as after Simulation_LLG_Start.
 * info (nullable) receives n_states run summaries. */
SPIRIT_API void Spirit_HIP_Simulation_LLG_Start_Batch( State ** states, int n_states, int solver_type,
                                                      int n_iterations SPIRIT_DEFAULT( -1 ),
                                                      int n_iterations_log SPIRIT_DEFAULT( -1 ),
                                                      Simulation_Run_Info * info SPIRIT_DEFAULT( 0 ) ) SPIRIT_NOEXCEPT;

/* For benchmarks and long single-shot loops: mode 1 (default) refreshes the host
 * spin mirror after every Simulation_SingleShot / Simulation_N_Shot call; mode 0
 * keeps the spins on the GPU until a System_* / Quantity_* getter, Simulation_Stop
 * or the end of the run asks for them. */
SPIRIT_API void Spirit_HIP_Set_Host_Sync( State * state, int mode ) SPIRIT_NOEXCEPT;

/* For parity tests with a given noise sequence: the next n_steps LLG steps of the
 * image use these standard normals (AoS, [n_steps][nos][3]) for the thermal field
 * instead of the engine's counter-based Philox generator.  Call after
 * Simulation_LLG_Start(..., singleshot=true).  NULL / 0 switches back. */
SPIRIT_API void Spirit_HIP_Set_Thermal_Noise( State * state, const double * noise, int n_steps,
                                              int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* For disordered samples and parameter fitting: per-site uniaxial anisotropy K_i [meV]
 * (nos values, copied) along `normal`, added to the per-basis anisotropy.  NULL clears it. */
SPIRIT_API void Spirit_HIP_Hamiltonian_Set_Site_Anisotropy( State * state, const double * K, const float normal[3],
                                                            int idx_image SPIRIT_DEFAULT( -1 ),
                                                            int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* For parity tests and tools: the full gradient dE/ds (AoS, nos*3, as the reference's
 * Hamiltonian::Gradient_and_Energy) and energy of the image's current spins, evaluated
 * on the GPU.  Returns 0 on success. */
SPIRIT_API int Spirit_HIP_Gradient_and_Energy( State * state, double * gradient, double * energy,
                                               int idx_image SPIRIT_DEFAULT( -1 ),
                                               int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* For benchmarks: mean wall time [ms] of one step in the last Simulation_N_Shot /
 * run batch of the image's LLG method (0 if none). */
SPIRIT_API double Spirit_HIP_Get_Last_Step_Time( State * state, int idx_image SPIRIT_DEFAULT( -1 ),
                                                 int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* For benchmarks: time kernels of the image's LLG method with HIP events on its stream.
 * Kernel names: "ddi_conv" (DDI convolution pass), "llg_stage" (fused field + solver stages).
 * Profile_Get returns the summed milliseconds and the launch count since the last reset. */
SPIRIT_API void Spirit_HIP_Profile_Enable( State * state, bool enable, int idx_image SPIRIT_DEFAULT( -1 ),
                                           int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
SPIRIT_API void Spirit_HIP_Profile_Get( State * state, const char * kernel, double * total_ms, int * count,
                                        int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
SPIRIT_API void Spirit_HIP_Profile_Reset( State * state, int idx_image SPIRIT_DEFAULT( -1 ),
                                          int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* Diagnostics: describes the DDI plan of the image's running LLG method (padded grid,
 * storage format of the kernel spectrum) into buffer; returns 1 if the image has one. */
SPIRIT_API int Spirit_HIP_DDI_Info( State * state, char * buffer, int length, int idx_image SPIRIT_DEFAULT( -1 ),
                                    int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* Name of the GPU the engine runs on (empty if none is visible). */
SPIRIT_API const char * Spirit_HIP_Device_Name( void ) SPIRIT_NOEXCEPT;

#endif
