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

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * gradient and energy as Spirit_HIP_Gradient_and_Energy, of spins the caller already holds
 * on the current device.  spins and gradient are device arrays of 3*nos doubles (AoS, as
 * System_Get_Spin_Directions), energy a device array of one double; stream is a hipStream_t
 * (NULL: the default stream).  The call allocates nothing: its tables, partial sums and
 * structure-of-arrays copies live in a fixed device array of the library (about 57 MiB per
 * GPU, held until the process ends), shared by every call on that GPU.  Limits: at most
 * SPIRIT_HIP_NOALLOC_MAX_SPINS spins, 128 exchange + DMI pairs, 512 quadruplets and no
 * dipolar interactions (ddi_method none); beyond them it logs an error and returns 1 without
 * touching the outputs.  Calls through this entry point hold a per-device lock until their
 * stream has finished, so they never overlap on one GPU.  Returns 0 on success. */
#define SPIRIT_HIP_NOALLOC_MAX_SPINS 1048576
SPIRIT_API int Spirit_HIP_Gradient_and_Energy_NoAlloc( State * state, const double * spins, double * gradient,
                                                      double * energy, void * stream,
                                                      int idx_image SPIRIT_DEFAULT( -1 ),
                                                      int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* The dense 3N x 3N embedding Hessian of the image's Hamiltonian (row-major, index
 * 3*ispin + component), evaluated on the GPU: the matrix Hamiltonian_Heisenberg::Hessian
 * fills (Hamiltonian_Heisenberg.cpp:1119-1245; EMA, MMF and HTST read it), including its
 * anisotropy row sums on the diagonal and the dipolar block over the positions without
 * periodic images.  hessian holds 9*nos*nos doubles.  Returns 0 on success. */
SPIRIT_API int Spirit_HIP_Hamiltonian_Get_Hessian( State * state, double * hessian, int idx_image SPIRIT_DEFAULT( -1 ),
                                                   int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* out = H v (AoS, 3*nos) for the same matrix, evaluated on the GPU without forming it.
 * Returns 0 on success. */
SPIRIT_API int Spirit_HIP_Hamiltonian_Hessian_Vector_Product( State * state, const double * v, double * out,
                                                              int idx_image SPIRIT_DEFAULT( -1 ),
                                                              int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

/* Per-spin energy contributions of the image's current spins (the reference's
 * Hamiltonian_Heisenberg::Energy_Contributions_per_Spin, Hamiltonian_Heisenberg.cpp:262-305,
 * which its C API reaches only through IO_Image_Write_Energy_per_Spin), evaluated on the GPU:
 * out[k * nos + i] for the contributions System_Get_Energy_Array_Names lists.  Also updates
 * the image's energy and System_Get_Energy_Array.  Returns the number of contributions
 * (out may be NULL to query it). */
SPIRIT_API int Spirit_HIP_Get_Energy_Contributions_per_Spin( State * state, double * out,
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

/* For parameter fitting and inverse design: a differentiable LLG run of one image (the
 * image's geometry, Hamiltonian and LLG parameters are copied at creation).  solver_type is
 * Solver_Depondt or Solver_Heun; n_steps steps of LLG dynamics (T >= 0, thermal noise from the
 * engine's counter-based generator); checkpoints every checkpoint_every steps bound the memory
 * of the backward pass, which recomputes each segment.
 * Parameters: J per exchange shell, D per DMI shell, B [T] (field magnitude), K per uniaxial
 * anisotropy entry, K_site per spin (along the per-site anisotropy axis, default z).
 * Spin arrays are AoS (nos*3); with on_device = 1 they are device pointers, else host pointers.
 * Parameter arrays and their gradients are host arrays; any may be NULL.  Functions returning
 * int return 0 on success (errors are logged). */
typedef struct Spirit_HIP_Adjoint Spirit_HIP_Adjoint;
SPIRIT_API Spirit_HIP_Adjoint * Spirit_HIP_Adjoint_Create( State * state, int solver_type, int n_steps,
                                                           int checkpoint_every, int idx_image SPIRIT_DEFAULT( -1 ),
                                                           int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
SPIRIT_API void Spirit_HIP_Adjoint_Destroy( Spirit_HIP_Adjoint * handle ) SPIRIT_NOEXCEPT;
SPIRIT_API int Spirit_HIP_Adjoint_Get_Sizes( Spirit_HIP_Adjoint * handle, int * nos, int * n_exchange_shells,
                                             int * n_dmi_shells, int * n_anisotropy ) SPIRIT_NOEXCEPT;
SPIRIT_API int Spirit_HIP_Adjoint_Set_Parameters( Spirit_HIP_Adjoint * handle, const double * J, const double * D,
                                                  const double * B, const double * K,
                                                  const double * K_site ) SPIRIT_NOEXCEPT;
SPIRIT_API int Spirit_HIP_Adjoint_Get_Parameters( Spirit_HIP_Adjoint * handle, double * J, double * D, double * B,
                                                  double * K, double * K_site ) SPIRIT_NOEXCEPT;
/* s_N of the run from s_0 (and the checkpoints the backward pass needs). */
SPIRIT_API int Spirit_HIP_Adjoint_Forward( Spirit_HIP_Adjoint * handle, const double * s0, double * sN,
                                           int on_device ) SPIRIT_NOEXCEPT;
/* The discrete adjoint of the last forward run: from dL/ds_N, dL/ds_0 and the parameter
 * gradients dJ, dD, dB, dK, dK_site. */
SPIRIT_API int Spirit_HIP_Adjoint_Backward( Spirit_HIP_Adjoint * handle, const double * dL_dsN, double * dL_ds0,
                                            double * dJ, double * dD, double * dB, double * dK, double * dK_site,
                                            int on_device ) SPIRIT_NOEXCEPT;
/* Tangent-linear run: s_N and its directional derivative ds_N along (ds0, dJ, dD, dB, dK, dK_site). */
SPIRIT_API int Spirit_HIP_Adjoint_Tangent( Spirit_HIP_Adjoint * handle, const double * s0, const double * ds0,
                                           const double * dJ, const double * dD, const double * dB, const double * dK,
                                           const double * dK_site, double * sN, double * dsN,
                                           int on_device ) SPIRIT_NOEXCEPT;

/* Name of the GPU the engine runs on (empty if none is visible). */
SPIRIT_API const char * Spirit_HIP_Device_Name( void ) SPIRIT_NOEXCEPT;

#endif
