/* Simulation: start, advance, stop and monitor the LLG and GNEB methods.
 * Replaces core/include/Spirit/Simulation.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_SIMULATION_H
#define SPIRIT_HIP_SIMULATION_H
#include "Spirit_Defines.h"

/* Simulation.h:33-54 */
#define Solver_VP 0
#define Solver_SIB 1
#define Solver_Depondt 2
#define Solver_Heun 3
#define Solver_RungeKutta4 4
#define Solver_LBFGS_OSO 5
#define Solver_LBFGS_Atlas 6
#define Solver_VP_OSO 7

/* Simulation.h:58-71 (member defaults are zero; C callers zero-initialise) */
typedef struct Simulation_Run_Info
{
    int total_iterations;
    int total_walltime;
    float total_ips;
    float max_torque;
    int n_history_iteration;
    int * history_iteration;
    int n_history_max_torque;
    float * history_max_torque;
    int n_history_energy;
    float * history_energy;
} Simulation_Run_Info;

/* Simulation.h:73 */
SPIRIT_API void free_run_info( Simulation_Run_Info info ) SPIRIT_NOEXCEPT;
/* Simulation.h:81 (Metropolis sweeps of colour classes on the GPU; one basis atom; no single shot) */
SPIRIT_API void Simulation_MC_Start( State * state, int n_iterations SPIRIT_DEFAULT( -1 ), int n_iterations_log SPIRIT_DEFAULT( -1 ), bool singleshot SPIRIT_DEFAULT( false ), Simulation_Run_Info * info SPIRIT_DEFAULT( 0 ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:86 */
SPIRIT_API void Simulation_LLG_Start( State * state, int solver_type, int n_iterations SPIRIT_DEFAULT( -1 ), int n_iterations_log SPIRIT_DEFAULT( -1 ), bool singleshot SPIRIT_DEFAULT( false ), Simulation_Run_Info * info SPIRIT_DEFAULT( 0 ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:91 */
SPIRIT_API void Simulation_GNEB_Start( State * state, int solver_type, int n_iterations SPIRIT_DEFAULT( -1 ), int n_iterations_log SPIRIT_DEFAULT( -1 ), bool singleshot SPIRIT_DEFAULT( false ), Simulation_Run_Info * info SPIRIT_DEFAULT( 0 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:96 (not part of the HIP engine: logs an error) */
SPIRIT_API void Simulation_MMF_Start( State * state, int solver_type, int n_iterations SPIRIT_DEFAULT( -1 ), int n_iterations_log SPIRIT_DEFAULT( -1 ), bool singleshot SPIRIT_DEFAULT( false ), Simulation_Run_Info * info SPIRIT_DEFAULT( 0 ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:101 (not part of the HIP engine: logs an error) */
SPIRIT_API void Simulation_EMA_Start( State * state, int n_iterations SPIRIT_DEFAULT( -1 ), int n_iterations_log SPIRIT_DEFAULT( -1 ), bool singleshot SPIRIT_DEFAULT( false ), Simulation_Run_Info * info SPIRIT_DEFAULT( 0 ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:111 */
SPIRIT_API void Simulation_SingleShot( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:119 */
SPIRIT_API void Simulation_N_Shot( State * state, int N, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:122 */
SPIRIT_API void Simulation_Stop( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:125 */
SPIRIT_API void Simulation_Stop_All( State * state ) SPIRIT_NOEXCEPT;
/* Simulation.h:139 */
SPIRIT_API float Simulation_Get_MaxTorqueComponent( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:146 */
SPIRIT_API void Simulation_Get_Chain_MaxTorqueComponents( State * state, float * torques, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:155 */
SPIRIT_API float Simulation_Get_MaxTorqueNorm( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:162 */
SPIRIT_API void Simulation_Get_Chain_MaxTorqueNorms( State * state, float * torques, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:171 */
SPIRIT_API float Simulation_Get_IterationsPerSecond( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:174 */
SPIRIT_API int Simulation_Get_Iteration( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:183 */
SPIRIT_API float Simulation_Get_Time( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:186 */
SPIRIT_API int Simulation_Get_Wall_Time( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:195 */
SPIRIT_API const char * Simulation_Get_Solver_Name( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:204 */
SPIRIT_API const char * Simulation_Get_Method_Name( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:212 */
SPIRIT_API bool Simulation_Running_On_Image( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:215 */
SPIRIT_API bool Simulation_Running_On_Chain( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Simulation.h:218 */
SPIRIT_API bool Simulation_Running_Anywhere_On_Chain( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
