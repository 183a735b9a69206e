/* Parameters_LLG: parameters of the LLG method.
 * Replaces core/include/Spirit/Parameters_LLG.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_PARAMETERS_LLG_H
#define SPIRIT_HIP_PARAMETERS_LLG_H
#include "Spirit_Defines.h"

/* Parameters_LLG.h:30 */
SPIRIT_API void Parameters_LLG_Set_Output_Tag( State * state, const char * tag, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:34 */
SPIRIT_API void Parameters_LLG_Set_Output_Folder( State * state, const char * folder, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:37 */
SPIRIT_API void Parameters_LLG_Set_Output_General( State * state, bool any, bool initial, bool final, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:70 */
SPIRIT_API void Parameters_LLG_Set_N_Iterations( State * state, int n_iterations, int n_iterations_log, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:85 */
SPIRIT_API void Parameters_LLG_Set_Direct_Minimization( State * state, bool direct, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:94 */
SPIRIT_API void Parameters_LLG_Set_Convergence( State * state, float convergence, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:97 */
SPIRIT_API void Parameters_LLG_Set_Time_Step( State * state, float dt, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:100 */
SPIRIT_API void Parameters_LLG_Set_Damping( State * state, float damping, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:104 */
SPIRIT_API void Parameters_LLG_Set_Non_Adiabatic_Damping( State * state, float beta, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:113 (a normal shorter than 0.9 becomes (0, 0, 1), as Parameters_LLG.cpp:332-336) */
SPIRIT_API void Parameters_LLG_Set_STT( State * state, bool use_gradient, float magnitude, const float normal[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:118 */
SPIRIT_API void Parameters_LLG_Set_Temperature( State * state, float T, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:126 (the direction is stored as given and normalised where the distribution is formed) */
SPIRIT_API void Parameters_LLG_Set_Temperature_Gradient( State * state, float inclination, const float direction[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:135 */
SPIRIT_API const char * Parameters_LLG_Get_Output_Tag( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:138 */
SPIRIT_API const char * Parameters_LLG_Get_Output_Folder( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:141 */
/* Parameters_LLG.h:49 (stored; energy files are not written by the HIP engine) */
SPIRIT_API void Parameters_LLG_Set_Output_Energy( State * state, bool energy_step, bool energy_archive, bool energy_spin_resolved, bool energy_divide_by_nos, bool energy_add_readability_lines, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:60: spin configuration files per log step (one file each, or appended to an archive) */
SPIRIT_API void Parameters_LLG_Set_Output_Configuration( State * state, bool configuration_step, bool configuration_archive, int configuration_filetype SPIRIT_DEFAULT( 3 ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:145 */
SPIRIT_API void Parameters_LLG_Get_Output_Energy( State * state, bool * energy_step, bool * energy_archive, bool * energy_spin_resolved, bool * energy_divide_by_nos, bool * energy_add_readability_lines, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:150 */
SPIRIT_API void Parameters_LLG_Get_Output_Configuration( State * state, bool * configuration_step, bool * configuration_archive, int * configuration_filetype, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
SPIRIT_API void Parameters_LLG_Get_Output_General( State * state, bool * any, bool * initial, bool * final, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:155 */
SPIRIT_API void Parameters_LLG_Get_N_Iterations( State * state, int * iterations, int * iterations_log, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:164 */
SPIRIT_API bool Parameters_LLG_Get_Direct_Minimization( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:167 */
SPIRIT_API float Parameters_LLG_Get_Convergence( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:170 */
SPIRIT_API float Parameters_LLG_Get_Time_Step( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:173 */
SPIRIT_API float Parameters_LLG_Get_Damping( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:176 */
SPIRIT_API float Parameters_LLG_Get_Non_Adiabatic_Damping( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:179 */
SPIRIT_API float Parameters_LLG_Get_Temperature( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:187 */
SPIRIT_API void Parameters_LLG_Get_Temperature_Gradient( State * state, float * inclination, float direction[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_LLG.h:197 */
SPIRIT_API void Parameters_LLG_Get_STT( State * state, bool * use_gradient, float * magnitude, float normal[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
