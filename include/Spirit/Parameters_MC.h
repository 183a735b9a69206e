/* Parameters_MC: Monte Carlo parameters.
 * Replaces core/include/Spirit/Parameters_MC.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_PARAMETERS_MC_H
#define SPIRIT_HIP_PARAMETERS_MC_H
#include "Spirit_Defines.h"

/* Parameters_MC.h:70 */
SPIRIT_API void Parameters_MC_Set_N_Iterations( State * state, int n_iterations, int n_iterations_log,
                                                int idx_image SPIRIT_DEFAULT( -1 ),
                                                int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_MC.h:79 */
SPIRIT_API void Parameters_MC_Set_Temperature( State * state, float T, int idx_image SPIRIT_DEFAULT( -1 ),
                                               int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_MC.h:89 */
SPIRIT_API void Parameters_MC_Set_Metropolis_Cone( State * state, bool cone, float cone_angle, bool adaptive_cone,
                                                   float target_acceptance_ratio, int idx_image SPIRIT_DEFAULT( -1 ),
                                                   int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_MC.h:94 */
SPIRIT_API void Parameters_MC_Set_Random_Sample( State * state, bool random_sample, int idx_image SPIRIT_DEFAULT( -1 ),
                                                 int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_MC.h:123 */
SPIRIT_API void Parameters_MC_Get_N_Iterations( State * state, int * iterations, int * iterations_log,
                                                int idx_image SPIRIT_DEFAULT( -1 ),
                                                int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_MC.h:132 */
SPIRIT_API float Parameters_MC_Get_Temperature( State * state, int idx_image SPIRIT_DEFAULT( -1 ),
                                                int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_MC.h:142 */
SPIRIT_API void Parameters_MC_Get_Metropolis_Cone( State * state, bool * cone, float * cone_angle,
                                                   bool * adaptive_cone, float * target_acceptance_ratio,
                                                   int idx_image SPIRIT_DEFAULT( -1 ),
                                                   int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_MC.h:147 */
SPIRIT_API bool Parameters_MC_Get_Random_Sample( State * state, int idx_image SPIRIT_DEFAULT( -1 ),
                                                 int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
