/* Transitions: initial guesses for a chain of images.
 * Replaces core/include/Spirit/Transitions.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_TRANSITIONS_H
#define SPIRIT_HIP_TRANSITIONS_H
#include "Spirit_Defines.h"

/* Transitions.h:29 */
SPIRIT_API void Transition_Homogeneous( State * state, int idx_1, int idx_2, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Transitions.h:37 */
SPIRIT_API void Transition_Homogeneous_Insert_Interpolated( State * state, int n_interpolate, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Transitions.h:47 */
SPIRIT_API void Transition_Add_Noise_Temperature( State * state, float temperature, int idx_1, int idx_2, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
