/* Configurations: set spin configurations of an image inside a spatial filter.
 * Replaces core/include/Spirit/Configurations.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_CONFIGURATIONS_H
#define SPIRIT_HIP_CONFIGURATIONS_H
#include "Spirit_Defines.h"

/* Configurations.h:18-34: pass position {0,0,0} and r_cut_rectangular {-1,-1,-1} for "everywhere" */
/* Configurations.h:42 */
SPIRIT_API void Configuration_To_Clipboard( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:45 */
SPIRIT_API void Configuration_From_Clipboard( State * state, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:62 */
SPIRIT_API void Configuration_Domain( State * state, const float direction[3], const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:68 */
SPIRIT_API void Configuration_PlusZ( State * state, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:74 */
SPIRIT_API void Configuration_MinusZ( State * state, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:80 */
SPIRIT_API void Configuration_Random( State * state, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), bool external SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:86 */
SPIRIT_API void Configuration_SpinSpiral( State * state, const char * direction_type, float q[3], float axis[3], float theta, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:103 */
SPIRIT_API void Configuration_Add_Noise_Temperature( State * state, float temperature, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:118 */
SPIRIT_API void Configuration_Skyrmion( State * state, float r, float order, float phase, bool upDown, bool achiral, bool rl, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:124-127 */
SPIRIT_API void Configuration_DW_Skyrmion( State * state, float dw_radius, float dw_width, float order, float phase, bool upDown, bool achiral, bool rl, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Configurations.h:130-133 */
SPIRIT_API void Configuration_Hopfion( State * state, float r, int order, const float position[3], const float r_cut_rectangular[3], float r_cut_cylindrical SPIRIT_DEFAULT( -1 ), float r_cut_spherical SPIRIT_DEFAULT( -1 ), bool inverted SPIRIT_DEFAULT( false ), const float normal[3] SPIRIT_DEFAULT( 0 ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
