/* Parameters_GNEB: parameters of the GNEB method.
 * Replaces core/include/Spirit/Parameters_GNEB.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_PARAMETERS_GNEB_H
#define SPIRIT_HIP_PARAMETERS_GNEB_H
#include "Spirit_Defines.h"

/* Parameters_GNEB.h:22-35 */
#define GNEB_IMAGE_NORMAL 0
#define GNEB_IMAGE_CLIMBING 1
#define GNEB_IMAGE_FALLING 2
#define GNEB_IMAGE_STATIONARY 3

/* Parameters_GNEB.h:50 */
SPIRIT_API void Parameters_GNEB_Set_Output_Tag( State * state, const char * tag, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:53 */
SPIRIT_API void Parameters_GNEB_Set_Output_Folder( State * state, const char * folder, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:57 */
SPIRIT_API void Parameters_GNEB_Set_Output_General( State * state, bool any, bool initial, bool final, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:87 */
SPIRIT_API void Parameters_GNEB_Set_N_Iterations( State * state, int n_iterations, int n_iterations_log, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:101 */
SPIRIT_API void Parameters_GNEB_Set_Convergence( State * state, float convergence, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:104 */
SPIRIT_API void Parameters_GNEB_Set_Spring_Constant( State * state, float spring_constant, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:108 */
SPIRIT_API void Parameters_GNEB_Set_Spring_Force_Ratio( State * state, float ratio, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:111 */
SPIRIT_API void Parameters_GNEB_Set_Path_Shortening_Constant( State * state, float path_shortening_constant, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:115 */
SPIRIT_API void Parameters_GNEB_Set_Moving_Endpoints( State * state, bool moving_endpoints, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:119 */
SPIRIT_API void Parameters_GNEB_Set_Translating_Endpoints( State * state, bool translating_endpoints, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:122 */
SPIRIT_API void Parameters_GNEB_Set_Equilibrium_Delta_Rx( State * state, float delta_Rx_left, float delta_Rx_right, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:127 */
SPIRIT_API void Parameters_GNEB_Set_Climbing_Falling( State * state, int image_type, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:134 */
SPIRIT_API void Parameters_GNEB_Set_Image_Type_Automatically( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:137 */
SPIRIT_API void Parameters_GNEB_Set_N_Energy_Interpolations( State * state, int n, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:145 */
SPIRIT_API const char * Parameters_GNEB_Get_Output_Tag( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:148 */
SPIRIT_API const char * Parameters_GNEB_Get_Output_Folder( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:151 */
SPIRIT_API void Parameters_GNEB_Get_Output_General( State * state, bool * any, bool * initial, bool * final, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:164 */
SPIRIT_API void Parameters_GNEB_Get_N_Iterations( State * state, int * iterations, int * iterations_log, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:172 */
SPIRIT_API float Parameters_GNEB_Get_Convergence( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:175 */
SPIRIT_API float Parameters_GNEB_Get_Spring_Constant( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:178 */
SPIRIT_API float Parameters_GNEB_Get_Spring_Force_Ratio( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:181 */
SPIRIT_API float Parameters_GNEB_Get_Path_Shortening_Constant( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:184 */
SPIRIT_API bool Parameters_GNEB_Get_Moving_Endpoints( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:187 */
SPIRIT_API bool Parameters_GNEB_Get_Translating_Endpoints( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:190 */
SPIRIT_API void Parameters_GNEB_Get_Equilibrium_Delta_Rx( State * state, float * delta_Rx_left, float * delta_Rx_right, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:198 */
SPIRIT_API int Parameters_GNEB_Get_Climbing_Falling( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Parameters_GNEB.h:201 */
SPIRIT_API int Parameters_GNEB_Get_N_Energy_Interpolations( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
