/* Chain: the chain of images used by GNEB.
 * Replaces core/include/Spirit/Chain.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_CHAIN_H
#define SPIRIT_HIP_CHAIN_H
#include "Spirit_Defines.h"

/* Chain.h:22 */
SPIRIT_API int Chain_Get_NOI( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:34 */
SPIRIT_API bool Chain_next_Image( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:41 */
SPIRIT_API bool Chain_prev_Image( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:48 */
SPIRIT_API bool Chain_Jump_To_Image( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:63 */
SPIRIT_API void Chain_Set_Length( State * state, int n_images, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:70 */
SPIRIT_API void Chain_Image_to_Clipboard( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:77 */
SPIRIT_API void Chain_Replace_Image( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:84 */
SPIRIT_API void Chain_Insert_Image_Before( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:91 */
SPIRIT_API void Chain_Insert_Image_After( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:98 */
SPIRIT_API void Chain_Push_Back( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:105 */
SPIRIT_API bool Chain_Delete_Image( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:114 */
SPIRIT_API bool Chain_Pop_Back( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:124 */
SPIRIT_API void Chain_Get_Rx( State * state, float * Rx, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:129 */
SPIRIT_API void Chain_Get_Rx_Interpolated( State * state, float * Rx_interpolated, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:134 */
SPIRIT_API void Chain_Get_Energy( State * state, float * energy, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:139 */
SPIRIT_API void Chain_Get_Energy_Interpolated( State * state, float * E_interpolated, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:151 */
SPIRIT_API void Chain_Update_Data( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Chain.h:156 */
SPIRIT_API void Chain_Setup_Data( State * state, int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
