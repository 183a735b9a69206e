/* State: creation, update and deletion of the opaque simulation state.
 * Replaces core/include/Spirit/State.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_STATE_H
#define SPIRIT_HIP_STATE_H
#include "Spirit_Defines.h"

/* State.h:23 */
SPIRIT_API State * State_Setup( const char * config_file SPIRIT_DEFAULT( "" ), bool quiet SPIRIT_DEFAULT( false ) ) SPIRIT_NOEXCEPT;
/* State.h:25 */
SPIRIT_API void State_Delete( State * state ) SPIRIT_NOEXCEPT;
/* State.h:81 */
SPIRIT_API void State_Update( State * state ) SPIRIT_NOEXCEPT;
/* State.h:87 */
SPIRIT_API void State_To_Config( State * state, const char * config_file, const char * comment SPIRIT_DEFAULT( "" ) ) SPIRIT_NOEXCEPT;
/* State.h:94 */
SPIRIT_API const char * State_DateTime( State * state ) SPIRIT_NOEXCEPT;

#endif
