/* Constants: the physical constants the engine uses.
 * Replaces core/include/Spirit/Constants.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_CONSTANTS_H
#define SPIRIT_HIP_CONSTANTS_H
#include "Spirit_Defines.h"

/* Constants.h:20 */
SPIRIT_API scalar Constants_mu_B( void ) SPIRIT_NOEXCEPT;
/* Constants.h:23 */
SPIRIT_API scalar Constants_mu_0( void ) SPIRIT_NOEXCEPT;
/* Constants.h:26 */
SPIRIT_API scalar Constants_k_B( void ) SPIRIT_NOEXCEPT;
/* Constants.h:29 */
SPIRIT_API scalar Constants_hbar( void ) SPIRIT_NOEXCEPT;
/* Constants.h:32 */
SPIRIT_API scalar Constants_mRy( void ) SPIRIT_NOEXCEPT;
/* Constants.h:35 */
SPIRIT_API scalar Constants_gamma( void ) SPIRIT_NOEXCEPT;
/* Constants.h:38 */
SPIRIT_API scalar Constants_g_e( void ) SPIRIT_NOEXCEPT;
/* Constants.h:41 */
SPIRIT_API scalar Constants_Pi( void ) SPIRIT_NOEXCEPT;

#endif
