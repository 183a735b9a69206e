/* System: access to one image's spins, effective field and energy.
 * Replaces core/include/Spirit/System.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_SYSTEM_H
#define SPIRIT_HIP_SYSTEM_H
#include "Spirit_Defines.h"

/* System.h:25 */
SPIRIT_API int System_Get_Index( State * state ) SPIRIT_NOEXCEPT;
/* System.h:28 */
SPIRIT_API int System_Get_NOS( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:35 -- borrowed AoS (nos,3) host mirror, refreshed from the GPU */
SPIRIT_API scalar * System_Get_Spin_Directions( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:42 */
SPIRIT_API scalar * System_Get_Effective_Field( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:49 (no eigenmodes in the HIP engine: returns NULL) */
SPIRIT_API scalar * System_Get_Eigenmode( State * state, int idx_mode, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:52 */
SPIRIT_API float System_Get_Rx( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:55 */
SPIRIT_API float System_Get_Energy( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:59 */
SPIRIT_API int System_Get_Energy_Array_Names( State * state, char * names, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:63 */
SPIRIT_API int System_Get_Energy_Array( State * state, float * energies, bool divide_by_nspins SPIRIT_DEFAULT( true ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:67 (no eigenmodes: no-op) */
SPIRIT_API void System_Get_Eigenvalues( State * state, float * eigenvalues, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:70 */
SPIRIT_API void System_Print_Energy_Array( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:73 */
SPIRIT_API void System_Update_Data( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* System.h:76 (no eigenmodes: logs an error) */
SPIRIT_API void System_Update_Eigenmodes( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
