/* Hamiltonian: parameters of the Heisenberg Hamiltonian (exchange, DMI, anisotropy, Zeeman, DDI).
 * Replaces core/include/Spirit/Hamiltonian.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_HAMILTONIAN_H
#define SPIRIT_HIP_HAMILTONIAN_H
#include "Spirit_Defines.h"

/* Hamiltonian.h:30-40 */
#define SPIRIT_CHIRALITY_BLOCH 1
#define SPIRIT_CHIRALITY_NEEL 2
#define SPIRIT_CHIRALITY_BLOCH_INVERSE -1
#define SPIRIT_CHIRALITY_NEEL_INVERSE -2

/* Hamiltonian.h:47-57 (FFT is the method the HIP engine evaluates; CUTOFF/FMM log a warning and are ignored) */
#define SPIRIT_DDI_METHOD_NONE 0
#define SPIRIT_DDI_METHOD_FFT 1
#define SPIRIT_DDI_METHOD_FMM 2
#define SPIRIT_DDI_METHOD_CUTOFF 3

/* Hamiltonian.h:65 */
SPIRIT_API void Hamiltonian_Set_Boundary_Conditions( State * state, const bool * periodical, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:69 */
SPIRIT_API void Hamiltonian_Set_Field( State * state, float magnitude, const float * normal, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:73 */
SPIRIT_API void Hamiltonian_Set_Anisotropy( State * state, float magnitude, const float * normal, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:78 */
SPIRIT_API void Hamiltonian_Set_Cubic_Anisotropy( State * state, float magnitude, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:81 */
SPIRIT_API void Hamiltonian_Set_Exchange( State * state, int n_shells, const float * jij, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:85 */
SPIRIT_API void Hamiltonian_Set_DMI( State * state, int n_shells, const float * dij, int chirality SPIRIT_DEFAULT( SPIRIT_CHIRALITY_BLOCH ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:100 */
SPIRIT_API void Hamiltonian_Set_DDI( State * state, int ddi_method, int n_periodic_images[3], float cutoff_radius SPIRIT_DEFAULT( 0 ), bool pb_zero_padding SPIRIT_DEFAULT( true ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:110 */
SPIRIT_API const char * Hamiltonian_Get_Name( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:114 */
SPIRIT_API void Hamiltonian_Get_Boundary_Conditions( State * state, bool * periodical, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:117 */
SPIRIT_API void Hamiltonian_Get_Field( State * state, float * magnitude, float * normal, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:121 */
SPIRIT_API void Hamiltonian_Get_Anisotropy( State * state, float * magnitude, float * normal, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:126 */
SPIRIT_API void Hamiltonian_Get_Cubic_Anisotropy( State * state, float * magnitude, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:134 */
SPIRIT_API void Hamiltonian_Get_Exchange_Shells( State * state, int * n_shells, float * jij, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:138 */
SPIRIT_API int Hamiltonian_Get_Exchange_N_Pairs( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:140 */
SPIRIT_API void Hamiltonian_Get_Exchange_Pairs( State * state, int idx[][2], int translations[][3], float * Jij, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:149 */
SPIRIT_API void Hamiltonian_Get_DMI_Shells( State * state, int * n_shells, float * dij, int * chirality, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:153 */
SPIRIT_API int Hamiltonian_Get_DMI_N_Pairs( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:164 */
SPIRIT_API void Hamiltonian_Get_DDI( State * state, int * ddi_method, int n_periodic_images[3], float * cutoff_radius, bool * pb_zero_padding, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Hamiltonian.h:172: the 3N x 3N embedding Hessian of anisotropy, exchange and DMI as triplets or a dense matrix */
SPIRIT_API void Hamiltonian_Write_Hessian( State * state, const char * filename, bool triplet_format SPIRIT_DEFAULT( true ), int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
