/* Geometry: the lattice of the spin systems.
 * Replaces core/include/Spirit/Geometry.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_GEOMETRY_H
#define SPIRIT_HIP_GEOMETRY_H
#include "Spirit_Defines.h"

/* Geometry.h:28-39 */
typedef enum
{
    Bravais_Lattice_Irregular   = 0,
    Bravais_Lattice_Rectilinear = 1,
    Bravais_Lattice_SC          = 2,
    Bravais_Lattice_Hex2D       = 3,
    Bravais_Lattice_Hex2D_60    = 4,
    Bravais_Lattice_Hex2D_120   = 5,
    Bravais_Lattice_HCP         = 6,
    Bravais_Lattice_BCC         = 7,
    Bravais_Lattice_FCC         = 8
} Bravais_Lattice_Type;

/* Geometry.h:51 */
SPIRIT_API void Geometry_Set_Bravais_Lattice_Type( State * state, Bravais_Lattice_Type lattice_type ) SPIRIT_NOEXCEPT;
/* Geometry.h:56 */
SPIRIT_API void Geometry_Set_N_Cells( State * state, int n_cells[3] ) SPIRIT_NOEXCEPT;
/* Geometry.h:62 */
SPIRIT_API void Geometry_Set_Cell_Atoms( State * state, int n_atoms, float ** atoms ) SPIRIT_NOEXCEPT;
/* Geometry.h:67 */
SPIRIT_API void Geometry_Set_mu_s( State * state, float mu_s, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:72 (defects are off in this build: logs a warning) */
SPIRIT_API void Geometry_Set_Cell_Atom_Types( State * state, int n_atoms, int * atom_types ) SPIRIT_NOEXCEPT;
/* Geometry.h:77 */
SPIRIT_API void Geometry_Set_Bravais_Vectors( State * state, float ta[3], float tb[3], float tc[3] ) SPIRIT_NOEXCEPT;
/* Geometry.h:82 */
SPIRIT_API void Geometry_Set_Lattice_Constant( State * state, float lattice_constant ) SPIRIT_NOEXCEPT;
/* Geometry.h:92 */
SPIRIT_API int Geometry_Get_NOS( State * state ) SPIRIT_NOEXCEPT;
/* Geometry.h:97 */
SPIRIT_API scalar * Geometry_Get_Positions( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:102 */
SPIRIT_API int * Geometry_Get_Atom_Types( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:108 */
SPIRIT_API void Geometry_Get_Bounds( State * state, float min[3], float max[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:113 */
SPIRIT_API void Geometry_Get_Center( State * state, float center[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:119 */
SPIRIT_API Bravais_Lattice_Type Geometry_Get_Bravais_Lattice_Type( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:124 */
SPIRIT_API void Geometry_Get_Bravais_Vectors( State * state, float a[3], float b[3], float c[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:130 */
SPIRIT_API int Geometry_Get_Dimensionality( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:135 */
SPIRIT_API void Geometry_Get_mu_s( State * state, float * mu_s, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:140 */
SPIRIT_API void Geometry_Get_N_Cells( State * state, int n_cells[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:150 */
SPIRIT_API void Geometry_Get_Cell_Bounds( State * state, float min[3], float max[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:155 */
SPIRIT_API int Geometry_Get_N_Cell_Atoms( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Geometry.h:160 */
SPIRIT_API int Geometry_Get_Cell_Atoms( State * state, scalar ** atoms, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
