/* Quantities: derived observables of an image.
 * Replaces core/include/Spirit/Quantities.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_QUANTITIES_H
#define SPIRIT_HIP_QUANTITIES_H
#include "Spirit_Defines.h"

/* Quantities.h:18 */
SPIRIT_API void Quantity_Get_Average_Spin( State * state, float s[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Quantities.h:21 */
SPIRIT_API void Quantity_Get_Magnetization( State * state, float m[3], int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Quantities.h:24 (2D lattices: solid-angle sum over the lattice's triangles) */
SPIRIT_API float Quantity_Get_Topological_Charge( State * state, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* Quantities.h:27-28: per triangle of a 2D lattice with one basis atom, its charge and three spin
   indices; returns the number of triangles (outputs may be null to ask for it) */
SPIRIT_API int Quantity_Get_Topological_Charge_Density( State * state, float * charge_density, int * triangle_indices, int idx_image SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
