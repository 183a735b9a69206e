/* Shared definitions of the HIP engine's C API (libSpirit.so).
 *
 * The reference generates its Spirit_Defines.h from CMake
 * (core/CMake/Spirit_Defines.h.in) with the `scalar` typedef of the build; this
 * engine is FP64 throughout, so scalar is double.  Default arguments exist for
 * C++ callers only, as in the reference (core/include/Spirit/DLL_Define_Export.h:16-26);
 * C and ctypes callers pass every argument. */
#pragma once
#ifndef SPIRIT_HIP_DEFINES_H
#define SPIRIT_HIP_DEFINES_H

typedef double scalar;

#ifdef __cplusplus
#define SPIRIT_API extern "C"
#define SPIRIT_DEFAULT( x ) = x
#define SPIRIT_NOEXCEPT noexcept
#else
#include <stdbool.h>
#define SPIRIT_API
#define SPIRIT_DEFAULT( x )
#define SPIRIT_NOEXCEPT
#endif

#define SPIRIT_BACKEND_HIP 1

struct State;
typedef struct State State;

#endif
