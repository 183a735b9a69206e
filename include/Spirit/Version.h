/* Version: build information (Spirit_Scalar_Type is "double"; the backend is HIP).
 * Replaces core/include/Spirit/Version.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_VERSION_H
#define SPIRIT_HIP_VERSION_H
#include "Spirit_Defines.h"

/* Version.h:6 */
SPIRIT_API int Spirit_Version_Major( void ) SPIRIT_NOEXCEPT;
/* Version.h:7 */
SPIRIT_API int Spirit_Version_Minor( void ) SPIRIT_NOEXCEPT;
/* Version.h:8 */
SPIRIT_API int Spirit_Version_Patch( void ) SPIRIT_NOEXCEPT;
/* Version.h:10 */
SPIRIT_API const char * Spirit_Version( void ) SPIRIT_NOEXCEPT;
/* Version.h:11 */
SPIRIT_API const char * Spirit_Version_Revision( void ) SPIRIT_NOEXCEPT;
/* Version.h:12 */
SPIRIT_API const char * Spirit_Version_Full( void ) SPIRIT_NOEXCEPT;
/* Version.h:14 */
SPIRIT_API const char * Spirit_Compiler( void ) SPIRIT_NOEXCEPT;
/* Version.h:15 */
SPIRIT_API const char * Spirit_Compiler_Version( void ) SPIRIT_NOEXCEPT;
/* Version.h:16 */
SPIRIT_API const char * Spirit_Compiler_Full( void ) SPIRIT_NOEXCEPT;
/* Version.h:18 */
SPIRIT_API const char * Spirit_Scalar_Type( void ) SPIRIT_NOEXCEPT;
/* Version.h:20 */
SPIRIT_API const char * Spirit_Defects( void ) SPIRIT_NOEXCEPT;
/* Version.h:21 */
SPIRIT_API const char * Spirit_Pinning( void ) SPIRIT_NOEXCEPT;
/* Version.h:23 */
SPIRIT_API const char * Spirit_Cuda( void ) SPIRIT_NOEXCEPT;
/* Version.h:24 */
SPIRIT_API const char * Spirit_OpenMP( void ) SPIRIT_NOEXCEPT;
/* Version.h:25 */
SPIRIT_API int Spirit_OpenMP_Get_Num_Threads( void ) SPIRIT_NOEXCEPT;
/* Version.h:26 */
SPIRIT_API const char * Spirit_Threads( void ) SPIRIT_NOEXCEPT;
/* Version.h:27 */
SPIRIT_API const char * Spirit_FFTW( void ) SPIRIT_NOEXCEPT;

#endif
