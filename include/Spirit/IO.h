/* IO: OVF 2.0 spin and chain files.
 * Replaces core/include/Spirit/IO.h (same names, argument meaning and error behaviour:
 * errors are logged and a default value is returned). */
#pragma once
#ifndef SPIRIT_HIP_IO_H
#define SPIRIT_HIP_IO_H
#include "Spirit_Defines.h"

/* IO.h:27-40: file formats of vector fields */
#define IO_Fileformat_OVF_bin 0
#define IO_Fileformat_OVF_bin4 1
#define IO_Fileformat_OVF_bin8 2
#define IO_Fileformat_OVF_text 3
#define IO_Fileformat_OVF_csv 4

/* IO.h:50 */
SPIRIT_API void IO_Positions_Write( State * state, const char * file, int format SPIRIT_DEFAULT( IO_Fileformat_OVF_bin ),
                                    const char * comment SPIRIT_DEFAULT( "-" ), int idx_image SPIRIT_DEFAULT( -1 ),
                                    int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* IO.h:60 */
SPIRIT_API int IO_N_Images_In_File( State * state, const char * file, int idx_image SPIRIT_DEFAULT( -1 ),
                                    int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* IO.h:63 */
SPIRIT_API void IO_Image_Read( State * state, const char * file, int idx_image_infile SPIRIT_DEFAULT( 0 ),
                               int idx_image_inchain SPIRIT_DEFAULT( -1 ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* IO.h:67 */
SPIRIT_API void IO_Image_Write( State * state, const char * file, int format SPIRIT_DEFAULT( IO_Fileformat_OVF_bin ),
                                const char * comment SPIRIT_DEFAULT( "-" ), int idx_image SPIRIT_DEFAULT( -1 ),
                                int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* IO.h:72 */
SPIRIT_API void IO_Image_Append( State * state, const char * file, int format SPIRIT_DEFAULT( IO_Fileformat_OVF_bin ),
                                 const char * comment SPIRIT_DEFAULT( "-" ), int idx_image SPIRIT_DEFAULT( -1 ),
                                 int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* IO.h:87 */
SPIRIT_API void IO_Chain_Read( State * state, const char * file, int start_image_infile SPIRIT_DEFAULT( 0 ),
                               int end_image_infile SPIRIT_DEFAULT( -1 ), int insert_idx SPIRIT_DEFAULT( 0 ),
                               int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* IO.h:92 */
SPIRIT_API void IO_Chain_Write( State * state, const char * file, int format SPIRIT_DEFAULT( IO_Fileformat_OVF_text ),
                                const char * comment SPIRIT_DEFAULT( "-" ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;
/* IO.h:97 */
SPIRIT_API void IO_Chain_Append( State * state, const char * file, int format SPIRIT_DEFAULT( IO_Fileformat_OVF_text ),
                                 const char * comment SPIRIT_DEFAULT( "-" ), int idx_chain SPIRIT_DEFAULT( -1 ) ) SPIRIT_NOEXCEPT;

#endif
