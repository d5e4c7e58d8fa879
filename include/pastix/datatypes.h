/**
 * @file pastix/datatypes.h
 *
 * Integer and opaque handle types of the PaStiX C API, MI355X engine.
 *
 * Same widths and names as the reference include/pastix/datatypes.h:20-53
 * (pastix_int_t is int64_t by default because the reference CMake option
 * PASTIX_INT64 defaults to ON, CMakeLists.txt:101-102), so that a caller
 * compiled against the reference headers links against this library.
 */
#ifndef PASTIX_AMD_DATATYPES_H
#define PASTIX_AMD_DATATYPES_H

#include <stdint.h>

#if defined(PASTIX_INT32)
typedef int32_t  pastix_int_t;
typedef uint32_t pastix_uint_t;
#define PASTIX_INT_MAX INT32_MAX
#else
typedef int64_t  pastix_int_t;
typedef uint64_t pastix_uint_t;
#define PASTIX_INT_MAX INT64_MAX
#endif

typedef double pastix_fixdbl_t;

/* Opaque solver instance (reference include/pastix/datatypes.h:137-138). */
struct pastix_data_s;
typedef struct pastix_data_s pastix_data_t;

#endif
