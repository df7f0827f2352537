/* Scalar typedefs of the AOCL-FFTZ C interface.  Names, widths and
 * signedness are identical to the reference's api/types.h:21-54 so existing
 * callers compile and link unchanged. */
#ifndef TYPES_H
#define TYPES_H

#include <stddef.h>
#include <stdint.h>

typedef int64_t FFTZ_INT64;
typedef int32_t FFTZ_INT32;
typedef ptrdiff_t FFTZ_INTP;
typedef uint64_t FFTZ_UINT64;
typedef uint32_t FFTZ_UINT32;
typedef size_t FFTZ_UINTP;
typedef char FFTZ_CHAR;
typedef unsigned char FFTZ_UCHAR;
typedef short FFTZ_SHORT;
typedef unsigned short FFTZ_USHORT;
typedef void FFTZ_VOID;
typedef float FFTZ_FLOAT32;
typedef float FFTZ_FLOAT;
typedef double FFTZ_FLOAT64;
typedef double FFTZ_DOUBLE;
typedef uint8_t FFTZ_UINT8;
typedef int8_t FFTZ_INT8;

#endif /* TYPES_H */
