/*
 * Integer type and assertion macro of the BLOPEX C API
 * (blopex_abstract/utilities/fortran_options.h:9-18).  BlopexInt defaults to
 * int; building with -DBlopexInt=long changes the ABI of every header here.
 */
#ifndef BLOPEX_AMD_FORTRAN_OPTIONS_H
#define BLOPEX_AMD_FORTRAN_OPTIONS_H

#ifndef BlopexInt
#define BlopexInt int
#endif

#ifndef BlopexAssert
#include <assert.h>
#define BlopexAssert assert
#endif

#endif
