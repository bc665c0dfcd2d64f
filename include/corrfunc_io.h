/* Catalog readers of the engine library (libcorrfunc_amd.so): fast-food
   (Fortran unformatted records) and ascii / csv column files.

   Same entry points, argument meaning and return convention as the
   reference's io/io.h:14-15; host code only (no GPU work). */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reads the first num_fields columns of `filename` into newly malloc'ed arrays
   of `size` (4: float, 8: double) bytes per element and stores them through the
   num_fields trailing `void **` arguments. Returns the number of rows, or -1 on
   error (message on stderr; nothing is allocated then). format "f...": fast-food
   file (a 5-int header record whose second int is the row count, a 9-float and a
   1-float record, then one record per column, float or double; columns of the
   other precision are converted); "a..." or "c...": whitespace/comma/tab
   separated text, a line whose first num_fields fields do not all parse as
   numbers is skipped with a warning. A missing `filename` whose `filename.gz`
   exists is first uncompressed in place with `gunzip`.
   Replaces read_positions (io/io.h:14, io/io.c:28-49). */
int64_t read_positions(const char *filename, const char *format, const size_t size, const int num_fields, ...)
    __attribute__((warn_unused_result));

/* As read_positions, the column pointers stored into data[0 .. num_fields-1].
   Replaces read_columns_into_array (io/io.h:15, io/io.c:52-284). */
int64_t read_columns_into_array(const char *filename, const char *format, const size_t size, const int num_fields,
                                void **data) __attribute__((warn_unused_result));

#ifdef __cplusplus
}
#endif
