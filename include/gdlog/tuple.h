// gdlog/tuple.h -- tuple types and hook typedefs of the gdlog C++ API.
//
// Same names and meaning as the reference (include/tuple.cuh:5-25): a column
// is `unsigned long` (8 bytes on LP64) on the host API and in hook
// signatures, a tuple is a pointer to `arity` columns.  The engine stores each
// tuple of arity <= 2 as one u64 key over u32 ids; load_relation checks that
// every id fits in 32 bits.
#pragma once

#include <cstdint>

using u64 = unsigned long long;
using u32 = unsigned long;

using column_type = u32;
using tuple_type = column_type *;
using tuple_size_t = u64;

// Device function pointers, obtained by the program with hipMemcpyFromSymbol
// exactly as the reference programs obtain them (test/tc.cu:126-130).
typedef void (*tuple_generator_hook)(tuple_type, tuple_type, tuple_type);
typedef void (*tuple_copy_hook)(tuple_type, tuple_type);
typedef bool (*tuple_predicate)(tuple_type);
