/* Compatibility prefix for compiling the reference with a modern libstdc++.
 * The reference declares `inline bool isnan(double)` unless `isnan` is a
 * macro (Euler3D/include/Commons.h:16-20), which clashes with the C++11
 * std::isnan brought into the global namespace by <cmath>.  Including <cmath>
 * first and defining isnan as the compiler builtin makes the reference take
 * its own `#ifndef isnan` branch.  No reference header, library or generated
 * file is replaced. */
#include <cmath>
#define isnan(x) __builtin_isnan(x)
