// Forced include for compiling the reference's host sources with g++: they
// call max/min unqualified and use size_t without including <cstddef>
// (src/host-arith/field/mini-gmp.cu:47). Standard-library names only.
#include <algorithm>
#include <cstddef>
#include <cstring>
using std::max;
using std::min;
