// C-callable shim over the reference's Lookup::boys (compiled from the reference
// tree into oracle/_ref/ only; used to write the Boys golden vectors).
#include "boys.h"
extern "C" double ref_lookup_boys(int n, double x) { return HartreeFock::Lookup::boys(n, x); }
