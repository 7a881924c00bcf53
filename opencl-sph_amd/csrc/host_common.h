// Internal declarations shared by the host-side translation units of
// libsph_hip.so (not part of the public C ABI).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sph_hip.h"

namespace sphhip {

// How a psdata field is indexed.  The builder records it from the conf text
// (a field whose last dimension is given by the word "pnum" is indexed by
// particle); the device runtime uses it to decide which fields follow the
// cell-sorted particle permutation.
enum class FieldKind : uint8_t {
    Other = 0,      // scalars, parameters, small tables
    Particle = 1,   // last dimension is pnum: one tuple per particle
    Cell = 2,       // gridcount / celloffset: one entry per grid cell
    Slot = 3,       // cellparticles: one entry per sorted slot
};

// Kinds of every field of the psdata whose blob starts at `blob`, or an empty
// vector when the psdata was not built by this library.
std::vector<FieldKind> field_kinds_for(const void *blob);

void set_build_status(int status);

}  // namespace sphhip
