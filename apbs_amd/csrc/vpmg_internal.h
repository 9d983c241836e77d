// vpmg_internal.h -- library-internal helpers shared by vpmg.cpp and server.cpp.
#pragma once
#include <cstddef>

namespace pmgint {

// Upper estimate of the device bytes one Vpmg holds from Vpmg_ctor through
// Vpmg_solve and the energy: fine maps, atoms, fillco temporaries (nref SAS
// reference points per atom for srfm mol/smol, 0 otherwise) and the
// multigrid hierarchy, including the caching pool's rounding.
size_t footprint_bytes(int nx, int ny, int nz, int nlev, bool newton, bool mixed, int natoms,
                       int nref);

}  // namespace pmgint
