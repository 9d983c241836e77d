// vpmg_internal.h -- library-internal helpers shared by vpmg.cpp and server.cpp.
#pragma once
#include <cstddef>

#include "../../include/apbs_pmg.h"

namespace pmgint {

// Upper estimate of the device bytes one Vpmg holds from Vpmg_ctor through
// Vpmg_solve and the energy: fine maps, atoms, fillco temporaries (nref SAS
// reference points per atom for srfm mol/smol, 0 otherwise) and the
// multigrid hierarchy, including the caching pool's rounding.
size_t footprint_bytes(int nx, int ny, int nz, int nlev, bool newton, bool mixed, int natoms,
                       int nref);

// Whether Vpmg_fillco may compute the MDH boundary on a second stream while
// the maps are built (default on).  The service turns it off: there other
// requests fill the device, and the extra stream per request only adds queue
// contention (measured 3% fewer served solves/s).
void set_side_stream(Vpmg* pmg, bool on);

}  // namespace pmgint
