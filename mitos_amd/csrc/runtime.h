// Library-wide runtime state: the stream used by the drop-in calls and
// device workspaces reused across calls.
#pragma once

#include "mitos_common.h"

namespace mitos {

// Stream of the drop-in (reference-named) entry points.  Every drop-in call
// is host-synchronous, as the reference's CPU functions are.
hipStream_t api_stream();
int api_sync();

// Per-device scratch buffer of at least `bytes`, grown on demand and kept
// until mitos_release_workspace().  Not thread-safe across host threads.
void* workspace(int slot, size_t bytes);

}  // namespace mitos
