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

// Device blocks of the drop-in grids and temporaries, cached by exact size:
// pool_free keeps a block for the next pool_alloc of that size on the same
// device (the reference's call order allocates and frees the same grid
// sizes again and again; hipMalloc/hipFree of multi-GB blocks cost far more
// than the kernels).  pool_free waits for the device first, so work queued
// on any stream has finished with the block.  mitos_release_workspace()
// returns the cached blocks; an allocation that fails releases them and
// tries once more.
void* pool_alloc(size_t bytes);
void pool_free(void* p);
// return the cached free blocks (before the large allocations of a fused
// pipeline or an ensemble)
void pool_trim();

}  // namespace mitos
