// rt_alloc.h -- device allocation epoch of the library.  Every device
// (re)allocation made by the library bumps it; a captured frame graph
// records the epoch it was captured in and is replayed only in that epoch,
// so it never runs on freed or moved buffers.
#pragma once

#include <stdint.h>

uint64_t rt_alloc_epoch();
void rt_alloc_bump();
