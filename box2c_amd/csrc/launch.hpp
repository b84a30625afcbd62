// Host-side launchers of the kernels in kernels.hip (each checks its launch shape and returns
// the launch status).
#pragma once

#include <hip/hip_runtime.h>

#include "b2mi_layout.h"

namespace b2mi
{

hipError_t launch_init_contacts( const DevWorld* worlds, const ContactInit* inits, int count, hipStream_t stream );
hipError_t launch_collide( const DevWorld* worlds, const int2* list, int count, ChangeRecord* changes, int* changeCount,
						   hipStream_t stream );
hipError_t launch_solve( const DevWorld* worlds, const WorkItem* items, int itemCount, int blockSize, int ldsBytes,
						 const int* bodyList, const int* slotContact, const int2* slotBodies, float* cf, int cfCapacity,
						 float4* priv, EnlargeRecord* enl, int* enlCount, int enlCapacity, hipStream_t stream );

} // namespace b2mi
