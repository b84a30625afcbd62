// Host-side launchers of the kernels in kernels.hip (each checks its launch shape and returns
// the launch status).
#pragma once

#include <hip/hip_runtime.h>

#include "b2mi_layout.h"

namespace b2mi
{

hipError_t launch_init_contacts( const DevWorld* worlds, const ContactInit* inits, int count, hipStream_t stream );
hipError_t launch_collide( const DevWorld* worlds, const int4* list, int count, ChangeRecord* changes, int* changeCount,
						   hipStream_t stream );
hipError_t launch_solve( int store, const DevWorld* worlds, const WorkItem* items, int itemOffset, int itemCount, int blockSize,
						 int ldsBytes, const int* bodyList, const int* slotContact, const int2* slotBodies, const int* ovfLevels,
						 float* cf, float4* priv, float2* mass, EnlargeRecord* enl, int* enlCount, int enlCapacity,
						 int levelRun, const int* abortFlag, hipStream_t stream );

// one very large HBM-store work item over a cooperative grid of `blocks` workgroups of
// kGridSolveBlock threads (k_solve_grid); sync holds two words it uses, error is set if a barrier
// wait exceeded its bound. solve_grid_capacity: the most workgroups one such launch may have (all
// resident at once), 0 when the device has no cooperative launches. levelRun: consecutive overflow
// levels of at most this many slots in all run as one cached run in workgroup 0 (0 = none).
int solve_grid_capacity();
hipError_t launch_solve_grid( const DevWorld* worlds, const WorkItem* items, int itemIndex, int blocks, const int* bodyList,
							  const int* slotContact, const int2* slotBodies, const int* ovfLevels, float* cf, float4* priv, float2* mass,
							  EnlargeRecord* enl, int* enlCount, int enlCapacity, unsigned* sync, int* error, int levelRun,
							  const int* abortFlag, hipStream_t stream );

// one phase of a very large work item over a grid (k_split_phase; phases in SplitPhase): `count`
// slots or bodies from `first` of phase, beside `jointCount` joint slots from jointFirst of
// jointPhase (a joint phase, or ignored when jointCount is 0)
hipError_t launch_split_phase( const DevWorld* worlds, const WorkItem* items, int itemIndex, int phase, int first, int count,
							   int jointPhase, int jointFirst, int jointCount, const int* bodyList, const int* slotContact,
							   const int2* slotBodies, float* cf, float4* priv, float2* mass, EnlargeRecord* enl, int* enlCount,
							   int enlCapacity, hipStream_t stream );
// overflow levels [levelBegin, levelBegin + levelCount) of one very large work item, at most
// kLevelCacheSlots slots in all, as one cached run in one workgroup (k_split_levels), for an
// overflow or joint level phase; levels is the device level list
hipError_t launch_split_levels( const DevWorld* worlds, const WorkItem* items, int itemIndex, int phase, const int* levels, int levelBegin,
								int levelCount, const int* bodyList, const int* slotContact, const int2* slotBodies, float* cf,
								float4* priv, float2* mass, hipStream_t stream );

// the continuous stage over `count` fast bodies (k_continuous)
hipError_t launch_continuous( const DevWorld* worlds, const ContinuousBody* list, int count, const int* noCollide, EnlargeRecord* enl,
							  int* enlCount, int enlCapacity, int* hits, int* error, hipStream_t stream );

// broadphase pair queries (broadphase.hip)
hipError_t launch_pair_erase( unsigned long long* table, uint32_t mask, const unsigned long long* keys, int count, hipStream_t s );
hipError_t launch_pair_insert( unsigned long long* table, uint32_t mask, const unsigned long long* keys, int count, int* error,
							   hipStream_t s );
hipError_t launch_mark_moved( uint32_t* bits, const int* moves, int count, bool set, hipStream_t s );
hipError_t launch_find_pairs( const DevBroadphase& bp, const int* moves, int moveCount, PairRecord* out, int capacity, int* counters,
							  PairSpill* spills, int spillCapacity, hipStream_t s );
// the spilled subtrees (`spilled` of them, as counted by launch_find_pairs)
hipError_t launch_find_pairs_spilled( const DevBroadphase& bp, const int* moves, int spilled, PairRecord* out, int capacity, int* counters,
									  const PairSpill* spills, int spillCapacity, hipStream_t s );
// the pair set's hash on the host (the same probe sequence as the kernels')
inline uint64_t pair_hash_host( uint64_t k )
{
	k ^= k >> 30;
	k *= 0xbf58476d1ce4e5b9ull;
	k ^= k >> 27;
	k *= 0x94d049bb133111ebull;
	k ^= k >> 31;
	return k;
}

// replay and adjoint (adjoint.hip)
int adjoint_scratch_floats( int bodyCount, int slotCount, int subSteps );
hipError_t launch_snapshot( const DevWorld* worlds, const WorkItem* items, int itemCount, const int* slotContact, DevContact* snap,
							hipStream_t stream );
hipError_t launch_gather_poses( const DevWorld* worlds, const int* simBase, const int* simCount, int worldCount, int maxSims,
								StepPose* out, hipStream_t stream );
// storeLdsBytes: 0 = the solver state (and the adjoint's body and slot adjoints) in HBM; else
// the dynamic LDS bytes of the launch's largest item, whose solver state and adjoints are in LDS
hipError_t launch_replay_step( const DevWorld* worlds, const WorkItem* items, int itemCount, int blockSize, const int* bodyList,
							   const int2* slotBodies, const DevContact* snap, const StepPose* poseIn, StepPose* poseOut,
							   const int* simBase, float* cf, int cfCapacity, float4* priv, float2* mass, int storeLdsBytes,
							   hipStream_t stream );
hipError_t launch_adjoint_step( const DevWorld* worlds, const WorkItem* items, int itemCount, int blockSize, const int* bodyList,
								const int2* slotBodies, const DevContact* snap, const StepPose* poseIn, StepPose* adj,
								const int* simBase, const float4* lamOut, float* lamPrev, const int* prevSlot, float* paramGrad,
								float* cf, int cfCapacity, float4* priv, float2* mass, float* scratch, const int* scratchOffset,
								int storeLdsBytes, hipStream_t stream );
hipError_t launch_link_slots( const int* slotContact, const int* slotWorld, int slotCount, const int* contactBase,
							  const int2* slotOfContact, int prevStep, int* prevSlot, hipStream_t stream );
hipError_t launch_mark_slots( const int* slotContact, const int* slotWorld, int slotCount, const int* contactBase, int step,
							  int2* slotOfContact, hipStream_t stream );
hipError_t launch_patch_params( const DevContact* src, DevContact* dst, const int* slotWorld, int slotCount, const float* params,
								hipStream_t stream );
hipError_t launch_fill_zero( float* p, size_t n, hipStream_t stream );
// poses between rows of 7 floats (the API's layout) and StepPose records
hipError_t launch_poses_from_rows( const float* rows, StepPose* poses, int n, hipStream_t stream );
hipError_t launch_rows_from_poses( const StepPose* poses, float* rows, int n, hipStream_t stream );

} // namespace b2mi
