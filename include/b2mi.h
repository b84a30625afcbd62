// Extensions of the MI355X step pipeline beyond the reference's public API.
//
// None of these has a counterpart in the reference (include/box2d/box2d.h); they are
// additions for three kinds of caller:
//  - inspection: tests and tools that compare contact/colour/manifold state with another
//    implementation of the same API (the dump records below are produced identically by
//    this library, by the CPU oracle and by the reference-side tools);
//  - batching: thousands of independent worlds advanced by one device launch per stage;
//  - differentiation: a device adjoint of the soft-step substep behind a ctypes binding.
#pragma once

#include "box2d/box2d.h"

#include <stdint.h>

// ---- inspection --------------------------------------------------------------------------

// One awake contact. Order of records: graph colours 0..11 (11 = overflow) in array order,
// then the awake non-touching set in array order.
typedef struct b2mi_ContactRecord
{
	int32_t contactId;
	int32_t shapeIdA;
	int32_t shapeIdB;
	int32_t colorIndex; // -1 for non-touching
	int32_t localIndex; // index inside its colour array or the non-touching array
	uint32_t simFlags;
	float friction;
	float restitution;
	b2Manifold manifold;
} b2mi_ContactRecord;

// One body, in body-id order (free slots skipped).
typedef struct b2mi_BodyRecord
{
	int32_t bodyId;
	int32_t type;
	int32_t setIndex;	// 0 static, 2 awake
	int32_t localIndex; // index in its solver set
	b2Transform transform;
	b2Vec2 center;
	b2Vec2 linearVelocity;
	float angularVelocity;
	float pad;
} b2mi_BodyRecord;

// Both return the number of records available; at most capacity are written.
B2_API int b2mi_World_DumpContacts( b2WorldId worldId, b2mi_ContactRecord* records, int capacity );
B2_API int b2mi_World_DumpBodies( b2WorldId worldId, b2mi_BodyRecord* records, int capacity );

// ---- batching ----------------------------------------------------------------------------

// A batch advances many independent worlds with one narrowphase launch and one solver
// launch per step. Returns a batch handle (> 0), or 0 when none is free.
B2_API int b2mi_CreateBatch( void );
B2_API void b2mi_DestroyBatch( int batch );

// Creates a world that belongs to the batch. Bodies and shapes are added with the regular
// API (b2CreateBody, b2CreatePolygonShape, ...). Batch worlds do not count against the
// 128-world limit of b2CreateWorld.
B2_API b2WorldId b2mi_Batch_CreateWorld( int batch, const b2WorldDef* def );
B2_API int b2mi_Batch_GetWorldCount( int batch );

// Steps every world of the batch (the batch counterpart of b2World_Step).
B2_API void b2mi_Batch_Step( int batch, float timeStep, int subStepCount );

// Solver launch shape: threads per workgroup (a multiple of 64 up to 1024, or 0 = each work item
// its own: the power of two at or above its largest colour group, items of one size launched
// together and the sizes side by side on separate streams) and the number of contact
// constraints packed into one work item (0 = 256).
B2_API void b2mi_Batch_Configure( int batch, int blockSize, int targetSlots );
B2_API void b2mi_World_Configure( b2WorldId worldId, int blockSize, int targetSlots );
B2_API int b2mi_Batch_GetWorkItemCount( int batch );

// Largest LDS footprint (bytes, at most 160 KiB less the 256 bytes the solver kernels keep for
// themselves; larger values are clamped) of a work item whose solver state is kept in LDS;
// larger work items keep it in HBM. -1 = that maximum, 0 = HBM for every work item.
// The results are bit-identical either way; only the speed differs.
B2_API void b2mi_Batch_SetLdsBudget( int batch, int maxBytes );
B2_API void b2mi_World_SetLdsBudget( b2WorldId worldId, int maxBytes );

// Largest LDS footprint (bytes) of a work item that keeps all its solver state in LDS (-1 = the
// LDS budget). Items above it but within the budget keep their bodies and impulses in LDS and
// read the constant constraint fields from HBM, a smaller footprint that lets more items share
// a CU. Results are bit-identical either way.
B2_API void b2mi_Batch_SetLdsFullBudget( int batch, int fullMaxBytes );

// World chunks (0 = one chunk; at most 3): the batch's narrowphase runs chunk after chunk and
// each chunk's solver starts as soon as its chunk's narrowphase is done, beside the next one.
// Results are bit-identical for every count.
B2_API void b2mi_Batch_SetWorldChunks( int batch, int chunks );

// A work item whose solver state is kept in HBM and that has at least `slots` constraint slots
// (one very large island, such as a 40,000-shape pile) runs over the whole GPU instead of in a
// single workgroup: one cooperative launch whose workgroups meet at a grid-wide barrier between
// dependent phases, or, on a device without cooperative launches, one launch per solver phase
// (stage, prepare, each substep's integrations and colour and overflow-level passes,
// restitution, store, finalize); results are the same bits either way. -1 = the default (4,096
// slots for a world stepped alone, 16,384 in a batch), 0 = never. Not in taped runs or chunked
// batches.
B2_API void b2mi_Batch_SetSplitSlots( int batch, int slots );
B2_API void b2mi_World_SetSplitSlots( b2WorldId worldId, int slots );

// Workgroups (of 256 threads) of such an item's cooperative launch: 0 = one thread per slot of
// its largest colour group, -1 = one launch per phase instead (the default: replayed as a launch
// graph it measured faster than the grid barrier on the base-100 pyramid); at most what the
// device holds at once. Results are the same bits for every count.
B2_API void b2mi_Batch_SetGridBlocks( int batch, int blocks );
B2_API void b2mi_World_SetGridBlocks( b2WorldId worldId, int blocks );

// Overflow levels (the serial chain of the constraints no colour could take, split into levels
// whose constraints share no body) of an item kept in HBM: consecutive levels of at most `slots`
// slots in all run as one run in one workgroup (in a grid-wide item, between two grid barriers or
// in one launch) with the bodies they touch and their records copied into LDS; a larger level
// runs on its own over all the item's threads. -1 = the default and largest value (128), 0 = no
// runs. Results are the same bits for every value.
B2_API void b2mi_Batch_SetLevelRunSlots( int batch, int slots );
B2_API void b2mi_World_SetLevelRunSlots( b2WorldId worldId, int slots );

// Bodies that continuous collision moved back to a time of impact (src/solver.c:993-1031), over
// all steps of the world's group: the world stepped alone, or its batch. No counterpart in the
// reference, which keeps no such count.
B2_API int b2mi_World_GetTimeOfImpactCount( b2WorldId worldId );

// Broadphase pair queries of a world stepped alone run on the device (one thread per moved
// proxy over uploaded copies of the trees, the contacts then created on the host in the
// reference's order) when at least `minMoves` proxies moved; -1 = always on the host, 0 = always
// on the device; default 2048. Results are the same bits either way.
B2_API void b2mi_World_SetDevicePairs( b2WorldId worldId, int minMoves );
// Accumulated over the device pair updates so far (9 values): host clock (ms since each update
// began) after the pair-set update, the uploads, the queries, the readback and the contact
// creation; then the pairs found, the moved proxies, the number of updates and the subtrees that
// queries over their visit budget handed to the workgroup pass.
B2_API void b2mi_World_GetDevicePairStats( b2WorldId worldId, float* stats );

// HIP graphs for the solver launches (on by default): when a step enqueues the same solver
// launches as the step before it, the sequence is captured as a graph and replayed while it stays
// the same. Results are the same bits either way. The stats give the captures and replays so far.
B2_API void b2mi_Batch_SetLaunchGraphs( int batch, int enabled );
B2_API void b2mi_World_SetLaunchGraphs( b2WorldId worldId, int enabled );
B2_API void b2mi_Batch_GetLaunchGraphStats( int batch, int* captures, int* replays );
B2_API void b2mi_World_GetLaunchGraphStats( b2WorldId worldId, int* captures, int* replays );

// Autotuner: steps the batch (or world) under each candidate launch shape (threads per
// workgroup, work-item packing target, LDS budget), trialSteps measured steps each after one
// rebuild step, then at the best shape under smaller limits (96, 64, 48, 32 KiB) on the
// footprint of an item that keeps all its state in LDS (b2mi_Batch_SetLdsFullBudget), and keeps
// the shape with the fastest solver kernel. Every shape gives bit-identical results, so these
// are ordinary simulation steps (5 x 3 x 4 + 4 candidates, i.e. 64 x (1 + trialSteps) steps).
// Returns the best solver-kernel time in milliseconds. GetLaunchShape reads the shape back
// (ldsFullBudget may be null).
B2_API float b2mi_Batch_Autotune( int batch, float timeStep, int subStepCount, int trialSteps );
B2_API float b2mi_World_Autotune( b2WorldId worldId, float timeStep, int subStepCount, int trialSteps );
B2_API void b2mi_Batch_GetLaunchShape( int batch, int* blockSize, int* targetSlots, int* ldsBudget, int* ldsFullBudget );

// Host clock of the last step, in ms since it began, at ten points: pair update started,
// collide lists started, narrowphase launched, narrowphase wait started, wait returned, solver
// setup started, solver launched, solver wait started, wait returned, enlarge processing done.
B2_API void b2mi_Batch_GetHostTimeline( int batch, float* marks );

// The solver launches of the last step, 4 ints each: first work item, work-item count, threads
// per workgroup, dynamic LDS bytes per workgroup. Returns the number of launches and writes at
// most `capacity` records.
B2_API int b2mi_Batch_GetSolverLaunches( int batch, int* records, int capacity );

// The work items of the last step's solver, in launch order, 4 ints each: constraint slots,
// bodies (kinematic shadows included), store (0 HBM, 1 bodies in LDS, 2 LDS) and how the item
// ran (0 one workgroup, 1 one grid launch per solver phase, 2 one cooperative grid launch).
// Returns the item count and writes at most `capacity` records.
B2_API int b2mi_Batch_GetWorkItems( int batch, int* records, int capacity );
B2_API int b2mi_World_GetWorkItems( b2WorldId worldId, int* records, int capacity );

// Device time of the last step's narrowphase and solver launches and the remaining host
// time, in milliseconds.
B2_API void b2mi_Batch_GetTiming( int batch, float* msCollide, float* msSolve, float* msHost );
B2_API void b2mi_World_GetTiming( b2WorldId worldId, float* msCollide, float* msSolve, float* msHost );

// ---- gradients of a batch run ------------------------------------------------------------
// A tape records every step of the batch from b2mi_Batch_BeginTape on: the solver's inputs of
// each step (work items, slot lists, a snapshot of every slot's contact with the impulses it
// inherited) and every checkpointInterval steps the poses of all bodies. b2mi_Batch_Backward
// then returns gradients of a loss through the whole taped run: it replays each checkpoint
// segment's solver steps and runs their hand-written adjoint in reverse. The narrowphase
// manifold of each step is a constant of that step (contact points, normals and the contact
// set do not move with the gradient). Requirements: dynamic bodies only, no more than 11
// graph colours (no overflow constraints), the same time step and substep count at every
// step, no bodies created or destroyed, no applied forces.
//
// Poses are 7 floats per body: center of mass (x, y), rotation (cos, sin), linear velocity
// (x, y), angular velocity; bodies in the order of b2mi_Batch_GetTapeBodies (world by world).
// One constraint slot of a taped step, as the solver saw it (export for checks and tools).
typedef struct b2mi_TapeSlot
{
	int world;		   // index of the world in the batch
	int contactId;	   // the contact's id in its world (stable across steps)
	int color;		   // graph colour (0..10); slots run colour by colour in ascending order
	int bodyA, bodyB;  // index in the tape body order, -1 for a static body
	int pointCount;
	uint32_t warmFrom; // per point (2 bits): the previous step's point it inherited impulses from, 3 = none
	float friction, restitution;
	float invMassA, invIA, invMassB, invIB;
	float normal[2];
	float anchorA[2][2], anchorB[2][2]; // relative to the body centers of mass
	float separation[2];
	float normalImpulse[2], tangentImpulse[2]; // inherited, before the warm-start scale
} b2mi_TapeSlot;

// Per-world step parameters of a taped run: h, 1/h, gravity x, gravity y, contact softness
// (bias rate, mass scale, impulse scale), static softness (same), restitution threshold,
// pushout velocity, max linear speed, max angular speed, warm-start scale, substep count.
#define B2MI_TAPE_WORLD_PARAMS 16

B2_API void b2mi_Batch_BeginTape( int batch, int checkpointInterval );
// Ends the tape. Its device memory stays with the batch for the next tape (freed with the batch).
B2_API void b2mi_Batch_EndTape( int batch );
B2_API int b2mi_Batch_GetTapeStepCount( int batch );
B2_API int b2mi_Batch_GetTapeBodies( int batch, b2BodyId* bodies, int capacity );
B2_API void b2mi_Batch_GetTapePoses( int batch, float* poses );
// The slots of taped step `step`, work item by work item (returns their count; writes them
// only when capacity holds them all).
B2_API int b2mi_Batch_GetTapeSlots( int batch, int step, b2mi_TapeSlot* slots, int capacity );
// B2MI_TAPE_WORLD_PARAMS floats per world.
B2_API void b2mi_Batch_GetTapeWorldParams( int batch, float* params );
// Runs the taped solver steps again from startPoses (the same manifolds and contact sets), with
// every contact's friction and restitution replaced by per-world values when worldParams
// (friction, restitution per world) is given; endPoses receives the final poses. This is the
// function b2mi_Batch_Backward differentiates.
B2_API void b2mi_Batch_ReplayTape( int batch, const float* startPoses, const float* worldParams, float* endPoses );
// b2mi_Batch_ReplayTape for callers that cannot allocate device memory (code under HIP graph
// capture, tight loops, harnesses that forbid hipMalloc); it has no counterpart in the
// reference. It computes the same bits as b2mi_Batch_ReplayTape, with every buffer the replay
// needs carved from one fixed device array of B2MI_REPLAY_SCRATCH_BYTES that each GPU holds
// until the process ends, and it calls no allocating routine. Returns 1, or 0 (through the assert
// callback and b2mi_GetLastError) when the taped run needs more than that array. Every call on one
// GPU uses that GPU's array: calls that overlap on one GPU race.
#define B2MI_REPLAY_SCRATCH_BYTES ( 32 * 1024 * 1024 )
B2_API int b2mi_Batch_ReplayTapeNoAlloc( int batch, const float* startPoses, const float* worldParams, float* endPoses );
// gradEnd: d(loss)/d(final poses); gradStart: d(loss)/d(poses when the tape began);
// gradParams: per world, d(loss)/d(friction) and d(loss)/d(restitution), where each contact's
// mixed coefficient counts as the world's parameter (exact when every shape of a world has the
// same friction, and restitution mixes to the bodies' value).
B2_API void b2mi_Batch_Backward( int batch, const float* gradEnd, float* gradStart, float* gradParams );

// ---- device ----------------------------------------------------------------------------
// The arc tangent the joint kernels use (the fdlibm float algorithm, so that joint angles round as
// the host C library's atan2f does), evaluated on the host over n argument pairs: out[i] =
// atan2(y[i], x[i]). No counterpart in the reference; for checking the kernels' angles.
B2_API void b2mi_Atan2( const float* y, const float* x, float* out, int n );

B2_API int b2mi_DeviceCount( void );
B2_API const char* b2mi_GetLastError( void );
