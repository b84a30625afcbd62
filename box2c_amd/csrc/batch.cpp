// Batched worlds: many independent worlds advanced together, one narrowphase launch and one
// solver launch per step for the whole batch. Each world keeps its own ids, trees, contact
// lists and colours exactly as a world created with b2CreateWorld would.
#include "world_internal.hpp"

#include <vector>

namespace b2mi
{
World* b2mi_internal_createWorld( const b2WorldDef* def, int firstSlot, int endSlot );
void b2mi_internal_destroyWorld( World* w );
const char* b2mi_internal_lastError();

constexpr int kMaxBatches = 256;
static StepGroup* g_batches[kMaxBatches] = {};
} // namespace b2mi

using namespace b2mi;

// The single-world counterpart of a batch: a world's own one-world step group.
static StepGroup* worldGroup( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	B2MI_CHECK( w->batchWorld == false );
	if ( w->group == nullptr )
	{
		w->group = new StepGroup();
		w->group->worlds.push_back( w );
		w->groupIndex = 0;
	}
	return w->group;
}

static void setShape( StepGroup* g, int blockSize, int targetSlots, int ldsBudget, int ldsFullBudget )
{
	g->blockSize = blockSize;
	g->targetSlots = targetSlots;
	g->ldsBudget = ldsBudget;
	g->ldsFullBudget = ldsFullBudget;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
}

// Autotuner of the solver launch shape. Every shape gives bit-identical results, so the trials
// are ordinary steps of the simulation: each candidate runs one step to rebuild its work items
// and then `trialSteps` measured steps; the candidate with the lowest solver-kernel time (device
// events) is kept. First every (threads per workgroup x work-item packing target x LDS budget)
// with all of an item's state in LDS whenever it fits the budget; then, at the best of those,
// smaller limits on the footprint of an item that keeps all its state in LDS (larger items keep
// only bodies and impulses there), which let more items share a CU.
static float autotune( StepGroup* g, float timeStep, int subStepCount, int trialSteps )
{
	static const int blocks[] = { 0, 128, 256, 512, 1024 };
	static const int targets[] = { 128, 256, 512 };
	static const int budgets[] = { 0, 64 * 1024, 96 * 1024, 160 * 1024 };
	static const int fullBudgets[] = { 96 * 1024, 64 * 1024, 48 * 1024, 32 * 1024 };
	float best = 1e30f;
	int bestBlock = 0, bestTarget = 0, bestBudget = -1, bestFull = -1;
	trialSteps = trialSteps < 1 ? 1 : trialSteps;
	auto trial = [&]( int block, int target, int budget, int full ) {
		setShape( g, block, target, budget, full );
		g->step( timeStep, subStepCount );
		float t = 1e30f;
		for ( int k = 0; k < trialSteps; ++k )
		{
			g->step( timeStep, subStepCount );
			t = g->msSolve < t ? g->msSolve : t;
		}
		if ( t < best )
		{
			best = t;
			bestBlock = block;
			bestTarget = target;
			bestBudget = budget;
			bestFull = full;
		}
	};
	for ( int budget : budgets )
		for ( int target : targets )
			for ( int block : blocks )
				trial( block, target, budget, -1 );
	const int budget = bestBudget;
	for ( int full : fullBudgets )
		if ( full < budget )
			trial( bestBlock, bestTarget, budget, full );
	setShape( g, bestBlock, bestTarget, bestBudget, bestFull );
	g->tunedBlock = bestBlock;
	g->tunedTarget = bestTarget;
	g->tunedBudget = bestBudget;
	g->tunedFullBudget = bestFull;
	return best;
}

static StepGroup* getBatch( int batch );

namespace b2mi
{
StepGroup* b2mi_internal_batch( int batch )
{
	return getBatch( batch );
}
} // namespace b2mi

static StepGroup* getBatch( int batch )
{
	B2MI_CHECK( 1 <= batch && batch <= kMaxBatches && g_batches[batch - 1] != nullptr );
	return g_batches[batch - 1];
}

extern "C" {

B2_API int b2mi_CreateBatch( void )
{
	for ( int i = 0; i < kMaxBatches; ++i )
		if ( g_batches[i] == nullptr )
		{
			g_batches[i] = new StepGroup();
			return i + 1;
		}
	return 0;
}

B2_API void b2mi_DestroyBatch( int batch )
{
	StepGroup* g = getBatch( batch );
	for ( World* w : g->worlds )
	{
		w->group = nullptr;
		b2mi_internal_destroyWorld( w );
	}
	g->worlds.clear();
	delete g;
	g_batches[batch - 1] = nullptr;
}

B2_API b2WorldId b2mi_Batch_CreateWorld( int batch, const b2WorldDef* def )
{
	StepGroup* g = getBatch( batch );
	World* w = b2mi_internal_createWorld( def, kMaxPlainWorlds, kMaxWorldSlots );
	if ( w == nullptr )
		return b2WorldId{ 0, 0 };
	w->batchWorld = true;
	w->group = g;
	w->mirrorEpoch = w->contactsEpoch = g->epoch;
	w->groupIndex = (int)g->worlds.size();
	g->worlds.push_back( w );
	g->structureDirty = true;
	g->allActive = true;
	return b2WorldId{ (uint16_t)( w->slot + 1 ), w->revision };
}

B2_API int b2mi_Batch_GetWorldCount( int batch )
{
	return (int)getBatch( batch )->worlds.size();
}

// blockSize: threads per workgroup (0 = per work item, from its largest colour group);
// targetSlots: islands are packed into one work item up to this many constraints (0 = 256).
B2_API void b2mi_Batch_Configure( int batch, int blockSize, int targetSlots )
{
	StepGroup* g = getBatch( batch );
	g->blockSize = blockSize;
	g->targetSlots = targetSlots;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
}

// maxBytes: largest LDS footprint of a work item kept in LDS (-1 = 160 KiB default, 0 = every
// work item keeps its solver state in HBM). Results are identical either way.
B2_API void b2mi_Batch_SetLdsBudget( int batch, int maxBytes )
{
	StepGroup* g = getBatch( batch );
	g->ldsBudget = maxBytes;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
}

// fullMaxBytes: largest LDS footprint of a work item that keeps all its solver state in LDS (-1 =
// the LDS budget); larger items within the budget keep bodies and impulses in LDS and read their
// constant constraint fields from HBM. Smaller footprints let more work items share a CU.
B2_API void b2mi_Batch_SetLdsFullBudget( int batch, int fullMaxBytes )
{
	StepGroup* g = getBatch( batch );
	g->ldsFullBudget = fullMaxBytes;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
}

// chunks: world chunks whose narrowphase and solver launches are pipelined (0 = one chunk).
// Results are identical for every count.
B2_API void b2mi_Batch_SetWorldChunks( int batch, int chunks )
{
	StepGroup* g = getBatch( batch );
	B2MI_CHECK( chunks >= 0 );
	g->chunksRequested = chunks;
	g->structureDirty = true;
	g->allActive = true;
}

B2_API void b2mi_Batch_Step( int batch, float timeStep, int subStepCount )
{
	StepGroup* g = getBatch( batch );
	for ( World* w : g->worlds )
		clearStepEvents( w );
	if ( timeStep == 0.0f || g->worlds.empty() )
		return;
	g->step( timeStep, subStepCount );
}

B2_API void b2mi_Batch_GetTiming( int batch, float* msCollide, float* msSolve, float* msHost )
{
	StepGroup* g = getBatch( batch );
	*msCollide = g->msCollide;
	*msSolve = g->msSolve;
	*msHost = g->msHost;
}

B2_API int b2mi_Batch_GetWorkItemCount( int batch )
{
	return getBatch( batch )->itemCount;
}

// the same solver settings for a single world (its own one-world group)
B2_API void b2mi_World_Configure( b2WorldId worldId, int blockSize, int targetSlots )
{
	StepGroup* g = worldGroup( worldId );
	setShape( g, blockSize, targetSlots, g->ldsBudget, g->ldsFullBudget );
}

// HBM-store work items of at least `slots` constraint slots run as one grid launch per solver
// phase instead of in one workgroup (-1 = the default, 0 = never)
static void setSplitSlots( StepGroup* g, int slots )
{
	g->splitSlots = slots;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
}

B2_API void b2mi_Batch_SetSplitSlots( int batch, int slots )
{
	setSplitSlots( getBatch( batch ), slots );
}

B2_API void b2mi_World_SetSplitSlots( b2WorldId worldId, int slots )
{
	setSplitSlots( worldGroup( worldId ), slots );
}

// workgroups of a very large item's cooperative launch (0 = automatic, -1 = one launch per phase)
static void setGridBlocks( StepGroup* g, int blocks )
{
	g->gridBlocks = blocks < 0 ? -1 : blocks;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
}

B2_API void b2mi_Batch_SetGridBlocks( int batch, int blocks )
{
	setGridBlocks( getBatch( batch ), blocks );
}

B2_API void b2mi_World_SetGridBlocks( b2WorldId worldId, int blocks )
{
	setGridBlocks( worldGroup( worldId ), blocks );
}

// HBM-store items run consecutive overflow levels of at most this many slots in all as one run
// with its bodies and records in LDS (-1 = the default, kLevelCacheSlots, which is also the
// largest; 0 = none)
static void setLevelRunSlots( StepGroup* g, int slots )
{
	g->levelRunSlots = slots < 0 ? kLevelCacheSlots : std::min( slots, kLevelCacheSlots );
	g->allActive = true;
}

B2_API void b2mi_Batch_SetLevelRunSlots( int batch, int slots )
{
	setLevelRunSlots( getBatch( batch ), slots );
}

B2_API void b2mi_World_SetLevelRunSlots( b2WorldId worldId, int slots )
{
	setLevelRunSlots( worldGroup( worldId ), slots );
}

// bodies that continuous collision moved back to a time of impact, over all steps of the world's
// group (the world alone, or its batch)
B2_API int b2mi_World_GetTimeOfImpactCount( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	return w->group != nullptr ? w->group->toiHits : 0;
}

B2_API void b2mi_World_GetDevicePairStats( b2WorldId worldId, float* stats )
{
	World* w = getWorld( worldId );
	for ( int i = 0; i < 9; ++i )
		stats[i] = w->devicePairs.phaseMs[i];
}

B2_API void b2mi_World_SetDevicePairs( b2WorldId worldId, int minMoves )
{
	getWorld( worldId )->devicePairsMin = minMoves < 0 ? -1 : minMoves;
}

B2_API void b2mi_Batch_SetLaunchGraphs( int batch, int enabled )
{
	getBatch( batch )->launchGraphs = enabled != 0;
}

B2_API void b2mi_World_SetLaunchGraphs( b2WorldId worldId, int enabled )
{
	worldGroup( worldId )->launchGraphs = enabled != 0;
}

static void graphStats( const StepGroup* g, int* captures, int* replays )
{
	*captures = g ? g->graphCaptures : 0;
	*replays = g ? g->graphReplays : 0;
}

B2_API void b2mi_Batch_GetLaunchGraphStats( int batch, int* captures, int* replays )
{
	graphStats( getBatch( batch ), captures, replays );
}

B2_API void b2mi_World_GetLaunchGraphStats( b2WorldId worldId, int* captures, int* replays )
{
	graphStats( getWorld( worldId )->group, captures, replays );
}

B2_API void b2mi_World_SetLdsBudget( b2WorldId worldId, int maxBytes )
{
	StepGroup* g = worldGroup( worldId );
	setShape( g, g->blockSize, g->targetSlots, maxBytes, g->ldsFullBudget );
}

B2_API float b2mi_World_Autotune( b2WorldId worldId, float timeStep, int subStepCount, int trialSteps )
{
	StepGroup* g = worldGroup( worldId );
	B2MI_CHECK( timeStep > 0.0f && subStepCount > 0 );
	return autotune( g, timeStep, subStepCount, trialSteps );
}

B2_API float b2mi_Batch_Autotune( int batch, float timeStep, int subStepCount, int trialSteps )
{
	StepGroup* g = getBatch( batch );
	B2MI_CHECK( timeStep > 0.0f && subStepCount > 0 );
	if ( g->worlds.empty() )
		return 0.0f;
	return autotune( g, timeStep, subStepCount, trialSteps );
}

// The solver launches of the last step: per launch, (first item, item count, threads per
// workgroup, dynamic LDS bytes). Returns the launch count; writes at most `capacity` records.
B2_API int b2mi_Batch_GetSolverLaunches( int batch, int* records, int capacity )
{
	StepGroup* g = getBatch( batch );
	int n = 0;
	auto put = [&]( int first, int count, int block, int lds ) {
		if ( records != nullptr && n < capacity )
		{
			records[4 * n + 0] = first;
			records[4 * n + 1] = count;
			records[4 * n + 2] = block;
			records[4 * n + 3] = lds;
		}
		n += 1;
	};
	for ( const LaunchClass& c : g->classes )
		put( c.first, c.count, g->blockSize > 0 ? g->blockSize : c.block, g->blockSize > 0 ? g->maxItemLds : c.lds );
	return n;
}

// The work items of the last solver launch, in launch order: per item (constraint slots, contacts
// and joints, bodies
// including kinematic shadows, store, how it ran: 0 one workgroup, 1 one grid launch per phase,
// 2 one cooperative grid launch).
static int workItems( StepGroup* g, int* records, int capacity )
{
	std::vector<int> mode( g->hItems.size(), 0 );
	for ( int i : g->splitItems )
		mode[i] = g->splitPerPhase ? 1 : 2;
	int n = 0;
	for ( size_t i = 0; i < g->hItems.size(); ++i )
	{
		const WorkItem& wi = g->hItems[i];
		if ( records != nullptr && n < capacity )
		{
			records[4 * n + 0] = item_slots( wi );
			records[4 * n + 1] = wi.bodyCount;
			records[4 * n + 2] = wi.store;
			records[4 * n + 3] = mode[i];
		}
		n += 1;
	}
	return n;
}

B2_API int b2mi_Batch_GetWorkItems( int batch, int* records, int capacity )
{
	return workItems( getBatch( batch ), records, capacity );
}

B2_API int b2mi_World_GetWorkItems( b2WorldId worldId, int* records, int capacity )
{
	World* w = getWorld( worldId );
	return w->group != nullptr ? workItems( w->group, records, capacity ) : 0;
}

// Host clock of the last step (ms since it began) at ten points: pair update started, collide
// lists started, narrowphase launched, narrowphase wait started, wait returned, solver setup
// started, solver launched, solver wait started, wait returned, enlarge processing done.
B2_API void b2mi_Batch_GetHostTimeline( int batch, float* marks )
{
	StepGroup* g = getBatch( batch );
	for ( int i = 0; i < 10; ++i )
		marks[i] = g->hostPhases[i];
}

B2_API void b2mi_Batch_GetLaunchShape( int batch, int* blockSize, int* targetSlots, int* ldsBudget, int* ldsFullBudget )
{
	StepGroup* g = getBatch( batch );
	*blockSize = g->blockSize;
	*targetSlots = g->targetSlots;
	*ldsBudget = g->ldsBudget;
	if ( ldsFullBudget != nullptr )
		*ldsFullBudget = g->ldsFullBudget;
}

B2_API void b2mi_World_GetTiming( b2WorldId worldId, float* msCollide, float* msSolve, float* msHost )
{
	World* w = getWorld( worldId );
	*msCollide = *msSolve = *msHost = 0.0f;
	if ( w->group )
	{
		*msCollide = w->group->msCollide;
		*msSolve = w->group->msSolve;
		*msHost = w->group->msHost;
	}
}

B2_API int b2mi_DeviceCount( void )
{
	int count = 0;
	if ( hipGetDeviceCount( &count ) != hipSuccess )
		return 0;
	return count;
}

B2_API const char* b2mi_GetLastError( void )
{
	return b2mi_internal_lastError();
}

} // extern "C"
