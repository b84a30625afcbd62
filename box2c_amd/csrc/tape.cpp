// Taped runs and the backward pass (the host side of adjoint.hip).
//
// While a tape is on, every step of the group records, right before its solver launch: the
// work items and slot lists (a new copy only when they were rebuilt), a snapshot of every
// slot's contact, and every `interval` steps a checkpoint of all body poses. The backward pass
// rebuilds the step-start poses of one checkpoint segment at a time by replaying the solver,
// then runs the adjoint of each step of the segment from the last to the first.
#include "b2mi.h"
#include "launch.hpp"
#include "world_internal.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace b2mi
{

struct TapeVersion
{
	int itemsVersion = -1;
	int itemCount = 0, slotCount = 0, bodyListCount = 0, block = 64;
	int usedSlots = 0; // slots that belong to work items
	bool adjointsInLds = true; // every item's solver state and adjoints fit the LDS limits
	int storeLdsBytes = 0;	   // LDS of the largest item's solver state (when adjointsInLds)
	WorkItem* dItems = nullptr;
	int* dBodyList = nullptr;
	int* dSlotContact = nullptr;
	int2* dSlotBodies = nullptr;
	int* dSlotWorld = nullptr;
	int* dScratchOffset = nullptr;
	size_t scratchFloats = 0;
	// host sources of the version's uploads, kept until the tape ends so the copies need no wait
	std::vector<int> hOffsets, hSlotWorld;
};

struct TapeStep
{
	int version;
	DevContact* dSnap;
	int checkpoint; // index into checkpoints, or -1
};

// Device memory of a step group's tapes: chunks of at least 64 MiB handed out in order
// (256-byte aligned). Ending a tape rewinds the arena without freeing it, so a recorded step
// allocates nothing once the chunks cover a run's length; the chunks are freed with the group.
struct TapeArena
{
	struct Chunk
	{
		char* base;
		size_t size;
	};
	struct Mark
	{
		size_t current, used;
	};
	static constexpr size_t kChunkBytes = (size_t)64 << 20;
	std::vector<Chunk> chunks;
	size_t current = 0, used = 0; // chunk in use and the bytes of it handed out

	template <typename T>
	T* take( size_t n )
	{
		const size_t bytes = ( sizeof( T ) * std::max<size_t>( n, 1 ) + 255 ) & ~(size_t)255;
		while ( current < chunks.size() && used + bytes > chunks[current].size )
		{
			current += 1;
			used = 0;
		}
		if ( current == chunks.size() )
		{
			Chunk c;
			// each new chunk at least doubles the arena: a long run takes few allocations
			size_t held = 0;
			for ( const Chunk& k : chunks )
				held += k.size;
			c.size = std::max( { bytes, kChunkBytes, held } );
			HIP_CHECK( hipMalloc( (void**)&c.base, c.size ) );
			chunks.push_back( c );
			used = 0;
		}
		T* p = reinterpret_cast<T*>( chunks[current].base + used );
		used += bytes;
		return p;
	}
	Mark mark() const
	{
		return Mark{ current, used };
	}
	void rewind( Mark m )
	{
		current = m.current;
		used = m.used;
	}
	void release()
	{
		for ( Chunk& c : chunks )
			(void)hipFree( c.base );
		chunks.clear();
		current = used = 0;
	}
};

// what one recorded step added, so a speculative record that the step voided can be taken back
struct TapeUndo
{
	TapeArena::Mark mark;
	size_t versions, checkpoints, steps;
};

struct Tape
{
	TapeUndo undo{};
	TapeArena* arena = nullptr; // the group's, rewound when the tape ends
	int interval = 10;
	float timeStep = 0.0f;
	int subSteps = 0;
	int savedTarget = 0;
	std::vector<int> simBase, simCount;
	int totalSims = 0, maxSims = 0;
	int* dSimBase = nullptr;
	int* dSimCount = nullptr;
	std::vector<TapeVersion> versions;
	std::vector<TapeStep> steps;
	std::vector<StepPose*> checkpoints;
	std::vector<b2BodyId> bodies; // tape order: world by world, awake-set order
};

void tapeFree( StepGroup* g )
{
	Tape* t = g->tape;
	if ( t == nullptr )
		return;
	// every kernel that reads the tape's memory has finished before the arena hands it out again
	(void)hipStreamSynchronize( stream() );
	t->arena->rewind( TapeArena::Mark{ 0, 0 } );
	g->targetSlots = t->savedTarget;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
	delete t;
	g->tape = nullptr;
}

void tapeRelease( StepGroup* g )
{
	tapeFree( g );
	if ( g->tapeArena )
	{
		g->tapeArena->release();
		delete g->tapeArena;
		g->tapeArena = nullptr;
	}
}

static void tapeBegin( StepGroup* g, int interval )
{
	tapeFree( g );
	Tape* t = new Tape();
	if ( g->tapeArena == nullptr )
		g->tapeArena = new TapeArena();
	t->arena = g->tapeArena;
	t->interval = interval > 0 ? interval : 10;
	t->savedTarget = g->targetSlots;
	for ( World* w : g->worlds )
	{
		t->simBase.push_back( t->totalSims );
		int n = (int)w->awakeSimBody.size();
		t->simCount.push_back( n );
		t->totalSims += n;
		t->maxSims = std::max( t->maxSims, n );
		for ( int i = 0; i < n; ++i )
		{
			const Body& b = w->bodies[w->awakeSimBody[i]];
			if ( b.type != b2_dynamicBody )
				fatal( "a taped run supports dynamic bodies only (no kinematic bodies)", __FILE__, __LINE__ );
			t->bodies.push_back( b2BodyId{ b.id + 1, (uint16_t)w->slot, b.revision } );
		}
	}
	hipStream_t s = stream();
	t->dSimBase = t->arena->take<int>( t->simBase.size() );
	t->dSimCount = t->arena->take<int>( t->simCount.size() );
	if ( !t->simBase.empty() )
	{
		HIP_CHECK( hipMemcpyAsync( t->dSimBase, t->simBase.data(), sizeof( int ) * t->simBase.size(), hipMemcpyHostToDevice, s ) );
		HIP_CHECK( hipMemcpyAsync( t->dSimCount, t->simCount.data(), sizeof( int ) * t->simCount.size(), hipMemcpyHostToDevice, s ) );
	}
	// one work item per world: the adjoint sums each world's parameter gradients in one workgroup
	g->targetSlots = 1 << 24;
	for ( World* w : g->worlds )
		w->topologyDirty = true;
	g->allActive = true;
	g->tape = t;
}

void tapeRecordStep( StepGroup* g, float timeStep, int subStepCount, int blockSize )
{
	Tape* t = g->tape;
	t->undo = TapeUndo{ t->arena->mark(), t->versions.size(), t->checkpoints.size(), t->steps.size() };
	hipStream_t s = stream();
	if ( t->steps.empty() )
	{
		t->timeStep = timeStep;
		t->subSteps = subStepCount;
	}
	else if ( t->timeStep != timeStep || t->subSteps != subStepCount )
		fatal( "a taped run needs the same time step and substep count at every step", __FILE__, __LINE__ );
	for ( size_t k = 0; k < g->worlds.size(); ++k )
		if ( (int)g->worlds[k]->awakeSimBody.size() != t->simCount[k] )
			fatal( "bodies were created or destroyed during a taped run", __FILE__, __LINE__ );

	if ( t->versions.empty() || t->versions.back().itemsVersion != g->itemsVersion )
	{
		TapeVersion v;
		v.itemsVersion = g->itemsVersion;
		// an LDS budget of 0 keeps the adjoint's state in HBM as well
		v.adjointsInLds = g->ldsBudget != 0;
		v.itemCount = g->itemCount;
		v.slotCount = g->slotCount;
		v.bodyListCount = g->bodyListCount;
		v.block = std::min( blockSize, kAdjointMaxBlock );
		std::vector<int> itemsOfWorld( g->worlds.size(), 0 );
		// world of every slot of the launch arrays; -1 for the unused slots between the worlds'
		// regions, whose device contents are undefined
		std::vector<int> slotWorld( std::max( g->slotCount, 1 ), -1 );
		std::vector<int> offsets( std::max( g->itemCount, 1 ), 0 );
		size_t scratch = 0;
		for ( int i = 0; i < g->itemCount; ++i )
		{
			const WorkItem& wi = g->hItems[i];
			if ( wi.groupStart[1] != wi.groupStart[0] )
				fatal( "a taped run does not support overflow constraints (more than 11 graph colours)", __FILE__, __LINE__ );
			if ( wi.ownedCount != wi.bodyCount )
				fatal( "a taped run does not support kinematic bodies", __FILE__, __LINE__ );
			if ( item_slots( wi ) != wi.groupStart[kGroups] )
				fatal( "a taped run does not support joints", __FILE__, __LINE__ );
			itemsOfWorld[wi.world] += 1;
			v.usedSlots += wi.groupStart[kGroups] - wi.groupStart[1];
			const int storeBytes = 40 * ( wi.bodyCount + 1 ) + 16 * cfQuads * wi.groupStart[kGroups];
			v.adjointsInLds = v.adjointsInLds && 12 * wi.bodyCount + 8 * wi.groupStart[kGroups] <= kAdjointLdsFloats &&
							  storeBytes <= kAdjointStoreLdsBytes;
			v.storeLdsBytes = std::max( v.storeLdsBytes, storeBytes );
			for ( int k = 0; k < wi.groupStart[kGroups]; ++k )
				slotWorld[wi.slotStart + k] = wi.world;
			B2MI_CHECK( scratch < ( (size_t)1 << 31 ) );
			offsets[i] = (int)scratch;
			scratch += (size_t)adjoint_scratch_floats( wi.bodyCount, wi.groupStart[kGroups], subStepCount );
		}
		v.storeLdsBytes = v.adjointsInLds ? std::max( v.storeLdsBytes, 16 ) : 0;
		for ( int n : itemsOfWorld )
			B2MI_CHECK( n <= 1 );
		v.scratchFloats = scratch;
		v.dItems = t->arena->take<WorkItem>( v.itemCount );
		v.dBodyList = t->arena->take<int>( v.bodyListCount );
		v.dSlotContact = t->arena->take<int>( v.slotCount );
		v.dSlotBodies = t->arena->take<int2>( v.slotCount );
		v.dSlotWorld = t->arena->take<int>( v.slotCount );
		v.dScratchOffset = t->arena->take<int>( v.itemCount );
		v.hOffsets = std::move( offsets );
		v.hSlotWorld = std::move( slotWorld );
		if ( v.itemCount )
		{
			HIP_CHECK( hipMemcpyAsync( v.dItems, g->dItems, sizeof( WorkItem ) * v.itemCount, hipMemcpyDeviceToDevice, s ) );
			HIP_CHECK( hipMemcpyAsync( v.dScratchOffset, v.hOffsets.data(), sizeof( int ) * v.itemCount, hipMemcpyHostToDevice, s ) );
		}
		if ( v.bodyListCount )
			HIP_CHECK( hipMemcpyAsync( v.dBodyList, g->dBodyList, sizeof( int ) * v.bodyListCount, hipMemcpyDeviceToDevice, s ) );
		if ( v.slotCount )
		{
			HIP_CHECK( hipMemcpyAsync( v.dSlotContact, g->dSlotContact, sizeof( int ) * v.slotCount, hipMemcpyDeviceToDevice, s ) );
			HIP_CHECK( hipMemcpyAsync( v.dSlotBodies, g->dSlotBodies, sizeof( int2 ) * v.slotCount, hipMemcpyDeviceToDevice, s ) );
			HIP_CHECK( hipMemcpyAsync( v.dSlotWorld, v.hSlotWorld.data(), sizeof( int ) * v.slotCount, hipMemcpyHostToDevice, s ) );
		}
		t->versions.push_back( std::move( v ) );
	}

	TapeStep st;
	st.version = (int)t->versions.size() - 1;
	st.checkpoint = -1;
	if ( t->steps.size() % (size_t)t->interval == 0 )
	{
		StepPose* cp = t->arena->take<StepPose>( t->totalSims );
		HIP_CHECK( launch_gather_poses( g->dWorlds, t->dSimBase, t->dSimCount, (int)g->worlds.size(), t->maxSims, cp, s ) );
		st.checkpoint = (int)t->checkpoints.size();
		t->checkpoints.push_back( cp );
	}
	st.dSnap = t->arena->take<DevContact>( g->slotCount );
	HIP_CHECK( launch_snapshot( g->dWorlds, g->dItems, g->itemCount, g->dSlotContact, st.dSnap, s ) );
	t->steps.push_back( st );
}

void tapeUndoStep( StepGroup* g )
{
	Tape* t = g->tape;
	const TapeUndo& u = t->undo;
	B2MI_CHECK( t->steps.size() == u.steps + 1 );
	// the discarded records' kernels are ahead of any later use of their memory on the stream
	t->steps.resize( u.steps );
	t->checkpoints.resize( u.checkpoints );
	t->versions.resize( u.versions );
	t->arena->rewind( u.mark );
}

// d(loss)/d(start poses) and d(loss)/d(friction, restitution per world) from d(loss)/d(end poses)
static void tapeBackward( StepGroup* g, const float* gradEnd, float* gradStart, float* gradParams )
{
	Tape* t = g->tape;
	hipStream_t s = stream();
	const int T = (int)t->steps.size();
	const int n = t->totalSims;
	const int W = (int)g->worlds.size();
	if ( T == 0 )
	{
		for ( int i = 0; i < 7 * n; ++i )
			gradStart[i] = gradEnd[i];
		for ( int i = 0; i < 2 * W; ++i )
			gradParams[i] = 0.0f;
		return;
	}
	HIP_CHECK( hipStreamSynchronize( s ) );

	int maxSlots = 1, maxBodies = 1;
	size_t maxScratch = 1;
	for ( const TapeVersion& v : t->versions )
	{
		maxSlots = std::max( maxSlots, v.slotCount );
		maxBodies = std::max( maxBodies, v.bodyListCount );
		maxScratch = std::max( maxScratch, v.scratchFloats );
	}
	std::vector<int> contactBase( W );
	size_t contactTotal = 0;
	for ( int k = 0; k < W; ++k )
	{
		contactBase[k] = (int)contactTotal;
		contactTotal += (size_t)std::max( g->worlds[k]->capContacts, 1 );
	}

	const int K = t->interval;
	// working memory of the pass from the tape's arena, taken back when the results are home
	const TapeArena::Mark scratchMark = t->arena->mark();
	StepPose* dAdj = t->arena->take<StepPose>( n );
	StepPose* dStates = t->arena->take<StepPose>( (size_t)( K + 1 ) * n );
	float* dParam = t->arena->take<float>( 2 * W );
	float* dCf = t->arena->take<float>( (size_t)cfFloatCount * maxSlots );
	float4* dPriv = t->arena->take<float4>( 2 * (size_t)maxBodies );
	float2* dMass = t->arena->take<float2>( maxBodies );
	float* dScratch = t->arena->take<float>( maxScratch );
	float* dLamA = t->arena->take<float>( 4 * (size_t)maxSlots );
	float* dLamB = t->arena->take<float>( 4 * (size_t)maxSlots );
	int* dPrevSlot = t->arena->take<int>( maxSlots );
	int* dContactBase = t->arena->take<int>( W );
	int2* dSlotOf = t->arena->take<int2>( contactTotal );

	std::vector<StepPose> h( n );
	for ( int i = 0; i < n; ++i )
	{
		const float* e = gradEnd + 7 * i;
		h[i] = StepPose{ e[0], e[1], e[2], e[3], e[4], e[5], e[6], 0.0f };
	}
	HIP_CHECK( hipMemcpyAsync( dAdj, h.data(), sizeof( StepPose ) * n, hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipMemcpyAsync( dContactBase, contactBase.data(), sizeof( int ) * W, hipMemcpyHostToDevice, s ) );
	HIP_CHECK( launch_fill_zero( dParam, 2 * (size_t)W, s ) );
	HIP_CHECK( launch_fill_zero( reinterpret_cast<float*>( dSlotOf ), 2 * contactTotal, s ) );
	HIP_CHECK( launch_fill_zero( dLamA, 4 * (size_t)maxSlots, s ) );

	float* lamOut = dLamA;
	float* lamPrev = dLamB;
	for ( int c = ( ( T - 1 ) / K ) * K; c >= 0; c -= K )
	{
		int end = std::min( c + K, T );
		const TapeStep& first = t->steps[c];
		B2MI_CHECK( first.checkpoint >= 0 );
		HIP_CHECK( hipMemcpyAsync( dStates, t->checkpoints[first.checkpoint], sizeof( StepPose ) * n, hipMemcpyDeviceToDevice, s ) );
		for ( int step = c; step < end - 1; ++step )
		{
			const TapeStep& ts = t->steps[step];
			const TapeVersion& v = t->versions[ts.version];
			HIP_CHECK( launch_replay_step( g->dWorlds, v.dItems, v.itemCount, v.block, v.dBodyList, v.dSlotBodies, ts.dSnap,
										   dStates + (size_t)( step - c ) * n, dStates + (size_t)( step - c + 1 ) * n, t->dSimBase, dCf,
										   maxSlots, dPriv, dMass, v.storeLdsBytes, s ) );
		}
		for ( int step = end - 1; step >= c; --step )
		{
			const TapeStep& ts = t->steps[step];
			const TapeVersion& v = t->versions[ts.version];
			// slots of the previous step, tagged step (= previous step index + 1), and the links to them
			int prevTag = -1;
			if ( step > 0 )
			{
				const TapeVersion& vp = t->versions[t->steps[step - 1].version];
				HIP_CHECK( launch_mark_slots( vp.dSlotContact, vp.dSlotWorld, vp.slotCount, dContactBase, step, dSlotOf, s ) );
				HIP_CHECK( launch_fill_zero( lamPrev, 4 * (size_t)vp.slotCount, s ) );
				prevTag = step;
			}
			HIP_CHECK( launch_link_slots( v.dSlotContact, v.dSlotWorld, v.slotCount, dContactBase, dSlotOf, prevTag, dPrevSlot, s ) );
			HIP_CHECK( launch_adjoint_step( g->dWorlds, v.dItems, v.itemCount, v.block, v.dBodyList, v.dSlotBodies, ts.dSnap,
											dStates + (size_t)( step - c ) * n, dAdj, t->dSimBase, reinterpret_cast<const float4*>( lamOut ),
											lamPrev, dPrevSlot, dParam, dCf, maxSlots, dPriv, dMass, dScratch, v.dScratchOffset,
											v.storeLdsBytes, s ) );
			std::swap( lamOut, lamPrev );
		}
	}
	HIP_CHECK( hipMemcpyAsync( h.data(), dAdj, sizeof( StepPose ) * n, hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipMemcpyAsync( gradParams, dParam, sizeof( float ) * 2 * W, hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	for ( int i = 0; i < n; ++i )
	{
		float* o = gradStart + 7 * i;
		o[0] = h[i].cx, o[1] = h[i].cy, o[2] = h[i].qc, o[3] = h[i].qs, o[4] = h[i].vx, o[5] = h[i].vy, o[6] = h[i].w;
	}
	t->arena->rewind( scratchMark );
}

// The taped solver steps again from the given start poses (n x 7), optionally with every
// contact's friction and restitution replaced by per-world values: the function whose
// derivatives the adjoint computes (same manifolds and contact sets).
static void tapeReplay( StepGroup* g, const float* startPoses, const float* worldParams, float* endPoses )
{
	Tape* t = g->tape;
	hipStream_t s = stream();
	const int T = (int)t->steps.size();
	const int n = t->totalSims;
	const int W = (int)g->worlds.size();
	int maxSlots = 1, maxBodies = 1;
	for ( const TapeVersion& v : t->versions )
	{
		maxSlots = std::max( maxSlots, v.slotCount );
		maxBodies = std::max( maxBodies, v.bodyListCount );
	}
	const TapeArena::Mark scratchMark = t->arena->mark();
	StepPose* dA = t->arena->take<StepPose>( n );
	StepPose* dB = t->arena->take<StepPose>( n );
	float* dCf = t->arena->take<float>( (size_t)cfFloatCount * maxSlots );
	float4* dPriv = t->arena->take<float4>( 2 * (size_t)maxBodies );
	float2* dMass = t->arena->take<float2>( maxBodies );
	DevContact* dPatched = t->arena->take<DevContact>( maxSlots );
	float* dParams = t->arena->take<float>( 2 * W );
	std::vector<StepPose> h( n );
	for ( int i = 0; i < n; ++i )
	{
		const float* e = startPoses + 7 * i;
		h[i] = StepPose{ e[0], e[1], e[2], e[3], e[4], e[5], e[6], 0.0f };
	}
	HIP_CHECK( hipMemcpyAsync( dA, h.data(), sizeof( StepPose ) * n, hipMemcpyHostToDevice, s ) );
	if ( worldParams )
		HIP_CHECK( hipMemcpyAsync( dParams, worldParams, sizeof( float ) * 2 * W, hipMemcpyHostToDevice, s ) );
	for ( int step = 0; step < T; ++step )
	{
		const TapeStep& ts = t->steps[step];
		const TapeVersion& v = t->versions[ts.version];
		const DevContact* snap = ts.dSnap;
		if ( worldParams )
		{
			HIP_CHECK( launch_patch_params( ts.dSnap, dPatched, v.dSlotWorld, v.slotCount, dParams, s ) );
			snap = dPatched;
		}
		HIP_CHECK( launch_replay_step( g->dWorlds, v.dItems, v.itemCount, v.block, v.dBodyList, v.dSlotBodies, snap, dA, dB, t->dSimBase,
									   dCf, maxSlots, dPriv, dMass, v.storeLdsBytes, s ) );
		std::swap( dA, dB );
	}
	HIP_CHECK( hipMemcpyAsync( h.data(), dA, sizeof( StepPose ) * n, hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	for ( int i = 0; i < n; ++i )
	{
		float* o = endPoses + 7 * i;
		o[0] = h[i].cx, o[1] = h[i].cy, o[2] = h[i].qc, o[3] = h[i].qs, o[4] = h[i].vx, o[5] = h[i].vy, o[6] = h[i].w;
	}
	t->arena->rewind( scratchMark );
}

// ---- the replay without allocation --------------------------------------------------------
// b2mi_Batch_ReplayTapeNoAlloc keeps everything tapeReplay allocates per call (pose buffers,
// constraint records, private body state, masses, patched contacts, parameters) in one fixed
// array of device memory, `scratch`. Each GPU holds its own copy of it until the process ends;
// its device address is looked up once per device and only that address is used on the host.
// Every call on a GPU reuses that GPU's array, so calls that overlap on one GPU (from several
// host threads, or on streams that are not ordered) race on it.
constexpr size_t kReplayScratchBytes = size_t( B2MI_REPLAY_SCRATCH_BYTES );
static __device__ __align__( 256 ) unsigned char scratch[kReplayScratchBytes];

static unsigned char* replayScratch()
{
	static std::mutex lock;
	static std::vector<void*> perDevice;
	int device = 0;
	HIP_CHECK( hipGetDevice( &device ) );
	std::lock_guard<std::mutex> guard( lock );
	if ( (int)perDevice.size() <= device )
		perDevice.resize( device + 1, nullptr );
	if ( perDevice[device] == nullptr )
	{
		void* p = nullptr;
		hipError_t e = hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) );
		if ( e != hipSuccess || p == nullptr )
		{
			fatal( "the replay scratch array has no device address", __FILE__, __LINE__ );
			return nullptr;
		}
		perDevice[device] = p;
	}
	return static_cast<unsigned char*>( perDevice[device] );
}

// the scratch layout of a replay: byte offsets of its buffers, each 256-byte aligned
struct ReplayLayout
{
	size_t rows, poseA, poseB, cf, priv, mass, patched, params, total;
};

static ReplayLayout replayLayout( const Tape* t, int worldCount, bool patched )
{
	int maxSlots = 1, maxBodies = 1;
	for ( const TapeVersion& v : t->versions )
	{
		maxSlots = std::max( maxSlots, v.slotCount );
		maxBodies = std::max( maxBodies, v.bodyListCount );
	}
	const size_t n = (size_t)std::max( t->totalSims, 1 );
	ReplayLayout L{};
	size_t at = 0;
	auto take = [&]( size_t bytes ) {
		size_t p = at;
		at += ( bytes + 255 ) / 256 * 256;
		return p;
	};
	L.rows = take( sizeof( float ) * 7 * n );
	L.poseA = take( sizeof( StepPose ) * n );
	L.poseB = take( sizeof( StepPose ) * n );
	L.cf = take( sizeof( float ) * cfFloatCount * (size_t)maxSlots );
	L.priv = take( sizeof( float4 ) * 2 * (size_t)maxBodies );
	L.mass = take( sizeof( float2 ) * (size_t)maxBodies );
	L.patched = patched ? take( sizeof( DevContact ) * (size_t)maxSlots ) : 0;
	L.params = take( sizeof( float ) * 2 * (size_t)std::max( worldCount, 1 ) );
	L.total = at;
	return L;
}

// tapeReplay's steps on buffers carved from the scratch array; no allocation of any kind
static bool tapeReplayNoAlloc( StepGroup* g, const float* startPoses, const float* worldParams, float* endPoses )
{
	Tape* t = g->tape;
	const int W = (int)g->worlds.size();
	const ReplayLayout L = replayLayout( t, W, worldParams != nullptr );
	if ( L.total > kReplayScratchBytes )
	{
		fatal( "b2mi_Batch_ReplayTapeNoAlloc: the taped run needs more scratch than B2MI_REPLAY_SCRATCH_BYTES; use "
			   "b2mi_Batch_ReplayTape",
			   __FILE__, __LINE__ );
		return false;
	}
	unsigned char* base = replayScratch();
	if ( base == nullptr )
		return false;
	hipStream_t s = stream();
	const int T = (int)t->steps.size();
	const int n = t->totalSims;
	int maxSlots = 1;
	for ( const TapeVersion& v : t->versions )
		maxSlots = std::max( maxSlots, v.slotCount );
	float* rows = reinterpret_cast<float*>( base + L.rows );
	StepPose* dA = reinterpret_cast<StepPose*>( base + L.poseA );
	StepPose* dB = reinterpret_cast<StepPose*>( base + L.poseB );
	float* dCf = reinterpret_cast<float*>( base + L.cf );
	float4* dPriv = reinterpret_cast<float4*>( base + L.priv );
	float2* dMass = reinterpret_cast<float2*>( base + L.mass );
	DevContact* dPatched = worldParams ? reinterpret_cast<DevContact*>( base + L.patched ) : nullptr;
	float* dParams = reinterpret_cast<float*>( base + L.params );
	if ( n > 0 )
		HIP_CHECK( hipMemcpyAsync( rows, startPoses, sizeof( float ) * 7 * (size_t)n, hipMemcpyHostToDevice, s ) );
	HIP_CHECK( launch_poses_from_rows( rows, dA, n, s ) );
	if ( worldParams )
		HIP_CHECK( hipMemcpyAsync( dParams, worldParams, sizeof( float ) * 2 * W, hipMemcpyHostToDevice, s ) );
	for ( int step = 0; step < T; ++step )
	{
		const TapeStep& ts = t->steps[step];
		const TapeVersion& v = t->versions[ts.version];
		const DevContact* snap = ts.dSnap;
		if ( worldParams )
		{
			HIP_CHECK( launch_patch_params( ts.dSnap, dPatched, v.dSlotWorld, v.slotCount, dParams, s ) );
			snap = dPatched;
		}
		HIP_CHECK( launch_replay_step( g->dWorlds, v.dItems, v.itemCount, v.block, v.dBodyList, v.dSlotBodies, snap, dA, dB, t->dSimBase,
									   dCf, maxSlots, dPriv, dMass, v.storeLdsBytes, s ) );
		std::swap( dA, dB );
	}
	HIP_CHECK( launch_rows_from_poses( dA, rows, n, s ) );
	if ( n > 0 )
		HIP_CHECK( hipMemcpyAsync( endPoses, rows, sizeof( float ) * 7 * (size_t)n, hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	return true;
}

StepGroup* b2mi_internal_batch( int batch );

} // namespace b2mi

using namespace b2mi;

extern "C" {

B2_API void b2mi_Batch_BeginTape( int batch, int checkpointInterval )
{
	tapeBegin( b2mi_internal_batch( batch ), checkpointInterval );
}

B2_API void b2mi_Batch_EndTape( int batch )
{
	tapeFree( b2mi_internal_batch( batch ) );
}

B2_API int b2mi_Batch_GetTapeStepCount( int batch )
{
	StepGroup* g = b2mi_internal_batch( batch );
	return g->tape ? (int)g->tape->steps.size() : 0;
}

B2_API int b2mi_Batch_GetTapeBodies( int batch, b2BodyId* bodies, int capacity )
{
	StepGroup* g = b2mi_internal_batch( batch );
	B2MI_CHECK( g->tape != nullptr );
	int n = (int)g->tape->bodies.size();
	for ( int i = 0; i < n && i < capacity; ++i )
		bodies[i] = g->tape->bodies[i];
	return n;
}

B2_API void b2mi_Batch_GetTapePoses( int batch, float* poses )
{
	StepGroup* g = b2mi_internal_batch( batch );
	Tape* t = g->tape;
	B2MI_CHECK( t != nullptr );
	hipStream_t s = stream();
	const TapeArena::Mark scratchMark = t->arena->mark();
	StepPose* d = t->arena->take<StepPose>( t->totalSims );
	if ( g->dWorlds == nullptr )
		fatal( "the batch has not been stepped", __FILE__, __LINE__ );
	HIP_CHECK( launch_gather_poses( g->dWorlds, t->dSimBase, t->dSimCount, (int)g->worlds.size(), t->maxSims, d, s ) );
	std::vector<StepPose> h( t->totalSims );
	HIP_CHECK( hipMemcpyAsync( h.data(), d, sizeof( StepPose ) * h.size(), hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	t->arena->rewind( scratchMark );
	for ( size_t i = 0; i < h.size(); ++i )
	{
		float* o = poses + 7 * i;
		o[0] = h[i].cx, o[1] = h[i].cy, o[2] = h[i].qc, o[3] = h[i].qs, o[4] = h[i].vx, o[5] = h[i].vy, o[6] = h[i].w;
	}
}

B2_API void b2mi_Batch_ReplayTape( int batch, const float* startPoses, const float* worldParams, float* endPoses )
{
	StepGroup* g = b2mi_internal_batch( batch );
	B2MI_CHECK( g->tape != nullptr );
	tapeReplay( g, startPoses, worldParams, endPoses );
}

B2_API int b2mi_Batch_ReplayTapeNoAlloc( int batch, const float* startPoses, const float* worldParams, float* endPoses )
{
	StepGroup* g = b2mi_internal_batch( batch );
	B2MI_CHECK( g->tape != nullptr );
	return tapeReplayNoAlloc( g, startPoses, worldParams, endPoses ) ? 1 : 0;
}

B2_API int b2mi_Batch_GetTapeSlots( int batch, int step, b2mi_TapeSlot* slots, int capacity )
{
	StepGroup* g = b2mi_internal_batch( batch );
	Tape* t = g->tape;
	B2MI_CHECK( t != nullptr && step >= 0 && step < (int)t->steps.size() );
	const TapeStep& ts = t->steps[step];
	const TapeVersion& v = t->versions[ts.version];
	int used = v.usedSlots; // slots of work items (the launch arrays also hold unused entries)
	if ( slots == nullptr || capacity < used )
		return used;
	hipStream_t s = stream();
	std::vector<WorkItem> items( v.itemCount );
	std::vector<int> bodyList( v.bodyListCount ), slotContact( v.slotCount );
	std::vector<int2> slotBodies( v.slotCount );
	std::vector<DevContact> snap( v.slotCount );
	if ( v.itemCount )
		HIP_CHECK( hipMemcpyAsync( items.data(), v.dItems, sizeof( WorkItem ) * v.itemCount, hipMemcpyDeviceToHost, s ) );
	if ( v.bodyListCount )
		HIP_CHECK( hipMemcpyAsync( bodyList.data(), v.dBodyList, sizeof( int ) * v.bodyListCount, hipMemcpyDeviceToHost, s ) );
	if ( v.slotCount )
	{
		HIP_CHECK( hipMemcpyAsync( slotContact.data(), v.dSlotContact, sizeof( int ) * v.slotCount, hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipMemcpyAsync( slotBodies.data(), v.dSlotBodies, sizeof( int2 ) * v.slotCount, hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipMemcpyAsync( snap.data(), ts.dSnap, sizeof( DevContact ) * v.slotCount, hipMemcpyDeviceToHost, s ) );
	}
	HIP_CHECK( hipStreamSynchronize( s ) );
	int out = 0;
	for ( const WorkItem& wi : items )
		for ( int gr = 1; gr < kGroups; ++gr )
			for ( int k = wi.groupStart[gr]; k < wi.groupStart[gr + 1]; ++k )
			{
				int slot = wi.slotStart + k;
				const DevContact& c = snap[slot];
				b2mi_TapeSlot& o = slots[out++];
				int2 lb = slotBodies[slot];
				o.world = wi.world;
				o.contactId = slotContact[slot];
				o.color = gr - 1;
				o.bodyA = lb.x < 0 ? -1 : t->simBase[wi.world] + bodyList[wi.bodyStart + lb.x];
				o.bodyB = lb.y < 0 ? -1 : t->simBase[wi.world] + bodyList[wi.bodyStart + lb.y];
				o.pointCount = c.manifold.pointCount;
				o.warmFrom = c.warmFrom;
				o.friction = c.friction;
				o.restitution = c.restitution;
				o.invMassA = c.invMassA;
				o.invIA = c.invIA;
				o.invMassB = c.invMassB;
				o.invIB = c.invIB;
				o.normal[0] = c.manifold.normalX;
				o.normal[1] = c.manifold.normalY;
				for ( int j = 0; j < 2; ++j )
				{
					const DevManifoldPoint& p = c.manifold.points[j];
					o.anchorA[j][0] = p.anchorAX;
					o.anchorA[j][1] = p.anchorAY;
					o.anchorB[j][0] = p.anchorBX;
					o.anchorB[j][1] = p.anchorBY;
					o.separation[j] = p.separation;
					o.normalImpulse[j] = p.normalImpulse;
					o.tangentImpulse[j] = p.tangentImpulse;
				}
			}
	return out;
}

B2_API void b2mi_Batch_GetTapeWorldParams( int batch, float* params )
{
	StepGroup* g = b2mi_internal_batch( batch );
	B2MI_CHECK( g->tape != nullptr && g->hWorlds.size() == g->worlds.size() );
	for ( size_t k = 0; k < g->hWorlds.size(); ++k )
	{
		const DevWorld& d = g->hWorlds[k];
		float* o = params + B2MI_TAPE_WORLD_PARAMS * k;
		o[0] = d.h, o[1] = d.inv_h, o[2] = d.gravityX, o[3] = d.gravityY;
		o[4] = d.contactSoft[0], o[5] = d.contactSoft[1], o[6] = d.contactSoft[2];
		o[7] = d.staticSoft[0], o[8] = d.staticSoft[1], o[9] = d.staticSoft[2];
		o[10] = d.restitutionThreshold, o[11] = d.pushout, o[12] = d.maxLinearSpeed, o[13] = d.maxAngularSpeed;
		o[14] = d.warmStartScale, o[15] = (float)d.subSteps;
	}
}

B2_API void b2mi_Batch_Backward( int batch, const float* gradEnd, float* gradStart, float* gradParams )
{
	StepGroup* g = b2mi_internal_batch( batch );
	B2MI_CHECK( g->tape != nullptr );
	tapeBackward( g, gradEnd, gradStart, gradParams );
}

} // extern "C"
