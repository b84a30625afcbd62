// Host-side world model of the MI355X step pipeline.
//
// The host keeps the serial, order-defining bookkeeping of the reference (ids, shape and
// contact lists, broadphase trees, pair set, move buffer, greedy graph colouring) and the
// device keeps every numeric array the step reads or writes (body states and sims, shape
// AABBs, contact manifolds). Host mirrors of the numeric arrays are downloaded lazily for
// the getters and uploaded again after a setter changed them.
#pragma once

#include "b2mi_layout.h"
#include "tree.hpp"

#include "box2d/box2d.h"
#include "b2mi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

namespace b2mi
{

constexpr int kColorCount = 12;
constexpr int kStaticSet = 0; // solver set indices (src/solver_set.h:12-19)
constexpr int kAwakeSet = 2;
constexpr int kOverflow = 11;
constexpr int kMaxPlainWorlds = 128; // src/core.h:129
constexpr int kMaxWorldSlots = 65535;
constexpr int kCookie = 1152023; // src/core.h:159
constexpr float kLinearSlop = 0.005f;
constexpr float kSpeculative = 4.0f * kLinearSlop;
constexpr float kAabbMargin = 0.1f;
constexpr float kHuge = 100000.0f;

void fatal( const char* what, const char* file, int line );
#define B2MI_CHECK( cond )                          \
	do                                              \
	{                                               \
		if ( !( cond ) )                            \
			::b2mi::fatal( #cond, __FILE__, __LINE__ ); \
	} while ( 0 )
#define HIP_CHECK( expr )                                                \
	do                                                                   \
	{                                                                    \
		hipError_t e_ = ( expr );                                        \
		if ( e_ != hipSuccess )                                          \
			::b2mi::fatal( hipGetErrorString( e_ ), __FILE__, __LINE__ ); \
	} while ( 0 )

struct IdPool
{
	std::vector<int> freeArray;
	int nextIndex = 0;
	int alloc()
	{
		if ( !freeArray.empty() )
		{
			int id = freeArray.back();
			freeArray.pop_back();
			return id;
		}
		return nextIndex++;
	}
	// src/id_pool.c:36-47: freed ids always go to the free array
	void free( int id )
	{
		freeArray.push_back( id );
	}
	int count() const
	{
		return nextIndex - (int)freeArray.size();
	}
};

struct BitSet
{
	std::vector<uint64_t> bits;
	void setCountAndClear( uint32_t bitCount )
	{
		bits.assign( ( bitCount + 63 ) / 64, 0 );
	}
	void set( uint32_t i )
	{
		uint32_t b = i / 64;
		if ( b >= bits.size() )
			bits.resize( b + 1, 0 );
		bits[b] |= uint64_t( 1 ) << ( i % 64 );
	}
	void clear( uint32_t i )
	{
		uint32_t b = i / 64;
		if ( b < bits.size() )
			bits[b] &= ~( uint64_t( 1 ) << ( i % 64 ) );
	}
	bool get( uint32_t i ) const
	{
		uint32_t b = i / 64;
		return b < bits.size() && ( ( bits[b] >> ( i % 64 ) ) & 1 );
	}
};

struct Body
{
	void* userData = nullptr;
	int setIndex = -1; // -1 free, 0 static, 2 awake
	int localIndex = -1;
	int headContactKey = -1;
	int contactCount = 0;
	int headShapeId = -1;
	int shapeCount = 0;
	int headJointKey = -1;
	int jointCount = 0;
	int id = -1;
	b2BodyType type = b2_staticBody;
	uint16_t revision = 0;
	bool fixedRotation = false;
	bool automaticMass = true;
};

struct Shape
{
	int id = -1;
	int bodyId = -1;
	int prevShapeId = -1;
	int nextShapeId = -1;
	b2ShapeType type = b2_circleShape;
	float density = 0, friction = 0, restitution = 0;
	b2AABB aabb{}, fatAABB{};
	int proxyKey = -1;
	b2Filter filter{};
	void* userData = nullptr;
	union
	{
		b2Capsule capsule;
		b2Circle circle;
		b2Polygon polygon;
		b2Segment segment;
	};
	uint16_t revision = 0;
	bool enableHitEvents = false;
	bool enableContactEvents = true;
	Shape() : polygon{} {}
};

struct ContactEdge
{
	int bodyId, prevKey, nextKey;
};

struct Contact
{
	int setIndex = -1; // -1 free, 2 awake
	int colorIndex = -1;
	int localIndex = -1;
	ContactEdge edges[2];
	int shapeIdA = -1, shapeIdB = -1;
	int contactId = -1;
	uint32_t flags = 0; // kContactTouching, kContactEvents
};

// A joint's host record (src/joint.h:182-205): its id, set, colour and place in that colour's
// list, and the edges of the two bodies' joint lists. Its parameters, impulses and prepared terms
// are a DevJoint, kept on the device by joint id (World::dJoints) with a host mirror.
struct JointEdge
{
	int bodyId, prevKey, nextKey;
};

struct Joint
{
	void* userData = nullptr;
	int setIndex = -1; // -1 free, 0 static (both bodies static), 2 awake
	int colorIndex = -1;
	int localIndex = -1;
	JointEdge edges[2];
	int jointId = -1;
	b2JointType type = b2_revoluteJoint;
	uint16_t revision = 0;
	float drawSize = 1.0f;
	bool collideConnected = false;
};

// Contact::flags (src/contact.h b2ContactFlags)
constexpr uint32_t kContactTouching = 0x1u;
constexpr uint32_t kContactEvents = 0x20u; // either shape had contact events enabled at creation

struct StepGroup;

// The device side of a world's pair queries: uploaded copies of the trees and shape filters, the
// pair set as an open-addressing table kept in step with the host's by the keys created and
// destroyed since it was last brought up to date, and the buffers of one query launch.
// The set of shape pair keys that have a contact (src/broad_phase.h pairSet): open addressing with
// linear probing over a power-of-two table kept at most half full, erase by backward shift (no
// tombstones). Only membership matters; the order of a walk over the keys is the table's.
class PairSet
{
public:
	size_t size() const
	{
		return count_;
	}
	bool count( uint64_t key ) const
	{
		if ( count_ == 0 )
			return false;
		for ( size_t i = home( key );; i = ( i + 1 ) & mask_ )
		{
			if ( slots_[i] == key )
				return true;
			if ( slots_[i] == kEmpty )
				return false;
		}
	}
	void insert( uint64_t key )
	{
		if ( 2 * ( count_ + 1 ) > slots_.size() )
			grow();
		size_t i = home( key );
		while ( slots_[i] != kEmpty )
		{
			if ( slots_[i] == key )
				return;
			i = ( i + 1 ) & mask_;
		}
		slots_[i] = key;
		count_ += 1;
	}
	void erase( uint64_t key )
	{
		if ( count_ == 0 )
			return;
		size_t i = home( key );
		while ( slots_[i] != key )
		{
			if ( slots_[i] == kEmpty )
				return;
			i = ( i + 1 ) & mask_;
		}
		// backward shift: move later keys of the run into the hole when their home allows it
		size_t j = i;
		for ( ;; )
		{
			j = ( j + 1 ) & mask_;
			if ( slots_[j] == kEmpty )
				break;
			const size_t h = home( slots_[j] );
			// the key at j may fill the hole at i unless its home lies cyclically in (i, j]
			if ( ( ( j - h ) & mask_ ) >= ( ( j - i ) & mask_ ) )
			{
				slots_[i] = slots_[j];
				i = j;
			}
		}
		slots_[i] = kEmpty;
		count_ -= 1;
	}
	template <typename F>
	void forEach( F&& f ) const
	{
		for ( uint64_t k : slots_ )
			if ( k != kEmpty )
				f( k );
	}

private:
	static constexpr uint64_t kEmpty = ~0ull; // never a key: shape ids are below 2^31
	size_t home( uint64_t k ) const
	{
		k ^= k >> 33;
		k *= 0xff51afd7ed558ccdull;
		k ^= k >> 33;
		return (size_t)k & mask_;
	}
	void grow()
	{
		std::vector<uint64_t> old;
		old.swap( slots_ );
		slots_.assign( old.empty() ? 1024 : 2 * old.size(), kEmpty );
		mask_ = slots_.size() - 1;
		count_ = 0;
		for ( uint64_t k : old )
			if ( k != kEmpty )
				insert( k );
	}
	std::vector<uint64_t> slots_;
	size_t mask_ = 0;
	size_t count_ = 0;
};

struct DeviceBroadphase
{
	DevTreeNode* nodes[3] = {};
	int capNodes[3] = {};
	uint64_t uploadedVersion[3] = { ~0ull, ~0ull, ~0ull };
	DevTreeNode* staging[3] = {}; // pinned host copies of the trees as uploaded
	int capStaging[3] = {};
	std::vector<int> walk; // the depth-first walk that numbers the leaves of a staged tree
	DevShapeFilter* shapes = nullptr;
	int capShapes = 0;
	uint64_t uploadedShapes = ~0ull;
	unsigned long long* table = nullptr;
	uint32_t tableSize = 0; // a power of two
	int tableTombstones = 0;
	bool tableValid = false;
	std::vector<uint64_t> pendingKeys; // pair keys created or destroyed since the table was updated
	uint32_t* moved = nullptr;
	int capMovedWords = 0;
	int* moves = nullptr;
	int capMoves = 0;
	unsigned long long* keys = nullptr;
	int capKeys = 0;
	PairRecord* pairs = nullptr;
	int capPairs = 0;
	PairSpill* spills = nullptr;
	int capSpills = 0;
	int* counters = nullptr;	 // [0] pairs found, [1] error, [2] spilled subtrees
	int* hostCounters = nullptr; // pinned
	std::vector<PairRecord> found;
	// accumulated host clock (ms since each update began) after: the pair-set update, the uploads,
	// the queries, the readback and the contact creation; then pairs found, moves, updates and
	// spilled subtrees
	float phaseMs[9] = {};
	~DeviceBroadphase();
};

// work items of one world, with body and slot indices local to these lists. Rebuilding only the
// items whose bodies' contacts changed retires them (live = 0) and appends their replacements, so
// the lists hold the retired items' entries as holes until the next full rebuild.
struct SolverLists
{
	std::vector<WorkItem> items;
	std::vector<char> live;		  // per item: 0 once retired
	std::vector<int> bodyList;	  // awake sim indices
	std::vector<int> slotContact; // contact ids
	std::vector<int2> slotBodies; // local body indices (-1 static)
	std::vector<int> ovfLevels;	  // overflow level boundaries of the items (relative to each item's slots)
	std::vector<int> itemOfSim;	  // per awake sim: the live item that owns it
	std::vector<int> posInSet;	  // per awake sim, -1 (scratch of a partial rebuild)
	int liveBodies = 0, liveSlots = 0; // entries of the live items
	int maxGroupSize = 0, maxItemLds = 0;
};

// a world's part of one of the step group's launch arrays
struct Region
{
	int base = 0, cap = 0;
};

struct World
{
	int slot = -1; // index in the world table (worldId.index1 - 1)
	uint16_t revision = 0;
	bool inUse = false;
	bool batchWorld = false;

	std::vector<Body> bodies;
	IdPool bodyIdPool;
	std::vector<int> staticSimBody; // static set: sim index -> body id
	std::vector<int> awakeSimBody;	// awake set: sim index -> body id

	// host mirrors of the numeric arrays (valid when mirrorValid)
	std::vector<DevBodySim> staticSims;
	std::vector<DevBodySim> awakeSims;
	std::vector<float> awakeState; // 8 floats per awake body
	std::vector<DevShape> devShapes;
	std::vector<DevContact> devContacts;
	std::vector<DevContactOut> devContactOut;
	bool mirrorValid = true; // and mirrorEpoch is the group's step epoch
	uint64_t mirrorEpoch = 0, contactsEpoch = 0;
	bool bodiesDirty = true; // host mirror of bodies/shapes must be uploaded
	bool contactsDownloaded = false;

	std::vector<Shape> shapes;
	IdPool shapeIdPool;

	std::vector<Contact> contacts;
	IdPool contactIdPool;
	std::vector<int> nonTouching; // awake non-touching contact ids (array order of the reference)
	std::vector<int> colors[kColorCount];
	BitSet colorBodySets[kColorCount];

	// joints: host records, each colour's joint list (awake joints; the order of the reference's
	// colour arrays, which fixes the overflow joints' serial order) and the static set's list
	std::vector<Joint> joints;
	IdPool jointIdPool;
	std::vector<int> jointColors[kColorCount];
	std::vector<int> staticJoints;
	// the joints' device records (by joint id) and their host mirror: the mirror is current when
	// jointsDirty (edits not yet uploaded) or jointsEpoch is the group's step epoch
	std::vector<DevJoint> devJoints;
	DevJoint* dJoints = nullptr;
	int capJoints = 0;
	bool jointsDirty = false;
	uint64_t jointsEpoch = 0;

	DynamicTree trees[3];
	BitSet moveSet;
	std::vector<int> moveArray;
	PairSet pairSet;
	uint64_t shapesVersion = 0; // shapes created, destroyed or refiltered
	// pair queries of a world stepped alone run on the device from this many moved proxies on (-1:
	// never); in a batch they run on the host pool, world by world
	int devicePairsMin = 2048;
	DeviceBroadphase devicePairs;

	// events of the last step (src/world.c:1366-1413). Begin/end touch events are recorded by the
	// state changes; move and hit events are gathered from the device state when first asked for,
	// or before an API call changes what they report (lazyEvents marks them as still to gather).
	std::vector<b2BodyMoveEvent> moveEvents;
	std::vector<b2SensorBeginTouchEvent> sensorBegin;
	std::vector<b2SensorEndTouchEvent> sensorEnd;
	std::vector<b2ContactBeginTouchEvent> contactBegin;
	std::vector<b2ContactEndTouchEvent> contactEnd;
	std::vector<b2ContactHitEvent> contactHit;
	bool lazyEvents = false;

	b2Vec2 gravity{ 0.0f, -10.0f };
	float hitEventThreshold = 1, restitutionThreshold = 1, contactPushoutVelocity = 3;
	float contactHertz = 30, contactDampingRatio = 10, jointHertz = 60, jointDampingRatio = 2;
	bool enableSleep = true, enableWarmStarting = true, enableContinuous = true;

	b2Profile profile{};
	bool topologyDirty = true; // bodies or launch settings changed: rebuild every work item
	std::vector<int> dirtySims; // awake sims whose contacts' colours changed: rebuild their items
	std::vector<int> posOfSim;	// per awake sim, -1 (scratch of the item builder)
	bool contactsDirty = true; // contacts created or destroyed since the narrowphase list was built

	// this world's solver lists (indices local to the lists) and narrowphase list, and the
	// regions they occupy in the step group's launch arrays
	SolverLists solver;
	std::vector<int> collideIds;
	Region bodyRegion, slotRegion, collideRegion;
	bool solverPending = true, collidePending = true; // region contents not yet uploaded
	int pendingBodyFrom = 0, pendingSlotFrom = 0;	  // solver-list entries from these on are not uploaded

	// device arrays
	float* dState = nullptr;
	DevBodySim* dSims = nullptr;
	DevBodySim* dStatics = nullptr;
	DevPose* dPoses = nullptr;		 // capAwake
	DevPose* dStaticPoses = nullptr; // capStatic
	DevShape* dShapes = nullptr;
	int* dBodyShapes = nullptr;
	DevContact* dContacts = nullptr;
	DevContactOut* dContactOut = nullptr;
	int capAwake = 0, capStatic = 0, capShapes = 0, capContacts = 0;

	StepGroup* group = nullptr;
	int groupIndex = -1;
	std::atomic<bool> activeQueued{ false }; // listed in group->active for the next step
	int countedShapes = 0;					  // shapes counted in group->shapeTotal

	// pending new contacts (device init) of the current step
	std::vector<ContactInit> pendingInits;

	~World();
};

struct Tape;	   // taped run of a step group (tape.cpp)
struct TapeArena; // device memory of the group's tapes (tape.cpp)

// items of one world chunk and one solver workgroup size: a contiguous range of the
// launch-ordered items
struct LaunchClass
{
	int first, count, block, lds, chunk, store;
};

constexpr int kAuxStreams = 3;
// A group of worlds can be stepped in up to this many chunks of consecutive worlds: the chunks'
// narrowphase launches run in turn on the group's stream and each chunk's solver as soon as its
// narrowphase is done, beside the narrowphase of the next chunk. Off by default: on 4096
// pyramid worlds the solver launches already fill the GPU and the overlap measured no gain.
constexpr int kMaxChunks = kAuxStreams;
inline int chunkCountFor( int worldCount, int requested )
{
	if ( requested > 0 )
		return std::min( std::min( requested, kMaxChunks ), std::max( worldCount, 1 ) );
	return 1;
}
inline int chunkOfWorld( int worldIndex, int worldCount, int chunks )
{
	return (int)( (int64_t)worldIndex * chunks / worldCount );
}
// device counters of a step: [1] enlarged shapes, [2] bodies at the sleep time, [3] set when a
// barrier wait of the grid solver exceeded its bound, [kChunkCounter + c] state changes of chunk c,
// [kContinuousCounter] bodies the continuous stage moved back to a time of impact and
// [kContinuousCounter + 1] set when one of its tree queries ran out of stack
constexpr int kChunkCounter = 4;
constexpr int kContinuousCounter = kChunkCounter + kMaxChunks;
constexpr int kCounterCount = kContinuousCounter + 2;

// What a step's solver launches bake in (see StepGroup::solveSignature).
struct SolveSignature
{
	int itemsVersion;
	const int* abortFlag;
	const void *worlds, *items, *bodyList, *slotContact, *slotBodies, *ovfLevels, *cons, *priv, *mass, *enlarge, *counters,
		*gridSync;
	int enlargeCapacity, blockSize, maxItemLds, gridBlocks, levelRun, splitPerPhase, chunks;
};

// a set of worlds advanced together by one narrowphase launch and one solver launch
struct StepGroup
{
	std::vector<World*> worlds;
	int blockSize = 0;		  // 0 = automatic
	int targetSlots = 0;	  // work-item packing target (constraint slots), 0 = automatic
	int ldsBudget = -1;		  // max LDS bytes per work item (-1 = kDefaultItemLdsBytes, 0 = HBM only)
	int tunedBlock = 0, tunedTarget = 0, tunedBudget = -1, tunedFullBudget = -1; // last autotuner choice
	int ldsFullBudget = -1;	  // max LDS bytes of an item keeping all its state in LDS (-1 = ldsBudget)
	int splitSlots = -1;	  // HBM-store items of at least this many slots run over the grid (-1 = kDefaultSplitSlots, 0 = never)
	int gridBlocks = -1;	  // workgroups of such an item's cooperative launch (0 = from its largest colour, -1 = one launch per phase)
	int levelRunSlots = kLevelCacheSlots; // HBM-store items: consecutive overflow levels of at most this many slots in all run as one cached run
	unsigned* dGridSync = nullptr; // barrier arrivals and vote of the cooperative grid solver
	int builtTarget = -1, builtBudget = -1, builtFullBudget = -1;				   // packing target and LDS budget of the current items
	bool structureDirty = true;

	// launch buffers (device)
	DevWorld* dWorlds = nullptr;
	int capWorlds = 0;
	int4* dCollideList = nullptr;
	int capCollide = 0;
	int collideCount = 0;
	bool layoutValid = false; // regions of the worlds in the launch arrays are assigned
	ChangeRecord* dChanges = nullptr;
	int capChanges = 0;
	int* dCounters = nullptr; // kCounterCount counters of a step (above)
	int* hCounters = nullptr; // pinned
	EnlargeRecord* dEnlarge = nullptr;
	int capEnlarge = 0;
	ContinuousBody* dContinuous = nullptr; // fast bodies of a continuous pass, and their joints' bodies
	int capContinuous = 0;
	int* dNoCollide = nullptr;
	int capNoCollide = 0;
	int toiHits = 0; // bodies moved back to a time of impact, over all steps
	ContactInit* dInits = nullptr;
	int capInits = 0;
	WorkItem* dItems = nullptr;
	int capItems = 0;
	std::vector<int> hOvfLevels; // the worlds' overflow level boundaries, concatenated
	int* dOvfLevels = nullptr;
	int capOvfLevels = 0;
	int itemCount = 0;
	int* dBodyList = nullptr;
	int capBodyList = 0;
	int* dSlotContact = nullptr;
	int2* dSlotBodies = nullptr;
	int capSlots = 0;
	float* dCons = nullptr;
	int capCons = 0;
	float4* dPriv = nullptr;
	int capPriv = 0;
	float2* dMass = nullptr; // inverse mass and inertia per body-list entry (HBM store)
	int capMass = 0;
	int slotCount = 0, bodyListCount = 0;
	int maxGroupSize = 0;
	int maxItemLds = 0; // dynamic LDS bytes of the solver launch (largest LDS-resident item)
	int itemsVersion = 0; // incremented by every work-item rebuild
	Tape* tape = nullptr; // recording for the adjoint, when on
	TapeArena* tapeArena = nullptr; // kept from one tape to the next, freed with the group

	// host staging
	std::vector<DevWorld> hWorlds;
	std::vector<WorkItem> hItems;
	std::vector<int> hBodyList;
	std::vector<int> hSlotContact;
	std::vector<int2> hSlotBodies;
	std::vector<int4> hCollide; // narrowphase list: (world, contact id, shape ids A and B), or -1
	std::vector<ContactInit> hInits;
	std::vector<ChangeRecord> hChanges;
	std::vector<EnlargeRecord> hEnlarge;

	// timing of the last step (ms)
	float msCollide = 0, msSolve = 0, msHost = 0;
	hipEvent_t events[4] = { nullptr, nullptr, nullptr, nullptr };
	std::vector<LaunchClass> classes;
	std::vector<int> splitItems; // items (indices into hItems) that run over the whole grid
	bool splitPerPhase = false;	 // those run as one launch per phase (no cooperative launches)
	int gridBlocksFor( const WorkItem& wi ) const;
	// the continuous stage of the fast bodies (or bullets) among hEnlarge's records; the shape
	// records it writes are appended to hEnlarge
	void runContinuous( bool bullets );
	hipError_t enqueueSplit( int item, hipStream_t s );
	hipError_t enqueueSolve( hipStream_t s, const hipStream_t* aux, const int* abortFlag );
	// the solver launches of a step replayed from a HIP graph while they stay the same
	bool launchGraphs = true;
	bool graphsRefused = false; // a capture failed: direct launches from then on
	hipGraphExec_t graphExec = nullptr;
	hipStream_t captureStreams[1 + kAuxStreams] = {}; // the group's own, for recording captures
	SolveSignature graphKey{}, lastSolveKey{};
	int graphCaptures = 0, graphReplays = 0;
	SolveSignature solveSignature( const int* abortFlag ) const;
	bool captureSolve( hipStream_t s, const int* abortFlag );
	// host clock (ms since the step began) at: pair update, collide-list build, collide launch,
	// narrowphase wait, wait returned, change processing done, solver launch, solver wait, wait
	// returned, enlarge processing done
	float hostPhases[10] = {};
	bool stepping = false; // inside step(): the API refuses edits of its worlds
	// Worlds the next step must visit on the host: those reached through the API since the last
	// step (any call that names the world or one of its bodies or shapes) and those whose
	// contacts or proxies the last step changed. Every world when allActive is set.
	std::vector<World*> active;
	std::mutex activeMutex;
	bool allActive = true;
	int specMisses = 0; // consecutive speculative solver launches voided by contact changes
	float lastTimeStep = -1.0f;
	int lastSubSteps = -1;
	int shapeTotal = 0;
	uint64_t epoch = 0;	   // steps taken; host mirrors of an older epoch are stale
	b2Profile profile{};   // of the last step, shared by the group's worlds
	void launchSolve( hipStream_t s, const int* abortFlag );
	void launchSolveChunk( int chunk, hipEvent_t ready, const int* abortFlag );
	void joinAux( hipStream_t s );
	int chunkCount = 1;							 // of the launch-ordered items
	int chunksRequested = 0;					 // 0 = automatic
	int chunkCollide[kMaxChunks + 1] = {};		 // collide-list range of each chunk
	hipEvent_t forkEvent = nullptr;
	hipEvent_t joinEvents[kAuxStreams] = {};
	hipEvent_t collideDone[kMaxChunks] = {};

	~StepGroup();
	void step( float timeStep, int subStepCount );
};

// lists the world for the host pass of its group's next step
inline void markActive( World* w )
{
	if ( w->group == nullptr || w->activeQueued.load( std::memory_order_relaxed ) )
		return;
	std::lock_guard<std::mutex> lock( w->group->activeMutex );
	if ( !w->activeQueued.load( std::memory_order_relaxed ) )
	{
		w->activeQueued.store( true, std::memory_order_relaxed );
		w->group->active.push_back( w );
	}
}

// a world whose group is stepping accepts no edits
inline bool worldLocked( const World* w )
{
	return w->group != nullptr && w->group->stepping;
}

// step epoch of a world's group (0 before it has one)
inline uint64_t stepEpoch( const World* w )
{
	return w->group != nullptr ? w->group->epoch : 0;
}

// step.cpp / world.cpp: the events of a world's last step
void clearStepEvents( World* w );
void gatherLazyEvents( World* w );

// tape.cpp
void tapeRecordStep( StepGroup* g, float timeStep, int subStepCount, int blockSize );
void tapeFree( StepGroup* g );	 // ends the tape; its device memory stays with the group
void tapeRelease( StepGroup* g ); // ends the tape and frees the group's tape memory
void tapeUndoStep( StepGroup* g ); // takes back the last tapeRecordStep (a voided speculative launch)

hipStream_t stream();
hipStream_t auxStream( int i ); // extra streams (default priority) for concurrent launches
World* getWorld( b2WorldId id );
World* worldAt( int slot );

// the public b2Manifold of a device contact (its manifold plus the output-only world points and
// normal velocities)
inline b2Manifold publicManifold( const DevContact& d, const DevContactOut& e )
{
	b2Manifold m;
	memset( &m, 0, sizeof( m ) );
	for ( int j = 0; j < 2; ++j )
	{
		const DevManifoldPoint& p = d.manifold.points[j];
		b2ManifoldPoint& o = m.points[j];
		o.point = b2Vec2{ e.pointX[j], e.pointY[j] };
		o.anchorA = b2Vec2{ p.anchorAX, p.anchorAY };
		o.anchorB = b2Vec2{ p.anchorBX, p.anchorBY };
		o.separation = p.separation;
		o.normalImpulse = p.normalImpulse;
		o.tangentImpulse = p.tangentImpulse;
		o.maxNormalImpulse = p.maxNormalImpulse;
		o.normalVelocity = e.normalVelocity[j];
		o.id = p.id;
		o.persisted = p.persisted != 0;
	}
	m.normal = b2Vec2{ d.manifold.normalX, d.manifold.normalY };
	m.pointCount = d.manifold.pointCount;
	return m;
}

// the narrowphase's view of a body sim
inline DevPose poseOf( const DevBodySim& s )
{
	DevPose p;
	memcpy( &p, &s, sizeof( DevPose ) );
	return p;
}

} // namespace b2mi
