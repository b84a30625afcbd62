// World / body / shape API and the host half of the step pipeline.
//
// Serial, order-defining work (ids, pairs, contact creation, graph colouring, tree
// maintenance) runs here exactly as the reference orders it; the data-parallel work
// (narrowphase, solver, finalize) runs in kernels.hip. Line references are to the reference.
#include "launch.hpp"
#include "world_internal.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace b2mi
{


// ---------------------------------------------------------------------------------------
// errors

static int defaultAssert( const char* condition, const char* fileName, int lineNumber )
{
	fprintf( stderr, "box2c_amd: %s (%s:%d)\n", condition, fileName, lineNumber );
	fflush( stderr );
	return 1;
}

static b2AssertFcn* g_assert = defaultAssert;
static std::string g_lastError;

void fatal( const char* what, const char* file, int line )
{
	g_lastError = what;
	if ( g_assert( what, file, line ) )
		abort();
}

// ---------------------------------------------------------------------------------------
// device context

static hipStream_t g_stream = nullptr;
static std::once_flag g_streamOnce;

hipStream_t stream()
{
	std::call_once( g_streamOnce, [] {
		int count = 0;
		hipError_t e = hipGetDeviceCount( &count );
		if ( e != hipSuccess || count == 0 )
			fatal( "no HIP device is available: the step pipeline runs only on the GPU", __FILE__, __LINE__ );
		HIP_CHECK( hipStreamCreateWithFlags( &g_stream, hipStreamNonBlocking ) );
	} );
	return g_stream;
}

hipStream_t auxStream( int i )
{
	static hipStream_t streams[kAuxStreams] = {};
	static std::once_flag once;
	std::call_once( once, [] {
		for ( hipStream_t& st : streams )
			HIP_CHECK( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
	} );
	B2MI_CHECK( 0 <= i && i < kAuxStreams );
	return streams[i];
}

template <typename T>
static void devGrow( T*& ptr, int& capacity, int needed, bool preserve, int preserveCount = 0 )
{
	if ( needed <= capacity )
		return;
	int cap = capacity > 0 ? capacity : 16;
	while ( cap < needed )
		cap *= 2;
	T* p = nullptr;
	HIP_CHECK( hipMalloc( (void**)&p, sizeof( T ) * (size_t)cap ) );
	if ( preserve && ptr != nullptr && preserveCount > 0 )
		HIP_CHECK( hipMemcpyAsync( p, ptr, sizeof( T ) * (size_t)preserveCount, hipMemcpyDeviceToDevice, stream() ) );
	if ( ptr != nullptr )
	{
		HIP_CHECK( hipStreamSynchronize( stream() ) );
		HIP_CHECK( hipFree( ptr ) );
	}
	ptr = p;
	capacity = cap;
}

template <typename T>
static void devFree( T*& ptr )
{
	if ( ptr )
		(void)hipFree( ptr );
	ptr = nullptr;
}

// ---------------------------------------------------------------------------------------
// world table

static World* g_worlds[kMaxWorldSlots] = {};
static uint16_t g_revisions[kMaxWorldSlots] = {};

// every API call reaches its world through here: the world is then visited by the host pass of
// its group's next step, whatever the call changed
World* worldAt( int slot )
{
	B2MI_CHECK( 0 <= slot && slot < kMaxWorldSlots && g_worlds[slot] != nullptr );
	World* w = g_worlds[slot];
	markActive( w );
	return w;
}

World* getWorld( b2WorldId id )
{
	B2MI_CHECK( 1 <= id.index1 );
	World* w = worldAt( id.index1 - 1 );
	B2MI_CHECK( w->revision == id.revision );
	return w;
}

World::~World()
{
	devFree( dState );
	devFree( dSims );
	devFree( dPoses );
	devFree( dStaticPoses );
	devFree( dStatics );
	devFree( dShapes );
	devFree( dBodyShapes );
	devFree( dContacts );
	devFree( dContactOut );
	devFree( dJoints );
}

StepGroup::~StepGroup()
{
	tapeRelease( this );
	devFree( dWorlds );
	devFree( dCollideList );
	devFree( dChanges );
	devFree( dCounters );
	if ( hCounters )
		(void)hipHostFree( hCounters );
	devFree( dEnlarge );
	devFree( dInits );
	devFree( dItems );
	devFree( dOvfLevels );
	devFree( dBodyList );
	devFree( dSlotContact );
	devFree( dSlotBodies );
	devFree( dCons );
	devFree( dPriv );
	devFree( dMass );
	devFree( dGridSync );
	if ( graphExec )
		(void)hipGraphExecDestroy( graphExec );
	for ( hipStream_t cs : captureStreams )
		if ( cs )
			(void)hipStreamDestroy( cs );
	for ( hipEvent_t e : events )
		if ( e )
			(void)hipEventDestroy( e );
	if ( forkEvent )
		(void)hipEventDestroy( forkEvent );
	for ( hipEvent_t e : collideDone )
		if ( e )
			(void)hipEventDestroy( e );
	for ( hipEvent_t e : joinEvents )
		if ( e )
			(void)hipEventDestroy( e );
}

static World* createWorld( const b2WorldDef* def, int firstSlot, int endSlot )
{
	B2MI_CHECK( def->internalValue == kCookie );
	int slot = -1;
	for ( int i = firstSlot; i < endSlot; ++i )
		if ( g_worlds[i] == nullptr )
		{
			slot = i;
			break;
		}
	if ( slot < 0 )
		return nullptr;
	World* w = new World();
	w->slot = slot;
	w->revision = g_revisions[slot];
	w->inUse = true;
	w->gravity = def->gravity;
	w->hitEventThreshold = def->hitEventThreshold;
	w->restitutionThreshold = def->restitutionThreshold;
	w->contactPushoutVelocity = def->contactPushoutVelocity;
	w->contactHertz = def->contactHertz;
	w->contactDampingRatio = def->contactDampingRatio;
	w->jointHertz = def->jointHertz;
	w->jointDampingRatio = def->jointDampingRatio;
	w->enableSleep = def->enableSleep;
	w->enableContinuous = def->enableContinous;
	g_worlds[slot] = w;
	return w;
}

static void destroyWorld( World* w )
{
	if ( w->group && w->group->worlds.size() == 1 && !w->batchWorld )
		delete w->group;
	g_revisions[w->slot] = w->revision + 1;
	g_worlds[w->slot] = nullptr;
	delete w;
}

// ---------------------------------------------------------------------------------------
// host mirrors

static void downloadMirror( World* w )
{
	if ( w->mirrorValid && w->mirrorEpoch == stepEpoch( w ) )
		return;
	hipStream_t s = stream();
	int na = (int)w->awakeSimBody.size();
	if ( na )
	{
		HIP_CHECK( hipMemcpyAsync( w->awakeSims.data(), w->dSims, sizeof( DevBodySim ) * na, hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipMemcpyAsync( w->awakeState.data(), w->dState, sizeof( float ) * 8 * na, hipMemcpyDeviceToHost, s ) );
	}
	int ns = (int)w->shapes.size();
	if ( ns )
		HIP_CHECK( hipMemcpyAsync( w->devShapes.data(), w->dShapes, sizeof( DevShape ) * ns, hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	for ( int i = 0; i < ns; ++i )
	{
		Shape& sh = w->shapes[i];
		if ( sh.id < 0 )
			continue;
		const DevShape& d = w->devShapes[i];
		sh.aabb = b2AABB{ { d.aabb[0], d.aabb[1] }, { d.aabb[2], d.aabb[3] } };
		sh.fatAABB = b2AABB{ { d.fatAABB[0], d.fatAABB[1] }, { d.fatAABB[2], d.fatAABB[3] } };
	}
	w->mirrorValid = true;
	w->mirrorEpoch = stepEpoch( w );
}

static void downloadContacts( World* w )
{
	if ( w->contactsDownloaded && w->contactsEpoch == stepEpoch( w ) )
		return;
	int n = w->contactIdPool.nextIndex;
	w->devContacts.resize( n );
	w->devContactOut.resize( n );
	if ( n && w->dContacts )
	{
		HIP_CHECK( hipMemcpyAsync( w->devContacts.data(), w->dContacts, sizeof( DevContact ) * n, hipMemcpyDeviceToHost, stream() ) );
		HIP_CHECK(
			hipMemcpyAsync( w->devContactOut.data(), w->dContactOut, sizeof( DevContactOut ) * n, hipMemcpyDeviceToHost, stream() ) );
		HIP_CHECK( hipStreamSynchronize( stream() ) );
	}
	w->contactsDownloaded = true;
	w->contactsEpoch = stepEpoch( w );
}

// The last step's move and hit events, gathered from the device state on first demand: a move
// event per awake body in awake order with its transform after finalize (src/solver.c:226-231),
// and for each touching contact with hit events, colour by colour in local order, the manifold
// point of the fastest approach above the threshold that took an impulse (src/solver.c:1670-1723).
// Begin/end touch events were recorded by the step's state changes.
void gatherLazyEvents( World* w )
{
	if ( !w->lazyEvents )
		return;
	w->lazyEvents = false;
	const uint16_t world0 = (uint16_t)w->slot;
	const int na = (int)w->awakeSimBody.size();
	w->moveEvents.resize( na );
	if ( na > 0 )
	{
		std::vector<DevPose> poses( na );
		HIP_CHECK( hipMemcpyAsync( poses.data(), w->dPoses, sizeof( DevPose ) * na, hipMemcpyDeviceToHost, stream() ) );
		HIP_CHECK( hipStreamSynchronize( stream() ) );
		for ( int i = 0; i < na; ++i )
		{
			const Body& b = w->bodies[w->awakeSimBody[i]];
			b2BodyMoveEvent& e = w->moveEvents[i];
			e.transform = b2Transform{ { poses[i].p[0], poses[i].p[1] }, { poses[i].q[0], poses[i].q[1] } };
			e.bodyId = b2BodyId{ b.id + 1, world0, b.revision };
			e.userData = b.userData;
			e.fellAsleep = false;
		}
	}
	bool hits = false;
	for ( const Shape& sh : w->shapes )
		if ( sh.id >= 0 && sh.enableHitEvents )
		{
			hits = true;
			break;
		}
	if ( !hits )
		return;
	downloadContacts( w );
	for ( int color = 0; color < kColorCount; ++color )
		for ( int contactId : w->colors[color] )
		{
			const DevContact& dc = w->devContacts[contactId];
			if ( ( dc.simFlags & kSimHitEvent ) == 0 )
				continue;
			const DevContactOut& out = w->devContactOut[contactId];
			b2ContactHitEvent event{};
			event.approachSpeed = w->hitEventThreshold;
			bool hit = false;
			for ( int k = 0; k < dc.manifold.pointCount; ++k )
			{
				const float approachSpeed = -out.normalVelocity[k];
				if ( approachSpeed > event.approachSpeed && dc.manifold.points[k].normalImpulse > 0.0f )
				{
					event.approachSpeed = approachSpeed;
					event.point = b2Vec2{ out.pointX[k], out.pointY[k] };
					hit = true;
				}
			}
			if ( !hit )
				continue;
			const Contact& c = w->contacts[contactId];
			const Shape& a = w->shapes[c.shapeIdA];
			const Shape& b = w->shapes[c.shapeIdB];
			event.normal = b2Vec2{ dc.manifold.normalX, dc.manifold.normalY };
			event.shapeIdA = b2ShapeId{ a.id + 1, world0, a.revision };
			event.shapeIdB = b2ShapeId{ b.id + 1, world0, b.revision };
			w->contactHit.push_back( event );
		}
}

// writes the host mirror (bodies, shapes, per-body shape lists) to the device
static void uploadBodies( World* w )
{
	hipStream_t s = stream();
	int na = (int)w->awakeSimBody.size(), nst = (int)w->staticSimBody.size(), ns = (int)w->shapes.size();
	// reallocate the device arrays at a grown capacity (uploads happen only after edits)
	auto grow = []( int n ) { return std::max( 16, n + n / 2 ); };
	if ( w->capAwake < na || w->dState == nullptr )
	{
		devFree( w->dState );
		devFree( w->dSims );
		devFree( w->dPoses );
		w->capAwake = grow( na );
		HIP_CHECK( hipMalloc( (void**)&w->dState, sizeof( float ) * 8 * (size_t)w->capAwake ) );
		HIP_CHECK( hipMalloc( (void**)&w->dSims, sizeof( DevBodySim ) * (size_t)w->capAwake ) );
		HIP_CHECK( hipMalloc( (void**)&w->dPoses, sizeof( DevPose ) * (size_t)w->capAwake ) );
	}
	if ( w->capStatic < nst || w->dStatics == nullptr )
	{
		devFree( w->dStatics );
		devFree( w->dStaticPoses );
		w->capStatic = grow( nst );
		HIP_CHECK( hipMalloc( (void**)&w->dStatics, sizeof( DevBodySim ) * (size_t)w->capStatic ) );
		HIP_CHECK( hipMalloc( (void**)&w->dStaticPoses, sizeof( DevPose ) * (size_t)w->capStatic ) );
	}
	if ( w->capShapes < ns || w->dShapes == nullptr )
	{
		devFree( w->dShapes );
		devFree( w->dBodyShapes );
		w->capShapes = grow( ns );
		HIP_CHECK( hipMalloc( (void**)&w->dShapes, sizeof( DevShape ) * (size_t)w->capShapes ) );
		HIP_CHECK( hipMalloc( (void**)&w->dBodyShapes, sizeof( int ) * (size_t)w->capShapes ) );
	}

	// per-body shape lists, head-first (the order finalize and the enlarge pass walk them)
	std::vector<int> bodyShapes;
	bodyShapes.reserve( ns );
	for ( int i = 0; i < na; ++i )
	{
		Body& b = w->bodies[w->awakeSimBody[i]];
		DevBodySim& sim = w->awakeSims[i];
		sim.shapeStart = (int)bodyShapes.size();
		for ( int id = b.headShapeId; id != -1; id = w->shapes[id].nextShapeId )
			bodyShapes.push_back( id );
		sim.shapeCount = (int)bodyShapes.size() - sim.shapeStart;
	}
	for ( int i = 0; i < ns; ++i )
	{
		Shape& sh = w->shapes[i];
		DevShape& d = w->devShapes[i];
		if ( sh.id < 0 )
			continue;
		const Body& b = w->bodies[sh.bodyId];
		d.body = b.setIndex == 0 ? ( b.localIndex | kStaticRef ) : b.localIndex;
	}
	std::vector<DevPose> poses( na ), staticPoses( nst );
	for ( int i = 0; i < na; ++i )
		poses[i] = poseOf( w->awakeSims[i] );
	for ( int i = 0; i < nst; ++i )
		staticPoses[i] = poseOf( w->staticSims[i] );
	if ( na )
	{
		HIP_CHECK( hipMemcpyAsync( w->dState, w->awakeState.data(), sizeof( float ) * 8 * na, hipMemcpyHostToDevice, s ) );
		HIP_CHECK( hipMemcpyAsync( w->dSims, w->awakeSims.data(), sizeof( DevBodySim ) * na, hipMemcpyHostToDevice, s ) );
		HIP_CHECK( hipMemcpyAsync( w->dPoses, poses.data(), sizeof( DevPose ) * na, hipMemcpyHostToDevice, s ) );
	}
	if ( nst )
	{
		HIP_CHECK( hipMemcpyAsync( w->dStatics, w->staticSims.data(), sizeof( DevBodySim ) * nst, hipMemcpyHostToDevice, s ) );
		HIP_CHECK( hipMemcpyAsync( w->dStaticPoses, staticPoses.data(), sizeof( DevPose ) * nst, hipMemcpyHostToDevice, s ) );
	}
	if ( ns )
		HIP_CHECK( hipMemcpyAsync( w->dShapes, w->devShapes.data(), sizeof( DevShape ) * ns, hipMemcpyHostToDevice, s ) );
	if ( !bodyShapes.empty() )
		HIP_CHECK( hipMemcpyAsync( w->dBodyShapes, bodyShapes.data(), sizeof( int ) * bodyShapes.size(), hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	w->bodiesDirty = false;
	if ( w->group )
		w->group->structureDirty = true;
}

// the host mirror is about to be modified by an API call (the last step's events are taken
// first, as the step left them)
static void beginEdit( World* w )
{
	B2MI_CHECK( worldLocked( w ) == false );
	gatherLazyEvents( w );
	downloadMirror( w );
	w->bodiesDirty = true;
}

static void updateBodyMass( World* w, Body* body );
void destroyContactImpl( World* w, Contact* c );
void destroyBodyJoints( World* w, Body* body );
int jointColorCount( const World* w, int color );

void b2mi_internal_beginEdit( World* w )
{
	beginEdit( w );
}
void b2mi_internal_uploadBodies( World* w )
{
	uploadBodies( w );
}
void b2mi_internal_updateBodyMass( World* w, Body* body )
{
	updateBodyMass( w, body );
}
void b2mi_internal_downloadMirror( World* w )
{
	downloadMirror( w );
}
void b2mi_internal_downloadContacts( World* w )
{
	downloadContacts( w );
}
World* b2mi_internal_worldOrNull( int slot )
{
	return ( 0 <= slot && slot < kMaxWorldSlots ) ? g_worlds[slot] : nullptr;
}
World* b2mi_internal_createWorld( const b2WorldDef* def, int firstSlot, int endSlot )
{
	return createWorld( def, firstSlot, endSlot );
}
void b2mi_internal_destroyWorld( World* w )
{
	destroyWorld( w );
}
const char* b2mi_internal_lastError()
{
	return g_lastError.c_str();
}

} // namespace b2mi

using namespace b2mi;

// ---------------------------------------------------------------------------------------
// defaults (src/types.c:8-60)

extern "C" {

B2_API void b2SetAssertFcn( b2AssertFcn* assertFcn )
{
	g_assert = assertFcn ? assertFcn : defaultAssert;
}

B2_API b2Version b2GetVersion( void )
{
	return b2Version{ 3, 0, 0 };
}

B2_API b2WorldDef b2DefaultWorldDef( void )
{
	b2WorldDef def{};
	def.gravity = { 0.0f, -10.0f };
	def.hitEventThreshold = 1.0f;
	def.restitutionThreshold = 1.0f;
	def.contactPushoutVelocity = 3.0f;
	def.contactHertz = 30.0f;
	def.contactDampingRatio = 10.0f;
	def.jointHertz = 60.0f;
	def.jointDampingRatio = 2.0f;
	def.enableSleep = true;
	def.enableContinous = true;
	def.internalValue = kCookie;
	return def;
}

B2_API b2BodyDef b2DefaultBodyDef( void )
{
	b2BodyDef def{};
	def.type = b2_staticBody;
	def.rotation = { 1.0f, 0.0f };
	def.sleepThreshold = 0.05f;
	def.gravityScale = 1.0f;
	def.enableSleep = true;
	def.isAwake = true;
	def.isEnabled = true;
	def.automaticMass = true;
	def.internalValue = kCookie;
	return def;
}

B2_API b2Filter b2DefaultFilter( void )
{
	return b2Filter{ 0x00000001u, 0xFFFFFFFFu, 0 };
}

B2_API b2ShapeDef b2DefaultShapeDef( void )
{
	b2ShapeDef def{};
	def.friction = 0.6f;
	def.density = 1.0f;
	def.filter = b2DefaultFilter();
	def.enableSensorEvents = true;
	def.enableContactEvents = true;
	def.internalValue = kCookie;
	return def;
}

// ---------------------------------------------------------------------------------------
// world API (src/world.c)

B2_API b2WorldId b2CreateWorld( const b2WorldDef* def )
{
	World* w = createWorld( def, 0, kMaxPlainWorlds );
	if ( w == nullptr )
		return b2WorldId{ 0, 0 };
	return b2WorldId{ (uint16_t)( w->slot + 1 ), w->revision };
}

B2_API void b2DestroyWorld( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	B2MI_CHECK( w->batchWorld == false && "destroy batch worlds through b2mi_DestroyBatch" );
	destroyWorld( w );
}

B2_API bool b2World_IsValid( b2WorldId id )
{
	if ( id.index1 < 1 || id.index1 > kMaxWorldSlots )
		return false;
	World* w = g_worlds[id.index1 - 1];
	return w != nullptr && w->revision == id.revision;
}

B2_API void b2World_EnableSleeping( b2WorldId worldId, bool flag )
{
	World* w = getWorld( worldId );
	w->enableSleep = flag;
	if ( w->group )
		w->group->structureDirty = true;
}

B2_API void b2World_EnableContinuous( b2WorldId worldId, bool flag )
{
	World* w = getWorld( worldId );
	w->enableContinuous = flag;
	if ( w->group )
		w->group->structureDirty = true;
}

B2_API void b2World_SetRestitutionThreshold( b2WorldId worldId, float value )
{
	getWorld( worldId )->restitutionThreshold = b2ClampFloat( value, 0.0f, FLT_MAX );
}

B2_API void b2World_SetHitEventThreshold( b2WorldId worldId, float value )
{
	World* w = getWorld( worldId );
	gatherLazyEvents( w ); // the last step's hit events used the threshold of that step
	w->hitEventThreshold = b2ClampFloat( value, 0.0f, FLT_MAX );
}

// src/world.c:1366-1413
B2_API b2BodyEvents b2World_GetBodyEvents( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	if ( worldLocked( w ) )
		return b2BodyEvents{};
	gatherLazyEvents( w );
	return b2BodyEvents{ w->moveEvents.data(), (int32_t)w->moveEvents.size() };
}

B2_API b2SensorEvents b2World_GetSensorEvents( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	if ( worldLocked( w ) )
		return b2SensorEvents{};
	return b2SensorEvents{ w->sensorBegin.data(), w->sensorEnd.data(), (int32_t)w->sensorBegin.size(), (int32_t)w->sensorEnd.size() };
}

B2_API b2ContactEvents b2World_GetContactEvents( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	if ( worldLocked( w ) )
		return b2ContactEvents{};
	gatherLazyEvents( w );
	return b2ContactEvents{ w->contactBegin.data(),		  w->contactEnd.data(),			  w->contactHit.data(),
							(int32_t)w->contactBegin.size(), (int32_t)w->contactEnd.size(), (int32_t)w->contactHit.size() };
}

// A custom filter or pre-solve callback would have to run inside the device narrowphase and
// pair kernels; the request is refused loudly instead of silently ignored.
B2_API void b2World_SetCustomFilterCallback( b2WorldId worldId, b2CustomFilterFcn* fcn, void* context )
{
	(void)worldId;
	(void)context;
	if ( fcn != nullptr )
		fatal( "custom filter callbacks are not supported by the device pipeline", __FILE__, __LINE__ );
}

B2_API void b2World_SetPreSolveCallback( b2WorldId worldId, b2PreSolveFcn* fcn, void* context )
{
	(void)worldId;
	(void)context;
	if ( fcn != nullptr )
		fatal( "pre-solve callbacks are not supported by the device pipeline", __FILE__, __LINE__ );
}

B2_API void b2World_SetGravity( b2WorldId worldId, b2Vec2 gravity )
{
	getWorld( worldId )->gravity = gravity;
}

B2_API b2Vec2 b2World_GetGravity( b2WorldId worldId )
{
	return getWorld( worldId )->gravity;
}

B2_API void b2World_SetContactTuning( b2WorldId worldId, float hertz, float dampingRatio, float pushVelocity )
{
	World* w = getWorld( worldId );
	w->contactHertz = b2ClampFloat( hertz, 0.0f, FLT_MAX );
	w->contactDampingRatio = b2ClampFloat( dampingRatio, 0.0f, FLT_MAX );
	w->contactPushoutVelocity = b2ClampFloat( pushVelocity, 0.0f, FLT_MAX );
}

B2_API void b2World_EnableWarmStarting( b2WorldId worldId, bool flag )
{
	getWorld( worldId )->enableWarmStarting = flag;
}

B2_API b2Profile b2World_GetProfile( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	return w->group ? w->group->profile : w->profile;
}

B2_API b2Counters b2World_GetCounters( b2WorldId worldId )
{
	World* w = getWorld( worldId );
	b2Counters c{};
	c.staticBodyCount = (int)w->staticSimBody.size();
	c.bodyCount = w->bodyIdPool.count();
	c.shapeCount = w->shapeIdPool.count();
	c.contactCount = w->contactIdPool.count();
	c.jointCount = w->jointIdPool.count();
	c.staticTreeHeight = w->trees[b2_staticBody].height();
	c.treeHeight = std::max( w->trees[b2_dynamicBody].height(), w->trees[b2_kinematicBody].height() );
	for ( int i = 0; i < kColorCount; ++i )
		c.colorCounts[i] = (int)w->colors[i].size() + jointColorCount( w, i );
	return c;
}

} // extern "C"

// ---------------------------------------------------------------------------------------
// bodies (src/body.c:178-318, 516-617)

namespace b2mi
{

static Body* getBody( World* w, b2BodyId id )
{
	int index = id.index1 - 1;
	B2MI_CHECK( 0 <= index && index < (int)w->bodies.size() );
	Body* b = &w->bodies[index];
	B2MI_CHECK( b->id == index && b->revision == id.revision );
	return b;
}

static DevBodySim& simOf( World* w, const Body* b )
{
	return b->setIndex == 0 ? w->staticSims[b->localIndex] : w->awakeSims[b->localIndex];
}

static void shapeExtent( const Shape& s, b2Vec2 localCenter, float* minExtent, float* maxExtent )
{
	// src/shape.c:551-619
	switch ( s.type )
	{
		case b2_capsuleShape:
		{
			float radius = s.capsule.radius;
			*minExtent = radius;
			b2Vec2 c1 = b2Sub( s.capsule.center1, localCenter ), c2 = b2Sub( s.capsule.center2, localCenter );
			*maxExtent = sqrtf( b2MaxFloat( b2LengthSquared( c1 ), b2LengthSquared( c2 ) ) ) + radius;
		}
		break;
		case b2_circleShape:
			*minExtent = s.circle.radius;
			*maxExtent = b2Length( b2Sub( s.circle.center, localCenter ) ) + s.circle.radius;
			break;
		case b2_polygonShape:
		{
			const b2Polygon& p = s.polygon;
			float minE = kHuge, maxSqr = 0.0f;
			for ( int i = 0; i < p.count; ++i )
			{
				b2Vec2 v = p.vertices[i];
				float planeOffset = b2Dot( p.normals[i], b2Sub( v, p.centroid ) );
				minE = b2MinFloat( minE, planeOffset );
				maxSqr = b2MaxFloat( maxSqr, b2LengthSquared( b2Sub( v, localCenter ) ) );
			}
			*minExtent = minE + p.radius;
			*maxExtent = sqrtf( maxSqr ) + p.radius;
		}
		break;
		case b2_segmentShape:
		{
			*minExtent = 0.0f;
			b2Vec2 c1 = b2Sub( s.segment.point1, localCenter ), c2 = b2Sub( s.segment.point2, localCenter );
			*maxExtent = sqrtf( b2MaxFloat( b2LengthSquared( c1 ), b2LengthSquared( c2 ) ) );
		}
		break;
		default:
			*minExtent = *maxExtent = 0.0f;
	}
}

static b2MassData shapeMass( const Shape& s )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return b2ComputeCapsuleMass( &s.capsule, s.density );
		case b2_circleShape:
			return b2ComputeCircleMass( &s.circle, s.density );
		case b2_polygonShape:
			return b2ComputePolygonMass( &s.polygon, s.density );
		default:
			return b2MassData{};
	}
}

static b2AABB shapeAABB( const Shape& s, b2Transform xf )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return b2ComputeCapsuleAABB( &s.capsule, xf );
		case b2_circleShape:
			return b2ComputeCircleAABB( &s.circle, xf );
		case b2_polygonShape:
			return b2ComputePolygonAABB( &s.polygon, xf );
		case b2_segmentShape:
			return b2ComputeSegmentAABB( &s.segment, xf );
		default:
			return b2AABB{ xf.p, xf.p };
	}
}

static b2Transform simTransform( const DevBodySim& s )
{
	return b2Transform{ { s.p[0], s.p[1] }, { s.q[0], s.q[1] } };
}

// src/body.c:516-617
static void updateBodyMass( World* w, Body* body )
{
	DevBodySim& sim = simOf( w, body );
	sim.mass = 0.0f;
	sim.invMass = 0.0f;
	sim.invI = 0.0f;
	float inertia = 0.0f;
	sim.localCenter[0] = sim.localCenter[1] = 0.0f;
	sim.minExtent = kHuge;
	sim.maxExtent = 0.0f;
	if ( body->type != b2_dynamicBody )
	{
		sim.center[0] = sim.p[0];
		sim.center[1] = sim.p[1];
		if ( body->type == b2_kinematicBody )
			for ( int id = body->headShapeId; id != -1; id = w->shapes[id].nextShapeId )
			{
				float mn, mx;
				shapeExtent( w->shapes[id], b2Vec2{ 0.0f, 0.0f }, &mn, &mx );
				sim.minExtent = b2MinFloat( sim.minExtent, mn );
				sim.maxExtent = b2MaxFloat( sim.maxExtent, mx );
			}
		return;
	}
	b2Vec2 localCenter = { 0.0f, 0.0f };
	for ( int id = body->headShapeId; id != -1; )
	{
		const Shape& s = w->shapes[id];
		id = s.nextShapeId;
		if ( s.density == 0.0f )
			continue;
		b2MassData md = shapeMass( s );
		sim.mass += md.mass;
		localCenter = b2MulAdd( localCenter, md.mass, md.center );
		inertia += md.rotationalInertia;
	}
	if ( sim.mass > 0.0f )
	{
		sim.invMass = 1.0f / sim.mass;
		localCenter = b2MulSV( sim.invMass, localCenter );
	}
	if ( inertia > 0.0f && body->fixedRotation == false )
	{
		inertia -= sim.mass * b2Dot( localCenter, localCenter );
		sim.invI = 1.0f / inertia;
	}
	else
	{
		inertia = 0.0f;
		sim.invI = 0.0f;
	}
	sim.pad = 0;
	b2Vec2 oldCenter = { sim.center[0], sim.center[1] };
	sim.localCenter[0] = localCenter.x;
	sim.localCenter[1] = localCenter.y;
	b2Vec2 c = b2TransformPoint( simTransform( sim ), localCenter );
	sim.center[0] = c.x;
	sim.center[1] = c.y;
	if ( body->setIndex == 2 )
	{
		float* st = &w->awakeState[8 * body->localIndex];
		b2Vec2 deltaLinear = b2CrossSV( st[2], b2Sub( c, oldCenter ) );
		st[0] = st[0] + deltaLinear.x;
		st[1] = st[1] + deltaLinear.y;
	}
	for ( int id = body->headShapeId; id != -1; id = w->shapes[id].nextShapeId )
	{
		float mn, mx;
		shapeExtent( w->shapes[id], localCenter, &mn, &mx );
		sim.minExtent = b2MinFloat( sim.minExtent, mn );
		sim.maxExtent = b2MaxFloat( sim.maxExtent, mx );
	}
	// inertia is not kept on the device; stash it for b2Body_GetInertiaTensor
	memcpy( &sim.pad, &inertia, sizeof( float ) );
}

static void bufferMove( World* w, int proxyKey )
{
	// src/broad_phase.h:76-83
	if ( !w->moveSet.get( (uint32_t)proxyKey ) )
	{
		w->moveSet.set( (uint32_t)proxyKey );
		w->moveArray.push_back( proxyKey );
	}
}

} // namespace b2mi

extern "C" {

B2_API b2BodyId b2CreateBody( b2WorldId worldId, const b2BodyDef* def )
{
	B2MI_CHECK( def->internalValue == kCookie );
	World* w = getWorld( worldId );
	if ( worldLocked( w ) )
		return b2BodyId{ 0, 0, 0 };
	if ( def->isEnabled == false )
		fatal( "disabled bodies are not supported", __FILE__, __LINE__ );
	if ( def->type != b2_staticBody && def->isAwake == false && def->enableSleep )
		fatal( "bodies created asleep are not supported", __FILE__, __LINE__ );
	beginEdit( w );

	int bodyId = w->bodyIdPool.alloc();
	DevBodySim sim{};
	sim.p[0] = def->position.x;
	sim.p[1] = def->position.y;
	sim.q[0] = def->rotation.c;
	sim.q[1] = def->rotation.s;
	sim.center[0] = sim.center0[0] = def->position.x;
	sim.center[1] = sim.center0[1] = def->position.y;
	sim.rotation0[0] = def->rotation.c;
	sim.rotation0[1] = def->rotation.s;
	sim.minExtent = kHuge;
	sim.linearDamping = def->linearDamping;
	sim.angularDamping = def->angularDamping;
	sim.gravityScale = def->gravityScale;
	sim.sleepThreshold = def->sleepThreshold;
	sim.flags = ( def->type == b2_dynamicBody ? kBodyDynamic : 0u ) | ( def->type == b2_kinematicBody ? kBodyKinematic : 0u ) |
				( def->isBullet ? kBodyBullet : 0u ) | ( def->enableSleep ? kBodyEnableSleep : 0u );
	int setIndex = def->type == b2_staticBody ? 0 : 2;
	int localIndex;
	if ( setIndex == 0 )
	{
		localIndex = (int)w->staticSims.size();
		w->staticSims.push_back( sim );
		w->staticSimBody.push_back( bodyId );
	}
	else
	{
		localIndex = (int)w->awakeSims.size();
		w->awakeSims.push_back( sim );
		w->awakeSimBody.push_back( bodyId );
		float st[8] = { def->linearVelocity.x, def->linearVelocity.y, def->angularVelocity, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f };
		w->awakeState.insert( w->awakeState.end(), st, st + 8 );
	}
	if ( bodyId == (int)w->bodies.size() )
		w->bodies.push_back( Body{} );
	Body& b = w->bodies[bodyId];
	uint16_t revision = b.revision;
	b = Body{};
	b.userData = def->userData;
	b.setIndex = setIndex;
	b.localIndex = localIndex;
	b.revision = revision + 1;
	b.id = bodyId;
	b.type = def->type;
	b.fixedRotation = def->fixedRotation;
	b.automaticMass = def->automaticMass;
	w->topologyDirty = true;
	w->contactsDirty = true;
	return b2BodyId{ bodyId + 1, (uint16_t)w->slot, b.revision };
}

B2_API bool b2Body_IsValid( b2BodyId id )
{
	if ( id.world0 >= kMaxWorldSlots || g_worlds[id.world0] == nullptr )
		return false;
	World* w = g_worlds[id.world0];
	if ( id.index1 < 1 || id.index1 > (int)w->bodies.size() )
		return false;
	const Body& b = w->bodies[id.index1 - 1];
	return b.setIndex != -1 && b.id == id.index1 - 1 && b.revision == id.revision;
}

#define B2MI_BODY( bodyId )                 \
	World* w = worldAt( ( bodyId ).world0 ); \
	Body* body = getBody( w, ( bodyId ) )

B2_API b2BodyType b2Body_GetType( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	return body->type;
}

B2_API void b2Body_SetUserData( b2BodyId bodyId, void* userData )
{
	B2MI_BODY( bodyId );
	gatherLazyEvents( w ); // the last step's move events carry the user data they had
	body->userData = userData;
}

// src/body.c:1654-1664
B2_API void b2Body_EnableHitEvents( b2BodyId bodyId, bool enableHitEvents )
{
	B2MI_BODY( bodyId );
	beginEdit( w );
	for ( int id = body->headShapeId; id != -1; id = w->shapes[id].nextShapeId )
	{
		w->shapes[id].enableHitEvents = enableHitEvents;
		if ( enableHitEvents )
			w->devShapes[id].flags |= kShapeHitEvents;
		else
			w->devShapes[id].flags &= ~kShapeHitEvents;
	}
}

B2_API void* b2Body_GetUserData( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	return body->userData;
}

B2_API b2Transform b2Body_GetTransform( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	return simTransform( simOf( w, body ) );
}

B2_API b2Vec2 b2Body_GetPosition( b2BodyId bodyId )
{
	return b2Body_GetTransform( bodyId ).p;
}

B2_API b2Rot b2Body_GetRotation( b2BodyId bodyId )
{
	return b2Body_GetTransform( bodyId ).q;
}

B2_API b2Vec2 b2Body_GetLocalPoint( b2BodyId bodyId, b2Vec2 worldPoint )
{
	return b2InvTransformPoint( b2Body_GetTransform( bodyId ), worldPoint );
}

B2_API b2Vec2 b2Body_GetWorldPoint( b2BodyId bodyId, b2Vec2 localPoint )
{
	return b2TransformPoint( b2Body_GetTransform( bodyId ), localPoint );
}

B2_API b2Vec2 b2Body_GetLinearVelocity( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return b2Vec2{ 0.0f, 0.0f };
	downloadMirror( w );
	const float* st = &w->awakeState[8 * body->localIndex];
	return b2Vec2{ st[0], st[1] };
}

B2_API float b2Body_GetAngularVelocity( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return 0.0f;
	downloadMirror( w );
	return w->awakeState[8 * body->localIndex + 2];
}

// The overflow levels of the work items count a kinematic body as changed by the overflow
// stages when one of its velocity components is -0 (v - 0 * P can turn it into +0); a new
// velocity of an awake kinematic body therefore rebuilds the items.
static void kinematicVelocitySet( World* w, const Body* body )
{
	if ( body->type == b2_kinematicBody )
		w->topologyDirty = true;
}

B2_API void b2Body_SetLinearVelocity( b2BodyId bodyId, b2Vec2 v )
{
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	float* st = &w->awakeState[8 * body->localIndex];
	st[0] = v.x;
	st[1] = v.y;
	kinematicVelocitySet( w, body );
}

B2_API void b2Body_SetAngularVelocity( b2BodyId bodyId, float av )
{
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	w->awakeState[8 * body->localIndex + 2] = av;
	kinematicVelocitySet( w, body );
}

// src/body.c:799-904
B2_API void b2Body_ApplyForce( b2BodyId bodyId, b2Vec2 force, b2Vec2 point, bool wake )
{
	(void)wake;
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	DevBodySim& sim = w->awakeSims[body->localIndex];
	sim.force[0] = sim.force[0] + force.x;
	sim.force[1] = sim.force[1] + force.y;
	sim.torque += b2Cross( b2Sub( point, b2Vec2{ sim.center[0], sim.center[1] } ), force );
}

B2_API void b2Body_ApplyForceToCenter( b2BodyId bodyId, b2Vec2 force, bool wake )
{
	(void)wake;
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	DevBodySim& sim = w->awakeSims[body->localIndex];
	sim.force[0] = sim.force[0] + force.x;
	sim.force[1] = sim.force[1] + force.y;
}

B2_API void b2Body_ApplyTorque( b2BodyId bodyId, float torque, bool wake )
{
	(void)wake;
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	w->awakeSims[body->localIndex].torque += torque;
}

B2_API void b2Body_ApplyLinearImpulse( b2BodyId bodyId, b2Vec2 impulse, b2Vec2 point, bool wake )
{
	(void)wake;
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	DevBodySim& sim = w->awakeSims[body->localIndex];
	float* st = &w->awakeState[8 * body->localIndex];
	b2Vec2 v = b2MulAdd( b2Vec2{ st[0], st[1] }, sim.invMass, impulse );
	st[0] = v.x;
	st[1] = v.y;
	st[2] = st[2] + sim.invI * b2Cross( b2Sub( point, b2Vec2{ sim.center[0], sim.center[1] } ), impulse );
}

B2_API void b2Body_ApplyLinearImpulseToCenter( b2BodyId bodyId, b2Vec2 impulse, bool wake )
{
	(void)wake;
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	DevBodySim& sim = w->awakeSims[body->localIndex];
	float* st = &w->awakeState[8 * body->localIndex];
	b2Vec2 v = b2MulAdd( b2Vec2{ st[0], st[1] }, sim.invMass, impulse );
	st[0] = v.x;
	st[1] = v.y;
}

B2_API void b2Body_ApplyAngularImpulse( b2BodyId bodyId, float impulse, bool wake )
{
	(void)wake;
	B2MI_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	beginEdit( w );
	float* st = &w->awakeState[8 * body->localIndex];
	st[2] = st[2] + w->awakeSims[body->localIndex].invI * impulse;
}

B2_API float b2Body_GetMass( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	return simOf( w, body ).mass;
}

B2_API float b2Body_GetInertiaTensor( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	float inertia;
	memcpy( &inertia, &simOf( w, body ).pad, sizeof( float ) );
	return body->type == b2_dynamicBody ? inertia : 0.0f;
}

B2_API b2Vec2 b2Body_GetLocalCenterOfMass( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	const DevBodySim& s = simOf( w, body );
	return b2Vec2{ s.localCenter[0], s.localCenter[1] };
}

B2_API b2Vec2 b2Body_GetWorldCenterOfMass( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	const DevBodySim& s = simOf( w, body );
	return b2Vec2{ s.center[0], s.center[1] };
}

B2_API b2MassData b2Body_GetMassData( b2BodyId bodyId )
{
	b2MassData m;
	m.mass = b2Body_GetMass( bodyId );
	m.center = b2Body_GetLocalCenterOfMass( bodyId );
	m.rotationalInertia = b2Body_GetInertiaTensor( bodyId );
	return m;
}

B2_API void b2Body_SetLinearDamping( b2BodyId bodyId, float v )
{
	B2MI_BODY( bodyId );
	beginEdit( w );
	simOf( w, body ).linearDamping = v;
}

B2_API float b2Body_GetLinearDamping( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	return simOf( w, body ).linearDamping;
}

B2_API void b2Body_SetAngularDamping( b2BodyId bodyId, float v )
{
	B2MI_BODY( bodyId );
	beginEdit( w );
	simOf( w, body ).angularDamping = v;
}

B2_API float b2Body_GetAngularDamping( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	return simOf( w, body ).angularDamping;
}

B2_API void b2Body_SetGravityScale( b2BodyId bodyId, float v )
{
	B2MI_BODY( bodyId );
	beginEdit( w );
	simOf( w, body ).gravityScale = v;
}

B2_API float b2Body_GetGravityScale( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	return simOf( w, body ).gravityScale;
}

B2_API bool b2Body_IsAwake( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	return body->setIndex == 2;
}

B2_API bool b2Body_IsBullet( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	return ( simOf( w, body ).flags & kBodyBullet ) != 0;
}

B2_API int b2Body_GetShapeCount( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	return body->shapeCount;
}

B2_API int b2Body_GetShapes( b2BodyId bodyId, b2ShapeId* shapeArray, int capacity )
{
	B2MI_BODY( bodyId );
	int n = 0;
	for ( int id = body->headShapeId; id != -1 && n < capacity; id = w->shapes[id].nextShapeId )
		shapeArray[n++] = b2ShapeId{ id + 1, bodyId.world0, w->shapes[id].revision };
	return n;
}

B2_API int b2Body_GetContactCapacity( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	return body->contactCount;
}

// src/body.c:447-489
B2_API int b2Body_GetContactData( b2BodyId bodyId, b2ContactData* contactData, int capacity )
{
	B2MI_BODY( bodyId );
	downloadContacts( w );
	int index = 0;
	for ( int key = body->headContactKey; key != -1 && index < capacity; )
	{
		const Contact& c = w->contacts[key >> 1];
		if ( c.flags & kContactTouching )
		{
			const Shape& a = w->shapes[c.shapeIdA];
			const Shape& b = w->shapes[c.shapeIdB];
			contactData[index].shapeIdA = b2ShapeId{ a.id + 1, bodyId.world0, a.revision };
			contactData[index].shapeIdB = b2ShapeId{ b.id + 1, bodyId.world0, b.revision };
			contactData[index].manifold = publicManifold( w->devContacts[c.contactId], w->devContactOut[c.contactId] );
			index += 1;
		}
		key = c.edges[key & 1].nextKey;
	}
	return index;
}

B2_API b2AABB b2Body_ComputeAABB( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	downloadMirror( w );
	if ( body->headShapeId == -1 )
	{
		b2Transform xf = simTransform( simOf( w, body ) );
		return b2AABB{ xf.p, xf.p };
	}
	b2AABB aabb = w->shapes[body->headShapeId].aabb;
	for ( int id = w->shapes[body->headShapeId].nextShapeId; id != -1; id = w->shapes[id].nextShapeId )
		aabb = b2AABB_Union( aabb, w->shapes[id].aabb );
	return aabb;
}

// src/body.c:683-744
B2_API void b2Body_SetTransform( b2BodyId bodyId, b2Vec2 position, b2Rot rotation )
{
	B2MI_BODY( bodyId );
	beginEdit( w );
	DevBodySim& sim = simOf( w, body );
	sim.p[0] = position.x;
	sim.p[1] = position.y;
	sim.q[0] = rotation.c;
	sim.q[1] = rotation.s;
	b2Vec2 c = b2TransformPoint( b2Transform{ position, rotation }, b2Vec2{ sim.localCenter[0], sim.localCenter[1] } );
	sim.center[0] = sim.center0[0] = c.x;
	sim.center[1] = sim.center0[1] = c.y;
	sim.rotation0[0] = rotation.c;
	sim.rotation0[1] = rotation.s;
	b2Transform xf{ position, rotation };
	for ( int id = body->headShapeId; id != -1; id = w->shapes[id].nextShapeId )
	{
		Shape& sh = w->shapes[id];
		b2AABB aabb = shapeAABB( sh, xf );
		aabb.lowerBound.x -= kSpeculative;
		aabb.lowerBound.y -= kSpeculative;
		aabb.upperBound.x += kSpeculative;
		aabb.upperBound.y += kSpeculative;
		sh.aabb = aabb;
		if ( b2AABB_Contains( sh.fatAABB, aabb ) == false )
		{
			b2AABB fat = { { aabb.lowerBound.x - kAabbMargin, aabb.lowerBound.y - kAabbMargin },
						   { aabb.upperBound.x + kAabbMargin, aabb.upperBound.y + kAabbMargin } };
			sh.fatAABB = fat;
			w->trees[sh.proxyKey & 3].moveProxy( sh.proxyKey >> 2, fat );
			bufferMove( w, sh.proxyKey );
		}
		DevShape& d = w->devShapes[id];
		d.aabb[0] = sh.aabb.lowerBound.x;
		d.aabb[1] = sh.aabb.lowerBound.y;
		d.aabb[2] = sh.aabb.upperBound.x;
		d.aabb[3] = sh.aabb.upperBound.y;
		d.fatAABB[0] = sh.fatAABB.lowerBound.x;
		d.fatAABB[1] = sh.fatAABB.lowerBound.y;
		d.fatAABB[2] = sh.fatAABB.upperBound.x;
		d.fatAABB[3] = sh.fatAABB.upperBound.y;
	}
}

} // extern "C"

extern "C" {

// src/body.c:320-440 (contacts, proxies, shapes, sim swap-remove)
B2_API void b2DestroyBody( b2BodyId bodyId )
{
	B2MI_BODY( bodyId );
	beginEdit( w );
	destroyBodyJoints( w, body );
	for ( int key = body->headContactKey; key != -1; )
	{
		Contact* c = &w->contacts[key >> 1];
		key = c->edges[key & 1].nextKey;
		destroyContactImpl( w, c );
	}
	for ( int id = body->headShapeId; id != -1; )
	{
		Shape& sh = w->shapes[id];
		id = sh.nextShapeId;
		if ( w->moveSet.get( (uint32_t)sh.proxyKey ) )
		{
			w->moveSet.clear( (uint32_t)sh.proxyKey );
			auto it = std::find( w->moveArray.begin(), w->moveArray.end(), sh.proxyKey );
			if ( it != w->moveArray.end() )
			{
				*it = w->moveArray.back();
				w->moveArray.pop_back();
			}
		}
		w->trees[sh.proxyKey & 3].destroyProxy( sh.proxyKey >> 2 );
		w->shapeIdPool.free( sh.id );
		w->shapesVersion += 1;
		sh.id = -1;
	}
	if ( body->setIndex == 0 )
	{
		int last = (int)w->staticSims.size() - 1;
		if ( body->localIndex < last )
		{
			w->staticSims[body->localIndex] = w->staticSims[last];
			w->staticSimBody[body->localIndex] = w->staticSimBody[last];
			w->bodies[w->staticSimBody[body->localIndex]].localIndex = body->localIndex;
		}
		w->staticSims.pop_back();
		w->staticSimBody.pop_back();
	}
	else
	{
		int last = (int)w->awakeSims.size() - 1;
		if ( body->localIndex < last )
		{
			w->awakeSims[body->localIndex] = w->awakeSims[last];
			memcpy( &w->awakeState[8 * body->localIndex], &w->awakeState[8 * last], 8 * sizeof( float ) );
			w->awakeSimBody[body->localIndex] = w->awakeSimBody[last];
			w->bodies[w->awakeSimBody[body->localIndex]].localIndex = body->localIndex;
		}
		w->awakeSims.pop_back();
		w->awakeState.resize( 8 * w->awakeSims.size() );
		w->awakeSimBody.pop_back();
	}
	w->bodyIdPool.free( body->id );
	body->setIndex = -1;
	body->localIndex = -1;
	body->id = -1;
	w->topologyDirty = true;
	w->contactsDirty = true;
}

} // extern "C"
