// Shapes, contacts, broadphase pairs, graph colouring, work-item building and the step of
// a group of worlds (one world for b2World_Step, many for a batch). Line references are to
// the reference.
#include "launch.hpp"
#include "world_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <climits>
#include <cstring>

namespace b2mi
{


// defined in world.cpp
void b2mi_internal_beginEdit( World* w );
void b2mi_internal_uploadBodies( World* w );
void b2mi_internal_updateBodyMass( World* w, Body* body );
void b2mi_internal_downloadMirror( World* w );
void b2mi_internal_downloadContacts( World* w );
World* b2mi_internal_worldOrNull( int slot );

template <typename T>
static void grow( T*& ptr, int& capacity, int needed )
{
	if ( needed <= capacity && ptr != nullptr )
		return;
	int cap = std::max( 64, capacity );
	while ( cap < needed )
		cap *= 2;
	if ( ptr )
	{
		HIP_CHECK( hipStreamSynchronize( stream() ) );
		HIP_CHECK( hipFree( ptr ) );
	}
	HIP_CHECK( hipMalloc( (void**)&ptr, sizeof( T ) * (size_t)cap ) );
	capacity = cap;
}

// ---------------------------------------------------------------------------------------
// shapes (src/shape.c:38-217)

static b2Vec2 shapeCentroid( const Shape& s )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return b2Lerp( s.capsule.center1, s.capsule.center2, 0.5f );
		case b2_circleShape:
			return s.circle.center;
		case b2_polygonShape:
			return s.polygon.centroid;
		case b2_segmentShape:
			return b2Lerp( s.segment.point1, s.segment.point2, 0.5f );
		default:
			return b2Vec2{ 0.0f, 0.0f };
	}
}

static void fillDevShape( DevShape& d, const Shape& s )
{
	memset( &d, 0, sizeof( d ) );
	d.type = s.type;
	d.flags = s.enableHitEvents ? kShapeHitEvents : 0u;
	d.friction = s.friction;
	d.restitution = s.restitution;
	switch ( s.type )
	{
		case b2_circleShape:
			d.vn[0][0] = s.circle.center.x;
			d.vn[0][1] = s.circle.center.y;
			d.radius = s.circle.radius;
			d.count = 1;
			break;
		case b2_capsuleShape:
			d.vn[0][0] = s.capsule.center1.x;
			d.vn[0][1] = s.capsule.center1.y;
			d.vn[1][0] = s.capsule.center2.x;
			d.vn[1][1] = s.capsule.center2.y;
			d.radius = s.capsule.radius;
			d.count = 2;
			break;
		case b2_segmentShape:
			d.vn[0][0] = s.segment.point1.x;
			d.vn[0][1] = s.segment.point1.y;
			d.vn[1][0] = s.segment.point2.x;
			d.vn[1][1] = s.segment.point2.y;
			d.count = 2;
			break;
		case b2_polygonShape:
			d.count = s.polygon.count;
			d.radius = s.polygon.radius;
			d.centroid[0] = s.polygon.centroid.x;
			d.centroid[1] = s.polygon.centroid.y;
			for ( int i = 0; i < s.polygon.count; ++i )
			{
				d.vert( i )[0] = s.polygon.vertices[i].x;
				d.vert( i )[1] = s.polygon.vertices[i].y;
				d.vert( i )[2] = s.polygon.normals[i].x;
				d.vert( i )[3] = s.polygon.normals[i].y;
			}
			break;
		default:
			break;
	}
	d.aabb[0] = s.aabb.lowerBound.x;
	d.aabb[1] = s.aabb.lowerBound.y;
	d.aabb[2] = s.aabb.upperBound.x;
	d.aabb[3] = s.aabb.upperBound.y;
	d.fatAABB[0] = s.fatAABB.lowerBound.x;
	d.fatAABB[1] = s.fatAABB.lowerBound.y;
	d.fatAABB[2] = s.fatAABB.upperBound.x;
	d.fatAABB[3] = s.fatAABB.upperBound.y;
}

static b2AABB computeShapeAABB( const Shape& s, b2Transform xf )
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

static b2ShapeId createShape( b2BodyId bodyId, const b2ShapeDef* def, const void* geometry, b2ShapeType type )
{
	B2MI_CHECK( def->internalValue == kCookie );
	if ( def->isSensor )
		fatal( "sensor shapes are not supported by the device pipeline", __FILE__, __LINE__ );
	World* w = worldAt( bodyId.world0 );
	if ( worldLocked( w ) )
		return b2ShapeId{ 0, 0, 0 };
	int bodyIndex = bodyId.index1 - 1;
	B2MI_CHECK( 0 <= bodyIndex && bodyIndex < (int)w->bodies.size() );
	Body* body = &w->bodies[bodyIndex];
	B2MI_CHECK( body->id == bodyIndex && body->revision == bodyId.revision );
	b2mi_internal_beginEdit( w );

	const DevBodySim& sim = body->setIndex == 0 ? w->staticSims[body->localIndex] : w->awakeSims[body->localIndex];
	b2Transform xf = { { sim.p[0], sim.p[1] }, { sim.q[0], sim.q[1] } };

	int shapeId = w->shapeIdPool.alloc();
	if ( shapeId == (int)w->shapes.size() )
	{
		w->shapes.push_back( Shape{} );
		w->devShapes.push_back( DevShape{} );
	}
	Shape& s = w->shapes[shapeId];
	uint16_t revision = s.revision;
	s = Shape{};
	switch ( type )
	{
		case b2_capsuleShape:
			s.capsule = *(const b2Capsule*)geometry;
			break;
		case b2_circleShape:
			s.circle = *(const b2Circle*)geometry;
			break;
		case b2_polygonShape:
			s.polygon = *(const b2Polygon*)geometry;
			break;
		case b2_segmentShape:
			s.segment = *(const b2Segment*)geometry;
			break;
		default:
			fatal( "unsupported shape type", __FILE__, __LINE__ );
	}
	s.id = shapeId;
	w->shapesVersion += 1;
	s.bodyId = body->id;
	s.type = type;
	s.density = def->density;
	s.friction = def->friction;
	s.restitution = def->restitution;
	s.filter = def->filter;
	s.userData = def->userData;
	s.enableHitEvents = def->enableHitEvents;
	s.enableContactEvents = def->enableContactEvents;
	s.revision = revision + 1;
	(void)shapeCentroid;

	// b2CreateShapeProxy / b2UpdateShapeAABBs (shape.c:38-59, 692-702)
	b2AABB aabb = computeShapeAABB( s, xf );
	aabb.lowerBound.x -= kSpeculative;
	aabb.lowerBound.y -= kSpeculative;
	aabb.upperBound.x += kSpeculative;
	aabb.upperBound.y += kSpeculative;
	s.aabb = aabb;
	float margin = body->type == b2_staticBody ? kSpeculative : kAabbMargin;
	s.fatAABB = b2AABB{ { aabb.lowerBound.x - margin, aabb.lowerBound.y - margin },
						{ aabb.upperBound.x + margin, aabb.upperBound.y + margin } };
	int proxyType = body->type;
	int proxyId = w->trees[proxyType].createProxy( s.fatAABB, s.filter.categoryBits, shapeId );
	s.proxyKey = ( proxyId << 2 ) | proxyType;
	if ( proxyType != b2_staticBody || def->forceContactCreation )
	{
		if ( !w->moveSet.get( (uint32_t)s.proxyKey ) )
		{
			w->moveSet.set( (uint32_t)s.proxyKey );
			w->moveArray.push_back( s.proxyKey );
		}
	}

	// head insertion in the body's shape list (shape.c:137-148)
	if ( body->headShapeId != -1 )
		w->shapes[body->headShapeId].prevShapeId = shapeId;
	s.prevShapeId = -1;
	s.nextShapeId = body->headShapeId;
	body->headShapeId = shapeId;
	body->shapeCount += 1;
	fillDevShape( w->devShapes[shapeId], s );
	if ( body->automaticMass )
		b2mi_internal_updateBodyMass( w, body );
	return b2ShapeId{ shapeId + 1, bodyId.world0, s.revision };
}

// ---------------------------------------------------------------------------------------
// contacts (src/contact.c:185-414, src/constraint_graph.c:64-214)

// which shape type pairs have a manifold function and which order is primary
// (src/contact.c:150-183); circle=0 capsule=1 segment=2 polygon=3
static bool manifoldFcn( int typeA, int typeB, bool* primary )
{
	static const signed char table[4][4] = {
		{ 1, -1, -1, -1 },
		{ 1, 1, -1, -1 },
		{ 1, 1, 0, 1 },
		{ 1, 1, -1, 1 },
	};
	if ( typeA < 0 || typeA > 3 || typeB < 0 || typeB > 3 )
		return false;
	int v = table[typeA][typeB];
	if ( v == 0 )
		return false;
	*primary = v > 0;
	return true;
}

static uint64_t pairKey( int a, int b )
{
	return a < b ? ( uint64_t( a ) << 32 | uint64_t( b ) ) : ( uint64_t( b ) << 32 | uint64_t( a ) );
}

static void createContact( World* w, Shape* shapeA, Shape* shapeB )
{
	bool primary;
	if ( !manifoldFcn( shapeA->type, shapeB->type, &primary ) )
		return;
	if ( !primary )
	{
		createContact( w, shapeB, shapeA );
		return;
	}
	int contactId = w->contactIdPool.alloc();
	if ( contactId == (int)w->contacts.size() )
		w->contacts.push_back( Contact{} );
	Contact& c = w->contacts[contactId];
	c.contactId = contactId;
	c.setIndex = 2;
	c.colorIndex = -1;
	c.localIndex = (int)w->nonTouching.size();
	c.shapeIdA = shapeA->id;
	c.shapeIdB = shapeB->id;
	c.flags = ( shapeA->enableContactEvents || shapeB->enableContactEvents ) ? kContactEvents : 0u;
	Body* bodies[2] = { &w->bodies[shapeA->bodyId], &w->bodies[shapeB->bodyId] };
	for ( int e = 0; e < 2; ++e )
	{
		Body* b = bodies[e];
		c.edges[e].bodyId = b->id;
		c.edges[e].prevKey = -1;
		c.edges[e].nextKey = b->headContactKey;
		int key = ( contactId << 1 ) | e;
		int head = b->headContactKey;
		if ( head != -1 )
			w->contacts[head >> 1].edges[head & 1].prevKey = key;
		b->headContactKey = key;
		b->contactCount += 1;
	}
	w->pairSet.insert( pairKey( shapeA->id, shapeB->id ) );
	if ( w->devicePairs.tableValid )
		w->devicePairs.pendingKeys.push_back( pairKey( shapeA->id, shapeB->id ) );
	w->nonTouching.push_back( contactId );

	ContactInit init{};
	init.world = w->groupIndex;
	init.contactId = contactId;
	init.shapeIdA = shapeA->id;
	init.shapeIdB = shapeB->id;
	init.friction = sqrtf( shapeA->friction * shapeB->friction );
	init.restitution = shapeA->restitution > shapeB->restitution ? shapeA->restitution : shapeB->restitution;
	w->pendingInits.push_back( init );
	w->contactsDirty = true;
}

// Marks the work items of a contact's bodies for rebuilding: its colour or the overflow order
// around it changed (begin or end of touching, destruction of a touching contact, or a move
// within the overflow list).
static void markContactBodies( World* w, const Contact& c )
{
	for ( int e = 0; e < 2; ++e )
	{
		const Body& b = w->bodies[c.edges[e].bodyId];
		if ( b.setIndex == kAwakeSet )
			w->dirtySims.push_back( b.localIndex );
	}
}

static void removeFromList( World* w, std::vector<int>& list, int localIndex )
{
	int last = (int)list.size() - 1;
	if ( localIndex < last )
	{
		list[localIndex] = list[last];
		w->contacts[list[localIndex]].localIndex = localIndex;
	}
	list.pop_back();
}

// src/constraint_graph.c:64-183 (colour choice only; the sim lives on the device by id)
static void addContactToGraph( World* w, Contact& c )
{
	int colorIndex = kOverflow;
	int bodyIdA = c.edges[0].bodyId, bodyIdB = c.edges[1].bodyId;
	bool staticA = w->bodies[bodyIdA].setIndex == 0, staticB = w->bodies[bodyIdB].setIndex == 0;
	if ( !staticA && !staticB )
	{
		for ( int i = 0; i < kOverflow; ++i )
		{
			BitSet& set = w->colorBodySets[i];
			if ( set.get( bodyIdA ) || set.get( bodyIdB ) )
				continue;
			set.set( bodyIdA );
			set.set( bodyIdB );
			colorIndex = i;
			break;
		}
	}
	else if ( !staticA )
	{
		for ( int i = 1; i < kOverflow; ++i )
		{
			BitSet& set = w->colorBodySets[i];
			if ( set.get( bodyIdA ) )
				continue;
			set.set( bodyIdA );
			colorIndex = i;
			break;
		}
	}
	else if ( !staticB )
	{
		for ( int i = 1; i < kOverflow; ++i )
		{
			BitSet& set = w->colorBodySets[i];
			if ( set.get( bodyIdB ) )
				continue;
			set.set( bodyIdB );
			colorIndex = i;
			break;
		}
	}
	c.colorIndex = colorIndex;
	c.localIndex = (int)w->colors[colorIndex].size();
	w->colors[colorIndex].push_back( c.contactId );
}

// src/constraint_graph.c:185-214. The overflow constraints run in their list's order, so a
// contact moved into the removed one's place there rebuilds its work item.
static void removeContactFromGraph( World* w, int bodyIdA, int bodyIdB, int colorIndex, int localIndex )
{
	if ( colorIndex != kOverflow )
	{
		w->colorBodySets[colorIndex].clear( bodyIdA );
		w->colorBodySets[colorIndex].clear( bodyIdB );
	}
	std::vector<int>& list = w->colors[colorIndex];
	const bool moves = localIndex < (int)list.size() - 1;
	removeFromList( w, list, localIndex );
	if ( colorIndex == kOverflow && moves )
		markContactBodies( w, w->contacts[list[localIndex]] );
}

// src/contact.c:335-414
bool jointsAllowCollision( const World* w, const Body& bodyA, const Body& bodyB );
void uploadJoints( World* w );

void destroyContactImpl( World* w, Contact* c )
{
	w->pairSet.erase( pairKey( c->shapeIdA, c->shapeIdB ) );
	if ( w->devicePairs.tableValid )
		w->devicePairs.pendingKeys.push_back( pairKey( c->shapeIdA, c->shapeIdB ) );
	int contactId = c->contactId;
	for ( int e = 0; e < 2; ++e )
	{
		ContactEdge& edge = c->edges[e];
		Body& b = w->bodies[edge.bodyId];
		if ( edge.prevKey != -1 )
			w->contacts[edge.prevKey >> 1].edges[edge.prevKey & 1].nextKey = edge.nextKey;
		if ( edge.nextKey != -1 )
			w->contacts[edge.nextKey >> 1].edges[edge.nextKey & 1].prevKey = edge.prevKey;
		int key = ( contactId << 1 ) | e;
		if ( b.headContactKey == key )
			b.headContactKey = edge.nextKey;
		b.contactCount -= 1;
	}
	bool wasTouching = c->colorIndex != -1;
	if ( wasTouching )
		removeContactFromGraph( w, c->edges[0].bodyId, c->edges[1].bodyId, c->colorIndex, c->localIndex );
	else
		removeFromList( w, w->nonTouching, c->localIndex );
	if ( wasTouching )
		markContactBodies( w, *c );
	c->contactId = -1;
	c->setIndex = -1;
	c->colorIndex = -1;
	c->localIndex = -1;
	w->contactIdPool.free( contactId );
	w->contactsDirty = true;
}

// ---------------------------------------------------------------------------------------
// broadphase pair update (src/broad_phase.c:162-436)

static bool shouldShapesCollide( b2Filter a, b2Filter b )
{
	if ( a.groupIndex == b.groupIndex && a.groupIndex != 0 )
		return a.groupIndex > 0;
	return ( a.maskBits & b.categoryBits ) != 0 && ( a.categoryBits & b.maskBits ) != 0;
}

struct MovePair
{
	int a, b, next;
};

class HostPool;
static void hostPoolRun( int n, const std::function<void( int )>& fn );

// src/broad_phase.c:278-331 (b2FindPairsTask) for moves [begin, end): the new pairs of each
// moved proxy, as a list through `next` that starts at heads[i] (most recent first, as the
// reference prepends to its per-proxy list); reads the trees and sets only
static void findPairs( const World* w, int begin, int end, std::vector<MovePair>& pairs, int* heads )
{
	for ( int i = begin; i < end; ++i )
	{
		int queryProxyKey = w->moveArray[i];
		int proxyType = queryProxyKey & 3;
		const TreeNode& node = w->trees[proxyType].nodes[queryProxyKey >> 2];
		b2AABB fatAABB = node.aabb;
		int queryShapeIndex = node.userData;
		int* head = &heads[i];
		auto query = [&]( int treeType ) {
			w->trees[treeType].query( fatAABB, b2_defaultMaskBits, [&]( int proxyId, int shapeId ) {
				int proxyKey = ( proxyId << 2 ) | treeType;
				if ( proxyKey == queryProxyKey )
					return true;
				if ( treeType != b2_staticBody && w->moveSet.get( (uint32_t)proxyKey ) && proxyKey < queryProxyKey )
					return true;
				if ( w->pairSet.count( pairKey( shapeId, queryShapeIndex ) ) )
					return true;
				int shapeIdA, shapeIdB;
				if ( proxyKey < queryProxyKey )
				{
					shapeIdA = shapeId;
					shapeIdB = queryShapeIndex;
				}
				else
				{
					shapeIdA = queryShapeIndex;
					shapeIdB = shapeId;
				}
				const Shape& sa = w->shapes[shapeIdA];
				const Shape& sb = w->shapes[shapeIdB];
				if ( sa.bodyId == sb.bodyId )
					return true;
				if ( !shouldShapesCollide( sa.filter, sb.filter ) )
					return true;
				const Body& ba = w->bodies[sa.bodyId];
				const Body& bb = w->bodies[sb.bodyId];
				if ( ba.type != b2_dynamicBody && bb.type != b2_dynamicBody )
					return true;
				// src/body.c:1731-1757: a joint between the bodies may turn their collision off
				if ( ( ba.jointCount > 0 || bb.jointCount > 0 ) && !jointsAllowCollision( w, ba, bb ) )
					return true;
				pairs.push_back( MovePair{ shapeIdA, shapeIdB, *head } );
				*head = (int)pairs.size() - 1;
				return true;
			} );
		};
		if ( proxyType == b2_dynamicBody )
		{
			query( b2_kinematicBody );
			query( b2_staticBody );
		}
		query( b2_dynamicBody );
	}
}

DeviceBroadphase::~DeviceBroadphase()
{
	for ( DevTreeNode*& p : nodes )
		if ( p )
			(void)hipFree( p );
	for ( DevTreeNode*& p : staging )
		if ( p )
			(void)hipHostFree( p );
	for ( void* p : { (void*)shapes, (void*)table, (void*)moved, (void*)moves, (void*)keys, (void*)pairs, (void*)spills, (void*)counters } )
		if ( p )
			(void)hipFree( p );
	if ( hostCounters )
		(void)hipHostFree( hostCounters );
}

// The device pair set brought up to the host's: rebuilt from it when it is not valid or would be
// over half full (counting tombstones), else the keys created or destroyed since the last update
// are erased (those the host no longer has) and inserted (those it has; a key already there is
// left as it is).
static void syncPairTable( World* w, hipStream_t s )
{
	DeviceBroadphase& D = w->devicePairs;
	const size_t live = w->pairSet.size();
	if ( D.tableValid && 2 * ( live + (size_t)D.tableTombstones + D.pendingKeys.size() ) < D.tableSize )
	{
		// each key touched since the last update (a key may repeat): erased where the host's set
		// no longer holds it, inserted where it does (both kernels take repeats)
		std::vector<uint64_t>& keys = D.pendingKeys;
		const int erase = (int)( std::partition( keys.begin(), keys.end(), [&]( uint64_t k ) { return w->pairSet.count( k ) == 0; } ) -
								 keys.begin() );
		const int n = (int)keys.size();
		if ( n > 0 )
		{
			grow( D.keys, D.capKeys, n );
			HIP_CHECK( hipMemcpyAsync( D.keys, keys.data(), sizeof( uint64_t ) * n, hipMemcpyHostToDevice, s ) );
			HIP_CHECK( launch_pair_erase( D.table, D.tableSize - 1, D.keys, erase, s ) );
			HIP_CHECK( launch_pair_insert( D.table, D.tableSize - 1, D.keys + erase, n - erase, D.counters + 1, s ) );
			// the key list is read by the copy: wait before it changes
			HIP_CHECK( hipStreamSynchronize( s ) );
		}
		D.tableTombstones += erase;
		keys.clear();
		return;
	}
	uint32_t size = 1024;
	while ( size < 4 * live )
		size *= 2;
	std::vector<unsigned long long> table( size, kPairEmpty );
	w->pairSet.forEach( [&]( uint64_t k ) {
		uint32_t i = (uint32_t)pair_hash_host( k ) & ( size - 1 );
		while ( table[i] != kPairEmpty )
			i = ( i + 1 ) & ( size - 1 );
		table[i] = k;
	} );
	if ( size != D.tableSize || D.table == nullptr )
	{
		if ( D.table )
		{
			HIP_CHECK( hipStreamSynchronize( s ) );
			HIP_CHECK( hipFree( D.table ) );
		}
		HIP_CHECK( hipMalloc( (void**)&D.table, sizeof( unsigned long long ) * size ) );
		D.tableSize = size;
	}
	HIP_CHECK( hipMemcpyAsync( D.table, table.data(), sizeof( unsigned long long ) * size, hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	D.tableTombstones = 0;
	D.tableValid = true;
	D.pendingKeys.clear();
}

// The device copy of tree t of world w (DevTreeNode form: leaves have child1 = -1 and child2 =
// their position in a query's depth-first order, child2's subtree first), uploaded on stream s
// when the tree changed since the last upload (the pinned staging buffer is rewritten only by the
// next upload of the same tree, after the stream has been synchronized).
static void syncDeviceTree( World* w, int t, hipStream_t s )
{
	DeviceBroadphase& D = w->devicePairs;
	const DynamicTree& tree = w->trees[t];
	const int n = (int)tree.nodes.size();
	if ( D.uploadedVersion[t] == tree.version && D.nodes[t] != nullptr && D.capNodes[t] >= n )
		return;
	grow( D.nodes[t], D.capNodes[t], std::max( n, 1 ) );
	if ( D.capStaging[t] < n )
	{
		// pinned, so the copy runs at full speed; the last update's copies have completed
		if ( D.staging[t] )
			HIP_CHECK( hipHostFree( D.staging[t] ) );
		D.capStaging[t] = std::max( n, 2 * D.capStaging[t] );
		HIP_CHECK( hipHostMalloc( (void**)&D.staging[t], sizeof( DevTreeNode ) * D.capStaging[t], hipHostMallocDefault ) );
	}
	DevTreeNode* staging = D.staging[t];
	for ( int i = 0; i < n; ++i )
	{
		const TreeNode& src = tree.nodes[i];
		staging[i] = DevTreeNode{ src.aabb.lowerBound.x, src.aabb.lowerBound.y, src.aabb.upperBound.x, src.aabb.upperBound.y,
								  src.height == 0 ? -1 : src.child1, src.height == 0 ? -1 : src.child2, src.userData, src.categoryBits };
	}
	int position = 0;
	D.walk.clear();
	if ( tree.root != -1 )
		D.walk.push_back( tree.root );
	while ( !D.walk.empty() )
	{
		const int id = D.walk.back();
		D.walk.pop_back();
		DevTreeNode& node = staging[id];
		if ( node.child1 == -1 )
		{
			node.child2 = position++;
			continue;
		}
		D.walk.push_back( node.child1 );
		D.walk.push_back( node.child2 );
	}
	if ( n > 0 )
		HIP_CHECK( hipMemcpyAsync( D.nodes[t], staging, sizeof( DevTreeNode ) * n, hipMemcpyHostToDevice, s ) );
	D.uploadedVersion[t] = tree.version;
}

// the device copy of world w's shape filters (by shape id), uploaded when shapes changed
static void syncShapeFilters( World* w, hipStream_t s )
{
	DeviceBroadphase& D = w->devicePairs;
	const int ns = (int)w->shapes.size();
	if ( D.uploadedShapes == w->shapesVersion && D.shapes != nullptr && D.capShapes >= ns )
		return;
	grow( D.shapes, D.capShapes, std::max( ns, 1 ) );
	std::vector<DevShapeFilter> filters( ns );
	for ( int i = 0; i < ns; ++i )
	{
		const Shape& sh = w->shapes[i];
		if ( sh.id < 0 )
			continue;
		filters[i] = DevShapeFilter{ sh.bodyId, (int)w->bodies[sh.bodyId].type, (uint32_t)sh.filter.categoryBits,
									 (uint32_t)sh.filter.maskBits, (int)sh.filter.groupIndex, 0 };
	}
	if ( ns > 0 )
		HIP_CHECK( hipMemcpyAsync( D.shapes, filters.data(), sizeof( DevShapeFilter ) * ns, hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	D.uploadedShapes = w->shapesVersion;
}

// src/broad_phase.c:333-436 with the queries on the device (k_find_pairs): the trees and shape
// filters are uploaded where they changed, the pair set brought up to date, and the new pairs come
// back with the index of the move that found them and their rank among that move's pairs; the
// contacts are then created here in the reference's order (moves in order, each move's pairs most
// recent first, as the reference walks its prepended lists).
static void updatePairsOnDevice( World* w )
{
	DeviceBroadphase& D = w->devicePairs;
	hipStream_t s = stream();
	const int moveCount = (int)w->moveArray.size();
	auto t0 = std::chrono::steady_clock::now();
	auto mark = [&]( int k ) { D.phaseMs[k] += std::chrono::duration<float, std::milli>( std::chrono::steady_clock::now() - t0 ).count(); };
	if ( D.counters == nullptr )
	{
		HIP_CHECK( hipMalloc( (void**)&D.counters, kPairCounters * sizeof( int ) ) );
		HIP_CHECK( hipHostMalloc( (void**)&D.hostCounters, kPairCounters * sizeof( int ), hipHostMallocDefault ) );
	}
	HIP_CHECK( hipMemsetAsync( D.counters, 0, kPairCounters * sizeof( int ), s ) );
	syncPairTable( w, s );
	mark( 0 );
	DevBroadphase bp{};
	int maxNodes = 1;
	for ( int t = 0; t < 3; ++t )
	{
		syncDeviceTree( w, t, s );
		maxNodes = std::max( maxNodes, (int)w->trees[t].nodes.size() );
		bp.nodes[t] = D.nodes[t];
		bp.roots[t] = w->trees[t].root;
	}
	syncShapeFilters( w, s );
	mark( 1 );
	// the moved-proxy bitmap covers every proxy key of the three trees (all zero between queries)
	const int words = ( 4 * maxNodes + 31 ) / 32;
	if ( D.moved == nullptr || D.capMovedWords < words )
	{
		grow( D.moved, D.capMovedWords, words );
		HIP_CHECK( hipMemsetAsync( D.moved, 0, sizeof( uint32_t ) * D.capMovedWords, s ) );
	}
	grow( D.moves, D.capMoves, moveCount );
	HIP_CHECK( hipMemcpyAsync( D.moves, w->moveArray.data(), sizeof( int ) * moveCount, hipMemcpyHostToDevice, s ) );
	bp.shapes = D.shapes;
	bp.moved = D.moved;
	bp.pairs = D.table;
	bp.pairMask = D.tableSize - 1;
	grow( D.pairs, D.capPairs, std::max( 1024, moveCount ) );
	grow( D.spills, D.capSpills, 4096 );
	HIP_CHECK( launch_mark_moved( D.moved, D.moves, moveCount, true, s ) );
	int found = 0;
	for ( int attempt = 0;; ++attempt )
	{
		HIP_CHECK( launch_find_pairs( bp, D.moves, moveCount, D.pairs, D.capPairs, D.counters, D.spills, D.capSpills, s ) );
		HIP_CHECK( hipMemcpyAsync( D.hostCounters, D.counters, kPairCounters * sizeof( int ), hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipStreamSynchronize( s ) );
		const int spilled = D.hostCounters[2];
		if ( spilled > 0 && spilled <= D.capSpills && D.hostCounters[1] == 0 )
		{
			HIP_CHECK( launch_find_pairs_spilled( bp, D.moves, spilled, D.pairs, D.capPairs, D.counters, D.spills, D.capSpills, s ) );
			HIP_CHECK( hipMemcpyAsync( D.hostCounters, D.counters, kPairCounters * sizeof( int ), hipMemcpyDeviceToHost, s ) );
			HIP_CHECK( hipStreamSynchronize( s ) );
			D.phaseMs[8] += (float)spilled;
		}
		if ( D.hostCounters[1] != 0 )
			fatal( "a device pair query overflowed its traversal stack or the pair table", __FILE__, __LINE__ );
		found = D.hostCounters[0];
		if ( found <= D.capPairs && spilled <= D.capSpills )
			break;
		B2MI_CHECK( attempt == 0 );
		// more pairs or spilled subtrees than the buffers hold: larger ones, and the queries once more
		grow( D.pairs, D.capPairs, found );
		grow( D.spills, D.capSpills, spilled );
		HIP_CHECK( hipMemsetAsync( D.counters, 0, kPairCounters * sizeof( int ), s ) );
	}
	mark( 2 );
	HIP_CHECK( launch_mark_moved( D.moved, D.moves, moveCount, false, s ) );
	D.found.resize( found );
	if ( found > 0 )
		HIP_CHECK( hipMemcpyAsync( D.found.data(), D.pairs, sizeof( PairRecord ) * found, hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	mark( 3 );
	std::sort( D.found.begin(), D.found.end(), []( const PairRecord& a, const PairRecord& b ) {
		return a.query != b.query ? a.query < b.query : a.rank > b.rank;
	} );
	const int ns = (int)w->shapes.size();
	for ( const PairRecord& p : D.found )
	{
		B2MI_CHECK( 0 <= p.shapeIdA && p.shapeIdA < ns && 0 <= p.shapeIdB && p.shapeIdB < ns );
		// the joint part of the pair filter (src/body.c:1731-1757), which the device query leaves
		// to the host
		const Body& ba = w->bodies[w->shapes[p.shapeIdA].bodyId];
		const Body& bb = w->bodies[w->shapes[p.shapeIdB].bodyId];
		if ( ( ba.jointCount > 0 || bb.jointCount > 0 ) && !jointsAllowCollision( w, ba, bb ) )
			continue;
		createContact( w, &w->shapes[p.shapeIdA], &w->shapes[p.shapeIdB] );
	}
	for ( int key : w->moveArray )
		w->moveSet.clear( (uint32_t)key );
	w->moveArray.clear();
	mark( 4 );
	D.phaseMs[5] += (float)found;
	D.phaseMs[6] += (float)moveCount;
	D.phaseMs[7] += 1.0f;
}

// src/broad_phase.c:333-436: the queries of the moved proxies (in parallel chunks when
// `parallel`), then contact creation in move order on this thread. A world stepped alone with many
// moved proxies queries on the device instead (updatePairsOnDevice).
static void updatePairs( World* w, bool parallel )
{
	int moveCount = (int)w->moveArray.size();
	if ( moveCount == 0 )
		return;
	if ( parallel && w->devicePairsMin >= 0 && moveCount >= w->devicePairsMin )
	{
		updatePairsOnDevice( w );
		return;
	}
	constexpr int kChunk = 128;
	int chunks = ( moveCount + kChunk - 1 ) / kChunk;
	std::vector<std::vector<MovePair>> pairs( chunks );
	std::vector<int> heads( moveCount, -1 );
	auto run = [&]( int c ) { findPairs( w, c * kChunk, std::min( moveCount, ( c + 1 ) * kChunk ), pairs[c], heads.data() ); };
	if ( parallel && chunks > 1 )
		hostPoolRun( chunks, run );
	else
		for ( int c = 0; c < chunks; ++c )
			run( c );
	for ( int i = 0; i < moveCount; ++i )
	{
		const std::vector<MovePair>& list = pairs[i / kChunk];
		for ( int p = heads[i]; p != -1; p = list[p].next )
			createContact( w, &w->shapes[list[p].a], &w->shapes[list[p].b] );
	}
	for ( int key : w->moveArray )
		w->moveSet.clear( (uint32_t)key );
	w->moveArray.clear();
}

// ---------------------------------------------------------------------------------------
// work items: unions of islands (connected components through touching contacts of the
// non-static, non-kinematic bodies), packed to a target size

static int findRoot( std::vector<int>& parent, int x )
{
	while ( parent[x] != x )
	{
		parent[x] = parent[parent[x]];
		x = parent[x];
	}
	return x;
}

// a body as a touching contact's solver reference: >= 0 a dynamic body (its awake sim index),
// <= -2 a kinematic body (-2 - its sim index), -1 a static body
static inline int awakeRef( const Body& b )
{
	return b.setIndex != kAwakeSet ? -1 : b.type == b2_dynamicBody ? b.localIndex : -2 - b.localIndex;
}

// A touching contact for the item builder: its id and its two bodies' references.
struct Src
{
	int id, a, b;
};

// Runs fn over `parts` index ranges of [0, n), on the host pool when `parallel` (the caller is
// then not a pool task), each range writing only its own outputs.
static void forRanges( bool parallel, int n, const std::function<void( int, int )>& fn )
{
	const int parts = parallel ? 64 : 1;
	auto part = [&]( int p ) { fn( (int)( (int64_t)n * p / parts ), (int)( (int64_t)n * ( p + 1 ) / parts ) ); };
	if ( parallel )
		hostPoolRun( parts, part );
	else
		part( 0 );
}

// The components of a set of awake bodies and their touching contacts, as the item builder takes
// them. Component k's bodies, in ascending sim index, are compBodies[compBodyStart[k] ..
// compBodyStart[k+1]); the contacts are src in colour-list order (colour c at colorStart[c] ..
// colorStart[c+1]); per colour they are ordered by component (a stable counting sort, so each
// component keeps the colour list's order): component k's contacts of colour c are
// src[bySlot[colorStart[c] + p]] for p in compStart[c][k] .. compStart[c][k+1], where compStart[c]
// starts at compStart.data() + c * (nc + 1).
struct ComponentSet
{
	int nc = 0;
	std::vector<int> compBodyStart, compBodies, compSlots;
	std::vector<Src> src;
	int colorStart[kColorCount + 1] = {};
	std::vector<int> compStart, bySlot;
	// the awake joints, laid out as the contacts (colour-list order, then by component); a joint
	// belongs to the component of its first dynamic body
	std::vector<Src> jsrc;
	int jColorStart[kColorCount + 1] = {};
	std::vector<int> jCompStart, jBySlot;
	int compSize( int k ) const
	{
		return compBodyStart[k + 1] - compBodyStart[k];
	}
};

// Fills the components of cs for the bodies `sims` (ascending sim indices) and the contacts
// already in cs.src / cs.colorStart: union-find over the dynamic-dynamic contacts, components in
// order of their smallest sim index (a root is the smallest position of its set, so it is numbered
// before the rest of it); a kinematic body is a component of its own. posOf maps each sim of
// `sims` to its position there.
static void makeComponents( ComponentSet& cs, const std::vector<int>& sims, const std::vector<int>& posOf, bool parallel )
{
	const int m = (int)sims.size();
	std::vector<int> parent( m );
	for ( int i = 0; i < m; ++i )
		parent[i] = i;
	for ( const std::vector<Src>* list : { &cs.src, &cs.jsrc } )
		for ( const Src& e : *list )
			if ( e.a >= 0 && e.b >= 0 )
			{
				int pa = posOf[e.a], pb = posOf[e.b];
				B2MI_CHECK( pa >= 0 && pb >= 0 );
				int ra = findRoot( parent, pa ), rb = findRoot( parent, pb );
				if ( ra != rb )
					parent[std::max( ra, rb )] = std::min( ra, rb );
			}
	std::vector<int> compOf( m );
	int nc = 0;
	for ( int i = 0; i < m; ++i )
	{
		int r = findRoot( parent, i );
		compOf[i] = r == i ? nc++ : compOf[r];
	}
	cs.nc = nc;
	cs.compBodyStart.assign( nc + 1, 0 );
	cs.compBodies.resize( m );
	cs.compSlots.assign( nc, 0 );
	for ( int i = 0; i < m; ++i )
		cs.compBodyStart[compOf[i] + 1] += 1;
	for ( int k = 0; k < nc; ++k )
		cs.compBodyStart[k + 1] += cs.compBodyStart[k];
	{
		std::vector<int> fill( cs.compBodyStart.begin(), cs.compBodyStart.end() - 1 );
		for ( int i = 0; i < m; ++i )
			cs.compBodies[fill[compOf[i]]++] = sims[i];
	}
	// a contact belongs to the component of its dynamic body
	const int ns = (int)cs.src.size();
	cs.compStart.assign( (size_t)kColorCount * ( nc + 1 ), 0 );
	cs.bySlot.resize( ns );
	std::vector<int> compOfSrc( ns );
	forRanges( parallel, ns, [&]( int lo, int hi ) {
		for ( int q = lo; q < hi; ++q )
		{
			int a = cs.src[q].a >= 0 ? cs.src[q].a : cs.src[q].b;
			B2MI_CHECK( a >= 0 && posOf[a] >= 0 );
			compOfSrc[q] = compOf[posOf[a]];
		}
	} );
	auto sortColor = [&]( int c ) {
		int* st = cs.compStart.data() + (size_t)c * ( nc + 1 );
		for ( int q = cs.colorStart[c]; q < cs.colorStart[c + 1]; ++q )
			st[compOfSrc[q] + 1] += 1;
		for ( int k = 0; k < nc; ++k )
			st[k + 1] += st[k];
		std::vector<int> fill( st, st + nc );
		for ( int q = cs.colorStart[c]; q < cs.colorStart[c + 1]; ++q )
			cs.bySlot[cs.colorStart[c] + fill[compOfSrc[q]]++] = q;
	};
	if ( parallel )
		hostPoolRun( kColorCount, sortColor );
	else
		for ( int c = 0; c < kColorCount; ++c )
			sortColor( c );
	for ( int c = 0; c < kColorCount; ++c )
	{
		const int* st = cs.compStart.data() + (size_t)c * ( nc + 1 );
		for ( int k = 0; k < nc; ++k )
			cs.compSlots[k] += st[k + 1] - st[k];
	}
	// the joints, sorted the same way (few; serial)
	const int nj = (int)cs.jsrc.size();
	cs.jCompStart.assign( (size_t)kColorCount * ( nc + 1 ), 0 );
	cs.jBySlot.resize( nj );
	std::vector<int> compOfJoint( nj );
	for ( int q = 0; q < nj; ++q )
	{
		int a = cs.jsrc[q].a >= 0 ? cs.jsrc[q].a : cs.jsrc[q].b;
		B2MI_CHECK( a >= 0 && posOf[a] >= 0 );
		compOfJoint[q] = compOf[posOf[a]];
	}
	for ( int c = 0; c < kColorCount; ++c )
	{
		int* st = cs.jCompStart.data() + (size_t)c * ( nc + 1 );
		for ( int q = cs.jColorStart[c]; q < cs.jColorStart[c + 1]; ++q )
			st[compOfJoint[q] + 1] += 1;
		for ( int k = 0; k < nc; ++k )
			st[k + 1] += st[k];
		std::vector<int> fill( st, st + nc );
		for ( int q = cs.jColorStart[c]; q < cs.jColorStart[c + 1]; ++q )
			cs.jBySlot[cs.jColorStart[c] + fill[compOfJoint[q]]++] = q;
		for ( int k = 0; k < nc; ++k )
			cs.compSlots[k] += st[k + 1] - st[k];
	}
}

// One work item built into lists of its own (body, slot and level indices relative to the item).
struct BuiltItem
{
	WorkItem wi;
	std::vector<int> bodies, slots, levels;
	std::vector<int2> slotBodies;
	int maxGroup = 0;
	int jointLevelAt = 0; // where the joints' overflow levels start in `levels`
};

// Packs the components of cs in order into items of up to `target` constraint slots (and 4 x
// target bodies) and builds each item: owned bodies, kinematic shadows, slots grouped as
// overflow first then colours 0..10, overflow levels, workgroup size, the next-colour table and
// the store for the LDS budget. posOfSim (one entry per awake sim, all -1) is the item-local
// position of each owned sim while its item is built: items own disjoint bodies, and each clears
// its entries when done.
static std::vector<BuiltItem> packItems( const World* w, const ComponentSet& cs, int target, int budget, int fullBudget,
										 std::vector<int>& posOfSim, bool parallel )
{
	std::vector<int2> spans;
	for ( int ci = 0; ci < cs.nc; )
	{
		int cj = ci;
		int slots = 0, bodies = 0;
		do
		{
			slots += cs.compSlots[cj];
			bodies += cs.compSize( cj );
			cj += 1;
		} while ( cj < cs.nc && slots + cs.compSlots[cj] <= target && bodies + cs.compSize( cj ) <= 4 * target );
		spans.push_back( make_int2( ci, cj ) );
		ci = cj;
	}
	const int ni = (int)spans.size();
	std::vector<BuiltItem> built( ni );
	auto buildItem = [&]( int item ) {
		BuiltItem& out = built[item];
		const int ci = spans[item].x, cj = spans[item].y;
		WorkItem& wi = out.wi;
		wi = WorkItem{};
		wi.world = w->groupIndex;
		// owned bodies
		for ( int k = ci; k < cj; ++k )
			for ( int q = cs.compBodyStart[k]; q < cs.compBodyStart[k + 1]; ++q )
			{
				posOfSim[cs.compBodies[q]] = (int)out.bodies.size();
				out.bodies.push_back( cs.compBodies[q] );
			}
		wi.ownedCount = (int)out.bodies.size();
		// local index of a body reference: owned bodies by position, kinematic bodies as read-only
		// shadows appended on first use, static bodies -1
		std::vector<std::pair<int, int>> shadowMap;
		auto localIndex = [&]( int r ) -> int {
			if ( r >= 0 )
			{
				// a dynamic body of the item's contacts is one of its own
				B2MI_CHECK( posOfSim[r] >= 0 );
				return posOfSim[r];
			}
			if ( r == -1 )
				return -1;
			int sim = -2 - r;
			for ( auto& pr : shadowMap )
				if ( pr.first == sim )
					return pr.second;
			int li = (int)out.bodies.size();
			out.bodies.push_back( sim );
			shadowMap.push_back( { sim, li } );
			return li;
		};
		int itemMaxGroup = 0;
		int groupSize[kGroups] = {};
		// group 0 = overflow (colour 11) then colours 0..10
		for ( int gidx = 0; gidx < kGroups; ++gidx )
		{
			int color = gidx == 0 ? kOverflow : gidx - 1;
			wi.groupStart[gidx] = (int)out.slots.size();
			const int* st = cs.compStart.data() + (size_t)color * ( cs.nc + 1 );
			for ( int k = ci; k < cj; ++k )
				for ( int p = st[k]; p < st[k + 1]; ++p )
				{
					const Src& e = cs.src[cs.bySlot[cs.colorStart[color] + p]];
					int la = localIndex( e.a );
					int lb = localIndex( e.b );
					out.slots.push_back( e.id );
					out.slotBodies.push_back( make_int2( la, lb ) );
				}
			groupSize[gidx] = (int)out.slots.size() - wi.groupStart[gidx];
		}
		wi.groupStart[kGroups] = (int)out.slots.size();
		// Overflow levels (src/contact_solver.c:212-345 and src/joint.c:1106-1163 run the overflow
		// group serially): a slot's level is one past the last level that changed one of its bodies,
		// and the slots are ordered by level, keeping their order within a level. Static bodies are
		// never changed; nor are kinematic ones, whose zero inverse mass and inertia leave their
		// velocity bits as they are, unless a component is -0 (then they count as changed).
		// The joints' levels come after the contacts' (all overflow joints run first).
		auto changed = [&]( int li ) {
			if ( li < 0 )
				return false;
			if ( li < wi.ownedCount )
				return true;
			const float* st = &w->awakeState[8 * (size_t)out.bodies[li]];
			for ( int c = 0; c < 3; ++c )
				if ( st[c] == 0.0f && std::signbit( st[c] ) )
					return true;
			return false;
		};
		// orders slots [first, end) by level and appends the level boundaries; returns the count
		auto makeLevels = [&]( int first, int end ) -> int {
			if ( end <= first )
				return 0;
			const int nb = (int)out.bodies.size();
			std::vector<int> last( nb, -1 ), level( end - first );
			int maxLevel = 0;
			for ( int k = first; k < end; ++k )
			{
				int2 lb = out.slotBodies[k];
				bool ca = changed( lb.x ), cb = changed( lb.y );
				int lv = 0;
				if ( ca )
					lv = std::max( lv, last[lb.x] + 1 );
				if ( cb )
					lv = std::max( lv, last[lb.y] + 1 );
				if ( ca )
					last[lb.x] = lv;
				if ( cb )
					last[lb.y] = lv;
				level[k - first] = lv;
				maxLevel = std::max( maxLevel, lv );
			}
			std::vector<int> order( end - first );
			for ( int k = 0; k < end - first; ++k )
				order[k] = k;
			std::stable_sort( order.begin(), order.end(), [&]( int a, int b ) { return level[a] < level[b]; } );
			std::vector<int> ids( end - first );
			std::vector<int2> bodies( end - first );
			for ( int k = 0; k < end - first; ++k )
			{
				ids[k] = out.slots[first + order[k]];
				bodies[k] = out.slotBodies[first + order[k]];
			}
			std::copy( ids.begin(), ids.end(), out.slots.begin() + first );
			std::copy( bodies.begin(), bodies.end(), out.slotBodies.begin() + first );
			const size_t before = out.levels.size();
			for ( int k = 0; k < end - first; ++k )
				if ( k == 0 || level[order[k]] != level[order[k - 1]] )
					out.levels.push_back( first + k );
			out.levels.push_back( end );
			B2MI_CHECK( (int)( out.levels.size() - before ) == maxLevel + 2 );
			return maxLevel + 1;
		};
		wi.ovfLevelCount = makeLevels( wi.groupStart[0], wi.groupStart[1] );
		// the joints, grouped as the contacts after them: a static body as -1 - its static sim index
		for ( int gidx = 0; gidx < kGroups; ++gidx )
		{
			int color = gidx == 0 ? kOverflow : gidx - 1;
			wi.jointGroupStart[gidx] = (int)out.slots.size();
			const int* st = cs.jCompStart.data() + (size_t)color * ( cs.nc + 1 );
			for ( int k = ci; k < cj; ++k )
				for ( int p = st[k]; p < st[k + 1]; ++p )
				{
					const Src& e = cs.jsrc[cs.jBySlot[cs.jColorStart[color] + p]];
					const Joint& jt = w->joints[e.id];
					int lb[2];
					for ( int side = 0; side < 2; ++side )
					{
						const Body& body = w->bodies[jt.edges[side].bodyId];
						lb[side] = body.setIndex == kStaticSet ? -1 - body.localIndex : localIndex( side == 0 ? e.a : e.b );
					}
					out.slots.push_back( e.id );
					out.slotBodies.push_back( make_int2( lb[0], lb[1] ) );
				}
			groupSize[gidx] += (int)out.slots.size() - wi.jointGroupStart[gidx];
		}
		wi.jointGroupStart[kGroups] = (int)out.slots.size();
		out.jointLevelAt = (int)out.levels.size();
		wi.jointLevelCount = makeLevels( wi.jointGroupStart[0], wi.jointGroupStart[1] );
		for ( int k = ci; k < cj; ++k )
			for ( int q = cs.compBodyStart[k]; q < cs.compBodyStart[k + 1]; ++q )
				posOfSim[cs.compBodies[q]] = -1;
		for ( int g = 0; g < kGroups; ++g )
			itemMaxGroup = std::max( itemMaxGroup, groupSize[g] );
		// per-item workgroup: one thread per slot of its largest group (a colour's contacts and
		// joints run in the same phase)
		wi.block = 64;
		while ( wi.block < itemMaxGroup && wi.block < kSolveMaxBlock )
			wi.block *= 2;
		wi.bodyCount = (int)out.bodies.size();
		wi.nextGroup[0] = wi.nextGroup[1] = 0;
		for ( int g = 1; g < kGroups; ++g )
			for ( int step = 1; step < kGroups; ++step )
			{
				int n = 1 + ( g - 1 + step ) % ( kGroups - 1 );
				if ( wi.groupStart[n] < wi.groupStart[n + 1] )
				{
					wi.nextGroup[g >> 3] |= n << ( 4 * ( g & 7 ) );
					break;
				}
			}
		{
			int bytes = itemLdsBytes( wi.bodyCount, wi.groupStart[kGroups] );
			int bodyBytes = itemBodyLdsBytes( wi.bodyCount, wi.groupStart[kGroups] );
			if ( bytes <= budget && bytes <= fullBudget )
			{
				wi.store = kStoreLds;
				wi.ldsBytes = bytes;
			}
			else if ( budget > 0 && bodyBytes <= budget )
			{
				wi.store = kStoreBodiesLds;
				wi.ldsBytes = bodyBytes;
			}
			else
			{
				wi.store = kStoreHbm;
				wi.ldsBytes = 0;
			}
		}
		out.maxGroup = std::max( itemMaxGroup, wi.ownedCount );
	};
	forRanges( parallel, ni, [&]( int lo, int hi ) {
		for ( int i = lo; i < hi; ++i )
			buildItem( i );
	} );
	return built;
}

// Appends built items to the world's lists (their offsets relative to the lists), records which
// item owns each of their bodies, and refreshes the lists' maxima over the live items.
static void appendItems( World* w, std::vector<BuiltItem>& built, bool parallel )
{
	SolverLists& L = w->solver;
	int nb = (int)L.bodyList.size(), ns = (int)L.slotContact.size(), nl = (int)L.ovfLevels.size();
	const int firstItem = (int)L.items.size();
	for ( BuiltItem& b : built )
	{
		b.wi.bodyStart = nb;
		b.wi.slotStart = ns;
		b.wi.ovfLevelBegin = nl;
		b.wi.jointLevelBegin = nl + b.jointLevelAt;
		nb += (int)b.bodies.size();
		ns += (int)b.slots.size();
		nl += (int)b.levels.size();
	}
	L.bodyList.resize( nb );
	L.slotContact.resize( ns );
	L.slotBodies.resize( ns );
	L.ovfLevels.resize( nl );
	L.items.resize( firstItem + built.size() );
	L.live.resize( firstItem + built.size(), 1 );
	forRanges( parallel, (int)built.size(), [&]( int lo, int hi ) {
		for ( int i = lo; i < hi; ++i )
		{
			const BuiltItem& b = built[i];
			std::copy( b.bodies.begin(), b.bodies.end(), L.bodyList.begin() + b.wi.bodyStart );
			std::copy( b.slots.begin(), b.slots.end(), L.slotContact.begin() + b.wi.slotStart );
			std::copy( b.slotBodies.begin(), b.slotBodies.end(), L.slotBodies.begin() + b.wi.slotStart );
			std::copy( b.levels.begin(), b.levels.end(), L.ovfLevels.begin() + b.wi.ovfLevelBegin );
			L.items[firstItem + i] = b.wi;
			for ( int k = 0; k < b.wi.ownedCount; ++k )
				L.itemOfSim[b.bodies[k]] = firstItem + i;
		}
	} );
	L.maxGroupSize = 0;
	L.maxItemLds = 0;
	for ( size_t i = 0; i < L.items.size(); ++i )
		if ( L.live[i] )
		{
			int maxGroup = L.items[i].ownedCount;
			for ( int g = 0; g < kGroups; ++g )
				maxGroup = std::max( maxGroup, L.items[i].groupStart[g + 1] - L.items[i].groupStart[g] + L.items[i].jointGroupStart[g + 1] -
													L.items[i].jointGroupStart[g] );
			L.maxGroupSize = std::max( L.maxGroupSize, maxGroup );
			L.maxItemLds = std::max( L.maxItemLds, L.items[i].ldsBytes );
		}
}

// Every index the solver kernel dereferences for the items [first, end) of the world's lists,
// checked on the host.
static void checkItems( const World* w, int first, int end, bool parallel )
{
	const SolverLists& L = w->solver;
	const int na = (int)w->awakeSimBody.size();
	const int nst = (int)w->staticSimBody.size();
	forRanges( parallel, end - first, [&]( int lo, int hi ) {
		for ( int it = first + lo; it < first + hi; ++it )
		{
			const WorkItem& wi = L.items[it];
			B2MI_CHECK( wi.bodyStart >= 0 && wi.bodyStart + wi.bodyCount <= (int)L.bodyList.size() );
			B2MI_CHECK( 0 <= wi.ownedCount && wi.ownedCount <= wi.bodyCount );
			for ( int i = 0; i < wi.bodyCount; ++i )
				B2MI_CHECK( 0 <= L.bodyList[wi.bodyStart + i] && L.bodyList[wi.bodyStart + i] < na );
			B2MI_CHECK( wi.slotStart >= 0 && wi.slotStart + item_slots( wi ) <= (int)L.slotContact.size() );
			B2MI_CHECK( wi.jointGroupStart[0] == wi.groupStart[kGroups] );
			for ( int k = 0; k < kGroups; ++k )
				B2MI_CHECK( wi.jointGroupStart[k] <= wi.jointGroupStart[k + 1] );
			for ( int k = wi.jointGroupStart[0]; k < item_slots( wi ); ++k )
			{
				// a joint's bodies: item-local, or -1 - a static sim index
				int2 lb = L.slotBodies[wi.slotStart + k];
				B2MI_CHECK( lb.x < wi.bodyCount && lb.y < wi.bodyCount && ( lb.x >= 0 || -1 - lb.x < nst ) && ( lb.y >= 0 || -1 - lb.y < nst ) );
				B2MI_CHECK( 0 <= L.slotContact[wi.slotStart + k] && L.slotContact[wi.slotStart + k] < w->jointIdPool.nextIndex );
			}
			B2MI_CHECK( wi.ldsBytes <= kMaxItemLdsBytes );
			B2MI_CHECK( ( wi.store == kStoreHbm && wi.ldsBytes == 0 ) ||
						( wi.store == kStoreBodiesLds && wi.ldsBytes == itemBodyLdsBytes( wi.bodyCount, wi.groupStart[kGroups] ) ) ||
						( wi.store == kStoreLds && wi.ldsBytes == itemLdsBytes( wi.bodyCount, wi.groupStart[kGroups] ) ) );
			for ( int k = 0; k < kGroups; ++k )
				B2MI_CHECK( wi.groupStart[k] <= wi.groupStart[k + 1] );
			for ( int g = 1; g < kGroups; ++g )
				B2MI_CHECK( 0 <= next_group( wi, g ) && next_group( wi, g ) < kGroups );
			for ( int k = 0; k < wi.groupStart[kGroups]; ++k )
			{
				int2 lb = L.slotBodies[wi.slotStart + k];
				B2MI_CHECK( -1 <= lb.x && lb.x < wi.bodyCount && -1 <= lb.y && lb.y < wi.bodyCount );
				B2MI_CHECK( 0 <= L.slotContact[wi.slotStart + k] && L.slotContact[wi.slotStart + k] < w->capContacts );
			}
		}
	} );
}

// All work items of one world: islands (connected components of dynamic bodies through touching
// contacts; static and kinematic bodies do not join islands) packed in order up to `target`
// constraint slots, each item's slots grouped as overflow first, then colours 0..10. Writes
// w->solver with indices local to its lists; runs on any host thread (touches this world only).
// `parallel`: the stages run over index ranges on the host pool (the caller is not a pool task:
// the pool does not nest); each range writes only its own outputs, so the lists are the same
// either way.
static void buildWorldItems( World* w, int target, int budget, int fullBudget, bool parallel )
{
	SolverLists& L = w->solver;
	L = SolverLists{};
	w->dirtySims.clear();
	const int na = (int)w->awakeSimBody.size();
	L.itemOfSim.assign( na, -1 );
	w->posOfSim.assign( na, -1 );
	w->topologyDirty = false;
	w->solverPending = true;
	w->pendingBodyFrom = w->pendingSlotFrom = 0;
	if ( na == 0 )
		return;
	// What the rest needs of each touching contact, in colour-list order (overflow last), gathered
	// by one pass over the contact array (a contact's colour and position in that colour's list
	// are stored in it), not by following the colour lists into the contact array.
	ComponentSet cs;
	cs.colorStart[0] = 0;
	for ( int c = 0; c < kColorCount; ++c )
		cs.colorStart[c + 1] = cs.colorStart[c] + (int)w->colors[c].size();
	cs.src.resize( cs.colorStart[kColorCount] );
	{
		std::atomic<int> placed{ 0 };
		forRanges( parallel, (int)w->contacts.size(), [&]( int lo, int hi ) {
			int n = 0;
			for ( int cid = lo; cid < hi; ++cid )
			{
				const Contact& ct = w->contacts[cid];
				if ( ct.contactId < 0 || ct.colorIndex < 0 )
					continue;
				const int c = ct.colorIndex, li = ct.localIndex;
				B2MI_CHECK( c < kColorCount && 0 <= li && li < (int)w->colors[c].size() && w->colors[c][li] == cid );
				cs.src[cs.colorStart[c] + li] =
					Src{ cid, awakeRef( w->bodies[ct.edges[0].bodyId] ), awakeRef( w->bodies[ct.edges[1].bodyId] ) };
				n += 1;
			}
			placed.fetch_add( n );
		} );
		// each list position names exactly one contact, so all of them were written
		B2MI_CHECK( placed.load() == cs.colorStart[kColorCount] );
	}
	// the awake joints in colour-list order
	cs.jColorStart[0] = 0;
	for ( int c = 0; c < kColorCount; ++c )
		cs.jColorStart[c + 1] = cs.jColorStart[c] + (int)w->jointColors[c].size();
	cs.jsrc.resize( cs.jColorStart[kColorCount] );
	for ( int c = 0; c < kColorCount; ++c )
		for ( size_t li = 0; li < w->jointColors[c].size(); ++li )
		{
			const Joint& jt = w->joints[w->jointColors[c][li]];
			cs.jsrc[cs.jColorStart[c] + li] =
				Src{ jt.jointId, awakeRef( w->bodies[jt.edges[0].bodyId] ), awakeRef( w->bodies[jt.edges[1].bodyId] ) };
		}
	std::vector<int> sims( na ), posOf( na );
	for ( int i = 0; i < na; ++i )
		sims[i] = posOf[i] = i;
	makeComponents( cs, sims, posOf, parallel );
	std::vector<BuiltItem> built = packItems( w, cs, target, budget, fullBudget, w->posOfSim, parallel );
	appendItems( w, built, parallel );
	checkItems( w, 0, (int)L.items.size(), parallel );
	L.liveBodies = (int)L.bodyList.size();
	L.liveSlots = (int)L.slotContact.size();
}

// Rebuilds only the work items that own a body of w->dirtySims: their bodies are regathered into
// components with the touching contacts they own (the contacts whose first dynamic body is one
// of them), packed into new items appended to the world's lists, and the old items retired.
// Items not touched keep their lists: a colour's slots share no non-static body, so the order
// of slots within a colour does not change any result, and each item's overflow slots keep the
// overflow list's order. Falls back to a full rebuild when the retired items' holes would make up
// more than half the lists.
static void rebuildDirtyItems( World* w, int target, int budget, int fullBudget, bool parallel )
{
	SolverLists& L = w->solver;
	const int na = (int)w->awakeSimBody.size();
	if ( (int)L.itemOfSim.size() != na || (int)w->posOfSim.size() != na )
	{
		buildWorldItems( w, target, budget, fullBudget, parallel );
		return;
	}
	// the items to retire
	std::vector<int> dead;
	for ( int s : w->dirtySims )
	{
		B2MI_CHECK( 0 <= s && s < na );
		int it = L.itemOfSim[s];
		if ( it >= 0 && L.live[it] )
		{
			L.live[it] = 0;
			dead.push_back( it );
		}
	}
	w->dirtySims.clear();
	if ( dead.empty() )
		return;
	// their owned bodies, ascending
	std::vector<int> sims;
	int deadBodies = 0, deadSlots = 0;
	for ( int it : dead )
	{
		const WorkItem& wi = L.items[it];
		for ( int i = 0; i < wi.ownedCount; ++i )
			sims.push_back( L.bodyList[wi.bodyStart + i] );
		deadBodies += wi.bodyCount;
		deadSlots += item_slots( wi );
	}
	if ( 2 * ( L.liveBodies - deadBodies ) < (int)L.bodyList.size() || 2 * ( L.liveSlots - deadSlots ) < (int)L.slotContact.size() )
	{
		buildWorldItems( w, target, budget, fullBudget, parallel );
		return;
	}
	std::sort( sims.begin(), sims.end() );
	std::vector<int>& posOf = L.posInSet; // all -1 between calls
	if ( (int)posOf.size() != na )
		posOf.assign( na, -1 );
	for ( int i = 0; i < (int)sims.size(); ++i )
		posOf[sims[i]] = i;
	// the touching contacts these bodies own, in colour-list order
	ComponentSet cs;
	struct Key
	{
		int color, local;
		Src s;
	};
	std::vector<Key> keys;
	for ( int sim : sims )
	{
		const Body& body = w->bodies[w->awakeSimBody[sim]];
		for ( int key = body.headContactKey; key != -1; )
		{
			const Contact& c = w->contacts[key >> 1];
			const int next = c.edges[key & 1].nextKey;
			if ( c.colorIndex >= 0 )
			{
				const Body& a = w->bodies[c.edges[0].bodyId];
				const Body& b = w->bodies[c.edges[1].bodyId];
				int ra = awakeRef( a ), rb = awakeRef( b );
				int owner = ra >= 0 ? ra : rb;
				if ( owner == sim )
					keys.push_back( Key{ c.colorIndex, c.localIndex, Src{ c.contactId, ra, rb } } );
			}
			key = next;
		}
	}
	std::sort( keys.begin(), keys.end(),
			   []( const Key& x, const Key& y ) { return x.color != y.color ? x.color < y.color : x.local < y.local; } );
	cs.src.resize( keys.size() );
	for ( int c = 0; c <= kColorCount; ++c )
		cs.colorStart[c] = 0;
	for ( size_t q = 0; q < keys.size(); ++q )
	{
		cs.src[q] = keys[q].s;
		cs.colorStart[keys[q].color + 1] += 1;
	}
	for ( int c = 0; c < kColorCount; ++c )
		cs.colorStart[c + 1] += cs.colorStart[c];
	// and the joints these bodies own (a joint's owner is its first dynamic body), in colour-list
	// order
	keys.clear();
	for ( int sim : sims )
	{
		const Body& body = w->bodies[w->awakeSimBody[sim]];
		for ( int key = body.headJointKey; key != -1; )
		{
			const Joint& jt = w->joints[key >> 1];
			const int next = jt.edges[key & 1].nextKey;
			if ( jt.setIndex == kAwakeSet )
			{
				int ra = awakeRef( w->bodies[jt.edges[0].bodyId] ), rb = awakeRef( w->bodies[jt.edges[1].bodyId] );
				if ( ( ra >= 0 ? ra : rb ) == sim )
					keys.push_back( Key{ jt.colorIndex, jt.localIndex, Src{ jt.jointId, ra, rb } } );
			}
			key = next;
		}
	}
	std::sort( keys.begin(), keys.end(),
			   []( const Key& x, const Key& y ) { return x.color != y.color ? x.color < y.color : x.local < y.local; } );
	cs.jsrc.resize( keys.size() );
	for ( int c = 0; c <= kColorCount; ++c )
		cs.jColorStart[c] = 0;
	for ( size_t q = 0; q < keys.size(); ++q )
	{
		cs.jsrc[q] = keys[q].s;
		cs.jColorStart[keys[q].color + 1] += 1;
	}
	for ( int c = 0; c < kColorCount; ++c )
		cs.jColorStart[c + 1] += cs.jColorStart[c];
	makeComponents( cs, sims, posOf, false );
	for ( int sim : sims )
		posOf[sim] = -1;
	std::vector<BuiltItem> built = packItems( w, cs, target, budget, fullBudget, w->posOfSim, false );
	const int firstNew = (int)L.items.size();
	const int bodyFrom = (int)L.bodyList.size(), slotFrom = (int)L.slotContact.size();
	appendItems( w, built, false );
	checkItems( w, firstNew, (int)L.items.size(), false );
	for ( const BuiltItem& b : built )
	{
		deadBodies -= (int)b.bodies.size();
		deadSlots -= (int)b.slots.size();
	}
	L.liveBodies -= deadBodies;
	L.liveSlots -= deadSlots;
	if ( !w->solverPending )
	{
		w->pendingBodyFrom = bodyFrom;
		w->pendingSlotFrom = slotFrom;
	}
	else
	{
		w->pendingBodyFrom = std::min( w->pendingBodyFrom, bodyFrom );
		w->pendingSlotFrom = std::min( w->pendingSlotFrom, slotFrom );
	}
	w->solverPending = true;
}

// Every contact of the world (touching ones in the colours, the others in nonTouching), in
// contact-id order: one sequential pass over the contact array. Ids are handed out as the pair
// update finds new pairs, moved proxy by moved proxy, so neighbouring ids tend to share bodies
// and shapes. The narrowphase's state changes are sorted by contact id afterwards, so this order
// does not affect results.
static void buildWorldCollide( World* w )
{
	w->collideIds.clear();
	const int n = (int)w->contacts.size();
	for ( int cid = 0; cid < n; ++cid )
		if ( w->contacts[cid].contactId >= 0 )
			w->collideIds.push_back( cid );
	size_t expected = w->nonTouching.size();
	for ( int c = 0; c < kColorCount; ++c )
		expected += w->colors[c].size();
	B2MI_CHECK( w->collideIds.size() == expected );
	w->contactsDirty = false;
	w->collidePending = true;
}

static int withSlack( int n )
{
	return n + n / 4 + 8;
}

// Assigns every world a region (with slack) in the launch arrays and sizes the device arrays.
static void layoutGroup( StepGroup* g )
{
	int bodies = 0, slots = 0, collide = 0;
	for ( World* w : g->worlds )
	{
		w->bodyRegion = Region{ bodies, withSlack( (int)w->solver.bodyList.size() ) };
		w->slotRegion = Region{ slots, withSlack( (int)w->solver.slotContact.size() ) };
		w->collideRegion = Region{ collide, withSlack( (int)w->collideIds.size() ) };
		bodies += w->bodyRegion.cap;
		slots += w->slotRegion.cap;
		collide += w->collideRegion.cap;
		w->solverPending = true;
		w->pendingBodyFrom = w->pendingSlotFrom = 0;
		w->collidePending = true;
	}
	g->bodyListCount = bodies;
	g->slotCount = slots;
	g->collideCount = collide;
	g->hBodyList.assign( bodies, 0 );
	g->hSlotContact.assign( slots, 0 );
	g->hSlotBodies.assign( slots, make_int2( -1, -1 ) );
	g->hCollide.assign( collide, make_int4( -1, -1, -1, -1 ) );
	grow( g->dBodyList, g->capBodyList, std::max( bodies, 1 ) );
	int slotCap = g->capSlots;
	grow( g->dSlotContact, slotCap, std::max( slots, 1 ) );
	grow( g->dSlotBodies, g->capSlots, std::max( slots, 1 ) );
	grow( g->dCollideList, g->capCollide, std::max( collide, 1 ) );
	grow( g->dChanges, g->capChanges, std::max( collide, 1 ) );
	if ( g->capCons < std::max( slots, 1 ) || g->dCons == nullptr )
	{
		if ( g->dCons )
		{
			HIP_CHECK( hipStreamSynchronize( stream() ) );
			HIP_CHECK( hipFree( g->dCons ) );
		}
		g->capCons = std::max( 64, slots + slots / 4 );
		HIP_CHECK( hipMalloc( (void**)&g->dCons, sizeof( float ) * (size_t)cfFloatCount * g->capCons ) );
	}
	grow( g->dPriv, g->capPriv, 2 * std::max( bodies, 1 ) );
	grow( g->dMass, g->capMass, std::max( bodies, 1 ) );
	g->layoutValid = true;
}

// Every range a solver launch dereferences, checked on the host against the device arrays
// before the items are uploaded: each item's body entries (body list, private state, masses)
// and slots (slot lists, constraint records of either record stride) lie inside its world's
// regions and the arrays' capacities, no two items share a body entry or a slot, overflow level
// boundaries lie inside the item's overflow group, and an LDS-store item's footprint fits the
// launch. A work item's own indices (local bodies of each slot, sim indices) are checked where
// it is built.
static void validateItems( const StepGroup* g )
{
	const int nw = (int)g->worlds.size();
	std::vector<int2> bodySpans, slotSpans;
	bodySpans.reserve( g->hItems.size() );
	slotSpans.reserve( g->hItems.size() );
	for ( const WorkItem& wi : g->hItems )
	{
		B2MI_CHECK( 0 <= wi.world && wi.world < nw );
		const World* w = g->worlds[wi.world];
		const int total = item_slots( wi );
		B2MI_CHECK( wi.groupStart[0] == 0 && total >= 0 && wi.bodyCount >= 0 && wi.jointGroupStart[0] == wi.groupStart[kGroups] );
		B2MI_CHECK( w->bodyRegion.base <= wi.bodyStart && wi.bodyStart + wi.bodyCount <= w->bodyRegion.base + w->bodyRegion.cap );
		B2MI_CHECK( wi.bodyStart + wi.bodyCount <= g->capBodyList && 2 * ( wi.bodyStart + wi.bodyCount ) <= g->capPriv &&
					wi.bodyStart + wi.bodyCount <= g->capMass );
		B2MI_CHECK( w->slotRegion.base <= wi.slotStart && wi.slotStart + total <= w->slotRegion.base + w->slotRegion.cap );
		// HBM records at (slotStart + k) * 7 quads, bodies-in-LDS records at slotStart * 7 + k * 5:
		// both end before quad (slotStart + total) * 7
		B2MI_CHECK( wi.slotStart + total <= g->capSlots && wi.slotStart + total <= g->capCons );
		if ( wi.groupStart[1] > wi.groupStart[0] )
		{
			B2MI_CHECK( wi.ovfLevelCount >= 1 && wi.ovfLevelBegin >= 0 &&
						wi.ovfLevelBegin + wi.ovfLevelCount + 1 <= (int)g->hOvfLevels.size() &&
						wi.ovfLevelBegin + wi.ovfLevelCount + 1 <= g->capOvfLevels );
			const int* lv = g->hOvfLevels.data() + wi.ovfLevelBegin;
			B2MI_CHECK( lv[0] == wi.groupStart[0] && lv[wi.ovfLevelCount] == wi.groupStart[1] );
			for ( int L = 0; L < wi.ovfLevelCount; ++L )
				B2MI_CHECK( lv[L] < lv[L + 1] );
		}
		if ( wi.jointGroupStart[1] > wi.jointGroupStart[0] )
		{
			B2MI_CHECK( wi.jointLevelCount >= 1 && wi.jointLevelBegin >= 0 &&
						wi.jointLevelBegin + wi.jointLevelCount + 1 <= (int)g->hOvfLevels.size() &&
						wi.jointLevelBegin + wi.jointLevelCount + 1 <= g->capOvfLevels );
			const int* lv = g->hOvfLevels.data() + wi.jointLevelBegin;
			B2MI_CHECK( lv[0] == wi.jointGroupStart[0] && lv[wi.jointLevelCount] == wi.jointGroupStart[1] );
			for ( int L = 0; L < wi.jointLevelCount; ++L )
				B2MI_CHECK( lv[L] < lv[L + 1] );
		}
		// the joint records the item's slots name lie inside its world's joint array
		for ( int k = wi.jointGroupStart[0]; k < total; ++k )
			B2MI_CHECK( 0 <= g->hSlotContact[wi.slotStart + k] && g->hSlotContact[wi.slotStart + k] < w->capJoints );
		B2MI_CHECK( wi.ldsBytes >= 0 && wi.ldsBytes <= kMaxItemLdsBytes );
		bodySpans.push_back( make_int2( wi.bodyStart, wi.bodyStart + wi.bodyCount ) );
		slotSpans.push_back( make_int2( wi.slotStart, wi.slotStart + total ) );
	}
	auto disjoint = []( std::vector<int2>& spans ) {
		std::sort( spans.begin(), spans.end(), []( int2 a, int2 b ) { return a.x < b.x; } );
		for ( size_t i = 1; i < spans.size(); ++i )
			B2MI_CHECK( spans[i - 1].y <= spans[i].x || spans[i - 1].y == spans[i - 1].x || spans[i].y == spans[i].x );
	};
	disjoint( bodySpans );
	disjoint( slotSpans );
	for ( const LaunchClass& c : g->classes )
	{
		B2MI_CHECK( c.first >= 0 && c.count > 0 && c.first + c.count <= (int)g->hItems.size() );
		for ( int i = c.first; i < c.first + c.count; ++i )
			B2MI_CHECK( g->hItems[i].store == c.store && g->hItems[i].ldsBytes <= c.lds && c.lds <= g->maxItemLds );
	}
}

template <typename T>
static void uploadSpan( T* dst, const std::vector<T>& src, int first, int last, hipStream_t s )
{
	if ( last > first )
		HIP_CHECK( hipMemcpyAsync( dst + first, src.data() + first, sizeof( T ) * (size_t)( last - first ), hipMemcpyHostToDevice, s ) );
}

// Writes the pending worlds' lists (those of `pending`, or of every world) into their regions
// (absolute indices) and uploads the span they cover; rebuilds the launch-ordered item array
// when any world's items changed.
static void publishLists( StepGroup* g, const std::vector<World*>* pending, bool itemsChanged )
{
	hipStream_t s = stream();
	int bFirst = INT32_MAX, bLast = 0, sFirst = INT32_MAX, sLast = 0, cFirst = INT32_MAX, cLast = 0;
	for ( World* w : pending ? *pending : g->worlds )
	{
		if ( w->solverPending )
		{
			// the entries not yet uploaded: all of them after a full rebuild or a new layout, the
			// appended items' after a partial rebuild
			const SolverLists& L = w->solver;
			Region br = w->bodyRegion, sr = w->slotRegion;
			const int b0 = w->pendingBodyFrom, s0 = w->pendingSlotFrom;
			B2MI_CHECK( (int)L.bodyList.size() <= br.cap && (int)L.slotContact.size() <= sr.cap );
			std::copy( L.bodyList.begin() + b0, L.bodyList.end(), g->hBodyList.begin() + br.base + b0 );
			std::copy( L.slotContact.begin() + s0, L.slotContact.end(), g->hSlotContact.begin() + sr.base + s0 );
			std::copy( L.slotBodies.begin() + s0, L.slotBodies.end(), g->hSlotBodies.begin() + sr.base + s0 );
			bFirst = std::min( bFirst, br.base + b0 );
			bLast = std::max( bLast, br.base + (int)L.bodyList.size() );
			sFirst = std::min( sFirst, sr.base + s0 );
			sLast = std::max( sLast, sr.base + (int)L.slotContact.size() );
			w->solverPending = false;
			w->pendingBodyFrom = (int)L.bodyList.size();
			w->pendingSlotFrom = (int)L.slotContact.size();
		}
		if ( w->collidePending )
		{
			Region cr = w->collideRegion;
			for ( int k = 0; k < cr.cap; ++k )
			{
				if ( k < (int)w->collideIds.size() )
				{
					const Contact& c = w->contacts[w->collideIds[k]];
					g->hCollide[cr.base + k] = make_int4( w->groupIndex, w->collideIds[k], c.shapeIdA, c.shapeIdB );
				}
				else
					g->hCollide[cr.base + k] = make_int4( -1, -1, -1, -1 );
			}
			cFirst = std::min( cFirst, cr.base );
			cLast = std::max( cLast, cr.base + cr.cap );
			w->collidePending = false;
		}
	}
	uploadSpan( g->dBodyList, g->hBodyList, bFirst, bLast, s );
	uploadSpan( g->dSlotContact, g->hSlotContact, sFirst, sLast, s );
	uploadSpan( g->dSlotBodies, g->hSlotBodies, sFirst, sLast, s );
	uploadSpan( g->dCollideList, g->hCollide, cFirst, cLast, s );

	if ( itemsChanged )
	{
		g->hItems.clear();
		g->hOvfLevels.clear();
		g->maxGroupSize = 0;
		g->maxItemLds = 0;
		for ( World* w : g->worlds )
		{
			const int levelBase = (int)g->hOvfLevels.size();
			g->hOvfLevels.insert( g->hOvfLevels.end(), w->solver.ovfLevels.begin(), w->solver.ovfLevels.end() );
			for ( size_t i = 0; i < w->solver.items.size(); ++i )
			{
				if ( !w->solver.live[i] )
					continue;
				WorkItem wi = w->solver.items[i];
				wi.bodyStart += w->bodyRegion.base;
				wi.slotStart += w->slotRegion.base;
				wi.ovfLevelBegin += levelBase;
				wi.jointLevelBegin += levelBase;
				g->hItems.push_back( wi );
			}
			g->maxGroupSize = std::max( g->maxGroupSize, w->solver.maxGroupSize );
			g->maxItemLds = std::max( g->maxItemLds, w->solver.maxItemLds );
		}
		// Grouped by workgroup size (one launch per size, largest size first) and longest first
		// within a size: the hardware dispatches workgroups in order, so the few large islands of an
		// uneven batch start at once and the small ones fill the CUs behind them. Items are
		// independent; their order does not change any result.
		// Items of one world chunk are consecutive (chunks in order).
		const int nw = (int)g->worlds.size(), chunks = chunkCountFor( nw, g->chunksRequested );
		g->chunkCount = chunks;
		auto chunkOf = [&]( const WorkItem& wi ) { return chunkOfWorld( wi.world, nw, chunks ); };
		std::stable_sort( g->hItems.begin(), g->hItems.end(), [&]( const WorkItem& a, const WorkItem& b ) {
			int ka = chunkOf( a ), kb = chunkOf( b );
			if ( ka != kb )
				return ka < kb;
			if ( a.block != b.block )
				return a.block > b.block;
			if ( a.store != b.store )
				return a.store < b.store;
			int ca = item_slots( a ) + a.bodyCount, cb = item_slots( b ) + b.bodyCount;
			return ca > cb;
		} );
		g->classes.clear();
		g->splitItems.clear();
		// very large HBM-store items run over the whole grid (enqueueSolve) rather
		// than in one workgroup; not in taped runs or chunked groups
		// by default a one-world group runs a large island over the grid from fewer slots than a batch,
		// whose other worlds' items share the CUs meanwhile
		const int splitSlots = g->splitSlots >= 0 ? g->splitSlots : g->worlds.size() == 1 ? kDefaultSplitSlotsWorld : kDefaultSplitSlots;
		const bool maySplit = splitSlots > 0 && g->tape == nullptr && chunks == 1;
		g->splitPerPhase = g->gridBlocks < 0 || solve_grid_capacity() == 0;
		if ( g->dGridSync == nullptr )
			HIP_CHECK( hipMalloc( (void**)&g->dGridSync, 2 * sizeof( unsigned ) ) );
		for ( int i = 0; i < (int)g->hItems.size(); ++i )
		{
			const WorkItem& wi = g->hItems[i];
			if ( maySplit && wi.store == kStoreHbm && item_slots( wi ) >= splitSlots )
			{
				g->splitItems.push_back( i );
				continue;
			}
			int chunk = chunkOf( wi );
			if ( g->classes.empty() || g->classes.back().block != wi.block || g->classes.back().chunk != chunk ||
				 g->classes.back().store != wi.store )
				g->classes.push_back( LaunchClass{ i, 0, wi.block, 0, chunk, wi.store } );
			LaunchClass& c = g->classes.back();
			c.count += 1;
			c.lds = std::max( c.lds, wi.ldsBytes );
		}
		g->itemCount = (int)g->hItems.size();
		g->itemsVersion += 1;
		grow( g->dItems, g->capItems, std::max( g->itemCount, 1 ) );
		if ( g->itemCount )
			HIP_CHECK( hipMemcpyAsync( g->dItems, g->hItems.data(), sizeof( WorkItem ) * g->itemCount, hipMemcpyHostToDevice, s ) );
		grow( g->dOvfLevels, g->capOvfLevels, std::max( (int)g->hOvfLevels.size(), 1 ) );
		validateItems( g );
		if ( !g->hOvfLevels.empty() )
			HIP_CHECK( hipMemcpyAsync( g->dOvfLevels, g->hOvfLevels.data(), sizeof( int ) * g->hOvfLevels.size(), hipMemcpyHostToDevice, s ) );
	}
	// the host vectors are read by the copies above: they stay unchanged until the next step,
	// which synchronizes the stream before it edits them
}

// src/solver.h:144-156
static void makeSoft( float hertz, float zeta, float h, float* out )
{
	if ( hertz == 0.0f )
	{
		out[0] = 0.0f;
		out[1] = 1.0f;
		out[2] = 0.0f;
		return;
	}
	float omega = 2.0f * b2_pi * hertz;
	float a1 = 2.0f * zeta + h * omega;
	float a2 = h * omega * a1;
	float a3 = 1.0f / ( 1.0f + a2 );
	out[0] = omega / a1;
	out[1] = a2 * a3;
	out[2] = a3;
}

static void fillDevWorld( DevWorld& d, World* w, float timeStep, int subStepCount )
{
	memset( &d, 0, sizeof( d ) );
	d.state = w->dState;
	d.sims = w->dSims;
	d.statics = w->dStatics;
	d.poses = w->dPoses;
	d.staticPoses = w->dStaticPoses;
	d.shapes = w->dShapes;
	d.bodyShapes = w->dBodyShapes;
	d.contacts = w->dContacts;
	d.contactOut = w->dContactOut;
	d.joints = w->dJoints;
	d.gravityX = w->gravity.x;
	d.gravityY = w->gravity.y;
	// src/world.c:717-746
	int sub = std::max( 1, subStepCount );
	d.subSteps = sub;
	d.dt = timeStep;
	d.inv_dt = timeStep > 0.0f ? 1.0f / timeStep : 0.0f;
	d.h = timeStep > 0.0f ? timeStep / sub : 0.0f;
	d.inv_h = timeStep > 0.0f ? sub * d.inv_dt : 0.0f;
	float contactHertz = b2MinFloat( w->contactHertz, 0.25f * d.inv_h );
	makeSoft( contactHertz, w->contactDampingRatio, d.h, d.contactSoft );
	makeSoft( 2.0f * contactHertz, w->contactDampingRatio, d.h, d.staticSoft );
	makeSoft( b2MinFloat( w->jointHertz, 0.125f * d.inv_h ), w->jointDampingRatio, d.h, d.jointSoft );
	d.restitutionThreshold = w->restitutionThreshold;
	d.pushout = w->contactPushoutVelocity;
	d.maxLinearSpeed = 4.0f * d.inv_dt;				  // b2_maxTranslation * inv_dt
	d.maxAngularSpeed = ( 0.25f * b2_pi ) * d.inv_dt; // b2_maxRotation * inv_dt
	d.warmStartScale = w->enableWarmStarting ? 1.0f : 0.0f;
	d.flags = ( w->enableWarmStarting ? kWorldWarmStart : 0u ) | ( w->enableContinuous ? kWorldContinuous : 0u ) |
			  ( w->enableSleep ? kWorldSleep : 0u );
}

// ensure the device contact array covers every allocated contact id
static void ensureContactCapacity( World* w )
{
	int needed = std::max( w->contactIdPool.nextIndex, 1 );
	if ( needed <= w->capContacts && w->dContacts )
		return;
	int cap = std::max( 64, w->capContacts );
	while ( cap < needed )
		cap *= 2;
	DevContact* p = nullptr;
	DevContactOut* q = nullptr;
	HIP_CHECK( hipMalloc( (void**)&p, sizeof( DevContact ) * (size_t)cap ) );
	HIP_CHECK( hipMalloc( (void**)&q, sizeof( DevContactOut ) * (size_t)cap ) );
	if ( w->dContacts && w->capContacts > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( p, w->dContacts, sizeof( DevContact ) * (size_t)w->capContacts, hipMemcpyDeviceToDevice, stream() ) );
		HIP_CHECK( hipMemcpyAsync( q, w->dContactOut, sizeof( DevContactOut ) * (size_t)w->capContacts, hipMemcpyDeviceToDevice,
								   stream() ) );
		HIP_CHECK( hipStreamSynchronize( stream() ) );
		HIP_CHECK( hipFree( w->dContacts ) );
		HIP_CHECK( hipFree( w->dContactOut ) );
	}
	w->dContacts = p;
	w->dContactOut = q;
	w->capContacts = cap;
	if ( w->group )
		w->group->structureDirty = true;
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

// Enlarge pass after finalize (src/solver.c:1741-1797 and 1877-1942 for fast shapes).
static void processEnlarged( World* w, const EnlargeRecord* recs, int count, bool bullets )
{
	if ( count == 0 )
		return;
	if ( bullets )
	{
		// src/solver.c:1889-1937: the bullets' shapes whose fat AABB grew (already moved)
		for ( int i = 0; i < count; ++i )
			if ( ( recs[i].flags & kFastEnlarged ) && ( recs[i].flags & kBulletFast ) )
			{
				Shape& s = w->shapes[recs[i].shapeId];
				s.fatAABB = { { recs[i].fatAABB[0], recs[i].fatAABB[1] }, { recs[i].fatAABB[2], recs[i].fatAABB[3] } };
				w->trees[b2_dynamicBody].enlargeProxy( s.proxyKey >> 2, s.fatAABB );
			}
		return;
	}
	// src/solver.c:1741-1797 in sim order, within a body the head-first shape list order (one sort
	// on (sim index, shape rank) in a buffer the thread keeps: a batch calls this for thousands of
	// worlds with a few records each): an enlarged shape's proxy grows and moves, a fast body's
	// shapes move (their AABBs come from the continuous stage)
	thread_local std::vector<std::pair<uint64_t, const EnlargeRecord*>> keyed;
	keyed.clear();
	const int awake = (int)w->awakeSimBody.size();
	for ( int i = 0; i < count; ++i )
	{
		const EnlargeRecord& r = recs[i];
		if ( !( r.flags & ( kEnlarged | kFastBody ) ) )
			continue;
		B2MI_CHECK( r.simIndex >= 0 && r.simIndex < awake );
		uint32_t rank = 0; // a fast body's one record first; an enlarged shape by its list position
		if ( r.flags & kEnlarged )
		{
			const Body& b = w->bodies[w->shapes[r.shapeId].bodyId];
			rank = 1;
			for ( int id = b.headShapeId; id != -1 && id != r.shapeId; id = w->shapes[id].nextShapeId )
				rank += 1;
		}
		keyed.push_back( { ( (uint64_t)(uint32_t)r.simIndex << 32 ) | rank, &r } );
	}
	std::sort( keyed.begin(), keyed.end(),
			   []( const std::pair<uint64_t, const EnlargeRecord*>& x, const std::pair<uint64_t, const EnlargeRecord*>& y ) {
				   return x.first < y.first;
			   } );
	for ( const auto& kr : keyed )
	{
		const EnlargeRecord* r = kr.second;
		if ( r->flags & kFastBody )
		{
			const Body& b = w->bodies[w->awakeSimBody[r->simIndex]];
			for ( int id = b.headShapeId; id != -1; id = w->shapes[id].nextShapeId )
				bufferMove( w, w->shapes[id].proxyKey );
			continue;
		}
		Shape& s = w->shapes[r->shapeId];
		s.fatAABB = { { r->fatAABB[0], r->fatAABB[1] }, { r->fatAABB[2], r->fatAABB[3] } };
		w->trees[s.proxyKey & 3].enlargeProxy( s.proxyKey >> 2, s.fatAABB );
		bufferMove( w, s.proxyKey );
	}
	// src/solver.c:1826-1873: the fast (non-bullet) shapes whose fat AABB grew in the continuous
	// stage, enlarged in the dynamic tree
	for ( int i = 0; i < count; ++i )
		if ( ( recs[i].flags & kFastEnlarged ) && !( recs[i].flags & kBulletFast ) )
		{
			Shape& s = w->shapes[recs[i].shapeId];
			s.fatAABB = { { recs[i].fatAABB[0], recs[i].fatAABB[1] }, { recs[i].fatAABB[2], recs[i].fatAABB[3] } };
			w->trees[b2_dynamicBody].enlargeProxy( s.proxyKey >> 2, s.fatAABB );
		}
}

// Island sleep (src/solver.c:1956-2001, src/solver_set.c:162-455) is not part of this
// pipeline, so a world with sleep enabled must stop at the first step after which the reference
// could put an island to sleep, rather than go on with results the reference would not give.
// The reference's islands persist and are split lazily, but each one is a union of whole
// components of the touching-contact graph over awake (dynamic and kinematic) bodies; it sleeps
// only when every body in it has reached the sleep time (src/solver.c:281-298). So if it can
// sleep, some component has all its bodies at the sleep time, which is what is checked here
// (called only after a step in which some body reached it).
static void checkIslandSleep( World* w )
{
	const int na = (int)w->awakeSimBody.size();
	if ( !w->enableSleep || na == 0 )
		return;
	std::vector<DevBodySim> sims( na );
	HIP_CHECK( hipMemcpyAsync( sims.data(), w->dSims, sizeof( DevBodySim ) * na, hipMemcpyDeviceToHost, stream() ) );
	HIP_CHECK( hipStreamSynchronize( stream() ) );
	std::vector<int> parent( na );
	for ( int i = 0; i < na; ++i )
		parent[i] = i;
	for ( int c = 0; c < kColorCount; ++c )
		for ( int id : w->colors[c] )
		{
			const Contact& ct = w->contacts[id];
			const Body& a = w->bodies[ct.edges[0].bodyId];
			const Body& b = w->bodies[ct.edges[1].bodyId];
			if ( a.setIndex != kAwakeSet || b.setIndex != kAwakeSet )
				continue;
			int ra = findRoot( parent, a.localIndex ), rb = findRoot( parent, b.localIndex );
			if ( ra != rb )
				parent[std::max( ra, rb )] = std::min( ra, rb );
		}
	// joints link islands too (src/island.c:298-380)
	for ( int c = 0; c < kColorCount; ++c )
		for ( int id : w->jointColors[c] )
		{
			const Joint& jt = w->joints[id];
			const Body& a = w->bodies[jt.edges[0].bodyId];
			const Body& b = w->bodies[jt.edges[1].bodyId];
			if ( a.setIndex != kAwakeSet || b.setIndex != kAwakeSet )
				continue;
			int ra = findRoot( parent, a.localIndex ), rb = findRoot( parent, b.localIndex );
			if ( ra != rb )
				parent[std::max( ra, rb )] = std::min( ra, rb );
		}
	std::vector<char> awake( na, 0 );
	for ( int i = 0; i < na; ++i )
		if ( sims[i].sleepTime < 0.5f ) // b2_timeToSleep, src/core.h:151
			awake[findRoot( parent, i )] = 1;
	for ( int i = 0; i < na; ++i )
		if ( findRoot( parent, i ) == i && !awake[i] )
			fatal( "an island has reached the sleep time: island sleep is not supported by the device pipeline "
				   "(create the world with b2WorldDef.enableSleep = false, or bodies with b2BodyDef.enableSleep = false)",
				   __FILE__, __LINE__ );
}

static float msSince( std::chrono::steady_clock::time_point t )
{
	return std::chrono::duration<float, std::milli>( std::chrono::steady_clock::now() - t ).count();
}

// Fork-join pool of host threads for per-world work; every task touches one world only. The
// calling thread takes part; with one task, or one hardware thread, everything runs inline.
// One caller at a time owns the pool: a call made while another thread's call holds it (worlds
// stepped from several threads, which the API allows for different worlds), or made from inside
// a task, runs its tasks inline on the calling thread, so concurrent callers never share the
// pool's task state.
class HostPool
{
public:
	static HostPool& get()
	{
		static HostPool pool;
		return pool;
	}

	void run( int n, const std::function<void( int )>& fn )
	{
		std::unique_lock<std::mutex> owner( owner_, std::defer_lock );
		if ( n <= 1 || threads_.empty() || insidePool() || !owner.try_lock() )
		{
			for ( int i = 0; i < n; ++i )
				fn( i );
			return;
		}
		struct Inside
		{
			Inside()
			{
				insidePool() = true;
			}
			~Inside()
			{
				insidePool() = false;
			}
		} inside;
		uint64_t g;
		{
			std::lock_guard<std::mutex> lk( m_ );
			task_ = &fn;
			count_ = n;
			finished_.store( 0, std::memory_order_relaxed );
			g = ++gen_;
			next_.store( g << 32, std::memory_order_release );
			published_.store( g, std::memory_order_release );
		}
		cv_.notify_all();
		work( g, fn, n );
		// the tasks the workers claimed: spin briefly, then sleep until the last one signals
		spinWhile( kCallerSpin, [&] { return finished_.load( std::memory_order_acquire ) < n; } );
		std::unique_lock<std::mutex> lk( m_ );
		done_.wait( lk, [&] { return finished_.load( std::memory_order_acquire ) >= n; } );
		task_ = nullptr;
	}

private:
	// How long a worker polls for the next job before it sleeps, and the caller for the claimed
	// tasks: a step's jobs come within a few milliseconds of each other, and waking a sleeping
	// thread costs tens of microseconds (the whole of a small job).
	static constexpr std::chrono::microseconds kWorkerSpin{ 2000 };
	static constexpr std::chrono::microseconds kCallerSpin{ 500 };

	template <class Pred>
	static void spinWhile( std::chrono::microseconds limit, Pred busy )
	{
		const auto until = std::chrono::steady_clock::now() + limit;
		while ( busy() )
			for ( int i = 0; i < 64; ++i )
				if ( !busy() )
					return;
				else if ( i == 63 && std::chrono::steady_clock::now() > until )
					return;
	}

	HostPool()
	{
		unsigned hw = std::thread::hardware_concurrency();
		int n = (int)std::min( hw > 0 ? hw : 1u, 16u ) - 1;
		for ( int i = 0; i < n; ++i )
			threads_.emplace_back( [this] { worker(); } );
	}
	~HostPool()
	{
		{
			std::lock_guard<std::mutex> lk( m_ );
			quit_ = true;
			published_.store( ~0ull, std::memory_order_release );
		}
		cv_.notify_all();
		for ( std::thread& t : threads_ )
			t.join();
	}
	// set on the thread that owns the pool for the duration of its call, and on the workers
	static bool& insidePool()
	{
		static thread_local bool inside = false;
		return inside;
	}
	// Runs tasks of job g while any are left: a run of consecutive tasks (about a quarter of one
	// thread's share, so the threads rarely meet on the counter) is claimed by advancing next_ =
	// (job << 32 | next index) with a compare-and-swap that fails once another job has begun, so a
	// late thread never takes a task, nor touches fn, of a job it was not given.
	void work( uint64_t g, const std::function<void( int )>& fn, int n )
	{
		const int chunk = std::max( 1, n / ( 4 * ( (int)threads_.size() + 1 ) ) );
		uint64_t cur = next_.load( std::memory_order_acquire );
		for ( ;; )
		{
			const int begin = (int)( cur & 0xffffffffu );
			if ( ( cur >> 32 ) != g || begin >= n )
				return;
			const int end = std::min( n, begin + chunk );
			if ( !next_.compare_exchange_weak( cur, ( g << 32 ) | (uint64_t)end, std::memory_order_acq_rel, std::memory_order_acquire ) )
				continue;
			for ( int i = begin; i < end; ++i )
				fn( i );
			if ( finished_.fetch_add( end - begin, std::memory_order_acq_rel ) + ( end - begin ) == n )
			{
				std::lock_guard<std::mutex> lk( m_ );
				done_.notify_one();
			}
			cur = next_.load( std::memory_order_acquire );
		}
	}
	void worker()
	{
		insidePool() = true;
		uint64_t seen = 0;
		for ( ;; )
		{
			spinWhile( kWorkerSpin, [&] { return published_.load( std::memory_order_acquire ) == seen; } );
			const std::function<void( int )>* fn;
			int n;
			uint64_t g;
			{
				std::unique_lock<std::mutex> lk( m_ );
				cv_.wait( lk, [&] { return quit_ || gen_ != seen; } );
				if ( quit_ )
					return;
				g = gen_;
				fn = task_;
				n = count_;
			}
			seen = g;
			if ( fn != nullptr )
				work( g, *fn, n );
		}
	}

	std::vector<std::thread> threads_;
	std::mutex owner_; // held by the caller whose tasks the pool runs
	std::mutex m_;
	std::condition_variable cv_, done_;
	const std::function<void( int )>* task_ = nullptr;
	int count_ = 0;
	uint64_t gen_ = 0;
	bool quit_ = false;
	std::atomic<uint64_t> next_{ 0 }, published_{ 0 };
	std::atomic<int> finished_{ 0 };
};

static void hostPoolRun( int n, const std::function<void( int )>& fn )
{
	HostPool::get().run( n, fn );
}

static void forEachWorld( StepGroup* g, const std::function<void( World* )>& fn )
{
	HostPool::get().run( (int)g->worlds.size(), [&]( int i ) { fn( g->worlds[i] ); } );
}

// serial state changes of one world in contact-id order (src/world.c:506-667)
// the last step's events go when the next step begins (src/world.c:685-692)
void clearStepEvents( World* w )
{
	w->moveEvents.clear();
	w->sensorBegin.clear();
	w->sensorEnd.clear();
	w->contactBegin.clear();
	w->contactEnd.clear();
	w->contactHit.clear();
	w->lazyEvents = false;
}

static void applyChanges( World* w, const ChangeRecord* recs, int count )
{
	const uint16_t world0 = (uint16_t)w->slot;
	for ( int i = 0; i < count; ++i )
	{
		const ChangeRecord& r = recs[i];
		Contact& c = w->contacts[r.contactId];
		int colorIndex = c.colorIndex, localIndex = c.localIndex;
		// the touch events of this change (src/world.c:544-551, 583-590, 632-639)
		const bool events = ( c.flags & kContactEvents ) != 0;
		const Shape& shapeA = w->shapes[c.shapeIdA];
		const Shape& shapeB = w->shapes[c.shapeIdB];
		const b2ShapeId shapeIdA{ shapeA.id + 1, world0, shapeA.revision }, shapeIdB{ shapeB.id + 1, world0, shapeB.revision };
		if ( r.flags & kSimDisjoint )
		{
			if ( events && ( c.flags & kContactTouching ) )
				w->contactEnd.push_back( b2ContactEndTouchEvent{ shapeIdA, shapeIdB } );
			c.flags &= ~kContactTouching;
			destroyContactImpl( w, &c );
		}
		else if ( r.flags & kSimStarted )
		{
			if ( events )
				w->contactBegin.push_back( b2ContactBeginTouchEvent{ shapeIdA, shapeIdB } );
			c.flags |= kContactTouching;
			addContactToGraph( w, c );
			removeFromList( w, w->nonTouching, localIndex );
			markContactBodies( w, c );
		}
		else if ( r.flags & kSimStopped )
		{
			if ( events )
				w->contactEnd.push_back( b2ContactEndTouchEvent{ shapeIdA, shapeIdB } );
			c.flags &= ~kContactTouching;
			int bodyIdA = c.edges[0].bodyId, bodyIdB = c.edges[1].bodyId;
			c.colorIndex = -1;
			c.localIndex = (int)w->nonTouching.size();
			w->nonTouching.push_back( c.contactId );
			removeContactFromGraph( w, bodyIdA, bodyIdB, colorIndex, localIndex );
			markContactBodies( w, c );
		}
	}
}

// records sorted by world: [first, last) of each world
template <typename R>
static void sortByWorld( std::vector<R>& recs, int worldCount )
{
	// stable counting sort on the world index (records of one world keep their order)
	std::vector<int> start( worldCount + 1, 0 );
	for ( const R& r : recs )
		start[r.world + 1] += 1;
	for ( int i = 0; i < worldCount; ++i )
		start[i + 1] += start[i];
	std::vector<R> sorted( recs.size() );
	for ( const R& r : recs )
		sorted[start[r.world]++] = r;
	recs.swap( sorted );
}

template <class R>
static void splitByWorld( const std::vector<R>& recs, int worldCount, std::vector<int>& first, std::vector<int>& last )
{
	first.assign( worldCount, 0 );
	last.assign( worldCount, 0 );
	int n = (int)recs.size();
	for ( int b = 0; b < n; )
	{
		int e = b;
		while ( e < n && recs[e].world == recs[b].world )
			e += 1;
		first[recs[b].world] = b;
		last[recs[b].world] = e;
		b = e;
	}
}

// Worlds picked out by a parallel pass over a group (one slot per world, claimed through an
// atomic counter), then put back in group order.
struct WorldPick
{
	std::vector<World*> list;
	std::atomic<int> count{ 0 };
	void reset( size_t n )
	{
		list.resize( n );
		count.store( 0 );
	}
	void add( World* w )
	{
		list[count.fetch_add( 1 )] = w;
	}
	void finish()
	{
		list.resize( count.load() );
		std::sort( list.begin(), list.end(), []( const World* a, const World* b ) { return a->groupIndex < b->groupIndex; } );
	}
};

// Solver launches of one step: one launch for every item at a fixed workgroup size, or one per
// item size class (the items' own sizes, blockSize 0), those on separate streams so that they
// run side by side. With `abortFlag` set every workgroup returns at once when the flag is
// nonzero (a speculative launch behind the narrowphase, void when contacts changed state).
#define B2MI_TRY( expr )                 \
	do                                   \
	{                                    \
		hipError_t e_ = ( expr );        \
		if ( e_ != hipSuccess )          \
			return e_;                   \
	} while ( 0 )

// Enqueues the solver launches of one step on stream s (and the extra streams `aux`, joined back
// into s); returns the first error instead of failing, so that a stream capture of it can be
// undone.
hipError_t StepGroup::enqueueSolve( hipStream_t s, const hipStream_t* aux, const int* abortFlag )
{
	for ( const LaunchClass& c : classes )
	{
		int block = blockSize > 0 ? blockSize : c.block;
		B2MI_CHECK( block >= 64 && block <= kSolveMaxBlock && block % 64 == 0 );
	}
	auto launch = [&]( const LaunchClass& c, hipStream_t cs ) {
		int block = blockSize > 0 ? blockSize : c.block;
		// (HBM-store items keep nothing in dynamic LDS; their kernel's static LDS holds the
		// overflow run cache)
		int lds = c.store == kStoreHbm ? 0 : blockSize > 0 ? maxItemLds : c.lds;
		return launch_solve( c.store, dWorlds, dItems, c.first, c.count, block, lds, dBodyList, dSlotContact, dSlotBodies, dOvfLevels,
							 dCons, dPriv, dMass, dEnlarge, dCounters + 1, capEnlarge, levelRunSlots, abortFlag, cs );
	};
	B2MI_CHECK( splitItems.empty() || abortFlag == nullptr || !splitPerPhase );
	for ( int item : splitItems )
	{
		if ( splitPerPhase )
			B2MI_TRY( enqueueSplit( item, s ) );
		else
			B2MI_TRY( launch_solve_grid( dWorlds, dItems, item, gridBlocksFor( hItems[item] ), dBodyList, dSlotContact, dSlotBodies,
										 dOvfLevels, dCons, dPriv, dMass, dEnlarge, dCounters + 1, capEnlarge, dGridSync, dCounters + 3,
										 levelRunSlots, abortFlag, s ) );
	}
	if ( classes.size() <= 1 )
	{
		if ( !classes.empty() )
			B2MI_TRY( launch( classes[0], s ) );
		return hipSuccess;
	}
	B2MI_TRY( hipEventRecord( forkEvent, s ) );
	int n = (int)classes.size();
	for ( int c = 0; c < n; ++c )
	{
		hipStream_t cs = c == 0 ? s : aux[( c - 1 ) % kAuxStreams];
		if ( c > 0 && c <= kAuxStreams )
			B2MI_TRY( hipStreamWaitEvent( cs, forkEvent, 0 ) );
		B2MI_TRY( launch( classes[c], cs ) );
	}
	for ( int c = 1; c < std::min( n, kAuxStreams + 1 ); ++c )
	{
		B2MI_TRY( hipEventRecord( joinEvents[c - 1], aux[c - 1] ) );
		B2MI_TRY( hipStreamWaitEvent( s, joinEvents[c - 1], 0 ) );
	}
	return hipSuccess;
}

// Everything a step's solver launches take as arguments or launch shape: two steps with the same
// signature enqueue the same launches.
SolveSignature StepGroup::solveSignature( const int* abortFlag ) const
{
	SolveSignature k;
	memset( &k, 0, sizeof( k ) );
	k.itemsVersion = itemsVersion;
	k.abortFlag = abortFlag;
	k.worlds = dWorlds;
	k.items = dItems;
	k.bodyList = dBodyList;
	k.slotContact = dSlotContact;
	k.slotBodies = dSlotBodies;
	k.ovfLevels = dOvfLevels;
	k.cons = dCons;
	k.priv = dPriv;
	k.mass = dMass;
	k.enlarge = dEnlarge;
	k.counters = dCounters;
	k.gridSync = dGridSync;
	k.enlargeCapacity = capEnlarge;
	k.blockSize = blockSize;
	k.maxItemLds = maxItemLds;
	k.gridBlocks = gridBlocks;
	k.levelRun = levelRunSlots;
	k.splitPerPhase = splitPerPhase ? 1 : 0;
	k.chunks = chunkCount;
	return k;
}

// The solver launches of one step: one launch for every item at a fixed workgroup size, or one per
// item size class (the items' own sizes, blockSize 0), those on separate streams so that they
// run side by side; very large items over the whole grid first. With `abortFlag` set every
// workgroup returns at once when the flag is nonzero (a speculative launch behind the
// narrowphase, void when contacts changed state). When a step enqueues the same launches as the
// step before it (the usual case once a scene's contacts have settled), the sequence is captured
// once as a HIP graph and replayed from then on, until anything it bakes in changes; a capture
// that the runtime refuses turns graphs off for the group. The capture is recorded on streams of
// the group's own (recording executes nothing), so that other threads stepping other groups on the
// shared streams meanwhile never enqueue into it; the replay runs on the shared stream s.
void StepGroup::launchSolve( hipStream_t s, const int* abortFlag )
{
	if ( forkEvent == nullptr )
	{
		HIP_CHECK( hipEventCreateWithFlags( &forkEvent, hipEventDisableTiming ) );
		for ( hipEvent_t& e : joinEvents )
			HIP_CHECK( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
		for ( hipEvent_t& e : collideDone )
			HIP_CHECK( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
	}
	// a graph pays for itself over a sequence of launches: per-phase launches of a split item or
	// launches on several streams; one plain launch goes straight to the stream
	const bool sequence = ( !splitItems.empty() && splitPerPhase ) || classes.size() > 1;
	if ( launchGraphs && !graphsRefused && sequence )
	{
		const SolveSignature key = solveSignature( abortFlag );
		if ( graphExec != nullptr && memcmp( &key, &graphKey, sizeof( key ) ) == 0 )
		{
			HIP_CHECK( hipGraphLaunch( graphExec, s ) );
			graphReplays += 1;
			return;
		}
		const bool repeat = memcmp( &key, &lastSolveKey, sizeof( key ) ) == 0;
		lastSolveKey = key;
		if ( repeat && captureSolve( s, abortFlag ) )
		{
			graphKey = key;
			HIP_CHECK( hipGraphLaunch( graphExec, s ) );
			graphReplays += 1;
			return;
		}
	}
	hipStream_t aux[kAuxStreams];
	for ( int i = 0; i < kAuxStreams; ++i )
		aux[i] = auxStream( i );
	HIP_CHECK( enqueueSolve( s, aux, abortFlag ) );
}

bool StepGroup::captureSolve( hipStream_t, const int* abortFlag )
{
	if ( graphExec != nullptr )
	{
		HIP_CHECK( hipGraphExecDestroy( graphExec ) );
		graphExec = nullptr;
	}
	if ( captureStreams[0] == nullptr )
		for ( hipStream_t& cs : captureStreams )
			HIP_CHECK( hipStreamCreateWithFlags( &cs, hipStreamNonBlocking ) );
	hipGraph_t graph = nullptr;
	hipStream_t cs = captureStreams[0];
	hipError_t e = hipStreamBeginCapture( cs, hipStreamCaptureModeThreadLocal );
	if ( e == hipSuccess )
	{
		hipError_t le = enqueueSolve( cs, captureStreams + 1, abortFlag );
		e = hipStreamEndCapture( cs, &graph );
		if ( le != hipSuccess )
			e = le;
	}
	if ( e == hipSuccess )
		e = hipGraphInstantiate( &graphExec, graph, nullptr, nullptr, 0 );
	if ( graph != nullptr )
		(void)hipGraphDestroy( graph );
	if ( e != hipSuccess )
	{
		(void)hipGetLastError();
		graphExec = nullptr;
		graphsRefused = true;
		return false;
	}
	graphCaptures += 1;
	return true;
}

// Workgroups of a very large item's cooperative launch: one thread per slot of its largest colour
// group, contacts and joints (or the set count), within what can be resident at once.
int StepGroup::gridBlocksFor( const WorkItem& wi ) const
{
	int largest = 1;
	for ( int g = 1; g < kGroups; ++g )
		largest = std::max( largest, wi.groupStart[g + 1] - wi.groupStart[g] + wi.jointGroupStart[g + 1] - wi.jointGroupStart[g] );
	int blocks = gridBlocks > 0 ? gridBlocks : ( largest + kGridSolveBlock - 1 ) / kGridSolveBlock;
	return std::max( 1, std::min( blocks, solve_grid_capacity() ) );
}

// The phases of one very large work item (solve_item's order) on stream s: each phase one grid
// launch, a colour's joints and contacts in one launch, and each run of consecutive overflow
// levels of at most levelRunSlots slots in all one single-workgroup launch with the run's bodies
// and records in LDS (k_split_levels) rather than a launch per level.
hipError_t StepGroup::enqueueSplit( int item, hipStream_t s )
{
	hipError_t status = hipSuccess;
	const WorkItem& wi = hItems[item];
	const DevWorld& hw = hWorlds[wi.world];
	auto run = [&]( int phase, int first, int count, int jointPhase = -1, int jointFirst = 0, int jointCount = 0 ) {
		if ( status == hipSuccess )
			status = launch_split_phase( dWorlds, dItems, item, phase, first, count, jointPhase, jointFirst, jointCount, dBodyList,
										 dSlotContact, dSlotBodies, dCons, dPriv, dMass, dEnlarge, dCounters + 1, capEnlarge, s );
	};
	const int total = wi.groupStart[kGroups];
	const int levelCount = wi.groupStart[1] > wi.groupStart[0] ? wi.ovfLevelCount : 0;
	B2MI_CHECK( levelCount == 0 || wi.ovfLevelBegin + levelCount < (int)hOvfLevels.size() );
	// the overflow joints in levels before the overflow contacts, the joints of a colour beside
	// the colour's contacts (they share no body with them)
	const int jointLevelCount = wi.jointGroupStart[1] > wi.jointGroupStart[0] ? wi.jointLevelCount : 0;
	B2MI_CHECK( jointLevelCount == 0 || wi.jointLevelBegin + jointLevelCount < (int)hOvfLevels.size() );
	// levels [begin, begin + count) of the level list: consecutive levels of at most
	// levelRunSlots slots in all as one cached run in one workgroup, a larger level over the grid
	const int runSlots = std::min( levelRunSlots, kLevelCacheSlots );
	auto levelsOf = [&]( int phase, int begin, int count ) {
		const int* lv = hOvfLevels.data() + begin;
		int L = 0;
		while ( L < count && status == hipSuccess )
		{
			if ( runSlots <= 0 || lv[L + 1] - lv[L] > runSlots )
			{
				run( phase, lv[L], lv[L + 1] - lv[L] );
				++L;
				continue;
			}
			int E = L + 1;
			while ( E < count && lv[E + 1] - lv[L] <= runSlots )
				++E;
			status = launch_split_levels( dWorlds, dItems, item, phase, dOvfLevels, begin + L, E - L, dBodyList, dSlotContact, dSlotBodies,
										  dCons, dPriv, dMass, s );
			L = E;
		}
	};
	auto overflow = [&]( int jointPhase, int phase ) {
		if ( jointPhase >= 0 )
			levelsOf( jointPhase, wi.jointLevelBegin, jointLevelCount );
		levelsOf( phase, wi.ovfLevelBegin, levelCount );
	};
	auto colours = [&]( int jointPhase, int phase ) {
		for ( int g = 1; g < kGroups; ++g )
		{
			const int joints = jointPhase >= 0 ? wi.jointGroupStart[g + 1] - wi.jointGroupStart[g] : 0;
			run( phase, wi.groupStart[g], wi.groupStart[g + 1] - wi.groupStart[g], jointPhase, wi.jointGroupStart[g], joints );
		}
	};
	run( kSplitStage, 0, wi.bodyCount );
	run( -1, 0, 0, kSplitJointPrepare, wi.jointGroupStart[0], wi.jointGroupStart[kGroups] - wi.jointGroupStart[0] );
	run( kSplitPrepare, 0, total );
	for ( int sub = 0; sub < hw.subSteps; ++sub )
	{
		run( kSplitIntegrateVelocities, 0, wi.bodyCount );
		overflow( kSplitJointWarmStart, kSplitOverflowWarmStart );
		if ( hw.flags & kWorldWarmStart )
			colours( kSplitJointWarmStart, kSplitWarmStart );
		overflow( kSplitJointSolve, kSplitOverflowSolve );
		colours( kSplitJointSolve, kSplitSolve );
		run( kSplitIntegratePositions, 0, wi.bodyCount );
		overflow( kSplitJointRelax, kSplitOverflowRelax );
		colours( kSplitJointRelax, kSplitRelax );
	}
	overflow( -1, kSplitOverflowRestitution );
	colours( -1, kSplitRestitution );
	run( kSplitStoreImpulses, 0, total );
	run( kSplitFinalize, 0, wi.ownedCount );
	return status;
}

// src/solver.c:1806-1937 on the device: every fast body (bullets = false) or bullet (true) of
// hEnlarge's kFastBody records sweeps against the device copies of its world's trees (the static
// tree; for a bullet also the kinematic and dynamic ones, as the enlarge pass left them). Each
// body's joints that do not collide their bodies name the bodies its query skips
// (b2ShouldBodiesCollide, src/body.c:1724-1760).
void StepGroup::runContinuous( bool bullets )
{
	hipStream_t s = stream();
	std::vector<ContinuousBody> list;
	std::vector<int> noCollide;
	std::vector<char> synced( worlds.size(), 0 );
	for ( const EnlargeRecord& r : hEnlarge )
	{
		if ( !( r.flags & kFastBody ) || ( ( r.flags & kBulletFast ) != 0 ) != bullets )
			continue;
		World* w = worlds[r.world];
		DeviceBroadphase& D = w->devicePairs;
		if ( !synced[r.world] )
		{
			syncDeviceTree( w, b2_staticBody, s );
			if ( bullets )
			{
				syncDeviceTree( w, b2_kinematicBody, s );
				syncDeviceTree( w, b2_dynamicBody, s );
			}
			syncShapeFilters( w, s );
			synced[r.world] = 1;
		}
		ContinuousBody cb{};
		cb.world = r.world;
		cb.simIndex = r.simIndex;
		cb.noCollideBegin = (int)noCollide.size();
		const Body& body = w->bodies[w->awakeSimBody[r.simIndex]];
		for ( int key = body.headJointKey; key != -1; )
		{
			const Joint& j = w->joints[key >> 1];
			const int other = j.edges[( key & 1 ) ^ 1].bodyId;
			if ( j.collideConnected == false )
			{
				const Body& ob = w->bodies[other];
				noCollide.push_back( ob.setIndex == 0 ? ( ob.localIndex | kStaticRef ) : ob.localIndex );
			}
			key = j.edges[key & 1].nextKey;
		}
		cb.noCollideEnd = (int)noCollide.size();
		for ( int t = 0; t < 3; ++t )
		{
			const bool used = t == b2_staticBody || bullets;
			cb.nodes[t] = used ? D.nodes[t] : nullptr;
			cb.roots[t] = used ? w->trees[t].root : -1;
		}
		cb.filters = D.shapes;
		list.push_back( cb );
	}
	if ( list.empty() )
		return;
	grow( dContinuous, capContinuous, (int)list.size() );
	grow( dNoCollide, capNoCollide, std::max( (int)noCollide.size(), 1 ) );
	HIP_CHECK( hipMemcpyAsync( dContinuous, list.data(), sizeof( ContinuousBody ) * list.size(), hipMemcpyHostToDevice, s ) );
	if ( !noCollide.empty() )
		HIP_CHECK( hipMemcpyAsync( dNoCollide, noCollide.data(), sizeof( int ) * noCollide.size(), hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipMemsetAsync( dCounters + kContinuousCounter, 0, 2 * sizeof( int ), s ) );
	HIP_CHECK( launch_continuous( dWorlds, dContinuous, (int)list.size(), dNoCollide, dEnlarge, dCounters + 1, capEnlarge,
								  dCounters + kContinuousCounter, dCounters + kContinuousCounter + 1, s ) );
	HIP_CHECK( hipMemcpyAsync( hCounters + 1, dCounters + 1, sizeof( int ), hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipMemcpyAsync( hCounters + kContinuousCounter, dCounters + kContinuousCounter, 2 * sizeof( int ), hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	if ( hCounters[kContinuousCounter + 1] != 0 )
		fatal( "a continuous-collision tree query exceeded its stack", __FILE__, __LINE__ );
	const int hits = hCounters[kContinuousCounter];
	toiHits += hits;
	if ( hits > 0 && tape != nullptr )
		fatal( "a taped run does not support a time of impact (a fast body that continuous collision moved back)", __FILE__, __LINE__ );
	const int total = hCounters[1];
	const int have = (int)hEnlarge.size();
	B2MI_CHECK( have <= total && total <= capEnlarge );
	if ( total > have )
	{
		hEnlarge.resize( total );
		HIP_CHECK( hipMemcpyAsync( hEnlarge.data() + have, dEnlarge + have, sizeof( EnlargeRecord ) * ( total - have ), hipMemcpyDeviceToHost,
								   s ) );
		HIP_CHECK( hipStreamSynchronize( s ) );
	}
}

// The solver launches of one world chunk, its size classes side by side: class j on extra
// stream j (mod the stream count), each stream first waiting for `ready` (recorded on the
// group's stream) and then joined back into it through joinEvents.
void StepGroup::launchSolveChunk( int chunk, hipEvent_t ready, const int* abortFlag )
{
	int j = 0;
	bool used[kAuxStreams] = {};
	for ( const LaunchClass& c : classes )
	{
		if ( c.chunk != chunk )
			continue;
		int block = blockSize > 0 ? blockSize : c.block;
		// (HBM-store items keep nothing in dynamic LDS; their kernel's static LDS holds the
		// overflow run cache)
		int lds = c.store == kStoreHbm ? 0 : blockSize > 0 ? maxItemLds : c.lds;
		B2MI_CHECK( block >= 64 && block <= kSolveMaxBlock && block % 64 == 0 );
		int k = j++ % kAuxStreams;
		hipStream_t as = auxStream( k );
		if ( !used[k] )
			HIP_CHECK( hipStreamWaitEvent( as, ready, 0 ) );
		used[k] = true;
		HIP_CHECK( launch_solve( c.store, dWorlds, dItems, c.first, c.count, block, lds, dBodyList, dSlotContact, dSlotBodies,
								 dOvfLevels, dCons, dPriv, dMass, dEnlarge, dCounters + 1, capEnlarge, levelRunSlots, abortFlag, as ) );
	}
	for ( int k = 0; k < kAuxStreams; ++k )
		if ( used[k] )
			HIP_CHECK( hipEventRecord( joinEvents[k], auxStream( k ) ) );
}

// makes the group's stream wait for every extra stream's last recorded join event
void StepGroup::joinAux( hipStream_t s )
{
	for ( int k = 0; k < kAuxStreams; ++k )
		HIP_CHECK( hipStreamWaitEvent( s, joinEvents[k], 0 ) );
}

void StepGroup::step( float timeStep, int subStepCount )
{
	auto t0 = std::chrono::steady_clock::now();
	// host phases of this step (ms), reported through b2World_GetProfile:
	//   pairs          pair update and contact creation
	//   broadphase     dynamic-tree rebuild (overlaps the narrowphase kernel)
	//   collide        narrowphase kernel (device events)
	//   buildIslands   begin/end-touch processing and work-item rebuild
	//   solve          solver kernel (device events)
	//   finalizeBodies enlarged-proxy processing
	float msPairs = 0.0f, msBroad = 0.0f, msChanges = 0.0f, msEnlarge = 0.0f;
	hipStream_t s = stream();
	B2MI_CHECK( stepping == false );
	for ( World* w : worlds )
		clearStepEvents( w );
	stepping = true;
	profile = b2Profile{};
	const int nw = (int)worlds.size();
	if ( structureDirty || (int)hWorlds.size() != nw )
	{
		hWorlds.resize( nw );
		structureDirty = true;
	}
	const int target = targetSlots > 0 ? targetSlots : 256;
	const int budget = ldsBudget < 0 ? kDefaultItemLdsBytes : std::min( ldsBudget, kMaxItemLdsBytes );
	const int fullBudget = ldsFullBudget < 0 ? budget : std::min( ldsFullBudget, budget );
	const bool reshape = target != builtTarget || budget != builtBudget || fullBudget != builtFullBudget;
	builtTarget = target;
	builtBudget = budget;
	builtFullBudget = fullBudget;
	// the worlds this step visits on the host: every world, or those listed since the last step
	std::vector<World*> visit;
	{
		std::lock_guard<std::mutex> lock( activeMutex );
		visit.swap( active );
	}
	for ( World* w : visit )
		w->activeQueued.store( false, std::memory_order_relaxed );
	const bool visitAll = allActive || structureDirty || reshape || timeStep != lastTimeStep || subStepCount != lastSubSteps;
	allActive = false;
	lastTimeStep = timeStep;
	lastSubSteps = subStepCount;
	const std::vector<World*>& pass = visitAll ? worlds : visit;

	// 1. per world in parallel: pair update and new contacts (src/broad_phase.c:333-436), the
	//    narrowphase list when contacts were created or destroyed, and the world's device
	//    parameters. Worlds that need uploads or device allocations finish serially after it.
	hostPhases[0] = msSince( t0 );
	auto tp = std::chrono::steady_clock::now();
	WorldPick serial, collideDirty, topology;
	serial.reset( nw );
	collideDirty.reset( nw );
	topology.reset( nw );
	std::atomic<bool> paramsChanged{ false }, regionFull{ false };
	std::atomic<int> shapeDelta{ 0 };
	auto refreshParams = [&]( World* w ) {
		DevWorld d;
		fillDevWorld( d, w, timeStep, subStepCount );
		DevWorld& h = hWorlds[w->groupIndex];
		if ( memcmp( &d, &h, sizeof( DevWorld ) ) != 0 )
		{
			h = d;
			paramsChanged.store( true );
		}
	};
	// one world alone takes the pool for its own pair queries
	const bool nestedPool = pass.size() == 1;
	HostPool::get().run( (int)pass.size(), [&]( int i ) {
		World* w = pass[i];
		updatePairs( w, nestedPool );
		if ( w->contactsDirty )
		{
			buildWorldCollide( w );
			collideDirty.add( w );
		}
		if ( (int)w->collideIds.size() > w->collideRegion.cap )
			regionFull.store( true );
		if ( reshape )
			w->topologyDirty = true;
		if ( w->topologyDirty || !w->dirtySims.empty() )
			topology.add( w );
		int ns = (int)w->shapes.size();
		if ( ns != w->countedShapes )
		{
			shapeDelta.fetch_add( ns - w->countedShapes );
			w->countedShapes = ns;
		}
		if ( w->bodiesDirty || w->jointsDirty || !w->pendingInits.empty() || w->contactIdPool.nextIndex > w->capContacts ||
			 w->dContacts == nullptr )
			serial.add( w );
		else
			refreshParams( w );
	} );
	serial.finish();
	collideDirty.finish();
	topology.finish();
	hInits.clear();
	for ( World* w : serial.list )
	{
		if ( w->bodiesDirty )
			b2mi_internal_uploadBodies( w );
		if ( w->jointsDirty )
			uploadJoints( w );
		ensureContactCapacity( w );
		hInits.insert( hInits.end(), w->pendingInits.begin(), w->pendingInits.end() );
		w->pendingInits.clear();
		refreshParams( w );
	}
	msPairs = msSince( tp );
	if ( structureDirty )
	{
		layoutValid = false;
		paramsChanged.store( true );
	}
	if ( paramsChanged.load() )
	{
		grow( dWorlds, capWorlds, std::max( nw, 1 ) );
		HIP_CHECK( hipMemcpyAsync( dWorlds, hWorlds.data(), sizeof( DevWorld ) * nw, hipMemcpyHostToDevice, s ) );
		structureDirty = false;
	}
	if ( dCounters == nullptr )
	{
		HIP_CHECK( hipMalloc( (void**)&dCounters, sizeof( int ) * kCounterCount ) );
		HIP_CHECK( hipHostMalloc( (void**)&hCounters, sizeof( int ) * kCounterCount, hipHostMallocDefault ) );
	}
	if ( !hInits.empty() )
	{
		grow( dInits, capInits, (int)hInits.size() );
		HIP_CHECK( hipMemcpyAsync( dInits, hInits.data(), sizeof( ContactInit ) * hInits.size(), hipMemcpyHostToDevice, s ) );
		HIP_CHECK( launch_init_contacts( dWorlds, dInits, (int)hInits.size(), s ) );
	}

	// 2. narrowphase over the worlds' contact lists
	hostPhases[1] = msSince( t0 );
	bool relayout = !layoutValid || regionFull.load();
	if ( relayout )
		layoutGroup( this );
	publishLists( this, relayout ? nullptr : &collideDirty.list, relayout );
	HIP_CHECK( hipMemsetAsync( dCounters, 0, sizeof( int ) * kCounterCount, s ) );
	shapeTotal += shapeDelta.load();
	// a record per enlarged shape, per fast body and per fast shape that grew
	grow( dEnlarge, capEnlarge, std::max( 2 * shapeTotal, 1 ) );
	// events bracket each kernel launch alone (the stream is idle while the host works)
	if ( events[0] == nullptr )
		for ( hipEvent_t& e : events )
			HIP_CHECK( hipEventCreate( &e ) );
	hipEvent_t e0 = events[0], e1 = events[1], e2 = events[2], e3 = events[3];
	hostPhases[2] = msSince( t0 );
	// When no work item needs rebuilding, the solver goes right behind the narrowphase: a step
	// in which no contact begins or ends touching (the usual one in a batch of settled worlds)
	// then needs no host round trip. Otherwise the launch returns at once and the solver runs
	// after the state changes below. Groups whose contacts change at most steps (one large
	// world) stop speculating after a miss and try again every 16th step.
	// A taped step records its snapshot ahead of the speculative launch (one world chunk only)
	// and takes the record back when the launch is voided.
	const bool speculate = topology.list.empty() && ( tape == nullptr || chunkCount == 1 ) && itemCount > 0 &&
						   ( splitItems.empty() || !splitPerPhase ) &&
						   ( specMisses == 0 || epoch % 16 == 0 );
	// World chunks: consecutive worlds, so each chunk's narrowphase list is one range and its
	// state changes go to that range of the change buffer with a counter of their own. A
	// chunk's speculative solver launch follows its narrowphase on the chunk's stream and is
	// voided by that chunk's changes only.
	const int chunks = chunkCount;
	B2MI_CHECK( chunks >= 1 && chunks <= kMaxChunks );
	for ( int c = 0; c <= chunks; ++c )
	{
		int wi = (int)( (int64_t)c * nw / chunks );
		chunkCollide[c] = wi < nw ? worlds[wi]->collideRegion.base : collideCount;
	}
	for ( int c = 0; c < chunks; ++c )
		B2MI_CHECK( 0 <= chunkCollide[c] && chunkCollide[c] <= chunkCollide[c + 1] && chunkCollide[c + 1] <= collideCount );
	if ( forkEvent == nullptr )
	{
		HIP_CHECK( hipEventCreateWithFlags( &forkEvent, hipEventDisableTiming ) );
		for ( hipEvent_t& e : joinEvents )
			HIP_CHECK( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
		for ( hipEvent_t& e : collideDone )
			HIP_CHECK( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
	}
	HIP_CHECK( hipEventRecord( e0, s ) );
	for ( int c = 0; c < chunks; ++c )
	{
		int first = chunkCollide[c], count = chunkCollide[c + 1] - chunkCollide[c];
		HIP_CHECK( launch_collide( dWorlds, dCollideList + first, count, dChanges + first, dCounters + kChunkCounter + c, s ) );
		if ( c == 0 )
		{
			// device times: e0-e1 the first chunk's narrowphase; e3-e2 the rest of the step (with
			// several chunks it holds their narrowphases too, beside the solver)
			HIP_CHECK( hipEventRecord( e1, s ) );
			if ( speculate )
				HIP_CHECK( hipEventRecord( e3, s ) );
		}
		if ( speculate && chunks > 1 )
		{
			HIP_CHECK( hipEventRecord( collideDone[c], s ) );
			launchSolveChunk( c, collideDone[c], dCounters + kChunkCounter + c );
		}
	}
	if ( speculate && chunks == 1 )
	{
		if ( tape )
			tapeRecordStep( this, timeStep, subStepCount, blockSize > 0 ? std::min( blockSize, kAdjointMaxBlock ) : kAdjointMaxBlock );
		launchSolve( s, dCounters + kChunkCounter );
	}
	if ( speculate && chunks > 1 )
		joinAux( s );
	if ( speculate )
		HIP_CHECK( hipEventRecord( e2, s ) );
	HIP_CHECK( hipMemcpyAsync( hCounters, dCounters, sizeof( int ) * kCounterCount, hipMemcpyDeviceToHost, s ) );

	// the tree rebuild overlaps the narrowphase (src/world.c:367-379, 420)
	auto tb = std::chrono::steady_clock::now();
	forEachWorld( this, []( World* w ) {
		w->trees[b2_dynamicBody].rebuild( false );
		w->trees[b2_kinematicBody].rebuild( false );
	} );
	msBroad = msSince( tb );
	hostPhases[3] = msSince( t0 );
	HIP_CHECK( hipStreamSynchronize( s ) );
	auto tc = std::chrono::steady_clock::now();
	hostPhases[4] = msSince( t0 );
	int changeCount = 0;
	bool chunkChanged[kMaxChunks] = {};
	for ( int c = 0; c < chunks; ++c )
	{
		int n = hCounters[kChunkCounter + c];
		B2MI_CHECK( 0 <= n && n <= chunkCollide[c + 1] - chunkCollide[c] );
		chunkChanged[c] = n > 0;
		changeCount += n;
	}
	if ( speculate )
		specMisses = changeCount > 0 ? specMisses + 1 : 0;
	if ( speculate && tape && changeCount > 0 )
		tapeUndoStep( this );
	std::vector<World*> changedTopology;
	if ( changeCount > 0 )
	{
		hChanges.resize( changeCount );
		int at = 0;
		for ( int c = 0; c < chunks; ++c )
		{
			int n = hCounters[kChunkCounter + c];
			if ( n > 0 )
				HIP_CHECK( hipMemcpyAsync( hChanges.data() + at, dChanges + chunkCollide[c], sizeof( ChangeRecord ) * n,
										   hipMemcpyDeviceToHost, s ) );
			at += n;
		}
		HIP_CHECK( hipStreamSynchronize( s ) );
		std::sort( hChanges.begin(), hChanges.end(), []( const ChangeRecord& a, const ChangeRecord& b ) {
			return a.world != b.world ? a.world < b.world : a.contactId < b.contactId;
		} );
		std::vector<int> first, last;
		splitByWorld( hChanges, nw, first, last );
		HostPool::get().run( nw, [&]( int i ) {
			if ( last[i] > first[i] )
				applyChanges( worlds[i], hChanges.data() + first[i], last[i] - first[i] );
		} );
		// the changed worlds: their items may need rebuilding now, their lists next step
		for ( int i = 0; i < nw; ++i )
			if ( last[i] > first[i] )
			{
				markActive( worlds[i] );
				if ( worlds[i]->topologyDirty || !worlds[i]->dirtySims.empty() )
					changedTopology.push_back( worlds[i] );
			}
	}

	// 3. solver (one workgroup per work item); rebuild the items of worlds whose graph changed.
	// After a speculative launch only the chunks with state changes run again: the others have
	// already advanced their worlds.
	hostPhases[5] = msSince( t0 );
	if ( !speculate || changeCount > 0 )
	{
		if ( !changedTopology.empty() )
		{
			// the state changes marked more worlds
			std::vector<World*>& list = topology.list;
			list.insert( list.end(), changedTopology.begin(), changedTopology.end() );
			std::sort( list.begin(), list.end(), []( const World* a, const World* b ) { return a->groupIndex < b->groupIndex; } );
			list.erase( std::unique( list.begin(), list.end() ), list.end() );
		}
		if ( !topology.list.empty() )
		{
			HostPool::get().run( (int)topology.list.size(),
								 [&]( int i ) {
									 World* w = topology.list[i];
									 if ( w->topologyDirty )
										 buildWorldItems( w, target, budget, fullBudget, topology.list.size() == 1 );
									 else
										 rebuildDirtyItems( w, target, budget, fullBudget, topology.list.size() == 1 );
								 } );
			bool full = false;
			for ( World* w : topology.list )
				full = full || (int)w->solver.bodyList.size() > w->bodyRegion.cap ||
					   (int)w->solver.slotContact.size() > w->slotRegion.cap;
			if ( full )
				layoutGroup( this );
			publishLists( this, full ? nullptr : &topology.list, true );
		}
		B2MI_CHECK( chunkCount == chunks );
		if ( tape )
			tapeRecordStep( this, timeStep, subStepCount, blockSize > 0 ? std::min( blockSize, kAdjointMaxBlock ) : kAdjointMaxBlock );
		hostPhases[6] = msSince( t0 );
		HIP_CHECK( hipEventRecord( e3, s ) );
		if ( chunks == 1 )
			launchSolve( s, nullptr );
		else
		{
			int used = 0;
			HIP_CHECK( hipEventRecord( forkEvent, s ) );
			for ( int c = 0; c < chunks; ++c )
				if ( !speculate || chunkChanged[c] )
				{
					launchSolveChunk( c, forkEvent, nullptr );
					used += 1;
				}
			B2MI_CHECK( used > 0 );
			joinAux( s );
		}
		HIP_CHECK( hipEventRecord( e2, s ) );
		HIP_CHECK( hipMemcpyAsync( hCounters + 1, dCounters + 1, 3 * sizeof( int ), hipMemcpyDeviceToHost, s ) );
		hostPhases[7] = msSince( t0 );
		HIP_CHECK( hipStreamSynchronize( s ) );
	}
	else
		hostPhases[6] = hostPhases[7] = hostPhases[5];
	msChanges = std::chrono::duration<float, std::milli>( std::chrono::steady_clock::now() - tc ).count();
	hostPhases[8] = msSince( t0 );
	if ( hCounters[3] != 0 )
		fatal( "a barrier wait of the grid solver exceeded its bound", __FILE__, __LINE__ );
	int enlargeCount = hCounters[1];
	B2MI_CHECK( enlargeCount <= capEnlarge );
	auto te = std::chrono::steady_clock::now();
	if ( enlargeCount > 0 )
	{
		hEnlarge.resize( enlargeCount );
		HIP_CHECK( hipMemcpyAsync( hEnlarge.data(), dEnlarge, sizeof( EnlargeRecord ) * enlargeCount, hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipStreamSynchronize( s ) );
		// continuous collision (src/solver.c:1806-1937): the fast bodies sweep against the static
		// trees on the device (no host state is involved), then the enlarge pass on the host, then
		// the bullets against the trees it updated
		bool bullets = false;
		for ( const EnlargeRecord& r : hEnlarge )
			bullets = bullets || ( r.flags & ( kFastBody | kBulletFast ) ) == ( kFastBody | kBulletFast );
		runContinuous( false );
		sortByWorld( hEnlarge, nw );
		std::vector<int> first, last;
		splitByWorld( hEnlarge, nw, first, last );
		HostPool::get().run( nw, [&]( int i ) {
			if ( last[i] > first[i] )
				processEnlarged( worlds[i], hEnlarge.data() + first[i], last[i] - first[i], false );
		} );
		if ( bullets )
		{
			const size_t done = hEnlarge.size();
			runContinuous( true );
			std::vector<EnlargeRecord> added( hEnlarge.begin() + done, hEnlarge.end() );
			hEnlarge.resize( done );
			sortByWorld( added, nw );
			splitByWorld( added, nw, first, last );
			HostPool::get().run( nw, [&]( int i ) {
				if ( last[i] > first[i] )
					processEnlarged( worlds[i], added.data() + first[i], last[i] - first[i], true );
			} );
			hEnlarge.insert( hEnlarge.end(), added.begin(), added.end() );
		}
		for ( const EnlargeRecord& r : hEnlarge )
			markActive( worlds[r.world] );
	}
	msEnlarge = msSince( te );
	if ( hCounters[2] > 0 )
		for ( World* w : worlds )
			checkIslandSleep( w );
	hostPhases[9] = msSince( t0 );
	HIP_CHECK( hipEventElapsedTime( &msCollide, e0, e1 ) );
	HIP_CHECK( hipEventElapsedTime( &msSolve, e3, e2 ) );
	float msTotal = std::chrono::duration<float, std::milli>( std::chrono::steady_clock::now() - t0 ).count();
	msHost = msTotal - msCollide - msSolve;
	profile.step = msTotal;
	profile.pairs = msPairs;
	profile.collide = msCollide;
	profile.solve = msSolve;
	profile.broadphase = msBroad;
	profile.buildIslands = msChanges;
	profile.finalizeBodies = msEnlarge;
	// device state moved on: host mirrors and downloaded contacts of every world are stale, and
	// the move and hit events of the step are there to gather
	epoch += 1;
	for ( World* w : worlds )
		w->lazyEvents = true;
	stepping = false;
}

} // namespace b2mi

using namespace b2mi;
extern "C" {

B2_API b2ShapeId b2CreateCircleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Circle* circle )
{
	return createShape( bodyId, def, circle, b2_circleShape );
}

B2_API b2ShapeId b2CreateCapsuleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Capsule* capsule )
{
	float lengthSqr = b2DistanceSquared( capsule->center1, capsule->center2 );
	if ( lengthSqr <= kLinearSlop * kLinearSlop )
	{
		b2Circle circle = { b2Lerp( capsule->center1, capsule->center2, 0.5f ), capsule->radius };
		return createShape( bodyId, def, &circle, b2_circleShape );
	}
	return createShape( bodyId, def, capsule, b2_capsuleShape );
}

B2_API b2ShapeId b2CreatePolygonShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Polygon* polygon )
{
	return createShape( bodyId, def, polygon, b2_polygonShape );
}

B2_API b2ShapeId b2CreateSegmentShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Segment* segment )
{
	float lengthSqr = b2DistanceSquared( segment->point1, segment->point2 );
	if ( lengthSqr <= kLinearSlop * kLinearSlop )
		return b2ShapeId{ 0, 0, 0 };
	return createShape( bodyId, def, segment, b2_segmentShape );
}

B2_API bool b2Shape_IsValid( b2ShapeId id )
{
	World* w = b2mi_internal_worldOrNull( id.world0 );
	if ( w == nullptr || id.index1 < 1 || id.index1 > (int)w->shapes.size() )
		return false;
	const Shape& s = w->shapes[id.index1 - 1];
	return s.id == id.index1 - 1 && s.revision == id.revision;
}

static Shape* getShape( World* w, b2ShapeId id )
{
	int index = id.index1 - 1;
	B2MI_CHECK( 0 <= index && index < (int)w->shapes.size() );
	Shape* s = &w->shapes[index];
	B2MI_CHECK( s->id == index && s->revision == id.revision );
	return s;
}

#define B2MI_SHAPE( shapeId )                 \
	World* w = worldAt( ( shapeId ).world0 ); \
	Shape* shape = getShape( w, ( shapeId ) )

B2_API b2ShapeType b2Shape_GetType( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->type;
}

B2_API b2BodyId b2Shape_GetBody( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	const Body& b = w->bodies[shape->bodyId];
	return b2BodyId{ b.id + 1, shapeId.world0, b.revision };
}

B2_API float b2Shape_GetDensity( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->density;
}

// Friction and restitution are mixed into a contact when it is created
// (src/contact.c:323-324), so changing them affects contacts created afterwards.
// src/shape.c:1010-1065: contact events apply to contacts created afterwards, hit events from the
// next step's narrowphase on
B2_API void b2Shape_EnableContactEvents( b2ShapeId shapeId, bool flag )
{
	B2MI_SHAPE( shapeId );
	if ( worldLocked( w ) )
		return;
	shape->enableContactEvents = flag;
}

B2_API bool b2Shape_AreContactEventsEnabled( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->enableContactEvents;
}

B2_API void b2Shape_EnableHitEvents( b2ShapeId shapeId, bool flag )
{
	B2MI_SHAPE( shapeId );
	if ( worldLocked( w ) )
		return;
	b2mi_internal_beginEdit( w );
	shape->enableHitEvents = flag;
	if ( flag )
		w->devShapes[shape->id].flags |= kShapeHitEvents;
	else
		w->devShapes[shape->id].flags &= ~kShapeHitEvents;
}

B2_API bool b2Shape_AreHitEventsEnabled( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->enableHitEvents;
}

B2_API void b2Shape_SetFriction( b2ShapeId shapeId, float friction )
{
	B2MI_SHAPE( shapeId );
	b2mi_internal_beginEdit( w );
	shape->friction = friction;
	w->devShapes[shape->id].friction = friction;
}

B2_API float b2Shape_GetFriction( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->friction;
}

B2_API void b2Shape_SetRestitution( b2ShapeId shapeId, float restitution )
{
	B2MI_SHAPE( shapeId );
	b2mi_internal_beginEdit( w );
	shape->restitution = restitution;
	w->devShapes[shape->id].restitution = restitution;
}

B2_API float b2Shape_GetRestitution( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->restitution;
}

B2_API b2Filter b2Shape_GetFilter( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->filter;
}

B2_API b2Circle b2Shape_GetCircle( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->circle;
}

B2_API b2Segment b2Shape_GetSegment( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->segment;
}

B2_API b2Capsule b2Shape_GetCapsule( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->capsule;
}

B2_API b2Polygon b2Shape_GetPolygon( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return shape->polygon;
}

B2_API b2AABB b2Shape_GetAABB( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	b2mi_internal_downloadMirror( w );
	return shape->aabb;
}

B2_API int b2Shape_GetContactCapacity( b2ShapeId shapeId )
{
	B2MI_SHAPE( shapeId );
	return w->bodies[shape->bodyId].contactCount;
}

B2_API int b2Shape_GetContactData( b2ShapeId shapeId, b2ContactData* contactData, int capacity )
{
	B2MI_SHAPE( shapeId );
	b2mi_internal_downloadContacts( w );
	const Body& body = w->bodies[shape->bodyId];
	int index = 0;
	for ( int key = body.headContactKey; key != -1 && index < capacity; )
	{
		const Contact& c = w->contacts[key >> 1];
		if ( ( c.shapeIdA == shape->id || c.shapeIdB == shape->id ) && ( c.flags & kContactTouching ) )
		{
			const Shape& a = w->shapes[c.shapeIdA];
			const Shape& b = w->shapes[c.shapeIdB];
			contactData[index].shapeIdA = b2ShapeId{ a.id + 1, shapeId.world0, a.revision };
			contactData[index].shapeIdB = b2ShapeId{ b.id + 1, shapeId.world0, b.revision };
			contactData[index].manifold = publicManifold( w->devContacts[c.contactId], w->devContactOut[c.contactId] );
			index += 1;
		}
		key = c.edges[key & 1].nextKey;
	}
	return index;
}

// ---- step ----

// src/world.c:676-776
B2_API void b2World_Step( b2WorldId worldId, float timeStep, int subStepCount )
{
	World* w = getWorld( worldId );
	if ( worldLocked( w ) )
		return;
	w->profile = b2Profile{};
	if ( w->group )
		w->group->profile = b2Profile{};
	clearStepEvents( w );
	if ( timeStep == 0.0f )
		return;
	if ( w->group == nullptr )
	{
		w->group = new StepGroup();
		w->group->worlds.push_back( w );
		w->groupIndex = 0;
	}
	B2MI_CHECK( w->batchWorld == false && "batch worlds are stepped with b2mi_Batch_Step" );
	w->group->step( timeStep, subStepCount );
}

// ---- inspection ----

B2_API int b2mi_World_DumpContacts( b2WorldId worldId, b2mi_ContactRecord* records, int capacity )
{
	World* w = getWorld( worldId );
	b2mi_internal_downloadContacts( w );
	int n = 0;
	auto emit = [&]( int id, int color, int local ) {
		if ( n < capacity )
		{
			b2mi_ContactRecord& r = records[n];
			const DevContact& d = w->devContacts[id];
			r.contactId = id;
			r.shapeIdA = d.shapeIdA;
			r.shapeIdB = d.shapeIdB;
			r.colorIndex = color;
			r.localIndex = local;
			r.simFlags = d.simFlags;
			r.friction = d.friction;
			r.restitution = d.restitution;
			r.manifold = publicManifold( d, w->devContactOut[id] );
		}
		n += 1;
	};
	for ( int c = 0; c < kColorCount; ++c )
		for ( int i = 0; i < (int)w->colors[c].size(); ++i )
			emit( w->colors[c][i], c, i );
	for ( int i = 0; i < (int)w->nonTouching.size(); ++i )
		emit( w->nonTouching[i], -1, i );
	return n;
}

B2_API int b2mi_World_DumpBodies( b2WorldId worldId, b2mi_BodyRecord* records, int capacity )
{
	World* w = getWorld( worldId );
	b2mi_internal_downloadMirror( w );
	int n = 0;
	for ( int id = 0; id < (int)w->bodies.size(); ++id )
	{
		const Body& b = w->bodies[id];
		if ( b.setIndex == -1 )
			continue;
		if ( n < capacity )
		{
			b2mi_BodyRecord& r = records[n];
			memset( &r, 0, sizeof( r ) );
			const DevBodySim& sim = b.setIndex == 0 ? w->staticSims[b.localIndex] : w->awakeSims[b.localIndex];
			r.bodyId = id;
			r.type = b.type;
			r.setIndex = b.setIndex;
			r.localIndex = b.localIndex;
			r.transform = b2Transform{ { sim.p[0], sim.p[1] }, { sim.q[0], sim.q[1] } };
			r.center = b2Vec2{ sim.center[0], sim.center[1] };
			if ( b.setIndex == 2 )
			{
				const float* st = &w->awakeState[8 * b.localIndex];
				r.linearVelocity = b2Vec2{ st[0], st[1] };
				r.angularVelocity = st[2];
			}
		}
		n += 1;
	}
	return n;
}

} // extern "C"
